// Forward declarations shared by the public headers (FastNoise2 Utility/Config.h).
#pragma once
#include "Export.h"
#include <FastSIMD/FeatureSet.h>

#define FASTNOISE_CALC_MIN_MAX true

namespace FastNoise
{
    class Generator;
    struct Metadata;

    template<typename T = Generator>
    class SmartNode;

    template<typename T = Generator>
    using SmartNodeArg = const SmartNode<const T>&;

    template<typename T>
    SmartNode<T> New( FastSIMD::FeatureSet maxFeatureSet = FastSIMD::FeatureSet::Max );
}

#include "SmartNode.h"
