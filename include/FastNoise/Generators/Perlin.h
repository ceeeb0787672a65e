// Perlin noise node (FastNoise2 include/FastNoise/Generators/Perlin.h).
#pragma once
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API Perlin : public virtual VariableRange<Seeded<ScalableGenerator>>
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
