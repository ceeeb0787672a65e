// Value noise node (FastNoise2 include/FastNoise/Generators/Value.h).
#pragma once
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API Value : public virtual VariableRange<Seeded<ScalableGenerator>>
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
