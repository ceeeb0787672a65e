// Simplex and SuperSimplex noise nodes (FastNoise2 include/FastNoise/Generators/Simplex.h).
#pragma once
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API Simplex : public virtual VariableRange<Seeded<ScalableGenerator>>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API SuperSimplex : public virtual VariableRange<Seeded<ScalableGenerator>>
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
