// Fractal domain warps (FastNoise2 include/FastNoise/Generators/DomainWarpFractal.h).
#pragma once
#include "Fractal.h"
#include "DomainWarp.h"

namespace FastNoise
{
    class FASTNOISE_API DomainWarpFractalProgressive : public virtual Fractal<DomainWarp>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API DomainWarpFractalIndependent : public virtual Fractal<DomainWarp>
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
