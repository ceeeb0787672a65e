// Simplex-lattice domain warps (FastNoise2 include/FastNoise/Generators/DomainWarpSimplex.h).
#pragma once
#include "Generator.h"
#include "DomainWarp.h"

namespace FastNoise
{
    enum class VectorizationScheme
    {
        OrthogonalGradientMatrix,
        GradientOuterProduct
    };

    constexpr static const char* kVectorizationScheme_Strings[] =
    {
        "Orthogonal Gradient Matrix",
        "Gradient Outer Product",
    };

    class FASTNOISE_API DomainWarpSimplex : public virtual DomainWarp
    {
    public:
        FASTNOISE_NODE_DECL

        void SetVectorizationScheme( VectorizationScheme value ) { mVectorizationScheme = value; }

    protected:
        VectorizationScheme mVectorizationScheme = VectorizationScheme::OrthogonalGradientMatrix;
    };

    class FASTNOISE_API DomainWarpSuperSimplex : public virtual DomainWarpSimplex
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
