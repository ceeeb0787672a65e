// Domain warp base and gradient warp (FastNoise2 include/FastNoise/Generators/DomainWarp.h).
#pragma once
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API DomainWarp : public virtual Seeded<ScalableGenerator>
    {
    public:
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetWarpAmplitude( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mWarpAmplitude, gen ); }
        void SetWarpAmplitude( float value ) { mWarpAmplitude = value; }

        template<Dim D>
        void SetAmplitudeScaling( float value ) { mAxisScale[(int)D] = value; }
        void SetAmplitudeScalingIdx( int d, float value ) { mAxisScale[d] = value; }

    protected:
        GeneratorSource mSource;
        HybridSource mWarpAmplitude = 50.0f;
        PerDimensionVariable<float> mAxisScale = 1.0f;

        friend struct DomainWarpAccess;
    };

    class FASTNOISE_API DomainWarpGradient : public virtual DomainWarp
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
