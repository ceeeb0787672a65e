// Basic generators and the shared node mixins, API-compatible with
// FastNoise2 include/FastNoise/Generators/BasicGenerators.h.
#pragma once
#include "Generator.h"

// Declares the two per-class overrides every concrete node defines.
#define FASTNOISE_NODE_DECL \
    const Metadata& GetMetadata() const override; \
    void Describe( Internal::NodeDesc& desc ) const override;

namespace FastNoise
{
    // Feature scale -> frequency (BasicGenerators.h:14-29)
    class FASTNOISE_API ScalableGenerator : public virtual Generator
    {
    public:
        void SetScale( float value )
        {
            mScale = value;
            mFrequency = 1.0f / value;
        }

    protected:
        float mScale = 100;
        float mFrequency = 1.0f / 100;
    };

    // Per-node seed offset (BasicGenerators.h:51-65)
    template<typename PARENT>
    class Seeded : public virtual PARENT
    {
    public:
        void SetSeedOffset( int value ) { mSeedOffset = value; }

    protected:
        int mSeedOffset = 0;
    };

    // Output range remap (BasicGenerators.h:89-113)
    template<typename PARENT>
    class VariableRange : public virtual PARENT
    {
    public:
        void SetOutputMin( float value )
        {
            mRangeScale += mRangeMin - value;
            mRangeMin = value;
        }

        void SetOutputMax( float value )
        {
            mRangeScale = ( value - mRangeMin );
        }

    protected:
        float mRangeMin = -1;
        float mRangeScale = 2;
    };

    class FASTNOISE_API Constant : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetValue( float value ) { mValue = value; }

    protected:
        float mValue = 1.0f;
    };

    class FASTNOISE_API White : public virtual VariableRange<Seeded<Generator>>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API Checkerboard : public virtual VariableRange<ScalableGenerator>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API SineWave : public virtual VariableRange<ScalableGenerator>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API Gradient : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL

        template<Dim D>
        void SetMultiplier( float multiplier ) { mMultiplier[(int)D] = multiplier; }
        template<Dim D>
        void SetOffset( float offset ) { mOffset[(int)D] = offset; }
        template<Dim D>
        void SetOffset( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mOffset[(int)D], gen ); }

        // metadata helpers
        void SetMultiplierIdx( int d, float v ) { mMultiplier[d] = v; }
        void SetOffsetIdx( int d, float v ) { mOffset[d] = v; }
        void SetOffsetIdx( int d, SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mOffset[d], gen ); }

    protected:
        PerDimensionVariable<float> mMultiplier = 0.0f;
        PerDimensionVariable<HybridSource> mOffset = 0.0f;
    };

    class FASTNOISE_API DistanceToPoint : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL

        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetDistanceFunction( DistanceFunction value ) { mDistanceFunction = value; }

        void SetMinkowskiP( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mMinkowskiP, gen ); }
        void SetMinkowskiP( float value ) { mMinkowskiP = value; }

        void SetPoint( float x, float y, float z = 0, float w = 0 )
        {
            mPoint[0] = x;
            mPoint[1] = y;
            mPoint[2] = z;
            mPoint[3] = w;
        }

        template<Dim D>
        void SetPoint( float value ) { mPoint[(int)D] = value; }

        template<Dim D>
        void SetPoint( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mPoint[(int)D], gen ); }

        void SetPointIdx( int d, float v ) { mPoint[d] = v; }
        void SetPointIdx( int d, SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mPoint[d], gen ); }

    protected:
        GeneratorSource mSource;
        HybridSource mMinkowskiP = 1.5f;
        DistanceFunction mDistanceFunction = DistanceFunction::EuclideanSquared;
        PerDimensionVariable<HybridSource> mPoint = 0.0f;
    };
}
