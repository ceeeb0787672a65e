// Modifier nodes (FastNoise2 include/FastNoise/Generators/Modifiers.h).
#pragma once
#include "Generator.h"
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API DomainScale : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetScaling( float value ) { mScale = value; }

    protected:
        GeneratorSource mSource;
        float mScale = 1.0f;
    };

    class FASTNOISE_API DomainOffset : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

        template<Dim D>
        void SetOffset( float value ) { mOffset[(int)D] = value; }
        template<Dim D>
        void SetOffset( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mOffset[(int)D], gen ); }

        void SetOffsetIdx( int d, float v ) { mOffset[d] = v; }
        void SetOffsetIdx( int d, SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mOffset[d], gen ); }

    protected:
        GeneratorSource mSource;
        PerDimensionVariable<HybridSource> mOffset = 0.0f;
    };

    class FASTNOISE_API DomainRotate : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

        void SetYaw( float value ) { mYawCos = std::cos( value ); mYawSin = std::sin( value ); CalculateRotation(); }
        void SetPitch( float value ) { mPitchCos = std::cos( value ); mPitchSin = std::sin( value ); CalculateRotation(); }
        void SetRoll( float value ) { mRollCos = std::cos( value ); mRollSin = std::sin( value ); CalculateRotation(); }

    protected:
        GeneratorSource mSource;
        float mYawCos = 1.0f;
        float mYawSin = 0.0f;
        float mPitchCos = 1.0f;
        float mPitchSin = 0.0f;
        float mRollCos = 1.0f;
        float mRollSin = 0.0f;

        float mXa = 1.0f;
        float mXb = 0.0f;
        float mXc = 0.0f;
        float mYa = 0.0f;
        float mYb = 1.0f;
        float mYc = 0.0f;
        float mZa = 0.0f;
        float mZb = 0.0f;
        float mZc = 1.0f;

        // Yaw-pitch-roll matrix (Modifiers.h:107-121 of the reference)
        void CalculateRotation()
        {
            mXa = mYawCos * mPitchCos;
            mXb = mYawCos * mPitchSin * mRollSin - mYawSin * mRollCos;
            mXc = mYawCos * mPitchSin * mRollCos + mYawSin * mRollSin;

            mYa = mYawSin * mPitchCos;
            mYb = mYawSin * mPitchSin * mRollSin + mYawCos * mRollCos;
            mYc = mYawSin * mPitchSin * mRollCos - mYawCos * mRollSin;

            mZa = -mPitchSin;
            mZb = mPitchCos * mRollSin;
            mZc = mPitchCos * mRollCos;
        }
    };

    class FASTNOISE_API SeedOffset : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetOffset( int value ) { mOffset = value; }

    protected:
        GeneratorSource mSource;
        int mOffset = 1;
    };

    class FASTNOISE_API Remap : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

        void SetFromMin( float value ) { mFromMin = value; }
        void SetFromMin( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mFromMin, gen ); }

        void SetFromMax( float value ) { mFromMax = value; }
        void SetFromMax( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mFromMax, gen ); }

        void SetToMin( float value ) { mToMin = value; }
        void SetToMin( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mToMin, gen ); }

        void SetToMax( float value ) { mToMax = value; }
        void SetToMax( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mToMax, gen ); }

        void SetClampOutput( Boolean value ) { mClampOutput = value; }
        void SetClampOutput( bool value ) { mClampOutput = value ? Boolean::True : Boolean::False; }

    protected:
        GeneratorSource mSource;
        HybridSource mFromMin = -1.0f;
        HybridSource mFromMax = 1.0f;
        HybridSource mToMin = 0.0f;
        HybridSource mToMax = 1.0f;
        Boolean mClampOutput = Boolean::False;
    };

    class FASTNOISE_API ConvertRGBA8 : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetMinMax( float min, float max ) { mMin = min; mMax = max; }
        void SetMin( float v ) { mMin = v; }
        void SetMax( float v ) { mMax = v; }

    protected:
        GeneratorSource mSource;
        float mMin = -1.0f;
        float mMax = 1.0f;
    };

    class FASTNOISE_API Terrace : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetStepCount( float stepCount ) { mStepCount = stepCount; mStepCountRecip = 1 / stepCount; }

        void SetSmoothness( float smoothness ) { mSmoothness = smoothness; if( smoothness != 0.0f ) mSmoothnessRecip = 1 + 1 / smoothness; }
        void SetSmoothness( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSmoothness, gen ); }

    protected:
        GeneratorSource mSource;
        HybridSource mSmoothness = 0.0f;
        float mSmoothnessRecip = 0.0f;
        float mStepCount = 1.0f;
        float mStepCountRecip = 1.0f;
    };

    class FASTNOISE_API PingPong : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

        void SetPingPongStrength( float value ) { mPingPongStrength = value; }
        void SetPingPongStrength( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mPingPongStrength, gen ); }

    protected:
        GeneratorSource mSource;
        HybridSource mPingPongStrength = 2.0f;
    };

    class FASTNOISE_API DomainAxisScale : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

        template<Dim D>
        void SetScaling( float value ) { mScale[(int)D] = value; }
        void SetScalingIdx( int d, float value ) { mScale[d] = value; }

    protected:
        GeneratorSource mSource;
        PerDimensionVariable<float> mScale = 1.0f;
    };

    class FASTNOISE_API AddDimension : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetNewDimensionPosition( float value ) { mNewDimensionPosition = value; }
        void SetNewDimensionPosition( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mNewDimensionPosition, gen ); }

    protected:
        GeneratorSource mSource;
        HybridSource mNewDimensionPosition = 0.0f;
    };

    class FASTNOISE_API RemoveDimension : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetRemoveDimension( Dim dimension ) { mRemoveDimension = dimension; }

    protected:
        GeneratorSource mSource;
        Dim mRemoveDimension = Dim::Y;
    };

    class FASTNOISE_API GeneratorCache : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

    protected:
        GeneratorSource mSource;
    };

    class FASTNOISE_API SignedSquareRoot : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

    protected:
        GeneratorSource mSource;
    };

    class FASTNOISE_API Abs : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }

    protected:
        GeneratorSource mSource;
    };

    enum class PlaneRotationType
    {
        ImproveXYPlanes,
        ImproveXZPlanes
    };

    class FASTNOISE_API DomainRotatePlane : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSource( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetRotationType( PlaneRotationType type ) { mRotationType = type; }

    protected:
        GeneratorSource mSource;
        PlaneRotationType mRotationType = PlaneRotationType::ImproveXYPlanes;
    };
}
