// Blend and operator nodes (FastNoise2 include/FastNoise/Generators/Blends.h).
#pragma once
#include "Generator.h"
#include "BasicGenerators.h"

namespace FastNoise
{
    class FASTNOISE_API OperatorSourceLHS : public virtual Generator
    {
    public:
        void SetLHS( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mLHS, gen ); }
        void SetRHS( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mRHS, gen ); }
        void SetRHS( float value ) { mRHS = value; }

    protected:
        GeneratorSource mLHS;
        HybridSource mRHS = 0.0f;
    };

    class FASTNOISE_API OperatorHybridLHS : public virtual Generator
    {
    public:
        void SetLHS( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mLHS, gen ); }
        void SetLHS( float value ) { mLHS = value; }
        void SetRHS( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mRHS, gen ); }
        void SetRHS( float value ) { mRHS = value; }

    protected:
        HybridSource mLHS = 0.0f;
        HybridSource mRHS = 0.0f;
    };

    class FASTNOISE_API Add : public virtual OperatorSourceLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Subtract : public virtual OperatorHybridLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Multiply : public virtual OperatorSourceLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Divide : public virtual OperatorHybridLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Modulus : public virtual OperatorHybridLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Min : public virtual OperatorSourceLHS { public: FASTNOISE_NODE_DECL };
    class FASTNOISE_API Max : public virtual OperatorSourceLHS { public: FASTNOISE_NODE_DECL };

    class FASTNOISE_API PowFloat : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetValue( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mValue, gen ); }
        void SetValue( float value ) { mValue = value; }
        void SetPow( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mPow, gen ); }
        void SetPow( float value ) { mPow = value; }

    protected:
        HybridSource mValue = 2.0f;
        HybridSource mPow = 2.0f;
    };

    class FASTNOISE_API PowInt : public virtual Generator
    {
    public:
        FASTNOISE_NODE_DECL
        void SetValue( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mValue, gen ); }
        void SetPow( int value ) { mPow = value; }

    protected:
        GeneratorSource mValue;
        int mPow = 2;
    };

    class FASTNOISE_API MinSmooth : public virtual OperatorSourceLHS
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSmoothness( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSmoothness, gen ); }
        void SetSmoothness( float value ) { mSmoothness = value; }

    protected:
        HybridSource mSmoothness = 0.1f;
    };

    class FASTNOISE_API MaxSmooth : public virtual OperatorSourceLHS
    {
    public:
        FASTNOISE_NODE_DECL
        void SetSmoothness( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSmoothness, gen ); }
        void SetSmoothness( float value ) { mSmoothness = value; }

    protected:
        HybridSource mSmoothness = 0.1f;
    };

    class FASTNOISE_API Fade : public virtual Generator
    {
    public:
        enum class Interpolation
        {
            Linear,
            Hermite,
            Quintic,
        };

        FASTNOISE_NODE_DECL
        void SetA( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mA, gen ); }
        void SetB( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mB, gen ); }

        void SetFade( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mFade, gen ); }
        void SetFade( float value ) { mFade = value; }

        void SetFadeMin( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mFadeMin, gen ); }
        void SetFadeMin( float value ) { mFadeMin = value; }

        void SetFadeMax( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mFadeMax, gen ); }
        void SetFadeMax( float value ) { mFadeMax = value; }

        void SetInterpolation( Interpolation interpolation ) { mInterpolation = interpolation; }

    protected:
        GeneratorSource mA;
        GeneratorSource mB;
        HybridSource mFade = 0;
        HybridSource mFadeMin = -1.f;
        HybridSource mFadeMax = 1.f;
        Interpolation mInterpolation = Interpolation::Linear;
    };
}
