// Cellular noise nodes (FastNoise2 include/FastNoise/Generators/Cellular.h).
#pragma once
#include "BasicGenerators.h"

namespace FastNoise
{
    template<typename PARENT = VariableRange<Seeded<ScalableGenerator>>>
    class Cellular : public virtual PARENT
    {
    public:
        void SetDistanceFunction( DistanceFunction value ) { mDistanceFunction = value; }

        void SetMinkowskiP( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mMinkowskiP, gen ); }
        void SetMinkowskiP( float value ) { mMinkowskiP = value; }

        void SetGridJitter( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mGridJitter, gen ); }
        void SetGridJitter( float value ) { mGridJitter = value; }

        void SetSizeJitter( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mSizeJitter, gen ); }
        void SetSizeJitter( float value ) { mSizeJitter = value; }

    protected:
        HybridSource mMinkowskiP = 1.5f;
        HybridSource mGridJitter = 1.0f;
        HybridSource mSizeJitter = 0.f;
        DistanceFunction mDistanceFunction = DistanceFunction::EuclideanSquared;
    };

    class FASTNOISE_API CellularValue : public virtual Cellular<>
    {
    public:
        FASTNOISE_NODE_DECL

        static const int kMaxDistanceCount = 4;

        void SetValueIndex( int value ) { mValueIndex = std::min( std::max( value, 0 ), kMaxDistanceCount - 1 ); }

    protected:
        int mValueIndex = 0;
    };

    class FASTNOISE_API CellularDistance : public virtual Cellular<>
    {
    public:
        FASTNOISE_NODE_DECL

        enum class ReturnType
        {
            Index0,
            Index0Add1,
            Index0Sub1,
            Index0Mul1,
            Index0Div1
        };

        static const int kMaxDistanceCount = 4;

        void SetDistanceIndex0( int value ) { mDistanceIndex0 = std::min( std::max( value, 0 ), kMaxDistanceCount - 1 ); }
        void SetDistanceIndex1( int value ) { mDistanceIndex1 = std::min( std::max( value, 0 ), kMaxDistanceCount - 1 ); }
        void SetReturnType( ReturnType value ) { mReturnType = value; }

    protected:
        ReturnType mReturnType = ReturnType::Index0;
        int mDistanceIndex0 = 0;
        int mDistanceIndex1 = 1;
    };

    class FASTNOISE_API CellularLookup : public virtual Cellular<Seeded<ScalableGenerator>>
    {
    public:
        FASTNOISE_NODE_DECL

        void SetLookup( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mLookup, gen ); }

    protected:
        GeneratorSource mLookup;
    };
}
