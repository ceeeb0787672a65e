// Fractal nodes (FastNoise2 include/FastNoise/Generators/Fractal.h).
#pragma once
#include "Generator.h"
#include "BasicGenerators.h"

namespace FastNoise
{
    template<typename T = Generator>
    class Fractal : public virtual Generator
    {
    public:
        void SetSource( SmartNodeArg<T> gen ) { this->SetSourceMemberVariable( mSource, gen ); }
        void SetGain( float value ) { mGain = value; }
        void SetGain( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mGain, gen ); }
        void SetWeightedStrength( float value ) { mWeightedStrength = value; }
        void SetWeightedStrength( SmartNodeArg<> gen ) { this->SetSourceMemberVariable( mWeightedStrength, gen ); }
        void SetOctaveCount( int value ) { mOctaves = value; }
        void SetLacunarity( float value ) { mLacunarity = value; }

    protected:
        GeneratorSourceT<T> mSource;
        HybridSource mGain = 0.5f;
        HybridSource mWeightedStrength = 0.0f;

        int mOctaves = 3;
        float mLacunarity = 2.0f;
    };

    class FASTNOISE_API FractalFBm : public virtual Fractal<>
    {
    public:
        FASTNOISE_NODE_DECL
    };

    class FASTNOISE_API FractalRidged : public virtual Fractal<>
    {
    public:
        FASTNOISE_NODE_DECL
    };
}
