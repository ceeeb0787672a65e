// Base node class of the MI355X backend, API-compatible with FastNoise2's
// FastNoise::Generator (include/FastNoise/Generators/Generator.h:173-450 of
// the reference): the batch generation entry points keep their exact
// signatures and write the caller's host buffer; the work runs as one
// compiled device program per tree on the GPU.
#pragma once
#include <cstdint>
#include <cmath>
#include <atomic>
#include <algorithm>

#include "FastNoise/Utility/Config.h"

namespace FastNoise
{
    enum class Dim
    {
        X, Y, Z, W,
        Count
    };

    constexpr static const char* kDim_Strings[] = { "X", "Y", "Z", "W" };

    enum class DistanceFunction
    {
        Euclidean,
        EuclideanSquared,
        Manhattan,
        Hybrid,
        MaxAxis,
        Minkowski,
    };

    constexpr static const char* kDistanceFunction_Strings[] =
    {
        "Euclidean", "Euclidean Squared", "Manhattan", "Hybrid", "Max Axis", "Minkowski",
    };

    enum class Boolean
    {
        False,
        True
    };

    constexpr static const char* kBoolean_Strings[] = { "False", "True" };

    static constexpr float kInfinity = __builtin_inff();

    // Range of the values written by a batch call (Generator.h:85-105).
    struct OutputMinMax
    {
        float min = kInfinity;
        float max = -kInfinity;

        OutputMinMax& operator<<( float v )
        {
            min = std::min( min, v );
            max = std::max( max, v );
            return *this;
        }

        OutputMinMax& operator<<( const OutputMinMax& v )
        {
            min = std::min( min, v.min );
            max = std::max( max, v.max );
            return *this;
        }
    };

    template<typename T>
    struct BaseSource
    {
        using Type = T;
        SmartNode<const T> base;

    protected:
        BaseSource() = default;
    };

    // Required generator input (Generator.h:131-133)
    template<typename T>
    struct GeneratorSourceT : BaseSource<T>
    { };

    // Input that is either a constant or a generator (Generator.h:142-148)
    template<typename T>
    struct HybridSourceT : BaseSource<T>
    {
        float constant;

        constexpr HybridSourceT( float f = 0.0f ) : constant( f ) { }
    };

    namespace Internal
    {
        struct NodeDesc;
    }

    class FASTNOISE_API Generator
    {
    public:
        Generator() = default;
        Generator( const Generator& ) = delete;
        Generator( Generator&& ) = delete;

        virtual ~Generator() = default;

        // Reports FastSIMD::FeatureSet::GPU: the device backend is the only one.
        virtual FastSIMD::FeatureSet GetActiveFeatureSet() const;

        virtual const Metadata& GetMetadata() const = 0;

        virtual OutputMinMax GenUniformGrid2D( float* out,
            float xOffset, float yOffset,
            int xCount, int yCount,
            float xStepSize, float yStepSize,
            int seed ) const;

        virtual OutputMinMax GenUniformGrid3D( float* out,
            float xOffset, float yOffset, float zOffset,
            int xCount, int yCount, int zCount,
            float xStepSize, float yStepSize, float zStepSize,
            int seed ) const;

        virtual OutputMinMax GenUniformGrid4D( float* out,
            float xOffset, float yOffset, float zOffset, float wOffset,
            int xCount, int yCount, int zCount, int wCount,
            float xStepSize, float yStepSize, float zStepSize, float wStepSize,
            int seed ) const;

        virtual OutputMinMax GenTileable2D( float* out,
            int xSize, int ySize,
            float xStepSize, float yStepSize,
            int seed ) const;

        virtual OutputMinMax GenPositionArray2D( float* out, int count,
            const float* xPosArray, const float* yPosArray,
            float xOffset, float yOffset, int seed ) const;

        virtual OutputMinMax GenPositionArray3D( float* out, int count,
            const float* xPosArray, const float* yPosArray, const float* zPosArray,
            float xOffset, float yOffset, float zOffset, int seed ) const;

        virtual OutputMinMax GenPositionArray4D( float* out, int count,
            const float* xPosArray, const float* yPosArray, const float* zPosArray, const float* wPosArray,
            float xOffset, float yOffset, float zOffset, float wOffset, int seed ) const;

        virtual float GenSingle2D( float x, float y, int seed ) const;
        virtual float GenSingle3D( float x, float y, float z, int seed ) const;
        virtual float GenSingle4D( float x, float y, float z, float w, int seed ) const;

        // Fills `desc` with this node's id, parameters and inputs; the tree
        // compiler walks a tree through it.
        virtual void Describe( Internal::NodeDesc& desc ) const = 0;

    protected:
        template<typename T>
        void SetSourceMemberVariable( BaseSource<T>& memberVariable, SmartNodeArg<T> gen )
        {
            static_assert( std::is_base_of<Generator, T>::value, "T must be child of FastNoise::Generator class" );
            memberVariable.base = gen;
        }

    private:
        friend FASTNOISE_API int32_t Internal::NodeRefAdd( const Generator*, int32_t );
        mutable std::atomic<int32_t> mReferences{ 0 };
    };

    using GeneratorSource = GeneratorSourceT<Generator>;
    using HybridSource = HybridSourceT<Generator>;

    // One value per axis (Generator.h:463-491)
    template<typename T>
    struct PerDimensionVariable
    {
        using Type = T;

        T varArray[(int)Dim::Count];

        template<typename U = T>
        constexpr PerDimensionVariable( U value = 0 )
        {
            for( T& element : varArray )
            {
                element = value;
            }
        }

        T& operator[]( size_t i ) { return varArray[i]; }
        const T& operator[]( size_t i ) const { return varArray[i]; }
    };
}
