// Test infrastructure only: node records of the CPU oracle.
//
// The records are plain data shared by every build of the evaluator (the
// scalar checker and the AVX2 / AVX-512 lane-16 builds live in different
// namespaces but read the same records).
#pragma once
#include <cstdint>
#include <cfloat>
#include <cmath>
#include "fn2_oracle_lanes.h"

namespace fn2o_data
{
    // Node ids in registration order (src/FastNoise/FastSIMD_Build.inl:86-143).
    enum NodeId
    {
        kConstant, kWhite, kCheckerboard, kSineWave, kGradient, kDistanceToPoint,
        kSimplex, kSuperSimplex, kPerlin, kValue,
        kCellularValue, kCellularDistance, kCellularLookup,
        kFractalFBm, kPingPong, kFractalRidged,
        kDomainWarpSimplex, kDomainWarpSuperSimplex, kDomainWarpGradient,
        kDomainWarpFractalProgressive, kDomainWarpFractalIndependent,
        kAdd, kSubtract, kMultiply, kDivide,
        kAbs, kMin, kMax, kMinSmooth, kMaxSmooth, kSignedSquareRoot, kPowFloat, kPowInt,
        kDomainScale, kDomainOffset, kDomainRotate, kDomainAxisScale,
        kSeedOffset, kConvertRGBA8, kGeneratorCache,
        kFade, kRemap, kTerrace, kAddDimension, kRemoveDimension,
        kModulus, kDomainRotatePlane,
        kNodeCount
    };

    struct Node;
    struct Hybrid
    {
        const Node* node = nullptr;
        float c = 0.0f;
    };

    // One record type for all nodes; each node type uses the fields its
    // reference class declares (defaults from the reference headers, set by
    // make_node in fn2_oracle.cpp).
    //   Cellular*:    i[0] distance function, i[1] value index / distance index 0,
    //                 i[2] distance index 1, i[3] return type;
    //                 hyb[0] Minkowski P, hyb[1] grid jitter, hyb[2] size jitter
    //   Fractal*:     hyb[0] gain, hyb[1] weighted strength, i[0] octaves, f[0] lacunarity
    //   DomainWarp*:  hyb[0] warp amplitude, f[0..3] amplitude scaling, i[0] vectorization scheme
    //   Operators:    src[0] LHS + hyb[0] RHS, or hyb[0] LHS + hyb[1] RHS; hyb[1] smoothness
    //   DomainRotate: f[0..5] yaw/pitch/roll cos,sin; f[6..14] rotation matrix Xa..Zc
    //   Terrace:      f[0] step count, f[1] its reciprocal, f[2] 1 + 1/smoothness, hyb[0] smoothness
    struct Node
    {
        int id = 0;
        // ScalableGenerator (BasicGenerators.h:20-28)
        float scale = 100.0f;
        float frequency = 1.0f / 100.0f;
        // Seeded (BasicGenerators.h:58-64)
        int seedOffset = 0;
        // VariableRange (BasicGenerators.h:96-112)
        float rangeMin = -1.0f;
        float rangeScale = 2.0f;
        // generator lookups: src[0] Source / LHS / A / Lookup / Value, src[1] B (Fade)
        const Node* src[2] = { nullptr, nullptr };
        Hybrid hyb[6];
        float f[16] = { 0 };
        int i[6] = { 0 };
    };

    Node make_node( int id );

    struct DriverArgs
    {
        int kind = 0;            // 0 grid, 1 position array, 2 tileable
        int dims = 2;
        const Node* root = nullptr;
        float* out = nullptr;
        int64_t total = 0;
        int count[4] = { 1, 1, 1, 1 };
        float offset[4] = { 0, 0, 0, 0 };
        float step[4] = { 1, 1, 1, 1 };
        const float* posArrays[4] = { nullptr, nullptr, nullptr, nullptr };
        int seed = 0;
    };

    // NaN-skipping running min/max (the reference's OutputMinMax, Generator.h:85-105)
    struct MinMax
    {
        float mn = INFINITY, mx = -INFINITY;
        void add( float v ) { if( v < mn ) mn = v; if( v > mx ) mx = v; }
    };
}

namespace FN2O_NS
{
    using namespace fn2o_data;
}
