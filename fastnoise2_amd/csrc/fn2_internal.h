// Internal view of a node used by the tree compiler: node id (registration
// order of FastNoise2, src/FastNoise/FastSIMD_Build.inl:86-143), inputs and
// the parameters in the form the device program needs.
#pragma once
#include <FastNoise/FastNoise.h>
#include <FastNoise/Metadata.h>

namespace FastNoise
{
    namespace Internal
    {
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

        struct HybridDesc
        {
            const Generator* node = nullptr;
            float c = 0.0f;
        };

        // Field use per node kind:
        //   lattice/cellular/warp: scale, frequency, seedOffset, rangeMin, rangeScale
        //   Cellular*: i[0] distance function, i[1] value index / distance index 0,
        //              i[2] distance index 1, i[3] return type; hyb[0] Minkowski P,
        //              hyb[1] grid jitter, hyb[2] size jitter; src[0] lookup
        //   Fractal*: src[0] source, hyb[0] gain, hyb[1] weighted strength, i[0] octaves, f[0] lacunarity
        //   DomainWarp*: src[0] source, hyb[0] warp amplitude, f[0..3] axis scale, i[0] scheme
        //   operators: src[0] LHS + hyb[0] RHS, or hyb[0] LHS + hyb[1] RHS, hyb[1] smoothness
        //   DomainRotate: f[0..5] yaw/pitch/roll cos,sin; f[6..14] matrix
        //   Terrace: f[0] step count, f[1] step reciprocal, f[2] smoothness reciprocal
        struct NodeDesc
        {
            int id = -1;
            float scale = 100.0f;
            float frequency = 0.01f;
            int seedOffset = 0;
            float rangeMin = -1.0f;
            float rangeScale = 2.0f;
            const Generator* src[2] = { nullptr, nullptr };
            HybridDesc hyb[6];
            float f[16] = {};
            int i[6] = {};
        };

        inline HybridDesc Hyb( const HybridSource& h ) { HybridDesc d; d.node = h.base.get(); d.c = h.constant; return d; }
    }
}
