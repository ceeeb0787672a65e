// Device program format shared by the tree compiler (host) and the HIP
// interpreter kernels (device).
//
// A node tree is compiled into a straight-line list of instructions over a
// per-point register file of kMaxRegs floats that lives in VGPRs (all register
// indices are wave-uniform, so the kernel indexes the file with GPR-indexing
// mode and never touches scratch memory). Seeds are wave-uniform: every seed
// in a tree is the call's seed plus a constant the compiler folds into the
// instruction (`seed` field). Node parameters live in a float constant pool
// (`k` indexes it); small integer parameters ride in `p[]`.
#pragma once
#if !defined( __HIPCC_RTC__ )
#include <stdint.h>
#else
// hipRTC has no <stdint.h>
typedef signed char int8_t;
typedef unsigned char uint8_t;
typedef short int16_t;
typedef unsigned short uint16_t;
typedef int int32_t;
typedef unsigned int uint32_t;
typedef long long int64_t;
typedef unsigned long long uint64_t;
#endif

namespace fn2
{
    static constexpr int kMaxRegs = 32;
    static constexpr int kMaxInstructions = 4096;

    enum Op : uint8_t
    {
        OP_END = 0,

        // ---- leaf noises. r[0] dst, r[1] first of dim consecutive position registers.
        // k: [frequency, rangeMin, rangeScale] (Checkerboard/SineWave/White use the same layout)
        OP_PERLIN, OP_VALUE, OP_SIMPLEX, OP_SUPERSIMPLEX,
        OP_WHITE, OP_CHECKERBOARD, OP_SINEWAVE,

        // ---- cellular. r[0] dst (OP_CELL_LOOKUP: first of dim dst registers holding
        // cell position * scale), r[1] positions, r[2] grid-jitter register,
        // r[3] size-jitter register (kNoReg when inactive).
        // p[0] distance function, p[1] value index / distance index 0, p[2] distance index 1, p[3] return type
        // k: [frequency, nativeMin, outScale, minkowskiP, rangeMin, scale]
        OP_CELL_VALUE, OP_CELL_DISTANCE, OP_CELL_LOOKUP,

        // ---- domain warps. r[0] dst squared strength, r[1] scaled positions (in),
        // r[2] positions to displace (in/out), r[3] amplitude register.
        // p[0] vectorization scheme. k: [axis scale x, y, z, w]
        OP_WARP_GRADIENT, OP_WARP_SIMPLEX, OP_WARP_SUPERSIMPLEX,

        // ---- register arithmetic, r[0] = f( r[1], r[2], r[3] )
        OP_CONST,       // r[0] = k[0]
        OP_MOV,         // r[0] = r[1]
        OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MIN, OP_MAX, OP_FMOD,
        OP_ABS,         // r[0] = |r[1]|
        OP_SCALE,       // r[0..dim) = r[1..dim) * k[0]          (DomainScale, fractal lacunarity)
        OP_SCALE_AXES,  // r[0..dim) = r[1..dim) * k[0..dim)     (DomainAxisScale, warp frequency)
        OP_MUL_VEC,     // r[0..dim) = r[1..dim) * r[2]           (warp-fractal frequency)
        OP_ADD_VEC,     // r[0..dim) = r[1..dim) + r[2..2+dim)    (DomainOffset)
        OP_SUB_VEC,     // r[0..dim) = r[1..dim) - r[2..2+dim)    (DistanceToPoint)
        OP_MOV_VEC,     // r[0..dim) = r[1..dim)
        OP_FBM_AMP,     // r[0] = r[0] * lerp( 1, (r[1]+1)*0.5, r[2] ) * r[3]   (Fractal.inl:28-29)
        OP_RIDGED_AMP,  // r[0] = r[0] * lerp( 1, 1 - r[1], r[2] ) * r[3]       (Fractal.inl:58-59)
        OP_WARP_AMP,    // r[0] = r[0] * lerp( 1, 1 - sqrt(r[1]), r[2] ) * r[3] (DomainWarpFractal.inl:26-30)
        OP_FBM_SUM,     // r[0] = r[0] + r[1] * r[2]
        OP_RIDGED_TERM, // r[0] = ( r[1] * -2 + 1 ) * r[2]   (p[0]=1: r[0] += that)
        OP_PINGPONG,    // r[0] = pingpong( r[1] * r[2] )
        OP_MINSMOOTH,   // r[0] = smooth min/max( r[1], r[2], smoothness r[3] ), p[0]=1 max
        OP_FADE,        // r[0] = lerp( r[1], r[2], fade(r[3], r[4], r[5]) ), p[0] interpolation
        OP_REMAP,       // r[0] = remap( r[1]; from r[2], r[3]; to r[4], r[5] ), p[0] clamp
        OP_TERRACE,     // r[0] = terrace( r[1] ), r[2] smoothness register or kNoReg; k: [stepCount, stepRecip, smoothRecip]
        OP_POWF,        // r[0] = pow( max( |r[1]|, FLT_MIN ), r[2] )
        OP_POWI,        // r[0] = r[1]^p[0]
        OP_SSQRT,       // r[0] = sqrt(|r[1]|) with the sign of r[1]
        OP_RGBA8,       // k: [min, max, 255/(max-min)]
        OP_ROTATE2,     // r[0..2) = 2D yaw rotation of r[1..3); k: [yawCos, yawSin]
        OP_ROTATE3,     // r[0..3) = matrix * r[1..4); k: 9 matrix entries
        OP_ROTATE_PLANE,// r[0..3) = plane rotation of r[1..4); p[0] type
        OP_GRADIENT_AXIS, // r[0] = fma( r[1] + r[2], k[0], r[0] )  (Gradient node, per axis)
        OP_DISTANCE,    // r[0] = distance( r[1..1+dim) ), p[0] function, p[1] 1 = take sqrt, r[2] Minkowski P register

        // ---- fused fractal over a leaf noise: loops octaves in registers
        // r[0] dst, r[1] positions. p[0] leaf op, p[1] octaves, p[2] 0 FBm / 1 Ridged
        // k: [leaf frequency, rangeMin, rangeScale, gain, weightedStrength, lacunarity]
        OP_FRACTAL_LEAF,

        OP_COUNT
    };

    static constexpr uint8_t kNoReg = 0xFF;

    struct Ins
    {
        uint8_t op;
        uint8_t dim;
        uint8_t p[4];
        uint8_t r[6];
        int32_t seed;   // added to the call's seed
        uint32_t k;     // constant pool index
    };
    static_assert( sizeof( Ins ) == 20, "instruction layout" );

    // Where the starting position of every point comes from.
    enum SourceKind : int32_t
    {
        SRC_GRID = 0,       // uniform grid: fma( index, step, offset )
        SRC_POSITIONS = 1,  // position arrays: offset + array value
        SRC_TILEABLE = 2,   // 2D index mapped onto a 4D torus
    };

    // Granlund-Montgomery division by a run-time invariant (Figure 4.1 of
    // "Division by Invariant Integers using Multiplication", 1994).
    struct FastDiv
    {
        uint32_t d, m, s1, s2;
    };
}

namespace fn2
{
    // Feature groups of the interpreter variants
    static constexpr int kFeatCell = 1;   // cellular ops
    static constexpr int kFeatWarp = 2;   // domain warps
    static constexpr int kFeatMisc = 4;   // white, checkerboard, sine wave, distance
}
