// The node-program interpreter of the generator kernels: per-point start
// position (uniform grid, position arrays or the tileable 4D torus of
// Generator.inl:79-433) and the wave-uniform instruction loop over a
// 32-entry register file (uniform indices -> GPR indexing, no scratch).
// Host-and-device code: the kernels in fn2_kernels.hip instantiate it.
#pragma once
// hipRTC (run-time compiled kernels) provides the runtime and the fixed-width types itself
#if !defined( __HIPCC_RTC__ )
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif
#include "fn2_program.h"
#include "fn2_device_cellwarp.h"
#include "fn2_kernel_args.h"

// The interpreter takes the pruned cell walk, as the specialised kernels do
// (on MI355X the warped cellular tree's interpreter kernel went from 0.633 to
// 0.461 ms on a 256^3 grid and from 0.672 to 0.563 ms on 16.8 M positions,
// with the same bits); defined false, it takes the full walk
// Specialised kernels share the per-cell work of a coherent wave through
// an LDS table (cellularDistancesTable); defined 0, they take the pruned
// walk. The position-array variant with the shared far-cell walk leaves the
// table out: its points are scattered, and the two LDS regions together
// would cost resident blocks.
// Specialised gradient warps try the wave-uniform lattice cell (defined 0:
// the per-lane corner loop only; long programs' kernels do)
#ifndef FN2_UNIFORM_WARP
#define FN2_UNIFORM_WARP 1
#endif
// Specialised fused fractals try the wave-uniform lattice cell of their
// leaf (simplex2<G, true>); defined 0, they take the per-lane path only
#ifndef FN2_UNIFORM_LATTICE
#define FN2_UNIFORM_LATTICE 1
#endif
#ifndef FN2_CELL_TABLE
#define FN2_CELL_TABLE 1
#endif
#ifndef FN2_PRUNED_CELL_WALK
#define FN2_PRUNED_CELL_WALK true
#endif

// Unroll factor of a fused fractal's octave loop; 0 leaves it to the compiler
#ifndef FN2_OCTAVE_UNROLL
#define FN2_OCTAVE_UNROLL 0
#endif

namespace fn2i
{
    using namespace fn2d;
    using fn2::Ins;

    FN2D uint32_t fastDiv( const fn2::FastDiv& d, uint32_t n )
    {
        uint32_t t1 = (uint32_t)( ( (uint64_t)d.m * n ) >> 32 );
        return ( t1 + ( ( n - t1 ) >> d.s1 ) ) >> d.s2;
    }

    // A register file whose registers lie STRIDE floats apart (one column of
    // an LDS array shared by the block's threads); indexes like a float*.
    template<int STRIDE>
    struct StridedRegs
    {
        float* base;
        FN2D float& operator[]( int i ) const { return base[i * STRIDE]; }
    };

    // order-preserving int key of a float (for atomicMin / atomicMax)
    FN2D int32_t orderedKey( float f )
    {
        int32_t b = __builtin_bit_cast( int32_t, f );
        return b >= 0 ? b : b ^ 0x7fffffff;
    }

    // Compile-time op selection. The interpreter passes OP = -1 (any op, chosen
    // at run time from the instruction); a specialised kernel passes the
    // instruction's own op, so only that op's code is instantiated.
    template<int OP, int... Ops>
    constexpr bool opIn() { return OP < 0 || ( ( OP == Ops ) || ... ); }

    // Lattice noise on explicit coordinates; LEAF as OP above. G = false
    // leaves out the out-of-range select of the float -> int conversions, for
    // coordinates shown to lie below kUnguardedCoord in magnitude.
    template<int D, int LEAF = -1, bool G = true, bool U = false>
    FN2D float leafRaw( int op, int32_t seed, const float* x )
    {
        if constexpr( opIn<LEAF, fn2::OP_PERLIN>() )
            if( LEAF >= 0 || op == fn2::OP_PERLIN )
            {
                if constexpr( D == 2 ) return perlin2<G>( seed, x[0], x[1] );
                else if constexpr( D == 3 ) return perlin3<G>( seed, x[0], x[1], x[2] );
                else return perlin4<G>( seed, x[0], x[1], x[2], x[3] );
            }
        if constexpr( opIn<LEAF, fn2::OP_VALUE>() )
            if( LEAF >= 0 || op == fn2::OP_VALUE )
            {
                if constexpr( D == 2 ) return value2<G>( seed, x[0], x[1] );
                else if constexpr( D == 3 ) return value3<G>( seed, x[0], x[1], x[2] );
                else return value4<G>( seed, x[0], x[1], x[2], x[3] );
            }
        if constexpr( opIn<LEAF, fn2::OP_SIMPLEX>() )
            if( LEAF >= 0 || op == fn2::OP_SIMPLEX )
            {
                if constexpr( D == 2 ) return simplex2<G, U>( seed, x[0], x[1] );
                else if constexpr( D == 3 ) return simplex3<G>( seed, x[0], x[1], x[2] );
                else return simplex4<G>( seed, x[0], x[1], x[2], x[3] );
            }
        if constexpr( opIn<LEAF, fn2::OP_SUPERSIMPLEX>() )
        {
            if constexpr( D == 2 ) return superSimplex2<G>( seed, x[0], x[1] );
            else if constexpr( D == 3 ) return superSimplex3<G>( seed, x[0], x[1], x[2] );
            else return superSimplex4<G>( seed, x[0], x[1], x[2], x[3] );
        }
        return 0.0f;
    }

    template<int D, int OP = -1, class RF = const float*>
    FN2D void leafAt( const Ins& in, const float* K, int32_t seed, RF R, int p, float& out )
    {
        const float freq = K[0];
        float x[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) x[a] = R[p + a] * freq;
        out = scaleOutput( leafRaw<D, OP>( in.op, seed, x ), K[1], K[2], K[3] );
    }

    // FractalFBm / FractalRidged over a lattice leaf with constant gain and
    // weighted strength (Fractal.inl:9-67); the octave loop stays in registers.
    // A specialised kernel knows the octave count; FN2_OCTAVE_UNROLL (0: the
    // compiler's choice) sets how far the loop is unrolled, a per-tree tuning
    // knob of the serving engine.
    template<int D, int LEAF, bool G, bool BRICK>
    FN2D float fractalOctaves( const Ins& in, const float* K, int32_t seed, const float* start )
    {
        const int leaf = in.p[0], octaves = in.p[1];
        const bool ridged = in.p[2] != 0;
        const float freq = K[0], nMin = K[1], oScale = K[2], rMin = K[3];
        const float gain = K[4], ws = K[5], lac = K[6];
        float pos[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) pos[a] = start[a];
        auto octave = [&]( int32_t s ) {
            float x[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) x[a] = pos[a] * freq;
            // specialised kernels walking grid bricks try the wave-uniform
            // lattice cell (rows of 64 rarely share one, and then pay the
            // check for nothing)
            return scaleOutput( leafRaw<D, LEAF, G, BRICK && ( LEAF >= 0 ) && FN2_UNIFORM_LATTICE>( leaf, s, x ), nMin, oScale, rMin );
        };
        float amp = 1.0f;
        float noise = octave( seed );
        if( ridged ) noise = fabs_( noise );
        float sum = ridged ? ( noise * -2.0f + 1.0f ) * amp : noise * amp;
#if FN2_OCTAVE_UNROLL > 0
#pragma unroll FN2_OCTAVE_UNROLL
#endif
        for( int i = 1; i < octaves; i++ )
        {
            seed = iadd( seed, 1 );
            // lerp( 1, v, 0 ) = fma( 0, v - 1, 1 ) is exactly 1 for a finite
            // v and NaN otherwise. A lattice leaf's octave value is finite or
            // NaN (its outputs are bounded), and a NaN octave has already made
            // the sum NaN for good, so with weighted strength zero the factor
            // changes no result and is left out (a specialised kernel knows
            // ws, so only one side is compiled)
            if( ws != 0.0f )
                amp *= ridged ? lerp( 1.0f, 1.0f - noise, ws ) : lerp( 1.0f, ( noise + 1.0f ) * 0.5f, ws );
            amp *= gain;
#pragma unroll
            for( int a = 0; a < D; a++ ) pos[a] = pos[a] * lac;
            noise = octave( seed );
            if( ridged )
            {
                noise = fabs_( noise );
                sum += ( noise * -2.0f + 1.0f ) * amp;
            }
            else sum += noise * amp;
        }
        return sum;
    }

    // Lattice coordinates below this magnitude keep every lattice's floor
    // argument (at most 2.24 times the largest coordinate, for the 4D skew)
    // far below 2^31, where the float -> int conversion needs no select
    constexpr float kUnguardedCoord = 268435456.0f;   // 2^28

    template<int D, int LEAF = -1, bool BRICK = false, class RF = const float*>
    FN2D float fractalLeaf( const Ins& in, const float* K, int32_t seed, RF R, int p )
    {
        float pos[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) pos[a] = R[p + a];
        if constexpr( LEAF >= 0 )
        {
            // A specialised kernel folds this bound to a literal: octave k
            // reads pos * lac^k * freq (rounded at each step, far inside the
            // margin), so |pos| < limit keeps every octave's coordinates below
            // kUnguardedCoord. The comparisons fail for NaN and infinities,
            // whose waves take the guarded loop, as does any wave when the
            // bound overflows.
            const int octaves = in.p[1];
            float growth = fabs_( K[0] );
            for( int i = 1; i < octaves; i++ ) growth *= fmaxf_( 1.0f, fabs_( K[6] ) );
            const float limit = kUnguardedCoord / growth;
            bool small = true;
#pragma unroll
            for( int a = 0; a < D; a++ ) small = small && fabs_( pos[a] ) < limit;
            if( waveAll( small ) ) return fractalOctaves<D, LEAF, false, BRICK>( in, K, seed, pos );
        }
        return fractalOctaves<D, LEAF, true, BRICK>( in, K, seed, pos );
    }

    // SLOTS: distance slots CellularDistance keeps (4, or fewer when the
    // instruction's distance indices are compile-time constants)
    template<int D, int OP = -1, int SLOTS = 4, class RF = float*>
    FN2D void cellular( const Ins& in, const float* K, int32_t seed, RF R )
    {
        CellParams cp;
        cp.seed = seed;
        cp.jitter = jitterConst<D>() * R[in.r[2]];
        cp.sizeJitterOn = in.r[3] != fn2::kNoReg;
        cp.sizeJitter = cp.sizeJitterOn ? R[in.r[3]] * ( D == 2 ? -1.f / 0x3ff : D == 3 ? -1.f / 0xffff : -1.f / 0xfffff ) : 0.0f;
        cp.distFunc = in.p[0];
        cp.minkowskiP = K[3];
        const float freq = K[0];
        float pos[D];
        const int p = in.r[1];
#pragma unroll
        for( int a = 0; a < D; a++ ) pos[a] = R[p + a] * freq;
        if constexpr( opIn<OP, fn2::OP_CELL_VALUE>() )
            if( in.op == fn2::OP_CELL_VALUE )
            {
#if defined( __HIP_DEVICE_COMPILE__ )
                if constexpr( ( OP >= 0 ) && FN2_CELL_TABLE && !FN2_SHARED_CELL_WALK && ( D == 2 || D == 3 ) )
                {
                    float v;
                    if( cellularValueTable<D>( cp, pos, in.p[1], v ) )
                    {
                        R[in.r[0]] = scaleOutput( v, K[1], K[2], K[4] );
                        return;
                    }
                }
#endif
                R[in.r[0]] = scaleOutput( cellularValue<D>( cp, pos, in.p[1] ), K[1], K[2], K[4] );
                return;
            }
        if constexpr( opIn<OP, fn2::OP_CELL_DISTANCE>() )
            if( in.op == fn2::OP_CELL_DISTANCE )
            {
                float dist[4];
                cellularDistances<D, SLOTS, ( OP >= 0 ) || FN2_PRUNED_CELL_WALK, ( OP >= 0 ) && FN2_CELL_TABLE && !FN2_SHARED_CELL_WALK>( cp, pos, dist );
                R[in.r[0]] = scaleOutput( cellularReturn( dist, in.p[0], in.p[1], in.p[2], in.p[3], 0.0f ), K[1], K[2], K[4] );
                return;
            }
        if constexpr( opIn<OP, fn2::OP_CELL_LOOKUP>() )
        {
            float cell[D];
#if defined( __HIP_DEVICE_COMPILE__ )
            bool done = false;
            if constexpr( ( OP >= 0 ) && FN2_CELL_TABLE && !FN2_SHARED_CELL_WALK && ( D == 2 || D == 3 ) )
                done = cellularLookupTable<D>( cp, pos, cell );
            if( !done )
#endif
                cellularLookupPos<D>( cp, pos, cell );
            const float scale = K[5];
#pragma unroll
            for( int a = 0; a < D; a++ ) R[in.r[0] + a] = cell[a] * scale;
        }
    }

    template<int D, int OP = -1, class RF = float*>
    FN2D void warpOp( const Ins& in, const float* K, int32_t seed, RF R )
    {
        float sp[D], out[D], axis[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) { sp[a] = R[in.r[1] + a]; out[a] = R[in.r[2] + a]; axis[a] = K[a]; }
        const float amp = R[in.r[3]];
        float s = 0.0f;
        // specialised kernels of grid-like requests try the wave-uniform
        // corner loop; the position-array variant and the interpreter skip
        // the check, which scattered points never pass
        if constexpr( OP == fn2::OP_WARP_GRADIENT ) s = warpGradient<D, !FN2_SHARED_CELL_WALK && FN2_UNIFORM_WARP>( seed, amp, axis, sp, out );
        else if constexpr( OP == fn2::OP_WARP_SIMPLEX ) s = warpSimplex<D>( seed, in.p[0], amp, axis, sp, out );
        else if constexpr( OP == fn2::OP_WARP_SUPERSIMPLEX ) s = warpSuperSimplex<D>( seed, in.p[0], amp, axis, sp, out );
        else if( in.op == fn2::OP_WARP_GRADIENT ) s = warpGradient<D>( seed, amp, axis, sp, out );
        else if( in.op == fn2::OP_WARP_SIMPLEX ) s = warpSimplex<D>( seed, in.p[0], amp, axis, sp, out );
        else s = warpSuperSimplex<D>( seed, in.p[0], amp, axis, sp, out );
#pragma unroll
        for( int a = 0; a < D; a++ ) R[in.r[2] + a] = out[a];
        if( in.r[0] != fn2::kNoReg ) R[in.r[0]] = s;
    }

    template<int D, class RF>
    FN2D void distanceOp( const Ins& in, RF R )
    {
        float d[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) d[a] = R[in.r[1] + a];
        float mp = in.r[2] != fn2::kNoReg ? R[in.r[2]] : 1.5f;
        R[in.r[0]] = calcDistance<D>( in.p[0], in.p[1] != 0, d, mp );
    }

    // Runs the program on one point. R[0..dims) hold the starting position.
    // DM: bit d set = d-dimensional ops compiled in; FT: feature groups
    // compiled in (kFeatCell, kFeatWarp, kFeatMisc). The host picks the
    // smallest variant that covers the program, which keeps the VGPR budget
    // (and with it occupancy) of lattice-only programs low.
    // One instruction. `K` points at the instruction's constants, `seed` is
    // the call seed plus the instruction's seed delta.
    // OP / AUX: -1 in the interpreter; a specialised kernel passes the
    // instruction's op and, for OP_FRACTAL_LEAF, its leaf op, for
    // OP_CELL_DISTANCE the number of distance slots it reads, so that only
    // the code the instruction runs is instantiated.
    // BRICK: the point is one lane of a wave-sized grid brick (specialised
    // kernels walking a grid in bricks)
    template<int DM, int FT, int OP = -1, int AUX = -1, bool BRICK = false, class RF = float*>
    FN2D void execIns( const Ins& in, const float* K, const int32_t seed, RF R )
    {
        {
            const int d = in.dim;
            switch( in.op )
            {
            case fn2::OP_PERLIN: case fn2::OP_VALUE: case fn2::OP_SIMPLEX: case fn2::OP_SUPERSIMPLEX:
                if constexpr( opIn<OP, fn2::OP_PERLIN, fn2::OP_VALUE, fn2::OP_SIMPLEX, fn2::OP_SUPERSIMPLEX>() )
                {
                    float v = 0.0f;
                    if( ( DM & 4 ) && d == 2 ) leafAt<2, OP>( in, K, seed, R, in.r[1], v );
                    else if( ( DM & 8 ) && d == 3 ) leafAt<3, OP>( in, K, seed, R, in.r[1], v );
                    else if( DM & 16 ) leafAt<4, OP>( in, K, seed, R, in.r[1], v );
                    R[in.r[0]] = v;
                }
                break;
            case fn2::OP_FRACTAL_LEAF:
                if constexpr( opIn<OP, fn2::OP_FRACTAL_LEAF>() )
                {
                    float v = 0.0f;
                    if( ( DM & 4 ) && d == 2 ) v = fractalLeaf<2, AUX, BRICK>( in, K, seed, R, in.r[1] );
                    else if( ( DM & 8 ) && d == 3 ) v = fractalLeaf<3, AUX, BRICK>( in, K, seed, R, in.r[1] );
                    else if( DM & 16 ) v = fractalLeaf<4, AUX, BRICK>( in, K, seed, R, in.r[1] );
                    R[in.r[0]] = v;
                }
                break;
            case fn2::OP_CELL_VALUE: case fn2::OP_CELL_DISTANCE: case fn2::OP_CELL_LOOKUP:
                if constexpr( ( FT & fn2::kFeatCell ) != 0 && opIn<OP, fn2::OP_CELL_VALUE, fn2::OP_CELL_DISTANCE, fn2::OP_CELL_LOOKUP>() )
                {
                    constexpr int slots = OP == fn2::OP_CELL_DISTANCE && AUX > 0 ? AUX : 4;
                    if( ( DM & 4 ) && d == 2 ) cellular<2, OP, slots>( in, K, seed, R );
                    else if( ( DM & 8 ) && d == 3 ) cellular<3, OP, slots>( in, K, seed, R );
                    else if( DM & 16 ) cellular<4, OP, slots>( in, K, seed, R );
                }
                break;
            case fn2::OP_WARP_GRADIENT: case fn2::OP_WARP_SIMPLEX: case fn2::OP_WARP_SUPERSIMPLEX:
                if constexpr( ( FT & fn2::kFeatWarp ) != 0 && opIn<OP, fn2::OP_WARP_GRADIENT, fn2::OP_WARP_SIMPLEX, fn2::OP_WARP_SUPERSIMPLEX>() )
                {
                    if( ( DM & 4 ) && d == 2 ) warpOp<2, OP>( in, K, seed, R );
                    else if( ( DM & 8 ) && d == 3 ) warpOp<3, OP>( in, K, seed, R );
                    else if( DM & 16 ) warpOp<4, OP>( in, K, seed, R );
                }
                break;
            case fn2::OP_WHITE:
            if constexpr( ( FT & fn2::kFeatMisc ) != 0 && opIn<OP, fn2::OP_WHITE>() )
            {
                // White (BasicGenerators.inl:45-59): hash of the raw position bits
                const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
                int32_t h = seed;
                for( int ax = 0; ax < d; ax++ )
                {
                    int32_t b = fbits( R[in.r[1] + ax] );
                    h ^= imul( b ^ ( b >> 16 ), prime[ax] );
                }
                R[in.r[0]] = scaleOutput( valueCoordFinal( h ), K[1], K[2], K[3] );
                break;
            }
            break;
            case fn2::OP_CHECKERBOARD:
            if constexpr( ( FT & fn2::kFeatMisc ) != 0 && opIn<OP, fn2::OP_CHECKERBOARD>() )
            {
                int32_t v = roundToInt( R[in.r[1]] * K[0] );
                for( int ax = 1; ax < d; ax++ ) v ^= roundToInt( R[in.r[1] + ax] * K[0] );
                R[in.r[0]] = scaleOutput( ibits( ishl( v & 1, 30 ) ), K[1], K[2], K[3] );
                break;
            }
            break;
            case fn2::OP_SINEWAVE:
            if constexpr( ( FT & fn2::kFeatMisc ) != 0 && opIn<OP, fn2::OP_SINEWAVE>() )
            {
                float r = sinf( R[in.r[1] + d - 1] * K[0] );
                for( int ax = d - 2; ax >= 0; ax-- ) r = sinf( R[in.r[1] + ax] * K[0] ) * r;
                R[in.r[0]] = scaleOutput( r, K[1], K[2], K[3] );
                break;
            }
            break;
            case fn2::OP_CONST: R[in.r[0]] = K[0]; break;
            case fn2::OP_MOV: R[in.r[0]] = R[in.r[1]]; break;
            case fn2::OP_ADD: R[in.r[0]] = R[in.r[1]] + R[in.r[2]]; break;
            case fn2::OP_SUB: R[in.r[0]] = R[in.r[1]] - R[in.r[2]]; break;
            case fn2::OP_MUL: R[in.r[0]] = R[in.r[1]] * R[in.r[2]]; break;
            case fn2::OP_DIV: R[in.r[0]] = R[in.r[1]] / R[in.r[2]]; break;
            case fn2::OP_MIN: R[in.r[0]] = fminf_( R[in.r[1]], R[in.r[2]] ); break;
            case fn2::OP_MAX: R[in.r[0]] = fmaxf_( R[in.r[1]], R[in.r[2]] ); break;
            case fn2::OP_FMOD: R[in.r[0]] = __builtin_fmodf( R[in.r[1]], R[in.r[2]] ); break;
            case fn2::OP_ABS: R[in.r[0]] = fabs_( R[in.r[1]] ); break;
            case fn2::OP_SCALE:
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax] * K[0];
                break;
            case fn2::OP_SCALE_AXES:
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax] * K[ax];
                break;
            case fn2::OP_MUL_VEC:
            {
                float s = R[in.r[2]];
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax] * s;
                break;
            }
            case fn2::OP_ADD_VEC:
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax] + R[in.r[2] + ax];
                break;
            case fn2::OP_SUB_VEC:
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax] - R[in.r[2] + ax];
                break;
            case fn2::OP_MOV_VEC:
                for( int ax = 0; ax < d; ax++ ) R[in.r[0] + ax] = R[in.r[1] + ax];
                break;
            case fn2::OP_FBM_AMP:
            {
                float amp = R[in.r[0]] * lerp( 1.0f, ( R[in.r[1]] + 1.0f ) * 0.5f, R[in.r[2]] );
                R[in.r[0]] = amp * R[in.r[3]];
                break;
            }
            case fn2::OP_RIDGED_AMP:
            {
                float amp = R[in.r[0]] * lerp( 1.0f, 1.0f - R[in.r[1]], R[in.r[2]] );
                R[in.r[0]] = amp * R[in.r[3]];
                break;
            }
            case fn2::OP_WARP_AMP:
            {
                float strength = __builtin_sqrtf( R[in.r[1]] );
                float amp = R[in.r[0]] * lerp( 1.0f, 1.0f - strength, R[in.r[2]] );
                R[in.r[0]] = amp * R[in.r[3]];
                break;
            }
            case fn2::OP_FBM_SUM: R[in.r[0]] = R[in.r[0]] + R[in.r[1]] * R[in.r[2]]; break;
            case fn2::OP_RIDGED_TERM:
            {
                float t = ( R[in.r[1]] * -2.0f + 1.0f ) * R[in.r[2]];
                R[in.r[0]] = in.p[0] ? R[in.r[0]] + t : t;
                break;
            }
            case fn2::OP_PINGPONG:
            {
                // PingPong (Modifiers.inl:189-200)
                float t = R[in.r[1]] * R[in.r[2]];
                t -= ffloor( t * 0.5f ) * 2.0f;
                R[in.r[0]] = t < 1.0f ? t : 2.0f - t;
                break;
            }
            case fn2::OP_MINSMOOTH:
            {
                // MinSmooth / MaxSmooth (Blends.inl:124-162)
                float x = R[in.r[1]], y = R[in.r[2]];
                if( in.p[0] ) { x = -x; y = -y; }
                float s = fmaxf_( 1.17549435e-38f, fabs_( R[in.r[3]] ) );
                float h = fmaxf_( s - fabs_( x - y ), 0.0f );
                h *= recip( s );
                float r = fnmadd( 1.0f / 6.0f, h * h * h * s, fminf_( x, y ) );
                R[in.r[0]] = in.p[0] ? -r : r;
                break;
            }
            case fn2::OP_FADE:
            {
                // Fade (Blends.inl:164-199)
                float fmin = R[in.r[4]], range = R[in.r[5]] - fmin;
                float f = ( R[in.r[3]] - fmin ) / range;
                f = fmaxf_( 0.0f, fminf_( 1.0f, f ) );
                if( in.p[0] == 1 ) f = interpHermite( f );
                else if( in.p[0] == 2 ) f = interpQuintic( f );
                f = range == 0.0f ? 0.5f : f;
                R[in.r[0]] = lerp( R[in.r[1]], R[in.r[2]], f );
                break;
            }
            case fn2::OP_REMAP:
            {
                // Remap (Modifiers.inl:81-99): true division
                float fromMin = R[in.r[2]], toMin = R[in.r[4]], toMax = R[in.r[5]];
                float r = toMin + ( ( R[in.r[1]] - fromMin ) / ( R[in.r[3]] - fromMin ) * ( toMax - toMin ) );
                if( in.p[0] ) { r = fmaxf_( r, toMin ); r = fminf_( r, toMax ); }
                R[in.r[0]] = r;
                break;
            }
            case fn2::OP_TERRACE:
            {
                // Terrace (Modifiers.inl:135-181)
                float s = R[in.r[1]] * K[0];
                float rounded = __builtin_rintf( s );
                bool dynamic = in.r[2] != fn2::kNoReg;
                float smooth = dynamic ? R[in.r[2]] : 0.0f;
                bool on = dynamic ? smooth != 0.0f : in.p[1] != 0;
                if( on )
                {
                    float diff = rounded - s;
                    bool neg = diff < 0.0f;
                    diff = 0.5f - fabs_( diff );
                    diff *= dynamic ? 1.0f + recip( smooth ) : K[2];
                    diff = fminf_( diff, 0.5f );
                    diff = neg ? 0.5f - diff : diff - 0.5f;
                    rounded += diff;
                }
                R[in.r[0]] = rounded * K[1];
                break;
            }
            case fn2::OP_POWF:
                R[in.r[0]] = powf( fmaxf_( fabs_( R[in.r[1]] ), 1.17549435e-38f ), R[in.r[2]] );
                break;
            case fn2::OP_POWI:
            {
                float v = R[in.r[1]], r = v * v;
                for( int i = 2; i < (int)in.p[0]; i++ ) r *= v;
                R[in.r[0]] = r;
                break;
            }
            case fn2::OP_SSQRT:
            {
                float v = R[in.r[1]];
                R[in.r[0]] = ibits( fbits( __builtin_sqrtf( fabs_( v ) ) ) | ( fbits( v ) & (int32_t)0x80000000u ) );
                break;
            }
            case fn2::OP_RGBA8:
            {
                float s = fminf_( R[in.r[1]], K[1] );
                s = fmaxf_( s, K[0] );
                s -= K[0];
                s *= K[2];
                int32_t b = roundToInt( s );
                R[in.r[0]] = ibits( (int32_t)( 255u << 24 ) | b | ishl( b, 8 ) | ishl( b, 16 ) );
                break;
            }
            case fn2::OP_ROTATE2:
            {
                // DomainRotate 2D yaw-only branch (Modifiers.inl:38-43)
                float x = R[in.r[1]], y = R[in.r[1] + 1];
                R[in.r[0]] = fnmadd( y, K[1], x * K[0] );
                R[in.r[0] + 1] = fmadd( x, K[1], y * K[0] );
                break;
            }
            case fn2::OP_ROTATE3:
            {
                // DomainRotate 3D matrix (Modifiers.inl:48-54)
                float x = R[in.r[1]], y = R[in.r[1] + 1], z = R[in.r[1] + 2];
                R[in.r[0]] = fmadd( x, K[0], fmadd( y, K[1], z * K[2] ) );
                R[in.r[0] + 1] = fmadd( x, K[3], fmadd( y, K[4], z * K[5] ) );
                R[in.r[0] + 2] = fmadd( x, K[6], fmadd( y, K[7], z * K[8] ) );
                break;
            }
            case fn2::OP_ROTATE_PLANE:
            {
                // DomainRotatePlane (Modifiers.inl:354-388), plain multiplies and adds
                float x = R[in.r[1]], y = R[in.r[1] + 1], z = R[in.r[1] + 2];
                float nx = x, ny = y, nz = z;
                if( in.p[0] == 0 )
                {
                    float xy = x + y, s2 = xy * -0.211324865405187f;
                    nz = z * 0.577350269189626f;
                    nx = x + s2 - nz;
                    ny = y + s2 - nz;
                    nz = nz + xy * 0.577350269189626f;
                }
                else if( in.p[0] == 1 )
                {
                    float xz = x + z, s2 = xz * -0.211324865405187f;
                    ny = y * 0.577350269189626f;
                    nx = x + s2 - ny;
                    nz = z + s2 - ny;
                    ny = ny + xz * 0.577350269189626f;
                }
                R[in.r[0]] = nx; R[in.r[0] + 1] = ny; R[in.r[0] + 2] = nz;
                break;
            }
            case fn2::OP_GRADIENT_AXIS:
                R[in.r[0]] = fmadd( R[in.r[1]] + R[in.r[2]], K[0], R[in.r[0]] );
                break;
            case fn2::OP_DISTANCE:
                if constexpr( ( FT & fn2::kFeatMisc ) != 0 && opIn<OP, fn2::OP_DISTANCE>() )
                {
                    if( ( DM & 4 ) && d == 2 ) distanceOp<2>( in, R );
                    else if( ( DM & 8 ) && d == 3 ) distanceOp<3>( in, R );
                    else if( DM & 16 ) distanceOp<4>( in, R );
                }
                break;
            default:
                break;
            }
        }
    }

    template<int DM, int FT, class RF>
    FN2D float runProgram( const fn2::GenArgs& a, RF R )
    {
        for( int pc = 0; pc < a.nIns; pc++ )
        {
            const Ins in = a.ins[pc];
            execIns<DM, FT>( in, a.konst + in.k, iadd( a.seed, in.seed ), R );
        }
        return R[a.resultReg];
    }

    // Starting position of point `idx` into R[0..dims): grid, position arrays
    // or the tileable torus (Generator.inl:79-433).
    // BASE: the job may be a part of a split request (a.base != 0); the
    // single-call kernels pass false (their base is always 0) and save the add
    template<bool BASE = true, class RF>
    FN2D void startPosition( const fn2::GenArgs& a, uint32_t idx, RF R )
    {
        if( a.kind == fn2::SRC_POSITIONS )
        {
            for( int ax = 0; ax < a.dims; ax++ ) R[ax] = a.offset[ax] + a.pos[ax][idx];
        }
        else if( a.kind == fn2::SRC_TILEABLE )
        {
            if( BASE ) idx += a.base;
            uint32_t yi = fastDiv( a.div[0], idx );
            uint32_t xi = idx - yi * a.div[0].d;
            // per-column and per-row torus coordinates, tabulated by the host
            R[0] = a.pos[0][xi];
            R[1] = a.pos[1][yi];
            R[2] = a.pos[2][xi];
            R[3] = a.pos[3][yi];
        }
        else
        {
            uint32_t rem = BASE ? idx + a.base : idx;
            for( int ax = 0; ax < a.dims - 1; ax++ )
            {
                uint32_t q = fastDiv( a.div[ax], rem );
                R[ax] = fmadd( (float)(int32_t)( rem - q * a.div[ax].d ), a.step[ax], a.offset[ax] );
                rem = q;
            }
            R[a.dims - 1] = fmadd( (float)(int32_t)rem, a.step[a.dims - 1], a.offset[a.dims - 1] );
        }
    }

    // The same for a program compiled for D-dimensional positions (specialised
    // kernels): every source writes the same D registers, from values computed
    // per source, so the register file never needs a run-time index. A
    // tileable call always runs a 4D program.
    template<int D, bool BASE = true, class RF>
    FN2D void startPositionD( const fn2::GenArgs& a, uint32_t idx, RF R )
    {
        float p[D];
        if( a.kind == fn2::SRC_POSITIONS )
        {
#pragma unroll
            for( int ax = 0; ax < D; ax++ ) p[ax] = a.offset[ax] + a.pos[ax][idx];
        }
        else if( D == 4 && a.kind == fn2::SRC_TILEABLE )
        {
            if( BASE ) idx += a.base;
            uint32_t yi = fastDiv( a.div[0], idx );
            uint32_t xi = idx - yi * a.div[0].d;
            p[0] = a.pos[0][xi];
            p[1] = a.pos[1][yi];
            p[D > 2 ? 2 : 0] = a.pos[2][xi];
            p[D > 3 ? 3 : 0] = a.pos[3][yi];
        }
        else
        {
            uint32_t rem = BASE ? idx + a.base : idx;
#pragma unroll
            for( int ax = 0; ax < D - 1; ax++ )
            {
                uint32_t q = fastDiv( a.div[ax], rem );
                p[ax] = fmadd( (float)(int32_t)( rem - q * a.div[ax].d ), a.step[ax], a.offset[ax] );
                rem = q;
            }
            p[D - 1] = fmadd( (float)(int32_t)rem, a.step[D - 1], a.offset[D - 1] );
        }
#pragma unroll
        for( int ax = 0; ax < D; ax++ ) R[ax] = p[ax];
    }
}
