// Device-side shared noise math for gfx950: hashing, gradient selection,
// interpolants and distance functions of FastNoise2's generator path
// (include/FastNoise/Generators/Utils.inl of FastNoise2 v1.1.1).
//
// Everything is one point per lane in scalar VALU code. Bit-exactness rules:
// this file is compiled with -ffp-contract=off, every fused multiply-add is
// an explicit __builtin_fmaf placed exactly where the reference writes
// FS::FMulAdd / FNMulAdd / FMulSub, division and sqrt are the correctly
// rounded IEEE operations, integer arithmetic wraps (done in uint32_t).
#pragma once
// hipRTC (run-time compiled kernels) provides the runtime and the fixed-width types itself
#if !defined( __HIPCC_RTC__ )
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

// Host-and-device so the same source also compiles as plain host code.
#define FN2D __host__ __device__ __forceinline__

namespace fn2d
{
    // Primes and hash multiplier (Utils.inl:6-19)
    static constexpr int32_t kPX = (int32_t)0xF797C5C7u;
    static constexpr int32_t kPY = (int32_t)0x6C060C89u;
    static constexpr int32_t kPZ = (int32_t)0x465FD04Fu;
    static constexpr int32_t kPW = (int32_t)0xF7A62279u;
    static constexpr int32_t kHashMul = (int32_t)0xB7E0A5F5u;

    static constexpr double kRoot2 = 1.4142135623730950488016887242097;
    static constexpr double kRoot3 = 1.7320508075688772935274463415059;
    static constexpr double kRoot5 = 2.2360679774997896964091736687313;
    static constexpr float kRoot2f = (float)kRoot2;
    static constexpr float kRoot3f = (float)kRoot3;
    static constexpr float kSkew2f = (float)( 1.0 / ( kRoot3 + 1.0 ) );
    static constexpr float kSkew4f = (float)( 1.0 / ( kRoot5 + 1.0 ) );
    static constexpr float kValueBounds = 2147483648.f;

    // ---- wrapping integer arithmetic
    FN2D int32_t iadd( int32_t a, int32_t b ) { return (int32_t)( (uint32_t)a + (uint32_t)b ); }
    FN2D int32_t isub( int32_t a, int32_t b ) { return (int32_t)( (uint32_t)a - (uint32_t)b ); }
    FN2D int32_t imul( int32_t a, int32_t b ) { return (int32_t)( (uint32_t)a * (uint32_t)b ); }
    FN2D int32_t ishl( int32_t a, int s ) { return (int32_t)( (uint32_t)a << s ); }
    FN2D int32_t isrl( int32_t a, int s ) { return (int32_t)( (uint32_t)a >> s ); }

    // ---- float helpers
    FN2D float fmadd( float a, float b, float c ) { return __builtin_fmaf( a, b, c ); }
    FN2D float fnmadd( float a, float b, float c ) { return __builtin_fmaf( -a, b, c ); }
    FN2D float fmsub( float a, float b, float c ) { return __builtin_fmaf( a, b, -c ); }
    FN2D float ffloor( float a ) { return __builtin_floorf( a ); }
    // v_max_f32 / v_min_f32: equal to the reference's FS::Max / FS::Min for every
    // non-NaN input (only the sign of an exact zero result may differ)
    FN2D float fmaxf_( float a, float b ) { return __builtin_fmaxf( a, b ); }
    FN2D float fminf_( float a, float b ) { return __builtin_fminf( a, b ); }
    FN2D float fabs_( float a ) { return __builtin_fabsf( a ); }
    FN2D int32_t fbits( float a ) { return __builtin_bit_cast( int32_t, a ); }
    FN2D float ibits( int32_t a ) { return __builtin_bit_cast( float, a ); }
    FN2D float fxor( float a, int32_t signBits ) { return ibits( fbits( a ) ^ signBits ); }
    // FS::Convert<int32_t> of an integral float with the x86 cvtps2dq result
    // for NaN and out-of-range inputs: INT_MIN. v_cvt_i32_f32 saturates
    // instead (NaN -> 0, >= 2^31 -> INT_MAX, < -2^31 -> INT_MIN), so only NaN
    // and a >= 2^31 need the select. The hardware instruction is issued
    // directly because an out-of-range C++ float->int conversion is undefined.
    // GUARD = false: the caller has shown |a| < 2^31, so the select is left out
    template<bool GUARD = true>
    FN2D int32_t cvtIntegral( float a )
    {
#if defined( __HIP_DEVICE_COMPILE__ )
        int32_t v;
        asm volatile( "v_cvt_i32_f32 %0, %1" : "=v"( v ) : "v"( a ) );
        if constexpr( !GUARD ) return v;
#else
        int32_t v = a >= -2147483648.0f && a < 2147483648.0f ? (int32_t)a : (int32_t)0x80000000u;
#endif
        return a < 2147483648.0f ? v : (int32_t)0x80000000u;
    }
    // floor() results are integral, so the truncating convert is exact
    template<bool GUARD = true>
    FN2D int32_t floorToInt( float a ) { return cvtIntegral<GUARD>( a ); }
    // FS::Convert<int32_t> of a non-integer: round to nearest even
    FN2D int32_t roundToInt( float a ) { return cvtIntegral( __builtin_rintf( a ) ); }
    FN2D float invSqrt( float a ) { return 1.0f / __builtin_sqrtf( a ); }

    // Correctly rounded sqrt and reciprocal for positive, normal, finite
    // arguments with normal results, built from the hardware estimates
    // (v_sqrt_f32 / v_rcp_f32, both within 1 ulp) plus fma corrections,
    // without the zero / denormal / infinity scaling of the general IEEE
    // routines. Used where the argument range is known: the cellular jitter
    // vector's squared length, which lies in [0.5, 2.2e6] because every
    // component is a half-integer. fn2SelfTestMath checks them against the
    // IEEE results over whole binades.
#if defined( __HIP_DEVICE_COMPILE__ )
    FN2D float sqrtPosRN( float x )
    {
        const float s = __builtin_amdgcn_sqrtf( x );
        const float down = ibits( fbits( s ) - 1 );
        const float up = ibits( fbits( s ) + 1 );
        // x - down * s <= 0: sqrt(x) lies below the midpoint of (down, s)
        const float rDown = __builtin_fmaf( -down, s, x );
        // x - up * s > 0: sqrt(x) lies above the midpoint of (s, up)
        const float rUp = __builtin_fmaf( -up, s, x );
        float r = rDown <= 0.0f ? down : s;
        return rUp > 0.0f ? up : r;
    }

    FN2D float recipPosRN( float b )
    {
        const float y = __builtin_amdgcn_rcpf( b );
        const float e = __builtin_fmaf( -b, y, 1.0f );   // exact residual
        return __builtin_fmaf( e, y, y );
    }

    // sqrtPosRN for arguments on which the hardware estimate is never above
    // the correctly rounded root, so only the step up needs testing: the
    // squared lengths of 3D and 4D cellular jitter vectors (N / 4 for a sum N
    // of three or four odd squares below 2^22), checked exhaustively on the
    // device by the math tests
    FN2D float sqrtPosRNFromBelow( float x )
    {
        const float s = __builtin_amdgcn_sqrtf( x );
        const float up = ibits( fbits( s ) + 1 );
        return __builtin_fmaf( -up, s, x ) > 0.0f ? up : s;
    }
#else
    FN2D float sqrtPosRN( float x ) { return __builtin_sqrtf( x ); }
    FN2D float recipPosRN( float b ) { return 1.0f / b; }
    FN2D float sqrtPosRNFromBelow( float x ) { return __builtin_sqrtf( x ); }
#endif
    // Square root for bounds that carry margins far above 1 ulp (the
    // cellular walks' pruning limits): on the device the hardware estimate
    // (v_sqrt_f32, within 1 ulp), one instruction instead of the correctly
    // rounded routine's range scaling and correction steps
#if defined( __HIP_DEVICE_COMPILE__ )
    FN2D float sqrtBound( float x ) { return __builtin_amdgcn_sqrtf( x ); }
#else
    FN2D float sqrtBound( float x ) { return __builtin_sqrtf( x ); }
#endif
    FN2D uint32_t umin( uint32_t a, uint32_t b ) { return a < b ? a : b; }
    FN2D uint32_t umax( uint32_t a, uint32_t b ) { return a > b ? a : b; }

    // true when the predicate holds on every active lane of the wave (a host
    // build evaluates one point at a time: its own predicate)
#if defined( __HIP_DEVICE_COMPILE__ )
    __device__ __forceinline__ bool waveAll( bool p ) { return __all( p ); }
#else
    inline bool waveAll( bool p ) { return p; }
#endif

    // 1 / sqrt( x ) rounded twice, as FS::InvSqrt under strict semantics
    FN2D float invSqrtPos( float x ) { return recipPosRN( sqrtPosRN( x ) ); }
    // the same for the squared length of a D-dimensional cellular jitter vector
    template<int D> FN2D float invSqrtCell( float x ) { return recipPosRN( D == 2 ? sqrtPosRN( x ) : sqrtPosRNFromBelow( x ) ); }
    FN2D float recip( float a ) { return 1.0f / a; }

    // a ^ b ^ c as one v_bitop3_b32: the compiler splits three-input xors
    // into two v_xor_b32 on gfx950 (it has no v_xor3_b32)
    FN2D int32_t ixor3( int32_t a, int32_t b, int32_t c )
    {
#if defined( __HIP_DEVICE_COMPILE__ )
        return (int32_t)__builtin_amdgcn_bitop3_b32( (uint32_t)a, (uint32_t)b, (uint32_t)c, 0x96 );
#else
        return a ^ b ^ c;
#endif
    }
    // the seed and the primed coordinates xored together (Utils.inl:221-250)
    FN2D int32_t hashIn( int32_t seed, int32_t x, int32_t y ) { return ixor3( seed, x, y ); }
    FN2D int32_t hashIn( int32_t seed, int32_t x, int32_t y, int32_t z ) { return ixor3( seed, x, y ^ z ); }
    FN2D int32_t hashIn( int32_t seed, int32_t x, int32_t y, int32_t z, int32_t w ) { return ixor3( ixor3( seed, x, y ), z, w ); }

    // ---- hashing (Utils.inl:221-250)
    FN2D int32_t hashFinal( int32_t h )
    {
        h = imul( h, kHashMul );
        return ( h >> 15 ) ^ h;   // arithmetic shift
    }
    FN2D int32_t hash2( int32_t seed, int32_t x, int32_t y ) { return hashFinal( hashIn( seed, x, y ) ); }
    FN2D int32_t hash3( int32_t seed, int32_t x, int32_t y, int32_t z ) { return hashFinal( hashIn( seed, x, y, z ) ); }
    FN2D int32_t hash4( int32_t seed, int32_t x, int32_t y, int32_t z, int32_t w ) { return hashFinal( hashIn( seed, x, y, z, w ) ); }
    FN2D int32_t hashHB2( int32_t seed, int32_t x, int32_t y ) { return imul( hashIn( seed, x, y ), kHashMul ); }
    FN2D int32_t hashHB3( int32_t seed, int32_t x, int32_t y, int32_t z ) { return imul( hashIn( seed, x, y, z ), kHashMul ); }
    FN2D int32_t hashHB4( int32_t seed, int32_t x, int32_t y, int32_t z, int32_t w ) { return imul( hashIn( seed, x, y, z, w ), kHashMul ); }
    FN2D float valueCoordFinal( int32_t h ) { return (float)imul( h, imul( h, kHashMul ) ); }
    FN2D float valueCoord2( int32_t seed, int32_t x, int32_t y ) { return valueCoordFinal( hashIn( seed, x, y ) ); }
    FN2D float valueCoord3( int32_t seed, int32_t x, int32_t y, int32_t z ) { return valueCoordFinal( hashIn( seed, x, y, z ) ); }
    FN2D float valueCoord4( int32_t seed, int32_t x, int32_t y, int32_t z, int32_t w ) { return valueCoordFinal( hashIn( seed, x, y, z, w ) ); }

    // ---- interpolants (Utils.inl:252-265)
    FN2D float lerp( float a, float b, float t ) { return fmadd( t, b - a, a ); }
    FN2D float interpHermite( float t ) { return t * t * fnmadd( t, 2.0f, 3.0f ); }
    FN2D float interpQuintic( float t ) { return t * t * t * fmadd( t, fmadd( t, 6.0f, -15.0f ), 10.0f ); }

    // ---- gradient dot products, generic branches (Utils.inl:121-219)
    // 2D: 8 gradients ( 1+R2, 1 ) ... chosen by hash bits 0..2
    FN2D float gradDot2( int32_t h, float fx, float fy )
    {
        fx = fxor( fx, ishl( h, 31 ) );
        fy = fxor( fy, ishl( h >> 1, 31 ) );
        bool swap = ( h & 4 ) != 0;
        float u = swap ? fy : fx;
        float v = swap ? fx : fy;
        return fmadd( 1.0f + kRoot2f, u, v );
    }
    // 3D: 12 edge gradients + 4 repeats (hash & 13 selects the pair, bits 0/1 the signs)
    FN2D float gradDot3( int32_t h, float fx, float fy, float fz )
    {
        int32_t h13 = h & 13;
        float u = h13 > 7 ? fy : fx;
        float v = h13 == 12 ? fx : fz;
        v = h13 < 2 ? fy : v;
        return fxor( u, ishl( h, 31 ) ) + fxor( v, ishl( h >> 1, 31 ) );
    }
    // 4D: 32 gradients from bits 0..4
    FN2D float gradDot4( int32_t h, float fx, float fy, float fz, float fw )
    {
        int32_t p = h & ( 3 << 3 );
        float a = p > 0 ? fx : fy;
        float b = ( h & 16 ) ? fy : fz;
        float c = p > ( 2 << 3 ) ? fz : fw;
        return fxor( a, ishl( h, 31 ) ) + fxor( b, ishl( h >> 1, 31 ) ) + fxor( c, ishl( h >> 2, 31 ) );
    }

    // ScaleOutput (BasicGenerators.inl:27-30): scale = (1/(max-min)) * rangeScale folded on the host
    FN2D float scaleOutput( float value, float nativeMin, float scale, float rangeMin )
    {
        return fmadd( scale, value - nativeMin, rangeMin );
    }

    // CalcDistance (Utils.inl:281-335) on a D-vector; doSqrt selects the Euclidean square root
    template<int D>
    FN2D float calcDistance( int func, bool doSqrt, const float* d, float minkowskiP )
    {
        switch( func )
        {
        case 0:
        case 1:
        default:
        {
            float s = d[0] * d[0];
#pragma unroll
            for( int a = 1; a < D; a++ ) s = fmadd( d[a], d[a], s );
            return ( func == 0 && doSqrt ) ? __builtin_sqrtf( s ) : s;
        }
        case 2:
        {
            float tail = fabs_( d[D - 1] );
#pragma unroll
            for( int a = D - 2; a >= 1; a-- ) tail = fabs_( d[a] ) + tail;
            return fabs_( d[0] ) + tail;
        }
        case 3:
        {
            float both = fmadd( d[0], d[0], fabs_( d[0] ) );
#pragma unroll
            for( int a = 1; a < D; a++ ) both = both + fmadd( d[a], d[a], fabs_( d[a] ) );
            return both;
        }
        case 4:
        {
            float m = fabs_( d[0] );
#pragma unroll
            for( int a = 1; a < D; a++ ) m = fmaxf_( fabs_( d[a] ), m );
            return m;
        }
        case 5:
        {
            float tail = powf( fabs_( d[D - 1] ), minkowskiP );
#pragma unroll
            for( int a = D - 2; a >= 1; a-- ) tail = powf( fabs_( d[a] ), minkowskiP ) + tail;
            return powf( powf( fabs_( d[0] ), minkowskiP ) + tail, recip( minkowskiP ) );
        }
        }
    }
}
