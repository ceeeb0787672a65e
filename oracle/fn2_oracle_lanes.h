// Test infrastructure only: the lane types of the CPU oracle.
//
// The oracle restates FastNoise2's generator path on the CPU. It evaluates
// N points at once through fixed-width "lane" vectors so that the same source
// serves two purposes: with N = 1 it is the scalar checker used by the
// parity tests, and with N = 16 (compiled with AVX2/AVX-512 flags) every
// lane operation is a vector instruction (GCC vector types, x86 intrinsics
// for fma / floor / rint / sqrt / float->int), which gives the host SIMD
// baseline timed by bench.py. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load the oracle; nothing in the product
// links it.
//
// Floating point rules (these define the oracle): no contraction anywhere
// (built with -ffp-contract=off), fmaf() only where the reference writes
// FS::FMulAdd / FS::FNMulAdd / FS::FMulSub (or plain mul+add when the
// "unfused" mode is selected), exact division and square root, floorf,
// float->int conversion rounding to nearest-even with the x86 cvtps2dq
// out-of-range result INT_MIN, min/max with the x86 minps/maxps operand
// order (a < b ? a : b).
#pragma once
#ifndef FN2O_NS
#define FN2O_NS fn2o
#endif
#include <cstdint>
#include <cstring>
#include <cmath>
#if defined( __AVX2__ ) || defined( __AVX512F__ )
#include <immintrin.h>
#endif

#ifdef FN2O_COUNTING
// Op-counting build (one lane): every vector operation the reference's SIMD
// code performs is tallied once per lane; FLOPs count fma as 2.
namespace fn2o_count
{
    extern thread_local uint64_t flop, valu;
}
#define FN2O_CF( n ) ( fn2o_count::flop += ( n ), fn2o_count::valu += 1 )
#define FN2O_CV() ( fn2o_count::valu += 1 )
#else
#define FN2O_CF( n ) ( (void)0 )
#define FN2O_CV() ( (void)0 )
#endif

#ifdef FN2O_TRACING
// Index-replay build (one lane): every discrete decision the evaluation takes
// (float -> int conversions, float comparisons, floor / round results, the
// position bits White hashes) is folded into a per-point fingerprint, so that
// two variants of the oracle agree on a point's fingerprint exactly when they
// took the same lattice cells, simplex vertices, cells and selections there.
namespace fn2o_trace
{
    extern thread_local uint64_t fp;
    extern thread_local uint64_t* out;   // fingerprint per point of the running call, or null
    inline void fold( uint32_t v ) { fp = ( fp ^ v ) * 0x100000001B3ull; }
}
#define FN2O_TRACE_I( iv ) do { for( int k_ = 0; k_ < N; k_++ ) fn2o_trace::fold( (uint32_t)( iv ).v[k_] ); } while( 0 )
#else
#define FN2O_TRACE_I( iv ) ( (void)0 )
#endif

#ifndef FN2O_INLINE
#define FN2O_INLINE inline __attribute__( ( always_inline ) )
#endif

namespace FN2O_NS
{
#define FN2O_LOOP for( int k = 0; k < N; k++ )

    // Lanes are GCC vector types, so a 16-lane value is one AVX-512 register
    // (two AVX2 registers) and lane-wise operators compile to vector
    // instructions; element access is still .v[k].
    template<int N> struct FV
    {
        typedef float T __attribute__( ( vector_size( N * 4 ) ) );
        T v;
    };
    template<int N> struct IV
    {
        typedef int32_t T __attribute__( ( vector_size( N * 4 ) ) );
        typedef uint32_t U __attribute__( ( vector_size( N * 4 ) ) );
        T v;
    };
    template<int N> FN2O_INLINE FV<N> mkf( typename FV<N>::T v ) { FV<N> r; r.v = v; return r; }
    template<int N> FN2O_INLINE IV<N> mki( typename IV<N>::T v ) { IV<N> r; r.v = v; return r; }
    template<int N> FN2O_INLINE typename IV<N>::U uv( IV<N> a ) { return (typename IV<N>::U)a.v; }

    static inline uint32_t bits_of( float f ) { uint32_t u; std::memcpy( &u, &f, 4 ); return u; }
    static inline float float_of( uint32_t u ) { float f; std::memcpy( &f, &u, 4 ); return f; }

    template<int N> FN2O_INLINE FV<N> fsplat( float a ) { FV<N> r; r.v = a - (typename FV<N>::T){}; return r; }
    template<int N> FN2O_INLINE IV<N> isplat( int32_t a ) { IV<N> r; r.v = a - (typename IV<N>::T){}; return r; }

    // ---- float arithmetic
    template<int N> FN2O_INLINE FV<N> operator+( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); return mkf<N>( a.v + b.v ); }
    template<int N> FN2O_INLINE FV<N> operator-( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); return mkf<N>( a.v - b.v ); }
    template<int N> FN2O_INLINE FV<N> operator*( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); return mkf<N>( a.v * b.v ); }
    template<int N> FN2O_INLINE FV<N> operator/( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); return mkf<N>( a.v / b.v ); }
    template<int N> FN2O_INLINE FV<N> operator-( FV<N> a ) { FN2O_CV(); return mkf<N>( -a.v ); }
    template<int N> FN2O_INLINE FV<N> operator+( FV<N> a, float b ) { return a + fsplat<N>( b ); }
    template<int N> FN2O_INLINE FV<N> operator-( FV<N> a, float b ) { return a - fsplat<N>( b ); }
    template<int N> FN2O_INLINE FV<N> operator*( FV<N> a, float b ) { return a * fsplat<N>( b ); }
    template<int N> FN2O_INLINE FV<N> operator-( float a, FV<N> b ) { return fsplat<N>( a ) - b; }
    template<int N> FN2O_INLINE FV<N> operator+( float a, FV<N> b ) { return fsplat<N>( a ) + b; }
    template<int N> FN2O_INLINE FV<N> operator*( float a, FV<N> b ) { return fsplat<N>( a ) * b; }
    template<int N> FN2O_INLINE FV<N>& operator+=( FV<N>& a, FV<N> b ) { a = a + b; return a; }
    template<int N> FN2O_INLINE FV<N>& operator-=( FV<N>& a, FV<N> b ) { a = a - b; return a; }
    template<int N> FN2O_INLINE FV<N>& operator*=( FV<N>& a, FV<N> b ) { a = a * b; return a; }

    // fused multiply-add, floor, round-to-nearest-even, sqrt: x86 vector
    // instructions in the 16-lane builds, libm per lane otherwise
    template<int N> FN2O_INLINE FV<N> vfma( FV<N> a, FV<N> b, FV<N> c )
    {
        FN2O_CF( 2 );
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) return mkf<N>( (typename FV<N>::T)_mm512_fmadd_ps( (__m512)a.v, (__m512)b.v, (__m512)c.v ) );
#endif
#if defined( __FMA__ ) && defined( __AVX2__ )
        if constexpr( N == 16 )
        {
            FV<N> r;
            __m256 lo = _mm256_fmadd_ps( _mm256_loadu_ps( (const float*)&a.v ), _mm256_loadu_ps( (const float*)&b.v ), _mm256_loadu_ps( (const float*)&c.v ) );
            __m256 hi = _mm256_fmadd_ps( _mm256_loadu_ps( (const float*)&a.v + 8 ), _mm256_loadu_ps( (const float*)&b.v + 8 ), _mm256_loadu_ps( (const float*)&c.v + 8 ) );
            _mm256_storeu_ps( (float*)&r.v, lo );
            _mm256_storeu_ps( (float*)&r.v + 8, hi );
            return r;
        }
#endif
        FV<N> r;
        FN2O_LOOP r.v[k] = fmaf( a.v[k], b.v[k], c.v[k] );
        return r;
    }
    template<int N> FN2O_INLINE FV<N> vfloor( FV<N> a )
    {
        FN2O_CV();
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) return mkf<N>( (typename FV<N>::T)_mm512_roundscale_ps( (__m512)a.v, _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC ) );
#endif
        FV<N> r;
        FN2O_LOOP r.v[k] = floorf( a.v[k] );
#ifdef FN2O_TRACING
        FN2O_LOOP fn2o_trace::fold( bits_of( r.v[k] ) );
#endif
        return r;
    }
    template<int N> FN2O_INLINE FV<N> vround( FV<N> a )
    {
        FN2O_CV();
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) return mkf<N>( (typename FV<N>::T)_mm512_roundscale_ps( (__m512)a.v, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC ) );
#endif
        FV<N> r;
        FN2O_LOOP r.v[k] = rintf( a.v[k] );
#ifdef FN2O_TRACING
        FN2O_LOOP fn2o_trace::fold( bits_of( r.v[k] ) );
#endif
        return r;
    }
    template<int N> FN2O_INLINE FV<N> vabs( FV<N> a ) { FN2O_CV(); return mkf<N>( (typename FV<N>::T)( (typename IV<N>::T)a.v & 0x7fffffff ) ); }
    template<int N> FN2O_INLINE FV<N> vsqrt( FV<N> a )
    {
        FN2O_CF( 1 );
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) return mkf<N>( (typename FV<N>::T)_mm512_sqrt_ps( (__m512)a.v ) );
#endif
        FV<N> r;
        FN2O_LOOP r.v[k] = sqrtf( a.v[k] );
        return r;
    }
    // x86 minps / maxps operand order: the second operand when unordered
    template<int N> FN2O_INLINE FV<N> vmin( FV<N> a, FV<N> b ) { FN2O_CV(); return mkf<N>( a.v < b.v ? a.v : b.v ); }
    template<int N> FN2O_INLINE FV<N> vmax( FV<N> a, FV<N> b ) { FN2O_CV(); return mkf<N>( a.v > b.v ? a.v : b.v ); }
    // FS::InvSqrt and FS::Reciprocal under strict semantics: correctly rounded 1/sqrt(x) and 1/x.
    template<int N> FN2O_INLINE FV<N> vinvsqrt( FV<N> a )
    {
        FN2O_CF( 2 );
        FV<N> s;
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) s = mkf<N>( (typename FV<N>::T)_mm512_sqrt_ps( (__m512)a.v ) );
        else
#endif
        {
            FN2O_LOOP s.v[k] = sqrtf( a.v[k] );
        }
        return mkf<N>( 1.0f / s.v );
    }
    template<int N> FN2O_INLINE FV<N> vrecip( FV<N> a ) { FN2O_CF( 1 ); return mkf<N>( 1.0f / a.v ); }

    // ---- bit casts and conversions
    template<int N> FN2O_INLINE IV<N> fbits( FV<N> a ) { return mki<N>( (typename IV<N>::T)a.v ); }
    template<int N> FN2O_INLINE FV<N> ibits( IV<N> a ) { return mkf<N>( (typename FV<N>::T)a.v ); }
    template<int N> FN2O_INLINE FV<N> cvtf( IV<N> a ) { FN2O_CV(); return mkf<N>( __builtin_convertvector( a.v, typename FV<N>::T ) ); }
    static inline int32_t cvti1( float f )
    {
        // x86 cvtps2dq: round to nearest even, INT_MIN when out of range or NaN
        if( !( f >= -2147483648.0f && f < 2147483648.0f ) ) return INT32_MIN;
        return (int32_t)rintf( f );
    }
    template<int N> FN2O_INLINE IV<N> cvti( FV<N> a )
    {
        FN2O_CV();
#if defined( __AVX512F__ )
        if constexpr( N == 16 ) return mki<N>( (typename IV<N>::T)_mm512_cvtps_epi32( (__m512)a.v ) );
#endif
        IV<N> r;
        FN2O_LOOP r.v[k] = cvti1( a.v[k] );
        FN2O_TRACE_I( r );
        return r;
    }
    // float xor float (sign manipulation), float | float
    template<int N> FN2O_INLINE FV<N> operator^( FV<N> a, FV<N> b ) { FN2O_CV(); return ibits( mki<N>( fbits( a ).v ^ fbits( b ).v ) ); }
    template<int N> FN2O_INLINE FV<N> operator|( FV<N> a, FV<N> b ) { FN2O_CV(); return ibits( mki<N>( fbits( a ).v | fbits( b ).v ) ); }

    // ---- integer arithmetic (two's complement wrapping: done on unsigned lanes)
    template<int N> FN2O_INLINE IV<N> operator+( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( (typename IV<N>::T)( uv( a ) + uv( b ) ) ); }
    template<int N> FN2O_INLINE IV<N> operator-( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( (typename IV<N>::T)( uv( a ) - uv( b ) ) ); }
    template<int N> FN2O_INLINE IV<N> operator*( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( (typename IV<N>::T)( uv( a ) * uv( b ) ) ); }
    template<int N> FN2O_INLINE IV<N> operator^( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v ^ b.v ); }
    template<int N> FN2O_INLINE IV<N> operator&( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v & b.v ); }
    template<int N> FN2O_INLINE IV<N> operator|( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v | b.v ); }
    template<int N> FN2O_INLINE IV<N> operator~( IV<N> a ) { FN2O_CV(); return mki<N>( ~a.v ); }
    template<int N> FN2O_INLINE IV<N> operator+( IV<N> a, int32_t b ) { return a + isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N> operator-( IV<N> a, int32_t b ) { return a - isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N> operator*( IV<N> a, int32_t b ) { return a * isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N> operator&( IV<N> a, int32_t b ) { return a & isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N> operator^( IV<N> a, int32_t b ) { return a ^ isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N> operator|( IV<N> a, int32_t b ) { return a | isplat<N>( b ); }
    template<int N> FN2O_INLINE IV<N>& operator+=( IV<N>& a, IV<N> b ) { a = a + b; return a; }
    template<int N> FN2O_INLINE IV<N>& operator*=( IV<N>& a, IV<N> b ) { a = a * b; return a; }
    template<int N> FN2O_INLINE IV<N>& operator^=( IV<N>& a, IV<N> b ) { a = a ^ b; return a; }
    // arithmetic shift right, shift left, logical shift right
    template<int N> FN2O_INLINE IV<N> sra( IV<N> a, int s ) { FN2O_CV(); return mki<N>( a.v >> s ); }
    template<int N> FN2O_INLINE IV<N> shl( IV<N> a, int s ) { FN2O_CV(); return mki<N>( (typename IV<N>::T)( uv( a ) << s ) ); }
    template<int N> FN2O_INLINE IV<N> srl( IV<N> a, int s ) { FN2O_CV(); return mki<N>( (typename IV<N>::T)( uv( a ) >> s ) ); }

    // ---- masks: int32 lanes holding 0 or -1
    template<int N> FN2O_INLINE IV<N> mlt( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r = mki<N>( (typename IV<N>::T)( a.v < b.v ) ); FN2O_TRACE_I( r ); return r; }
    template<int N> FN2O_INLINE IV<N> mgt( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r = mki<N>( (typename IV<N>::T)( a.v > b.v ) ); FN2O_TRACE_I( r ); return r; }
    template<int N> FN2O_INLINE IV<N> mge( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r = mki<N>( (typename IV<N>::T)( a.v >= b.v ) ); FN2O_TRACE_I( r ); return r; }
    template<int N> FN2O_INLINE IV<N> meq( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r = mki<N>( (typename IV<N>::T)( a.v == b.v ) ); FN2O_TRACE_I( r ); return r; }
    template<int N> FN2O_INLINE IV<N> mne( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r = mki<N>( (typename IV<N>::T)( a.v != b.v ) ); FN2O_TRACE_I( r ); return r; }
    template<int N> FN2O_INLINE IV<N> mlt( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v < b.v ); }
    template<int N> FN2O_INLINE IV<N> mgt( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v > b.v ); }
    template<int N> FN2O_INLINE IV<N> mge( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v >= b.v ); }
    template<int N> FN2O_INLINE IV<N> meq( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v == b.v ); }
    template<int N> FN2O_INLINE IV<N> mne( IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( a.v != b.v ); }

    // FS::Select( m, a, b ) = m ? a : b
    template<int N> FN2O_INLINE FV<N> sel( IV<N> m, FV<N> a, FV<N> b ) { FN2O_CV(); return mkf<N>( m.v != 0 ? a.v : b.v ); }
    template<int N> FN2O_INLINE IV<N> sel( IV<N> m, IV<N> a, IV<N> b ) { FN2O_CV(); return mki<N>( m.v != 0 ? a.v : b.v ); }
    // FS::SelectHighBit( v, a, b ) = sign bit of v ? a : b
    template<int N> FN2O_INLINE FV<N> selhb( IV<N> m, FV<N> a, FV<N> b ) { FN2O_CV(); return mkf<N>( m.v < 0 ? a.v : b.v ); }
    // Masked helpers (FS::Masked*, FS::InvMasked*)
    template<int N> FN2O_INLINE FV<N> masked( IV<N> m, FV<N> a ) { return sel( m, a, fsplat<N>( 0.0f ) ); }
    template<int N> FN2O_INLINE IV<N> masked( IV<N> m, IV<N> a ) { return sel( m, a, isplat<N>( 0 ) ); }
    template<int N> FN2O_INLINE IV<N> invmasked( IV<N> m, IV<N> a ) { return sel( m, isplat<N>( 0 ), a ); }
    template<int N> FN2O_INLINE FV<N> madd( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a + b, a ); }
    template<int N> FN2O_INLINE IV<N> madd( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a + b, a ); }
    template<int N> FN2O_INLINE FV<N> msub( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a - b, a ); }
    template<int N> FN2O_INLINE FV<N> mmul( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a * b, a ); }
    template<int N> FN2O_INLINE FV<N> imadd( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a, a + b ); }
    template<int N> FN2O_INLINE IV<N> imadd( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a, a + b ); }
    template<int N> FN2O_INLINE FV<N> imsub( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a, a - b ); }
    template<int N> FN2O_INLINE IV<N> imsub( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a, a - b ); }
    template<int N> FN2O_INLINE FV<N> minc( IV<N> m, FV<N> a ) { return sel( m, a + 1.0f, a ); }
    template<int N> FN2O_INLINE IV<N> minc( IV<N> m, IV<N> a ) { return sel( m, a + 1, a ); }
    template<int N> FN2O_INLINE FV<N> mdec( IV<N> m, FV<N> a ) { return sel( m, a - 1.0f, a ); }
    template<int N> FN2O_INLINE IV<N> mdec( IV<N> m, IV<N> a ) { return sel( m, a - 1, a ); }
    template<int N> FN2O_INLINE bool any( IV<N> m ) { bool r = false; FN2O_LOOP r |= m.v[k] != 0; return r; }
}
