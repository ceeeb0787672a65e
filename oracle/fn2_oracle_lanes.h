// Test infrastructure only: the lane types of the CPU oracle.
//
// The oracle restates FastNoise2's generator path on the CPU. It evaluates
// N points at once through fixed-width "lane" arrays so that the same source
// serves two purposes: with N = 1 it is the scalar checker used by the
// parity tests, and with N = 16 (compiled with AVX2/AVX-512 flags) gcc turns
// every per-lane loop into vector instructions, which gives the host SIMD
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

namespace FN2O_NS
{
#define FN2O_LOOP for( int k = 0; k < N; k++ )

    template<int N> struct FV { float v[N]; };
    template<int N> struct IV { int32_t v[N]; };

    static inline uint32_t bits_of( float f ) { uint32_t u; std::memcpy( &u, &f, 4 ); return u; }
    static inline float float_of( uint32_t u ) { float f; std::memcpy( &f, &u, 4 ); return f; }

    template<int N> inline FV<N> fsplat( float a ) { FV<N> r; FN2O_LOOP r.v[k] = a; return r; }
    template<int N> inline IV<N> isplat( int32_t a ) { IV<N> r; FN2O_LOOP r.v[k] = a; return r; }

    // ---- float arithmetic
    template<int N> inline FV<N> operator+( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] + b.v[k]; return r; }
    template<int N> inline FV<N> operator-( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] - b.v[k]; return r; }
    template<int N> inline FV<N> operator*( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] * b.v[k]; return r; }
    template<int N> inline FV<N> operator/( FV<N> a, FV<N> b ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] / b.v[k]; return r; }
    template<int N> inline FV<N> operator-( FV<N> a ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = -a.v[k]; return r; }
    template<int N> inline FV<N> operator+( FV<N> a, float b ) { return a + fsplat<N>( b ); }
    template<int N> inline FV<N> operator-( FV<N> a, float b ) { return a - fsplat<N>( b ); }
    template<int N> inline FV<N> operator*( FV<N> a, float b ) { return a * fsplat<N>( b ); }
    template<int N> inline FV<N> operator-( float a, FV<N> b ) { return fsplat<N>( a ) - b; }
    template<int N> inline FV<N> operator+( float a, FV<N> b ) { return fsplat<N>( a ) + b; }
    template<int N> inline FV<N> operator*( float a, FV<N> b ) { return fsplat<N>( a ) * b; }
    template<int N> inline FV<N>& operator+=( FV<N>& a, FV<N> b ) { a = a + b; return a; }
    template<int N> inline FV<N>& operator-=( FV<N>& a, FV<N> b ) { a = a - b; return a; }
    template<int N> inline FV<N>& operator*=( FV<N>& a, FV<N> b ) { a = a * b; return a; }

    template<int N> inline FV<N> vfma( FV<N> a, FV<N> b, FV<N> c ) { FN2O_CF( 2 ); FV<N> r; FN2O_LOOP r.v[k] = fmaf( a.v[k], b.v[k], c.v[k] ); return r; }
    template<int N> inline FV<N> vfloor( FV<N> a ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = floorf( a.v[k] ); return r; }
    template<int N> inline FV<N> vround( FV<N> a ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = rintf( a.v[k] ); return r; }
    template<int N> inline FV<N> vabs( FV<N> a ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = fabsf( a.v[k] ); return r; }
    template<int N> inline FV<N> vsqrt( FV<N> a ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = sqrtf( a.v[k] ); return r; }
    template<int N> inline FV<N> vmin( FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] < b.v[k] ? a.v[k] : b.v[k]; return r; }
    template<int N> inline FV<N> vmax( FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = a.v[k] > b.v[k] ? a.v[k] : b.v[k]; return r; }
    // FS::InvSqrt and FS::Reciprocal under strict semantics: correctly rounded 1/sqrt(x) and 1/x.
    template<int N> inline FV<N> vinvsqrt( FV<N> a ) { FN2O_CF( 2 ); FV<N> r; FN2O_LOOP r.v[k] = 1.0f / sqrtf( a.v[k] ); return r; }
    template<int N> inline FV<N> vrecip( FV<N> a ) { FN2O_CF( 1 ); FV<N> r; FN2O_LOOP r.v[k] = 1.0f / a.v[k]; return r; }

    // ---- bit casts and conversions
    template<int N> inline IV<N> fbits( FV<N> a ) { IV<N> r; FN2O_LOOP r.v[k] = (int32_t)bits_of( a.v[k] ); return r; }
    template<int N> inline FV<N> ibits( IV<N> a ) { FV<N> r; FN2O_LOOP r.v[k] = float_of( (uint32_t)a.v[k] ); return r; }
    template<int N> inline FV<N> cvtf( IV<N> a ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = (float)a.v[k]; return r; }
    static inline int32_t cvti1( float f )
    {
        // x86 cvtps2dq: round to nearest even, INT_MIN when out of range or NaN
        if( !( f >= -2147483648.0f && f < 2147483648.0f ) ) return INT32_MIN;
        return (int32_t)rintf( f );
    }
    template<int N> inline IV<N> cvti( FV<N> a ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = cvti1( a.v[k] ); return r; }
    // float xor float (sign manipulation), float | float
    template<int N> inline FV<N> operator^( FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = float_of( bits_of( a.v[k] ) ^ bits_of( b.v[k] ) ); return r; }
    template<int N> inline FV<N> operator|( FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = float_of( bits_of( a.v[k] ) | bits_of( b.v[k] ) ); return r; }

    // ---- integer arithmetic (two's complement wrapping)
    template<int N> inline IV<N> operator+( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = (int32_t)( (uint32_t)a.v[k] + (uint32_t)b.v[k] ); return r; }
    template<int N> inline IV<N> operator-( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = (int32_t)( (uint32_t)a.v[k] - (uint32_t)b.v[k] ); return r; }
    template<int N> inline IV<N> operator*( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = (int32_t)( (uint32_t)a.v[k] * (uint32_t)b.v[k] ); return r; }
    template<int N> inline IV<N> operator^( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] ^ b.v[k]; return r; }
    template<int N> inline IV<N> operator&( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] & b.v[k]; return r; }
    template<int N> inline IV<N> operator|( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] | b.v[k]; return r; }
    template<int N> inline IV<N> operator~( IV<N> a ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = ~a.v[k]; return r; }
    template<int N> inline IV<N> operator+( IV<N> a, int32_t b ) { return a + isplat<N>( b ); }
    template<int N> inline IV<N> operator-( IV<N> a, int32_t b ) { return a - isplat<N>( b ); }
    template<int N> inline IV<N> operator*( IV<N> a, int32_t b ) { return a * isplat<N>( b ); }
    template<int N> inline IV<N> operator&( IV<N> a, int32_t b ) { return a & isplat<N>( b ); }
    template<int N> inline IV<N> operator^( IV<N> a, int32_t b ) { return a ^ isplat<N>( b ); }
    template<int N> inline IV<N> operator|( IV<N> a, int32_t b ) { return a | isplat<N>( b ); }
    template<int N> inline IV<N>& operator+=( IV<N>& a, IV<N> b ) { a = a + b; return a; }
    template<int N> inline IV<N>& operator*=( IV<N>& a, IV<N> b ) { a = a * b; return a; }
    template<int N> inline IV<N>& operator^=( IV<N>& a, IV<N> b ) { a = a ^ b; return a; }
    // arithmetic shift right, shift left, logical shift right
    template<int N> inline IV<N> sra( IV<N> a, int s ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] >> s; return r; }
    template<int N> inline IV<N> shl( IV<N> a, int s ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = (int32_t)( (uint32_t)a.v[k] << s ); return r; }
    template<int N> inline IV<N> srl( IV<N> a, int s ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = (int32_t)( (uint32_t)a.v[k] >> s ); return r; }

    // ---- masks: int32 lanes holding 0 or -1
    template<int N> inline IV<N> mlt( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] < b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mgt( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] > b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mge( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] >= b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> meq( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] == b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mne( FV<N> a, FV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] != b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mlt( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] < b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mgt( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] > b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mge( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] >= b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> meq( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] == b.v[k] ? -1 : 0; return r; }
    template<int N> inline IV<N> mne( IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = a.v[k] != b.v[k] ? -1 : 0; return r; }

    // FS::Select( m, a, b ) = m ? a : b
    template<int N> inline FV<N> sel( IV<N> m, FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = m.v[k] ? a.v[k] : b.v[k]; return r; }
    template<int N> inline IV<N> sel( IV<N> m, IV<N> a, IV<N> b ) { FN2O_CV(); IV<N> r; FN2O_LOOP r.v[k] = m.v[k] ? a.v[k] : b.v[k]; return r; }
    // FS::SelectHighBit( v, a, b ) = sign bit of v ? a : b
    template<int N> inline FV<N> selhb( IV<N> m, FV<N> a, FV<N> b ) { FN2O_CV(); FV<N> r; FN2O_LOOP r.v[k] = m.v[k] < 0 ? a.v[k] : b.v[k]; return r; }
    // Masked helpers (FS::Masked*, FS::InvMasked*)
    template<int N> inline FV<N> masked( IV<N> m, FV<N> a ) { return sel( m, a, fsplat<N>( 0.0f ) ); }
    template<int N> inline IV<N> masked( IV<N> m, IV<N> a ) { return sel( m, a, isplat<N>( 0 ) ); }
    template<int N> inline IV<N> invmasked( IV<N> m, IV<N> a ) { return sel( m, isplat<N>( 0 ), a ); }
    template<int N> inline FV<N> madd( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a + b, a ); }
    template<int N> inline IV<N> madd( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a + b, a ); }
    template<int N> inline FV<N> msub( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a - b, a ); }
    template<int N> inline FV<N> mmul( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a * b, a ); }
    template<int N> inline FV<N> imadd( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a, a + b ); }
    template<int N> inline IV<N> imadd( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a, a + b ); }
    template<int N> inline FV<N> imsub( IV<N> m, FV<N> a, FV<N> b ) { return sel( m, a, a - b ); }
    template<int N> inline IV<N> imsub( IV<N> m, IV<N> a, IV<N> b ) { return sel( m, a, a - b ); }
    template<int N> inline FV<N> minc( IV<N> m, FV<N> a ) { return sel( m, a + 1.0f, a ); }
    template<int N> inline IV<N> minc( IV<N> m, IV<N> a ) { return sel( m, a + 1, a ); }
    template<int N> inline FV<N> mdec( IV<N> m, FV<N> a ) { return sel( m, a - 1.0f, a ); }
    template<int N> inline IV<N> mdec( IV<N> m, IV<N> a ) { return sel( m, a - 1, a ); }
    template<int N> inline bool any( IV<N> m ) { bool r = false; FN2O_LOOP r |= m.v[k] != 0; return r; }
}
