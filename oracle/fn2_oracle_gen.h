// Test infrastructure only: per-node dispatch of the CPU oracle
// (the reference's virtual Gen calls, Generator.inl:19-26, 48-67).
#pragma once
#include <cmath>
#include "fn2_oracle_eval.h"

namespace FN2O_NS
{
    template<int N, bool FUSED, bool AVX2>
    template<int D>
    FV<N> Eval<N, FUSED, AVX2>::gen( const Node* np, IV<N> seed, Pos<D, N> p )
    {
        const Node& n = *np;
        switch( n.id )
        {
        case kConstant: return S( n.f[0] );
        case kWhite: return white<D>( n, seed, p );
        case kCheckerboard: return checkerboard<D>( n, p );
        case kSineWave: return sineWave<D>( n, p );
        case kGradient: return gradient<D>( n, seed, p );
        case kDistanceToPoint: return distanceToPoint<D>( n, seed, p );
        case kSimplex: case kSuperSimplex: case kPerlin: case kValue: return leafNoise<D>( n, seed, p );
        case kCellularValue: return cellularValue<D>( n, seed, p );
        case kCellularDistance: return cellularDistance<D>( n, seed, p );
        case kCellularLookup: return cellularLookup<D>( n, seed, p );
        case kFractalFBm: return fbm<D>( n, seed, p );
        case kPingPong: return pingPong<D>( n, seed, p );
        case kFractalRidged: return ridged<D>( n, seed, p );
        case kDomainWarpSimplex: case kDomainWarpSuperSimplex: case kDomainWarpGradient: return domainWarp<D>( n, seed, p );
        case kDomainWarpFractalProgressive: return domainWarpFractal<D, false>( n, seed, p );
        case kDomainWarpFractalIndependent: return domainWarpFractal<D, true>( n, seed, p );
        // Blends.inl:3-122
        case kAdd: return gen<D>( n.src[0], seed, p ) + hybrid<D>( n.hyb[0], seed, p );
        case kSubtract: return hybrid<D>( n.hyb[0], seed, p ) - hybrid<D>( n.hyb[1], seed, p );
        case kMultiply: return gen<D>( n.src[0], seed, p ) * hybrid<D>( n.hyb[0], seed, p );
        case kDivide: return hybrid<D>( n.hyb[0], seed, p ) / hybrid<D>( n.hyb[1], seed, p );
        case kModulus: return vfmod( hybrid<D>( n.hyb[0], seed, p ), hybrid<D>( n.hyb[1], seed, p ) );
        case kMin: return vmin( gen<D>( n.src[0], seed, p ), hybrid<D>( n.hyb[0], seed, p ) );
        case kMax: return vmax( gen<D>( n.src[0], seed, p ), hybrid<D>( n.hyb[0], seed, p ) );
        case kMinSmooth: return smoothMinMax<D, false>( n, seed, p );
        case kMaxSmooth: return smoothMinMax<D, true>( n, seed, p );
        case kPowFloat:
        {
            F v = vmax( vabs( hybrid<D>( n.hyb[0], seed, p ) ), S( FLT_MIN ) );
            return vpow( v, hybrid<D>( n.hyb[1], seed, p ) );
        }
        case kPowInt:
        {
            F v = gen<D>( n.src[0], seed, p );
            F r = v * v;
            for( int i = 2; i < n.i[0]; i++ ) r *= v;
            return r;
        }
        case kFade: return fade<D>( n, seed, p );
        // Modifiers.inl
        case kAbs: return vabs( gen<D>( n.src[0], seed, p ) );
        case kSignedSquareRoot:
        {
            F v = gen<D>( n.src[0], seed, p );
            return fastLengthSqrt( vabs( v ) ) | ibits( fbits( v ) & (int32_t)0x80000000u );
        }
        case kDomainScale:
            for( int a = 0; a < D; a++ ) p.c[a] = p.c[a] * S( n.f[0] );
            return gen<D>( n.src[0], seed, p );
        case kDomainOffset:
        {
            Pos<D, N> o = p;
            for( int a = 0; a < D; a++ ) o.c[a] = o.c[a] + hybrid<D>( n.hyb[a], seed, p );
            return gen<D>( n.src[0], seed, o );
        }
        case kDomainAxisScale:
            for( int a = 0; a < D; a++ ) p.c[a] = p.c[a] * S( n.f[a] );
            return gen<D>( n.src[0], seed, p );
        case kSeedOffset: return gen<D>( n.src[0], seed + n.i[0], p );
        case kConvertRGBA8: return convertRGBA8<D>( n, seed, p );
        case kGeneratorCache: return gen<D>( n.src[0], seed, p );
        case kRemap: return remap<D>( n, seed, p );
        case kTerrace: return terrace<D>( n, seed, p );
        case kDomainRotate:
        {
            if constexpr( D == 2 )
            {
                if( n.f[3] == 0.0f && n.f[5] == 0.0f )
                {
                    Pos<2, N> r;
                    r.c[0] = fnmadd( p.c[1], S( n.f[1] ), p.c[0] * S( n.f[0] ) );
                    r.c[1] = fmadd( p.c[0], S( n.f[1] ), p.c[1] * S( n.f[0] ) );
                    return gen<2>( n.src[0], seed, r );
                }
                return gen<3>( n.src[0], seed, rotate3( n, p.c[0], p.c[1], S( 0 ) ) );
            }
            else if constexpr( D == 3 )
                return gen<3>( n.src[0], seed, rotate3( n, p.c[0], p.c[1], p.c[2] ) );
            else
                return gen<4>( n.src[0], seed, p );
        }
        case kAddDimension:
        {
            if constexpr( D == 4 ) return gen<4>( n.src[0], seed, p );
            else
            {
                Pos<D + 1, N> q;
                for( int a = 0; a < D; a++ ) q.c[a] = p.c[a];
                q.c[D] = hybrid<D>( n.hyb[0], seed, p );
                return gen<D + 1>( n.src[0], seed, q );
            }
        }
        case kRemoveDimension:
        {
            if constexpr( D == 2 ) return gen<2>( n.src[0], seed, p );
            else
            {
                int rm = n.i[0];
                if( rm < 0 || rm >= D ) return gen<D>( n.src[0], seed, p );
                Pos<D - 1, N> q;
                for( int a = 0, b = 0; a < D; a++ ) if( a != rm ) q.c[b++] = p.c[a];
                return gen<D - 1>( n.src[0], seed, q );
            }
        }
        case kDomainRotatePlane:
        {
            if constexpr( D == 2 )
            {
                Pos<3, N> q;
                q.c[0] = p.c[0]; q.c[1] = p.c[1]; q.c[2] = S( 0 );
                rotatePlane( 0, q.c[0], q.c[1], q.c[2] );
                return gen<3>( n.src[0], seed, q );
            }
            else
            {
                rotatePlane( n.i[0], p.c[0], p.c[1], p.c[2] );
                return gen<D>( n.src[0], seed, p );
            }
        }
        default: return S( 0 );
        }
    }
}
