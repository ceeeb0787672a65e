// Test infrastructure only: the node evaluator of the CPU oracle.
//
// Eval<N, FUSED, AVX2> evaluates a node tree on N points at a time, following
// the reference node by node (citations per function; paths relative to
// include/FastNoise/Generators/). Seeds are carried per lane like the
// reference's int32v even though every lane holds the same value.
#pragma once
#include "fn2_oracle_nodes.h"

namespace FN2O_NS
{
    // Primes and hash multiplier (Utils.inl:6-19)
    static constexpr int32_t kPrimeX = (int32_t)0xF797C5C7u;
    static constexpr int32_t kPrimeY = (int32_t)0x6C060C89u;
    static constexpr int32_t kPrimeZ = (int32_t)0x465FD04Fu;
    static constexpr int32_t kPrimeW = (int32_t)0xF7A62279u;
    static constexpr int32_t kHashA = (int32_t)0xB7E0A5F5u;
    static constexpr int32_t kPrimes[4] = { kPrimeX, kPrimeY, kPrimeZ, kPrimeW };

    static constexpr double kRoot2 = 1.4142135623730950488016887242097;
    static constexpr double kRoot3 = 1.7320508075688772935274463415059;
    static constexpr double kRoot5 = 2.2360679774997896964091736687313;
    static constexpr float kValueBounds = 2147483648.f;
    static constexpr float kRoot2f = (float)kRoot2;
    static constexpr float kSkew2f = (float)( 1.0 / ( kRoot3 + 1.0 ) );
    static constexpr float kSkew4f = (float)( 1.0 / ( kRoot5 + 1.0 ) );

    // Cellular jitter constants (Cellular.inl:11-14)
    static constexpr float kJitter2D = 0.437016f;
    static constexpr float kJitter3D = 0.396144f;
    static constexpr float kJitter4D = 0.366025f;

    template<int D, int N> struct Pos { FV<N> c[D]; };

    template<int N, bool FUSED, bool AVX2>
    struct Eval
    {
        using F = FV<N>;
        using I = IV<N>;

        static F S( float a ) { return fsplat<N>( a ); }
        static I SI( int32_t a ) { return isplat<N>( a ); }

        // FS::FMulAdd / FNMulAdd / FMulSub sites
        static F fmadd( F a, F b, F c ) { if constexpr( FUSED ) return vfma( a, b, c ); else return a * b + c; }
        static F fnmadd( F a, F b, F c ) { if constexpr( FUSED ) return vfma( -a, b, c ); else return c - a * b; }
        static F fmsub( F a, F b, F c ) { if constexpr( FUSED ) return vfma( a, b, -c ); else return a * b - c; }

        // ---------------- shared math (Utils.inl) ----------------
        // HashPrimes (Utils.inl:221-230): arithmetic >> 15
        template<typename... P> static I hashPrimes( I seed, P... p )
        {
            I h = seed;
            h ^= ( p ^ ... );
            h = h * kHashA;
            return sra( h, 15 ) ^ h;
        }
        // HashPrimesHB (Utils.inl:232-240)
        template<typename... P> static I hashPrimesHB( I seed, P... p )
        {
            I h = seed;
            h ^= ( p ^ ... );
            return h * kHashA;
        }
        // GetValueCoord (Utils.inl:242-250): h * (h * A)
        template<typename... P> static F valueCoord( I seed, P... p )
        {
            I h = seed;
            h ^= ( p ^ ... );
            h = h * ( h * kHashA );
            return cvtf( h );
        }
        static F lerp( F a, F b, F t ) { return fmadd( t, b - a, a ); }          // Utils.inl:252-255
        static F interpHermite( F t ) { return t * t * fnmadd( t, S( 2 ), S( 3 ) ); } // Utils.inl:257-260
        static F interpQuintic( F t )                                            // Utils.inl:262-265
        {
            return t * t * t * fmadd( t, fmadd( t, S( 6 ), S( -15 ) ), S( 10 ) );
        }
        static F fastLengthSqrt( F d ) { return vsqrt( d ); }                   // Utils.inl:267-279, strict branch

        // GetGradientDotPerlin 2D (Utils.inl:161-191)
        static F gradDot2( I hash, F fX, F fY )
        {
            if constexpr( AVX2 )
            {
                // _mm256_permutevar8x32_ps uses the low three bits of hash
                const float a = 1 + kRoot2f;
                static const float tx[8] = { a, -a, a, -a, 1, -1, 1, -1 };
                static const float ty[8] = { 1, 1, -1, -1, a, a, -a, -a };
                F gX, gY;
                for( int k = 0; k < N; k++ ) { gX.v[k] = tx[hash.v[k] & 7]; gY.v[k] = ty[hash.v[k] & 7]; }
                return fmadd( gX, fX, fY * gY );
            }
            else
            {
                fX = fX ^ ibits( shl( hash, 31 ) );
                fY = fY ^ ibits( shl( sra( hash, 1 ), 31 ) );
                F u = selhb( shl( hash, 29 ), fY, fX );
                F v = selhb( shl( hash, 29 ), fX, fY );
                return fmadd( S( 1.0f + kRoot2f ), u, v );
            }
        }
        // GetGradientDotCommon 3D (Utils.inl:121-159), generic branch
        static F gradDot3( I hash, F fX, F fY, F fZ )
        {
            I hasha13 = hash & 13;
            I gt7 = mgt( hasha13, SI( 7 ) );
            F u = sel( gt7, fY, fX );
            F v = sel( meq( hasha13, SI( 12 ) ), fX, fZ );
            v = sel( mlt( hasha13, SI( 2 ) ), fY, v );
            F h1 = ibits( shl( hash, 31 ) );
            F h2 = ibits( shl( sra( hash, 1 ), 31 ) );
            return ( u ^ h1 ) + ( v ^ h2 );
        }
        // GetGradientDotPerlin 4D (Utils.inl:193-219), generic branch
        static F gradDot4( I hash, F fX, F fY, F fZ, F fW )
        {
            I p = hash & ( 3 << 3 );
            F a = sel( mgt( p, SI( 0 ) ), fX, fY );
            F b = selhb( shl( hash, 27 ), fY, fZ );
            F c = sel( mgt( p, SI( 2 << 3 ) ), fZ, fW );
            F aS = ibits( shl( hash, 31 ) );
            F bS = ibits( shl( sra( hash, 1 ), 31 ) );
            F cS = ibits( shl( sra( hash, 2 ), 31 ) );
            return ( a ^ aS ) + ( b ^ bS ) + ( c ^ cS );
        }

        // ScaleOutput (BasicGenerators.inl:27-30)
        static F scaleOutput( const Node& n, F value, float nativeMin, float nativeMax )
        {
            return fmadd( S( 1.0f / ( nativeMax - nativeMin ) ) * S( n.rangeScale ), value - S( nativeMin ), S( n.rangeMin ) );
        }

        // CalcDistance (Utils.inl:281-335)
        template<bool DO_SQRT, int D>
        static F calcDistance( int distFunc, const Hybrid& minkowskiP, I seed, const Pos<D, N>& d )
        {
            switch( distFunc )
            {
            default:
            case 0: // Euclidean
                if constexpr( DO_SQRT )
                {
                    F s = d.c[0] * d.c[0];
                    for( int a = 1; a < D; a++ ) s = fmadd( d.c[a], d.c[a], s );
                    return fastLengthSqrt( s );
                }
                [[fallthrough]];
            case 1: // EuclideanSquared
            {
                F s = d.c[0] * d.c[0];
                for( int a = 1; a < D; a++ ) s = fmadd( d.c[a], d.c[a], s );
                return s;
            }
            case 2: // Manhattan: |x| + ( |y| + |z| ... ) (right fold)
            {
                F tail = vabs( d.c[D - 1] );
                for( int a = D - 2; a >= 1; a-- ) tail = vabs( d.c[a] ) + tail;
                return vabs( d.c[0] ) + tail;
            }
            case 3: // Hybrid
            {
                F both = fmadd( d.c[0], d.c[0], vabs( d.c[0] ) );
                for( int a = 1; a < D; a++ ) both = both + fmadd( d.c[a], d.c[a], vabs( d.c[a] ) );
                return both;
            }
            case 4: // MaxAxis
            {
                F m = vabs( d.c[0] );
                for( int a = 1; a < D; a++ ) m = vmax( vabs( d.c[a] ), m );
                return m;
            }
            case 5: // Minkowski, FS::Pow taken as powf
            {
                F p = hybrid<D>( minkowskiP, seed, d );
                F tail = vpow( vabs( d.c[D - 1] ), p );
                for( int a = D - 2; a >= 1; a-- ) tail = vpow( vabs( d.c[a] ), p ) + tail;
                return vpow( vpow( vabs( d.c[0] ), p ) + tail, vrecip( p ) );
            }
            }
        }
        static F vpow( F a, F b ) { F r; for( int k = 0; k < N; k++ ) r.v[k] = powf( a.v[k], b.v[k] ); return r; }

        // ---------------- graph evaluation (Generator.inl:48-67) ----------------
        template<int D> static F hybrid( const Hybrid& h, I seed, const Pos<D, N>& p )
        {
            if( h.node ) return gen<D>( h.node, seed, p );
            return S( h.c );
        }

        template<int D> static F gen( const Node* n, I seed, Pos<D, N> p );

        // ---------------- lattice noises ----------------
        // Perlin 2D/3D/4D (Perlin.inl:7-125)
        static F perlin( const Node& n, I seed, Pos<2, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY;
            F xf0 = xs = x - xs, yf0 = ys = y - ys;
            F xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f;
            xs = interpQuintic( xs );
            ys = interpQuintic( ys );
            F value = lerp(
                lerp( gradDot2( hashPrimes( seed, x0, y0 ), xf0, yf0 ), gradDot2( hashPrimes( seed, x1, y0 ), xf1, yf0 ), xs ),
                lerp( gradDot2( hashPrimes( seed, x0, y1 ), xf0, yf1 ), gradDot2( hashPrimes( seed, x1, y1 ), xf1, yf1 ), xs ), ys );
            constexpr float kB = 1.726796627044677734375f;
            return scaleOutput( n, value, -kB, kB );
        }
        static F perlin( const Node& n, I seed, Pos<3, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y ), zs = vfloor( z );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY, z0 = cvti( zs ) * kPrimeZ;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY, z1 = z0 + kPrimeZ;
            F xf0 = xs = x - xs, yf0 = ys = y - ys, zf0 = zs = z - zs;
            F xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f, zf1 = zf0 - 1.0f;
            xs = interpQuintic( xs );
            ys = interpQuintic( ys );
            zs = interpQuintic( zs );
            F value = lerp( lerp(
                lerp( gradDot3( hashPrimes( seed, x0, y0, z0 ), xf0, yf0, zf0 ), gradDot3( hashPrimes( seed, x1, y0, z0 ), xf1, yf0, zf0 ), xs ),
                lerp( gradDot3( hashPrimes( seed, x0, y1, z0 ), xf0, yf1, zf0 ), gradDot3( hashPrimes( seed, x1, y1, z0 ), xf1, yf1, zf0 ), xs ), ys ),
                lerp(
                lerp( gradDot3( hashPrimes( seed, x0, y0, z1 ), xf0, yf0, zf1 ), gradDot3( hashPrimes( seed, x1, y0, z1 ), xf1, yf0, zf1 ), xs ),
                lerp( gradDot3( hashPrimes( seed, x0, y1, z1 ), xf0, yf1, zf1 ), gradDot3( hashPrimes( seed, x1, y1, z1 ), xf1, yf1, zf1 ), xs ), ys ), zs );
            constexpr double kB = 1.0363423824310302734375;
            return scaleOutput( n, value, (float)-kB, (float)kB );
        }
        static F perlin( const Node& n, I seed, Pos<4, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency ), w = p.c[3] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y ), zs = vfloor( z ), ws = vfloor( w );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY, z0 = cvti( zs ) * kPrimeZ, w0 = cvti( ws ) * kPrimeW;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY, z1 = z0 + kPrimeZ, w1 = w0 + kPrimeW;
            F xf0 = xs = x - xs, yf0 = ys = y - ys, zf0 = zs = z - zs, wf0 = ws = w - ws;
            F xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f, zf1 = zf0 - 1.0f, wf1 = wf0 - 1.0f;
            xs = interpQuintic( xs );
            ys = interpQuintic( ys );
            zs = interpQuintic( zs );
            ws = interpQuintic( ws );
            auto g = [&]( I xp, I yp, I zp, I wp, F a, F b, F c, F d ) { return gradDot4( hashPrimes( seed, xp, yp, zp, wp ), a, b, c, d ); };
            F value = lerp( lerp( lerp(
                lerp( g( x0, y0, z0, w0, xf0, yf0, zf0, wf0 ), g( x1, y0, z0, w0, xf1, yf0, zf0, wf0 ), xs ),
                lerp( g( x0, y1, z0, w0, xf0, yf1, zf0, wf0 ), g( x1, y1, z0, w0, xf1, yf1, zf0, wf0 ), xs ), ys ),
                lerp(
                lerp( g( x0, y0, z1, w0, xf0, yf0, zf1, wf0 ), g( x1, y0, z1, w0, xf1, yf0, zf1, wf0 ), xs ),
                lerp( g( x0, y1, z1, w0, xf0, yf1, zf1, wf0 ), g( x1, y1, z1, w0, xf1, yf1, zf1, wf0 ), xs ), ys ), zs ),
                lerp( lerp(
                lerp( g( x0, y0, z0, w1, xf0, yf0, zf0, wf1 ), g( x1, y0, z0, w1, xf1, yf0, zf0, wf1 ), xs ),
                lerp( g( x0, y1, z0, w1, xf0, yf1, zf0, wf1 ), g( x1, y1, z0, w1, xf1, yf1, zf0, wf1 ), xs ), ys ),
                lerp(
                lerp( g( x0, y0, z1, w1, xf0, yf0, zf1, wf1 ), g( x1, y0, z1, w1, xf1, yf0, zf1, wf1 ), xs ),
                lerp( g( x0, y1, z1, w1, xf0, yf1, zf1, wf1 ), g( x1, y1, z1, w1, xf1, yf1, zf1, wf1 ), xs ), ys ), zs ), ws );
            constexpr float kB = 1.33858621120452880859375f;
            return scaleOutput( n, value, -kB, kB );
        }

        // Value 2D/3D/4D (Value.inl:7-96)
        static F value( const Node& n, I seed, Pos<2, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY;
            xs = interpHermite( x - xs );
            ys = interpHermite( y - ys );
            return scaleOutput( n, lerp(
                lerp( valueCoord( seed, x0, y0 ), valueCoord( seed, x1, y0 ), xs ),
                lerp( valueCoord( seed, x0, y1 ), valueCoord( seed, x1, y1 ), xs ), ys ),
                -kValueBounds, kValueBounds );
        }
        static F value( const Node& n, I seed, Pos<3, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y ), zs = vfloor( z );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY, z0 = cvti( zs ) * kPrimeZ;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY, z1 = z0 + kPrimeZ;
            xs = interpHermite( x - xs );
            ys = interpHermite( y - ys );
            zs = interpHermite( z - zs );
            return scaleOutput( n, lerp( lerp(
                lerp( valueCoord( seed, x0, y0, z0 ), valueCoord( seed, x1, y0, z0 ), xs ),
                lerp( valueCoord( seed, x0, y1, z0 ), valueCoord( seed, x1, y1, z0 ), xs ), ys ),
                lerp(
                lerp( valueCoord( seed, x0, y0, z1 ), valueCoord( seed, x1, y0, z1 ), xs ),
                lerp( valueCoord( seed, x0, y1, z1 ), valueCoord( seed, x1, y1, z1 ), xs ), ys ), zs ),
                -kValueBounds, kValueBounds );
        }
        static F value( const Node& n, I seed, Pos<4, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency ), w = p.c[3] * S( n.frequency );
            F xs = vfloor( x ), ys = vfloor( y ), zs = vfloor( z ), ws = vfloor( w );
            I x0 = cvti( xs ) * kPrimeX, y0 = cvti( ys ) * kPrimeY, z0 = cvti( zs ) * kPrimeZ, w0 = cvti( ws ) * kPrimeW;
            I x1 = x0 + kPrimeX, y1 = y0 + kPrimeY, z1 = z0 + kPrimeZ, w1 = w0 + kPrimeW;
            xs = interpHermite( x - xs );
            ys = interpHermite( y - ys );
            zs = interpHermite( z - zs );
            ws = interpHermite( w - ws );
            auto v = [&]( I a, I b, I c, I d ) { return valueCoord( seed, a, b, c, d ); };
            return scaleOutput( n, lerp( lerp( lerp(
                lerp( v( x0, y0, z0, w0 ), v( x1, y0, z0, w0 ), xs ),
                lerp( v( x0, y1, z0, w0 ), v( x1, y1, z0, w0 ), xs ), ys ),
                lerp(
                lerp( v( x0, y0, z1, w0 ), v( x1, y0, z1, w0 ), xs ),
                lerp( v( x0, y1, z1, w0 ), v( x1, y1, z1, w0 ), xs ), ys ), zs ),
                lerp( lerp(
                lerp( v( x0, y0, z0, w1 ), v( x1, y0, z0, w1 ), xs ),
                lerp( v( x0, y1, z0, w1 ), v( x1, y1, z0, w1 ), xs ), ys ),
                lerp(
                lerp( v( x0, y0, z1, w1 ), v( x1, y0, z1, w1 ), xs ),
                lerp( v( x0, y1, z1, w1 ), v( x1, y1, z1, w1 ), xs ), ys ), zs ), ws ),
                -kValueBounds, kValueBounds );
        }

        // Simplex 2D (Simplex.inl:7-63)
        static F simplex( const Node& n, I seed, Pos<2, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency );
            constexpr double kSkew2 = 1.0 / ( kRoot3 + 1.0 );
            constexpr double kUnskew2 = -1.0 / ( kRoot3 + 3.0 );
            constexpr double kFalloff = 0.5;
            F skewDelta = S( (float)kSkew2 ) * ( x + y );
            F xSk = x + skewDelta, ySk = y + skewDelta;
            F xB = vfloor( xSk ), yB = vfloor( ySk );
            F dxS = xSk - xB, dyS = ySk - yB;
            I xP = cvti( xB ) * kPrimeX, yP = cvti( yB ) * kPrimeY;
            I xGEy = mge( dxS, dyS );
            F unskewDelta = S( (float)kUnskew2 ) * ( dxS + dyS );
            F dx0 = dxS + unskewDelta, dy0 = dyS + unskewDelta;
            F dx1 = minc( ~xGEy, dx0 ) - (float)( kUnskew2 + 1 );
            F dy1 = minc( xGEy, dy0 ) - (float)( kUnskew2 + 1 );
            F dx2 = dx0 - (float)( kUnskew2 * 2 + 1 );
            F dy2 = dy0 - (float)( kUnskew2 * 2 + 1 );
            F f0 = fnmadd( dx0, dx0, fnmadd( dy0, dy0, S( (float)kFalloff ) ) );
            F f1 = fnmadd( dx1, dx1, fnmadd( dy1, dy1, S( (float)kFalloff ) ) );
            F f2 = f0 + fmadd( unskewDelta, S( (float)( -4.0 * ( kRoot3 + 2.0 ) / ( kRoot3 + 3.0 ) ) ), S( (float)( -2.0 / 3.0 ) ) );
            f0 = vmax( f0, S( 0 ) ); f1 = vmax( f1, S( 0 ) ); f2 = vmax( f2, S( 0 ) );
            f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2;
            F g0 = gradDot2( hashPrimes( seed, xP, yP ), dx0, dy0 );
            F g1 = gradDot2( hashPrimes( seed, madd( xGEy, xP, SI( kPrimeX ) ), imadd( xGEy, yP, SI( kPrimeY ) ) ), dx1, dy1 );
            F g2 = gradDot2( hashPrimes( seed, xP + kPrimeX, yP + kPrimeY ), dx2, dy2 );
            F value = fmadd( g0, f0, fmadd( g1, f1, g2 * f2 ) );
            constexpr float kB = 0.0261208079755306243896484375f;
            return scaleOutput( n, value, -kB, kB );
        }
        // Simplex 3D (Simplex.inl:65-142)
        static F simplex( const Node& n, I seed, Pos<3, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency );
            constexpr double kSkew3 = 1.0 / 3.0;
            constexpr double kReflectUnskew3 = -1.0 / 2.0;
            constexpr double kFalloff = 0.6;
            F skewDelta = S( (float)kSkew3 ) * ( x + y + z );
            F xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta;
            F xB = vfloor( xSk ), yB = vfloor( ySk ), zB = vfloor( zSk );
            F dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB;
            I xP = cvti( xB ) * kPrimeX, yP = cvti( yB ) * kPrimeY, zP = cvti( zB ) * kPrimeZ;
            I xGEy = mge( dxS, dyS ), yGEz = mge( dyS, dzS ), xGEz = mge( dxS, dzS );
            F unskewDelta = S( (float)kReflectUnskew3 ) * ( dxS + dyS + dzS );
            F dx0 = dxS + unskewDelta, dy0 = dyS + unskewDelta, dz0 = dzS + unskewDelta;
            I maskX1 = xGEy & xGEz;
            I maskY1 = yGEz & ~xGEy;      // FS::BitwiseAndNot( a, b ) = a & ~b
            I maskZ1 = xGEz | yGEz;       // inverted
            I nMaskX2 = xGEy | xGEz;      // inverted
            I nMaskY2 = xGEy & ~yGEz;
            I nMaskZ2 = xGEz & yGEz;
            F dx3 = dx0 - (float)( kReflectUnskew3 * 3 + 1 );
            F dy3 = dy0 - (float)( kReflectUnskew3 * 3 + 1 );
            F dz3 = dz0 - (float)( kReflectUnskew3 * 3 + 1 );
            F dx1 = msub( maskX1, dx3, S( 1 ) );
            F dy1 = msub( maskY1, dy3, S( 1 ) );
            F dz1 = imsub( maskZ1, dz3, S( 1 ) );
            F dx2 = minc( ~nMaskX2, dx0 );
            F dy2 = minc( nMaskY2, dy0 );
            F dz2 = minc( nMaskZ2, dz0 );
            const F fr = S( (float)kFalloff );
            F f0 = fnmadd( dz0, dz0, fnmadd( dy0, dy0, fnmadd( dx0, dx0, fr ) ) );
            F f1 = fnmadd( dz1, dz1, fnmadd( dy1, dy1, fnmadd( dx1, dx1, fr ) ) );
            F f2 = fnmadd( dz2, dz2, fnmadd( dy2, dy2, fnmadd( dx2, dx2, fr ) ) );
            F f3 = f0 - ( unskewDelta + S( (float)( 3.0 / 4.0 ) ) );
            f0 = vmax( f0, S( 0 ) ); f1 = vmax( f1, S( 0 ) ); f2 = vmax( f2, S( 0 ) ); f3 = vmax( f3, S( 0 ) );
            f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2; f3 *= f3; f3 *= f3;
            F g0 = gradDot3( hashPrimes( seed, xP, yP, zP ), dx0, dy0, dz0 );
            F g1 = gradDot3( hashPrimes( seed, madd( maskX1, xP, SI( kPrimeX ) ), madd( maskY1, yP, SI( kPrimeY ) ), imadd( maskZ1, zP, SI( kPrimeZ ) ) ), dx1, dy1, dz1 );
            F g2 = gradDot3( hashPrimes( seed, madd( nMaskX2, xP, SI( kPrimeX ) ), imadd( nMaskY2, yP, SI( kPrimeY ) ), imadd( nMaskZ2, zP, SI( kPrimeZ ) ) ), dx2, dy2, dz2 );
            F g3 = gradDot3( hashPrimes( seed, xP + kPrimeX, yP + kPrimeY, zP + kPrimeZ ), dx3, dy3, dz3 );
            F value = fmadd( g3, f3, fmadd( g2, f2, fmadd( g1, f1, g0 * f0 ) ) );
            constexpr float kB = 0.030586399137973785400390625f;
            return scaleOutput( n, value, -kB, kB );
        }
        // Simplex 4D (Simplex.inl:144-286)
        static F simplex( const Node& n, I seed, Pos<4, N> p )
        {
            seed = seed + n.seedOffset;
            F x = p.c[0] * S( n.frequency ), y = p.c[1] * S( n.frequency ), z = p.c[2] * S( n.frequency ), w = p.c[3] * S( n.frequency );
            constexpr double kSkew4 = 1.0 / ( kRoot5 + 1.0 );
            constexpr double kUnskew4 = -1.0 / ( kRoot5 + 5.0 );
            constexpr double kFalloff = 0.6;
            F skewDelta = S( (float)kSkew4 ) * ( x + y + z + w );
            F xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta, wSk = w + skewDelta;
            F xB = vfloor( xSk ), yB = vfloor( ySk ), zB = vfloor( zSk ), wB = vfloor( wSk );
            F dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB, dwS = wSk - wB;
            I xP = cvti( xB ) * kPrimeX, yP = cvti( yB ) * kPrimeY, zP = cvti( zB ) * kPrimeZ, wP = cvti( wB ) * kPrimeW;
            F unskewDelta = S( (float)kUnskew4 ) * ( dxS + dyS + dzS + dwS );
            F dx0 = dxS + unskewDelta, dy0 = dyS + unskewDelta, dz0 = dzS + unskewDelta, dw0 = dwS + unskewDelta;
            I rX = SI( 0 ), rY = SI( 0 ), rZ = SI( 0 ), rW = SI( 0 );
            I m;
            m = mge( dx0, dy0 ); rX = minc( m, rX ); rY = minc( ~m, rY );
            m = mge( dx0, dz0 ); rX = minc( m, rX ); rZ = minc( ~m, rZ );
            m = mge( dx0, dw0 ); rX = minc( m, rX ); rW = minc( ~m, rW );
            m = mge( dy0, dz0 ); rY = minc( m, rY ); rZ = minc( ~m, rZ );
            m = mge( dy0, dw0 ); rY = minc( m, rY ); rW = minc( ~m, rW );
            m = mge( dz0, dw0 ); rZ = minc( m, rZ ); rW = minc( ~m, rW );
            I mX1 = mgt( rX, SI( 2 ) ), mY1 = mgt( rY, SI( 2 ) ), mZ1 = mgt( rZ, SI( 2 ) ), mW1 = mgt( rW, SI( 2 ) );
            I mX2 = mgt( rX, SI( 1 ) ), mY2 = mgt( rY, SI( 1 ) ), mZ2 = mgt( rZ, SI( 1 ) ), mW2 = mgt( rW, SI( 1 ) );
            I mX3 = mgt( rX, SI( 0 ) ), mY3 = mgt( rY, SI( 0 ) ), mZ3 = mgt( rZ, SI( 0 ) ), mW3 = mgt( rW, SI( 0 ) );
            const float u1 = (float)kUnskew4, u2 = (float)( kUnskew4 * 2 ), u3 = (float)( kUnskew4 * 3 ), u4 = (float)( kUnskew4 * 4 + 1 );
            F dx1 = msub( mX1, dx0, S( 1 ) ) - u1, dy1 = msub( mY1, dy0, S( 1 ) ) - u1, dz1 = msub( mZ1, dz0, S( 1 ) ) - u1, dw1 = msub( mW1, dw0, S( 1 ) ) - u1;
            F dx2 = msub( mX2, dx0, S( 1 ) ) - u2, dy2 = msub( mY2, dy0, S( 1 ) ) - u2, dz2 = msub( mZ2, dz0, S( 1 ) ) - u2, dw2 = msub( mW2, dw0, S( 1 ) ) - u2;
            F dx3 = msub( mX3, dx0, S( 1 ) ) - u3, dy3 = msub( mY3, dy0, S( 1 ) ) - u3, dz3 = msub( mZ3, dz0, S( 1 ) ) - u3, dw3 = msub( mW3, dw0, S( 1 ) ) - u3;
            F dx4 = dx0 - u4, dy4 = dy0 - u4, dz4 = dz0 - u4, dw4 = dw0 - u4;
            const F fr = S( (float)kFalloff );
            auto fo = [&]( F a, F b, F c, F d ) { return fnmadd( d, d, fnmadd( c, c, fnmadd( b, b, fnmadd( a, a, fr ) ) ) ); };
            F f0 = fo( dx0, dy0, dz0, dw0 ), f1 = fo( dx1, dy1, dz1, dw1 ), f2 = fo( dx2, dy2, dz2, dw2 ), f3 = fo( dx3, dy3, dz3, dw3 );
            F f4 = f0 + fmadd( unskewDelta, S( (float)( -4.0 * ( kRoot5 + 3.0 ) / ( kRoot5 + 5.0 ) ) ), S( (float)( -4.0 / 5.0 ) ) );
            f0 = vmax( f0, S( 0 ) ); f1 = vmax( f1, S( 0 ) ); f2 = vmax( f2, S( 0 ) ); f3 = vmax( f3, S( 0 ) ); f4 = vmax( f4, S( 0 ) );
            f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2; f3 *= f3; f3 *= f3; f4 *= f4; f4 *= f4;
            F g0 = gradDot4( hashPrimes( seed, xP, yP, zP, wP ), dx0, dy0, dz0, dw0 );
            F g1 = gradDot4( hashPrimes( seed, madd( mX1, xP, SI( kPrimeX ) ), madd( mY1, yP, SI( kPrimeY ) ), madd( mZ1, zP, SI( kPrimeZ ) ), madd( mW1, wP, SI( kPrimeW ) ) ), dx1, dy1, dz1, dw1 );
            F g2 = gradDot4( hashPrimes( seed, madd( mX2, xP, SI( kPrimeX ) ), madd( mY2, yP, SI( kPrimeY ) ), madd( mZ2, zP, SI( kPrimeZ ) ), madd( mW2, wP, SI( kPrimeW ) ) ), dx2, dy2, dz2, dw2 );
            F g3 = gradDot4( hashPrimes( seed, madd( mX3, xP, SI( kPrimeX ) ), madd( mY3, yP, SI( kPrimeY ) ), madd( mZ3, zP, SI( kPrimeZ ) ), madd( mW3, wP, SI( kPrimeW ) ) ), dx3, dy3, dz3, dw3 );
            F g4 = gradDot4( hashPrimes( seed, xP + kPrimeX, yP + kPrimeY, zP + kPrimeZ, wP + kPrimeW ), dx4, dy4, dz4, dw4 );
            F value = fmadd( g0, f0, fmadd( g1, f1, fmadd( g2, f2, fmadd( g3, f3, g4 * f4 ) ) ) );
            constexpr float kB = 0.036730043590068817138671875f;
            return scaleOutput( n, value, -kB, kB );
        }

#include "fn2_oracle_eval_supersimplex.inc"
#include "fn2_oracle_eval_cellular.inc"
#include "fn2_oracle_eval_warp.inc"
#include "fn2_oracle_eval_misc.inc"
    };
}
