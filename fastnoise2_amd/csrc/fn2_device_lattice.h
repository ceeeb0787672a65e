// Device lattice noises for gfx950: Perlin, Value, Simplex and SuperSimplex
// (OpenSimplex2S) in 2D/3D/4D, one point per lane, following FastNoise2 v1.1.1
// include/FastNoise/Generators/{Perlin,Value,Simplex}.inl operation by
// operation. The SuperSimplex lattice walks take a per-vertex functor so the
// DomainWarpSuperSimplex warp reuses them with a vector contribution.
#pragma once
#include "fn2_device_math.h"

namespace fn2d
{
    struct LeafParams
    {
        float freq, nativeMin, outScale, rangeMin;
    };

    // ---------------- Perlin (Perlin.inl:7-125) ----------------
    template<bool G = true>
    FN2D float perlin2( int32_t seed, float x, float y )
    {
        float xs = ffloor( x ), ys = ffloor( y );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY );
        float xf0 = x - xs, yf0 = y - ys;
        float xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f;
        float xq = interpQuintic( xf0 ), yq = interpQuintic( yf0 );
        return lerp( lerp( gradDot2( hash2( seed, x0, y0 ), xf0, yf0 ), gradDot2( hash2( seed, x1, y0 ), xf1, yf0 ), xq ),
                     lerp( gradDot2( hash2( seed, x0, y1 ), xf0, yf1 ), gradDot2( hash2( seed, x1, y1 ), xf1, yf1 ), xq ), yq );
    }
    template<bool G = true>
    FN2D float perlin3( int32_t seed, float x, float y, float z )
    {
        float xs = ffloor( x ), ys = ffloor( y ), zs = ffloor( z );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY ), z0 = imul( floorToInt<G>( zs ), kPZ );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY ), z1 = iadd( z0, kPZ );
        float xf0 = x - xs, yf0 = y - ys, zf0 = z - zs;
        float xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f, zf1 = zf0 - 1.0f;
        float xq = interpQuintic( xf0 ), yq = interpQuintic( yf0 ), zq = interpQuintic( zf0 );
        float l0 = lerp( lerp( gradDot3( hash3( seed, x0, y0, z0 ), xf0, yf0, zf0 ), gradDot3( hash3( seed, x1, y0, z0 ), xf1, yf0, zf0 ), xq ),
                         lerp( gradDot3( hash3( seed, x0, y1, z0 ), xf0, yf1, zf0 ), gradDot3( hash3( seed, x1, y1, z0 ), xf1, yf1, zf0 ), xq ), yq );
        float l1 = lerp( lerp( gradDot3( hash3( seed, x0, y0, z1 ), xf0, yf0, zf1 ), gradDot3( hash3( seed, x1, y0, z1 ), xf1, yf0, zf1 ), xq ),
                         lerp( gradDot3( hash3( seed, x0, y1, z1 ), xf0, yf1, zf1 ), gradDot3( hash3( seed, x1, y1, z1 ), xf1, yf1, zf1 ), xq ), yq );
        return lerp( l0, l1, zq );
    }
    template<bool G = true>
    FN2D float perlin4( int32_t seed, float x, float y, float z, float w )
    {
        float xs = ffloor( x ), ys = ffloor( y ), zs = ffloor( z ), ws = ffloor( w );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY ), z0 = imul( floorToInt<G>( zs ), kPZ ), w0 = imul( floorToInt<G>( ws ), kPW );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY ), z1 = iadd( z0, kPZ ), w1 = iadd( w0, kPW );
        float xf0 = x - xs, yf0 = y - ys, zf0 = z - zs, wf0 = w - ws;
        float xf1 = xf0 - 1.0f, yf1 = yf0 - 1.0f, zf1 = zf0 - 1.0f, wf1 = wf0 - 1.0f;
        float xq = interpQuintic( xf0 ), yq = interpQuintic( yf0 ), zq = interpQuintic( zf0 ), wq = interpQuintic( wf0 );
        float v[2];
#pragma unroll
        for( int wi = 0; wi < 2; wi++ )
        {
            int32_t wp = wi ? w1 : w0;
            float wf = wi ? wf1 : wf0;
            float zl[2];
#pragma unroll
            for( int zi = 0; zi < 2; zi++ )
            {
                int32_t zp = zi ? z1 : z0;
                float zf = zi ? zf1 : zf0;
                zl[zi] = lerp( lerp( gradDot4( hash4( seed, x0, y0, zp, wp ), xf0, yf0, zf, wf ), gradDot4( hash4( seed, x1, y0, zp, wp ), xf1, yf0, zf, wf ), xq ),
                               lerp( gradDot4( hash4( seed, x0, y1, zp, wp ), xf0, yf1, zf, wf ), gradDot4( hash4( seed, x1, y1, zp, wp ), xf1, yf1, zf, wf ), xq ), yq );
            }
            v[wi] = lerp( zl[0], zl[1], zq );
        }
        return lerp( v[0], v[1], wq );
    }

    // ---------------- Value (Value.inl:7-96) ----------------
    template<bool G = true>
    FN2D float value2( int32_t seed, float x, float y )
    {
        float xs = ffloor( x ), ys = ffloor( y );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY );
        float xh = interpHermite( x - xs ), yh = interpHermite( y - ys );
        return lerp( lerp( valueCoord2( seed, x0, y0 ), valueCoord2( seed, x1, y0 ), xh ),
                     lerp( valueCoord2( seed, x0, y1 ), valueCoord2( seed, x1, y1 ), xh ), yh );
    }
    template<bool G = true>
    FN2D float value3( int32_t seed, float x, float y, float z )
    {
        float xs = ffloor( x ), ys = ffloor( y ), zs = ffloor( z );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY ), z0 = imul( floorToInt<G>( zs ), kPZ );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY ), z1 = iadd( z0, kPZ );
        float xh = interpHermite( x - xs ), yh = interpHermite( y - ys ), zh = interpHermite( z - zs );
        float l0 = lerp( lerp( valueCoord3( seed, x0, y0, z0 ), valueCoord3( seed, x1, y0, z0 ), xh ),
                         lerp( valueCoord3( seed, x0, y1, z0 ), valueCoord3( seed, x1, y1, z0 ), xh ), yh );
        float l1 = lerp( lerp( valueCoord3( seed, x0, y0, z1 ), valueCoord3( seed, x1, y0, z1 ), xh ),
                         lerp( valueCoord3( seed, x0, y1, z1 ), valueCoord3( seed, x1, y1, z1 ), xh ), yh );
        return lerp( l0, l1, zh );
    }
    template<bool G = true>
    FN2D float value4( int32_t seed, float x, float y, float z, float w )
    {
        float xs = ffloor( x ), ys = ffloor( y ), zs = ffloor( z ), ws = ffloor( w );
        int32_t x0 = imul( floorToInt<G>( xs ), kPX ), y0 = imul( floorToInt<G>( ys ), kPY ), z0 = imul( floorToInt<G>( zs ), kPZ ), w0 = imul( floorToInt<G>( ws ), kPW );
        int32_t x1 = iadd( x0, kPX ), y1 = iadd( y0, kPY ), z1 = iadd( z0, kPZ ), w1 = iadd( w0, kPW );
        float xh = interpHermite( x - xs ), yh = interpHermite( y - ys ), zh = interpHermite( z - zs ), wh = interpHermite( w - ws );
        float v[2];
#pragma unroll
        for( int wi = 0; wi < 2; wi++ )
        {
            int32_t wp = wi ? w1 : w0;
            float zl[2];
#pragma unroll
            for( int zi = 0; zi < 2; zi++ )
            {
                int32_t zp = zi ? z1 : z0;
                zl[zi] = lerp( lerp( valueCoord4( seed, x0, y0, zp, wp ), valueCoord4( seed, x1, y0, zp, wp ), xh ),
                               lerp( valueCoord4( seed, x0, y1, zp, wp ), valueCoord4( seed, x1, y1, zp, wp ), xh ), yh );
            }
            v[wi] = lerp( zl[0], zl[1], zh );
        }
        return lerp( v[0], v[1], wh );
    }

    // ---------------- Simplex (Simplex.inl:7-286) ----------------
    // Returns the un-scaled value; the three falloffs and gradients are combined
    // in the reference's order for each dimension.
    // U: try the wave-uniform cell first. When every active lane of a wave
    // lies in the same skewed lattice cell (grid bricks at low octaves), the
    // cell's vertex hashes are the same on every lane: they are computed from
    // the first lane's cell with plain xors, which the compiler keeps on the
    // scalar unit, and each lane only picks its middle vertex. The values
    // are those of the per-lane path.
    template<bool G = true, bool U = false>
    FN2D float simplex2( int32_t seed, float x, float y )
    {
        constexpr double kSkew2 = 1.0 / ( kRoot3 + 1.0 );
        constexpr double kUnskew2 = -1.0 / ( kRoot3 + 3.0 );
        float skewDelta = (float)kSkew2 * ( x + y );
        float xSk = x + skewDelta, ySk = y + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk );
        float dxS = xSk - xB, dyS = ySk - yB;
        const int32_t xi = floorToInt<G>( xB ), yi = floorToInt<G>( yB );
        bool xGEy = dxS >= dyS;
        float unskewDelta = (float)kUnskew2 * ( dxS + dyS );
        float dx0 = dxS + unskewDelta, dy0 = dyS + unskewDelta;
        float dx1 = ( xGEy ? dx0 : dx0 + 1.0f ) - (float)( kUnskew2 + 1 );
        float dy1 = ( xGEy ? dy0 + 1.0f : dy0 ) - (float)( kUnskew2 + 1 );
        float dx2 = dx0 - (float)( kUnskew2 * 2 + 1 );
        float dy2 = dy0 - (float)( kUnskew2 * 2 + 1 );
        float f0 = fnmadd( dx0, dx0, fnmadd( dy0, dy0, 0.5f ) );
        float f1 = fnmadd( dx1, dx1, fnmadd( dy1, dy1, 0.5f ) );
        float f2 = f0 + fmadd( unskewDelta, (float)( -4.0 * ( kRoot3 + 2.0 ) / ( kRoot3 + 3.0 ) ), (float)( -2.0 / 3.0 ) );
        f0 = fmaxf_( f0, 0.0f ); f1 = fmaxf_( f1, 0.0f ); f2 = fmaxf_( f2, 0.0f );
        f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2;
#if defined( __HIP_DEVICE_COMPILE__ )
        if constexpr( U )
        {
            const int32_t xu = __builtin_amdgcn_readfirstlane( xi ), yu = __builtin_amdgcn_readfirstlane( yi );
            const int32_t su = __builtin_amdgcn_readfirstlane( seed );
            if( waveAll( xi == xu && yi == yu && seed == su ) )
            {
                const int32_t xP = imul( xu, kPX ), yP = imul( yu, kPY );
                const int32_t x1 = iadd( xP, kPX ), y1 = iadd( yP, kPY );
                const int32_t h0 = hashFinal( su ^ xP ^ yP ), h2 = hashFinal( su ^ x1 ^ y1 );
                const int32_t hX = hashFinal( su ^ x1 ^ yP ), hY = hashFinal( su ^ xP ^ y1 );
                float g0 = gradDot2( h0, dx0, dy0 );
                float g1 = gradDot2( xGEy ? hX : hY, dx1, dy1 );
                float g2 = gradDot2( h2, dx2, dy2 );
                return fmadd( g0, f0, fmadd( g1, f1, g2 * f2 ) );
            }
        }
#endif
        int32_t xP = imul( xi, kPX ), yP = imul( yi, kPY );
        float g0 = gradDot2( hash2( seed, xP, yP ), dx0, dy0 );
        float g1 = gradDot2( hash2( seed, xGEy ? iadd( xP, kPX ) : xP, xGEy ? yP : iadd( yP, kPY ) ), dx1, dy1 );
        float g2 = gradDot2( hash2( seed, iadd( xP, kPX ), iadd( yP, kPY ) ), dx2, dy2 );
        return fmadd( g0, f0, fmadd( g1, f1, g2 * f2 ) );
    }
    template<bool G = true>
    FN2D float simplex3( int32_t seed, float x, float y, float z )
    {
        constexpr double kReflectUnskew3 = -1.0 / 2.0;
        float skewDelta = (float)( 1.0 / 3.0 ) * ( x + y + z );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB;
        int32_t xP = imul( floorToInt<G>( xB ), kPX ), yP = imul( floorToInt<G>( yB ), kPY ), zP = imul( floorToInt<G>( zB ), kPZ );
        bool xGEy = dxS >= dyS, yGEz = dyS >= dzS, xGEz = dxS >= dzS;
        float unskewDelta = (float)kReflectUnskew3 * ( dxS + dyS + dzS );
        float dx0 = dxS + unskewDelta, dy0 = dyS + unskewDelta, dz0 = dzS + unskewDelta;
        // vertex 1 = the largest axis; vertex 2 = all but the smallest axis
        bool x1m = xGEy && xGEz, y1m = yGEz && !xGEy, z1inv = xGEz || yGEz;
        bool x2inv = xGEy || xGEz, y2m = xGEy && !yGEz, z2m = xGEz && yGEz;
        const float k3 = (float)( kReflectUnskew3 * 3 + 1 );
        float dx3 = dx0 - k3, dy3 = dy0 - k3, dz3 = dz0 - k3;
        float dx1 = x1m ? dx3 - 1.0f : dx3;
        float dy1 = y1m ? dy3 - 1.0f : dy3;
        float dz1 = z1inv ? dz3 : dz3 - 1.0f;
        float dx2 = x2inv ? dx0 : dx0 + 1.0f;
        float dy2 = y2m ? dy0 + 1.0f : dy0;
        float dz2 = z2m ? dz0 + 1.0f : dz0;
        const float fr = (float)0.6;
        float f0 = fnmadd( dz0, dz0, fnmadd( dy0, dy0, fnmadd( dx0, dx0, fr ) ) );
        float f1 = fnmadd( dz1, dz1, fnmadd( dy1, dy1, fnmadd( dx1, dx1, fr ) ) );
        float f2 = fnmadd( dz2, dz2, fnmadd( dy2, dy2, fnmadd( dx2, dx2, fr ) ) );
        float f3 = f0 - ( unskewDelta + (float)( 3.0 / 4.0 ) );
        f0 = fmaxf_( f0, 0.0f ); f1 = fmaxf_( f1, 0.0f ); f2 = fmaxf_( f2, 0.0f ); f3 = fmaxf_( f3, 0.0f );
        f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2; f3 *= f3; f3 *= f3;
        float g0 = gradDot3( hash3( seed, xP, yP, zP ), dx0, dy0, dz0 );
        float g1 = gradDot3( hash3( seed, x1m ? iadd( xP, kPX ) : xP, y1m ? iadd( yP, kPY ) : yP, z1inv ? zP : iadd( zP, kPZ ) ), dx1, dy1, dz1 );
        float g2 = gradDot3( hash3( seed, x2inv ? iadd( xP, kPX ) : xP, y2m ? yP : iadd( yP, kPY ), z2m ? zP : iadd( zP, kPZ ) ), dx2, dy2, dz2 );
        float g3 = gradDot3( hash3( seed, iadd( xP, kPX ), iadd( yP, kPY ), iadd( zP, kPZ ) ), dx3, dy3, dz3 );
        return fmadd( g3, f3, fmadd( g2, f2, fmadd( g1, f1, g0 * f0 ) ) );
    }
    template<bool G = true>
    FN2D float simplex4( int32_t seed, float x, float y, float z, float w )
    {
        constexpr double kSkew4 = 1.0 / ( kRoot5 + 1.0 );
        constexpr double kUnskew4 = -1.0 / ( kRoot5 + 5.0 );
        float skewDelta = (float)kSkew4 * ( x + y + z + w );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta, wSk = w + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk ), wB = ffloor( wSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB, dwS = wSk - wB;
        int32_t xP = imul( floorToInt<G>( xB ), kPX ), yP = imul( floorToInt<G>( yB ), kPY ), zP = imul( floorToInt<G>( zB ), kPZ ), wP = imul( floorToInt<G>( wB ), kPW );
        float unskewDelta = (float)kUnskew4 * ( dxS + dyS + dzS + dwS );
        float d0[4] = { dxS + unskewDelta, dyS + unskewDelta, dzS + unskewDelta, dwS + unskewDelta };
        // rank of each axis among the four (ties resolved by axis order)
        int rank[4] = { 0, 0, 0, 0 };
#pragma unroll
        for( int a = 0; a < 4; a++ )
#pragma unroll
            for( int b = a + 1; b < 4; b++ )
            {
                if( d0[a] >= d0[b] ) rank[a]++; else rank[b]++;
            }
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        const int32_t base[4] = { xP, yP, zP, wP };
        const float fr = (float)0.6;
        float value = 0.0f;
        float f0 = fnmadd( d0[3], d0[3], fnmadd( d0[2], d0[2], fnmadd( d0[1], d0[1], fnmadd( d0[0], d0[0], fr ) ) ) );
        float f4 = f0 + fmadd( unskewDelta, (float)( -4.0 * ( kRoot5 + 3.0 ) / ( kRoot5 + 5.0 ) ), (float)( -4.0 / 5.0 ) );
        // vertices 4, 3, 2, 1 then 0 to reproduce FMulAdd( g0, f0, FMulAdd( g1, f1, ... g4 * f4 ) )
        {
            float f = fmaxf_( f4, 0.0f ); f *= f; f *= f;
            const float u4 = (float)( kUnskew4 * 4 + 1 );
            float g = gradDot4( hash4( seed, iadd( xP, kPX ), iadd( yP, kPY ), iadd( zP, kPZ ), iadd( wP, kPW ) ), d0[0] - u4, d0[1] - u4, d0[2] - u4, d0[3] - u4 );
            value = g * f;
        }
#pragma unroll
        for( int v = 3; v >= 1; v-- )
        {
            const float u = (float)( kUnskew4 * v );
            float d[4];
            int32_t hp[4];
#pragma unroll
            for( int a = 0; a < 4; a++ )
            {
                bool m = rank[a] > 3 - v;
                d[a] = ( m ? d0[a] - 1.0f : d0[a] ) - u;
                hp[a] = m ? iadd( base[a], prime[a] ) : base[a];
            }
            float f = fnmadd( d[3], d[3], fnmadd( d[2], d[2], fnmadd( d[1], d[1], fnmadd( d[0], d[0], fr ) ) ) );
            f = fmaxf_( f, 0.0f ); f *= f; f *= f;
            value = fmadd( gradDot4( hash4( seed, hp[0], hp[1], hp[2], hp[3] ), d[0], d[1], d[2], d[3] ), f, value );
        }
        {
            float f = fmaxf_( f0, 0.0f ); f *= f; f *= f;
            value = fmadd( gradDot4( hash4( seed, xP, yP, zP, wP ), d0[0], d0[1], d0[2], d0[3] ), f, value );
        }
        return value;
    }

    // ---------------- SuperSimplex walks (Simplex.inl:291-936) ----------------
    // vertex( hash, d[], falloff^4, first )
    template<bool G = true, typename VTX>
    FN2D void superWalk2( int32_t seed, float x, float y, VTX&& vertex )
    {
        constexpr double kSkew2 = 1.0 / ( kRoot3 + 1.0 );
        constexpr double kUnskew2 = -1.0 / ( kRoot3 + 3.0 );
        const float kFall = (float)( 2.0 / 3.0 );
        float skewDelta = (float)kSkew2 * ( x + y );
        float xSk = x + skewDelta, ySk = y + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk );
        float dxS = xSk - xB, dyS = ySk - yB;
        int32_t xP = imul( floorToInt<G>( xB ), kPX ), yP = imul( floorToInt<G>( yB ), kPY );
        bool fXY = dxS + dyS > 1.0f;
        float boundary = fXY ? -1.0f : 0.0f;
        bool fX = fmadd( dxS, -2.0f, dyS ) < boundary;
        bool fY = fmadd( dyS, -2.0f, dxS ) < boundary;
        float unskewDelta = (float)kUnskew2 * ( dxS + dyS );
        float dxB = dxS + unskewDelta, dyB = dyS + unskewDelta;
        float fb0 = fnmadd( dxB, dxB, fnmadd( dyB, dyB, kFall ) );
        {
            float f = fb0; f *= f; f *= f;
            float d[2] = { dxB, dyB };
            vertex( hash2( seed, xP, yP ), d, f, true );
        }
        {
            float f = fmadd( unskewDelta, (float)( -4.0 * ( kRoot3 + 2.0 ) / ( kRoot3 + 3.0 ) ), fb0 - kFall );
            f *= f; f *= f;
            const float o = (float)( 2 * kUnskew2 + 1 );
            float d[2] = { dxB - o, dyB - o };
            vertex( hash2( seed, iadd( xP, kPX ), iadd( yP, kPY ) ), d, f, false );
        }
        float xyDelta = fXY ? (float)( kUnskew2 + 1 ) : (float)-kUnskew2;
        dxB -= xyDelta;
        dyB -= xyDelta;
        const int32_t kPX2 = imul( kPX, 2 ), kPY2 = imul( kPY, 2 );
        {
            int32_t hx = fX ? iadd( xP, kPX2 ) : xP;
            hx = fXY ? hx : isub( hx, kPX );
            int32_t hy = fXY ? iadd( yP, kPY ) : yP;
            float dx = dxB - ( fX ? (float)( 1 + 2 * kUnskew2 ) : -1.0f );
            float dy = fX ? dyB - (float)( 2 * kUnskew2 ) : dyB;
            float f = fmaxf_( fnmadd( dx, dx, fnmadd( dy, dy, kFall ) ), 0.0f );
            f *= f; f *= f;
            float d[2] = { dx, dy };
            vertex( hash2( seed, hx, hy ), d, f, false );
        }
        {
            int32_t hx = fXY ? iadd( xP, kPX ) : xP;
            int32_t hy = fY ? iadd( yP, kPY2 ) : yP;
            hy = fXY ? hy : isub( hy, kPY );
            float dx = fY ? dxB - (float)( 2 * kUnskew2 ) : dxB;
            float dy = dyB - ( fY ? (float)( 1 + 2 * kUnskew2 ) : -1.0f );
            float f = fmaxf_( fnmadd( dx, dx, fnmadd( dy, dy, kFall ) ), 0.0f );
            f *= f; f *= f;
            float d[2] = { dx, dy };
            vertex( hash2( seed, hx, hy ), d, f, false );
        }
    }

    FN2D int32_t signBitIf( bool m ) { return m ? (int32_t)0x80000000u : 0; }

    template<bool G = true, typename VTX>
    FN2D void superWalk3( int32_t seed, float x, float y, float z, VTX&& vertex )
    {
        constexpr double kReflectUnskew3 = -1.0 / 2.0;
        constexpr double kTwiceUnskew3 = -1.0 / 4.0;
        const float kDistA = (float)( 3.0 / 4.0 );
        float skewDelta = (float)( 1.0 / 3.0 ) * ( x + y + z );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB;
        {
            float tud = (float)kTwiceUnskew3 * ( dxS + dyS + dzS );
            float xN = dxS + tud, yN = dyS + tud, zN = dzS + tud, xyzN = -tud;
            float best = 0.375f;
            int32_t bits = xyzN > best ? -1 : 0;
            best = fmaxf_( best, xyzN );
            auto consider = [&]( float score, int32_t b ) { bits = score > best ? b : bits; best = fmaxf_( best, score ); };
            consider( xN, 0b001 ); consider( yN, 0b010 ); consider( zN, 0b100 );
            best += 0.125f - xyzN;
            consider( -zN, 0b011 ); consider( -yN, 0b101 ); consider( -xN, 0b110 );
            if( bits & 1 ) { xB += 1.0f; dxS -= 1.0f; }
            if( bits & 2 ) { yB += 1.0f; dyS -= 1.0f; }
            if( bits & 4 ) { zB += 1.0f; dzS -= 1.0f; }
        }
        int32_t xP = imul( floorToInt<G>( xB ), kPX ), yP = imul( floorToInt<G>( yB ), kPY ), zP = imul( floorToInt<G>( zB ), kPZ );
        float sum = dxS + dyS + dzS;
        float tud = (float)kTwiceUnskew3 * sum;
        float xN = dxS + tud, yN = dyS + tud, zN = dzS + tud, xyzN = -tud;
        float unskewDelta = (float)kReflectUnskew3 * sum;
        float dxB = dxS + unskewDelta, dyB = dyS + unskewDelta, dzB = dzS + unskewDelta;
        float coordSum = (float)( 1 + 3 * kReflectUnskew3 ) * sum;
        float stemA, stemB;
        {
            float fb = fnmadd( dzB, dzB, fnmadd( dyB, dyB, fnmadd( dxB, dxB, kDistA ) ) ) * 0.5f;
            stemA = fb - (float)( 3.0 / 4.0 * 0.5 );
            stemB = fb - (float)( 1.0 * 0.5 );
            float d[3] = { dxB, dyB, dzB };
            vertex( hash3( seed, xP, yP, zP ), d, ( fb * fb ) * ( fb * fb ), true );
        }
        auto pk = [&]( bool m, int32_t b, int32_t p ) { return iadd( b, m ? -p : p ); };
        {   // <1,1,1>
            bool sm = xyzN < 0.0f;
            float off = fxor( (float)( 3 * kReflectUnskew3 + 1 ), signBitIf( sm ) );
            float fb = fmaxf_( fmadd( off, coordSum, stemA ), 0.0f );
            float d[3] = { dxB - off, dyB - off, dzB - off };
            vertex( hash3( seed, pk( sm, xP, kPX ), pk( sm, yP, kPY ), pk( sm, zP, kPZ ) ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <1,1,0>
            bool sm = xyzN < zN;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)( 2 * kReflectUnskew3 ), s );
            float fb = fminf_( fxor( dzB, s ) - stemB, 0.0f );
            float d[3] = { dxB, dyB, dzB - off0 };
            vertex( hash3( seed, pk( sm, xP, kPX ), pk( sm, yP, kPY ), zP ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <1,0,1>
            bool sm = xyzN < yN;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)( 2 * kReflectUnskew3 ), s );
            float fb = fminf_( fxor( dyB, s ) - stemB, 0.0f );
            float d[3] = { dxB, dyB - off0, dzB };
            vertex( hash3( seed, pk( sm, xP, kPX ), yP, pk( sm, zP, kPZ ) ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <0,1,1>
            bool sm = xyzN < xN;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)( 2 * kReflectUnskew3 ), s );
            float fb = fminf_( fxor( dxB, s ) - stemB, 0.0f );
            float d[3] = { dxB - off0, dyB, dzB };
            vertex( hash3( seed, xP, pk( sm, yP, kPY ), pk( sm, zP, kPZ ) ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <1,0,0>
            bool sm = xN < 0.0f;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)kReflectUnskew3, s );
            float fb = fmaxf_( fmadd( off0, coordSum, stemA ) + fxor( dxB, s ), 0.0f );
            float d[3] = { dxB + off0, dyB - off0, dzB - off0 };
            vertex( hash3( seed, pk( sm, xP, kPX ), yP, zP ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <0,1,0>
            bool sm = yN < 0.0f;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)kReflectUnskew3, s );
            float fb = fmaxf_( fmadd( off0, coordSum, stemA ) + fxor( dyB, s ), 0.0f );
            float d[3] = { dxB - off0, dyB + off0, dzB - off0 };
            vertex( hash3( seed, xP, pk( sm, yP, kPY ), zP ), d, ( fb * fb ) * ( fb * fb ), false );
        }
        {   // <0,0,1>
            bool sm = zN < 0.0f;
            int32_t s = signBitIf( sm );
            float off0 = fxor( (float)kReflectUnskew3, s );
            float fb = fmaxf_( fmadd( off0, coordSum, stemA ) + fxor( dzB, s ), 0.0f );
            float d[3] = { dxB - off0, dyB - off0, dzB + off0 };
            vertex( hash3( seed, xP, yP, pk( sm, zP, kPZ ) ), d, ( fb * fb ) * ( fb * fb ), false );
        }
    }

    template<bool G = true, typename VTX>
    FN2D void superWalk4( int32_t seed, float x, float y, float z, float w, VTX&& vertex )
    {
        constexpr double kSkew4 = 1.0 / ( kRoot5 + 1.0 );
        constexpr double kUnskew4 = -1.0 / ( kRoot5 + 5.0 );
        constexpr double kTwiceUnskew4 = -1.0 / 5.0;
        const float kDistA = (float)( 4.0 / 5.0 );
        float skewDelta = (float)kSkew4 * ( x + y + z + w );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta, wSk = w + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk ), wB = ffloor( wSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB, dwS = wSk - wB;
        {
            float tud = (float)kTwiceUnskew4 * ( dxS + dyS + dzS + dwS );
            float xN = dxS + tud, yN = dyS + tud, zN = dzS + tud, wN = dwS + tud, xyzwN = -tud;
            float best = 0.6f - xyzwN;
            int32_t bits = 0.2f > best ? -1 : 0;
            best = fmaxf_( best, 0.2f );
            auto consider = [&]( float score, int32_t b ) { bits = score > best ? b : bits; best = fmaxf_( best, score ); };
            consider( -wN, 0b0111 ); consider( -zN, 0b1011 ); consider( -yN, 0b1101 ); consider( -xN, 0b1110 );
            best += xyzwN - 0.2f;
            consider( xN, 0b0001 ); consider( yN, 0b0010 ); consider( zN, 0b0100 ); consider( wN, 0b1000 );
            best += 0.2f - xN;
            consider( yN, 0b0011 ); consider( zN, 0b0101 ); consider( wN, 0b1001 );
            best += xN;
            consider( yN + zN, 0b0110 );
            best -= wN;
            consider( yN, 0b1010 ); consider( zN, 0b1100 );
            if( bits & 1 ) { xB += 1.0f; dxS -= 1.0f; }
            if( bits & 2 ) { yB += 1.0f; dyS -= 1.0f; }
            if( bits & 4 ) { zB += 1.0f; dzS -= 1.0f; }
            if( bits & 8 ) { wB += 1.0f; dwS -= 1.0f; }
        }
        int32_t P[4] = { imul( floorToInt<G>( xB ), kPX ), imul( floorToInt<G>( yB ), kPY ), imul( floorToInt<G>( zB ), kPZ ), imul( floorToInt<G>( wB ), kPW ) };
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        float sum = dxS + dyS + dzS + dwS;
        float tud = (float)kTwiceUnskew4 * sum;
        float Nrm[4] = { dxS + tud, dyS + tud, dzS + tud, dwS + tud };
        float xyzwN = -tud;
        float unskewDelta = (float)kUnskew4 * sum;
        float dB[4] = { dxS + unskewDelta, dyS + unskewDelta, dzS + unskewDelta, dwS + unskewDelta };
        float coordSum = (float)( 1 + 4 * kUnskew4 ) * sum;
        float stemA, stemB;
        {
            float fb = fnmadd( dB[3], dB[3], fnmadd( dB[2], dB[2], fnmadd( dB[1], dB[1], fnmadd( dB[0], dB[0], kDistA ) ) ) ) * 0.5f;
            stemA = fb - (float)( 4.0 / 5.0 * 0.5 );
            stemB = fb - (float)( 6.0 / 5.0 * 0.5 );
            vertex( hash4( seed, P[0], P[1], P[2], P[3] ), dB, ( fb * fb ) * ( fb * fb ), true );
        }
        // One vertex: axes in `mask` move by +-1 with sign from `sm`.
        auto emit = [&]( bool sm, int mask, float offIn, float offOut, float fbPre ) {
            int32_t s = signBitIf( sm );
            float o1 = fxor( offIn, s ), o0 = fxor( offOut, s );
            float d[4];
            int32_t hp[4];
#pragma unroll
            for( int a = 0; a < 4; a++ )
            {
                bool in = ( mask >> a ) & 1;
                d[a] = dB[a] - ( in ? o1 : o0 );
                hp[a] = in ? iadd( P[a], sm ? -prime[a] : prime[a] ) : P[a];
            }
            float fb = fmaxf_( fbPre, 0.0f );
            vertex( hash4( seed, hp[0], hp[1], hp[2], hp[3] ), d, ( fb * fb ) * ( fb * fb ), false );
        };
        {   // <1,1,1,1>
            bool sm = xyzwN < 0.0f;
            float off = fxor( (float)( 4 * kUnskew4 + 1 ), signBitIf( sm ) );
            float fbPre = fmadd( off, coordSum, stemA );
            emit( sm, 0xF, (float)( 4 * kUnskew4 + 1 ), (float)( 4 * kUnskew4 + 1 ), fbPre );
        }
        // <1,1,1,0> class: the excluded axis e, sign from xyzwN < N[e]
#pragma unroll
        for( int e = 3; e >= 0; e-- )
        {
            bool sm = xyzwN < Nrm[e];
            int32_t s = signBitIf( sm );
            float o1 = fxor( (float)( 3 * kUnskew4 + 1 ), s );
            float fbPre = fmadd( o1, coordSum, stemB ) - fxor( dB[e], s );
            emit( sm, 0xF & ~( 1 << e ), (float)( 3 * kUnskew4 + 1 ), (float)( 3 * kUnskew4 ), fbPre );
        }
        // single-axis and pair vertices in the reference order
        auto single = [&]( int a ) {
            bool sm = Nrm[a] < 0.0f;
            int32_t s = signBitIf( sm );
            float o0 = fxor( (float)kUnskew4, s );
            float fbPre = fmadd( o0, coordSum, stemA ) + fxor( dB[a], s );
            emit( sm, 1 << a, (float)( kUnskew4 + 1 ), (float)kUnskew4, fbPre );
        };
        auto pair = [&]( int a, int b ) {
            bool sm = Nrm[a] < -Nrm[b];
            int32_t s = signBitIf( sm );
            float o0 = fxor( (float)( 2 * kUnskew4 ), s );
            float fbPre = fmadd( o0, coordSum, stemB ) + fxor( dB[a] + dB[b], s );
            emit( sm, ( 1 << a ) | ( 1 << b ), (float)( 2 * kUnskew4 + 1 ), (float)( 2 * kUnskew4 ), fbPre );
        };
        single( 0 ); pair( 0, 1 ); pair( 0, 2 ); pair( 0, 3 );
        single( 1 ); pair( 1, 2 ); pair( 1, 3 );
        single( 2 ); pair( 2, 3 );
        single( 3 );
    }

    template<bool G = true>
    FN2D float superSimplex2( int32_t seed, float x, float y )
    {
        float value = 0.0f;
        superWalk2<G>( seed, x, y, [&]( int32_t h, const float* d, float f, bool first ) {
            float g = gradDot2( h, d[0], d[1] );
            value = first ? f * g : fmadd( f, g, value );
        } );
        return value;
    }
    template<bool G = true>
    FN2D float superSimplex3( int32_t seed, float x, float y, float z )
    {
        float value = 0.0f;
        superWalk3<G>( seed, x, y, z, [&]( int32_t h, const float* d, float f, bool first ) {
            float g = gradDot3( h, d[0], d[1], d[2] );
            value = first ? f * g : fmadd( f, g, value );
        } );
        return value;
    }
    template<bool G = true>
    FN2D float superSimplex4( int32_t seed, float x, float y, float z, float w )
    {
        float value = 0.0f;
        superWalk4<G>( seed, x, y, z, w, [&]( int32_t h, const float* d, float f, bool first ) {
            float g = gradDot4( h, d[0], d[1], d[2], d[3] );
            value = first ? f * g : fmadd( f, g, value );
        } );
        return value;
    }

    // Bounding constants of ScaleOutput per leaf (Perlin.inl:32,71,122; Simplex.inl:60,139,283,374,574,933; Value: 2^31)
    template<int OP, int D> struct Bound;
}
