// Device cellular noises and domain warps for gfx950, one point per lane,
// following FastNoise2 v1.1.1 include/FastNoise/Generators/Cellular.inl,
// DomainWarp.inl and DomainWarpSimplex.inl operation by operation.
#pragma once
#include "fn2_device_lattice.h"

namespace fn2d
{
    // ---------------- cellular (Cellular.inl) ----------------
    template<int D> FN2D float jitterConst() { return D == 2 ? 0.437016f : D == 3 ? 0.396144f : 0.366025f; }

    template<int D> FN2D int32_t cellHash( int32_t seed, const int32_t* c )
    {
        if constexpr( D == 2 ) return hashHB2( seed, c[0], c[1] );
        else if constexpr( D == 3 ) return hashHB3( seed, c[0], c[1], c[2] );
        else return hashHB4( seed, c[0], c[1], c[2], c[3] );
    }
    // jittered cell-point direction from the hash (Cellular.inl:55-56, 141-143, 239-242)
    template<int D> FN2D void cellVector( int32_t h, float* v )
    {
        if constexpr( D == 2 )
        {
            v[0] = (float)( h & 0x7ff ) - 0x7ff / 2.0f;
            v[1] = (float)isrl( h, 21 ) - 0x7ff / 2.0f;
        }
        else if constexpr( D == 3 )
        {
            v[0] = (float)( h & 0x3ff ) - 0x3ff / 2.0f;
            v[1] = (float)( ( h >> 11 ) & 0x3ff ) - 0x3ff / 2.0f;
            v[2] = (float)isrl( h, 22 ) - 0x3ff / 2.0f;
        }
        else
        {
            v[0] = (float)( h & 0xff ) - 0xff / 2.0f;
            v[1] = (float)( ( h >> 8 ) & 0xff ) - 0xff / 2.0f;
            v[2] = (float)( ( h >> 16 ) & 0xff ) - 0xff / 2.0f;
            v[3] = (float)isrl( h, 24 ) - 0xff / 2.0f;
        }
    }
    template<int D> FN2D float cellMagSq( const float* v )
    {
        float s = v[D - 1] * v[D - 1];
#pragma unroll
        for( int a = D - 2; a >= 0; a-- ) s = fmadd( v[a], v[a], s );
        return s;
    }
    template<int D> FN2D float sizeJitterBits( int32_t h )
    {
        if constexpr( D == 2 ) return (float)( ( h >> 11 ) & 0x3ff );
        else if constexpr( D == 3 ) return (float)( h & 0xffff );
        else return (float)( h & 0xfffff );
    }

    struct CellParams
    {
        int32_t seed;          // seed + seed offset
        float jitter;          // kJitter * grid jitter
        bool sizeJitterOn;
        float sizeJitter;      // size jitter * -1/0x3ff..., only when sizeJitterOn
        int distFunc;
        float minkowskiP;
    };

    // 3^D neighbourhood with the reference nesting (x outermost). `OFFSET_RUNS`
    // keeps the running cell coordinate as an offset from the point
    // (CellularValue), otherwise as the absolute cell coordinate.
    template<int D, bool OFFSET_RUNS, typename CELL>
    FN2D void cellWalk( const CellParams& cp, const float* pos, CELL&& cell )
    {
        int32_t base[D];
        float basef[D];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            int32_t c = iadd( roundToInt( pos[a] ), -1 );
            basef[a] = (float)c;
            if constexpr( OFFSET_RUNS ) basef[a] = basef[a] - pos[a];
        }
        base[0] = imul( iadd( roundToInt( pos[0] ), -1 ), kPX );
        if( D > 1 ) base[1] = imul( iadd( roundToInt( pos[1] ), -1 ), kPY );
        if constexpr( D > 2 ) base[2] = imul( iadd( roundToInt( pos[2] ), -1 ), kPZ );
        if constexpr( D > 3 ) base[3] = imul( iadd( roundToInt( pos[3] ), -1 ), kPW );
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        int32_t hc[4];
        float fc[4];
        hc[0] = base[0];
        fc[0] = basef[0];
        for( int i0 = 0; i0 < 3; i0++ )
        {
            hc[1] = base[1];
            fc[1] = basef[1];
            for( int i1 = 0; i1 < 3; i1++ )
            {
                if constexpr( D > 2 ) { hc[2] = base[2]; fc[2] = basef[2]; }
#pragma unroll
                for( int i2 = 0; i2 < ( D > 2 ? 3 : 1 ); i2++ )
                {
                    if constexpr( D > 3 ) { hc[3] = base[3]; fc[3] = basef[3]; }
#pragma unroll
                    for( int i3 = 0; i3 < ( D > 3 ? 3 : 1 ); i3++ )
                    {
                        int32_t h = cellHash<D>( cp.seed, hc );
                        float v[D];
                        cellVector<D>( h, v );
                        float invMag = cp.jitter * invSqrtCell<D>( cellMagSq<D>( v ) );
                        cell( h, v, invMag, fc );
                        if constexpr( D > 3 ) { fc[3] = fc[3] + 1.0f; hc[3] = iadd( hc[3], prime[3] ); }
                    }
                    if constexpr( D > 2 ) { fc[2] = fc[2] + 1.0f; hc[2] = iadd( hc[2], prime[2] ); }
                }
                fc[1] = fc[1] + 1.0f; hc[1] = iadd( hc[1], prime[1] );
            }
            fc[0] = fc[0] + 1.0f; hc[0] = iadd( hc[0], prime[0] );
        }
    }

    // CellularValue (Cellular.inl:17-292): hash of the valueIndex-th closest cell
    template<int D>
    FN2D float cellularValue( const CellParams& cp, const float* pos, int valueIndex )
    {
        float dist[4] = { __builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff() };
        int32_t vh[4] = { 0, 0, 0, 0 };
        cellWalk<D, true>( cp, pos, [&]( int32_t h, const float* v, float invMag, const float* fc ) {
            float d[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) d[a] = fmadd( v[a], invMag, fc[a] );
            float nd = calcDistance<D>( cp.distFunc, D == 2, d, cp.minkowskiP );
            if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( h ), 1.0f );
            int32_t nh = h;
#pragma unroll
            for( int i = 0; i < 4; i++ )
            {
                if( i > valueIndex ) break;
                bool closer = nd < dist[i];
                float ld = dist[i];
                int32_t lh = vh[i];
                dist[i] = closer ? nd : dist[i];
                vh[i] = closer ? nh : vh[i];
                nd = closer ? ld : nd;
                nh = closer ? lh : nh;
            }
        } );
        int32_t sel = vh[0];
#pragma unroll
        for( int i = 1; i < 4; i++ ) sel = valueIndex == i ? vh[i] : sel;
        return (float)sel;
    }

    // The 3^D cells of the walk seen from each point's own side: per axis,
    // column 1 is the point's lattice cell, column 2 the neighbour on the
    // side the point lies (the near neighbour) and column 0 the other (far)
    // one. Cells come in classes of (near axes n, far axes f), ordered by
    // their typical squared distance for a point uniform in its cell, about
    // 0.56 n + 1.56 f, so that every lane reaches its far cells at the same
    // step (measured on MI355X against a centre-first order of the same
    // cells: 16 % faster on scattered positions, 3 % on grids). Entry k
    // holds the columns (0..2 per axis, 2 bits each) and the class in bits
    // 8 and up.
    template<int D>
    struct MirroredCellOrder
    {
        static constexpr int kCells = D == 2 ? 9 : D == 3 ? 27 : 81;
        uint16_t e[kCells];
        constexpr MirroredCellOrder() : e()
        {
            int pn[15] = {}, pf[15] = {}, np = 0;
            for( int f = 0; f <= D; f++ )
                for( int n = 0; n + f <= D; n++ )
                {
                    pn[np] = n;
                    pf[np] = f;
                    np++;
                }
            // insertion sort of the (n, f) classes by 9 n + 25 f, fewer far axes first on ties
            for( int i = 1; i < np; i++ )
                for( int j = i; j > 0; j-- )
                {
                    const int a = 9 * pn[j - 1] + 25 * pf[j - 1], b = 9 * pn[j] + 25 * pf[j];
                    if( a < b || ( a == b && pf[j - 1] <= pf[j] ) ) break;
                    const int tn = pn[j], tf = pf[j];
                    pn[j] = pn[j - 1];
                    pf[j] = pf[j - 1];
                    pn[j - 1] = tn;
                    pf[j - 1] = tf;
                }
            int m = 0;
            for( int cls = 0; cls < np; cls++ )
                for( int k = 0; k < kCells; k++ )
                {
                    int n = 0, f = 0, code = 0, t = k;
                    for( int a = 0; a < D; a++, t /= 3 )
                    {
                        n += t % 3 == 2;
                        f += t % 3 == 0;
                        code |= ( t % 3 ) << ( 2 * a );
                    }
                    if( n == pn[cls] && f == pf[cls] ) e[m++] = (uint16_t)( code | ( cls << 8 ) );
                }
        }
    };

    // Distance functions that sum squared axis distances (Euclidean and
    // Euclidean squared; unknown values take the same branch of calcDistance)
    FN2D bool squaredDistance( int func ) { return func < 2 || func > 5; }

    // CellularDistance for squared distances without size jitter, visiting the
    // cells near side first (MirroredCellOrder) and leaving out, for the whole
    // wave at once, every
    // cell that cannot change the kept slots. A cell's feature point lies
    // within |jitter| of its lattice point, so its distance is at least
    // |lattice point - position| - |jitter|; when that exceeds the largest kept
    // distance (with margins of 2^-10, far above the rounding of either side
    // and the 1 ulp of the hardware square root the bound takes) the
    // insertion network would leave every kept slot as it is. The kept
    // slots are the k smallest distances, whatever the order cells come in, so
    // the result is the one of the full walk.
    template<int D, int SLOTS>
    FN2D void cellularDistancesPruned( const CellParams& cp, const float* pos, float* dist )
    {
        using Order = MirroredCellOrder<D>;
        constexpr Order order;
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        int32_t hx[D][3];
        float rel[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t c = iadd( roundToInt( pos[a] ), -1 );
            // the walk's hash coordinates and cell positions, stepped the same way
            hx[a][0] = imul( c, prime[a] );
            hx[a][1] = iadd( hx[a][0], prime[a] );
            hx[a][2] = iadd( hx[a][1], prime[a] );
            float fc = (float)c;
            rel[a][0] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][1] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][2] = fc - pos[a];
            // a point below its lattice point has its near neighbour at -1:
            // columns 0 and 2 trade places (the same values, per lane)
            const bool below = rel[a][1] > 0.0f;
            const int32_t h0 = hx[a][0], h2 = hx[a][2];
            const float r0 = rel[a][0], r2 = rel[a][2];
            hx[a][0] = below ? h2 : h0;
            hx[a][2] = below ? h0 : h2;
            rel[a][0] = below ? r2 : r0;
            rel[a][2] = below ? r0 : r2;
        }
        uint32_t key[4] = { 0x7f800000u, 0x7f800000u, 0x7f800000u, 0x7f800000u };   // +inf
        const float jitterBound = fabs_( cp.jitter ) * ( 1.0f + 0x1p-16f );
        float limit = __builtin_inff();
#pragma unroll
        for( int k = 0; k < Order::kCells; k++ )
        {
            const int code = order.e[k], cls = code >> 8;
            int o[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) o[a] = ( code >> ( 2 * a ) ) & 3;
            // before SLOTS cells, slot SLOTS - 1 is still +inf and no cell can
            // be left out
            if( k >= SLOTS )
            {
                if( ( order.e[k - 1] >> 8 ) != cls )
                {
                    // squared bound beyond which a cell cannot reach the kept slots
                    const float r = sqrtBound( ibits( (int32_t)key[SLOTS - 1] ) * ( 1.0f + 0x1p-10f ) ) + jitterBound + 0x1p-10f;
                    limit = r * r;
                }
                float pp = rel[0][o[0]] * rel[0][o[0]];
#pragma unroll
                for( int a = 1; a < D; a++ ) pp = fmadd( rel[a][o[a]], rel[a][o[a]], pp );
                if( waveAll( pp > limit ) ) continue;
            }
            int32_t hc[D];
            float d[D], v[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) hc[a] = hx[a][o[a]];
            const int32_t h = cellHash<D>( cp.seed, hc );
            cellVector<D>( h, v );
            const float invMag = cp.jitter * invSqrtCell<D>( cellMagSq<D>( v ) );
#pragma unroll
            for( int a = 0; a < D; a++ ) d[a] = fmadd( v[a], invMag, rel[a][o[a]] );
            const float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
            // Squared distances are +0 or more (or NaN), and on such floats
            // the unsigned order of the bit patterns is the float order with NaN
            // above +inf: unsigned min / max give the results of the float
            // min / max steps above (a NaN distance never displaces a slot, and
            // no slot becomes NaN), without canonicalising their operands.
            const uint32_t nb = (uint32_t)fbits( nd );
            if constexpr( SLOTS > 3 ) key[3] = umax( umin( key[3], nb ), key[2] );
            if constexpr( SLOTS > 2 ) key[2] = umax( umin( key[2], nb ), key[1] );
            if constexpr( SLOTS > 1 ) key[1] = umax( umin( key[1], nb ), key[0] );
            key[0] = umin( key[0], nb );
        }
#pragma unroll
        for( int k = 0; k < 4; k++ ) dist[k] = ibits( (int32_t)key[k] );
    }

#if defined( __HIP_DEVICE_COMPILE__ )
    // Per-wave table of the table walk (cellularDistancesTable): for each
    // cell of a 5^D box, its jitter vector and 1 / |vector|, one region per
    // wave of a block of at most 256 threads.
    struct alignas( 16 ) CellEntry
    {
        float v[3];      // jitter vector (the third is 0 in 2D)
        float invMag;    // 1 / |vector| (each lane scales it by its own jitter)
    };
    template<int D>
    __device__ __forceinline__ CellEntry* cellTable()
    {
        static_assert( D == 2 || D == 3, "cell table: 2D or 3D" );
        __shared__ CellEntry waves[4][D == 2 ? 25 : 125];
        return waves[threadIdx.x >> 6];
    }

    // each cell's hash, beside the table (CellularValue)
    template<int D>
    __device__ __forceinline__ int32_t* cellHashTable()
    {
        __shared__ int32_t waves[4][D == 2 ? 25 : 125];
        return waves[threadIdx.x >> 6];
    }

    // The table walks' common part. c[a]: the lane's first walk cell on
    // axis a (the rounded position - 1); first[a]: the first active lane's.
    // When every active lane has c within one of first on every axis and
    // the same seed, fills the wave's table for the box of cells first - 1
    // .. first + 3 (table coordinate t = cell - first + 1) and returns true;
    // else returns false, having written nothing. HASH: also keep each
    // cell's hash (CellularValue returns it).
    template<int D, bool HASH>
    __device__ __forceinline__ bool fillCellTable( const CellParams& cp, const float* pos, int32_t* c, int32_t* first )
    {
        constexpr int kEntries = D == 2 ? 25 : 125;
        const int32_t prime[3] = { kPX, kPY, kPZ };
        bool near = cp.seed == __builtin_amdgcn_readfirstlane( cp.seed );
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            c[a] = iadd( roundToInt( pos[a] ), -1 );
            first[a] = __builtin_amdgcn_readfirstlane( c[a] );
            // c - first in {-1, 0, 1}, in wrapping arithmetic
            near = near && (uint32_t)c[a] - (uint32_t)first[a] + 1u <= 2u;
        }
        if( !waveAll( near ) ) return false;

        CellEntry* tbl = cellTable<D>();
        const uint64_t active = __ballot( 1 );
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi( (uint32_t)( active >> 32 ), __builtin_amdgcn_mbcnt_lo( (uint32_t)active, 0u ) );
        const uint32_t lanes = (uint32_t)__builtin_popcountll( active );
        // the previous use of the table by this wave is over
        __builtin_amdgcn_fence( __ATOMIC_ACQ_REL, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        for( uint32_t e = rank; e < (uint32_t)kEntries; e += lanes )
        {
            int32_t hc[D];
            uint32_t rest = e;
#pragma unroll
            for( int a = 0; a < D; a++ )
            {
                const int32_t t = (int32_t)( rest % 5u );
                rest /= 5u;
                hc[a] = imul( iadd( first[a], t - 1 ), prime[a] );
            }
            const int32_t h = cellHash<D>( cp.seed, hc );
            float v[D];
            cellVector<D>( h, v );
            CellEntry entry;
            entry.v[0] = v[0];
            entry.v[1] = v[1];
            entry.v[2] = D > 2 ? v[D - 1] : 0.0f;
            entry.invMag = invSqrtCell<D>( cellMagSq<D>( v ) );
            tbl[e] = entry;
            if constexpr( HASH ) cellHashTable<D>()[e] = h;
        }
        __builtin_amdgcn_fence( __ATOMIC_ACQ_REL, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        return true;
    }

    // CellularDistance, squared distances, no size jitter, 2D/3D: the pruned
    // walk with the per-cell work shared by the wave. When every active
    // lane's walk starts within one cell of the first active lane's on every
    // axis, the 3^D cells of all lanes lie in a box of 5^D cells, and the
    // lanes fill the wave's table with each box cell's hash-derived jitter
    // vector and inverse length (the reference's per-cell work,
    // Cellular.inl:391-409, computed once per wave instead of once per lane
    // and cell). Each lane then walks its own cells as the pruned walk does,
    // reading the vector and the inverse length from the table and scaling
    // the length by its own jitter (a node may set it per point); its
    // distance is the same fused multiply-add of the same operands, so the
    // slots end bit for bit as the pruned walk's. Returns false, having done
    // nothing, for a wave whose lanes lie further apart (scattered positions)
    // or differ in seed.
    template<int D, int SLOTS>
    __device__ __forceinline__ bool cellularDistancesTable( const CellParams& cp, const float* pos, float* dist )
    {
        using Order = MirroredCellOrder<D>;
        constexpr Order order;
        int32_t c[D], first[D];
        if( !fillCellTable<D, false>( cp, pos, c, first ) ) return false;
        const CellEntry* tbl = cellTable<D>();

        // the lane's mirrored columns: table offsets and cell positions
        int32_t T[D][3];
        float rel[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t scale = a == 0 ? 1 : a == 1 ? 5 : 25;
            const int32_t t0 = (int32_t)( (uint32_t)c[a] - (uint32_t)first[a] + 1u );   // 0..2
            T[a][0] = t0 * scale;
            T[a][1] = T[a][0] + scale;
            T[a][2] = T[a][1] + scale;
            float fc = (float)c[a];
            rel[a][0] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][1] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][2] = fc - pos[a];
            const bool below = rel[a][1] > 0.0f;
            const int32_t i0 = T[a][0], i2 = T[a][2];
            const float r0 = rel[a][0], r2 = rel[a][2];
            T[a][0] = below ? i2 : i0;
            T[a][2] = below ? i0 : i2;
            rel[a][0] = below ? r2 : r0;
            rel[a][2] = below ? r0 : r2;
        }
        uint32_t key[4] = { 0x7f800000u, 0x7f800000u, 0x7f800000u, 0x7f800000u };   // +inf
        const float jitterBound = fabs_( cp.jitter ) * ( 1.0f + 0x1p-16f );
        float limit = __builtin_inff();
#pragma unroll
        for( int k = 0; k < Order::kCells; k++ )
        {
            const int code = order.e[k], cls = code >> 8;
            int o[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) o[a] = ( code >> ( 2 * a ) ) & 3;
            if( k >= SLOTS )
            {
                if( ( order.e[k - 1] >> 8 ) != cls )
                {
                    const float r = sqrtBound( ibits( (int32_t)key[SLOTS - 1] ) * ( 1.0f + 0x1p-10f ) ) + jitterBound + 0x1p-10f;
                    limit = r * r;
                }
                float pp = rel[0][o[0]] * rel[0][o[0]];
#pragma unroll
                for( int a = 1; a < D; a++ ) pp = fmadd( rel[a][o[a]], rel[a][o[a]], pp );
                if( waveAll( pp > limit ) ) continue;
            }
            int32_t idx = T[0][o[0]];
#pragma unroll
            for( int a = 1; a < D; a++ ) idx += T[a][o[a]];
            const CellEntry cell = tbl[idx];
            const float invMag = cp.jitter * cell.invMag;
            float d[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) d[a] = fmadd( cell.v[a], invMag, rel[a][o[a]] );
            const float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
            const uint32_t nb = (uint32_t)fbits( nd );
            if constexpr( SLOTS > 3 ) key[3] = umax( umin( key[3], nb ), key[2] );
            if constexpr( SLOTS > 2 ) key[2] = umax( umin( key[2], nb ), key[1] );
            if constexpr( SLOTS > 1 ) key[1] = umax( umin( key[1], nb ), key[0] );
            key[0] = umin( key[0], nb );
        }
#pragma unroll
        for( int k = 0; k < 4; k++ ) dist[k] = ibits( (int32_t)key[k] );
        return true;
    }

    // CellularDistance where the pruned walk does not apply (distance
    // functions other than the squared ones, size jitter) on the wave's cell
    // table: the full walk of cellularDistances, same order and the same
    // float min / max steps, with each cell's hash, jitter vector and inverse
    // length from the table. Returns false, having done nothing, where
    // fillCellTable does.
    template<int D, int SLOTS>
    __device__ __forceinline__ bool cellularDistancesTableAll( const CellParams& cp, const float* pos, float* dist )
    {
        int32_t c[D], first[D];
        if( !fillCellTable<D, true>( cp, pos, c, first ) ) return false;
        const CellEntry* tbl = cellTable<D>();
        const int32_t* htbl = cellHashTable<D>();
        int32_t T[D];
        float rel[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t scale = a == 0 ? 1 : a == 1 ? 5 : 25;
            T[a] = (int32_t)( (uint32_t)c[a] - (uint32_t)first[a] + 1u ) * scale;   // the lane's first cell
            float fc = (float)c[a];
            rel[a][0] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][1] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][2] = fc - pos[a];
        }
        dist[0] = dist[1] = dist[2] = dist[3] = __builtin_inff();
#pragma unroll
        for( int i0 = 0; i0 < 3; i0++ )
#pragma unroll
            for( int i1 = 0; i1 < 3; i1++ )
#pragma unroll
                for( int i2 = 0; i2 < ( D > 2 ? 3 : 1 ); i2++ )
                {
                    int32_t idx = T[0] + i0 + T[1] + 5 * i1;
                    if constexpr( D > 2 ) idx += T[D - 1] + 25 * i2;
                    const CellEntry cell = tbl[idx];
                    const float invMag = cp.jitter * cell.invMag;
                    const int ii[3] = { i0, i1, i2 };
                    float d[D];
#pragma unroll
                    for( int a = 0; a < D; a++ ) d[a] = fmadd( cell.v[a], invMag, rel[a][ii[a]] );
                    float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
                    if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( htbl[idx] ), 1.0f );
                    if constexpr( SLOTS > 3 ) dist[3] = fmaxf_( fminf_( dist[3], nd ), dist[2] );
                    if constexpr( SLOTS > 2 ) dist[2] = fmaxf_( fminf_( dist[2], nd ), dist[1] );
                    if constexpr( SLOTS > 1 ) dist[1] = fmaxf_( fminf_( dist[1], nd ), dist[0] );
                    dist[0] = fminf_( dist[0], nd );
                }
        return true;
    }

    // CellularLookup's closest cell (Cellular.inl:568-807) on the wave's
    // cell table: the walk of cellularLookupPos, same order (the first of
    // equal distances stays) and steps, with each cell's hash, jitter vector
    // and inverse length from the table. Returns false, having done nothing,
    // where fillCellTable does.
    template<int D>
    __device__ __forceinline__ bool cellularLookupTable( const CellParams& cp, const float* pos, float* cellPos )
    {
        int32_t c[D], first[D];
        if( !fillCellTable<D, true>( cp, pos, c, first ) ) return false;
        const CellEntry* tbl = cellTable<D>();
        const int32_t* htbl = cellHashTable<D>();
        int32_t T[D];
        float F[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t scale = a == 0 ? 1 : a == 1 ? 5 : 25;
            T[a] = (int32_t)( (uint32_t)c[a] - (uint32_t)first[a] + 1u ) * scale;   // the lane's first cell
            F[a][0] = (float)c[a];
            F[a][1] = F[a][0] + 1.0f;
            F[a][2] = F[a][1] + 1.0f;
        }
        float distance = 3.40282347e+38f;
#pragma unroll
        for( int a = 0; a < D; a++ ) cellPos[a] = 0.0f;
#pragma unroll
        for( int i0 = 0; i0 < 3; i0++ )
#pragma unroll
            for( int i1 = 0; i1 < 3; i1++ )
#pragma unroll
                for( int i2 = 0; i2 < ( D > 2 ? 3 : 1 ); i2++ )
                {
                    int32_t idx = T[0] + i0 + T[1] + 5 * i1;
                    if constexpr( D > 2 ) idx += T[D - 1] + 25 * i2;
                    const CellEntry cell = tbl[idx];
                    const float invMag = cp.jitter * cell.invMag;
                    const int ii[3] = { i0, i1, i2 };
                    float local[D], d[D];
#pragma unroll
                    for( int a = 0; a < D; a++ ) { local[a] = fmadd( cell.v[a], invMag, F[a][ii[a]] ); d[a] = local[a] - pos[a]; }
                    float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
                    if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( htbl[idx] ), 1.0f );
                    bool closer = nd < distance;
                    distance = fminf_( nd, distance );
#pragma unroll
                    for( int a = 0; a < D; a++ ) cellPos[a] = closer ? local[a] : cellPos[a];
                }
        return true;
    }

    // CellularValue (Cellular.inl:17-292) on the wave's cell table: the
    // reference's walk, x outermost, over the same cells with the same
    // running offsets, since the value of tied distances depends on the
    // order (the first of equal distances stays); only each cell's hash,
    // jitter vector and inverse length come from the table. Returns false,
    // having done nothing, where fillCellTable does.
    template<int D>
    __device__ __forceinline__ bool cellularValueTable( const CellParams& cp, const float* pos, int valueIndex, float& result )
    {
        int32_t c[D], first[D];
        if( !fillCellTable<D, true>( cp, pos, c, first ) ) return false;
        const CellEntry* tbl = cellTable<D>();
        const int32_t* htbl = cellHashTable<D>();
        int32_t T[D];
        float F[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t scale = a == 0 ? 1 : a == 1 ? 5 : 25;
            T[a] = (int32_t)( (uint32_t)c[a] - (uint32_t)first[a] + 1u ) * scale;   // the lane's first cell
            F[a][0] = (float)c[a] - pos[a];
            F[a][1] = F[a][0] + 1.0f;
            F[a][2] = F[a][1] + 1.0f;
        }
        float dist[4] = { __builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff() };
        int32_t vh[4] = { 0, 0, 0, 0 };
#pragma unroll
        for( int i0 = 0; i0 < 3; i0++ )
#pragma unroll
            for( int i1 = 0; i1 < 3; i1++ )
#pragma unroll
                for( int i2 = 0; i2 < ( D > 2 ? 3 : 1 ); i2++ )
                {
                    int32_t idx = T[0] + i0 + T[1] + 5 * i1;
                    if constexpr( D > 2 ) idx += T[D - 1] + 25 * i2;
                    const CellEntry cell = tbl[idx];
                    const int32_t h = htbl[idx];
                    const float invMag = cp.jitter * cell.invMag;
                    const int ii[3] = { i0, i1, i2 };
                    float d[D];
#pragma unroll
                    for( int a = 0; a < D; a++ ) d[a] = fmadd( cell.v[a], invMag, F[a][ii[a]] );
                    float nd = calcDistance<D>( cp.distFunc, D == 2, d, cp.minkowskiP );
                    if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( h ), 1.0f );
                    int32_t nh = h;
#pragma unroll
                    for( int i = 0; i < 4; i++ )
                    {
                        if( i > valueIndex ) break;
                        bool closer = nd < dist[i];
                        float ld = dist[i];
                        int32_t lh = vh[i];
                        dist[i] = closer ? nd : dist[i];
                        vh[i] = closer ? nh : vh[i];
                        nd = closer ? ld : nd;
                        nh = closer ? lh : nh;
                    }
                }
        int32_t sel = vh[0];
#pragma unroll
        for( int i = 1; i < 4; i++ ) sel = valueIndex == i ? vh[i] : sel;
        result = (float)sel;
        return true;
    }
#endif

#ifndef FN2_SHARED_CELL_WALK
#define FN2_SHARED_CELL_WALK 0
#endif
#if defined( __HIP_DEVICE_COMPILE__ )
    // Per-wave scratch of the shared far-cell walk (cellularDistancesShared),
    // one region per wave of a block of at most 256 threads: each lane's
    // mirrored cell columns (3 x D offsets, 3 x D hash coordinates), its kept
    // slots, and a queue of (lane, cell) items.
    constexpr int kShareQueue = 192;
    constexpr int kShareLaneWords = 3 * 3 * 2 + 2;
    struct CellShare
    {
        uint32_t lane[64 * kShareLaneWords];
        uint32_t queue[kShareQueue];
    };
    __device__ __forceinline__ CellShare& cellShare()
    {
        __shared__ CellShare waves[4];
        return waves[threadIdx.x >> 6];
    }

    // The distance of the cell in mirrored columns o[] of a point whose
    // columns are rel / hx ([axis * 3 + column]) (the walk body of the
    // reference, Cellular.inl:391-409)
    template<int D, typename REL, typename HX>
    __device__ __forceinline__ uint32_t shareCellKey( const CellParams& cp, const int* o, REL rel, HX hx )
    {
        int32_t hc[D];
        float d[D], v[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) hc[a] = hx( a, o[a] );
        const int32_t h = cellHash<D>( cp.seed, hc );
        cellVector<D>( h, v );
        const float invMag = cp.jitter * invSqrtCell<D>( cellMagSq<D>( v ) );
#pragma unroll
        for( int a = 0; a < D; a++ ) d[a] = fmadd( v[a], invMag, rel( a, o[a] ) );
        return (uint32_t)fbits( calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP ) );
    }

    // CellularDistance, squared distances, no size jitter, 2D/3D, one or two
    // kept slots: the cells of the point's own octant (columns 1 and 2 of
    // every axis, the near side) first, for every lane; then only the far
    // cells each lane still needs, under the same bound as the pruned walk
    // (a far cell whose lattice point lies beyond the kept slot's distance
    // plus the jitter cannot change the slots). Where the lanes of a wave lie
    // close together (grid bricks), the union of their needs is small and the
    // wave walks it like the pruned walk does. Where they are scattered
    // (position arrays), the union covers most far cells while each lane
    // needs few; the lanes then queue their (lane, cell) items in LDS, every
    // lane of the wave takes queued items in turn, and merges each distance
    // into its owner's slots with LDS atomic minimums: slot 0 takes the
    // minimum and slot 1 the minimum of what slot 0 gave up, which keeps the
    // two smallest of all distances offered in any order. The slots end as
    // those of the full walk, bit for bit.
    // Only the kernel variant that position-array requests use compiles it
    // in (FN2_SHARED_CELL_WALK, 3D): on MI355X it took the warped cellular
    // tree over 16.8 M scattered positions from 0.334 to 0.311 ms, but the
    // same tree on a 256^3 grid from 0.247 to 0.265 ms, where bricks already
    // keep the wave's union small and its 23 KB of LDS per block and the
    // extra scalar work cost occupancy and issue slots.
    template<int D, int SLOTS>
    __device__ __forceinline__ void cellularDistancesShared( const CellParams& cp, const float* pos, float* dist )
    {
        static_assert( D >= 2 && D <= 3 && SLOTS >= 1 && SLOTS <= 2, "shared walk: 2D/3D, one or two slots" );
        using Order = MirroredCellOrder<D>;
        constexpr Order order;
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        int32_t hx[D][3];
        float rel[D][3];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            const int32_t c = iadd( roundToInt( pos[a] ), -1 );
            hx[a][0] = imul( c, prime[a] );
            hx[a][1] = iadd( hx[a][0], prime[a] );
            hx[a][2] = iadd( hx[a][1], prime[a] );
            float fc = (float)c;
            rel[a][0] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][1] = fc - pos[a];
            fc = fc + 1.0f;
            rel[a][2] = fc - pos[a];
            // column 2 is the near neighbour, column 0 the far one
            const bool below = rel[a][1] > 0.0f;
            const int32_t h0 = hx[a][0], h2 = hx[a][2];
            const float r0 = rel[a][0], r2 = rel[a][2];
            hx[a][0] = below ? h2 : h0;
            hx[a][2] = below ? h0 : h2;
            rel[a][0] = below ? r2 : r0;
            rel[a][2] = below ? r0 : r2;
        }
        auto relR = [&]( int a, int col ) { return rel[a][col]; };
        auto hxR = [&]( int a, int col ) { return hx[a][col]; };
        uint32_t key[2] = { 0x7f800000u, 0x7f800000u };   // +inf
        auto insert = [&]( uint32_t nb ) {
            if constexpr( SLOTS > 1 ) key[1] = umax( umin( key[1], nb ), key[0] );
            key[0] = umin( key[0], nb );
        };
        // the near octant, for every lane
#pragma unroll
        for( int k = 0; k < Order::kCells; k++ )
        {
            const int code = order.e[k] & 0xff;
            bool nearCell = true;
            int o[D];
#pragma unroll
            for( int a = 0; a < D; a++ )
            {
                o[a] = ( code >> ( 2 * a ) ) & 3;
                nearCell = nearCell && o[a] != 0;
            }
            if( nearCell ) insert( shareCellKey<D>( cp, o, relR, hxR ) );
        }
        // the far cells this lane can still need, as a mask over cell codes
        const float jitterBound = fabs_( cp.jitter ) * ( 1.0f + 0x1p-16f );
        const float r = sqrtBound( ibits( (int32_t)key[SLOTS - 1] ) * ( 1.0f + 0x1p-10f ) ) + jitterBound + 0x1p-10f;
        const float limit = r * r;
        // `need` per lane, `any` their union over the wave's active lanes
        // (wave-uniform), both as masks over cell codes
        uint64_t need = 0, any = 0;
#pragma unroll
        for( int k = 0; k < Order::kCells; k++ )
        {
            const int code = order.e[k] & 0xff;
            bool nearCell = true;
            float pp = 0.0f;
#pragma unroll
            for( int a = 0; a < D; a++ )
            {
                const int col = ( code >> ( 2 * a ) ) & 3;
                nearCell = nearCell && col != 0;
                pp = a == 0 ? rel[a][col] * rel[a][col] : fmadd( rel[a][col], rel[a][col], pp );
            }
            if( nearCell ) continue;
            const bool mine = !( pp > limit );
            need |= mine ? 1ull << code : 0ull;
            any |= __ballot( mine ) ? 1ull << code : 0ull;
        }
        if( __builtin_popcountll( any ) <= 5 )
        {
            // few far cells for the whole wave: walk them together
#pragma unroll
            for( int k = 0; k < Order::kCells; k++ )
            {
                const int code = order.e[k] & 0xff;
                if( !( any & ( 1ull << code ) ) ) continue;   // wave-uniform
                int o[D];
#pragma unroll
                for( int a = 0; a < D; a++ ) o[a] = ( code >> ( 2 * a ) ) & 3;
                insert( shareCellKey<D>( cp, o, relR, hxR ) );
            }
        }
        else
        {
            CellShare& sh = cellShare();
            const uint32_t lane = __lane_id();
            uint32_t* mine = sh.lane + lane * kShareLaneWords;
#pragma unroll
            for( int a = 0; a < D; a++ )
#pragma unroll
                for( int col = 0; col < 3; col++ )
                {
                    mine[a * 3 + col] = (uint32_t)fbits( rel[a][col] );
                    mine[9 + a * 3 + col] = (uint32_t)hx[a][col];
                }
            mine[18] = key[0];
            mine[19] = key[1];
            const uint64_t active = __ballot( 1 );
            const uint32_t nActive = (uint32_t)__builtin_popcountll( active );
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi( (uint32_t)( active >> 32 ), __builtin_amdgcn_mbcnt_lo( (uint32_t)active, 0u ) );
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
            __builtin_amdgcn_wave_barrier();
            // every active lane takes queued items in turn and merges each
            // distance into its owner's slots
            auto drain = [&]( uint32_t count ) {
                __builtin_amdgcn_fence( __ATOMIC_ACQ_REL, "wavefront" );
                __builtin_amdgcn_wave_barrier();
                for( uint32_t q = rank; q < count; q += nActive )
                {
                    const uint32_t item = sh.queue[q];
                    const uint32_t* owner = sh.lane + ( item >> 8 ) * kShareLaneWords;
                    int o[D];
#pragma unroll
                    for( int a = 0; a < D; a++ ) o[a] = ( item >> ( 2 * a ) ) & 3;
                    auto orel = [&]( int a, int col ) { return ibits( (int32_t)owner[a * 3 + col] ); };
                    auto ohx = [&]( int a, int col ) { return (int32_t)owner[9 + a * 3 + col]; };
                    const uint32_t nb = shareCellKey<D>( cp, o, orel, ohx );
                    uint32_t* slots = sh.lane + ( item >> 8 ) * kShareLaneWords + 18;
                    const uint32_t gave = atomicMin( &slots[0], nb );
                    if constexpr( SLOTS > 1 ) atomicMin( &slots[1], umax( gave, nb ) );
                }
                __builtin_amdgcn_fence( __ATOMIC_ACQ_REL, "wavefront" );
                __builtin_amdgcn_wave_barrier();
            };
            uint32_t queued = 0;   // wave-uniform
            for( ;; )
            {
                const bool has = need != 0;
                const uint64_t b = __ballot( has );
                if( !b ) break;
                const uint32_t n = (uint32_t)__builtin_popcountll( b );
                if( queued + n > (uint32_t)kShareQueue )
                {
                    drain( queued );
                    queued = 0;
                }
                if( has )
                {
                    const uint32_t code = (uint32_t)__builtin_ctzll( need );
                    need &= need - 1;
                    const uint32_t at = queued + __builtin_amdgcn_mbcnt_hi( (uint32_t)( b >> 32 ), __builtin_amdgcn_mbcnt_lo( (uint32_t)b, 0u ) );
                    sh.queue[at] = ( lane << 8 ) | code;
                }
                queued += n;
            }
            drain( queued );
            key[0] = mine[18];
            key[1] = mine[19];
            __builtin_amdgcn_wave_barrier();
        }
        dist[0] = ibits( (int32_t)key[0] );
        dist[1] = SLOTS > 1 ? ibits( (int32_t)key[1] ) : __builtin_inff();
        dist[2] = dist[3] = __builtin_inff();
    }
#endif

    // CellularDistance (Cellular.inl:294-529): the four smallest distances.
    // Slot k of the insertion network depends only on slots below k, so a
    // caller that reads slots < SLOTS only (distance indices known when the
    // kernel is compiled) may keep just those: they hold the same values.
    // PRUNED: take the pruned walk where it applies (specialised kernels; the
    // interpreter keeps the one walk, whose registers it can afford)
    // TABLE: try the table walk first (specialised kernels: its table is
    // static LDS, which the interpreter's register file would lose)
    template<int D, int SLOTS = 4, bool PRUNED = false, bool TABLE = false>
    FN2D void cellularDistances( const CellParams& cp, const float* pos, float* dist )
    {
        if constexpr( PRUNED )
        {
            if( squaredDistance( cp.distFunc ) && !cp.sizeJitterOn )
            {
#if defined( __HIP_DEVICE_COMPILE__ )
                if constexpr( TABLE && ( D == 2 || D == 3 ) )
                    if( cellularDistancesTable<D, SLOTS>( cp, pos, dist ) ) return;
#endif
#if defined( __HIP_DEVICE_COMPILE__ ) && FN2_SHARED_CELL_WALK
                if constexpr( D == 3 && SLOTS <= 2 )
                {
                    cellularDistancesShared<D, SLOTS>( cp, pos, dist );
                    return;
                }
#endif
                cellularDistancesPruned<D, SLOTS>( cp, pos, dist );
                return;
            }
        }
#if defined( __HIP_DEVICE_COMPILE__ )
        if constexpr( TABLE && ( D == 2 || D == 3 ) )
            if( cellularDistancesTableAll<D, SLOTS>( cp, pos, dist ) ) return;
#endif
        dist[0] = dist[1] = dist[2] = dist[3] = __builtin_inff();
        cellWalk<D, false>( cp, pos, [&]( int32_t h, const float* v, float invMag, const float* fc ) {
            float d[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) d[a] = fmadd( v[a], invMag, fc[a] - pos[a] );
            float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
            if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( h ), 1.0f );
            if constexpr( SLOTS > 3 ) dist[3] = fmaxf_( fminf_( dist[3], nd ), dist[2] );
            if constexpr( SLOTS > 2 ) dist[2] = fmaxf_( fminf_( dist[2], nd ), dist[1] );
            if constexpr( SLOTS > 1 ) dist[1] = fmaxf_( fminf_( dist[1], nd ), dist[0] );
            dist[0] = fminf_( dist[0], nd );
        } );
    }
    // GetReturn (Cellular.inl:531-565)
    FN2D float cellularReturn( float* dist, int distFunc, int i0, int i1, int returnType, float maxDist )
    {
        float a = dist[0], b = dist[0];
#pragma unroll
        for( int i = 1; i < 4; i++ ) { a = i0 == i ? dist[i] : a; b = i1 == i ? dist[i] : b; }
        if( distFunc == 0 )
        {
            a *= invSqrt( a );
            // the same slot twice takes the root twice, as the reference does
            b = i0 == i1 ? a * invSqrt( a ) : b * invSqrt( b );
            if( i0 == i1 ) a = b;
        }
        (void)maxDist;
        switch( returnType )
        {
        case 1: return a + b;
        case 2: return fabs_( a - b );
        case 3: return a * b;
        case 4: return a * recip( b );
        default: return a;
        }
    }

    // CellularLookup (Cellular.inl:568-807): position of the closest cell
    template<int D>
    FN2D void cellularLookupPos( const CellParams& cp, const float* pos, float* cellPos )
    {
        float distance = 3.40282347e+38f;
#pragma unroll
        for( int a = 0; a < D; a++ ) cellPos[a] = 0.0f;
        cellWalk<D, false>( cp, pos, [&]( int32_t h, const float* v, float invMag, const float* fc ) {
            float local[D], d[D];
#pragma unroll
            for( int a = 0; a < D; a++ ) { local[a] = fmadd( v[a], invMag, fc[a] ); d[a] = local[a] - pos[a]; }
            float nd = calcDistance<D>( cp.distFunc, false, d, cp.minkowskiP );
            if( cp.sizeJitterOn ) nd *= fnmadd( cp.sizeJitter, sizeJitterBits<D>( h ), 1.0f );
            bool closer = nd < distance;
            distance = fminf_( nd, distance );
#pragma unroll
            for( int a = 0; a < D; a++ ) cellPos[a] = closer ? local[a] : cellPos[a];
        } );
    }

    // ---------------- DomainWarpGradient::Warp (DomainWarp.inl:33-206) ----------------
    template<int D, bool UNIFORM>
    FN2D float warpGradientCorners( int32_t seed, float amp, const float* axisScale, const int32_t* c0, const float* t0, const float* t1, float* out )
    {
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        const float normalise = D == 2 ? 1.0f / ( 0xffff / 2.0f ) : D == 3 ? 1.0f / ( 0x3ff / 2.0f ) : 1.0f / ( 0xff / 2.0f );
        float warp[D];
#pragma unroll
        for( int corner = 0; corner < ( 1 << D ); corner++ )
        {
            int32_t pc[D];
            float contrib = ( corner & 1 ) ? t1[0] : t0[0];
#pragma unroll
            for( int a = 0; a < D; a++ ) pc[a] = ( corner >> a ) & 1 ? iadd( c0[a], prime[a] ) : c0[a];
#pragma unroll
            for( int a = 1; a < D; a++ ) contrib = contrib * ( ( corner >> a ) & 1 ? t1[a] : t0[a] );
            int32_t h;
            if constexpr( UNIFORM )
            {
                // the same hash as cellHash with plain xors, which the
                // compiler keeps on the scalar unit for wave-uniform operands
                int32_t x = seed;
#pragma unroll
                for( int a = 0; a < D; a++ ) x ^= pc[a];
                h = imul( x, kHashMul );
            }
            else h = cellHash<D>( seed, pc );
            float comp[D];
            if constexpr( D == 2 )
            {
                comp[0] = (float)( h & 0xffff );
                comp[1] = (float)isrl( h, 16 );
            }
            else if constexpr( D == 3 )
            {
                comp[0] = (float)( h & 0x3ff );
                comp[1] = (float)( ( h >> 11 ) & 0x3ff );
                comp[2] = (float)isrl( h, 22 );
            }
            else
            {
                comp[0] = (float)( h & 0xff );
                comp[1] = (float)( ( h >> 8 ) & 0xff );
                comp[2] = (float)( ( h >> 16 ) & 0xff );
                comp[3] = (float)isrl( h, 24 );
            }
#pragma unroll
            for( int a = 0; a < D; a++ ) warp[a] = corner == 0 ? contrib * comp[a] : fmadd( contrib, comp[a], warp[a] );
        }
#pragma unroll
        for( int a = 0; a < D; a++ ) warp[a] = fmsub( warp[a], normalise, 1.0f );
#pragma unroll
        for( int a = 0; a < D; a++ ) out[a] = fmadd( warp[a], amp * axisScale[a], out[a] );
        float r = warp[D - 1] * warp[D - 1];
#pragma unroll
        for( int a = D - 2; a >= 0; a-- ) r = fmadd( warp[a], warp[a], r );
        return r;
    }

    // UNIFORM_CELL: when every active lane of a wave lies in the same lattice
    // cell (grid bricks under a low-frequency warp), the corner hashes and
    // their bit fields are the same on every lane: the wave takes the corner
    // loop on the first lane's cell, whose integer work the compiler moves to
    // the scalar unit; only the conversions and each lane's weighted sums
    // stay vector work. The values are those of the per-lane loop. Measured
    // on MI355X: the warped cellular 256^3 grid 0.2088 -> 0.1938 ms;
    // scattered positions, which never pass the check, were 1.3 % slower
    // with it, so the position-array variant and the interpreter leave it
    // out.
    template<int D, bool UNIFORM_CELL = false>
    FN2D float warpGradient( int32_t seed, float amp, const float* axisScale, const float* sp, float* out )
    {
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        int32_t cell[D];
        float t1[D], t0[D];
#pragma unroll
        for( int a = 0; a < D; a++ )
        {
            float fl = ffloor( sp[a] );
            cell[a] = floorToInt( fl );
            t1[a] = interpHermite( sp[a] - fl );
        }
#pragma unroll
        for( int a = 0; a < D; a++ ) t0[a] = 1.0f - t1[a];
#if defined( __HIP_DEVICE_COMPILE__ )
        if constexpr( UNIFORM_CELL )
        {
            bool same = seed == __builtin_amdgcn_readfirstlane( seed );
#pragma unroll
            for( int a = 0; a < D; a++ ) same = same && cell[a] == __builtin_amdgcn_readfirstlane( cell[a] );
            if( waveAll( same ) )
            {
                int32_t c0[D];
#pragma unroll
                for( int a = 0; a < D; a++ ) c0[a] = imul( __builtin_amdgcn_readfirstlane( cell[a] ), prime[a] );
                return warpGradientCorners<D, true>( __builtin_amdgcn_readfirstlane( seed ), amp, axisScale, c0, t0, t1, out );
            }
        }
#endif
        int32_t c0[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) c0[a] = imul( cell[a], prime[a] );
        return warpGradientCorners<D, false>( seed, amp, axisScale, c0, t0, t1, out );
    }

    // ---------------- simplex warp vector contributions (DomainWarpSimplex.inl:6-724) ----------------
    // Orthogonal gradient matrix, 2D (:193-235, generic branch)
    FN2D void orthoVec2( int32_t h31, const float* f, float m, float* v )
    {
        int32_t index = imul( isrl( h31, 1 ), 12 >> 2 );
        int32_t ofThree = isrl( index, 30 );
        int32_t signX = ishl( index >> 28, 31 ), signY = ishl( index >> 29, 31 );
        float masked = index >= 0 ? kSkew2f : 0.0f;
        float gX = fxor( ofThree != 0 ? masked + 1.0f : masked, signX );
        float gY = fxor( ofThree != 1 ? masked + 1.0f : masked, signY );
        v[0] = fmadd( m, fmadd( f[1], gY, f[0] * gX ), v[0] );
        m = fxor( m, ishl( h31, 31 ) );
        v[1] = fmadd( m, fmsub( f[1], gX, f[0] * gY ), v[1] );
    }
    // Orthogonal gradient matrix, 3D (:237-316, generic branch)
    FN2D void orthoVec3( int32_t h31, const float* f, float m, float* v )
    {
        const float kA = 2.224744871391589f, kB = -0.224744871391589f, kC = -1.0f, kDE = 1.0f, kF = 2.0f;
        int32_t hs = isrl( h31, 1 );
        int32_t ifb = imul( hs, (int32_t)0xAAAAAAABu );
        int32_t ip2hb = ishl( ifb, 29 );
        int32_t ip3 = isrl( imul( hs, 3 ), 30 );
        int32_t s0 = ishl( ifb, 31 );
        int32_t s1 = ishl( ifb, 30 ) & (int32_t)0x80000000u;
        bool notYZ = ifb < 0, notXY = ishl( ifb, 1 ) < 0;
        float aX = notYZ ? kC : fxor( kA, s0 );
        float aY = ( notYZ || notXY ) ? ( notXY ? fxor( kB, s1 ) : fxor( kA, s0 ) ) : kC;
        float aZ = notXY ? kC : fxor( kB, s1 );
        float valueA = fmadd( aZ, f[2], fmadd( f[1], aY, f[0] * aX ) );
        float bX = notYZ ? kC : fxor( kB, s0 );
        float bY = ( notYZ || notXY ) ? ( notXY ? fxor( kA, s1 ) : fxor( kB, s0 ) ) : kC;
        float bZ = notXY ? kC : fxor( kA, s1 );
        float valueB = fmadd( bZ, f[2], fmadd( f[1], bY, f[0] * bX ) );
        float cX = notYZ ? kF : fxor( kDE, s0 );
        float cY = ( notYZ || notXY ) ? ( notXY ? fxor( kDE, s1 ) : fxor( kDE, s0 ) ) : kF;
        float cZ = notXY ? kF : fxor( kDE, s1 );
        float valueC = fmadd( cZ, f[2], fmadd( f[1], cY, f[0] * cX ) );
        float valueAB = ip2hb < 0 ? valueB : valueA;
        float valueBA = ip2hb < 0 ? valueA : valueB;
        m = fxor( m, ishl( h31, 31 ) );
        v[0] = fmadd( m, ip3 == 0 ? valueC : valueAB, v[0] );
        v[1] = fmadd( m, ip3 == 1 ? valueC : ( ip3 == 2 ? valueBA : valueAB ), v[1] );
        v[2] = fmadd( m, ip3 == 2 ? valueC : valueBA, v[2] );
    }
    // Orthogonal gradient matrix, 4D (:318-694, generic branch)
    FN2D void orthoVec4( int32_t h31, const float* f, float m, float* v )
    {
        const float kCPINAB = -0.375999676691291f, kCPUFAB = 0.222726847849776f;
        const float kCPIPD = -kSkew4f, kCPUD = kSkew4f;
        const float kDPTSAB = -0.124000323308709f, kDPTSD = 0.190983005625053f;
        const float kDSTE = kSkew4f, kDPABC = 0.437016024448821f, kDUFD = -kRoot2f;
        int32_t hs = isrl( h31, 2 );
        int32_t ib = imul( hs, 20 >> 2 );
        int32_t ip3 = imul( hs, (int32_t)0xD5555556u ) >> 29;
        int32_t ip8 = ib >> 24;
        int32_t iPos = ib & ( 0x03 << 27 );
        int32_t iNeg = ( ( ib >> 2 ) & ( 0x03 << 27 ) ) ^ iPos;
        // the reference compares against (0x10 << 27) == INT_MIN with a signed
        // compare, which holds for every value
        const bool extra = true;
        bool single = iPos == iNeg;
        if( ib < ( 0x04 << 27 ) ) iPos = -1;
        float offAB = extra ? kDPTSAB + kDSTE : kDPTSAB;
        float cINAB = single ? kCPINAB + offAB : kCPINAB;
        float cUFAB = single ? kCPUFAB + offAB : kCPUFAB;
        float cINC = single ? cINAB : cINAB + kDPABC;
        float cUFC = single ? cUFAB : cUFAB + kDPABC;
        float offD = extra ? kDPTSD + kDSTE : kDPTSD;
        float cIPD = single ? kCPIPD + offD : kCPIPD;
        float cUD = single ? kCPUD + offD : kCPUD;
        int32_t counter = -1;
        float vA = 0.0f, vB = 0.0f, vC = 0.0f, vD = 0.0f;
#pragma unroll
        for( int axis = 0; axis < 4; axis++ )
        {
            bool ip = iPos == ( axis << 27 );
            bool idx = ip || iNeg == ( axis << 27 );
            float iAB = ip ? cINAB + 1.0f : cINAB;
            float iC = ip ? cINC + 1.0f : cINC;
            float iD = ip ? cIPD : cIPD - 1.0f;
            float uA, uB, uC;
            if( axis == 0 ) { uA = cUFAB + kDUFD; uB = cUFAB; uC = cUFC; }
            else if( axis == 1 ) { uA = counter == -2 ? cUFAB + kDUFD : cUFAB; uB = counter == -1 ? cUFAB + kDUFD : cUFAB; uC = cUFC; }
            else if( axis == 2 ) { uA = counter == -3 ? cUFAB + kDUFD : cUFAB; uB = counter == -2 ? cUFAB + kDUFD : cUFAB; uC = counter == -1 ? cUFC + kDUFD : cUFC; }
            else { uA = cUFAB; uB = counter == -3 ? cUFAB + kDUFD : cUFAB; uC = counter == -2 ? cUFC + kDUFD : cUFC; }
            float gA = idx ? iAB : uA, gB = idx ? iAB : uB, gC = idx ? iC : uC, gD = idx ? iD : cUD;
            if( axis == 0 ) { vA = gA * f[0]; vB = gB * f[0]; vC = gC * f[0]; vD = gD * f[0]; }
            else { vA = fmadd( gA, f[axis], vA ); vB = fmadd( gB, f[axis], vB ); vC = fmadd( gC, f[axis], vC ); vD = fmadd( gD, f[axis], vD ); }
            if( axis < 3 && idx ) counter -= 1;
        }
        int32_t iX = ip8 >> 1;
        int32_t iY = iadd( iadd( iX, 1 ), ip3 );
        int32_t iZ = ip8 & 1;
        iZ = iadd( iadd( iX, 1 ), iZ >= ip3 ? iZ + 1 : iZ );
        int32_t iS = iadd( iadd( iX, iY ), iZ );
        auto pick = [&]( int32_t i ) {
            return ishl( i, 31 ) < 0 ? ( ishl( i, 30 ) < 0 ? vD : vB ) : ( ishl( i, 30 ) < 0 ? vC : vA );
        };
        m = fxor( m, ishl( h31, 31 ) );
        v[0] = fmadd( m, pick( iX ), v[0] );
        v[1] = fmadd( m, pick( iY ), v[1] );
        v[2] = fmadd( m, pick( iZ ), v[2] );
        float w = ishl( iS, 31 ) < 0 ? ( ishl( iS, 30 ) < 0 ? vD : vB ) : ( ishl( iS, 30 ) < 0 ? vA : vC );
        v[3] = fmadd( m, w, v[3] );
    }
    // Gradient outer product, 2D (:6-60, generic branch)
    FN2D void outerVec2( int32_t h31, const float* f, float m, float* v )
    {
        int32_t hs = isrl( h31, 1 );
        int32_t ig = imul( hs, 12 >> 2 );
        int32_t io = imul( hs, (int32_t)0xAAAAAAABu ) & (int32_t)0xC0000003u;
        {
            bool sw = ishl( ig, 2 ) < 0;
            float u = sw ? f[1] : f[0];
            float w = sw ? f[0] : f[1];
            float a = u * ( ig < 0 ? 2.0f : kRoot3f );
            float b = fxor( w, ishl( ig >> 30, 31 ) );
            m *= fxor( ig >= 0 ? a + b : a, ishl( ( ig >> 28 ) ^ io, 31 ) );
        }
        {
            float a = m * ( io < 0 ? 2.0f : kRoot3f );
            float b = fxor( io >= 0 ? m : 0.0f, ishl( io >> 30, 31 ) );
            bool sw = ishl( io, 30 ) < 0;
            v[0] = v[0] + ( sw ? b : a );
            v[1] = v[1] + ( sw ? a : b );
        }
    }
    // Gradient outer product, 3D (:62-116, generic branch)
    FN2D void outerVec3( int32_t h31, const float* f, float m, float* v )
    {
        int32_t hs = isrl( h31, 1 );
        int32_t ig = isrl( imul( hs, 12 >> 2 ), 28 );
        int32_t io = imul( hs, (int32_t)0xAAAAAAABu ) & (int32_t)0xC0000003u;
        io = io | ( io >> 28 );
        {
            float u = ig >= 8 ? f[1] : f[0];
            float w = ig >= 4 ? f[2] : f[1];
            m *= fxor( u, ishl( ig, 31 ) ) + fxor( w, ishl( ig >> 1, 31 ) );
        }
        io &= 0xF;
        float s0 = fxor( m, ishl( io, 31 ) );
        float s1 = fxor( m, ishl( io >> 1, 31 ) );
        bool notYZ = io < 8, notXY = io >= 4;
        v[0] = notYZ ? v[0] + s0 : v[0];
        v[2] = notXY ? v[2] + s1 : v[2];
        v[1] = ( notYZ && notXY ) ? v[1] : v[1] + ( notXY ? s0 : s1 );
    }
    // Gradient outer product, 4D (:118-191, generic branch)
    FN2D void outerGrad4( int32_t index, float* g )
    {
        int32_t iA = index & ( 0x03 << 27 );
        int32_t iB = ( ( index >> 2 ) & ( 0x07 << 27 ) ) ^ iA;
        bool extra = iB >= ( 0x04 << 27 );
        bool equal = iA == iB;
        if( equal ) iA = -1;
        float neutral = ( equal || extra ) ? ( extra ? kSkew4f * -1.0f : kSkew4f ) : 0.0f;
#pragma unroll
        for( int a = 0; a < 4; a++ )
        {
            float t = iA == ( a << 27 ) ? neutral - 1.0f : neutral;
            g[a] = iB == ( a << 27 ) ? t + 1.0f : t;
        }
    }
    FN2D void outerVec4( int32_t h, const float* f, float m, float* v )
    {
        int32_t hs = isrl( h, 2 );
        int32_t ig = imul( hs, 20 >> 2 );
        int32_t io = imul( imul( hs, (int32_t)0xCCCCCCCDu ) & (int32_t)0xE0000003u, 3 | ( 1 << 27 ) );
        float g[4];
        outerGrad4( ig, g );
        m *= fmadd( g[3], f[3], fmadd( g[2], f[2], fmadd( g[1], f[1], g[0] * f[0] ) ) );
        outerGrad4( io, g );
#pragma unroll
        for( int a = 0; a < 4; a++ ) v[a] = fmadd( m, g[a], v[a] );
    }

    template<int D> FN2D void vecContribution( int scheme, int32_t h, const float* f, float m, float* v )
    {
        if constexpr( D == 2 ) { if( scheme == 1 ) outerVec2( h, f, m, v ); else orthoVec2( h, f, m, v ); }
        else if constexpr( D == 3 ) { if( scheme == 1 ) outerVec3( h, f, m, v ); else orthoVec3( h, f, m, v ); }
        else { if( scheme == 1 ) outerVec4( h, f, m, v ); else orthoVec4( h, f, m, v ); }
    }

    // Scale by bounding * axis scale, displace, return the squared length
    // (DomainWarpSimplex.inl:823-832, 919-930, 1073-1086). 2D returns
    // FMulAdd( y, y, x * x ), 3D/4D accumulate from x.
    template<int D> FN2D float warpFinish( float bounding, const float* axisScale, float amp, float* v, float* out )
    {
#pragma unroll
        for( int a = 0; a < D; a++ ) v[a] = v[a] * ( bounding * axisScale[a] );
#pragma unroll
        for( int a = 0; a < D; a++ ) out[a] = fmadd( v[a], amp, out[a] );
        float r = v[0] * v[0];
#pragma unroll
        for( int a = 1; a < D; a++ ) r = fmadd( v[a], v[a], r );
        return r;
    }

    // DomainWarpSimplex::Warp (DomainWarpSimplex.inl:772-1087)
    template<int D> FN2D float warpSimplex( int32_t seed, int scheme, float amp, const float* axisScale, const float* sp, float* out );

    template<> FN2D float warpSimplex<2>( int32_t seed, int scheme, float amp, const float* axisScale, const float* sp, float* out )
    {
        constexpr double kSkew2 = 1.0 / ( kRoot3 + 1.0 );
        constexpr double kUnskew2 = -1.0 / ( kRoot3 + 3.0 );
        float x = sp[0], y = sp[1];
        float skewDelta = (float)kSkew2 * ( x + y );
        float xSk = x + skewDelta, ySk = y + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk );
        float dxS = xSk - xB, dyS = ySk - yB;
        int32_t xP = imul( floorToInt( xB ), kPX ), yP = imul( floorToInt( yB ), kPY );
        bool xGEy = dxS >= dyS;
        float unskewDelta = (float)kUnskew2 * ( dxS + dyS );
        float d0[2] = { dxS + unskewDelta, dyS + unskewDelta };
        float d1[2] = { ( xGEy ? d0[0] : d0[0] + 1.0f ) - (float)( kUnskew2 + 1 ), ( xGEy ? d0[1] + 1.0f : d0[1] ) - (float)( kUnskew2 + 1 ) };
        float d2[2] = { d0[0] - (float)( kUnskew2 * 2 + 1 ), d0[1] - (float)( kUnskew2 * 2 + 1 ) };
        float f0 = fnmadd( d0[0], d0[0], fnmadd( d0[1], d0[1], 0.5f ) );
        float f1 = fnmadd( d1[0], d1[0], fnmadd( d1[1], d1[1], 0.5f ) );
        float f2 = f0 + fmadd( unskewDelta, (float)( -4.0 * ( kRoot3 + 2.0 ) / ( kRoot3 + 3.0 ) ), (float)( -2.0 / 3.0 ) );
        f0 = fmaxf_( f0, 0.0f ); f1 = fmaxf_( f1, 0.0f ); f2 = fmaxf_( f2, 0.0f );
        f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2;
        float v[2] = { 0.0f, 0.0f };
        vecContribution<2>( scheme, hash2( seed, xP, yP ), d0, f0, v );
        vecContribution<2>( scheme, hash2( seed, xGEy ? iadd( xP, kPX ) : xP, xGEy ? yP : iadd( yP, kPY ) ), d1, f1, v );
        vecContribution<2>( scheme, hash2( seed, iadd( xP, kPX ), iadd( yP, kPY ) ), d2, f2, v );
        const float kBound = scheme == 1 ? (float)( 49.918426513671875 / 2.0 ) : 70.1480577066486f;
        v[0] = v[0] * ( kBound * axisScale[0] );
        v[1] = v[1] * ( kBound * axisScale[1] );
        out[0] = fmadd( v[0], amp, out[0] );
        out[1] = fmadd( v[1], amp, out[1] );
        return fmadd( v[1], v[1], v[0] * v[0] );
    }
    template<> FN2D float warpSimplex<3>( int32_t seed, int scheme, float amp, const float* axisScale, const float* sp, float* out )
    {
        constexpr double kReflectUnskew3 = -1.0 / 2.0;
        float x = sp[0], y = sp[1], z = sp[2];
        float skewDelta = (float)( 1.0 / 3.0 ) * ( x + y + z );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB;
        int32_t xP = imul( floorToInt( xB ), kPX ), yP = imul( floorToInt( yB ), kPY ), zP = imul( floorToInt( zB ), kPZ );
        bool xGEy = dxS >= dyS, yGEz = dyS >= dzS, xGEz = dxS >= dzS;
        float unskewDelta = (float)kReflectUnskew3 * ( dxS + dyS + dzS );
        float d0[3] = { dxS + unskewDelta, dyS + unskewDelta, dzS + unskewDelta };
        bool x1m = xGEy && xGEz, y1m = yGEz && !xGEy, z1inv = xGEz || yGEz;
        bool x2inv = xGEy || xGEz, y2m = xGEy && !yGEz, z2m = xGEz && yGEz;
        const float k3 = (float)( kReflectUnskew3 * 3 + 1 );
        float d3[3] = { d0[0] - k3, d0[1] - k3, d0[2] - k3 };
        float d1[3] = { x1m ? d3[0] - 1.0f : d3[0], y1m ? d3[1] - 1.0f : d3[1], z1inv ? d3[2] : d3[2] - 1.0f };
        float d2[3] = { x2inv ? d0[0] : d0[0] + 1.0f, y2m ? d0[1] + 1.0f : d0[1], z2m ? d0[2] + 1.0f : d0[2] };
        const float fr = (float)0.6;
        float f0 = fnmadd( d0[2], d0[2], fnmadd( d0[1], d0[1], fnmadd( d0[0], d0[0], fr ) ) );
        float f1 = fnmadd( d1[2], d1[2], fnmadd( d1[1], d1[1], fnmadd( d1[0], d1[0], fr ) ) );
        float f2 = fnmadd( d2[2], d2[2], fnmadd( d2[1], d2[1], fnmadd( d2[0], d2[0], fr ) ) );
        float f3 = f0 - ( unskewDelta + (float)( 3.0 / 4.0 ) );
        f0 = fmaxf_( f0, 0.0f ); f1 = fmaxf_( f1, 0.0f ); f2 = fmaxf_( f2, 0.0f ); f3 = fmaxf_( f3, 0.0f );
        f0 *= f0; f0 *= f0; f1 *= f1; f1 *= f1; f2 *= f2; f2 *= f2; f3 *= f3; f3 *= f3;
        float v[3] = { 0.0f, 0.0f, 0.0f };
        vecContribution<3>( scheme, hash3( seed, xP, yP, zP ), d0, f0, v );
        vecContribution<3>( scheme, hash3( seed, x1m ? iadd( xP, kPX ) : xP, y1m ? iadd( yP, kPY ) : yP, z1inv ? zP : iadd( zP, kPZ ) ), d1, f1, v );
        vecContribution<3>( scheme, hash3( seed, x2inv ? iadd( xP, kPX ) : xP, y2m ? yP : iadd( yP, kPY ), z2m ? zP : iadd( zP, kPZ ) ), d2, f2, v );
        vecContribution<3>( scheme, hash3( seed, iadd( xP, kPX ), iadd( yP, kPY ), iadd( zP, kPZ ) ), d3, f3, v );
        if( scheme == 1 )
        {
            float delta = (float)( -2.0 / 2.0 ) * ( v[0] + v[1] + v[2] );
            v[0] += delta; v[1] += delta; v[2] += delta;
        }
        const float kBound = scheme == 1 ? (float)( 32.69428253173828125 / 1.4142135623730951 ) : 16.281631889139874f;
        return warpFinish<3>( kBound, axisScale, amp, v, out );
    }
    template<> FN2D float warpSimplex<4>( int32_t seed, int scheme, float amp, const float* axisScale, const float* sp, float* out )
    {
        constexpr double kSkew4 = 1.0 / ( kRoot5 + 1.0 );
        constexpr double kUnskew4 = -1.0 / ( kRoot5 + 5.0 );
        float x = sp[0], y = sp[1], z = sp[2], w = sp[3];
        float skewDelta = (float)kSkew4 * ( x + y + z + w );
        float xSk = x + skewDelta, ySk = y + skewDelta, zSk = z + skewDelta, wSk = w + skewDelta;
        float xB = ffloor( xSk ), yB = ffloor( ySk ), zB = ffloor( zSk ), wB = ffloor( wSk );
        float dxS = xSk - xB, dyS = ySk - yB, dzS = zSk - zB, dwS = wSk - wB;
        const int32_t base[4] = { imul( floorToInt( xB ), kPX ), imul( floorToInt( yB ), kPY ), imul( floorToInt( zB ), kPZ ), imul( floorToInt( wB ), kPW ) };
        const int32_t prime[4] = { kPX, kPY, kPZ, kPW };
        float unskewDelta = (float)kUnskew4 * ( dxS + dyS + dzS + dwS );
        float d0[4] = { dxS + unskewDelta, dyS + unskewDelta, dzS + unskewDelta, dwS + unskewDelta };
        int rank[4] = { 0, 0, 0, 0 };
#pragma unroll
        for( int a = 0; a < 4; a++ )
#pragma unroll
            for( int b = a + 1; b < 4; b++ )
            {
                if( d0[a] >= d0[b] ) rank[a]++; else rank[b]++;
            }
        const float fr = (float)0.6;
        float v[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
        float f0 = fnmadd( d0[3], d0[3], fnmadd( d0[2], d0[2], fnmadd( d0[1], d0[1], fnmadd( d0[0], d0[0], fr ) ) ) );
        float f4 = f0 + fmadd( unskewDelta, (float)( -4.0 * ( kRoot5 + 3.0 ) / ( kRoot5 + 5.0 ) ), (float)( -4.0 / 5.0 ) );
        {
            float f = fmaxf_( f0, 0.0f ); f *= f; f *= f;
            vecContribution<4>( scheme, hash4( seed, base[0], base[1], base[2], base[3] ), d0, f, v );
        }
#pragma unroll
        for( int vi = 1; vi <= 3; vi++ )
        {
            const float u = (float)( kUnskew4 * vi );
            float d[4];
            int32_t hp[4];
#pragma unroll
            for( int a = 0; a < 4; a++ )
            {
                bool m = rank[a] > 3 - vi;
                d[a] = ( m ? d0[a] - 1.0f : d0[a] ) - u;
                hp[a] = m ? iadd( base[a], prime[a] ) : base[a];
            }
            float f = fnmadd( d[3], d[3], fnmadd( d[2], d[2], fnmadd( d[1], d[1], fnmadd( d[0], d[0], fr ) ) ) );
            f = fmaxf_( f, 0.0f ); f *= f; f *= f;
            vecContribution<4>( scheme, hash4( seed, hp[0], hp[1], hp[2], hp[3] ), d, f, v );
        }
        {
            const float u4 = (float)( kUnskew4 * 4 + 1 );
            float d[4] = { d0[0] - u4, d0[1] - u4, d0[2] - u4, d0[3] - u4 };
            float f = fmaxf_( f4, 0.0f ); f *= f; f *= f;
            vecContribution<4>( scheme, hash4( seed, iadd( base[0], kPX ), iadd( base[1], kPY ), iadd( base[2], kPZ ), iadd( base[3], kPW ) ), d, f, v );
        }
        const float kBound = scheme == 1 ? (float)( 33.653125584827855 / 1.4142135623730951 ) : 30.88161777516092f;
        return warpFinish<4>( kBound, axisScale, amp, v, out );
    }

    // DomainWarpSuperSimplex::Warp (DomainWarpSimplex.inl:1131-1836); no seed offset
    template<int D> FN2D float warpSuperSimplex( int32_t seed, int scheme, float amp, const float* axisScale, const float* sp, float* out )
    {
        float v[D];
#pragma unroll
        for( int a = 0; a < D; a++ ) v[a] = 0.0f;
        auto vtx = [&]( int32_t h, const float* d, float f, bool ) { vecContribution<D>( scheme, h, d, f, v ); };
        if constexpr( D == 2 )
        {
            superWalk2( seed, sp[0], sp[1], vtx );
            const float kBound = scheme == 1 ? (float)( 9.28993664146183 / 2.0 ) : 12.814453124999995f;
            v[0] = v[0] * ( kBound * axisScale[0] );
            v[1] = v[1] * ( kBound * axisScale[1] );
            out[0] = fmadd( v[0], amp, out[0] );
            out[1] = fmadd( v[1], amp, out[1] );
            return fmadd( v[1], v[1], v[0] * v[0] );
        }
        else if constexpr( D == 3 )
        {
            superWalk3( seed, sp[0], sp[1], sp[2], vtx );
            if( scheme == 1 )
            {
                float delta = (float)( -2.0 / 3.0 ) * ( v[0] + v[1] + v[2] );
                v[0] += delta; v[1] += delta; v[2] += delta;
            }
            const float kBound = scheme == 1 ? (float)( 144.736422163332608 / 1.4142135623730951 ) : 37.63698669623629f;
            return warpFinish<3>( kBound, axisScale, amp, v, out );
        }
        else
        {
            superWalk4( seed, sp[0], sp[1], sp[2], sp[3], vtx );
            const float kBound = scheme == 1 ? (float)( 115.21625311930542 / 1.4142135623730951 ) : 48.80058117543753f;
            return warpFinish<4>( kBound, axisScale, amp, v, out );
        }
    }
}
