// Test infrastructure only: the grid / position-array / tileable drivers of
// the CPU oracle (Generator.inl:79-433), run on N lanes at a time and split
// over host threads in contiguous index slabs.
#pragma once
#include <thread>
#include <vector>
#include <cmath>
#include "fn2_oracle_gen.h"

namespace FN2O_NS
{
    template<int N, bool FUSED, bool AVX2>
    struct Run
    {
        using E = Eval<N, FUSED, AVX2>;
        using F = FV<N>;
        using I = IV<N>;

        template<int D>
        static void block( const DriverArgs& a, int64_t begin, int64_t end, MinMax& mm )
        {
            // per-axis index of the current point, x innermost (Generator.inl:447-456):
            // decomposed once per slab, then advanced by carries like AxisReset
            int32_t axis[4] = { 0, 0, 0, 0 };
            if( a.kind != 1 )
            {
                int64_t rem = begin;
                for( int d = 0; d < D; d++ )
                {
                    axis[d] = (int32_t)( rem % a.count[d] );
                    rem /= a.count[d];
                }
            }
            for( int64_t base = begin; base < end; base += N )
            {
                Pos<D, N> p;
                for( int k = 0; k < N; k++ )
                {
                    const bool live = base + k < end;
                    int64_t idx = live ? base + k : end - 1;
                    if( a.kind == 1 )
                    {
                        // GenPositionArray: offset + array value (Generator.inl:265-266, 295-297, 327-330)
                        for( int d = 0; d < D; d++ ) p.c[d].v[k] = a.offset[d] + a.posArrays[d][idx];
                    }
                    else
                    {
                        for( int d = 0; d < D; d++ ) p.c[d].v[k] = (float)axis[d];
                        if( live && base + k + 1 < end )
                        {
                            for( int d = 0; d < D; d++ )
                            {
                                if( ++axis[d] < a.count[d] || d == D - 1 ) break;
                                axis[d] = 0;
                            }
                        }
                    }
                }
                if( a.kind == 0 )
                {
                    // FMulAdd( Convert<float>( idx ), step, offset ) (Generator.inl:104-105, 160-162, 225-228)
                    for( int d = 0; d < D; d++ ) p.c[d] = E::fmadd( p.c[d], E::S( a.step[d] ), E::S( a.offset[d] ) );
                }
#ifdef FN2O_TRACING
                fn2o_trace::fp = 0xCBF29CE484222325ull;
#endif
                F v = E::template gen<D>( a.root, E::SI( a.seed ), p );
#ifdef FN2O_TRACING
                if( fn2o_trace::out ) fn2o_trace::out[base] = fn2o_trace::fp;
#endif
                for( int k = 0; k < N && base + k < end; k++ )
                {
                    a.out[base + k] = v.v[k];
                    mm.add( v.v[k] );
                }
            }
        }

        // GenTileable2D (Generator.inl:370-433): cosf/sinf for FS::Cos/FS::Sin
        static void tileable( const DriverArgs& a, int64_t begin, int64_t end, MinMax& mm )
        {
            const float pi2Recip = 0.15915493667f;
            const float xSizePi = (float)a.count[0] * pi2Recip;
            const float ySizePi = (float)a.count[1] * pi2Recip;
            const float xFreq = xSizePi * a.step[0];
            const float yFreq = ySizePi * a.step[1];
            const float xMul = 1 / xSizePi;
            const float yMul = 1 / ySizePi;
            for( int64_t base = begin; base < end; base += N )
            {
                Pos<4, N> p;
                for( int k = 0; k < N; k++ )
                {
                    int64_t idx = base + k < end ? base + k : end - 1;
                    float xF = (float)(int32_t)( idx % a.count[0] ) * xMul;
                    float yF = (float)(int32_t)( idx / a.count[0] ) * yMul;
                    p.c[0].v[k] = cosf( xF ) * xFreq;
                    p.c[1].v[k] = cosf( yF ) * yFreq;
                    p.c[2].v[k] = sinf( xF ) * xFreq;
                    p.c[3].v[k] = sinf( yF ) * yFreq;
                }
#ifdef FN2O_TRACING
                fn2o_trace::fp = 0xCBF29CE484222325ull;
#endif
                F v = E::template gen<4>( a.root, E::SI( a.seed ), p );
#ifdef FN2O_TRACING
                if( fn2o_trace::out ) fn2o_trace::out[base] = fn2o_trace::fp;
#endif
                for( int k = 0; k < N && base + k < end; k++ )
                {
                    a.out[base + k] = v.v[k];
                    mm.add( v.v[k] );
                }
            }
        }

        static void range( const DriverArgs& a, int64_t begin, int64_t end, MinMax& mm )
        {
            if( a.kind == 2 ) { tileable( a, begin, end, mm ); return; }
            switch( a.dims )
            {
            case 2: block<2>( a, begin, end, mm ); break;
            case 3: block<3>( a, begin, end, mm ); break;
            default: block<4>( a, begin, end, mm ); break;
            }
        }

        static void run( const DriverArgs& a, int threads, float* minmax )
        {
            MinMax total;
            if( a.total > 0 )
            {
                if( threads < 1 ) threads = 1;
                int64_t per = ( a.total + threads - 1 ) / threads;
                per = ( per + N - 1 ) / N * N;
                std::vector<MinMax> mms( threads );
                std::vector<std::thread> pool;
                for( int t = 0; t < threads; t++ )
                {
                    int64_t b = (int64_t)t * per, e = b + per < a.total ? b + per : a.total;
                    if( b >= e ) break;
                    if( threads == 1 ) range( a, b, e, mms[t] );
                    else pool.emplace_back( [&a, b, e, &mms, t]() { range( a, b, e, mms[t] ); } );
                }
                for( auto& th : pool ) th.join();
                for( auto& m : mms ) if( m.mn <= m.mx ) { total.add( m.mn ); total.add( m.mx ); }
            }
            if( minmax ) { minmax[0] = total.mn; minmax[1] = total.mx; }
        }

        template<int D> static float single( const Node* root, const float* pos, int seed )
        {
            Pos<D, N> p;
            for( int d = 0; d < D; d++ ) p.c[d] = E::S( pos[d] );
            return E::template gen<D>( root, E::SI( seed ), p ).v[0];
        }
    };

    // Dispatch on the two mode bits.
    template<int N>
    inline void run_mode( const DriverArgs& a, int mode, int threads, float* minmax )
    {
        switch( mode & 3 )
        {
        case 0: Run<N, true, false>::run( a, threads, minmax ); break;
        case 1: Run<N, false, false>::run( a, threads, minmax ); break;
        case 2: Run<N, true, true>::run( a, threads, minmax ); break;
        default: Run<N, false, true>::run( a, threads, minmax ); break;
        }
    }
}
