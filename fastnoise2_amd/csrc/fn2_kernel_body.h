// Bodies of the generator kernels, shared by the interpreter kernels
// (fn2_kernels.hip) and the run-time specialised kernels (fn2_jit.cpp).
// `prog( args, R )` evaluates one point whose starting position is in R and
// returns its value: either the interpreter loop or a fixed, inlined program.
//
// Output range: every lane keeps a running min/max in registers; a block
// reduces them through wave shuffles and LDS and publishes one atomic pair.
// Same-address atomics serialise in L2, so the number of blocks, not of
// waves, bounds their cost: the single-call kernel runs a grid-stride loop
// over a grid sized to fill the GPU, and the batched kernel gives each block
// a tile of several points per lane.
#pragma once
#include "fn2_interp.h"

namespace fn2k
{
    using namespace fn2d;
    using namespace fn2i;

    // Output stores of whole wave rows (64 consecutive floats): each value
    // is written once and never read back by the kernel, so they go out as
    // streaming (non-temporal) stores
    __device__ __forceinline__ void storeOut( float* p, float v ) { __builtin_nontemporal_store( v, p ); }

    // BLOCK: the most threads a block of the kernel has (its launch bound);
    // the block may have fewer, any multiple of 64.
    template<int BLOCK>
    __device__ __forceinline__ void blockMinMax( float mn, float mx, int32_t* record, int stride, int slots )
    {
        int32_t* keys = record + ( blockIdx.x & (uint32_t)( slots - 1 ) ) * stride;
        __shared__ float sMin[BLOCK / 64], sMax[BLOCK / 64];
#pragma unroll
        for( int off = 32; off > 0; off >>= 1 )
        {
            mn = fminf( mn, __shfl_xor( mn, off ) );
            mx = fmaxf( mx, __shfl_xor( mx, off ) );
        }
        const int wave = threadIdx.x >> 6;
        if( ( threadIdx.x & 63 ) == 0 )
        {
            sMin[wave] = mn;
            sMax[wave] = mx;
        }
        __syncthreads();
        if( threadIdx.x == 0 )
        {
            const int waves = (int)( blockDim.x >> 6 );
#pragma unroll
            for( int w = 1; w < BLOCK / 64; w++ )
            {
                if( w >= waves ) break;
                mn = fminf( mn, sMin[w] );
                mx = fmaxf( mx, sMax[w] );
            }
            if( mn <= mx )
            {
                atomicMin( &keys[0], orderedKey( mn ) );
                atomicMax( &keys[1], orderedKey( mx ) );
            }
        }
    }

    // A compiler fence over the first N registers of a specialised program's
    // register file: after it their values count as unknown, so the compiler
    // cannot reuse a computation of one instruction in another (on long
    // programs that kept hundreds of values live and left one wave per SIMD).
    // It emits no instruction.
    template<int N>
    __device__ __forceinline__ void registerFence( float* R )
    {
#pragma unroll
        for( int j = 0; j < N; j++ ) asm volatile( "" : "+v"( R[j] ) );
    }

    // The register file of one point of a specialised program (every register
    // written before it is read, all indices compile-time constants): a fresh
    // private array per point, which lives in VGPRs and leaves the compiler no
    // values to carry between points.
    struct RegFile
    {
        float R[fn2::kMaxRegs];
    };

    // The interpreter's register files: indexed at run time, so they live in
    // LDS, one column per thread (consecutive threads in consecutive banks),
    // zeroed once. The launch sizes the LDS to the program's register count,
    // so short programs keep the occupancy their VGPRs allow. Neither VGPR
    // indexing nor scratch memory, whose spills reach HBM on long programs.
    template<int BLOCK>
    __device__ __forceinline__ StridedRegs<BLOCK> ldsRegs( int regs )
    {
        extern __shared__ float fn2RegisterFile[];   // regs * BLOCK floats
        for( int r = 0; r < regs; r++ ) fn2RegisterFile[r * BLOCK + threadIdx.x] = 0.0f;
        return StridedRegs<BLOCK>{ fn2RegisterFile + threadIdx.x };
    }

    // A grid request walked in wave-sized bricks (8 x 8 in 2D, 4 x 4 x 4 in
    // 3D) instead of runs of 64 along x: the lanes of a wave then sit close
    // together in every axis, so they agree more often on which cells a pruned
    // cellular walk may leave out. Positions are computed as the row-major
    // walk computes them; each lane writes its own output element.
    template<int BLOCK, int D, class Prog>
    __device__ __forceinline__ void generateBricks( const fn2::GenArgs& a, Prog prog, float& mn, float& mx )
    {
        constexpr uint32_t BX = D == 2 ? 8 : 4, BY = D == 2 ? 8 : 4, BZ = D == 2 ? 1 : 4;
        const uint32_t nx = a.div[0].d;
        const uint32_t ny = D == 2 ? (uint32_t)a.total / nx : a.div[1].d;
        const uint32_t nz = D == 2 ? 1u : (uint32_t)a.total / ( nx * ny );
        const uint32_t tx = ( nx + BX - 1 ) / BX, ty = ( ny + BY - 1 ) / BY, tz = ( nz + BZ - 1 ) / BZ;
        const uint32_t bricks = tx * ty * tz;
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t lx = lane % BX, ly = ( lane / BX ) % BY, lz = lane / ( BX * BY );
        constexpr uint32_t WAVES = BLOCK / 64;
        // the brick index is the same on every lane of a wave: scalar arithmetic
        const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane( (int)( blockIdx.x * WAVES + threadIdx.x / 64 ) );
        for( uint32_t t = first; t < bricks; t += gridDim.x * WAVES )
        {
            const uint32_t ix = t % tx, rest = t / tx, iy = rest % ty, iz = rest / ty;
            const uint32_t x = ix * BX + lx, y = iy * BY + ly, z = iz * BZ + lz;
            if( x < nx && y < ny && z < nz )
            {
                RegFile fresh;
                fresh.R[0] = fmadd( (float)(int32_t)x, a.step[0], a.offset[0] );
                fresh.R[1] = fmadd( (float)(int32_t)y, a.step[1], a.offset[1] );
                if constexpr( D == 3 ) fresh.R[2] = fmadd( (float)(int32_t)z, a.step[2], a.offset[2] );
                // the program knows it runs in bricks (fn2i::fractalLeaf),
                // unless it is one copy for both walks (long programs)
                float v;
                if constexpr( Prog::kBrickAware ) v = prog.template eval<true>( a, fresh.R );
                else v = prog( a, fresh.R );
                // a plain store: a wave writes 16-byte pieces of 16 lines
                // here, which the L2 merges with its neighbours' pieces;
                // streaming stores went out piece by piece (2.8 times the
                // output's bytes on the warped cellular 256^3 grid)
                a.out[( z * ny + y ) * nx + x] = v;
                if( v == v )
                {
                    mn = fminf( mn, v );
                    mx = fmaxf( mx, v );
                }
            }
        }
    }

    // One request: grid-stride loop over all of its points.
    template<int BLOCK, class Prog>
    __device__ __forceinline__ void generateSingle( const fn2::GenArgs& a, Prog prog )
    {
        float mn = __builtin_inff(), mx = -__builtin_inff();
        auto emit = [&]( uint32_t i, float v ) {
            storeOut( a.out + i, v );
            if( v == v )
            {
                mn = fminf( mn, v );
                mx = fmaxf( mx, v );
            }
        };
        const uint32_t total = (uint32_t)a.total;
        const uint32_t stride = gridDim.x * BLOCK;
        if constexpr( Prog::kFresh )
        {
            bool bricks = false;
            if constexpr( Prog::kBricks && ( Prog::kDims == 2 || Prog::kDims == 3 ) )
            {
                if( a.kind == fn2::SRC_GRID && a.dims == Prog::kDims )
                {
                    generateBricks<BLOCK, Prog::kDims>( a, prog, mn, mx );
                    bricks = true;
                }
            }
            if( !bricks )
            {
                for( uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < total; i += stride )
                {
                    RegFile fresh;
                    startPositionD<Prog::kDims, false>( a, i, fresh.R );
                    emit( i, prog( a, fresh.R ) );
                }
            }
        }
        else
        {
            const auto R = ldsRegs<BLOCK>( a.regs );
            for( uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < total; i += stride )
            {
                startPosition<false>( a, i, R );
                emit( i, prog( a, R ) );
            }
        }
        if( a.minmax ) blockMinMax<BLOCK>( mn, mx, a.minmax, a.rangeStride, a.rangeSlots );
    }

    // One tile of a batched request: points [first, first + threads * ppt)
    // of job a (clipped to its size), where threads is the block's size (at
    // most BLOCK; the interpreter's register file in LDS is laid out for
    // BLOCK) and ppt the job's points per lane; the tile's range goes to the
    // job's key pair.
    template<int BLOCK, class Prog>
    __device__ __forceinline__ void generateTile( const fn2::GenArgs& a, uint32_t first, Prog prog )
    {
        float mn = __builtin_inff(), mx = -__builtin_inff();
        auto emit = [&]( uint32_t i, float v ) {
            storeOut( a.out + i, v );
            if( v == v )
            {
                mn = fminf( mn, v );
                mx = fmaxf( mx, v );
            }
        };
        const uint32_t total = (uint32_t)a.total;
        const uint32_t base = first + threadIdx.x;
        const uint32_t threads = blockDim.x;
        const int ppt = a.ppt;
        if constexpr( Prog::kFresh )
        {
#pragma unroll 1
            for( int p = 0; p < ppt; p++ )
            {
                const uint32_t i = base + p * threads;
                if( i < total )
                {
                    RegFile fresh;
                    startPositionD<Prog::kDims>( a, i, fresh.R );
                    emit( i, prog( a, fresh.R ) );
                }
            }
        }
        else
        {
            const auto R = ldsRegs<BLOCK>( a.regs );
#pragma unroll 1
            for( int p = 0; p < ppt; p++ )
            {
                const uint32_t i = base + p * threads;
                if( i < total )
                {
                    startPosition( a, i, R );
                    emit( i, prog( a, R ) );
                }
            }
        }
        if( a.minmax ) blockMinMax<BLOCK>( mn, mx, a.minmax, a.rangeStride, a.rangeSlots );
    }

    // Batched requests: block b runs tile (blockJob[b], blockBase[b]).
    template<int BLOCK, class Prog>
    __device__ __forceinline__ void generateBatch( const fn2::GenArgs* __restrict__ jobs, const int32_t* __restrict__ blockJob,
                                                   const uint32_t* __restrict__ blockBase, Prog prog )
    {
        // the job is read through a pointer (uniform address -> scalar loads)
        // rather than copied into a private struct
        generateTile<BLOCK>( jobs[blockJob[blockIdx.x]], blockBase[blockIdx.x], prog );
    }

    // Publishes a wave's range of one job to slot `slot` of its range record:
    // a wave reduction, then global atomics of one wave-uniform value (the
    // compiler issues them from one lane; it leaves generic-pointer atomics
    // per lane). No-op without a record.
    __device__ __forceinline__ void waveMinMax( float mn, float mx, int32_t* record, int slot, int stride )
    {
        if( !record ) return;
        int32_t* keys = record + slot * stride;
#pragma unroll
        for( int off = 32; off > 0; off >>= 1 )
        {
            mn = fminf( mn, __shfl_xor( mn, off ) );
            mx = fmaxf( mx, __shfl_xor( mx, off ) );
        }
        // a wave without a number leaves the keys as they are
        const int32_t kMin = mn <= mx ? orderedKey( mn ) : 0x7fffffff;
        const int32_t kMax = mn <= mx ? orderedKey( mx ) : (int32_t)0x80000000u;
        auto* g = (__attribute__( ( address_space( 1 ) ) ) int32_t*)keys;
        __hip_atomic_fetch_min( g, __builtin_amdgcn_readfirstlane( kMin ), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
        __hip_atomic_fetch_max( g + 1, __builtin_amdgcn_readfirstlane( kMax ), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }

    // Work-queue form of a batch. The unit of work is a quarter of a tile:
    // item w covers the lanes (w % W) * 64 .. + 63 of tile w / W (W = waves
    // per block) exactly as that wave of block w / W runs it in
    // generateBatch. Wave g of the grid (G waves) runs item g first. Items
    // beyond G are split into min(G, kQueueCounters) contiguous parts
    // (fn2::queuePart); wave g then pulls the items of its part from that
    // part's counter (zero at launch, kQueueCounterStride apart) until the
    // part is done. A
    // batch that fits the grid takes no atomic at all, and a large one
    // spreads its pulls over as many addresses (one counter for every
    // wave queued up at about 5 ns a pull, nearly doubling a large batch).
    // A wave publishes each item's range to its slot of the job's record
    // (a running range across items, published on a change of job, kept 8
    // more VGPRs live in the loop: a wave per SIMD on the serving tree). The
    // loop has no barrier and no branch that only some lanes take around its
    // back edge: all lanes take part in the pull (lane 0 adds one, the
    // others zero) and in publishing a range. The counters only grow, so
    // every wave leaves after at most a part's items + 1 pulls and no wave
    // waits for another.
    template<int BLOCK, class Prog>
    __device__ __forceinline__ void generateQueue( const fn2::GenArgs* __restrict__ jobs, const int32_t* __restrict__ tileJob,
                                                   const uint32_t* __restrict__ tileBase, uint32_t nTiles, uint32_t* counter, Prog prog )
    {
        constexpr uint32_t kWaves = BLOCK / 64;
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t items = nTiles * kWaves;
        const uint32_t g = (uint32_t)__builtin_amdgcn_readfirstlane( (int)( blockIdx.x * kWaves + ( threadIdx.x >> 6 ) ) );
        const uint32_t waves = gridDim.x * kWaves;
        // this wave's part of the items beyond the first round: [first, end)
        const fn2::QueuePart q = fn2::queuePart( g, waves, items );
        const uint32_t first = q.first, end = q.end;
        uint32_t* myCounter = counter + q.part * fn2::kQueueCounterStride;
        for( uint32_t w = g; w < items; )
        {
            const uint32_t t = w / kWaves;
            const fn2::GenArgs& a = jobs[tileJob[t]];
            float mn = __builtin_inff(), mx = -__builtin_inff();
            const uint32_t total = (uint32_t)a.total;
            const uint32_t base = tileBase[t] + ( w % kWaves ) * 64 + lane;
            const int ppt = a.ppt;
            auto emit = [&]( uint32_t i, float v ) {
                storeOut( a.out + i, v );
                if( v == v )
                {
                    mn = fminf( mn, v );
                    mx = fmaxf( mx, v );
                }
            };
            if constexpr( Prog::kFresh )
            {
#pragma unroll 1
                for( int p = 0; p < ppt; p++ )
                {
                    const uint32_t i = base + p * BLOCK;
                    if( i < total )
                    {
                        RegFile fresh;
                        startPositionD<Prog::kDims>( a, i, fresh.R );
                        emit( i, prog( a, fresh.R ) );
                    }
                }
            }
            else
            {
                const auto R = ldsRegs<BLOCK>( a.regs );   // each item starts from a cleared register file, as a block does
#pragma unroll 1
                for( int p = 0; p < ppt; p++ )
                {
                    const uint32_t i = base + p * BLOCK;
                    if( i < total )
                    {
                        startPosition( a, i, R );
                        emit( i, prog( a, R ) );
                    }
                }
            }
            waveMinMax( mn, mx, a.minmax, (int)( g & (uint32_t)( a.rangeSlots - 1 ) ), a.rangeStride );
            if( first >= end ) break;   // no items in this wave's part
            const uint32_t pulled = atomicAdd( myCounter, lane == 0 ? 1u : 0u );
            w = first + (uint32_t)__builtin_amdgcn_readfirstlane( (int)pulled );
            if( w >= end ) break;
        }
    }
}
