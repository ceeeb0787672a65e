// HIP kernels for gfx950: the node-tree program interpreter and the output
// range reduction.
//
// One lane evaluates one point: it forms the point's position (uniform grid,
// position arrays or the tileable 4D torus of Generator.inl:79-433), runs the
// wave-uniform instruction list of the compiled tree on a 32-entry register
// file held in VGPRs (uniform indices -> GPR indexing, no scratch), writes one
// float with a coalesced store and folds it into a block min/max that is
// published with one atomic pair per block.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fn2_interp.h"

using namespace fn2d;
using namespace fn2i;

namespace
{
    template<int BLOCK>
    __device__ __forceinline__ void reduceMinMax( float mn, float mx, int32_t* keys )
    {
        // wave reduce, then one atomic pair per wave
#pragma unroll
        for( int off = 32; off > 0; off >>= 1 )
        {
            mn = fminf( mn, __shfl_xor( mn, off ) );
            mx = fmaxf( mx, __shfl_xor( mx, off ) );
        }
        if( ( threadIdx.x & 63 ) == 0 && mn <= mx )
        {
            atomicMin( &keys[0], orderedKey( mn ) );
            atomicMax( &keys[1], orderedKey( mx ) );
        }
    }
}

template<int BLOCK, int DM, int FT>
__global__ void __launch_bounds__( BLOCK ) fn2_generate( fn2::GenArgs a )
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const bool live = i < (uint32_t)a.total;
    const uint32_t idx = live ? i : (uint32_t)a.total - 1;
    float R[fn2::kMaxRegs];
#pragma unroll
    for( int r = 0; r < fn2::kMaxRegs; r++ ) R[r] = 0.0f;
    startPosition( a, idx, R );
    float v = runProgram<DM, FT>( a, R );
    if( live ) a.out[i] = v;
    if( a.minmax )
    {
        bool ok = live && v == v;
        reduceMinMax<BLOCK>( ok ? v : INFINITY, ok ? v : -INFINITY, a.minmax );
    }
}

// Batched form for the serving layer: every block looks up its job in a
// per-block table, so requests with different trees, sizes and sources run in
// one launch. `jobs` holds one GenArgs per request; `blockJob[b]` and
// `blockBase[b]` map block b to (request, first point).
template<int BLOCK, int DM, int FT>
__global__ void __launch_bounds__( BLOCK ) fn2_generate_batch( const fn2::GenArgs* __restrict__ jobs, const int32_t* __restrict__ blockJob, const uint32_t* __restrict__ blockBase )
{
    // read the job through a pointer (uniform address -> scalar loads) rather
    // than copying it into a private struct
    const fn2::GenArgs& a = jobs[blockJob[blockIdx.x]];
    const uint32_t i = blockBase[blockIdx.x] + threadIdx.x;
    const bool live = i < (uint32_t)a.total;
    const uint32_t idx = live ? i : (uint32_t)a.total - 1;
    float R[fn2::kMaxRegs];
#pragma unroll
    for( int r = 0; r < fn2::kMaxRegs; r++ ) R[r] = 0.0f;
    startPosition( a, idx, R );
    float v = runProgram<DM, FT>( a, R );
    if( live ) a.out[i] = v;
    if( a.minmax )
    {
        bool ok = live && v == v;
        reduceMinMax<BLOCK>( ok ? v : INFINITY, ok ? v : -INFINITY, a.minmax );
    }
}

__global__ void fn2_minmax_init( int32_t* keys, int n )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < n )
    {
        keys[2 * i] = 0x7fffffff;          // key of +NaN region: above every float
        keys[2 * i + 1] = (int32_t)0x80000000u;
    }
}

extern "C" {

// Launchers used by the host backend (fn2_backend.cpp). Return hipError_t.
// Kernel variant table: (dimension mask, feature mask) pairs compiled here.
// The host picks the first variant whose masks cover the program's.
#define FN2_VARIANTS( X ) \
    X( 4, 0 ) X( 8, 0 ) X( 16, 0 ) \
    X( 4, 7 ) X( 8, 7 ) X( 16, 7 ) \
    X( 28, 7 )

int fn2_launch_generate( const fn2::GenArgs* args, int dimMask, int featMask, hipStream_t stream )
{
    if( args->total <= 0 ) return 0;
    const int block = 256;
    unsigned grid = (unsigned)( ( (int64_t)args->total + block - 1 ) / block );
#define FN2_TRY( DMV, FTV ) \
    if( ( dimMask & ~DMV ) == 0 && ( featMask & ~FTV ) == 0 ) \
    { \
        hipLaunchKernelGGL( ( fn2_generate<256, DMV, FTV> ), dim3( grid ), dim3( block ), 0, stream, *args ); \
        return (int)hipGetLastError(); \
    }
    FN2_VARIANTS( FN2_TRY )
#undef FN2_TRY
    return (int)hipErrorInvalidValue;
}

int fn2_launch_generate_batch( const fn2::GenArgs* dJobs, const int32_t* dBlockJob, const uint32_t* dBlockBase, int nBlocks, int dimMask, int featMask, hipStream_t stream )
{
    if( nBlocks <= 0 ) return 0;
#define FN2_TRY( DMV, FTV ) \
    if( ( dimMask & ~DMV ) == 0 && ( featMask & ~FTV ) == 0 ) \
    { \
        hipLaunchKernelGGL( ( fn2_generate_batch<256, DMV, FTV> ), dim3( nBlocks ), dim3( 256 ), 0, stream, dJobs, dBlockJob, dBlockBase ); \
        return (int)hipGetLastError(); \
    }
    FN2_VARIANTS( FN2_TRY )
#undef FN2_TRY
    return (int)hipErrorInvalidValue;
}

int fn2_launch_minmax_init( int32_t* keys, int n, hipStream_t stream )
{
    if( n <= 0 ) return 0;
    hipLaunchKernelGGL( fn2_minmax_init, dim3( ( n + 63 ) / 64 ), dim3( 64 ), 0, stream, keys, n );
    return (int)hipGetLastError();
}

}
