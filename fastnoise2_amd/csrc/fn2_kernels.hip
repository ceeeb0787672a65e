// HIP kernels for gfx950: the node-tree program interpreter and the output
// range reduction.
//
// One lane evaluates one point at a time: it forms the point's position
// (uniform grid, position arrays or the tileable 4D torus of
// Generator.inl:79-433), runs the wave-uniform instruction list of the
// compiled tree on a 32-entry register file held in VGPRs (uniform indices ->
// GPR indexing, no scratch) and writes one float with a coalesced store.
//
// Output range: every lane keeps a running min/max in registers; a block
// reduces them through wave shuffles and LDS and publishes one atomic pair
// to the slot of its index in the call's range record (fn2_kernel_args.h).
// Same-address atomics queue one behind the other, so the slots spread them;
// the number of blocks still bounds their cost: the single-call kernel runs a
// grid-stride loop over a grid sized to fill the GPU (a few thousand
// blocks), and the batched kernel gives each block a tile of several points
// per lane.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include "fn2_kernel_body.h"
#include "fn2_kernel_args.h"
#include "fn2_launch.h"

using namespace fn2d;
using namespace fn2i;

// The interpreter as a program: runs the request's instruction list.
template<int DM, int FT>
struct Interpreted
{
    static constexpr bool kFresh = false;   // indexes the register file at run time (kept in LDS)
    static constexpr int kDims = 0;         // a.dims at run time
    static constexpr bool kBricks = false;  // row-major walk
    template<class RF>
    __device__ __forceinline__ float operator()( const fn2::GenArgs& g, RF R ) const { return runProgram<DM, FT>( g, R ); }
};

// Single call of the interpreter: grid-stride loop over all points of one request.
template<int BLOCK, int DM, int FT>
__global__ void __launch_bounds__( BLOCK ) fn2_generate( fn2::GenArgs a )
{
    fn2k::generateSingle<BLOCK>( a, Interpreted<DM, FT>() );
}

// Batched form for the serving layer: every block looks up its request in a
// per-block table, so requests with different trees, sizes and sources run in
// one launch (points per lane chosen per tree by the serving engine's tuner).
template<int BLOCK, int DM, int FT>
__global__ void __launch_bounds__( BLOCK ) fn2_generate_batch( const fn2::GenArgs* __restrict__ jobs, const int32_t* __restrict__ blockJob, const uint32_t* __restrict__ blockBase )
{
    fn2k::generateBatch<BLOCK>( jobs, blockJob, blockBase, Interpreted<DM, FT>() );
}

// Persistent form for the serving layer: one launch of resident blocks per
// batch pulls the batch's tiles (any tree, size and source) from a device
// counter that is zero when the launch starts.
template<int BLOCK, int DM, int FT>
__global__ void __launch_bounds__( BLOCK ) fn2_generate_queue( const fn2::GenArgs* __restrict__ jobs, const int32_t* __restrict__ tileJob,
                                                              const uint32_t* __restrict__ tileBase, uint32_t nTiles, uint32_t* counter )
{
    fn2k::generateQueue<BLOCK>( jobs, tileJob, tileBase, nTiles, counter, Interpreted<DM, FT>() );
}

__global__ void fn2_selftest_math( int op, const float* in, float* out, int n )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < n )
    {
        float x = in[i];
        out[i] = op == 0 ? sqrtPosRN( x ) : op == 1 ? recipPosRN( x ) : op == 2 ? invSqrtPos( x ) : op == 3 ? invSqrt( x )
               : op == 4 ? __builtin_amdgcn_sqrtf( x ) : sqrtPosRNFromBelow( x );
    }
}

// Consumer of a chunk's (s + 2)^3 density block (x fastest): over the s^3
// interior voxels, counts the solid ones (value <= iso) and their faces
// toward a non-solid neighbour (value > iso), the quads of a bloxel mesh
// (MeshNoisePreview.cpp:457-520 of the reference). counts = { quads, solid }.
__global__ void __launch_bounds__( 256 ) fn2_bloxel_faces( const float* __restrict__ d, int g, float iso, uint32_t* counts )
{
    const uint32_t s = (uint32_t)g - 2, total = s * s * s;
    const uint32_t gg = (uint32_t)g * (uint32_t)g;
    uint32_t quads = 0, solid = 0;
    for( uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256 )
    {
        const uint32_t x = i % s, y = ( i / s ) % s, z = i / ( s * s );
        const uint32_t idx = ( z + 1 ) * gg + ( y + 1 ) * (uint32_t)g + ( x + 1 );
        if( d[idx] <= iso )
        {
            solid++;
            quads += ( d[idx + 1] > iso ) + ( d[idx - 1] > iso ) + ( d[idx + g] > iso ) + ( d[idx - g] > iso ) +
                     ( d[idx + gg] > iso ) + ( d[idx - gg] > iso );
        }
    }
    __shared__ uint32_t sq[4], ss[4];
#pragma unroll
    for( int off = 32; off > 0; off >>= 1 )
    {
        quads += __shfl_xor( quads, off );
        solid += __shfl_xor( solid, off );
    }
    if( ( threadIdx.x & 63 ) == 0 )
    {
        sq[threadIdx.x >> 6] = quads;
        ss[threadIdx.x >> 6] = solid;
    }
    __syncthreads();
    if( threadIdx.x == 0 )
    {
        atomicAdd( &counts[0], sq[0] + sq[1] + sq[2] + sq[3] );
        atomicAdd( &counts[1], ss[0] + ss[1] + ss[2] + ss[3] );
    }
}

// Resets n consecutive range records (`slots` slots each, `stride` ints
// apart) and zeroes nc work-queue counters
__global__ void fn2_minmax_init( int32_t* keys, int n, int slots, int stride, uint32_t* counters, int nc )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < n * slots )
    {
        int32_t* k = keys + i * stride;   // slot i % slots of record i / slots
        k[0] = 0x7fffffff;          // key of +NaN region: above every float
        k[1] = (int32_t)0x80000000u;
    }
    if( i < nc ) counters[i] = 0;
}

namespace
{
    constexpr int kBlock = 256;

    // Facts measured once per process; threads that race to measure one
    // store the same value.
    int CUs()
    {
        static std::atomic<int> cus{ 0 };
        int n = cus.load( std::memory_order_relaxed );
        if( n == 0 )
        {
            int dev = 0;
            if( hipGetDevice( &dev ) != hipSuccess || hipDeviceGetAttribute( &n, hipDeviceAttributeMultiprocessorCount, dev ) != hipSuccess || n <= 0 )
            {
                (void)hipGetLastError();
                n = 256;
            }
            cus.store( n, std::memory_order_relaxed );
        }
        return n;
    }

    // Blocks of a single-call (grid-stride) kernel. With the kernel's resident
    // blocks per CU known: up to 4 rounds of what the GPU holds at once, one
    // point per thread below that (measured on MI355X: 4 rounds 2-5 % faster
    // than 2 on the cellular and serving trees, the same on FBm; with the
    // 64-slot range record, one point per thread against at least two: the
    // 1 M-point tileable call 0.0332 -> 0.0327 ms, the 262 k-point Perlin
    // grid and large calls unchanged; at least four: tileable 0.0344 ms,
    // Perlin 4.9 against 4.4 us). Without it, 16 blocks per CU.
    unsigned GridFor( int64_t total, int blocksPerCU = 0, int rounds = 4 )
    {
        if( blocksPerCU <= 0 )
        {
            const int64_t need = ( total + kBlock - 1 ) / kBlock;
            const int64_t cap = (int64_t)CUs() * 16;
            return (unsigned)( need < cap ? need : cap );
        }
        const int64_t need = std::max<int64_t>( 1, ( total + kBlock - 1 ) / kBlock );
        const int64_t cap = (int64_t)CUs() * blocksPerCU * rounds;
        return (unsigned)( need < cap ? need : cap );
    }

    // Resident blocks per CU of an interpreter variant at a register-file size
    template<class K>
    int ResidentBlocks( K kernel, size_t lds )
    {
        int n = 0;
        if( hipOccupancyMaxActiveBlocksPerMultiprocessor( &n, kernel, kBlock, lds ) != hipSuccess )
        {
            (void)hipGetLastError();
            return 0;
        }
        return n;
    }

    // Resident blocks per CU of a kernel at a register-file size, from a
    // per-kernel cache indexed by the size
    template<class K>
    int CachedResident( std::atomic<int>* cache, int regs, K kernel, size_t lds )
    {
        int r = cache[regs].load( std::memory_order_relaxed );
        if( r == 0 )
        {
            r = std::max( 1, ResidentBlocks( kernel, lds ) );
            cache[regs].store( r, std::memory_order_relaxed );
        }
        return r;
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

// blocks of a single-call launch over `total` points (also used for the
// run-time specialised kernels)
unsigned fn2_grid_for( int64_t total, int blocksPerCU )
{
    return GridFor( total, blocksPerCU );
}

// threads per block of the batched kernel: a job with ppt points per lane
// covers kBlock * ppt points per block
int fn2_batch_block_threads( void )
{
    return kBlock;
}

int fn2_launch_generate( const fn2::GenArgs* args, int dimMask, int featMask, hipStream_t stream )
{
    if( args->total <= 0 ) return 0;
    // the register file in LDS: one column of args->regs floats per thread
    if( args->regs < 1 || args->regs > fn2::kMaxRegs ) return (int)hipErrorInvalidValue;
    const size_t lds = sizeof( float ) * args->regs * kBlock;
#define FN2_TRY( DMV, FTV ) \
    if( ( dimMask & ~DMV ) == 0 && ( featMask & ~FTV ) == 0 ) \
    { \
        static std::atomic<int> resident[fn2::kMaxRegs + 1];   /* per register count, zero at start */ \
        const int bpc = CachedResident( resident, args->regs, fn2_generate<kBlock, DMV, FTV>, lds ); \
        const unsigned grid = GridFor( args->total, bpc, 4 ); \
        hipLaunchKernelGGL( ( fn2_generate<kBlock, DMV, FTV> ), dim3( grid ), dim3( kBlock ), lds, stream, *args ); \
        return (int)hipGetLastError(); \
    }
    FN2_VARIANTS( FN2_TRY )
#undef FN2_TRY
    return (int)hipErrorInvalidValue;
}

// regs: the largest register count of the jobs in the launch
int fn2_launch_generate_batch( const fn2::GenArgs* dJobs, const int32_t* dBlockJob, const uint32_t* dBlockBase, int nBlocks, int dimMask, int featMask,
                               int regs, hipStream_t stream )
{
    if( nBlocks <= 0 ) return 0;
    if( regs < 1 || regs > fn2::kMaxRegs ) return (int)hipErrorInvalidValue;
    const size_t lds = sizeof( float ) * regs * kBlock;
#define FN2_TRY( DMV, FTV ) \
    if( ( dimMask & ~DMV ) == 0 && ( featMask & ~FTV ) == 0 ) \
    { \
        hipLaunchKernelGGL( ( fn2_generate_batch<kBlock, DMV, FTV> ), dim3( nBlocks ), dim3( kBlock ), lds, stream, dJobs, dBlockJob, dBlockBase ); \
        return (int)hipGetLastError(); \
    }
    FN2_VARIANTS( FN2_TRY )
#undef FN2_TRY
    return (int)hipErrorInvalidValue;
}

// The tiles of a batch through the work-queue kernel: a grid of at most the
// resident blocks of the variant (every block is on a CU from the start and
// pulls tiles until none is left). regs: the largest register count of the
// jobs; counter: kQueueCounterInts device words, zero when the launch starts
// (the serving engine zeroes them with the batch's range keys,
// fn2_launch_batch_init).
int fn2_launch_generate_queue( const fn2::GenArgs* dJobs, const int32_t* dTileJob, const uint32_t* dTileBase, int nTiles, int dimMask, int featMask,
                               int regs, uint32_t* dCounter, hipStream_t stream )
{
    if( nTiles <= 0 ) return 0;
    if( regs < 1 || regs > fn2::kMaxRegs || !dCounter ) return (int)hipErrorInvalidValue;
    const size_t lds = sizeof( float ) * regs * kBlock;
#define FN2_TRY( DMV, FTV ) \
    if( ( dimMask & ~DMV ) == 0 && ( featMask & ~FTV ) == 0 ) \
    { \
        static std::atomic<int> resident[fn2::kMaxRegs + 1];   /* per register count, zero at start */ \
        const int bpc = CachedResident( resident, regs, fn2_generate_queue<kBlock, DMV, FTV>, lds ); \
        const unsigned grid = (unsigned)std::min<int64_t>( nTiles, (int64_t)CUs() * bpc ); \
        hipLaunchKernelGGL( ( fn2_generate_queue<kBlock, DMV, FTV> ), dim3( grid ), dim3( kBlock ), lds, stream, dJobs, dTileJob, dTileBase, \
                            (uint32_t)nTiles, dCounter ); \
        return (int)hipGetLastError(); \
    }
    FN2_VARIANTS( FN2_TRY )
#undef FN2_TRY
    return (int)hipErrorInvalidValue;
}

int fn2_launch_selftest_math( int op, const float* in, float* out, int n, hipStream_t stream )
{
    if( n <= 0 ) return 0;
    hipLaunchKernelGGL( fn2_selftest_math, dim3( ( n + 255 ) / 256 ), dim3( 256 ), 0, stream, op, in, out, n );
    return (int)hipGetLastError();
}

int fn2_launch_bloxel_faces( const float* density, int gen, float iso, uint32_t* counts, hipStream_t stream )
{
    // the face count of a block reaches 6 (gen - 2)^3, counted in uint32:
    // it stays below 2^32 for gen <= 896 (6 * 894^3 = 4,287,101,904), which
    // also keeps every voxel and neighbour index inside the gen^3 block
    if( !density || !counts || gen < 3 || gen > 896 ) return (int)hipErrorInvalidValue;
    const int64_t s = gen - 2, total = s * s * s;
    // at most a block per CU: every block adds into the same two counters,
    // and those atomics queue one behind the other (a 64^3 chunk would
    // otherwise take 1024 blocks)
    const unsigned blocks = std::max( 1u, std::min<unsigned>( GridFor( total ), (unsigned)CUs() ) );
    hipLaunchKernelGGL( fn2_bloxel_faces, dim3( blocks ), dim3( 256 ), 0, stream, density, gen, iso, counts );
    return (int)hipGetLastError();
}

int fn2_launch_minmax_init( int32_t* keys, int n, int slots, int stride, hipStream_t stream )
{
    if( n <= 0 ) return 0;
    if( stride < 2 || slots < 1 || ( slots & ( slots - 1 ) ) != 0 ) return (int)hipErrorInvalidValue;
    const int m = n * slots;
    hipLaunchKernelGGL( fn2_minmax_init, dim3( ( m + 63 ) / 64 ), dim3( 64 ), 0, stream, keys, n, slots, stride, (uint32_t*)nullptr, 0 );
    return (int)hipGetLastError();
}

// The same, also zeroing nc work-queue counters (one launch per batch)
int fn2_launch_batch_init( int32_t* keys, int n, int slots, int stride, uint32_t* counters, int nc, hipStream_t stream )
{
    if( n < 0 || nc < 0 || stride < 2 || slots < 1 || ( slots & ( slots - 1 ) ) != 0 || ( nc > 0 && !counters ) )
        return (int)hipErrorInvalidValue;
    const int m = std::max( n * slots, nc );
    if( m == 0 ) return 0;
    hipLaunchKernelGGL( fn2_minmax_init, dim3( ( m + 63 ) / 64 ), dim3( 64 ), 0, stream, keys, n, slots, stride, counters, nc );
    return (int)hipGetLastError();
}

}
