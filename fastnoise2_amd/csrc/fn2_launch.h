// Host-side launchers of the kernels in fn2_kernels.hip (C linkage; each
// returns a hipError_t). fn2_kernels.hip includes this header too, so a
// definition that drifts from its declaration fails to compile.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include "fn2_kernel_args.h"

extern "C" {
// one interpreter launch over args->total points
int fn2_launch_generate( const fn2::GenArgs* args, int dimMask, int featMask, hipStream_t stream );
// blocks of a single-call launch over `total` points (also used for the
// run-time specialised kernels)
unsigned fn2_grid_for( int64_t total, int blocksPerCU );
// interpreter launch over a batch of jobs (block b runs job dBlockJob[b])
int fn2_launch_generate_batch( const fn2::GenArgs* dJobs, const int32_t* dBlockJob, const uint32_t* dBlockBase, int nBlocks, int dimMask, int featMask,
                               int regs, hipStream_t stream );
// interpreter work-queue launch over nTiles tiles of a batch
int fn2_launch_generate_queue( const fn2::GenArgs* dJobs, const int32_t* dTileJob, const uint32_t* dTileBase, int nTiles, int dimMask, int featMask,
                               int regs, uint32_t* dCounter, hipStream_t stream );
// resets n consecutive range records of `slots` slots, `stride` ints apart
int fn2_launch_minmax_init( int32_t* keys, int n, int slots, int stride, hipStream_t stream );
// the same, also zeroing nc work-queue counters
int fn2_launch_batch_init( int32_t* keys, int n, int slots, int stride, uint32_t* counters, int nc, hipStream_t stream );
// threads per block of the batched interpreter kernel
int fn2_batch_block_threads( void );
// device-side checks of the math helpers
int fn2_launch_selftest_math( int op, const float* in, float* out, int n, hipStream_t stream );
// faces of a bloxel mesh of a gen^3 density block
int fn2_launch_bloxel_faces( const float* density, int gen, float iso, uint32_t* counts, hipStream_t stream );
}
