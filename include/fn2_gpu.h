/*
 * Extensions of the MI355X FastNoise2 backend beyond FastNoise2's C ABI.
 *
 * These entry points have no counterpart in the reference's
 * include/FastNoise/FastNoise_C.h. They exist for callers that keep their
 * data on the GPU: generation into device buffers on a caller stream,
 * device memory helpers, timing, program introspection and the
 * continuous-batching serving engine. Plain C types only.
 */
#ifndef FN2_GPU_H
#define FN2_GPU_H

#include <stddef.h>
#include <stdint.h>
#include "FastNoise/Utility/Export.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Where the points of one generation call come from. */
enum fn2SourceKind
{
    FN2_SOURCE_GRID = 0,        /* GenUniformGrid2D/3D/4D */
    FN2_SOURCE_POSITIONS = 1,   /* GenPositionArray2D/3D/4D: pos[] are device pointers */
    FN2_SOURCE_TILEABLE = 2,    /* GenTileable2D (count[0], count[1] = sizes) */
};

/* One generation call in device form. Grid: count[]/offset[]/step[] per axis.
 * Positions: count[0] points, pos[axis] device arrays, offset[] added. */
typedef struct fn2GenSpec
{
    int kind;
    int dims;
    int count[4];
    float offset[4];
    float step[4];
    const float* pos[4];
    int seed;
} fn2GenSpec;

/* Generates into device memory dOut on `stream` (a hipStream_t, NULL = the
 * library stream). outputMinMax: host float[2] (the call then waits for the
 * result) or NULL (fully asynchronous). Returns 0 or a nonzero error code;
 * fn2GetLastError() describes it. */
FASTNOISE_API int fn2GenerateDevice( const void* node, const fn2GenSpec* spec, float* dOut, float* outputMinMax, void* stream );

/* fn2GenerateDevice for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, hosts that forbid hipMalloc
 * after start-up. No counterpart in the reference.
 *
 * The call allocates no device memory: the compiled program, the range
 * record and GenTileable2D's torus table go into one device array of
 * fn2NoAllocScratchBytes() bytes that each GPU holds until the process ends.
 * It runs the tree's specialised kernel when earlier calls have loaded it
 * (it never compiles or loads one itself), else the interpreter kernel: the
 * same bits either way.
 * Limits: at most FN2_NOALLOC_MAX_CONSTANTS constants in the compiled tree
 * and xSize + ySize <= FN2_NOALLOC_MAX_TILE_AXES for tileable specs; larger
 * requests fail with an error and launch nothing. Positions must be device
 * arrays. stream: a hipStream_t, or NULL for the calling thread's default
 * stream (the call then waits for its work).
 *
 * Every call on a GPU rewrites that GPU's array: calls that overlap on one
 * GPU race and must be serialised (one stream, or synchronised between).
 * Returns 0 or a nonzero error code; fn2GetLastError() describes it. */
#define FN2_NOALLOC_MAX_CONSTANTS 16384
#define FN2_NOALLOC_MAX_TILE_AXES 32768
FASTNOISE_API int fn2GenerateDeviceNoAlloc( const void* node, const fn2GenSpec* spec, float* dOut, float* outputMinMax, void* stream );
/* Bytes of the per-GPU array fn2GenerateDeviceNoAlloc works in. */
FASTNOISE_API size_t fn2NoAllocScratchBytes( void );

/* Times `iters` back-to-back generator launches (each writes all points of
 * `spec` to dOut and folds them into the call's min/max) after `warmup`
 * launches, with HIP events on the launch stream. msPerLaunch: average event
 * time per launch; wallMsPerLaunch: host wall time per launch from the first
 * launch to the completion of the last. */
FASTNOISE_API int fn2BenchGenerate( const void* node, const fn2GenSpec* spec, float* dOut, int warmup, int iters, double* msPerLaunch, double* wallMsPerLaunch );

/* Device memory helpers (hipMalloc / hipFree / hipMemcpy, kind: 0 H2D, 1 D2H, 2 D2D). */
FASTNOISE_API void* fn2Malloc( size_t bytes );
FASTNOISE_API void fn2Free( void* p );
/* Pinned (page-locked) host memory (hipHostMalloc / hipHostFree): outputs
 * in it are written by DMA with no staging copy. */
FASTNOISE_API void* fn2MallocHost( size_t bytes );
FASTNOISE_API void fn2FreeHost( void* p );
FASTNOISE_API int fn2Memcpy( void* dst, const void* src, size_t bytes, int kind );
FASTNOISE_API int fn2DeviceSynchronize( void );
FASTNOISE_API int fn2DeviceCount( void );

/* Last error of this thread ("" when none). */
FASTNOISE_API const char* fn2GetLastError( void );

/* Runs one of the library's device math routines over dIn[0..n) into dOut
 * (device arrays): 0 sqrt, 1 reciprocal, 2 1/sqrt (the fast correctly rounded
 * forms the cellular nodes use for positive normal arguments), 3 the IEEE
 * 1/sqrt, 4 the bare hardware square root estimate, 5 the square root that
 * corrects that estimate upward only (3D / 4D cellular jitter lengths). Lets
 * a test compare them with IEEE results on the installed GPU. */
FASTNOISE_API int fn2SelfTestMath( int op, const float* dIn, float* dOut, int n );

/* Run-time specialisation of trees (hipRTC): a tree used repeatedly gets a
 * kernel in which its instructions are compile-time constants (bit-identical
 * results, no interpretation). OFF: interpreter only. ASYNC (default):
 * compile in the background on first use, interpreter meanwhile. SYNC:
 * compile on first use. The FN2_JIT environment variable (off / async /
 * sync) sets the initial mode. fn2SetJitMode returns the previous mode.
 * Programs longer than FN2_JIT_MAX_INSTRUCTIONS (environment, default 512)
 * stay on the interpreter: their compile time grows to tens of seconds. */
enum fn2JitMode
{
    FN2_JIT_OFF = 0,
    FN2_JIT_ASYNC = 1,
    FN2_JIT_SYNC = 2,
};
FASTNOISE_API int fn2SetJitMode( int mode );
/* Writes and compiles (hipRTC, no GPU needed) a tree's specialised kernel
 * without loading it: 0 on success (codeBytes = size of the code object),
 * nonzero with the compiler log in fn2GetLastError otherwise. */
FASTNOISE_API int fn2JitCompileCheck( const void* node, int dims, int* codeBytes );
/* Copies the specialised kernel's HIP source for a tree into buf (at most
 * bufSize bytes, NUL-terminated) and returns the source's full length, or -1
 * when the tree does not compile. */
FASTNOISE_API int fn2JitSource( const void* node, int dims, char* buf, int bufSize );
/* Waits until every queued specialisation has finished compiling. */
FASTNOISE_API int fn2JitWait( void );
/* State of a tree's specialised kernel in `dims` dimensions: -1 never
 * requested, 0 queued, 1 ready, 2 failed (fn2GetLastError holds the
 * compiler log), 3 kept on the interpreter (over the instruction limit),
 * -2 bad arguments. */
FASTNOISE_API int fn2GetJitState( const void* node, int dims );
/* The same for a variant of the specialised kernel (same results, other
 * compile-time choices): bits 0-1 the fused fractals' octave-loop unroll
 * (0 the compiler's choice, 1 none, 2 by two, 3 by four), bit 2 the shared
 * far-cell walk that position-array requests of 3D CellularDistance trees use. */
#define FN2_JIT_VARIANT_SHARED_CELL_WALK 4
FASTNOISE_API int fn2GetJitVariantState( const void* node, int dims, int variant );
/* Selects the octave-unroll variant (0..3 as above) that single calls of a
 * tree in `dims` dimensions run once it is compiled (the default kernel
 * meanwhile); the serving engine's tuner picks its own. Same results for
 * every variant. 0 on success. */
FASTNOISE_API int fn2SetJitVariant( const void* node, int dims, int variant );
/* Resources of a loaded specialised kernel (single-call entry point): VGPRs
 * per lane and scratch bytes per lane as the runtime reports them, and its
 * resident 256-thread blocks per CU. 0, 1 when that variant is not loaded,
 * -1 on bad arguments. */
FASTNOISE_API int fn2GetJitKernelInfo( const void* node, int dims, int variant, int* vgprs, int* scratchBytes, int* blocksPerCU );

/* Compiled-program facts for a tree evaluated in `dims` dimensions. */
FASTNOISE_API int fn2GetProgramInfo( const void* node, int dims, int* instructions, int* registers, int* dimMask, int* featureMask );

/* ------------------------------------------------------------------------
 * Serving engine: continuous batching of chunk requests.
 * ------------------------------------------------------------------------ */

/* Batching policies */
enum fn2ServeMode
{
    FN2_SERVE_PLAIN = 0,    /* one launch + copy per request (the per-chunk baseline) */
    FN2_SERVE_BATCHED = 1,  /* requests packed into one launch per batch, overlapped with host work */
    FN2_SERVE_GRAPH = 2,    /* batched, each batch replayed through a captured hipGraph */
    FN2_SERVE_PERSISTENT = 3, /* each batch through work-queue kernels: resident blocks pull tiles from a device counter */
};

/* Creates an engine. hbmBudgetBytes bounds the device bytes held by queued and
 * running requests (outputs + positions); maxBatchPoints bounds the points of
 * one batch. Returns NULL on failure. */
FASTNOISE_API void* fn2ServeCreate( size_t hbmBudgetBytes, int maxBatchPoints, int mode );
FASTNOISE_API void fn2ServeDestroy( void* engine );

/* Registers a node tree (encoded string, FastNoise2 wire format) and returns
 * its id (>= 0) or -1. Trees stay resident until the engine is destroyed. */
FASTNOISE_API int fn2ServeAddTree( void* engine, const char* encodedTree );

/* Submits a request: tree id, generation spec (grid or tileable; host
 * positions are not accepted), output buffer of the spec's point count.
 * The output may be pageable host memory (staged through pinned memory),
 * pinned host memory (filled straight from the device) or device memory
 * (a device-to-device copy). A request larger than one batch
 * (maxBatchPoints) runs as several parts in successive batches and completes
 * when the last part has. Returns a request id (>= 0) or -1 on bad arguments.
 * The call does not block on the GPU. */
FASTNOISE_API int64_t fn2ServeSubmit( void* engine, int treeId, const fn2GenSpec* spec, float* hostOut );

/* Waits for a request; fills its output buffer and min/max. latencyMs:
 * submit-to-completion time. Returns 0 or an error code. */
FASTNOISE_API int fn2ServeWait( void* engine, int64_t requestId, float* outputMinMax, double* latencyMs );

/* Waits for everything submitted so far. */
FASTNOISE_API int fn2ServeDrain( void* engine );
/* Non-blocking form of fn2ServeWait: 1 when the request is done (it is then
 * retired and its range and latency written), 0 while it is pending, -1 on
 * failure or an unknown id. */
FASTNOISE_API int fn2ServePoll( void* engine, int64_t requestId, float* outputMinMax, double* latencyMs );

/* Engine counters: batches launched, requests completed, device bytes held. */
FASTNOISE_API int fn2ServeStats( void* engine, int64_t* batches, int64_t* completed, int64_t* peakBytes );

/* Lets the tuner also time the octave-unroll variants of trees with fused
 * fractals (compiled in the background once the tree's tile is tuned on its
 * specialised kernel; fn2ServeTreeTile reports the choice). Off by default:
 * on the mixed trees it found 2-8 % on some trees and no throughput overall,
 * for three more compilations per tree and dimensionality. 0 on success. */
FASTNOISE_API int fn2ServeTuneOctaveUnroll( void* engine, int enable );

/* Persistent mode: work-queue launches so far, how many of them had more
 * items than their resident waves (so that waves pulled further items from
 * the launch's counters), and the pulls made (counter increments). */
FASTNOISE_API int fn2ServeQueueStats( void* engine, int64_t* launches, int64_t* pulledLaunches, int64_t* pulls );

/* Points per lane the engine's tuner chose for a tree's batched launches in
 * `dims` dimensions (1, 2, 4 or 8; 0 = not tuned yet, -1 = bad arguments). The
 * tuner times a 2^20-point request of the tree at each candidate on the
 * first request of that tree and dimensionality. */
FASTNOISE_API int fn2ServeTreeTuning( void* engine, int treeId, int dims );
/* The whole tile the tuner chose for a tree in `dims` dimensions: threads per
 * block of its specialised batch kernel (64, 128 or 256; the interpreter and
 * the persistent mode use 256), points per lane, and the unroll factor of
 * its fused fractals' octave loops (0 = the compiler's choice, else 1, 2 or
 * 4; tried once the variants have compiled in the background). Every tile
 * gives the same bits. Returns 0, 1 when not tuned yet, -1 on bad
 * arguments. */
FASTNOISE_API int fn2ServeTreeTile( void* engine, int treeId, int dims, int* threadsPerBlock, int* pointsPerLane, int* octaveUnroll );

/* Replays a request trace against the engine from the calling thread:
 * request i (tree treeIds[i], spec specs[i]) is submitted at arrivalSec[i]
 * seconds after the start (open loop; NULL = as fast as possible), with at
 * most maxOutstanding requests in flight (closed-loop bound). Outputs go to a
 * pool of pinned host buffers, which the engine fills straight from the
 * device, or with deviceOutputs != 0 to a pool of device buffers (consumers
 * on the GPU, no PCIe copy). latencyMs[i] receives request i's submit-to-done
 * time, wallSec the time from the start to the last completion. */
FASTNOISE_API int fn2ServeReplay( void* engine, int count, const int* treeIds, const fn2GenSpec* specs, const double* arrivalSec,
                                  int maxOutstanding, int deviceOutputs, double* latencyMs, double* wallSec );

/* ------------------------------------------------------------------------
 * Chunked-world client: streams a world of (size + 2)^3 density blocks
 * around a moving camera through a serving engine, nearest chunks first, and
 * consumes each block on the GPU (range culling against an iso surface, then
 * a count of the exposed voxel faces a bloxel mesher would emit). Only the
 * per-chunk summary reaches the host. Destroy the world before its engine.
 * ------------------------------------------------------------------------ */
enum fn2ChunkState
{
    FN2_CHUNK_AIR = 0,      /* every value above the iso surface */
    FN2_CHUNK_SOLID = 1,    /* every value below it */
    FN2_CHUNK_SURFACE = 2,  /* straddles it: faces counted */
};

typedef struct fn2ChunkResult
{
    int32_t x, y, z;        /* chunk origin (a multiple of the chunk size) */
    int32_t state;          /* fn2ChunkState */
    float min, max;         /* value range of the whole block, margins included */
    uint32_t quads;         /* solid voxels' faces toward non-solid neighbours */
    uint32_t solidVoxels;   /* voxels at or below the iso surface (interior) */
    double latencyMs;       /* request to consumer done */
    uint32_t version;       /* generation the chunk belongs to */
} fn2ChunkResult;

/* Chunk origins within loadRange of a camera, nearest first (three int32 per
 * chunk, at most maxChunks written); returns the full count or -1. */
FASTNOISE_API int fn2WorldChunksInRange( const float* camera, int chunkSize, float loadRange, int32_t* chunkOrigins, int maxChunks );
/* A world over tree treeId of the engine, with at most maxInFlight chunks
 * requested at once (each holds a device density block); chunkSize 1..894
 * (the face count of a block stays within 32 bits). */
FASTNOISE_API void* fn2WorldCreate( void* engine, int treeId, int chunkSize, float isoSurface, float loadRange, int seed, int maxInFlight );
FASTNOISE_API void fn2WorldDestroy( void* world );
/* Unloads chunks beyond 1.1 x the load range and requests the missing ones in
 * range, nearest first, while blocks are free; returns the number requested. */
FASTNOISE_API int fn2WorldUpdate( void* world, float cameraX, float cameraY, float cameraZ );
/* Writes up to maxResults finished chunks (oldest first) and returns their
 * number; with waitAll it waits for every chunk in flight. Chunks of a
 * superseded version are dropped, never returned. */
FASTNOISE_API int fn2WorldPoll( void* world, fn2ChunkResult* results, int maxResults, int waitAll );
/* Starts a new generation (a teleport or a changed tree): chunks in flight
 * are dropped when they finish and every chunk is requested again. */
FASTNOISE_API uint32_t fn2WorldNewVersion( void* world );
FASTNOISE_API int fn2WorldStats( void* world, int64_t* requested, int64_t* completed, int64_t* dropped, int* loaded, int* inFlight );

#ifdef __cplusplus
}
#endif

#endif
