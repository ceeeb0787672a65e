// Serving engine: continuous batching of chunk-generation requests.
//
// Requests (tree, grid or tileable spec, host output buffer) are queued by
// fn2ServeSubmit from any thread. A dispatcher thread admits queued requests
// FIFO into the next free batch slot as long as the slot's point capacity
// allows, writes one job record per request (GenArgs pointing at the
// request's output range in the slot's device arena and at its own min/max
// key pair) plus a block -> (job, first point) table, and launches the batch
// kernel once per kernel variant present in the batch, followed by the copy
// of the slot's outputs and keys into pinned host memory. Slots run on their
// own streams, so a batch formed from requests that arrived while the
// previous one was on the GPU overlaps with it. A completion thread retires
// slots in launch order: it copies each request's values into the caller's
// buffer, decodes its range and wakes the waiters.
//
// Device memory: slots * (capacity * 4 bytes + tables) is allocated once at
// creation and must fit the HBM budget given there; requests wait in the
// host queue until a slot frees (admission). Trees are compiled once per
// dimensionality and stay resident.
//
// FN2_SERVE_PLAIN runs the same requests one launch + copy + synchronise at a
// time (the per-chunk API path) for comparison. FN2_SERVE_GRAPH records each
// batch's table upload, launches and copies into a HIP graph and replays the
// graph when the next batch has the same shape (same job count, block count
// and variant groups); job records and tables change, the topology does not.
// FN2_SERVE_PERSISTENT runs each batch through work-queue kernels: a grid of
// resident blocks pulls tiles from a device counter, one launch per
// specialised kernel in the batch and one interpreter launch for all other
// trees together.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fn2_backend.h"
#include "fn2_copy_pool.h"
#include "fn2_internal.h"
#include "fn2_jit.h"

#include "fn2_launch.h"

namespace
{
    using Clock = std::chrono::steady_clock;
    using fn2::GenArgs;

    // threads per block of the batched kernel; requests start on multiples of it
    const int kThreads = fn2_batch_block_threads();
    // the tuner's range record: compact
    constexpr int kTuneRangeInts = fn2::kRangeSlots * fn2::kRangeStrideCompact;
    // the tuner's candidates: points per lane, threads per block (the
    // specialised batch kernel; the interpreter's register file in LDS is laid
    // out for kThreads), and its calibration size
    constexpr int kPptCandidates[] = { 1, 2, 4, 8 };
    constexpr int kThreadCandidates[] = { 64, 128, 256 };
    constexpr int kTunePoints = 1 << 20;
    constexpr int kSlots = 3;
    constexpr int kMaxJobs = 4096;

    struct Tree
    {
        FastNoise::SmartNode<> node;
        const fn2::DeviceProgram* prog[5] = {};
        std::string compileError[5];
        // the tuned tile of the batched kernel, per dimensionality: points per
        // lane, threads per block, and the variant of the specialised kernel
        // (its octave-loop unroll, fn2_jit.h)
        int ppt[5] = {};
        int threads[5] = { 256, 256, 256, 256, 256 };
        int variant[5] = {};
        bool tunedSpecialised[5] = {};   // tile timed on the tree's specialised batch kernel
        int unrollStage[5] = {};         // 0 not started, 1 variants requested, 2 decided
        float tileMs[5] = {};            // the tuned tile's time on the default specialised kernel
        bool tuning[5] = {};             // a submitter is timing this dimensionality now
    };

    // A tile of the batched kernel: threads per block x points per lane
    struct Tile
    {
        int threads = 256;
        int ppt = 0;
        float ms = 0.0f;   // the tuner's best time for its 2^20-point request
    };

    // points per lane of a request whose tree is being tuned by another
    // submitter (any value gives the same bits)
    constexpr int kUntunedPpt = 2;

    struct Request
    {
        int64_t id = 0;
        int tree = 0;
        fn2GenSpec spec{};
        float* hostOut = nullptr;
        int64_t points = 0;
        const fn2::DeviceProgram* prog = nullptr;
        GenArgs args{};
        Clock::time_point tSubmit, tDone;
        // placement in its batch slot
        size_t outOffset = 0;
        int job = 0;
        int ppt = 1;
        int threads = 256;   // threads per block when a specialised kernel runs it
        int variant = 0;     // variant of the specialised kernel
        const fn2::JitKernel* jit = nullptr;
        // where the values go: through the slot's pinned staging buffer (pageable
        // caller memory), straight from the arena into pinned caller memory, or
        // into caller device memory (a device-to-device copy)
        enum Dest { kStaged, kPinned, kDevice } dest = kStaged;
        // results
        bool done = false;
        int status = 0;
        float mm[2] = { 0, 0 };
        // a part of a request split across batches (its parent owns the id),
        // and a split request's parts still running
        Request* parent = nullptr;
        int pending = 0;
    };

    struct Slot
    {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        float* dOut = nullptr;          // capacity floats
        float* hOut = nullptr;          // pinned
        int32_t* dKeys = nullptr;       // a range record per job, then the persistent mode's work counters (kQueueCounterInts per launch)
        uint32_t* dCounter = nullptr;
        int32_t* hKeys = nullptr;       // pinned
        GenArgs* dJobs = nullptr;
        int32_t* dBlockJob = nullptr;
        uint32_t* dBlockBase = nullptr;
        uint8_t* hTables = nullptr;     // pinned: jobs | blockJob | blockBase
        size_t tableBytes = 0;
        std::vector<Request*> reqs;
        int64_t points = 0;
        bool busy = false;
        // graph mode
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<std::array<int64_t, 4>> graphShape;   // (kernel key, blocks, register file, threads) per launch
        int graphJobs = -1;
        // persistent mode: work-queue launches of the batch in flight, whose
        // counters come back with the range keys
        int queueLaunches = 0;
    };

    size_t AlignUp( size_t v, size_t a ) { return ( v + a - 1 ) / a * a; }

    class Engine
    {
    public:
        int mode;
        int jobStride = fn2::kRangeStrideCompact;   // ints between a request's range slots
        int jobRangeInts = fn2::kRangeSlots * fn2::kRangeStrideCompact;
        int64_t capacity;       // points per batch
        size_t budget;
        std::string initError;

        Engine( size_t hbmBudget, int maxBatchPoints, int m ) : mode( m ), capacity( maxBatchPoints ), budget( hbmBudget )
        {
            // A request's range record: the engine reads one back per
            // request, so compact; the work-queue kernels publish a range per
            // wave (four times as many atomics as a block-per-tile launch),
            // which queue up within one line, so they get a line per slot
            jobStride = mode == FN2_SERVE_PERSISTENT ? fn2::kRangeStrideSpread : fn2::kRangeStrideCompact;
            jobRangeInts = fn2::kRangeSlots * jobStride;
        }

        bool Init()
        {
            if( capacity < kThreads ) capacity = kThreads;
            capacity = (int64_t)AlignUp( (size_t)capacity, kThreads );
            maxBlocks = (int)( capacity / kThreadCandidates[0] ) + kMaxJobs;   // the smallest tiles
            size_t jobsBytes = AlignUp( sizeof( GenArgs ) * kMaxJobs, 256 );
            size_t bjBytes = AlignUp( sizeof( int32_t ) * maxBlocks, 256 );
            size_t bbBytes = AlignUp( sizeof( uint32_t ) * maxBlocks, 256 );
            // range records, then (persistent mode) the counters of up to a launch per job
            const size_t keyBytes = ( jobRangeInts + ( mode == FN2_SERVE_PERSISTENT ? fn2::kQueueCounterInts : 0 ) ) * sizeof( int32_t ) * kMaxJobs;
            size_t perSlot = (size_t)capacity * 4 + jobsBytes + bjBytes + bbBytes + keyBytes;
            int nSlots = mode == FN2_SERVE_PLAIN ? 1 : kSlots;
            if( perSlot * nSlots > budget )
            {
                initError = "HBM budget too small for the batch capacity (needs " + std::to_string( perSlot * nSlots ) + " bytes)";
                return false;
            }
            slots.resize( nSlots );
            for( Slot& s : slots )
            {
                s.tableBytes = jobsBytes + bjBytes + bbBytes;
                if( !fn2::HipOk( hipStreamCreateWithFlags( &s.stream, hipStreamNonBlocking ), "serve stream" ) ||
                    !fn2::HipOk( hipEventCreateWithFlags( &s.done, hipEventDisableTiming ), "serve event" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&s.dOut, (size_t)capacity * 4 ), "serve arena" ) ||
                    !fn2::HipOk( hipHostMalloc( (void**)&s.hOut, (size_t)capacity * 4, hipHostMallocDefault ), "serve staging" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&s.dKeys, keyBytes ), "serve keys" ) ||
                    !fn2::HipOk( hipHostMalloc( (void**)&s.hKeys, keyBytes, hipHostMallocDefault ), "serve keys" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&s.dJobs, s.tableBytes ), "serve tables" ) ||
                    !fn2::HipOk( hipHostMalloc( (void**)&s.hTables, s.tableBytes, hipHostMallocDefault ), "serve tables" ) )
                {
                    initError = fn2::LastError();
                    return false;
                }
                s.dCounter = (uint32_t*)( s.dKeys + (size_t)jobRangeInts * kMaxJobs );
                s.dBlockJob = (int32_t*)( (uint8_t*)s.dJobs + jobsBytes );
                s.dBlockBase = (uint32_t*)( (uint8_t*)s.dJobs + jobsBytes + bjBytes );
            }
            deviceBytes = (int64_t)( perSlot * nSlots );
            {
                int dev = 0, cus = 0;
                if( hipGetDevice( &dev ) == hipSuccess && hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) == hipSuccess && cus > 0 )
                {
                    fillBlocks = 2 * cus;
                    this->cus = cus;
                }
                (void)hipGetLastError();
            }
            if( mode != FN2_SERVE_PLAIN )
            {
                size_t tb = (size_t)kTunePoints / kThreadCandidates[0] + 1;   // blocks of the smallest tile
                if( !fn2::HipOk( hipStreamCreateWithFlags( &tuneStream, hipStreamNonBlocking ), "tune stream" ) ||
                    !fn2::HipOk( hipEventCreate( &tuneEv0 ), "tune event" ) || !fn2::HipOk( hipEventCreate( &tuneEv1 ), "tune event" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&tuneOut, (size_t)kTunePoints * 4 ), "tune arena" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&tuneJob, sizeof( GenArgs ) ), "tune job" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&tuneKeys, kTuneRangeInts * sizeof( int32_t ) ), "tune keys" ) ||
                    !fn2::HipOk( hipMalloc( (void**)&tuneBJ, tb * sizeof( int32_t ) * 2 ), "tune tables" ) )
                {
                    initError = fn2::LastError();
                    return false;
                }
                deviceBytes += (int64_t)( (size_t)kTunePoints * 4 + sizeof( GenArgs ) + kTuneRangeInts * 4 + tb * 8 );
            }
            copies.reset( new fn2::CopyPool( 3 ) );
            dispatcher = std::thread( [this] { DispatchLoop(); } );
            completer = std::thread( [this] { CompleteLoop(); } );
            return true;
        }

        ~Engine()
        {
            {
                std::lock_guard<std::mutex> lk( mu );
                stop = true;
            }
            cvWork.notify_all();
            cvDone.notify_all();
            cvSlot.notify_all();
            if( dispatcher.joinable() ) dispatcher.join();
            if( completer.joinable() ) completer.join();
            for( Slot& s : slots )
            {
                if( s.stream ) (void)hipStreamSynchronize( s.stream );
                if( s.exec ) (void)hipGraphExecDestroy( s.exec );
                if( s.graph ) (void)hipGraphDestroy( s.graph );
                if( s.dOut ) (void)hipFree( s.dOut );
                if( s.hOut ) (void)hipHostFree( s.hOut );
                if( s.dKeys ) (void)hipFree( s.dKeys );
                if( s.hKeys ) (void)hipHostFree( s.hKeys );
                if( s.dJobs ) (void)hipFree( s.dJobs );
                if( s.hTables ) (void)hipHostFree( s.hTables );
                if( s.done ) (void)hipEventDestroy( s.done );
                if( s.stream ) (void)hipStreamDestroy( s.stream );
            }
            for( Request* q : queue )
                if( q->parent ) delete q;   // parts never launched
            for( auto& kv : requests ) delete kv.second;
            if( tuneStream ) (void)hipStreamSynchronize( tuneStream );
            if( tuneOut ) (void)hipFree( tuneOut );
            if( tuneJob ) (void)hipFree( tuneJob );
            if( tuneKeys ) (void)hipFree( tuneKeys );
            if( tuneBJ ) (void)hipFree( tuneBJ );
            if( tuneEv0 ) (void)hipEventDestroy( tuneEv0 );
            if( tuneEv1 ) (void)hipEventDestroy( tuneEv1 );
            if( tuneStream ) (void)hipStreamDestroy( tuneStream );
        }

        int TreePpt( int treeId, int dims )
        {
            std::lock_guard<std::mutex> lk( mu );
            if( treeId < 0 || treeId >= (int)trees.size() || dims < 2 || dims > 4 ) return -1;
            return trees[treeId]->ppt[dims];
        }

        int TreeTile( int treeId, int dims, int* threads, int* ppt, int* unroll )
        {
            std::lock_guard<std::mutex> lk( mu );
            if( treeId < 0 || treeId >= (int)trees.size() || dims < 2 || dims > 4 ) return -1;
            const Tree& t = *trees[treeId];
            if( threads ) *threads = t.threads[dims];
            if( ppt ) *ppt = t.ppt[dims];
            if( unroll ) *unroll = fn2::kJitUnrollFactor[t.variant[dims] & fn2::kJitUnrollMask];
            return t.ppt[dims] ? 0 : 1;
        }

        int AddTree( const char* encoded )
        {
            FastNoise::SmartNode<> node = FastNoise::NewFromEncodedNodeTree( encoded );
            if( !node )
            {
                fn2::SetError( "encoded node tree does not decode" );
                return -1;
            }
            std::lock_guard<std::mutex> lk( mu );
            trees.emplace_back( new Tree );
            trees.back()->node = node;
            return (int)trees.size() - 1;
        }

        int64_t Submit( int treeId, const fn2GenSpec& spec, float* hostOut )
        {
            if( !hostOut )
            {
                fn2::SetError( "null output buffer" );
                return -1;
            }
            if( spec.kind == FN2_SOURCE_POSITIONS )
            {
                fn2::SetError( "the serving engine takes grid and tileable requests" );
                return -1;
            }
            std::unique_ptr<Request> r( new Request );
            r->tree = treeId;
            r->spec = spec;
            r->hostOut = hostOut;
            {
                hipPointerAttribute_t attr;
                if( hipPointerGetAttributes( &attr, hostOut ) == hipSuccess )
                    r->dest = attr.type == hipMemoryTypeHost ? Request::kPinned : attr.type == hipMemoryTypeDevice ? Request::kDevice : Request::kStaged;
                else (void)hipGetLastError();
            }
            std::string err;
            if( !fn2::FillSpecArgs( spec, r->args, r->points, err ) )
            {
                fn2::SetError( err );
                return -1;
            }
            const fn2::DeviceProgram* prog = nullptr;
            {
                std::lock_guard<std::mutex> lk( mu );
                if( treeId < 0 || treeId >= (int)trees.size() )
                {
                    fn2::SetError( "unknown tree id" );
                    return -1;
                }
                Tree& t = *trees[treeId];
                int d = r->args.dims;
                if( !t.prog[d] && t.compileError[d].empty() )
                {
                    if( !fn2::CompileForDevice( t.node.get(), d, t.prog[d], nullptr ) ) t.compileError[d] = fn2::LastError();
                }
                if( !t.prog[d] )
                {
                    fn2::SetError( t.compileError[d] );
                    return -1;
                }
                prog = t.prog[d];
                r->ppt = t.ppt[d];
                r->threads = t.threads[d];
                r->variant = t.variant[d];
            }
            // One submitter per tree and dimensionality tunes (the flag is
            // taken under the lock); others meanwhile run at the current or
            // a default points per lane rather than time the tree again.
            const int d = r->args.dims;
            if( r->ppt == 0 )
            {
                bool mine = false;
                if( mode == FN2_SERVE_PLAIN ) r->ppt = 1;
                else
                {
                    std::lock_guard<std::mutex> lk( mu );
                    mine = !trees[treeId]->tuning[d];
                    trees[treeId]->tuning[d] = true;
                }
                if( mine )
                {
                    Tile best = Tune( prog, d, nullptr, false );
                    std::lock_guard<std::mutex> lk( mu );
                    trees[treeId]->tuning[d] = false;
                    if( best.ppt <= 0 ) return -1;
                    trees[treeId]->ppt[d] = best.ppt;
                    r->ppt = best.ppt;
                }
                else if( r->ppt == 0 )
                {
                    r->ppt = kUntunedPpt;
                }
            }
            if( mode != FN2_SERVE_PLAIN )
            {
                // the first timing ran the interpreter; once the tree's
                // specialised kernel is loaded, time that once more
                // (looked up outside the engine lock: in sync mode the lookup
                // compiles, which must not hold up the dispatcher)
                const fn2::JitKernel* loaded = fn2::GetJitKernel( prog );
                bool retune = false;
                {
                    std::lock_guard<std::mutex> lk( mu );
                    retune = loaded && !trees[treeId]->tunedSpecialised[d] && !trees[treeId]->tuning[d];
                    if( retune ) trees[treeId]->tuning[d] = true;
                }
                const fn2::JitKernel* k = retune ? loaded : nullptr;
                if( k )
                {
                    // the tile over threads per block and points per lane,
                    // except in persistent mode, whose work-queue kernel pulls
                    // quarter tiles of full blocks
                    Tile best = Tune( prog, d, k, mode != FN2_SERVE_PERSISTENT );
                    std::lock_guard<std::mutex> lk( mu );
                    Tree& t = *trees[treeId];
                    t.tuning[d] = false;
                    t.tunedSpecialised[d] = true;
                    if( best.ppt > 0 )
                    {
                        t.ppt[d] = best.ppt;
                        t.threads[d] = best.threads;
                        t.tileMs[d] = best.ms;
                    }
                    // a fused fractal's octave loop (when enabled): compile the
                    // unroll variants in the background, time them once loaded
                    if( tuneUnroll && mode != FN2_SERVE_PERSISTENT && fn2::JitHasFusedFractal( *prog ) ) t.unrollStage[d] = 1;
                    else t.unrollStage[d] = 2;
                    r->ppt = t.ppt[d];
                    r->threads = t.threads[d];
                }
                TuneUnroll( treeId, d, prog, r.get() );
            }
            r->prog = prog;
            r->args.ins = prog->dIns;
            r->args.konst = prog->dKonst;
            r->args.regs = prog->regs;
            r->args.nIns = (int32_t)prog->ins.size();
            r->args.resultReg = prog->resultReg;
            r->tSubmit = Clock::now();
            // A request beyond a batch slot's capacity is split into parts of
            // at most `capacity` points, each a job over points base +
            // [0, part) of the whole grid or image, written to its own range
            // of the destination. The request completes when every part has,
            // with the parts' ranges merged; the parts go through the queue
            // like requests, so batches of other requests interleave.
            std::vector<Request*> parts;
            if( r->points > capacity )
            {
                for( int64_t b = 0; b < r->points; b += capacity )
                {
                    Request* c = new Request( *r );
                    c->parent = r.get();
                    c->points = std::min<int64_t>( capacity, r->points - b );
                    c->args.total = (int32_t)c->points;
                    c->args.base = (uint32_t)b;
                    c->hostOut = r->hostOut + b;
                    parts.push_back( c );
                }
                r->pending = (int)parts.size();
                r->mm[0] = __builtin_inff();
                r->mm[1] = -__builtin_inff();
            }
            Request* raw = r.release();
            {
                std::lock_guard<std::mutex> lk( mu );
                raw->id = nextId++;
                requests[raw->id] = raw;
                if( parts.empty() ) queue.push_back( raw );
                for( Request* c : parts ) queue.push_back( c );
            }
            cvWork.notify_one();
            return raw->id;
        }

        int Wait( int64_t id, float* mm, double* latencyMs )
        {
            std::unique_lock<std::mutex> lk( mu );
            auto it = requests.find( id );
            if( it == requests.end() )
            {
                fn2::SetError( "unknown request id" );
                return 1;
            }
            Request* r = it->second;
            cvDone.wait( lk, [&] { return r->done || stop; } );
            if( !r->done )
            {
                fn2::SetError( "engine stopped" );
                return 1;
            }
            if( mm ) { mm[0] = r->mm[0]; mm[1] = r->mm[1]; }
            if( latencyMs ) *latencyMs = std::chrono::duration<double, std::milli>( r->tDone - r->tSubmit ).count();
            int status = r->status;
            requests.erase( it );
            delete r;
            if( status ) fn2::SetError( "request failed on the device" );
            return status;
        }

        // Non-blocking Wait: 1 done (the request is retired), 0 still pending, -1 failed
        int Poll( int64_t id, float* mm, double* latencyMs )
        {
            {
                std::lock_guard<std::mutex> lk( mu );
                auto it = requests.find( id );
                if( it == requests.end() )
                {
                    fn2::SetError( "unknown request id" );
                    return -1;
                }
                if( !it->second->done ) return 0;
            }
            return Wait( id, mm, latencyMs ) ? -1 : 1;
        }

        int Drain()
        {
            std::unique_lock<std::mutex> lk( mu );
            cvDone.wait( lk, [&] { return ( queue.empty() && inflight == 0 ) || stop; } );
            return 0;
        }

        void Stats( int64_t* b, int64_t* c, int64_t* p )
        {
            std::lock_guard<std::mutex> lk( mu );
            if( b ) *b = batches;
            if( c ) *c = completed;
            if( p ) *p = deviceBytes;
        }

        // The octave-unroll stage of the tuner: off by default. On the mixed
        // trees it found 2-8 % on three of the five trees with fused fractals
        // and nothing on the others, with the same saturated throughput,
        // while it costs three more background compilations per tree and
        // dimensionality.
        void SetTuneUnroll( bool on )
        {
            std::lock_guard<std::mutex> lk( mu );
            tuneUnroll = on;
        }

        void QueueStats( int64_t* launches, int64_t* pulledLaunches, int64_t* pulls )
        {
            std::lock_guard<std::mutex> lk( mu );
            if( launches ) *launches = queueLaunches;
            if( pulledLaunches ) *pulledLaunches = queuePulledLaunches;
            if( pulls ) *pulls = queuePulls;
        }

    private:
        // Per-tree tuning of the batched kernel: times a 2^20-point request of
        // the program at each candidate tile on the tuning stream and keeps
        // the fastest (a candidate must beat the best so far by 3 %; the
        // candidates run from small to large tiles). Runs in the thread that
        // submits the tree's first request of a dimensionality (on the
        // interpreter, points per lane only) and again once the tree's
        // specialised kernel `jit` is loaded (on its batch kernel, with the
        // threads per block as well when `threadsToo`). The persistent mode's
        // work-queue kernel runs the same per-lane loop over the same tiles,
        // so the batch kernel's timing stands in for it.
        Tile Tune( const fn2::DeviceProgram* prog, int dims, const fn2::JitKernel* jit, bool threadsToo )
        {
            std::lock_guard<std::mutex> lk( tuneMu );
            GenArgs a{};
            int64_t n = 0;
            Tile best;
            if( !TuneArgs( prog, dims, a, n ) ) return best;
            best.ms = 1e30f;
            for( int threads : kThreadCandidates )
            {
                if( threads != kThreads && !( jit && threadsToo ) ) continue;
                for( int ppt : kPptCandidates )
                {
                    float ms = TimeTile( prog, a, n, jit, threads, ppt );
                    if( ms < 0 ) return Tile();
                    if( ms < best.ms * 0.97f )
                    {
                        best.ms = ms;
                        best.ppt = ppt;
                        best.threads = threads;
                    }
                }
            }
            return best;
        }

        // The octave-unroll stage of a tree's tuning (after its tile is tuned on
        // the default specialised kernel): requests the unroll variants, and
        // once every one has finished compiling, times the loaded ones at the
        // tuned tile and keeps the fastest if it beats the default by 3 %.
        // Called on every submit of the tree and dimensionality until decided;
        // never blocks on a compilation in async mode.
        void TuneUnroll( int treeId, int d, const fn2::DeviceProgram* prog, Request* r )
        {
            int threads, ppt;
            float baseMs;
            {
                std::lock_guard<std::mutex> lk( mu );
                Tree& t = *trees[treeId];
                if( t.unrollStage[d] != 1 || t.tuning[d] ) return;
                threads = t.threads[d];
                ppt = t.ppt[d];
                baseMs = t.tileMs[d];
            }
            const fn2::JitKernel* k[4] = {};
            for( int v = 1; v < 4; v++ )
            {
                k[v] = fn2::GetJitKernel( prog, v );
                if( !k[v] && fn2::JitVariantState( prog, v ) == 0 ) return;   // still compiling
            }
            {
                std::lock_guard<std::mutex> lk( mu );
                Tree& t = *trees[treeId];
                if( t.unrollStage[d] != 1 || t.tuning[d] ) return;
                t.tuning[d] = true;
            }
            int bestV = 0;
            float bestMs = baseMs;
            {
                std::lock_guard<std::mutex> lk( tuneMu );
                GenArgs a{};
                int64_t n = 0;
                if( TuneArgs( prog, d, a, n ) )
                    for( int v = 1; v < 4; v++ )
                    {
                        if( !k[v] ) continue;
                        float ms = TimeTile( prog, a, n, k[v], threads, ppt );
                        if( ms >= 0 && ms < bestMs * 0.97f )
                        {
                            bestMs = ms;
                            bestV = v;
                        }
                    }
            }
            std::lock_guard<std::mutex> lk( mu );
            Tree& t = *trees[treeId];
            t.variant[d] = bestV;
            t.unrollStage[d] = 2;
            t.tuning[d] = false;
            r->variant = bestV;
        }

        // The tuner's request: a 2^20-point grid of the program in `dims`
        // dimensions, writing into the tuning arena
        bool TuneArgs( const fn2::DeviceProgram* prog, int dims, GenArgs& a, int64_t& n )
        {
            fn2GenSpec spec{};
            spec.kind = FN2_SOURCE_GRID;
            spec.dims = dims;
            int side = dims == 2 ? 1024 : dims == 3 ? 128 : 32;
            for( int ax = 0; ax < dims; ax++ )
            {
                spec.count[ax] = side;
                spec.step[ax] = 1.0f;
            }
            if( dims == 3 ) spec.count[2] = kTunePoints / ( side * side );
            std::string err;
            if( !fn2::FillSpecArgs( spec, a, n, err ) )
            {
                fn2::SetError( err );
                return false;
            }
            a.ins = prog->dIns;
            a.konst = prog->dKonst;
            a.regs = prog->regs;
            a.nIns = (int32_t)prog->ins.size();
            a.resultReg = prog->resultReg;
            a.out = tuneOut;
            a.minmax = tuneKeys;
            a.rangeStride = fn2::kRangeStrideCompact;
            a.rangeSlots = fn2::kRangeSlots;
            return true;
        }

        // Best of two timed runs (after one that warms caches) of the tuner's
        // request at one tile, on the specialised batch kernel `jit` or the
        // interpreter; -1 on a HIP error. Called with tuneMu held.
        float TimeTile( const fn2::DeviceProgram* prog, GenArgs a, int64_t n, const fn2::JitKernel* jit, int threads, int ppt )
        {
            a.ppt = ppt;
            const int64_t tile = (int64_t)threads * ppt;
            const int nb = (int)( ( n + tile - 1 ) / tile );
            std::vector<int32_t> bj( nb, 0 );
            std::vector<uint32_t> bb( nb );
            for( int b = 0; b < nb; b++ ) bb[b] = (uint32_t)( b * tile );
            // the table holds kTunePoints / 64 + 1 entries per column
            const size_t tb = (size_t)kTunePoints / kThreadCandidates[0] + 1;
            int32_t* dBJ = tuneBJ;
            uint32_t* dBB = (uint32_t*)( tuneBJ + tb );
            if( !fn2::HipOk( fn2::UploadSync( tuneJob, &a, sizeof( a ) ), "tune upload" ) ||
                !fn2::HipOk( fn2::UploadSync( dBJ, bj.data(), nb * sizeof( int32_t ) ), "tune upload" ) ||
                !fn2::HipOk( fn2::UploadSync( dBB, bb.data(), nb * sizeof( uint32_t ) ), "tune upload" ) )
                return -1.0f;
            float ms = 1e30f;
            for( int rep = 0; rep < 3; rep++ )
            {
                const fn2::GenArgs* jobs = tuneJob;
                const int32_t* bjp = dBJ;
                const uint32_t* bbp = dBB;
                void* params[] = { (void*)&jobs, (void*)&bjp, (void*)&bbp };
                if( !fn2::HipOk( (hipError_t)fn2_launch_minmax_init( tuneKeys, 1, fn2::kRangeSlots, fn2::kRangeStrideCompact, tuneStream ), "tune launch" ) ||
                    !fn2::HipOk( hipEventRecord( tuneEv0, tuneStream ), "tune event" ) ||
                    !fn2::HipOk( jit ? hipModuleLaunchKernel( jit->batch, (unsigned)nb, 1, 1, (unsigned)threads, 1, 1, 0, tuneStream, params, nullptr )
                                     : (hipError_t)fn2_launch_generate_batch( tuneJob, dBJ, dBB, nb, prog->dimMask, prog->featMask, prog->regs, tuneStream ),
                                 "tune launch" ) ||
                    !fn2::HipOk( hipEventRecord( tuneEv1, tuneStream ), "tune event" ) ||
                    !fn2::HipOk( hipEventSynchronize( tuneEv1 ), "tune run" ) )
                    return -1.0f;
                float t = 0;
                if( !fn2::HipOk( hipEventElapsedTime( &t, tuneEv0, tuneEv1 ), "tune timing" ) ) return -1.0f;
                if( rep > 0 ) ms = std::min( ms, t );   // first run warms caches
            }
            return ms;
        }

        std::mutex tuneMu;
        hipStream_t tuneStream = nullptr;
        hipEvent_t tuneEv0 = nullptr, tuneEv1 = nullptr;
        float* tuneOut = nullptr;
        GenArgs* tuneJob = nullptr;
        int32_t* tuneKeys = nullptr;
        int32_t* tuneBJ = nullptr;

        std::mutex mu;
        std::condition_variable cvWork, cvDone, cvSlot;
        bool stop = false;
        std::deque<Request*> queue;
        std::unordered_map<int64_t, Request*> requests;
        std::vector<std::unique_ptr<Tree>> trees;
        std::vector<Slot> slots;
        std::deque<int> launched;      // slot indices in launch order
        int inflight = 0;
        int64_t nextId = 0;
        int64_t batches = 0;
        int64_t completed = 0;
        int64_t deviceBytes = 0;
        // persistent mode: work-queue launches, those in which waves pulled
        // items beyond the first round, and the pulls (counter increments)
        int64_t queueLaunches = 0, queuePulledLaunches = 0, queuePulls = 0;
        bool tuneUnroll = false;
        int maxBlocks = 0;
        int fillBlocks = 512;   // blocks that fill the GPU twice over (2 per CU)
        int cus = 256;          // compute units of the device
        int nextSlot = 0;
        std::thread dispatcher, completer;
        std::unique_ptr<fn2::CopyPool> copies;

        // Fills slot s with queued requests (FIFO) up to its capacity; the plain
        // mode takes one request per launch. Called with mu held.
        void Gather( Slot& s )
        {
            s.reqs.clear();
            s.points = 0;
            while( !queue.empty() && (int)s.reqs.size() < kMaxJobs )
            {
                if( mode == FN2_SERVE_PLAIN && !s.reqs.empty() ) break;
                Request* r = queue.front();
                // requests start on whole blocks of threads (1 KiB) in the arena
                int64_t span = ( r->points + kThreads - 1 ) / kThreads * kThreads;
                if( s.points + span > capacity ) break;
                queue.pop_front();
                r->outOffset = (size_t)s.points;
                s.points += span;
                s.reqs.push_back( r );
            }
        }

        // Kernel of a job: the tree's specialised batch kernel when it is loaded
        // (key = its address), else the interpreter variant (key < 256).
        static int64_t KernelKey( const fn2::DeviceProgram* p, const fn2::JitKernel* k )
        {
            return k ? (int64_t)(uintptr_t)k : p->dimMask * 16 + p->featMask;
        }

        bool Launch( Slot& s )
        {
            // jobs sorted by kernel so that each kernel's blocks are contiguous
            std::vector<Request*> order( s.reqs );
            for( Request* r : order )
            {
                r->jit = fn2::GetJitKernel( r->prog, r->variant );
                if( !r->jit && r->variant ) r->jit = fn2::GetJitKernel( r->prog );
                // threads per block: the tuned tile's on a specialised batch
                // kernel; the interpreter's, and the work-queue kernels', are fixed
                if( !r->jit || mode == FN2_SERVE_PERSISTENT ) r->threads = kThreads;
            }
            std::stable_sort( order.begin(), order.end(), []( Request* a, Request* b ) {
                const int64_t ka = KernelKey( a->prog, a->jit ), kb = KernelKey( b->prog, b->jit );
                return ka != kb ? ka < kb : a->threads < b->threads;
            } );
            GenArgs* hJobs = (GenArgs*)s.hTables;
            size_t jobsBytes = AlignUp( sizeof( GenArgs ) * kMaxJobs, 256 );
            size_t bjBytes = AlignUp( sizeof( int32_t ) * maxBlocks, 256 );
            int32_t* hBJ = (int32_t*)( s.hTables + jobsBytes );
            uint32_t* hBB = (uint32_t*)( s.hTables + jobsBytes + bjBytes );
            // (kernel key, blocks, register-file size, threads per block) per
            // launch; the register file (interpreter launches only) sizes the
            // launch's LDS
            std::vector<std::array<int64_t, 4>> shape;
            std::vector<const fn2::JitKernel*> kernels;   // per entry of shape
            // A batch too small to fill the GPU at the tuned points per lane
            // runs with fewer points per lane (more blocks): the tuned value
            // maximises throughput, not the latency of a lone small request.
            int pptCap = 1;
            while( pptCap < 8 && (int64_t)pptCap * 2 * kThreads * fillBlocks <= s.points ) pptCap *= 2;
            int nb = 0;
            for( size_t j = 0; j < order.size(); j++ )
            {
                Request* r = order[j];
                r->job = (int)j;
                GenArgs a = r->args;
                a.out = s.dOut + r->outOffset;
                const int ppt = std::min( r->ppt, pptCap );
                a.ppt = ppt;
                a.minmax = s.dKeys + (size_t)jobRangeInts * j;
                a.rangeStride = jobStride;
                a.rangeSlots = fn2::kRangeSlots;
                hJobs[j] = a;
                const int64_t tile = (int64_t)r->threads * ppt;
                int rb = (int)( ( r->points + tile - 1 ) / tile );
                for( int b = 0; b < rb; b++ )
                {
                    hBJ[nb + b] = (int32_t)j;
                    hBB[nb + b] = (uint32_t)( b * tile );
                }
                int64_t key = KernelKey( r->prog, r->jit );
                const int64_t regs = r->jit ? 0 : r->prog->regs;
                if( shape.empty() || shape.back()[0] != key || shape.back()[3] != r->threads )
                {
                    shape.push_back( { key, rb, regs, r->threads } );
                    kernels.push_back( r->jit );
                }
                else
                {
                    shape.back()[1] += rb;
                    shape.back()[2] = std::max( shape.back()[2], regs );
                }
                nb += rb;
            }
            if( mode == FN2_SERVE_PERSISTENT )
            {
                // the interpreter jobs (sorted first: their keys are below 256)
                // in one work-queue launch of the variant covering them all
                size_t ni = 0;
                int64_t dm = 0, fm = 0, regs = 0, blocks = 0;
                for( ; ni < shape.size() && !kernels[ni]; ni++ )
                {
                    dm |= shape[ni][0] / 16;
                    fm |= shape[ni][0] % 16;
                    regs = std::max( regs, shape[ni][2] );
                    blocks += shape[ni][1];
                }
                if( ni > 1 )
                {
                    shape.erase( shape.begin() + 1, shape.begin() + ni );
                    kernels.erase( kernels.begin() + 1, kernels.begin() + ni );
                    shape[0] = { dm * 16 + fm, blocks, regs, kThreads };
                }
            }
            const int nJobs = (int)order.size();
            // the blocks of the tables that are used: copy jobs, then block tables
            if( mode == FN2_SERVE_GRAPH )
            {
                if( s.exec && ( s.graphJobs != nJobs || s.graphShape != shape ) )
                {
                    (void)hipGraphExecDestroy( s.exec );
                    (void)hipGraphDestroy( s.graph );
                    s.exec = nullptr;
                    s.graph = nullptr;
                }
                if( !s.exec )
                {
                    if( !fn2::HipOk( hipStreamBeginCapture( s.stream, hipStreamCaptureModeThreadLocal ), "begin capture" ) ) return false;
                    bool ok = Enqueue( s, shape, kernels, nJobs, nb );
                    hipGraph_t g = nullptr;
                    hipError_t e = hipStreamEndCapture( s.stream, &g );
                    if( !ok || !fn2::HipOk( e, "end capture" ) ) return false;
                    s.graph = g;
                    if( !fn2::HipOk( hipGraphInstantiate( &s.exec, s.graph, nullptr, nullptr, 0 ), "graph instantiate" ) ) return false;
                    s.graphJobs = nJobs;
                    s.graphShape = shape;
                }
                if( !fn2::HipOk( hipGraphLaunch( s.exec, s.stream ), "graph launch" ) ) return false;
            }
            else if( !Enqueue( s, shape, kernels, nJobs, nb ) ) return false;
            s.queueLaunches = mode == FN2_SERVE_PERSISTENT ? (int)shape.size() : 0;
            if( !Copies( s, nJobs ) ) return false;
            if( !fn2::HipOk( hipEventRecord( s.done, s.stream ), "serve event" ) ) return false;
            s.reqs.swap( order );
            return true;
        }

        // Stream work of one batch that depends only on its shape (recorded
        // once per shape in graph mode): table upload, range-key reset, one
        // batch launch per kernel.
        bool Enqueue( Slot& s, const std::vector<std::array<int64_t, 4>>& shape, const std::vector<const fn2::JitKernel*>& kernels,
                      int nJobs, int nb )
        {
            size_t jobsBytes = AlignUp( sizeof( GenArgs ) * kMaxJobs, 256 );
            size_t bjBytes = AlignUp( sizeof( int32_t ) * maxBlocks, 256 );
            size_t usedJob = sizeof( GenArgs ) * nJobs;
            if( !fn2::HipOk( hipMemcpyAsync( s.dJobs, s.hTables, usedJob, hipMemcpyHostToDevice, s.stream ), "upload jobs" ) ) return false;
            if( !fn2::HipOk( hipMemcpyAsync( s.dBlockJob, s.hTables + jobsBytes, sizeof( int32_t ) * nb, hipMemcpyHostToDevice, s.stream ), "upload blocks" ) ) return false;
            if( !fn2::HipOk( hipMemcpyAsync( s.dBlockBase, s.hTables + jobsBytes + bjBytes, sizeof( uint32_t ) * nb, hipMemcpyHostToDevice, s.stream ), "upload blocks" ) ) return false;
            if( mode == FN2_SERVE_PERSISTENT )
            {
                // range keys and one zero counter per work-queue launch, in one kernel
                if( !fn2::HipOk( (hipError_t)fn2_launch_batch_init( s.dKeys, nJobs, fn2::kRangeSlots, jobStride, s.dCounter, (int)shape.size() * fn2::kQueueCounterInts, s.stream ), "range reset" ) )
                    return false;
            }
            else if( !fn2::HipOk( (hipError_t)fn2_launch_minmax_init( s.dKeys, nJobs, fn2::kRangeSlots, jobStride, s.stream ), "range reset" ) ) return false;
            int first = 0;
            for( size_t gi = 0; gi < shape.size(); gi++ )
            {
                const auto& g = shape[gi];
                const int32_t* bj = s.dBlockJob + first;
                const uint32_t* bb = s.dBlockBase + first;
                if( kernels[gi] && mode == FN2_SERVE_PERSISTENT )
                {
                    const fn2::GenArgs* jobs = s.dJobs;
                    uint32_t nTiles = (uint32_t)g[1];
                    uint32_t* counter = s.dCounter + gi * fn2::kQueueCounterInts;
                    const int64_t resident = (int64_t)cus * std::max( 1, kernels[gi]->queueBlocksPerCU );
                    const unsigned grid = (unsigned)std::min<int64_t>( g[1], resident );
                    void* params[] = { (void*)&jobs, (void*)&bj, (void*)&bb, (void*)&nTiles, (void*)&counter };
                    if( !fn2::HipOk( hipModuleLaunchKernel( kernels[gi]->queue, grid, 1, 1, kThreads, 1, 1, 0, s.stream, params, nullptr ),
                                     "specialised work-queue launch" ) )
                        return false;
                }
                else if( kernels[gi] )
                {
                    const fn2::GenArgs* jobs = s.dJobs;
                    void* params[] = { (void*)&jobs, (void*)&bj, (void*)&bb };
                    if( !fn2::HipOk( hipModuleLaunchKernel( kernels[gi]->batch, (unsigned)g[1], 1, 1, (unsigned)g[3], 1, 1, 0, s.stream, params, nullptr ),
                                     "specialised batch launch" ) )
                        return false;
                }
                else
                {
                    int dm = (int)( g[0] / 16 ), fm = (int)( g[0] % 16 );
                    if( mode == FN2_SERVE_PERSISTENT )
                    {
                        if( !fn2::HipOk( (hipError_t)fn2_launch_generate_queue( s.dJobs, bj, bb, (int)g[1], dm, fm, (int)g[2], s.dCounter + gi * fn2::kQueueCounterInts, s.stream ),
                                         "work-queue launch" ) )
                            return false;
                    }
                    else if( !fn2::HipOk( (hipError_t)fn2_launch_generate_batch( s.dJobs, bj, bb, (int)g[1], dm, fm, (int)g[2], s.stream ), "batch launch" ) )
                        return false;
                }
                first += (int)g[1];
            }
            return true;
        }

        // Copies of one batch's results, issued on the slot's stream after its
        // launches (never recorded: destinations change from batch to batch).
        bool Copies( Slot& s, int nJobs )
        {
            size_t stagedEnd = 0;   // the staged requests lie in [0, stagedEnd) of the arena
            for( Request* r : s.reqs )
            {
                const size_t bytes = (size_t)r->points * 4;
                if( r->dest == Request::kStaged ) stagedEnd = std::max( stagedEnd, r->outOffset + (size_t)r->points );
                else if( !fn2::HipOk( hipMemcpyAsync( r->hostOut, s.dOut + r->outOffset, bytes,
                                                      r->dest == Request::kDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s.stream ),
                                      "copy output" ) )
                    return false;
            }
            if( stagedEnd && !fn2::HipOk( hipMemcpyAsync( s.hOut, s.dOut, stagedEnd * 4, hipMemcpyDeviceToHost, s.stream ), "copy outputs" ) ) return false;
            if( !fn2::HipOk( hipMemcpyAsync( s.hKeys, s.dKeys, jobRangeInts * sizeof( int32_t ) * nJobs, hipMemcpyDeviceToHost, s.stream ), "copy ranges" ) ) return false;
            // the work-queue counters: how many pulls each launch made
            if( s.queueLaunches &&
                !fn2::HipOk( hipMemcpyAsync( s.hKeys + (size_t)jobRangeInts * kMaxJobs, s.dCounter,
                                             sizeof( uint32_t ) * fn2::kQueueCounterInts * s.queueLaunches, hipMemcpyDeviceToHost, s.stream ),
                             "copy queue counters" ) )
                return false;
            return true;
        }

        void DispatchLoop()
        {
            for( ;; )
            {
                int si;
                {
                    std::unique_lock<std::mutex> lk( mu );
                    cvWork.wait( lk, [&] { return stop || ( !queue.empty() && !slots[nextSlot].busy ); } );
                    if( stop ) return;
                    si = nextSlot;
                    Gather( slots[si] );
                    slots[si].busy = true;
                    inflight++;
                    nextSlot = ( nextSlot + 1 ) % (int)slots.size();
                }
                Slot& s = slots[si];
                bool ok = Launch( s );
                {
                    std::lock_guard<std::mutex> lk( mu );
                    if( !ok )
                    {
                        for( Request* r : s.reqs ) r->status = 1;
                    }
                    launched.push_back( ok ? si : -1 - si );
                    batches++;
                }
                cvSlot.notify_all();
            }
        }

        void CompleteLoop()
        {
            for( ;; )
            {
                int code;
                {
                    std::unique_lock<std::mutex> lk( mu );
                    cvSlot.wait( lk, [&] { return stop || !launched.empty(); } );
                    if( launched.empty() ) return;
                    code = launched.front();
                    launched.pop_front();
                }
                int si = code >= 0 ? code : -1 - code;
                Slot& s = slots[si];
                bool ok = code >= 0 && fn2::HipOk( hipEventSynchronize( s.done ), "serve batch" );
                if( ok )
                {
                    std::vector<fn2::CopyPool::Seg> segs;
                    segs.reserve( s.reqs.size() );
                    for( Request* r : s.reqs )
                        if( r->dest == Request::kStaged ) segs.push_back( { r->hostOut, s.hOut + r->outOffset, (size_t)r->points } );
                    copies->Run( segs );
                    for( Request* r : s.reqs )
                    {
                        if( !fn2::DecodeRange( s.hKeys + (size_t)r->job * jobRangeInts, fn2::kRangeSlots, jobStride, r->mm[0], r->mm[1] ) )
                        {
                            r->mm[0] = __builtin_inff();
                            r->mm[1] = -__builtin_inff();
                        }
                    }
                    int64_t pulled = 0, pulls = 0;
                    const uint32_t* qc = (const uint32_t*)( s.hKeys + (size_t)jobRangeInts * kMaxJobs );
                    for( int l = 0; l < s.queueLaunches; l++ )
                    {
                        int64_t n = 0;
                        for( int c = 0; c < fn2::kQueueCounters; c++ ) n += qc[l * fn2::kQueueCounterInts + c * fn2::kQueueCounterStride];
                        pulls += n;
                        pulled += n > 0;
                    }
                    std::lock_guard<std::mutex> lk( mu );
                    queueLaunches += s.queueLaunches;
                    queuePulledLaunches += pulled;
                    queuePulls += pulls;
                }
                auto now = Clock::now();
                {
                    std::lock_guard<std::mutex> lk( mu );
                    for( Request* r : s.reqs )
                    {
                        if( !ok ) r->status = 1;
                        r->tDone = now;
                        if( Request* p = r->parent )
                        {
                            // a part: fold it into its request
                            if( r->status ) p->status = 1;
                            p->mm[0] = std::min( p->mm[0], r->mm[0] );
                            p->mm[1] = std::max( p->mm[1], r->mm[1] );
                            if( --p->pending == 0 )
                            {
                                p->tDone = now;
                                p->done = true;
                                completed++;
                            }
                            delete r;
                            continue;
                        }
                        r->done = true;
                        completed++;
                    }
                    s.reqs.clear();
                    s.busy = false;
                    inflight--;
                }
                cvDone.notify_all();
                cvWork.notify_all();
            }
        }
    };
}

extern "C" {

FASTNOISE_API void* fn2ServeCreate( size_t hbmBudgetBytes, int maxBatchPoints, int mode )
{
    fn2::ClearError();
    if( !fn2::EnsureDevice() ) return nullptr;
    if( mode < FN2_SERVE_PLAIN || mode > FN2_SERVE_PERSISTENT || maxBatchPoints <= 0 )
    {
        fn2::SetError( "bad serving engine arguments" );
        return nullptr;
    }
    Engine* e = new Engine( hbmBudgetBytes, maxBatchPoints, mode );
    if( !e->Init() )
    {
        fn2::SetError( e->initError );
        delete e;
        return nullptr;
    }
    return e;
}

FASTNOISE_API void fn2ServeDestroy( void* engine )
{
    delete static_cast<Engine*>( engine );
}

FASTNOISE_API int fn2ServeAddTree( void* engine, const char* encodedTree )
{
    fn2::ClearError();
    if( !engine || !encodedTree ) { fn2::SetError( "null argument" ); return -1; }
    return static_cast<Engine*>( engine )->AddTree( encodedTree );
}

FASTNOISE_API int64_t fn2ServeSubmit( void* engine, int treeId, const fn2GenSpec* spec, float* hostOut )
{
    fn2::ClearError();
    if( !engine || !spec ) { fn2::SetError( "null argument" ); return -1; }
    return static_cast<Engine*>( engine )->Submit( treeId, *spec, hostOut );
}

FASTNOISE_API int fn2ServeWait( void* engine, int64_t requestId, float* outputMinMax, double* latencyMs )
{
    fn2::ClearError();
    if( !engine ) { fn2::SetError( "null engine" ); return 1; }
    return static_cast<Engine*>( engine )->Wait( requestId, outputMinMax, latencyMs );
}

FASTNOISE_API int fn2ServePoll( void* engine, int64_t requestId, float* outputMinMax, double* latencyMs )
{
    fn2::ClearError();
    if( !engine ) { fn2::SetError( "null engine" ); return -1; }
    return static_cast<Engine*>( engine )->Poll( requestId, outputMinMax, latencyMs );
}

FASTNOISE_API int fn2ServeDrain( void* engine )
{
    fn2::ClearError();
    if( !engine ) { fn2::SetError( "null engine" ); return 1; }
    return static_cast<Engine*>( engine )->Drain();
}

FASTNOISE_API int fn2ServeStats( void* engine, int64_t* batches, int64_t* completed, int64_t* peakBytes )
{
    if( !engine ) return 1;
    static_cast<Engine*>( engine )->Stats( batches, completed, peakBytes );
    return 0;
}

FASTNOISE_API int fn2ServeTreeTuning( void* engine, int treeId, int dims )
{
    if( !engine ) return -1;
    return static_cast<Engine*>( engine )->TreePpt( treeId, dims );
}

FASTNOISE_API int fn2ServeTuneOctaveUnroll( void* engine, int enable )
{
    if( !engine ) return 1;
    static_cast<Engine*>( engine )->SetTuneUnroll( enable != 0 );
    return 0;
}

FASTNOISE_API int fn2ServeQueueStats( void* engine, int64_t* launches, int64_t* pulledLaunches, int64_t* pulls )
{
    if( !engine ) return 1;
    static_cast<Engine*>( engine )->QueueStats( launches, pulledLaunches, pulls );
    return 0;
}

FASTNOISE_API int fn2ServeTreeTile( void* engine, int treeId, int dims, int* threadsPerBlock, int* pointsPerLane, int* octaveUnroll )
{
    if( !engine ) return -1;
    return static_cast<Engine*>( engine )->TreeTile( treeId, dims, threadsPerBlock, pointsPerLane, octaveUnroll );
}

FASTNOISE_API int fn2ServeReplay( void* engine, int count, const int* treeIds, const fn2GenSpec* specs, const double* arrivalSec,
                                  int maxOutstanding, int deviceOutputs, double* latencyMs, double* wallSec )
{
    fn2::ClearError();
    if( !engine || count < 0 || !treeIds || !specs || !latencyMs || maxOutstanding < 1 )
    {
        fn2::SetError( "bad replay arguments" );
        return 1;
    }
    Engine* e = static_cast<Engine*>( engine );
    // output buffers: one per outstanding request, sized for the largest request
    int64_t maxPoints = 0;
    for( int i = 0; i < count; i++ )
    {
        fn2::GenArgs a{};
        int64_t n = 0;
        std::string err;
        if( !fn2::FillSpecArgs( specs[i], a, n, err ) )
        {
            fn2::SetError( err );
            return 1;
        }
        maxPoints = std::max( maxPoints, n );
    }
    // pinned host buffers, as a serving front end would hold (the engine
    // copies into them directly), or device buffers for consumers on the GPU
    std::vector<float*> pool( (size_t)maxOutstanding, nullptr );
    struct PoolFree
    {
        std::vector<float*>& p;
        bool device;
        ~PoolFree()
        {
            for( float* b : p )
                if( b ) (void)( device ? hipFree( b ) : hipHostFree( b ) );
        }
    } poolFree{ pool, deviceOutputs != 0 };
    const size_t poolBytes = (size_t)std::max<int64_t>( maxPoints, 1 ) * 4;
    for( float*& b : pool )
    {
        hipError_t err = deviceOutputs ? hipMalloc( (void**)&b, poolBytes ) : hipHostMalloc( (void**)&b, poolBytes, hipHostMallocDefault );
        if( !fn2::HipOk( err, "replay buffers" ) ) return 1;
    }
    std::deque<std::pair<int64_t, int>> pending;   // (request id, pool slot), in submit order
    std::vector<int> freeSlots;
    for( int k = maxOutstanding - 1; k >= 0; k-- ) freeSlots.push_back( k );
    std::vector<int> reqIndex;
    auto t0 = Clock::now();
    int status = 0;
    int done = 0;
    auto retireOne = [&]() {
        auto p = pending.front();
        pending.pop_front();
        double lat = 0;
        if( e->Wait( p.first, nullptr, &lat ) ) status = 1;
        latencyMs[reqIndex[done]] = lat;
        done++;
        freeSlots.push_back( p.second );
    };
    for( int i = 0; i < count; i++ )
    {
        if( arrivalSec )
        {
            auto due = t0 + std::chrono::duration_cast<Clock::duration>( std::chrono::duration<double>( arrivalSec[i] ) );
            // retire what has completed while waiting for the arrival time
            while( !pending.empty() && Clock::now() < due && freeSlots.empty() ) retireOne();
            std::this_thread::sleep_until( due );
        }
        while( freeSlots.empty() ) retireOne();
        int slot = freeSlots.back();
        freeSlots.pop_back();
        int64_t id = e->Submit( treeIds[i], specs[i], pool[slot] );
        if( id < 0 ) return 1;
        pending.emplace_back( id, slot );
        reqIndex.push_back( i );
    }
    while( !pending.empty() ) retireOne();
    if( wallSec ) *wallSec = std::chrono::duration<double>( Clock::now() - t0 ).count();
    return status;
}

}
