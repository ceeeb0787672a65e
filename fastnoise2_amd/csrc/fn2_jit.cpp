// Run-time specialised generator kernels (hipRTC).
//
// The interpreter kernels run any compiled tree, fetching each instruction
// and indexing a 32-entry register file. For a tree that is used repeatedly,
// this module writes HIP source in which the tree's instructions are
// compile-time constants: every `execIns` call folds to the one operation it
// names, register indices become constants, and the register file becomes
// ordinary VGPRs. The source includes the same device headers as the
// interpreter (embedded in the library at build time), so the arithmetic is
// identical and the results are bit-identical.
//
// Compilation runs on a background thread (mode async, the default): calls
// keep using the interpreter until the tree's kernel is loaded. Mode sync
// compiles on first use; mode off disables specialisation. The mode comes
// from fn2SetJitMode or the FN2_JIT environment variable (off / async / sync).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fn2_backend.h"
#include "fn2_jit.h"

namespace fn2
{
    namespace
    {
        struct JitHeader
        {
            const char* name;
            const char* text;
        };

        const JitHeader kJitHeaders[] = {
#include "fn2_jit_sources.inc"
        };

        const char* kJitOptions[] = {
            "--offload-arch=gfx950", "-O3", "-std=c++17",
            "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
            "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        };

        enum State { kQueued, kReady, kFailed, kTooLarge };

        struct Entry
        {
            State state = kQueued;
            JitKernel kernel;
            std::string error;
        };

        // Constants one instruction may read past its pool index (the device
        // pool carries the same padding)
        constexpr int kJitConstWindow = 16;

        // Programs at least this long get register fences between instructions,
        // one every kJitFenceEvery instructions (both overridable for
        // experiments: FN2_JIT_FENCE_MIN, FN2_JIT_FENCE_EVERY)
        int JitFenceMin()
        {
            static const int v = [] { const char* e = getenv( "FN2_JIT_FENCE_MIN" ); return e && atoi( e ) > 0 ? atoi( e ) : 128; }();
            return v;
        }
        int JitFenceEvery()
        {
            static const int v = [] { const char* e = getenv( "FN2_JIT_FENCE_EVERY" ); return e && atoi( e ) > 0 ? atoi( e ) : 1; }();
            return v;
        }

        // The feature group an op needs compiled in (see fn2::kFeat*)
        int OpFeature( int op )
        {
            switch( op )
            {
            case OP_CELL_VALUE: case OP_CELL_DISTANCE: case OP_CELL_LOOKUP: return kFeatCell;
            case OP_WARP_GRADIENT: case OP_WARP_SIMPLEX: case OP_WARP_SUPERSIMPLEX: return kFeatWarp;
            case OP_WHITE: case OP_CHECKERBOARD: case OP_SINEWAVE: case OP_DISTANCE: return kFeatMisc;
            default: return 0;
            }
        }

        // Compile-time detail of an instruction beyond its op: the leaf op of a
        // fused fractal, the distance slots a CellularDistance reads
        int Aux( const Ins& in )
        {
            if( in.op == OP_FRACTAL_LEAF ) return in.p[0];
            if( in.op == OP_CELL_DISTANCE ) return std::min( 3, std::max<int>( in.p[1], in.p[2] ) ) + 1;
            return -1;
        }

        // Whether the program holds a CellularDistance that takes the pruned
        // cell walk (squared distance, no size jitter): such programs walk grid
        // requests in bricks, which lets whole waves leave out more cells
        // programs of at least this many instructions get the lean source
        constexpr size_t kJitLeanMin = 128;

        bool PrunedCellular( const DeviceProgram& p )
        {
            for( const Ins& in : p.ins )
                if( in.op == OP_CELL_DISTANCE && ( in.p[0] < 2 || in.p[0] > 5 ) && in.r[3] == kNoReg ) return true;
            return false;
        }

        // a fused 2D Simplex fractal, whose octaves share their lattice cell
        // across a wave more often in 8 x 8 bricks than in rows of 64: with
        // the wave-uniform cell path, bricks took FBm(Simplex) 4096^2 from
        // 0.1640 to 0.1575 ms on MI355X (bricks alone 0.1655, the uniform
        // path on rows 0.1695)
        bool UniformSimplex2D( const DeviceProgram& p )
        {
            for( const Ins& in : p.ins )
                if( in.op == OP_FRACTAL_LEAF && in.dim == 2 && in.p[0] == OP_SIMPLEX ) return true;
            return false;
        }

        // Each instruction becomes an `execIns` instantiated for its own op,
        // dimensionality and (fused fractal) leaf op, so the compiler sees only
        // the code that instruction runs.
        std::string Source( const DeviceProgram& p, int variant )
        {
            std::ostringstream s;
            if( variant & kJitUnrollMask ) s << "#define FN2_OCTAVE_UNROLL " << kJitUnrollFactor[variant & kJitUnrollMask] << "\n";
            if( variant & kJitSharedWalk ) s << "#define FN2_SHARED_CELL_WALK 1\n";
            // long programs leave out the code that is duplicated per
            // instruction for coherent waves (cell tables, wave-uniform warp
            // cells, a brick-specific program copy): measured with hipcc at
            // -O3, the 355-instruction mixed tree compiled in 73 s without
            // them and 210 s with them, which a serving engine's first
            // requests would wait out on the interpreter
            const bool lean = p.ins.size() >= (size_t)kJitLeanMin;
            if( lean ) s << "#define FN2_CELL_TABLE 0\n#define FN2_UNIFORM_WARP 0\n";
            s << "#include \"fn2_kernel_body.h\"\n"
                 "using namespace fn2d;\nusing namespace fn2i;\n"
                 "namespace {\nstruct Prog {\n"
                 "static constexpr bool kFresh = true;\n"
                 "static constexpr int kDims = " << p.dims << ";\n"
                 "static constexpr bool kBricks = " << ( PrunedCellular( p ) || UniformSimplex2D( p ) ? "true" : "false" ) << ";\n"
                 "static constexpr bool kBrickAware = " << ( lean ? "false" : "true" ) << ";\n"
                 "template<bool BRICK = false>\n"
                 "__device__ __forceinline__ float eval( const fn2::GenArgs& a, float* R ) const {\n";
            // long programs get a register fence between instructions: measured
            // on MI355X, a 355-instruction tree needs 256 VGPRs without (one wave
            // per SIMD, 2.46 ms on 128^3) and 92 with (5 waves, 1.63 ms); a
            // 145-instruction tree runs 10 % faster with them, while some
            // 40-instruction trees lose 8 % to them
            const bool fence = p.ins.size() >= (size_t)JitFenceMin();
            const int every = JitFenceEvery();
            for( const Ins& in : p.ins )
            {
                const size_t i = (size_t)( &in - &p.ins[0] );
                if( fence && i > 0 && i % (size_t)every == 0 ) s << "fn2k::registerFence<" << p.regs << ">( R );\n";
                s << "{ constexpr fn2::Ins in = { " << (int)in.op << ", " << (int)in.dim << ", { ";
                for( int i = 0; i < 4; i++ ) s << (int)in.p[i] << ( i < 3 ? ", " : " }, { " );
                for( int i = 0; i < 6; i++ ) s << (int)in.r[i] << ( i < 5 ? ", " : " }, " );
                s << "(int32_t)" << (uint32_t)in.seed << "u, " << in.k << "u };\n";
                // the instruction's constant window as literals (exact bit patterns)
                s << "  constexpr float k[" << kJitConstWindow << "] = { ";
                for( int i = 0; i < kJitConstWindow; i++ )
                {
                    uint32_t bits = 0;
                    if( in.k + i < p.konst.size() ) memcpy( &bits, &p.konst[in.k + i], 4 );
                    s << "__builtin_bit_cast( float, 0x" << std::hex << bits << std::dec << "u )" << ( i + 1 < kJitConstWindow ? ", " : " };\n" );
                }
                s << "  execIns<" << ( in.dim >= 2 && in.dim <= 4 ? 1 << in.dim : 4 ) << ", " << OpFeature( in.op ) << ", "
                  << (int)in.op << ", " << Aux( in )
                  << ", BRICK>( in, k, iadd( a.seed, (int32_t)" << (uint32_t)in.seed << "u ), R ); }\n";
            }
            s << "return R[" << p.resultReg << "];\n}\n"
                 "__device__ __forceinline__ float operator()( const fn2::GenArgs& a, float* R ) const { return eval<false>( a, R ); }\n"
                 "};\n}\n"
                 "extern \"C\" __global__ void __launch_bounds__( 256 ) fn2_jit_single( fn2::GenArgs a )\n"
                 "{ fn2k::generateSingle<256>( a, Prog() ); }\n"
                 "extern \"C\" __global__ void __launch_bounds__( 256 ) fn2_jit_batch( const fn2::GenArgs* __restrict__ jobs, "
                 "const int32_t* __restrict__ bj, const uint32_t* __restrict__ bb )\n"
                 "{ fn2k::generateBatch<256>( jobs, bj, bb, Prog() ); }\n"
                 "extern \"C\" __global__ void __launch_bounds__( 256 ) fn2_jit_queue( const fn2::GenArgs* __restrict__ jobs, "
                 "const int32_t* __restrict__ tj, const uint32_t* __restrict__ tb, uint32_t nTiles, uint32_t* counter )\n"
                 "{ fn2k::generateQueue<256>( jobs, tj, tb, nTiles, counter, Prog() ); }\n";
            return s.str();
        }

        bool CompileCode( const DeviceProgram& p, int variant, std::vector<char>& code, std::string& err )
        {
            std::string src = Source( p, variant );
            std::vector<const char*> hdrs, names;
            for( const JitHeader& h : kJitHeaders )
            {
                hdrs.push_back( h.text );
                names.push_back( h.name );
            }
            hiprtcProgram prog;
            if( hiprtcCreateProgram( &prog, src.c_str(), "fn2_jit.hip", (int)hdrs.size(), hdrs.data(), names.data() ) != HIPRTC_SUCCESS )
            {
                err = "hiprtcCreateProgram failed";
                return false;
            }
            hiprtcResult r = hiprtcCompileProgram( prog, (int)( sizeof( kJitOptions ) / sizeof( kJitOptions[0] ) ), kJitOptions );
            if( r != HIPRTC_SUCCESS )
            {
                size_t n = 0;
                hiprtcGetProgramLogSize( prog, &n );
                std::string log( n, '\0' );
                if( n ) hiprtcGetProgramLog( prog, &log[0] );
                err = "hipRTC compile failed: " + log.substr( 0, 2000 );
                hiprtcDestroyProgram( &prog );
                return false;
            }
            size_t size = 0;
            hiprtcGetCodeSize( prog, &size );
            code.resize( size );
            hiprtcGetCode( prog, code.data() );
            hiprtcDestroyProgram( &prog );
            return true;
        }

        bool Compile( const DeviceProgram& p, int variant, JitKernel& k, std::string& err )
        {
            std::vector<char> code;
            if( !CompileCode( p, variant, code, err ) ) return false;
            hipModule_t mod = nullptr;
            hipError_t e = hipModuleLoadData( &mod, code.data() );
            if( e == hipSuccess ) e = hipModuleGetFunction( &k.single, mod, "fn2_jit_single" );
            if( e == hipSuccess ) e = hipModuleGetFunction( &k.batch, mod, "fn2_jit_batch" );
            if( e == hipSuccess ) e = hipModuleGetFunction( &k.queue, mod, "fn2_jit_queue" );
            if( e != hipSuccess )
            {
                (void)hipGetLastError();
                err = std::string( "loading the specialised kernel: " ) + hipGetErrorString( e );
                return false;
            }
            k.module = mod;
            if( hipModuleOccupancyMaxActiveBlocksPerMultiprocessor( &k.blocksPerCU, k.single, 256, 0 ) != hipSuccess )
            {
                (void)hipGetLastError();
                k.blocksPerCU = 0;
            }
            if( hipModuleOccupancyMaxActiveBlocksPerMultiprocessor( &k.queueBlocksPerCU, k.queue, 256, 0 ) != hipSuccess )
            {
                (void)hipGetLastError();
                k.queueBlocksPerCU = 0;
            }
            if( hipFuncGetAttribute( &k.singleRegs, HIP_FUNC_ATTRIBUTE_NUM_REGS, k.single ) != hipSuccess ||
                hipFuncGetAttribute( &k.singleScratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k.single ) != hipSuccess )
            {
                (void)hipGetLastError();
                k.singleRegs = k.singleScratch = -1;
            }
            return true;
        }

        // One per process, never destroyed (see Manager)
        class JitManager
        {
        public:
            JitManager()
            {
                const char* env = getenv( "FN2_JIT" );
                if( env )
                {
                    if( !strcmp( env, "off" ) || !strcmp( env, "0" ) ) mode = FN2_JIT_OFF;
                    else if( !strcmp( env, "sync" ) ) mode = FN2_JIT_SYNC;
                }
                const char* lim = getenv( "FN2_JIT_MAX_INSTRUCTIONS" );
                if( lim && atoi( lim ) > 0 ) maxInstructions = atoi( lim );
                const char* th = getenv( "FN2_JIT_THREADS" );
                if( th && atoi( th ) > 0 ) threads = std::min( 16, atoi( th ) );
            }


            // Stops the background compiler at exit (registered with atexit):
            // no queued compilation starts, and the ones in progress finish, so
            // no worker is inside the compiler libraries while they are torn
            // down. Idle workers are not woken: they stay blocked until the
            // process ends, so none of them exits, and runs the HIP runtime's
            // per-thread clean-up, while the runtime's own state is destroyed.
            void Shutdown()
            {
                std::unique_lock<std::mutex> lk( mu );
                stop = true;
                cv.wait( lk, [&] { return busy == 0; } );
                for( std::thread& w : workers )
                    if( w.joinable() ) w.detach();
            }

            int mode = FN2_JIT_ASYNC;
            // longer programs stay on the interpreter: compile time grows
            // faster than linearly (about 15 s for 355 instructions on the
            // MI355X host, most of it in LLVM's value-range propagation)
            size_t maxInstructions = 512;

            const JitKernel* Get( const DeviceProgram* p, int variant )
            {
                if( mode == FN2_JIT_OFF || !p ) return nullptr;
                const Key key{ p, variant };
                std::unique_lock<std::mutex> lk( mu );
                auto it = entries.find( key );
                if( it == entries.end() )
                {
                    it = entries.emplace( key, Entry() ).first;
                    if( p->ins.size() > maxInstructions )
                    {
                        it->second.state = kTooLarge;
                        return nullptr;
                    }
                    if( mode == FN2_JIT_SYNC )
                    {
                        lk.unlock();
                        Entry e;
                        e.state = Compile( *p, variant, e.kernel, e.error ) ? kReady : kFailed;
                        lk.lock();
                        it = entries.find( key );
                        it->second = e;
                        cv.notify_all();
                    }
                    else
                    {
                        if( stop ) return nullptr;
                        queue.push_back( key );
                        // compilations are independent: a few run side by side
                        if( (int)workers.size() < threads && (int)workers.size() < (int)queue.size() + busy )
                            workers.emplace_back( [this] { Work(); } );
                        cv.notify_all();
                        return nullptr;
                    }
                }
                else if( mode == FN2_JIT_SYNC && it->second.state == kQueued )
                {
                    cv.wait( lk, [&] { return entries[key].state != kQueued; } );
                    it = entries.find( key );
                }
                return it->second.state == kReady ? &it->second.kernel : nullptr;
            }

            const JitKernel* Loaded( const DeviceProgram* p, int variant )
            {
                if( mode == FN2_JIT_OFF || !p ) return nullptr;
                std::lock_guard<std::mutex> lk( mu );
                auto it = entries.find( Key{ p, variant } );
                return it != entries.end() && it->second.state == kReady ? &it->second.kernel : nullptr;
            }

            int State( const DeviceProgram* p, int variant, std::string* err )
            {
                std::lock_guard<std::mutex> lk( mu );
                auto it = entries.find( Key{ p, variant } );
                if( it == entries.end() ) return -1;
                if( err ) *err = it->second.error;
                return it->second.state;
            }

            void Wait()
            {
                std::unique_lock<std::mutex> lk( mu );
                cv.wait( lk, [&] { return queue.empty() && busy == 0; } );
            }

        private:
            void Work()
            {
                for( ;; )
                {
                    Key key;
                    {
                        std::unique_lock<std::mutex> lk( mu );
                        // after Shutdown, an idle worker waits until the process ends
                        cv.wait( lk, [&] { return !stop && !queue.empty(); } );
                        key = queue.front();
                        queue.pop_front();
                        busy++;
                    }
                    Entry e;
                    e.state = Compile( *key.p, key.variant, e.kernel, e.error ) ? kReady : kFailed;
                    {
                        std::lock_guard<std::mutex> lk( mu );
                        entries[key] = e;
                        busy--;
                    }
                    cv.notify_all();
                }
            }

            struct Key
            {
                const DeviceProgram* p = nullptr;
                int variant = 0;
                bool operator==( const Key& o ) const { return p == o.p && variant == o.variant; }
            };
            struct KeyHash
            {
                size_t operator()( const Key& k ) const { return std::hash<const void*>()( k.p ) * 31u + (size_t)k.variant; }
            };

            std::mutex mu;
            std::condition_variable cv;
            std::unordered_map<Key, Entry, KeyHash> entries;
            std::deque<Key> queue;
            std::vector<std::thread> workers;
            int threads = 4;
            int busy = 0;
            bool stop = false;
        };

        JitManager* gManager = nullptr;

        void ShutdownAtExit()
        {
            if( gManager ) gManager->Shutdown();
        }

        JitManager& Manager()
        {
            // lives for the process (its kernels stay loaded); the workers
            // stop taking compilations at exit
            static JitManager* m = [] {
                gManager = new JitManager;
                std::atexit( ShutdownAtExit );
                return gManager;
            }();
            return *m;
        }
    }

    const JitKernel* GetJitKernel( const DeviceProgram* p, int variant )
    {
        return Manager().Get( p, variant );
    }

    const JitKernel* LoadedJitKernel( const DeviceProgram* p, int variant )
    {
        return Manager().Loaded( p, variant );
    }

    int JitVariantState( const DeviceProgram* p, int variant )
    {
        return Manager().State( p, variant, nullptr );
    }

    bool JitHasFusedFractal( const DeviceProgram& p )
    {
        for( const Ins& in : p.ins )
            if( in.op == OP_FRACTAL_LEAF ) return true;
        return false;
    }

    bool JitSharedWalkApplies( const DeviceProgram& p )
    {
        for( const Ins& in : p.ins )
            if( in.op == OP_CELL_DISTANCE && in.dim == 3 && ( in.p[0] < 2 || in.p[0] > 5 ) && in.r[3] == kNoReg && Aux( in ) <= 2 )
                return true;
        return false;
    }

    std::string JitSourceFor( const DeviceProgram& p, int variant )
    {
        return Source( p, variant );
    }
}

extern "C" {

FASTNOISE_API int fn2SetJitMode( int mode )
{
    if( mode < FN2_JIT_OFF || mode > FN2_JIT_SYNC ) return -1;
    int old = fn2::Manager().mode;
    fn2::Manager().mode = mode;
    return old;
}

FASTNOISE_API int fn2JitWait( void )
{
    fn2::Manager().Wait();
    return 0;
}

FASTNOISE_API int fn2JitCompileCheck( const void* node, int dims, int* codeBytes )
{
    fn2::ClearError();
    if( !node ) { fn2::SetError( "null node" ); return 1; }
    fn2::Program prog;
    std::string err;
    if( !fn2::CompileTree( static_cast<const FastNoise::Generator*>( node ), dims, prog, err ) )
    {
        fn2::SetError( "cannot compile node tree: " + err );
        return 1;
    }
    fn2::DeviceProgram dp;
    dp.ins = prog.ins;
    dp.konst = prog.konst;
    dp.resultReg = prog.resultReg;
    dp.dims = prog.dims;
    std::vector<char> code;
    if( !fn2::CompileCode( dp, 0, code, err ) )
    {
        fn2::SetError( err );
        return 1;
    }
    if( codeBytes ) *codeBytes = (int)code.size();
    return 0;
}

FASTNOISE_API int fn2JitSource( const void* node, int dims, char* buf, int bufSize )
{
    fn2::ClearError();
    if( !node ) { fn2::SetError( "null node" ); return -1; }
    fn2::Program prog;
    std::string err;
    if( !fn2::CompileTree( static_cast<const FastNoise::Generator*>( node ), dims, prog, err ) )
    {
        fn2::SetError( "cannot compile node tree: " + err );
        return -1;
    }
    fn2::DeviceProgram dp;
    dp.ins = prog.ins;
    dp.konst = prog.konst;
    dp.resultReg = prog.resultReg;
    dp.dims = prog.dims;
    std::string src = fn2::Source( dp, 0 );
    if( buf && bufSize > 0 )
    {
        size_t n = std::min( src.size(), (size_t)bufSize - 1 );
        memcpy( buf, src.data(), n );
        buf[n] = 0;
    }
    return (int)src.size();
}

FASTNOISE_API int fn2GetJitVariantState( const void* node, int dims, int variant )
{
    fn2::ClearError();
    if( !node || variant < 0 || variant > ( fn2::kJitUnrollMask | fn2::kJitSharedWalk ) ) return -2;
    const fn2::DeviceProgram* dp = nullptr;
    if( !fn2::EnsureDevice() || !fn2::CompileForDevice( static_cast<const FastNoise::Generator*>( node ), dims, dp, nullptr ) ) return -2;
    std::string err;
    int st = fn2::Manager().State( dp, variant, &err );
    if( st == 2 ) fn2::SetError( err );
    return st;
}

FASTNOISE_API int fn2SetJitVariant( const void* node, int dims, int variant )
{
    fn2::ClearError();
    if( !node || variant < 0 || variant > fn2::kJitUnrollMask )
    {
        fn2::SetError( "bad node or variant" );
        return 1;
    }
    const fn2::DeviceProgram* dp = nullptr;
    if( !fn2::EnsureDevice() || !fn2::CompileForDevice( static_cast<const FastNoise::Generator*>( node ), dims, dp, nullptr ) ) return 1;
    dp->variant.store( variant, std::memory_order_relaxed );
    return 0;
}

FASTNOISE_API int fn2GetJitKernelInfo( const void* node, int dims, int variant, int* vgprs, int* scratchBytes, int* blocksPerCU )
{
    fn2::ClearError();
    if( !node || variant < 0 || variant > ( fn2::kJitUnrollMask | fn2::kJitSharedWalk ) ) return -1;
    const fn2::DeviceProgram* dp = nullptr;
    if( !fn2::EnsureDevice() || !fn2::CompileForDevice( static_cast<const FastNoise::Generator*>( node ), dims, dp, nullptr ) ) return -1;
    if( fn2::Manager().State( dp, variant, nullptr ) != 1 ) return 1;
    const fn2::JitKernel* k = fn2::GetJitKernel( dp, variant );
    if( !k ) return 1;
    if( vgprs ) *vgprs = k->singleRegs;
    if( scratchBytes ) *scratchBytes = k->singleScratch;
    if( blocksPerCU ) *blocksPerCU = k->blocksPerCU;
    return 0;
}

FASTNOISE_API int fn2GetJitState( const void* node, int dims )
{
    return fn2GetJitVariantState( node, dims, 0 );
}

}
