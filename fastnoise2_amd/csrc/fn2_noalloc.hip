// fn2GenerateDeviceNoAlloc: the generation call of fn2GenerateDevice for
// callers that must not allocate device memory (code under HIP graph capture,
// tight serving loops, hosts that forbid hipMalloc after start-up).
//
// Everything the call needs beyond the caller's positions and output lives in
// one file-scope device array, `scratch`, that every GPU holds for the life of
// the process:
//
//   [ range record | program instructions | constant pool | tileable table ]
//
// A tree whose specialised kernel earlier calls have loaded runs on that
// kernel: its instructions and constants are literals, so it reads no
// program from the array. Otherwise the program runs on the interpreter
// kernel; this call never compiles or loads a specialised kernel (a
// code-object load allocates). Results are the same bits either way.
//
// Each call overwrites the array of the GPU it runs on. Calls that overlap on
// one GPU (two host threads, or two streams whose work is not ordered) race on
// it: serialise them, or use fn2GenerateDevice, whose calls each take their
// own buffers.
#include <hip/hip_runtime.h>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fn2_backend.h"
#include "fn2_gpu.h"
#include "fn2_jit.h"
#include <FastNoise/FastNoise.h>

#include "fn2_launch.h"

namespace
{
    using namespace fn2;

    constexpr size_t Align256( size_t n ) { return ( n + 255 ) & ~(size_t)255; }

    constexpr size_t kKeysBytes = Align256( kCallRangeInts * sizeof( int32_t ) );
    constexpr size_t kInsBytes = Align256( (size_t)kMaxInstructions * sizeof( Ins ) );
    constexpr size_t kKonstBytes = Align256( ( (size_t)FN2_NOALLOC_MAX_CONSTANTS + 16 ) * sizeof( float ) );
    constexpr size_t kTileBytes = Align256( 2 * (size_t)FN2_NOALLOC_MAX_TILE_AXES * sizeof( float ) );

    constexpr size_t kKeysOff = 0;
    constexpr size_t kInsOff = kKeysOff + kKeysBytes;
    constexpr size_t kKonstOff = kInsOff + kInsBytes;
    constexpr size_t kTileOff = kKonstOff + kKonstBytes;
    constexpr size_t kScratchBytes = kTileOff + kTileBytes;
    static_assert( kScratchBytes <= 512 * 1024, "the scratch array stays modest" );

    static __device__ __align__( 256 ) unsigned char scratch[kScratchBytes];

    // Device address of `scratch` per GPU, looked up once per device with that
    // device current. Host code uses these addresses, never the array's name
    // (which on the host is only a stand-in for the symbol).
    constexpr int kMaxDevices = 64;
    std::mutex gScratchMu;
    unsigned char* gScratch[kMaxDevices] = {};

    unsigned char* ScratchForCurrentDevice( std::string& err )
    {
        int dev = 0;
        hipError_t e = hipGetDevice( &dev );
        if( e != hipSuccess )
        {
            (void)hipGetLastError();
            err = std::string( "hipGetDevice: " ) + hipGetErrorString( e );
            return nullptr;
        }
        if( dev < 0 || dev >= kMaxDevices )
        {
            err = "device index beyond the 64 devices the no-allocation entry point keeps addresses for";
            return nullptr;
        }
        std::lock_guard<std::mutex> lk( gScratchMu );
        if( !gScratch[dev] )
        {
            void* p = nullptr;
            e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
            if( e != hipSuccess || !p )
            {
                (void)hipGetLastError();
                err = std::string( "hipGetSymbolAddress(scratch): " ) + hipGetErrorString( e );
                return nullptr;
            }
            gScratch[dev] = static_cast<unsigned char*>( p );
        }
        return gScratch[dev];
    }
}

extern "C" {

FASTNOISE_API size_t fn2NoAllocScratchBytes( void )
{
    return kScratchBytes;
}

FASTNOISE_API int fn2GenerateDeviceNoAlloc( const void* node, const fn2GenSpec* spec, float* dOut, float* outputMinMax, void* stream )
{
    ClearError();
    if( !node || !spec ) { SetError( "null argument" ); return 1; }
    if( !EnsureDevice() ) return 1;
    std::string err;
    unsigned char* base = ScratchForCurrentDevice( err );
    if( !base ) { SetError( err ); return 1; }

    if( spec->kind == FN2_SOURCE_TILEABLE &&
        ( spec->count[0] < 0 || spec->count[1] < 0 || (int64_t)spec->count[0] + spec->count[1] > FN2_NOALLOC_MAX_TILE_AXES ) )
    {
        SetError( "tileable size beyond the no-allocation limit: xSize + ySize must be at most " +
                  std::to_string( FN2_NOALLOC_MAX_TILE_AXES ) );
        return 1;
    }
    float* dTile = reinterpret_cast<float*>( base + kTileOff );
    GenArgs a{};
    int64_t total = 0;
    if( !FillSpecArgs( *spec, a, total, err, spec->kind == FN2_SOURCE_TILEABLE ? dTile : nullptr ) ) { SetError( err ); return 1; }
    if( total == 0 )
    {
        if( outputMinMax ) { outputMinMax[0] = FastNoise::kInfinity; outputMinMax[1] = -FastNoise::kInfinity; }
        return 0;
    }
    if( !dOut ) { SetError( "null output buffer" ); return 1; }

    Program p;
    if( !CompileTree( static_cast<const FastNoise::Generator*>( node ), a.dims, p, err ) )
    {
        SetError( "cannot compile node tree: " + err );
        return 1;
    }
    int regs = 1;
    if( !ValidateProgram( p, regs, err ) ) { SetError( err ); return 1; }
    if( p.konst.size() > (size_t)FN2_NOALLOC_MAX_CONSTANTS )
    {
        SetError( "compiled tree holds " + std::to_string( p.konst.size() ) + " constants, beyond the no-allocation limit of " +
                  std::to_string( FN2_NOALLOC_MAX_CONSTANTS ) );
        return 1;
    }

    // the caller's stream, else this thread's default stream (never the
    // legacy null stream, which would order against every blocking stream)
    hipStream_t st = stream ? (hipStream_t)stream : hipStreamPerThread;

    // a specialised kernel of this tree that is already loaded (looked up
    // only: nothing is created, compiled or queued here)
    const JitKernel* jit = nullptr;
    if( const DeviceProgram* dp = FindDeviceProgram( p ) )
    {
        const int chosen = dp->variant.load( std::memory_order_relaxed );
        jit = LoadedJitKernel( dp, 0 );
        if( chosen )
            if( const JitKernel* v = LoadedJitKernel( dp, chosen ) ) jit = v;
        if( a.kind == SRC_POSITIONS && JitSharedWalkApplies( *dp ) )
            if( const JitKernel* v = LoadedJitKernel( dp, chosen | kJitSharedWalk ) ) jit = v;
    }
    Ins* dIns = reinterpret_cast<Ins*>( base + kInsOff );
    float* dKonst = reinterpret_cast<float*>( base + kKonstOff );
    int32_t* dKeys = reinterpret_cast<int32_t*>( base + kKeysOff );

    // Program and tables are copied from host memory on the call's stream, so
    // the kernel below sees them; the constant pool is padded by 16 zeros,
    // as in the device program cache, for windows that start at its end.
    if( !jit )
    {
        if( !p.ins.empty() && !HipOk( hipMemcpyAsync( dIns, p.ins.data(), p.ins.size() * sizeof( Ins ), hipMemcpyHostToDevice, st ), "copy program" ) ) return 1;
        p.konst.resize( p.konst.size() + 16, 0.0f );
        if( !HipOk( hipMemcpyAsync( dKonst, p.konst.data(), p.konst.size() * sizeof( float ), hipMemcpyHostToDevice, st ), "copy constants" ) ) return 1;
    }
    if( spec->kind == FN2_SOURCE_TILEABLE )
    {
        std::vector<float> t( 2 * ( (size_t)spec->count[0] + spec->count[1] ) );
        FillTileTable( spec->count[0], spec->count[1], spec->step[0], spec->step[1], t.data() );
        if( !HipOk( hipMemcpyAsync( dTile, t.data(), t.size() * sizeof( float ), hipMemcpyHostToDevice, st ), "copy tileable table" ) ) return 1;
    }

    a.ins = dIns;
    a.konst = dKonst;
    a.regs = regs;
    a.nIns = (int32_t)p.ins.size();
    a.resultReg = p.resultReg;
    a.out = dOut;
    a.minmax = outputMinMax ? dKeys : nullptr;
    // the range record starts empty: a fill kernel, not a memset, so that the
    // call can be recorded into a graph
    if( outputMinMax && !HipOk( (hipError_t)fn2_launch_minmax_init( dKeys, 1, kCallRangeSlots, kRangeStrideSpread, st ), "launch min/max init" ) ) return 1;
    if( jit )
    {
        GenArgs copy = a;
        void* params[] = { &copy };
        if( !HipOk( hipModuleLaunchKernel( jit->single, fn2_grid_for( total, jit->blocksPerCU ), 1, 1, 256, 1, 1, 0, st, params, nullptr ),
                    "launch specialised generator" ) )
            return 1;
    }
    else if( !HipOk( (hipError_t)fn2_launch_generate( &a, p.dimMask, p.featMask, st ), "launch generator" ) ) return 1;
    if( outputMinMax )
    {
        int32_t keys[kCallRangeInts];
        if( !HipOk( hipMemcpyAsync( keys, dKeys, sizeof( keys ), hipMemcpyDeviceToHost, st ), "read min/max" ) ) return 1;
        if( !HipOk( hipStreamSynchronize( st ), "generator kernel" ) ) return 1;
        if( !DecodeRange( keys, kCallRangeSlots, kRangeStrideSpread, outputMinMax[0], outputMinMax[1] ) )
        {
            outputMinMax[0] = FastNoise::kInfinity;
            outputMinMax[1] = -FastNoise::kInfinity;
        }
    }
    else if( !stream )
    {
        // finish on the thread's stream: the next call on this GPU rewrites the array
        if( !HipOk( hipStreamSynchronize( st ), "generator kernel" ) ) return 1;
    }
    return 0;
}

}
