// Host side of the MI355X backend: device set-up, the compiled-program cache,
// per-call contexts and FastNoise::Generator's batch entry points
// (Generator.h:173-450 of the reference), which here compile the tree, launch
// one interpreter kernel over all points and copy the result back.
//
// Caller buffers may be host memory (copied through a device staging buffer)
// or device memory (hipMalloc'd: the kernel reads and writes them in place).
// Errors never fall back to another implementation: the call fills its output
// with NaN, returns a NaN range, records a message for fn2GetLastError() and
// prints it once to stderr.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "fn2_backend.h"
#include "fn2_copy_pool.h"
#include "fn2_internal.h"
#include "fn2_jit.h"
#include "fn2_gpu.h"

#include "fn2_launch.h"

namespace fn2
{
    namespace
    {
        thread_local std::string tLastError;

        std::mutex gMu;
        std::once_flag gInitOnce;
        int gInitStatus = -1;   // 0 ok
        std::string gInitError;

        std::mutex gPrintMu;
        std::unordered_set<std::string> gPrinted;

        std::unordered_map<uint64_t, std::vector<DeviceProgram*>> gPrograms;
        size_t gProgramCount = 0;
        constexpr size_t kMaxCachedPrograms = 8192;

        std::vector<CallContext*> gContexts;

        std::mutex gTileMu;
        std::map<std::tuple<int, int, uint32_t, uint32_t>, float*> gTileTables;
        constexpr size_t kMaxTileTables = 1024;

        // Uploads of programs and tables go through one non-blocking stream:
        // never through the legacy default stream, which would conflict with a
        // serving engine capturing a graph on another thread.
        std::mutex gUploadMu;
        hipStream_t gUploadStream = nullptr;

        void DeviceInit()
        {
            int n = 0;
            hipError_t e = hipGetDeviceCount( &n );
            if( e != hipSuccess || n <= 0 )
            {
                (void)hipGetLastError();
                gInitError = "no HIP device available (the FastNoise2 MI355X backend runs on the GPU only)";
                gInitStatus = 1;
                return;
            }
            if( hipStreamCreateWithFlags( &gUploadStream, hipStreamNonBlocking ) != hipSuccess )
            {
                (void)hipGetLastError();
                gInitError = "cannot create the upload stream";
                gInitStatus = 1;
                return;
            }
            gInitStatus = 0;
        }
    }

    void SetError( const std::string& msg )
    {
        tLastError = msg;
        std::lock_guard<std::mutex> lk( gPrintMu );
        if( gPrinted.insert( msg ).second )
        {
            fprintf( stderr, "FastNoise2 MI355X backend error: %s\n", msg.c_str() );
        }
    }

    void ClearError() { tLastError.clear(); }
    const char* LastError() { return tLastError.c_str(); }

    bool EnsureDevice()
    {
        std::call_once( gInitOnce, DeviceInit );
        if( gInitStatus != 0 )
        {
            SetError( gInitError );
            return false;
        }
        return true;
    }

    bool HipOk( hipError_t e, const char* what )
    {
        if( e == hipSuccess ) return true;
        SetError( std::string( what ) + ": " + hipGetErrorString( e ) );
        return false;
    }

    hipError_t UploadSync( void* dst, const void* src, size_t bytes )
    {
        std::lock_guard<std::mutex> lk( gUploadMu );
        hipError_t e = hipMemcpyAsync( dst, src, bytes, hipMemcpyHostToDevice, gUploadStream );
        if( e == hipSuccess ) e = hipStreamSynchronize( gUploadStream );
        return e;
    }

    hipError_t LaunchProgram( const GenArgs& a, const DeviceProgram* dp, hipStream_t stream, bool* specialised )
    {
        if( a.total <= 0 ) return hipSuccess;
        const int chosen = dp->variant.load( std::memory_order_relaxed );
        const JitKernel* k = GetJitKernel( dp );
        if( chosen )
            if( const JitKernel* v = GetJitKernel( dp, chosen ) ) k = v;
        // scattered positions: the variant with the shared far-cell walk,
        // once it is loaded
        if( a.kind == SRC_POSITIONS && JitSharedWalkApplies( *dp ) )
            if( const JitKernel* v = GetJitKernel( dp, chosen | kJitSharedWalk ) ) k = v;
        if( k )
        {
            if( specialised ) *specialised = true;
            GenArgs copy = a;
            void* params[] = { &copy };
            return hipModuleLaunchKernel( k->single, fn2_grid_for( a.total, k->blocksPerCU ), 1, 1, 256, 1, 1, 0, stream, params, nullptr );
        }
        if( specialised ) *specialised = false;
        return (hipError_t)fn2_launch_generate( &a, dp->dimMask, dp->featMask, stream );
    }

    FastDiv MakeFastDiv( uint32_t d )
    {
        // Granlund & Montgomery 1994, Figure 4.1: l = ceil(log2 d),
        // m = floor( 2^32 (2^l - d) / d ) + 1, shifts min(l,1) and max(l-1,0)
        FastDiv f;
        f.d = d;
        uint32_t l = 0;
        while( l < 32 && ( (uint64_t)1 << l ) < d ) l++;
        f.m = (uint32_t)( ( ( (uint64_t)1 << 32 ) * ( ( (uint64_t)1 << l ) - d ) ) / d + 1 );
        f.s1 = l < 1 ? l : 1;
        f.s2 = l > 0 ? l - 1 : 0;
        return f;
    }

    // Registers operand k of an instruction spans: a vector operand covers
    // the instruction's dimensionality (2 or 3 for the rotations), a scalar
    // one register, an unused one none (the layouts of fn2_program.h)
    int OperandExtent( const Ins& in, int k )
    {
        const int d = std::max<int>( 1, in.dim );
        switch( in.op )
        {
        case OP_PERLIN: case OP_VALUE: case OP_SIMPLEX: case OP_SUPERSIMPLEX:
        case OP_WHITE: case OP_CHECKERBOARD: case OP_SINEWAVE: case OP_FRACTAL_LEAF:
            return k == 0 ? 1 : k == 1 ? d : 0;
        case OP_CELL_VALUE: case OP_CELL_DISTANCE:
            return k == 1 ? d : k <= 3 ? 1 : 0;
        case OP_CELL_LOOKUP:
            return k <= 1 ? d : k <= 3 ? 1 : 0;
        case OP_WARP_GRADIENT: case OP_WARP_SIMPLEX: case OP_WARP_SUPERSIMPLEX:
            return k == 1 || k == 2 ? d : k <= 3 ? 1 : 0;
        case OP_CONST: return k == 0 ? 1 : 0;
        case OP_MOV: case OP_ABS: case OP_POWI: case OP_SSQRT: case OP_RGBA8: return k <= 1 ? 1 : 0;
        case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MIN: case OP_MAX: case OP_FMOD:
        case OP_FBM_SUM: case OP_RIDGED_TERM: case OP_PINGPONG: case OP_POWF: case OP_TERRACE: case OP_GRADIENT_AXIS:
            return k <= 2 ? 1 : 0;
        case OP_FBM_AMP: case OP_RIDGED_AMP: case OP_WARP_AMP: case OP_MINSMOOTH: return k <= 3 ? 1 : 0;
        case OP_FADE: case OP_REMAP: return 1;
        case OP_SCALE: case OP_SCALE_AXES: case OP_MOV_VEC: return k <= 1 ? d : 0;
        case OP_MUL_VEC: return k <= 1 ? d : k == 2 ? 1 : 0;
        case OP_ADD_VEC: case OP_SUB_VEC: return k <= 2 ? d : 0;
        case OP_ROTATE2: return k <= 1 ? 2 : 0;
        case OP_ROTATE3: case OP_ROTATE_PLANE: return k <= 1 ? 3 : 0;
        case OP_DISTANCE: return k == 1 ? d : k == 0 || k == 2 ? 1 : 0;
        default: return 4;
        }
    }

    bool ValidateProgram( const Program& p, int& regs, std::string& err )
    {
        // every register an instruction touches (vector operands span dim
        // registers) must lie below the program's register count, itself at
        // most the register file; every constant-pool window (at most 16
        // floats) must lie inside the uploaded pool
        regs = std::max( 1, std::max( p.regsUsed, p.dims ) );
        if( regs > kMaxRegs )
        {
            err = "compiled program uses more registers than the register file holds";
            return false;
        }
        for( const Ins& in : p.ins )
        {
            if( in.op == OP_END || in.op >= OP_COUNT || in.dim > 4 )
            {
                err = "compiled program has an invalid instruction";
                return false;
            }
            for( int r = 0; r < 6; r++ )
            {
                const int ext = OperandExtent( in, r );
                if( in.r[r] != kNoReg && ext > 0 && in.r[r] + ext > regs )
                {
                    err = "compiled program addresses a register outside its register file";
                    return false;
                }
            }
            if( in.k > p.konst.size() )
            {
                err = "compiled program addresses constants outside its pool";
                return false;
            }
        }
        if( p.resultReg < 0 || p.resultReg >= kMaxRegs || p.ins.size() > (size_t)kMaxInstructions )
        {
            err = "compiled program is malformed";
            return false;
        }
        return true;
    }

    static bool SameProgram( const DeviceProgram* d, const Program& p )
    {
        return d->resultReg == p.resultReg && d->dims == p.dims && d->ins.size() == p.ins.size() && d->konst.size() == p.konst.size() &&
               memcmp( d->ins.data(), p.ins.data(), p.ins.size() * sizeof( Ins ) ) == 0 &&
               memcmp( d->konst.data(), p.konst.data(), p.konst.size() * sizeof( float ) ) == 0;
    }

    const DeviceProgram* FindDeviceProgram( const Program& p )
    {
        std::lock_guard<std::mutex> lk( gMu );
        auto it = gPrograms.find( p.hash );
        if( it == gPrograms.end() ) return nullptr;
        for( DeviceProgram* d : it->second )
            if( SameProgram( d, p ) ) return d;
        return nullptr;
    }

    const DeviceProgram* GetDeviceProgram( const Program& p )
    {
        std::lock_guard<std::mutex> lk( gMu );
        auto& bucket = gPrograms[p.hash];
        for( DeviceProgram* d : bucket )
        {
            if( SameProgram( d, p ) ) return d;
        }
        if( gProgramCount >= kMaxCachedPrograms )
        {
            SetError( "device program cache full (more than 8192 distinct compiled trees)" );
            return nullptr;
        }
        int regs = 1;
        std::string err;
        if( !ValidateProgram( p, regs, err ) )
        {
            SetError( err );
            return nullptr;
        }
        DeviceProgram* d = new DeviceProgram;
        d->ins = p.ins;
        d->konst = p.konst;
        d->resultReg = p.resultReg;
        d->dimMask = p.dimMask;
        d->featMask = p.featMask;
        d->dims = p.dims;
        d->regs = regs;
        size_t insBytes = std::max<size_t>( 1, p.ins.size() ) * sizeof( Ins );
        // padded so that a window starting at the last constant stays in bounds
        size_t kBytes = ( p.konst.size() + 16 ) * sizeof( float );
        if( !HipOk( hipMalloc( (void**)&d->dIns, insBytes ), "hipMalloc(program)" ) ||
            !HipOk( hipMalloc( (void**)&d->dKonst, kBytes ), "hipMalloc(constants)" ) )
        {
            delete d;
            return nullptr;
        }
        if( !p.ins.empty() && !HipOk( UploadSync( d->dIns, p.ins.data(), p.ins.size() * sizeof( Ins ) ), "upload program" ) ) return nullptr;
        std::vector<float> padded( p.konst );
        padded.resize( p.konst.size() + 16, 0.0f );
        if( !HipOk( UploadSync( d->dKonst, padded.data(), padded.size() * sizeof( float ) ), "upload constants" ) ) return nullptr;
        bucket.push_back( d );
        gProgramCount++;
        return d;
    }

    bool CompileForDevice( const FastNoise::Generator* root, int dims, const DeviceProgram*& dp, Program* hostCopy )
    {
        Program p;
        std::string err;
        if( !CompileTree( root, dims, p, err ) )
        {
            SetError( "cannot compile node tree: " + err );
            return false;
        }
        dp = GetDeviceProgram( p );
        if( hostCopy ) *hostCopy = std::move( p );
        return dp != nullptr;
    }

    CallContext* AcquireContext()
    {
        {
            std::lock_guard<std::mutex> lk( gMu );
            if( !gContexts.empty() )
            {
                CallContext* c = gContexts.back();
                gContexts.pop_back();
                return c;
            }
        }
        CallContext* c = new CallContext;
        if( !HipOk( hipStreamCreateWithFlags( &c->stream, hipStreamNonBlocking ), "hipStreamCreate" ) ||
            !HipOk( hipMalloc( (void**)&c->dKeys, kCallRangeInts * sizeof( int32_t ) ), "hipMalloc(min/max)" ) ||
            !HipOk( hipHostMalloc( (void**)&c->hKeys, kCallRangeInts * sizeof( int32_t ), hipHostMallocDefault ), "hipHostMalloc" ) ||
            !HipOk( hipEventCreate( &c->ev0 ), "hipEventCreate" ) || !HipOk( hipEventCreate( &c->ev1 ), "hipEventCreate" ) )
        {
            return nullptr;
        }
        return c;
    }

    void ReleaseContext( CallContext* c )
    {
        if( !c ) return;
        std::lock_guard<std::mutex> lk( gMu );
        gContexts.push_back( c );
    }

    bool CallContext::Reserve( float*& buf, size_t& cap, size_t floats )
    {
        if( floats <= cap ) return true;
        if( buf )
        {
            (void)hipStreamSynchronize( stream );
            (void)hipFree( buf );
            buf = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>( floats, 1024 );
        if( !HipOk( hipMalloc( (void**)&buf, want * sizeof( float ) ), "hipMalloc(staging)" ) ) return false;
        cap = want;
        return true;
    }

    // Host copies of the staged entry points: one pool for the process, used
    // by one call at a time (the copies are bound by memory bandwidth).
    std::mutex gHostCopyMu;

    CopyPool& HostCopies()
    {
        static CopyPool* pool = new CopyPool( 3 );   // three workers plus the caller
        return *pool;
    }

    void HostCopy( float* dst, const float* src, int64_t count )
    {
        std::lock_guard<std::mutex> lk( gHostCopyMu );
        HostCopies().Run( { { dst, src, (size_t)count } } );
    }

    bool CallContext::ReserveStage()
    {
        if( hStage ) return true;
        for( int i = 0; i < kStageSlots; i++ )
        {
            if( !evStage[i] && !HipOk( hipEventCreateWithFlags( &evStage[i], hipEventDisableTiming ), "hipEventCreate" ) ) return false;
        }
        return HipOk( hipHostMalloc( (void**)&hStage, (size_t)kStageSlots * kStageChunkFloats * sizeof( float ), hipHostMallocDefault ),
                      "hipHostMalloc(staging ring)" );
    }

    bool CallContext::StagedToHost( float* dst, const float* dSrc, int64_t count )
    {
        if( !ReserveStage() ) return false;
        const int64_t chunks = ( count + kStageChunkFloats - 1 ) / kStageChunkFloats;
        auto issue = [&]( int64_t k ) {
            const int64_t b = k * kStageChunkFloats, m = std::min( kStageChunkFloats, count - b );
            float* slot = hStage + ( k % kStageSlots ) * kStageChunkFloats;
            return HipOk( hipMemcpyAsync( slot, dSrc + b, (size_t)m * sizeof( float ), hipMemcpyDeviceToHost, stream ), "copy output" ) &&
                   HipOk( hipEventRecord( evStage[k % kStageSlots], stream ), "hipEventRecord" );
        };
        for( int64_t k = 0; k < std::min<int64_t>( chunks, kStageSlots ); k++ )
        {
            if( !issue( k ) ) return false;
        }
        for( int64_t k = 0; k < chunks; k++ )
        {
            if( !HipOk( hipEventSynchronize( evStage[k % kStageSlots] ), "generator kernel" ) ) return false;
            const int64_t b = k * kStageChunkFloats, m = std::min( kStageChunkFloats, count - b );
            HostCopy( dst + b, hStage + ( k % kStageSlots ) * kStageChunkFloats, m );
            // the slot is free again: the chunk kStageSlots ahead goes into it
            if( k + kStageSlots < chunks && !issue( k + kStageSlots ) ) return false;
        }
        return true;
    }

    bool CallContext::StagedToDevice( float* dDst, const float* src, int64_t count )
    {
        if( !ReserveStage() ) return false;
        const int64_t chunks = ( count + kStageChunkFloats - 1 ) / kStageChunkFloats;
        for( int64_t k = 0; k < chunks; k++ )
        {
            const int s = (int)( k % kStageSlots );
            // the slot's previous upload (of this call, or of the axis before)
            // must have left it; an event never recorded is complete
            if( !HipOk( hipEventSynchronize( evStage[s] ), "copy positions" ) ) return false;
            const int64_t b = k * kStageChunkFloats, m = std::min( kStageChunkFloats, count - b );
            float* slot = hStage + (int64_t)s * kStageChunkFloats;
            HostCopy( slot, src + b, m );
            if( !HipOk( hipMemcpyAsync( dDst + b, slot, (size_t)m * sizeof( float ), hipMemcpyHostToDevice, stream ), "copy positions" ) ||
                !HipOk( hipEventRecord( evStage[s], stream ), "hipEventRecord" ) )
            {
                return false;
            }
        }
        return true;
    }

    bool IsDevicePointer( const void* p )
    {
        if( !p ) return false;
        hipPointerAttribute_t attr;
        hipError_t e = hipPointerGetAttributes( &attr, p );
        if( e != hipSuccess )
        {
            (void)hipGetLastError();
            return false;
        }
        return attr.type == hipMemoryTypeDevice;
    }

    float DecodeKey( int32_t k )
    {
        int32_t b = k >= 0 ? k : k ^ 0x7fffffff;
        float f;
        memcpy( &f, &b, 4 );
        return f;
    }

    bool DecodeRange( const int32_t* record, int slots, int stride, float& min, float& max )
    {
        int32_t kMin = 0x7fffffff, kMax = (int32_t)0x80000000u;
        for( int s = 0; s < slots; s++ )
        {
            kMin = std::min( kMin, record[s * stride] );
            kMax = std::max( kMax, record[s * stride + 1] );
        }
        if( kMin == 0x7fffffff && kMax == (int32_t)0x80000000u ) return false;
        min = DecodeKey( kMin );
        max = DecodeKey( kMax );
        return true;
    }

    // Points a spec describes (0 when a count is negative or the product overflows int)
    int64_t SpecPoints( const fn2GenSpec& s )
    {
        int64_t n = 1;
        int axes = s.kind == FN2_SOURCE_POSITIONS ? 1 : s.kind == FN2_SOURCE_TILEABLE ? 2 : s.dims;
        if( axes < 1 || axes > 4 ) return 0;
        for( int d = 0; d < axes; d++ )
        {
            if( s.count[d] < 0 ) return 0;
            n *= s.count[d];   // both factors < 2^31: no overflow
            if( n > 0x7fffffff ) return 0;
        }
        return n;
    }

    void FillTileTable( int xSize, int ySize, float xStep, float yStep, float* t )
    {
        const float pi2Recip = 0.15915493667f;
        const float xSizePi = (float)xSize * pi2Recip;
        const float ySizePi = (float)ySize * pi2Recip;
        const float xFreq = xSizePi * xStep;
        const float yFreq = ySizePi * yStep;
        const float xMul = 1 / xSizePi;
        const float yMul = 1 / ySizePi;
        for( int i = 0; i < xSize; i++ )
        {
            float f = (float)i * xMul;
            t[i] = cosf( f ) * xFreq;
            t[xSize + i] = sinf( f ) * xFreq;
        }
        for( int j = 0; j < ySize; j++ )
        {
            float f = (float)j * yMul;
            t[2 * xSize + j] = cosf( f ) * yFreq;
            t[2 * xSize + ySize + j] = sinf( f ) * yFreq;
        }
    }

    // GenTileable2D maps column x and row y onto a 4D torus (Generator.inl:370-433
    // of the reference): x' = cos( x * xMul ) * xFreq, z' = sin( x * xMul ) * xFreq
    // and the same for y with w'. Those coordinates depend on the column or the
    // row alone, so they are tabulated once per (size, step) on the host and
    // kept on the device: [cos x | sin x] (2 * xSize), then [cos y | sin y].
    const float* TileTable( int xSize, int ySize, float xStep, float yStep, std::string& err )
    {
        uint32_t sx, sy;
        memcpy( &sx, &xStep, 4 );
        memcpy( &sy, &yStep, 4 );
        auto key = std::make_tuple( xSize, ySize, sx, sy );
        std::lock_guard<std::mutex> lk( gTileMu );
        auto it = gTileTables.find( key );
        if( it != gTileTables.end() ) return it->second;
        if( gTileTables.size() >= kMaxTileTables )
        {
            err = "more than 1024 distinct tileable sizes in one process";
            return nullptr;
        }
        std::vector<float> t( 2 * ( (size_t)xSize + ySize ) );
        FillTileTable( xSize, ySize, xStep, yStep, t.data() );
        float* d = nullptr;
        hipError_t e = hipMalloc( (void**)&d, t.size() * sizeof( float ) );
        if( e == hipSuccess ) e = UploadSync( d, t.data(), t.size() * sizeof( float ) );
        if( e != hipSuccess )
        {
            err = std::string( "tileable table upload: " ) + hipGetErrorString( e );
            return nullptr;
        }
        gTileTables.emplace( key, d );
        return d;
    }

    bool FillSpecArgs( const fn2GenSpec& s, GenArgs& a, int64_t& total, std::string& err, const float* tileTable )
    {
        a.rangeStride = kRangeStrideSpread;   // single calls; the serving engine packs its records
        a.rangeSlots = kCallRangeSlots;
        a.kind = s.kind;
        a.dims = s.dims;
        a.seed = s.seed;
        if( s.kind == FN2_SOURCE_TILEABLE )
        {
            if( s.dims != 4 && s.dims != 2 ) { err = "tileable generation is 2D"; return false; }
            a.dims = 4;
            if( s.count[0] < 0 || s.count[1] < 0 ) { err = "negative size"; return false; }
            total = (int64_t)s.count[0] * s.count[1];
            if( total > 0x7fffffff ) { err = "more than 2^31-1 points in one call"; return false; }
            if( total > 0 )
            {
                a.div[0] = MakeFastDiv( (uint32_t)s.count[0] );
                const float* tab = tileTable ? tileTable : TileTable( s.count[0], s.count[1], s.step[0], s.step[1], err );
                if( !tab ) return false;
                a.pos[0] = tab;
                a.pos[1] = tab + 2 * s.count[0];
                a.pos[2] = tab + s.count[0];
                a.pos[3] = tab + 2 * s.count[0] + s.count[1];
            }
        }
        else if( s.kind == FN2_SOURCE_GRID )
        {
            if( s.dims < 2 || s.dims > 4 ) { err = "grid dimensions must be 2, 3 or 4"; return false; }
            total = 1;
            for( int d = 0; d < s.dims; d++ )
            {
                if( s.count[d] < 0 ) { err = "negative count"; return false; }
                total *= s.count[d];   // both factors < 2^31: no overflow
                if( total > 0x7fffffff ) { err = "more than 2^31-1 points in one call"; return false; }
                a.offset[d] = s.offset[d];
                a.step[d] = s.step[d];
            }
            if( total > 0 )
            {
                for( int d = 0; d < s.dims - 1; d++ ) a.div[d] = MakeFastDiv( (uint32_t)s.count[d] );
            }
        }
        else if( s.kind == FN2_SOURCE_POSITIONS )
        {
            if( s.dims < 2 || s.dims > 4 ) { err = "position dimensions must be 2, 3 or 4"; return false; }
            if( s.count[0] < 0 ) { err = "negative count"; return false; }
            total = s.count[0];
            for( int d = 0; d < s.dims; d++ )
            {
                a.offset[d] = s.offset[d];
                a.pos[d] = s.pos[d];
            }
        }
        else
        {
            err = "unknown source kind";
            return false;
        }
        if( total > 0x7fffffff )
        {
            err = "more than 2^31-1 points in one call";
            return false;
        }
        a.total = (int32_t)total;
        return true;
    }

    // One batch call of the public API: host or device buffers, synchronous.
    OutputRange GenerateHost( const FastNoise::Generator* root, const fn2GenSpec& spec, float* out, const float* const* hostPos )
    {
        ClearError();
        OutputRange r;
        GenArgs a{};
        int64_t total = 0;
        std::string err;
        // points the caller's buffer holds: 0 when the spec is invalid or too large
        const int64_t fillCount = SpecPoints( spec );
        auto fail = [&]( const std::string& msg ) -> OutputRange
        {
            SetError( msg );
            if( out && fillCount > 0 && !IsDevicePointer( out ) )
            {
                for( int64_t i = 0; i < fillCount; i++ ) out[i] = __builtin_nanf( "" );
            }
            OutputRange bad;
            bad.min = bad.max = __builtin_nanf( "" );
            bad.ok = false;
            return bad;
        };
        if( !EnsureDevice() ) return fail( gInitError );
        if( !FillSpecArgs( spec, a, total, err ) ) return fail( err );
        if( total == 0 ) return r;
        if( !out ) return fail( "null output buffer" );

        const DeviceProgram* dp = nullptr;
        if( !CompileForDevice( root, a.dims, dp, nullptr ) ) return fail( tLastError );

        CallContext* c = AcquireContext();
        if( !c ) return fail( tLastError );
        struct Releaser { CallContext* c; ~Releaser() { ReleaseContext( c ); } } rel{ c };

        a.ins = dp->dIns;
        a.konst = dp->dKonst;
        a.regs = dp->regs;
        a.nIns = (int32_t)dp->ins.size();
        a.resultReg = dp->resultReg;
        a.minmax = c->dKeys;

        const bool devOut = IsDevicePointer( out );
        if( devOut ) a.out = out;
        else
        {
            if( !c->Reserve( c->dOut, c->outCap, (size_t)total ) ) return fail( tLastError );
            a.out = c->dOut;
        }
        if( spec.kind == FN2_SOURCE_POSITIONS )
        {
            bool allDev = true;
            for( int d = 0; d < spec.dims; d++ ) allDev = allDev && IsDevicePointer( hostPos[d] );
            if( allDev )
            {
                for( int d = 0; d < spec.dims; d++ ) a.pos[d] = hostPos[d];
            }
            else
            {
                if( !c->Reserve( c->dPos, c->posCap, (size_t)total * spec.dims ) ) return fail( tLastError );
                for( int d = 0; d < spec.dims; d++ )
                {
                    if( !hostPos[d] ) return fail( "null position array" );
                    float* dst = c->dPos + (size_t)d * total;
                    if( total >= kStagedMinFloats && !IsDevicePointer( hostPos[d] ) )
                    {
                        if( !c->StagedToDevice( dst, hostPos[d], total ) ) return fail( tLastError );
                    }
                    else if( !HipOk( hipMemcpyAsync( dst, hostPos[d], (size_t)total * sizeof( float ), hipMemcpyDefault, c->stream ), "copy positions" ) )
                    {
                        return fail( tLastError );
                    }
                    a.pos[d] = dst;
                }
            }
        }

        if( !HipOk( (hipError_t)fn2_launch_minmax_init( c->dKeys, 1, kCallRangeSlots, kRangeStrideSpread, c->stream ), "launch min/max init" ) ) return fail( tLastError );
        if( !HipOk( LaunchProgram( a, dp, c->stream, nullptr ), "launch generator" ) ) return fail( tLastError );
        if( !HipOk( hipMemcpyAsync( c->hKeys, c->dKeys, kCallRangeInts * sizeof( int32_t ), hipMemcpyDeviceToHost, c->stream ), "read min/max" ) ) return fail( tLastError );
        if( !devOut )
        {
            if( total >= kStagedMinFloats )
            {
                if( !c->StagedToHost( out, a.out, total ) ) return fail( tLastError );
            }
            else if( !HipOk( hipMemcpyAsync( out, a.out, (size_t)total * sizeof( float ), hipMemcpyDeviceToHost, c->stream ), "copy output" ) )
            {
                return fail( tLastError );
            }
        }
        if( !HipOk( hipStreamSynchronize( c->stream ), "generator kernel" ) ) return fail( tLastError );

        if( !DecodeRange( c->hKeys, kCallRangeSlots, kRangeStrideSpread, r.min, r.max ) )
        {
            r.min = FastNoise::kInfinity;
            r.max = -FastNoise::kInfinity;
        }
        return r;
    }
}

using namespace fn2;

namespace FastNoise
{
    static OutputMinMax ToMinMax( const OutputRange& r )
    {
        OutputMinMax m;
        m.min = r.min;
        m.max = r.max;
        return m;
    }

    // Generator.h:217-221 of the reference
    OutputMinMax Generator::GenUniformGrid2D( float* out, float xOffset, float yOffset, int xCount, int yCount, float xStepSize, float yStepSize, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_GRID; s.dims = 2; s.seed = seed;
        s.count[0] = xCount; s.count[1] = yCount;
        s.offset[0] = xOffset; s.offset[1] = yOffset;
        s.step[0] = xStepSize; s.step[1] = yStepSize;
        return ToMinMax( GenerateHost( this, s, out, nullptr ) );
    }

    // Generator.h:240-244 of the reference
    OutputMinMax Generator::GenUniformGrid3D( float* out, float xOffset, float yOffset, float zOffset, int xCount, int yCount, int zCount, float xStepSize, float yStepSize, float zStepSize, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_GRID; s.dims = 3; s.seed = seed;
        s.count[0] = xCount; s.count[1] = yCount; s.count[2] = zCount;
        s.offset[0] = xOffset; s.offset[1] = yOffset; s.offset[2] = zOffset;
        s.step[0] = xStepSize; s.step[1] = yStepSize; s.step[2] = zStepSize;
        return ToMinMax( GenerateHost( this, s, out, nullptr ) );
    }

    // Generator.h:266-270 of the reference
    OutputMinMax Generator::GenUniformGrid4D( float* out, float xOffset, float yOffset, float zOffset, float wOffset, int xCount, int yCount, int zCount, int wCount, float xStepSize, float yStepSize, float zStepSize, float wStepSize, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_GRID; s.dims = 4; s.seed = seed;
        s.count[0] = xCount; s.count[1] = yCount; s.count[2] = zCount; s.count[3] = wCount;
        s.offset[0] = xOffset; s.offset[1] = yOffset; s.offset[2] = zOffset; s.offset[3] = wOffset;
        s.step[0] = xStepSize; s.step[1] = yStepSize; s.step[2] = zStepSize; s.step[3] = wStepSize;
        return ToMinMax( GenerateHost( this, s, out, nullptr ) );
    }

    // Generator.h:285-287 of the reference
    OutputMinMax Generator::GenTileable2D( float* out, int xSize, int ySize, float xStepSize, float yStepSize, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_TILEABLE; s.dims = 2; s.seed = seed;
        s.count[0] = xSize; s.count[1] = ySize;
        s.step[0] = xStepSize; s.step[1] = yStepSize;
        return ToMinMax( GenerateHost( this, s, out, nullptr ) );
    }

    // Generator.h:301-304 of the reference
    OutputMinMax Generator::GenPositionArray2D( float* out, int count, const float* xPosArray, const float* yPosArray, float xOffset, float yOffset, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_POSITIONS; s.dims = 2; s.seed = seed;
        s.count[0] = count;
        s.offset[0] = xOffset; s.offset[1] = yOffset;
        const float* p[4] = { xPosArray, yPosArray, nullptr, nullptr };
        return ToMinMax( GenerateHost( this, s, out, p ) );
    }

    // Generator.h:320-323 of the reference
    OutputMinMax Generator::GenPositionArray3D( float* out, int count, const float* xPosArray, const float* yPosArray, const float* zPosArray, float xOffset, float yOffset, float zOffset, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_POSITIONS; s.dims = 3; s.seed = seed;
        s.count[0] = count;
        s.offset[0] = xOffset; s.offset[1] = yOffset; s.offset[2] = zOffset;
        const float* p[4] = { xPosArray, yPosArray, zPosArray, nullptr };
        return ToMinMax( GenerateHost( this, s, out, p ) );
    }

    // Generator.h:342-345 of the reference
    OutputMinMax Generator::GenPositionArray4D( float* out, int count, const float* xPosArray, const float* yPosArray, const float* zPosArray, const float* wPosArray, float xOffset, float yOffset, float zOffset, float wOffset, int seed ) const
    {
        fn2GenSpec s{};
        s.kind = FN2_SOURCE_POSITIONS; s.dims = 4; s.seed = seed;
        s.count[0] = count;
        s.offset[0] = xOffset; s.offset[1] = yOffset; s.offset[2] = zOffset; s.offset[3] = wOffset;
        const float* p[4] = { xPosArray, yPosArray, zPosArray, wPosArray };
        return ToMinMax( GenerateHost( this, s, out, p ) );
    }

    // GenSingle*: one point through the position path. The offset is -0.0f,
    // the additive identity, so the point reaches the tree bit for bit
    // (Generator.h:351-368 of the reference).
    float Generator::GenSingle2D( float x, float y, int seed ) const
    {
        float v = 0;
        GenPositionArray2D( &v, 1, &x, &y, -0.0f, -0.0f, seed );
        return v;
    }

    float Generator::GenSingle3D( float x, float y, float z, int seed ) const
    {
        float v = 0;
        GenPositionArray3D( &v, 1, &x, &y, &z, -0.0f, -0.0f, -0.0f, seed );
        return v;
    }

    float Generator::GenSingle4D( float x, float y, float z, float w, int seed ) const
    {
        float v = 0;
        GenPositionArray4D( &v, 1, &x, &y, &z, &w, -0.0f, -0.0f, -0.0f, -0.0f, seed );
        return v;
    }
}

// ---------------------------------------------------------------------------
// Device-side extension API (fn2_gpu.h)
// ---------------------------------------------------------------------------
extern "C" {

FASTNOISE_API const char* fn2GetLastError( void )
{
    return fn2::LastError();
}

FASTNOISE_API int fn2DeviceCount( void )
{
    int n = 0;
    if( hipGetDeviceCount( &n ) != hipSuccess )
    {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

FASTNOISE_API void* fn2Malloc( size_t bytes )
{
    ClearError();
    if( !EnsureDevice() ) return nullptr;
    void* p = nullptr;
    if( !HipOk( hipMalloc( &p, bytes ), "hipMalloc" ) ) return nullptr;
    return p;
}

FASTNOISE_API void fn2Free( void* p )
{
    if( p ) (void)hipFree( p );
}

FASTNOISE_API void* fn2MallocHost( size_t bytes )
{
    ClearError();
    if( !EnsureDevice() ) return nullptr;
    void* p = nullptr;
    if( !HipOk( hipHostMalloc( &p, bytes, hipHostMallocDefault ), "hipHostMalloc" ) ) return nullptr;
    return p;
}

FASTNOISE_API void fn2FreeHost( void* p )
{
    if( p ) (void)hipHostFree( p );
}

FASTNOISE_API int fn2Memcpy( void* dst, const void* src, size_t bytes, int kind )
{
    ClearError();
    if( !EnsureDevice() ) return 1;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    return HipOk( hipMemcpy( dst, src, bytes, k ), "hipMemcpy" ) ? 0 : 1;
}

FASTNOISE_API int fn2DeviceSynchronize( void )
{
    ClearError();
    if( !EnsureDevice() ) return 1;
    return HipOk( hipDeviceSynchronize(), "hipDeviceSynchronize" ) ? 0 : 1;
}

FASTNOISE_API int fn2GetProgramInfo( const void* node, int dims, int* instructions, int* registers, int* dimMask, int* featureMask )
{
    ClearError();
    if( !node ) { SetError( "null node" ); return 1; }
    Program p;
    std::string err;
    if( !CompileTree( static_cast<const FastNoise::Generator*>( node ), dims, p, err ) )
    {
        SetError( "cannot compile node tree: " + err );
        return 1;
    }
    if( instructions ) *instructions = (int)p.ins.size();
    if( registers ) *registers = p.regsUsed;
    if( dimMask ) *dimMask = p.dimMask;
    if( featureMask ) *featureMask = p.featMask;
    return 0;
}

FASTNOISE_API int fn2GenerateDevice( const void* node, const fn2GenSpec* spec, float* dOut, float* outputMinMax, void* stream )
{
    ClearError();
    if( !node || !spec ) { SetError( "null argument" ); return 1; }
    GenArgs a{};
    int64_t total = 0;
    std::string err;
    if( !EnsureDevice() ) return 1;
    if( !FillSpecArgs( *spec, a, total, err ) ) { SetError( err ); return 1; }
    if( total == 0 )
    {
        if( outputMinMax ) { outputMinMax[0] = FastNoise::kInfinity; outputMinMax[1] = -FastNoise::kInfinity; }
        return 0;
    }
    const DeviceProgram* dp = nullptr;
    if( !CompileForDevice( static_cast<const FastNoise::Generator*>( node ), a.dims, dp, nullptr ) ) return 1;
    CallContext* c = AcquireContext();
    if( !c ) return 1;
    struct Releaser { CallContext* c; ~Releaser() { ReleaseContext( c ); } } rel{ c };
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    a.ins = dp->dIns;
    a.konst = dp->dKonst;
    a.regs = dp->regs;
    a.nIns = (int32_t)dp->ins.size();
    a.resultReg = dp->resultReg;
    a.out = dOut;
    a.minmax = outputMinMax ? c->dKeys : nullptr;
    if( outputMinMax && !HipOk( (hipError_t)fn2_launch_minmax_init( c->dKeys, 1, kCallRangeSlots, kRangeStrideSpread, st ), "launch min/max init" ) ) return 1;
    if( !HipOk( LaunchProgram( a, dp, st, nullptr ), "launch generator" ) ) return 1;
    if( outputMinMax )
    {
        if( !HipOk( hipMemcpyAsync( c->hKeys, c->dKeys, kCallRangeInts * sizeof( int32_t ), hipMemcpyDeviceToHost, st ), "read min/max" ) ) return 1;
        if( !HipOk( hipStreamSynchronize( st ), "generator kernel" ) ) return 1;
        if( !DecodeRange( c->hKeys, kCallRangeSlots, kRangeStrideSpread, outputMinMax[0], outputMinMax[1] ) )
        {
            outputMinMax[0] = FastNoise::kInfinity;
            outputMinMax[1] = -FastNoise::kInfinity;
        }
    }
    else if( !stream )
    {
        // the library stream belongs to the context that is released below:
        // finish the work before another call can pick it up
        if( !HipOk( hipStreamSynchronize( st ), "generator kernel" ) ) return 1;
    }
    return 0;
}

FASTNOISE_API int fn2BenchGenerate( const void* node, const fn2GenSpec* spec, float* dOut, int warmup, int iters, double* msPerLaunch, double* wallMsPerLaunch )
{
    ClearError();
    if( !node || !spec || iters <= 0 ) { SetError( "bad argument" ); return 1; }
    GenArgs a{};
    int64_t total = 0;
    std::string err;
    if( !EnsureDevice() ) return 1;
    if( !FillSpecArgs( *spec, a, total, err ) ) { SetError( err ); return 1; }
    const DeviceProgram* dp = nullptr;
    if( !CompileForDevice( static_cast<const FastNoise::Generator*>( node ), a.dims, dp, nullptr ) ) return 1;
    CallContext* c = AcquireContext();
    if( !c ) return 1;
    struct Releaser { CallContext* c; ~Releaser() { ReleaseContext( c ); } } rel{ c };
    a.ins = dp->dIns;
    a.konst = dp->dKonst;
    a.regs = dp->regs;
    a.nIns = (int32_t)dp->ins.size();
    a.resultReg = dp->resultReg;
    a.out = dOut;
    a.minmax = c->dKeys;
    for( int i = 0; i < warmup; i++ )
    {
        if( !HipOk( (hipError_t)fn2_launch_minmax_init( c->dKeys, 1, kCallRangeSlots, kRangeStrideSpread, c->stream ), "launch" ) ) return 1;
        if( !HipOk( LaunchProgram( a, dp, c->stream, nullptr ), "launch" ) ) return 1;
    }
    if( !HipOk( hipStreamSynchronize( c->stream ), "warmup" ) ) return 1;
    // The range keys are reset once: every timed launch writes the same
    // output, so the keys accumulate the same min/max, and the timed region
    // holds generator launches only.
    if( !HipOk( (hipError_t)fn2_launch_minmax_init( c->dKeys, 1, kCallRangeSlots, kRangeStrideSpread, c->stream ), "launch" ) ) return 1;
    if( !HipOk( hipStreamSynchronize( c->stream ), "range reset" ) ) return 1;
    auto t0 = std::chrono::steady_clock::now();
    if( !HipOk( hipEventRecord( c->ev0, c->stream ), "event" ) ) return 1;
    for( int i = 0; i < iters; i++ )
    {
        if( !HipOk( LaunchProgram( a, dp, c->stream, nullptr ), "launch" ) ) return 1;
    }
    if( !HipOk( hipEventRecord( c->ev1, c->stream ), "event" ) ) return 1;
    if( !HipOk( hipStreamSynchronize( c->stream ), "timed launches" ) ) return 1;
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0;
    if( !HipOk( hipEventElapsedTime( &ms, c->ev0, c->ev1 ), "hipEventElapsedTime" ) ) return 1;
    if( msPerLaunch ) *msPerLaunch = ms / iters;
    if( wallMsPerLaunch ) *wallMsPerLaunch = std::chrono::duration<double, std::milli>( t1 - t0 ).count() / iters;
    return 0;
}

FASTNOISE_API int fn2SelfTestMath( int op, const float* dIn, float* dOut, int n )
{
    ClearError();
    if( op < 0 || op > 5 || n < 0 ) { SetError( "bad self-test arguments" ); return 1; }
    if( !EnsureDevice() ) return 1;
    if( !HipOk( (hipError_t)fn2_launch_selftest_math( op, dIn, dOut, n, nullptr ), "self-test launch" ) ) return 1;
    return HipOk( hipDeviceSynchronize(), "self-test" ) ? 0 : 1;
}

}
