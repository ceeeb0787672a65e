// Shared pieces of the host backend (fn2_backend.cpp) used by the C API and
// the serving engine.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <string>
#include <vector>
#include "fn2_compiler.h"
#include "fn2_kernel_args.h"
#include "fn2_gpu.h"

namespace FastNoise { class Generator; }

namespace fn2
{
    // A compiled program resident in device memory (kept for the process).
    struct DeviceProgram
    {
        std::vector<Ins> ins;
        std::vector<float> konst;
        int resultReg = 0;
        int dimMask = 0;
        int featMask = 0;
        int dims = 0;    // of the starting positions (programs are cached per dimensionality)
        int regs = 1;    // registers used: every operand lies below this
        Ins* dIns = nullptr;
        float* dKonst = nullptr;
        // the specialised-kernel variant single calls run (fn2_jit.h; same
        // results): set by fn2SetJitVariant
        mutable std::atomic<int> variant{ 0 };
    };

    // Host buffers of at least this many floats move through the pinned
    // staging ring of a call context in chunks: the DMA of one chunk overlaps
    // the multi-threaded copy of the previous one between pinned and pageable
    // memory (a single pageable hipMemcpy runs at about 10 GB/s).
    constexpr int kStageSlots = 3;
    constexpr int64_t kStageChunkFloats = 1 << 21;   // 8 MiB
    constexpr int64_t kStagedMinFloats = 1 << 18;    // 1 MiB

    // Per-call stream and staging buffers, pooled.
    struct CallContext
    {
        hipStream_t stream = nullptr;
        float* dOut = nullptr;
        size_t outCap = 0;
        float* dPos = nullptr;
        size_t posCap = 0;
        int32_t* dKeys = nullptr;   // one range record, slots kRangeStrideSpread ints apart
        int32_t* hKeys = nullptr;   // pinned, the same size
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        float* hStage = nullptr;    // pinned ring: kStageSlots chunks of kStageChunkFloats
        hipEvent_t evStage[kStageSlots] = {};

        bool Reserve( float*& buf, size_t& cap, size_t floats );
        bool ReserveStage();
        // device -> pageable host and pageable host -> device through the ring,
        // on the context's stream; both return when the host side is done
        bool StagedToHost( float* dst, const float* dSrc, int64_t count );
        bool StagedToDevice( float* dDst, const float* src, int64_t count );
    };

    struct OutputRange
    {
        float min = __builtin_inff();
        float max = -__builtin_inff();
        bool ok = true;
    };

    void SetError( const std::string& msg );
    void ClearError();
    const char* LastError();
    bool EnsureDevice();
    bool HipOk( hipError_t e, const char* what );
    // synchronous host-to-device copy on the library's non-blocking upload stream
    hipError_t UploadSync( void* dst, const void* src, size_t bytes );

    FastDiv MakeFastDiv( uint32_t d );
    float DecodeKey( int32_t k );
    // The range a record holds: false when no number was published to it
    // (every output NaN, or none), else its min and max
    bool DecodeRange( const int32_t* record, int slots, int stride, float& min, float& max );
    // ints of a single call's range record
    constexpr int kCallRangeInts = kCallRangeSlots * kRangeStrideSpread;
    // Launch arguments of a spec (everything but the program and the output).
    // A tileable spec uses `tileTable` (device floats laid out as FillTileTable
    // writes them) when it is given, else the process's cached table of its
    // size and step, which the first call creates.
    bool FillSpecArgs( const fn2GenSpec& s, GenArgs& a, int64_t& total, std::string& err, const float* tileTable = nullptr );
    // GenTileable2D's torus coordinates for xSize columns and ySize rows, on
    // the host: [cos x | sin x] (2 * xSize floats), then [cos y | sin y]
    void FillTileTable( int xSize, int ySize, float xStep, float yStep, float* table );

    // Checks a compiled program before it reaches a kernel: instructions,
    // register operands (regs = the register count it needs) and constant
    // windows; false with `err` set when it is malformed
    bool ValidateProgram( const Program& p, int& regs, std::string& err );
    const DeviceProgram* GetDeviceProgram( const Program& p );
    // The cached device program of `p` if one exists, else null (never creates one)
    const DeviceProgram* FindDeviceProgram( const Program& p );
    bool CompileForDevice( const FastNoise::Generator* root, int dims, const DeviceProgram*& dp, Program* hostCopy );

    // One generator launch: the tree's specialised kernel when it is loaded,
    // the interpreter otherwise (`specialised` reports which).
    hipError_t LaunchProgram( const GenArgs& a, const DeviceProgram* dp, hipStream_t stream, bool* specialised );

    CallContext* AcquireContext();
    void ReleaseContext( CallContext* c );
    bool IsDevicePointer( const void* p );

    OutputRange GenerateHost( const FastNoise::Generator* root, const fn2GenSpec& spec, float* out, const float* const* hostPos );
}
