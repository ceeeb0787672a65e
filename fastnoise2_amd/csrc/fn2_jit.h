// Run-time specialised generator kernels (fn2_jit.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace fn2
{
    struct DeviceProgram;

    struct JitKernel
    {
        hipModule_t module = nullptr;
        hipFunction_t single = nullptr;   // fn2_jit_single( GenArgs )
        hipFunction_t batch = nullptr;    // fn2_jit_batch( jobs, blockJob, blockBase )
        hipFunction_t queue = nullptr;    // fn2_jit_queue( jobs, tileJob, tileBase, nTiles, counter )
        int blocksPerCU = 0;              // resident 256-thread blocks of `single` per CU
        int queueBlocksPerCU = 0;         // the same for `queue`
        int singleRegs = 0;               // VGPRs per lane of `single`, as the runtime reports them
        int singleScratch = 0;            // private (scratch) bytes per lane of `single`
    };

    // Variants of a program's specialised kernel: compile-time choices that
    // change its speed, never its results. Bits 0-1: the fused fractals'
    // octave-loop unroll (0 the compiler's choice, then 1, 2, 4); bit 2: the
    // shared far-cell walk of CellularDistance (for scattered positions).
    constexpr int kJitUnrollMask = 3;
    constexpr int kJitSharedWalk = 4;
    constexpr int kJitUnrollFactor[4] = { 0, 1, 2, 4 };

    // The specialised kernel of a program (variant `variant`) when it is
    // loaded, else null (and, in async mode, it is queued for compilation).
    const JitKernel* GetJitKernel( const DeviceProgram* p, int variant = 0 );

    // The variant's kernel if it is already loaded, else null; never queues a
    // compilation (for callers that must not allocate)
    const JitKernel* LoadedJitKernel( const DeviceProgram* p, int variant );

    // Compilation state of a variant: -1 never requested, 0 queued or
    // compiling, 1 loaded, 2 failed, 3 too long to specialise.
    int JitVariantState( const DeviceProgram* p, int variant );

    // Whether the program holds a fused fractal (the unroll variants differ)
    bool JitHasFusedFractal( const DeviceProgram& p );
    // Whether the shared far-cell walk variant differs from the default: a 3D
    // CellularDistance with squared distances, no size jitter and at most two
    // kept distance slots
    bool JitSharedWalkApplies( const DeviceProgram& p );

    // HIP source written for a program (exposed for diagnostics).
    std::string JitSourceFor( const DeviceProgram& p, int variant = 0 );
}
