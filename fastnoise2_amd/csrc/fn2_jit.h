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
    };

    // The specialised kernel of a program when it is loaded, else null (and,
    // in async mode, the program is queued for compilation).
    const JitKernel* GetJitKernel( const DeviceProgram* p );

    // HIP source written for a program (exposed for diagnostics).
    std::string JitSourceFor( const DeviceProgram& p );
}
