// Kernel argument block of the generator kernels (host and device).
#pragma once
#if !defined( __HIPCC_RTC__ )
#include <stdint.h>
#endif
#include "fn2_program.h"

namespace fn2
{
    struct GenArgs
    {
        const Ins* ins;          // device program
        const float* konst;      // device constant pool
        int32_t nIns;
        int32_t resultReg;
        int32_t dims;            // dimensionality of the starting position
        int32_t kind;            // SourceKind
        int32_t seed;
        int32_t total;           // points (the reference counts in int)
        float* out;              // device output, `total` floats
        int32_t* minmax;         // device ordered-int keys {min, max}, or null
        FastDiv div[3];          // grid: divisors by x, y, z count; tileable: div[0] = x size
        float offset[4];
        float step[4];
        const float* pos[4];     // position arrays (device); tileable: cos x, cos y, sin x, sin y tables
        int32_t ppt;             // batched kernel: points per lane of each block of this job
        int32_t regs;            // registers the program uses (the interpreter's file holds as many)
    };
}
