// Test infrastructure only: lane-16 build of the CPU oracle, compiled twice
// (AVX2+FMA and AVX-512 flags, each in its own namespace) for the host SIMD
// baseline. The caller checks the CPU features before calling in.
#include "fn2_oracle_run.h"

namespace fn2o_simd
{
    void FN2O_SIMD_FN( const fn2o_data::DriverArgs& a, int mode, int threads, float* minmax )
    {
        FN2O_NS::run_mode<16>( a, mode, threads, minmax );
    }
}
