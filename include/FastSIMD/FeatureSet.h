// Feature-set type used in FastNoise2's public signatures.
//
// In the reference this enum comes from the external FastSIMD library
// (FastSIMD::FeatureSet, pinned by src/CMakeLists.txt:2-6 of FastNoise2) and
// selects a CPU SIMD level. Here it selects the execution backend of a node:
// every value at or above GPU runs on the GPU (the default, Max), anything
// lower is rejected at generation time with an error. The numeric values of
// the reference's enum are not part of its headers, so these are this
// library's own.
#pragma once
#include <cstdint>

namespace FastSIMD
{
    enum class FeatureSet : uint32_t
    {
        Invalid = 0,
        Scalar = 1,
        GPU = 0x1000,
        Max = 0xFFFFFFFF,
    };

    constexpr FeatureSet FeatureSetDefault() { return FeatureSet::GPU; }
}
