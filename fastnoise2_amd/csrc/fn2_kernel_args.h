// Kernel argument block of the generator kernels (host and device).
#pragma once
#if !defined( __HIPCC_RTC__ )
#include <stdint.h>
#endif
#include "fn2_program.h"

namespace fn2
{
    // A range record: the output range of one call as ordered-int keys, kept
    // in `slots` {min, max} pairs `stride` ints apart (GenArgs::rangeSlots,
    // a power of two, and GenArgs::rangeStride). A block (or a wave)
    // publishes to the slot of its index, so that its atomics do not queue
    // behind every other block's on one address: device-wide atomics on one
    // 128-byte line complete one after another, about 11 ns each on MI355X.
    // The range is the smallest min key and the largest max key over the
    // slots. Measured (1 slot -> 16 compact slots -> 16 slots on their own
    // lines): Perlin 64^3 15.0 -> 9.1 -> 4.4 us, a 1 M-point 4D call 56 ->
    // 41 -> 40 us; 512 blocks storing 1 MB took 3.22 us with 16 lines, 2.69
    // with 64 and 2.64 with 256 (2.22 without a record). Single calls take
    // kCallRangeSlots spread slots; the serving engine, which reads a
    // record back per request, kRangeSlots compact (or, persistent, spread)
    // ones.
    constexpr int kRangeSlots = 16;
    constexpr int kCallRangeSlots = 64;
    constexpr int kRangeStrideSpread = 32;    // a slot per 128-byte line
    constexpr int kRangeStrideCompact = 2;    // slots side by side

    // A work-queue launch's counters: kQueueCounters of them, each on its own
    // 128-byte line (kQueueCounterStride uint32 apart); counter c hands out
    // the c-th sixteenth of the items beyond the first one per wave.
    constexpr int kQueueCounters = 16;
    constexpr int kQueueCounterStride = 32;
    constexpr int kQueueCounterInts = kQueueCounters * kQueueCounterStride;

    // The work-queue partition of a launch of `waves` waves over `items`
    // items: wave g runs item g first; the items beyond the first round are
    // split into min(waves, kQueueCounters) contiguous parts, one per
    // counter, and wave g pulls the items of part g % parts, [first, end),
    // from that part's counter. Every part has a wave for any grid size.
    struct QueuePart
    {
        uint32_t part, first, end;
    };
#if defined( __HIPCC__ ) || defined( __HIPCC_RTC__ )
    __host__ __device__
#endif
    inline QueuePart queuePart( uint32_t g, uint32_t waves, uint32_t items )
    {
        const uint32_t parts = waves < (uint32_t)kQueueCounters ? waves : (uint32_t)kQueueCounters;
        const uint32_t extra = items > waves ? items - waves : 0;
        const uint32_t chunk = ( extra + parts - 1 ) / parts;
        QueuePart q;
        q.part = g % parts;
        q.first = waves + q.part * chunk;
        q.end = q.first + chunk < items ? q.first + chunk : items;
        return q;
    }

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
        int32_t* minmax;         // device range record (kRangeInts ordered-int keys), or null
        FastDiv div[3];          // grid: divisors by x, y, z count; tileable: div[0] = x size
        float offset[4];
        float step[4];
        const float* pos[4];     // position arrays (device); tileable: cos x, cos y, sin x, sin y tables
        int32_t ppt;             // batched kernel: points per lane of each block of this job
        int32_t regs;            // registers the program uses (the interpreter's file holds as many)
        int32_t rangeStride;     // ints between the slots of the range record
        int32_t rangeSlots;      // slots of the range record (a power of two)
        // index of this job's first point in its request: a request split
        // into several jobs evaluates points base + [0, total) of the whole
        // grid or image (position arrays are offset by the host instead)
        uint32_t base;
    };
}
