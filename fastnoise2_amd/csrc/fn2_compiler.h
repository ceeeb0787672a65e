// Tree compiler: node tree -> flat device program (see fn2_program.h).
#pragma once
#include <string>
#include <vector>
#include <cstdint>
#include "fn2_program.h"

namespace FastNoise { class Generator; }

namespace fn2
{
    struct Program
    {
        std::vector<Ins> ins;
        std::vector<float> konst;
        int resultReg = 0;
        int dimMask = 0;     // bit d set: d-dimensional ops used
        int featMask = 0;    // kFeatCell | kFeatWarp | kFeatMisc
        int regsUsed = 0;
        int dims = 0;        // dimensionality of the starting positions
        uint64_t hash = 0;   // of ins + konst + dims, the device cache key
    };

    // Compiles `root` for evaluation at `dims`-dimensional positions that the
    // kernel places in registers 0..dims-1. Returns false with `error` set
    // when the tree cannot be expressed (missing source, over 32 live
    // registers, an input this backend does not support).
    bool CompileTree( const FastNoise::Generator* root, int dims, Program& out, std::string& error );
}
