// Kernel-side argument block of the batched gapless extension kernel.
#ifndef TOPAZ_AMD_GAPLESS_KERNELS_H
#define TOPAZ_AMD_GAPLESS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// One seed: query number, offset in the query, and a text position (or a
// suffix-array slot when GaplessArgs::sa is set).  16 bytes.
struct SeedRec {
    int32_t query;
    int32_t offset;
    int64_t position;
};

struct GaplessArgs {
    const uint8_t* text;        // t_size bytes, proteins separated by '\n'
    int64_t t_size;
    const int64_t* sa;          // may be null: positions are text positions
    const uint8_t* q_chars;     // queries, internal alphabet characters
    const int64_t* q_off;
    const int32_t* q_len;
    const SeedRec* seeds;
    int64_t n_seeds;
    const int8_t* text_map;     // 128 entries: text byte -> matrix index
    const int8_t* query_map;    // 128 entries: query byte -> matrix index
    const int8_t* matrix;       // 25 x 25
    int32_t* out_score;
    int32_t* out_length;
};

}  // namespace tpz

extern "C" hipError_t tpz_launch_gapless(const tpz::GaplessArgs* args, int grid,
                                         hipStream_t stream);

#endif
