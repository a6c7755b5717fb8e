// Kernel-side argument blocks of the GPU seed walk: the copy that lays each
// walk stream out contiguously, and the walk of one query per wavefront.
#ifndef TOPAZ_AMD_WALK_KERNELS_H
#define TOPAZ_AMD_WALK_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// n elements (scores and proteins) of one round's output, from element src
// of round `round`'s arrays to element dst of the walk's arrays.
struct CompactJob {
    int64_t src;
    int64_t dst;
    int32_t n;
    int32_t round;
};

struct CompactArgs {
    const CompactJob* jobs;
    int64_t n_jobs;
    const int16_t* const* round_vals;      // per round: its output arrays
    const uint32_t* const* round_prot;
    int16_t* vals;
    uint32_t* prot;
};

// Per query: its streams [q_stream[q], q_stream[q + 1]) in creation order,
// each stream's seeds at s_base (length s_len, cut short where s_open), the
// hit list's capacity, its WalkMem at byte q_mem[q] of mem (walk_mem_layout
// of the query's stream count, capacity and max_score), and room for
// capacity hits at entry q_out[q] of out_index / out_score.  Results: the
// hit count, the walk status (WalkStatus) and {pops, runs, counted}.
struct WalkArgs {
    const int64_t* s_base;
    const int32_t* s_len;
    const uint8_t* s_open;
    const int16_t* vals;
    const uint32_t* prot;
    const int32_t* q_stream;
    const int32_t* q_capacity;
    const int64_t* q_mem;
    char* mem;
    const int64_t* q_out;
    uint32_t* out_index;
    int32_t* out_score;
    int32_t* out_n;
    int32_t* out_status;
    int64_t* out_stats;
    int32_t n_queries;
    int32_t limit;             // number_of_seeds
    int32_t max_score;
};

}  // namespace tpz

extern "C" hipError_t tpz_launch_compact(const tpz::CompactArgs* args, hipStream_t stream);
extern "C" hipError_t tpz_launch_walk(const tpz::WalkArgs* args, hipStream_t stream);

#endif
