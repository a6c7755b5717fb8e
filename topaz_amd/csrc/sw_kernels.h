// Kernel-side argument blocks of the batched Smith-Waterman kernels.
#ifndef TOPAZ_AMD_SW_KERNELS_H
#define TOPAZ_AMD_SW_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// Device pointers of one forward pass over a batch of (query, subject) pairs.
// Residues are substitution-matrix indices 0..24; subjects start on 16-byte
// boundaries of s_res and are padded with residue 25 to the next one, and at
// least 16 bytes follow the last subject.  A task is a range of entries
// [task_begin[t], task_begin[t+1]) of pair_index (or of the pair numbers
// themselves when pair_index is null) whose pairs belong to at most four
// queries, listed in task_query (4 per task, -1 past the last).  The first
// min(128, entries) entries of a task start the runs of its wavefront's 128
// lane halves (entry p0 + k: lane k % 64, half k / 64); entry_rec chains each
// run's further entries, whose pairs follow along the columns (item_cols(n)
// = max(8, n rounded up to 4) columns per pair); task_cols gives the columns
// the task's wavefront walks (the longest run).  A work item is either one
// task worked by a whole workgroup, or up to four tasks, one per wavefront.
// Every pair's best goes to out_key[entry] (zero before the launch) as a
// (score, -column, -row) key; sw_decode_kernel turns the keys into out_score,
// out_ref and out_read.
struct SwKernelArgs {
    const int8_t* q_res;
    const int64_t* q_off;
    const int32_t* q_len;
    const int8_t* s_res;
    int64_t s_pad;                 // s_res + s_pad: 16 residues 25 (padding columns)
    const int64_t* s_off;
    const int32_t* s_len;
    const int32_t* pair_query;
    const int32_t* pair_subject;
    const int32_t* pair_index;     // may be null
    const int4* entry_rec;         // per entry: subject offset / 16, subject length,
                                   // next entry of its run (-1: none), query slot (0..3)
    const int32_t* task_cols;      // per task: columns its wavefront walks, a multiple of 4
    unsigned long long* out_key;   // per entry: best key of its pair; zero before the launch
    const int32_t* task_begin;     // task t = entries [task_begin[t], task_begin[t+1])
    const int32_t* task_query;     // 4 query numbers per task, -1 = none
    const int32_t* item_task;      // 4 entries per work item (see sw_forward_kernel)
    int32_t n_items;
    int32_t* task_counter;         // null: workgroup b runs item b
    int32_t* error;                // set when a cross-wavefront wait gave up
    const int8_t* matrix;          // 25 x 25, W[s * 25 + q]
    int32_t gap_open;
    int32_t gap_extend;
    uint32_t* workspace;           // ws_words_per_slot words per workgroup
    int64_t ws_words_per_slot;
    int32_t* out_score;
    int32_t* out_ref;
    int32_t* out_read;
};

}  // namespace tpz

extern "C" hipError_t tpz_launch_fill_i32(int32_t* p, int32_t v, int32_t n, hipStream_t stream);

// Workgroups of sw_forward_kernel<rows, waves> resident per CU (registers
// and LDS); 0 on error.
extern "C" int tpz_sw_forward_groups_per_cu(int rows, int waves);

// out_score, out_ref and out_read of every entry's pair from out_key.
extern "C" hipError_t tpz_launch_sw_decode(const tpz::SwKernelArgs* args, int32_t n_entries,
                                           hipStream_t stream);

// rows: query rows per chunk, 16 or 32; waves: wavefronts per workgroup, 4
// or 8 (work items then have `waves` entries).
extern "C" hipError_t tpz_launch_sw_forward(const tpz::SwKernelArgs* args, int grid, int rows,
                                            int waves, hipStream_t stream);

#endif
