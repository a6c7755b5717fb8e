// Kernel-side argument block of the one-pair-per-wavefront forward
// Smith-Waterman kernel (the plain schedule's kernel).
#ifndef TOPAZ_AMD_SW_PAIR_KERNELS_H
#define TOPAZ_AMD_SW_PAIR_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

// Residues are substitution-matrix indices 0..24, W[s * 25 + q].  Workgroup
// b works pair pairs[b] (or pair b when pairs is null).  A query longer
// than one pass (64 x rows) needs 2 * s_len words of workspace per pair, at
// workspace + ws_off[b].
struct PairArgs {
    const int8_t* q_res;
    const int64_t* q_off;
    const int32_t* q_len;
    const int8_t* s_res;
    const int64_t* s_off;
    const int32_t* s_len;
    const int32_t* pair_query;
    const int32_t* pair_subject;
    const int32_t* pairs;
    int32_t n_list;
    const int8_t* matrix;
    int32_t gap_open, gap_extend;
    int32_t* workspace;            // may be null when every query fits one pass
    const int64_t* ws_off;
    int32_t* out_score;
    int32_t* out_ref;
    int32_t* out_read;
};

}  // namespace tpz

// Query rows per lane for a query of m rows: 8, 16 or 32 (one pass holds
// 64 times as many rows).
extern "C" int tpz_sw_pair_rows(int32_t m);
extern "C" hipError_t tpz_launch_sw_pair(const tpz::PairArgs* a, int rows, hipStream_t stream);

#endif
