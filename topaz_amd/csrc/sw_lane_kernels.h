// Kernel-side arguments of the striped-lane Smith-Waterman kernel: the
// reference's own lane recurrence, for gap settings with gap_open <=
// gap_extend, where it differs from textbook affine-gap alignment.
#ifndef TOPAZ_AMD_SW_LANE_KERNELS_H
#define TOPAZ_AMD_SW_LANE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_kernels.h"

namespace tpz {

struct LaneArgs {
    const int8_t* q_res;       // matrix rows of the queries' residues
    const int64_t* q_off;
    const int32_t* q_len;
    const int8_t* s_res;       // and of the subjects'
    const int64_t* s_off;
    const int32_t* s_len;
    const int32_t* pair_query;
    const int32_t* pair_subject;
    const int32_t* pairs;      // the pair numbers this launch covers (n_list of them); null: 0 .. n_list-1
    int64_t n_list;
    const int8_t* matrix;      // 25 x 25
    int32_t gap_open, gap_extend, bias;
    // Reverse pass of the statistics (ssw.c:900-941) when set: per pair
    // read_end1, ref_end1 and score1 of its forward alignment.  The query is
    // then the reversed prefix q[read_end1 .. 0], the columns ref_end1 .. 0,
    // the lane count 16 when bias + score1 < 255 (else 8), and the scan stops
    // after the first column whose maximum equals score1.
    const int32_t* read_end;
    const int32_t* ref_end;
    const int32_t* score;
    int32_t* workspace;        // 3 * ws_rows words per group of 16 lanes
    int64_t ws_rows;           // >= the query (or prefix) length rounded up to 16
    // forward: score (32767 where the word pass saturates too), ref_end1,
    // read_end1; reverse: out_ref = ref_begin1, out_read = read_begin1
    int32_t* out_score;
    int32_t* out_ref;
    int32_t* out_read;
};

// Lanes a launch runs at once: 4 pairs per 64-lane wavefront.
constexpr int kLaneGroupsPerBlock = 4;

// Rows a pair of query (or prefix) length m needs in the workspace.
inline int64_t lane_rows(int64_t m) { return (m + 15) / 16 * 16; }

// The lane kernel's arguments over the forward kernel's device arrays
// (results included).
inline LaneArgs lane_args(const SwKernelArgs& a, int32_t bias, int32_t* ws, int64_t ws_rows) {
    LaneArgs la{};
    la.q_res = a.q_res; la.q_off = a.q_off; la.q_len = a.q_len;
    la.s_res = a.s_res; la.s_off = a.s_off; la.s_len = a.s_len;
    la.pair_query = a.pair_query; la.pair_subject = a.pair_subject;
    la.matrix = a.matrix;
    la.gap_open = a.gap_open; la.gap_extend = a.gap_extend; la.bias = bias;
    la.workspace = ws; la.ws_rows = ws_rows;
    la.out_score = a.out_score; la.out_ref = a.out_ref; la.out_read = a.out_read;
    return la;
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_sw_lane(const tpz::LaneArgs* args, int blocks, hipStream_t stream);

#endif
