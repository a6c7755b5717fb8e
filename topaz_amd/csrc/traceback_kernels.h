// Kernel-side argument blocks of the traceback-statistics kernels.
#ifndef TOPAZ_AMD_TRACEBACK_KERNELS_H
#define TOPAZ_AMD_TRACEBACK_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpz {

struct ReverseArgs {
    const int8_t* q_res;
    const int64_t* q_off;
    const int8_t* s_res;
    const int64_t* s_off;
    const int32_t* pair_query;
    const int32_t* pair_subject;
    const int32_t* read_end;       // read_end1 of each hit
    const int32_t* ref_end;        // ref_end1
    const int32_t* score;          // score1
    int32_t n_pairs;
    const int8_t* matrix;
    int32_t gap_open, gap_extend, bias;
    int32_t* workspace;            // 3 * (ref_end1 + 1) words per hit
    const int64_t* ws_off;
    int32_t* out_read_begin;
    int32_t* out_ref_begin;
    const int32_t* hits;           // hit numbers the launch covers (n_list of them); null: all
    int32_t n_list;
};

// Query rows per lane of the reverse-pass kernel for a hit whose (lane-
// padded) prefix has `rows` rows: the fewest of 8, 16, 32 holding it in one
// pass of 64 lanes; longer prefixes take passes of 32 rows per lane.
inline int reverse_rows_per_lane(int rows) { return rows <= 512 ? 8 : rows <= 1024 ? 16 : 32; }

struct BandArgs {
    const int8_t* q_res;
    const int64_t* q_off;
    const int8_t* s_res;
    const int64_t* s_off;
    const int32_t* pair_query;
    const int32_t* pair_subject;
    const int32_t* read_begin;
    const int32_t* read_end;
    const int32_t* ref_begin;
    const int32_t* ref_end;
    const int32_t* score;
    int32_t n_pairs;
    const int8_t* matrix;
    int32_t gap_open, gap_extend;
    int32_t* workspace;            // three band rows (H, E, H of the row) of
                                   // band_row_words() words each, per listed hit
    const int64_t* ws_off;
    int32_t* out_band;             // final band half-width, -1: size guard
    int8_t* dirs;                  // direction buffers, zeroed
    const int64_t* dir_off;
    const int64_t* dir_size;       // bytes of each hit's buffer
    int32_t* out_rc, *out_length, *out_gapopen, *out_mismatch, *out_identical;
    const int32_t* hits;           // hit numbers the launch covers (n_list of them); null: all
    int32_t n_list;
};

// Words of one band row in the workspace of a hit whose rectangle is
// ref_len x read_len.  The band half-width stays below ref_len + read_len
// until the search stops, and a row needs 2w + 3 slots.
__host__ __device__ inline int32_t band_row_words(int32_t ref_len, int32_t read_len) {
    return 4 * (ref_len + read_len) + 8;
}

// Band rows the wavefront kernels hold in LDS: a band wider than this goes
// to the one-thread-per-hit kernels (band status -2 from the search).
constexpr int kBandSlots = 2048;

}  // namespace tpz

// one wavefront per listed hit, all of them of rows-per-lane class rpl
extern "C" hipError_t tpz_launch_sw_reverse(const tpz::ReverseArgs* a, int rpl, hipStream_t stream);
// wave: one wavefront per hit with the band rows in LDS; otherwise one
// thread per hit with the rows in the workspace
extern "C" hipError_t tpz_launch_band_search(const tpz::BandArgs* a, int wave, hipStream_t stream);
extern "C" hipError_t tpz_launch_band_trace(const tpz::BandArgs* a, int wave, hipStream_t stream);

#endif
