// Batched Smith-Waterman on gfx950 -- score and end positions.
//
// Computes, for many (query, subject) pairs, what TOPAZ's ssw_align returns
// (ssw.c:835-897 with score_size 2): the best local alignment score, the
// first subject column reaching it (ref_end1) and the lowest query row of
// that column reaching it (read_end1).
//
// The reference runs a 16-lane byte pass (sw_sse2_byte, ssw.c:127-332) and,
// when that saturates at 255 - bias, an 8-lane word pass (sw_sse2_word,
// ssw.c:358-536) that gives up at 32767.  Both compute, for gapO > gapE, the
// same H in every cell as textbook affine-gap local alignment:
//
//   H(i,j) = max(0, H(i-1,j-1) + W(s_i, q_j), E(i,j), F(i,j))
//   E(i+1,j) = max(E(i,j) - gapE, H(i,j) - gapO)      (gap along the subject)
//   F(i,j+1) = max(F(i,j) - gapE, H(i,j) - gapO)      (gap along the query)
//
// (their lazy-F loop drops a vertical-then-horizontal gap corner that crosses
// a lane boundary, but the horizontal-then-vertical corner of equal cost is
// always kept, so no cell value changes; the host refuses gapO <= gapE,
// where that no longer holds).  One exact pass in saturating int16 therefore
// gives the byte result when the score is below 255 - bias, the word result
// below 32767, and the reference's NULL at 32767.
//
// Mapping (inter-task): one wavefront = one query and up to 128 subjects;
// lane l holds two pairs, one in each 16-bit half of every VGPR, so each
// VOP3P instruction advances two DP cells.  Query rows are processed in
// chunks of 32 consecutive rows held in registers (H, E) while the lane walks
// the subject; a chunk's bottom row (H, F) is written per column to a
// per-wave workspace that the next chunk reads back in place.  The last
// chunk is padded with rows scoring -8192, which never reach a column's
// running maximum.
//
// End positions: within a chunk, a column whose max rises strictly above the
// chunk's best so far is an "event"; on an event the lane's 32 H values of
// that half are saved to LDS.  At the chunk end the saved column gives the
// lowest row; chunks combine by (larger score, then smaller column, then the
// earlier chunk), which is the reference's first-column / lowest-row rule.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_kernels.h"

namespace tpz {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));

constexpr int kRows = 32;          // query rows per chunk, held in registers
constexpr int kNRes = 26;          // 25 matrix residues + 1 padding residue
constexpr int kPadResidue = 25;
constexpr int kPadScore = -8192;   // score of padding rows / columns

__device__ __forceinline__ s2 add_sat(s2 a, s2 b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ s2 vmax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s2 usub(s2 a, s2 b) {
    return __builtin_bit_cast(s2, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, a),
                                                                __builtin_bit_cast(u2, b)));
}
__device__ __forceinline__ uint32_t bits(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2 from_bits(uint32_t b) { return __builtin_bit_cast(s2, b); }

// LDS of one wavefront.
struct WaveLds {
    int16_t prof[kRows * kNRes];           // W(s, q[j0 + r]) at [r * 26 + s]
    uint16_t saved[2][kRows][64];          // H column of the chunk best, per half
};

// Best of the chunks finished so far, per half of the lane.
struct LaneBest {
    int score[2];
    int ref[2];
    int read[2];
};

__device__ __forceinline__ uint32_t load4(const int8_t* s, int i, int n) {
    // subjects start on 4-byte boundaries and are padded with residue 25
    return i < n ? *reinterpret_cast<const uint32_t*>(s + i) : 0x19191919u;
}

// Event bookkeeping for one finished column: if a half's column max rises
// strictly above the chunk best, remember the column and save that half of
// the column's H values.
__device__ __forceinline__ void column_event(WaveLds& lds, const s2 (&Hc)[kRows], s2 cm,
                                             s2& bc, int col, int& ci_lo, int& ci_hi) {
    const int lane = threadIdx.x;
    if (cm.x > bc.x) {
        ci_lo = col;
#pragma unroll
        for (int r = 0; r < kRows; ++r) lds.saved[0][r][lane] = (uint16_t)Hc[r].x;
    }
    if (cm.y > bc.y) {
        ci_hi = col;
#pragma unroll
        for (int r = 0; r < kRows; ++r) lds.saved[1][r][lane] = (uint16_t)Hc[r].y;
    }
    bc = vmax(bc, cm);
}

// One chunk of kRows rows [j0, j0 + kRows) over all subject columns, two
// columns at a time: the second column runs one row behind the first, so
// every step advances two independent dependency chains.
__device__ __forceinline__ void run_chunk(const SwKernelArgs& a, WaveLds& lds,
                                          uint32_t* ws, int ncols,
                                          const int8_t* s_lo, int n_lo,
                                          const int8_t* s_hi, int n_hi,
                                          bool first, bool store_bottom, int j0,
                                          LaneBest& best) {
    const int lane = threadIdx.x;
    const s2 go = s2{(short)a.gap_open, (short)a.gap_open};
    const s2 ge = s2{(short)a.gap_extend, (short)a.gap_extend};
    const s2 zero = s2{0, 0};
    s2 E[kRows], H[kRows], Ha[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) { E[r] = zero; H[r] = zero; }
    s2 bc = zero;                 // chunk best per half
    int ci_lo = -1, ci_hi = -1;
    s2 hdiag = zero;              // H(i-1, j0-1)

    uint32_t* wsH = ws;                         // [col][lane]
    uint32_t* wsF = ws + (size_t)ncols * 64;

    uint32_t res_lo = load4(s_lo, 0, n_lo), res_hi = load4(s_hi, 0, n_hi);
    uint32_t ta_h = 0, ta_f = 0, tb_h = 0, tb_f = 0;
    if (!first) {
        ta_h = wsH[lane]; ta_f = wsF[lane];
        tb_h = wsH[64 + lane]; tb_f = wsF[64 + lane];
    }

    // ncols is even (the host pads subjects with residue 25 columns)
    for (int i = 0; i < ncols; i += 2) {
        uint32_t na_h = 0, na_f = 0, nb_h = 0, nb_f = 0;
        if (!first && i + 2 < ncols) {
            na_h = wsH[(size_t)(i + 2) * 64 + lane];
            na_f = wsF[(size_t)(i + 2) * 64 + lane];
            nb_h = wsH[(size_t)(i + 3) * 64 + lane];
            nb_f = wsF[(size_t)(i + 3) * 64 + lane];
        }
        const int sh = (i & 3) * 8;
        const int ra_lo = (res_lo >> sh) & 0xff, ra_hi = (res_hi >> sh) & 0xff;
        const int rb_lo = (res_lo >> (sh + 8)) & 0xff, rb_hi = (res_hi >> (sh + 8)) & 0xff;
        if ((i & 3) == 2) {
            res_lo = load4(s_lo, i + 2, n_lo);
            res_hi = load4(s_hi, i + 2, n_hi);
        }
        const int16_t* pa_lo = lds.prof + ra_lo;
        const int16_t* pa_hi = lds.prof + ra_hi;
        const int16_t* pb_lo = lds.prof + rb_lo;
        const int16_t* pb_hi = lds.prof + rb_hi;

        s2 fa = from_bits(ta_f), fb = from_bits(tb_f);
        s2 da = hdiag;                  // H(i-1, j0-1)
        s2 db = from_bits(ta_h);        // H(i, j0-1)
        hdiag = from_bits(tb_h);        // H(i+1, j0-1)
        s2 cma = zero, cmb = zero;

        // column a, row 0
        {
            s2 w = s2{pa_lo[0], pa_hi[0]};
            s2 h = vmax(vmax(add_sat(da, w), E[0]), fa);
            da = H[0];
            Ha[0] = h;
            cma = h;
            s2 hg = usub(h, go);
            E[0] = vmax(usub(E[0], ge), hg);
            fa = vmax(usub(fa, ge), hg);
        }
#pragma unroll
        for (int r = 1; r < kRows; ++r) {
            // column a, row r
            {
                s2 w = s2{pa_lo[r * kNRes], pa_hi[r * kNRes]};
                s2 h = vmax(vmax(add_sat(da, w), E[r]), fa);
                da = H[r];
                Ha[r] = h;
                cma = vmax(cma, h);
                s2 hg = usub(h, go);
                E[r] = vmax(usub(E[r], ge), hg);
                fa = vmax(usub(fa, ge), hg);
            }
            // column b, row r - 1
            {
                const int q = r - 1;
                s2 w = s2{pb_lo[q * kNRes], pb_hi[q * kNRes]};
                s2 h = vmax(vmax(add_sat(db, w), E[q]), fb);
                db = Ha[q];
                H[q] = h;
                cmb = vmax(cmb, h);
                s2 hg = usub(h, go);
                E[q] = vmax(usub(E[q], ge), hg);
                fb = vmax(usub(fb, ge), hg);
            }
        }
        // column b, last row
        {
            const int q = kRows - 1;
            s2 w = s2{pb_lo[q * kNRes], pb_hi[q * kNRes]};
            s2 h = vmax(vmax(add_sat(db, w), E[q]), fb);
            H[q] = h;
            cmb = vmax(cmb, h);
            s2 hg = usub(h, go);
            E[q] = vmax(usub(E[q], ge), hg);
            fb = vmax(usub(fb, ge), hg);
        }
        if (store_bottom) {
            wsH[(size_t)i * 64 + lane] = bits(Ha[kRows - 1]);
            wsF[(size_t)i * 64 + lane] = bits(fa);
            wsH[(size_t)(i + 1) * 64 + lane] = bits(H[kRows - 1]);
            wsF[(size_t)(i + 1) * 64 + lane] = bits(fb);
        }
        column_event(lds, Ha, cma, bc, i, ci_lo, ci_hi);
        column_event(lds, H, cmb, bc, i + 1, ci_lo, ci_hi);
        ta_h = na_h; ta_f = na_f; tb_h = nb_h; tb_f = nb_f;
    }

    // Lowest row of the saved column equal to the chunk best, then merge:
    // larger score wins; on an equal score the smaller column; on an equal
    // column the earlier chunk (already held).
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int b = half ? bc.y : bc.x;
        const int ci = half ? ci_hi : ci_lo;
        if (b <= 0) continue;
        int row = kRows;
#pragma unroll
        for (int r = kRows - 1; r >= 0; --r)
            if ((int)(int16_t)lds.saved[half][r][lane] == b) row = r;
        if (b > best.score[half] || (b == best.score[half] && ci < best.ref[half])) {
            best.score[half] = b;
            best.ref[half] = ci;
            best.read[half] = j0 + row;
        }
    }
}

// Substitution scores of rows j0 .. j0+31 for all 26 residues; rows past
// the query end score kPadScore.
__device__ __forceinline__ void build_profile(const SwKernelArgs& a, WaveLds& lds,
                                              const int8_t* q, int m, int j0) {
    const int lane = threadIdx.x;
    for (int e = lane; e < kRows * kNRes; e += 64) {
        const int r = e / kNRes, s = e - r * kNRes;
        const int j = j0 + r;
        int w = kPadScore;
        if (s < kPadResidue && j < m) w = a.matrix[s * 25 + q[j]];
        lds.prof[e] = (int16_t)w;
    }
}

// One task = one query and up to 128 of its pairs.  With a task counter the
// grid is persistent: each workgroup (one wavefront) pulls tasks until none
// are left and owns workspace slot blockIdx.x; without one, workgroup b runs
// task b.
__global__ void __launch_bounds__(64)
sw_forward_kernel(SwKernelArgs a) {
    __shared__ WaveLds lds;
    const int lane = threadIdx.x;
    uint32_t* ws = a.workspace + (size_t)blockIdx.x * a.ws_words_per_slot;
    for (;;) {
        int task;
        if (a.task_counter) {
            int t0 = 0;
            if (lane == 0) t0 = atomicAdd(a.task_counter, 1);
            task = __builtin_amdgcn_readfirstlane(t0);
        } else {
            task = blockIdx.x;
        }
        if (task >= a.n_tasks) break;
        const int p0 = a.task_begin[task], p1 = a.task_begin[task + 1];
        const int np = p1 - p0;
        const bool has_lo = lane < np, has_hi = lane + 64 < np;
        const int k_lo = p0 + (has_lo ? lane : 0);
        const int k_hi = p0 + (has_hi ? lane + 64 : 0);
        const int pair_lo = a.pair_index ? a.pair_index[k_lo] : k_lo;
        const int pair_hi = a.pair_index ? a.pair_index[k_hi] : k_hi;
        const int qid = a.pair_query[pair_lo];
        const int8_t* q = a.q_res + a.q_off[qid];
        const int m = a.q_len[qid];
        const int sid_lo = a.pair_subject[pair_lo], sid_hi = a.pair_subject[pair_hi];
        const int8_t* s_lo = a.s_res + a.s_off[sid_lo];
        const int8_t* s_hi = a.s_res + a.s_off[sid_hi];
        const int n_lo = has_lo ? a.s_len[sid_lo] : 0;
        const int n_hi = has_hi ? a.s_len[sid_hi] : 0;
        int ncols = n_lo > n_hi ? n_lo : n_hi;
        for (int off = 32; off >= 1; off >>= 1) {
            const int o = __shfl_xor(ncols, off);
            ncols = o > ncols ? o : ncols;
        }
        ncols = __builtin_amdgcn_readfirstlane(ncols);
        ncols += ncols & 1;            // columns go in pairs; the pad column scores kPadScore

        // score 0: ref_end1 -1 and read_end1 0, or -1 for an empty query
        // (ssw.c:148-149, 288-297)
        LaneBest best;
        for (int h = 0; h < 2; ++h) {
            best.score[h] = 0; best.ref[h] = -1; best.read[h] = m > 0 ? 0 : -1;
        }
        for (int j0 = 0; j0 < m; j0 += kRows) {
            __syncthreads();
            build_profile(a, lds, q, m, j0);
            __syncthreads();
            run_chunk(a, lds, ws, ncols, s_lo, n_lo, s_hi, n_hi, j0 == 0,
                      j0 + kRows < m, j0, best);
        }

        if (has_lo) {
            a.out_score[pair_lo] = best.score[0] >= 32767 ? 32767 : best.score[0];
            a.out_ref[pair_lo] = best.ref[0];
            a.out_read[pair_lo] = best.read[0];
        }
        if (has_hi) {
            a.out_score[pair_hi] = best.score[1] >= 32767 ? 32767 : best.score[1];
            a.out_ref[pair_hi] = best.ref[1];
            a.out_read[pair_hi] = best.read[1];
        }
        if (!a.task_counter) break;
    }
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_sw_forward(const tpz::SwKernelArgs* args, int grid,
                                            hipStream_t stream) {
    hipLaunchKernelGGL(tpz::sw_forward_kernel, dim3(grid), dim3(64), 0, stream, *args);
    return hipGetLastError();
}
