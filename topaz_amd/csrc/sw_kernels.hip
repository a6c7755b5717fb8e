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
// Mapping (inter-task): one wavefront = up to 128 pairs of up to four
// queries; lane l holds two pairs, one in each 16-bit half of every VGPR, so
// each VOP3P instruction advances two DP cells.  Query rows are processed in
// chunks of 32 consecutive rows held in registers (H, E) while the lane walks
// the subject; a chunk's bottom row (H, F) is written per column to a
// per-wave workspace that the next chunk reads back in place.  The profile
// of the chunk (32 rows x 26 residues) of each of the task's queries sits in
// LDS, residue-major: per column a lane half reads the 32 scores of its
// subject residue in its own query's profile with four 16-byte loads, and one
// v_perm per row packs the two halves' scores.  Packing pairs of several
// queries into a wavefront therefore costs nothing in the inner loop.  Rows past
// a query's end score -8192: a value there stays below a cell of the same
// chunk already counted in the running maximum, so it is never a maximum.
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

constexpr int kNRes = 26;          // 25 matrix residues + 1 padding residue
constexpr int kQ = 4;              // queries per task
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
// R (16 or 32) is the number of query rows per chunk, held in registers.
// The profile is residue-major: W(s, q_k[j0 + r]) at [(k * 26 + s) * (R + 8)
// + r], so a lane reads the R scores of its subject residue with R / 8
// 16-byte loads.  The 16-byte pad makes the row stride 20 dwords (R = 32) or
// 12 (R = 16): sixteen different residues then start on sixteen different
// groups of four banks.
template <int R>
struct WaveLds {
    static constexpr int kStride = R + 8;
    alignas(16) int16_t prof[kQ * kNRes * kStride];
    uint16_t saved[2][R][64];              // H column of the chunk best, per half
};

// The R profile scores of one residue: rows 2i and 2i+1 in dword i.
template <int R>
struct ProfCol {
    uint32_t d[R / 2];
};

template <int R>
__device__ __forceinline__ void load_col(ProfCol<R>& c, const int16_t* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < R / 8; ++i) {
        const uint4 v = q[i];
        c.d[4 * i] = v.x; c.d[4 * i + 1] = v.y; c.d[4 * i + 2] = v.z; c.d[4 * i + 3] = v.w;
    }
}

// Rows 8q .. 8q + 7 of a profile column (one 16-byte load).
template <int R>
__device__ __forceinline__ void load_quarter(ProfCol<R>& c, const int16_t* p, int q) {
    const uint4 v = reinterpret_cast<const uint4*>(p)[q];
    c.d[4 * q] = v.x; c.d[4 * q + 1] = v.y; c.d[4 * q + 2] = v.z; c.d[4 * q + 3] = v.w;
}

// Row r of the low-half and high-half columns as one packed pair.
template <int R>
__device__ __forceinline__ s2 row_w(const ProfCol<R>& lo, const ProfCol<R>& hi, int r) {
    return from_bits(__builtin_amdgcn_perm(hi.d[r / 2], lo.d[r / 2],
                                           (r & 1) ? 0x07060302u : 0x05040100u));
}

// Best of the chunks finished so far, per half of the lane; chunk is the
// chunk index of that best (ties on score and column go to the earlier one).
struct LaneBest {
    int score[2];
    int ref[2];
    int read[2];
    int chunk[2];
};

// Wavefronts per workgroup (W, 4 or 8) is a template parameter of the kernel.
constexpr int kSyncCols = 8;       // progress is published every 8 columns
constexpr int kAhead = 16;         // a consumer runs at least 16 columns behind
constexpr int kSpinLimit = 1 << 24;
// s_waitcnt vmcnt(0) with expcnt and lgkmcnt left at their maxima (gfx9
// encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitVmem = 0x0f70;

// Where a chunk reads its top boundary and writes its bottom one, and how it
// synchronises with the wavefront that produces its top (another wavefront
// of the same workgroup working on the chunk above).
struct ChunkIO {
    const uint32_t* top;      // always a valid buffer
    bool has_top;             // false: first chunk (zero boundary)
    uint32_t* bottom;         // always a valid buffer
    bool has_bottom;          // false: last chunk (its bottom row is not needed)
    int* wait_ctr;            // producer progress in LDS, or null
    int wait_base;            // producer progress at column 0 of this chunk's top
    int* pub_ctr;             // own progress in LDS, or null
    int pub_base;
    int* error;               // LDS flag: a wait timed out
    int* g_error;             // global flag for the host
};

__device__ __forceinline__ bool wait_for(const ChunkIO& io, int need) {
    int spins = 0;
    for (;;) {
        const int v = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(io.wait_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (v >= need) break;
        const int e = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(io.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (e) return false;
        if (++spins > kSpinLimit) {
            __hip_atomic_store(io.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (threadIdx.x % 64 == 0) atomicExch(io.g_error, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

__device__ __forceinline__ void publish(const ChunkIO& io, int value) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (threadIdx.x % 64 == 0)
        __hip_atomic_store(io.pub_ctr, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int R>
__device__ __forceinline__ s2 tree_max(const s2 (&v)[R]) {
    s2 t[R / 2];
#pragma unroll
    for (int r = 0; r < R / 2; ++r) t[r] = vmax(v[2 * r], v[2 * r + 1]);
#pragma unroll
    for (int n = R / 4; n >= 1; n /= 2)
#pragma unroll
        for (int r = 0; r < n; ++r) t[r] = vmax(t[2 * r], t[2 * r + 1]);
    return t[0];
}

// Event bookkeeping for one finished column: if a half's column max rises
// strictly above the chunk best, remember the column and save that half of
// the column's H values.
template <int R>
__device__ __forceinline__ void column_event(WaveLds<R>& lds, const s2 (&Hc)[R], s2 cm,
                                             s2& bc, int col, int& ci_lo, int& ci_hi) {
    const int lane = threadIdx.x % 64;
    if (cm.x > bc.x) {
        ci_lo = col;
#pragma unroll
        for (int r = 0; r < R; ++r) lds.saved[0][r][lane] = (uint16_t)Hc[r].x;
    }
    if (cm.y > bc.y) {
        ci_hi = col;
#pragma unroll
        for (int r = 0; r < R; ++r) lds.saved[1][r][lane] = (uint16_t)Hc[r].y;
    }
    bc = vmax(bc, cm);
}

// Top boundary (H, F) of columns c and c + 1 of one lane.
struct TopPair {
    uint32_t h0, f0, h1, f1;
};

// The loads are unconditional, so that every iteration issues the same
// vector memory operations and the compiler's wait counts stay exact: past
// the last column they read column ncols - 2 again (the values are not
// used), and the first chunk (keep == 0) reads columns 0 and 1 of its own
// buffer over and over (cache hits) and the consumer masks the values to
// zero.
__device__ __forceinline__ TopPair load_top(const uint32_t* tH, const uint32_t* tF, int c,
                                            int ncols, uint32_t keep, int lane) {
    int cc = c < ncols - 2 ? c : (ncols >= 2 ? ncols - 2 : 0);
    cc = keep ? cc : 0;
    TopPair p;
    p.h0 = tH[(size_t)cc * 64 + lane];
    p.f0 = tF[(size_t)cc * 64 + lane];
    p.h1 = tH[(size_t)(cc + 1) * 64 + lane];
    p.f1 = tF[(size_t)(cc + 1) * 64 + lane];
    return p;
}

// Residues c .. c + 3 of a subject (c a multiple of 4), residue 25 past its
// end.  The load itself is unconditional: its address is clamped to the
// subject's last dword (subjects start on 4-byte boundaries, are padded with
// residue 25 to the next one, and the residue array has at least 4 bytes
// after the last subject), so no exec-masked load feeds a phi.
__device__ __forceinline__ uint32_t load_res4(const int8_t* s, int c, int n) {
    const int last = n > 0 ? ((n - 1) & ~3) : 0;
    const uint32_t w = *reinterpret_cast<const uint32_t*>(s + (c < last ? c : last));
    return c < n ? w : 0x19191919u;
}

// One chunk of R rows [j0, j0 + R) over all subject columns, two
// columns at a time: the second column runs one row behind the first, so
// every step advances two independent dependency chains.  Returns false if
// a wait on the producer of the top boundary gave up.
//
// The loop body covers four columns (two column pairs).  A column pair
// reads its profile columns in quarters of 8 rows, each loaded one quarter
// ahead of its rows: quarter 0 during the previous pair (as soon as that
// pair's rows 0-7 are done, into the same registers), quarter 1 when the
// pair starts, quarter k + 2 when the pair's rows of quarter k are done.  So
// no column pair starts by waiting for LDS, and at most three quarters of
// profile columns are live at a time.  The next pair's residues come from
// the residue dword of the current four columns, or of the next four,
// loaded at the top of the iteration.
//
// Each iteration reloads the prefetched top boundary and the residue dword
// right after their last use, into the same variables: no loaded value is
// copied at the back edge before its load was waited for, so no iteration
// waits for the loads or stores it has just issued.
template <int R>
__device__ __forceinline__ bool run_chunk(const SwKernelArgs& a, WaveLds<R>& lds, const ChunkIO& io,
                                          int ncols, const int8_t* s_lo, int n_lo,
                                          const int8_t* s_hi, int n_hi, int off_lo, int off_hi,
                                          int j0, int chunk, LaneBest& best) {
    constexpr int kS = WaveLds<R>::kStride;
    constexpr int kQuarters = R / 8;
    const int lane = threadIdx.x % 64;
    const s2 go = s2{(short)a.gap_open, (short)a.gap_open};
    const s2 ge = s2{(short)a.gap_extend, (short)a.gap_extend};
    const s2 zero = s2{0, 0};
    s2 E[R], H[R], Ha[R];
#pragma unroll
    for (int r = 0; r < R; ++r) { E[r] = zero; H[r] = zero; }
    s2 bc = zero;                 // chunk best per half
    int ci_lo = -1, ci_hi = -1;
    s2 hdiag = zero;              // H(i-1, j0-1)

    // all ones when the chunk has a top boundary, zero for the first chunk
    const uint32_t keep = (uint32_t)__builtin_amdgcn_readfirstlane(io.has_top ? -1 : 0);
    const uint32_t* tH = io.top;                         // [col][lane]
    const uint32_t* tF = io.top + (size_t)ncols * 64;
    uint32_t* bH = io.bottom;
    uint32_t* bF = io.bottom + (size_t)ncols * 64;
    // the last chunk writes every column pair into columns 0 and 1 (already
    // read by then, and read by no one after): the stores stay unconditional
    // but land in one cached line instead of going to memory
    const int bstep = __builtin_amdgcn_readfirstlane(io.has_bottom ? 1 : 0);

    int waited = ncols < kAhead ? ncols : kAhead;       // columns known final
    if (io.wait_ctr && !wait_for(io, io.wait_base + waited)) return false;
    uint32_t res_lo = load_res4(s_lo, 0, n_lo), res_hi = load_res4(s_hi, 0, n_hi);
    TopPair T = load_top(tH, tF, 0, ncols, keep, lane);
    // Let these first loads land here: a load still pending at the loop
    // entry makes the compiler wait for all vector memory at the top of every
    // iteration.
    __builtin_amdgcn_s_waitcnt(kWaitVmem);

    // profile columns of the current column pair (a = even, b = odd column),
    // starting with quarter 0 of columns 0 and 1
    ProfCol<R> pa_lo, pa_hi, pb_lo, pb_hi;
    load_quarter<R>(pa_lo, lds.prof + (off_lo + (res_lo & 0xff)) * kS, 0);
    load_quarter<R>(pa_hi, lds.prof + (off_hi + (res_hi & 0xff)) * kS, 0);
    load_quarter<R>(pb_lo, lds.prof + (off_lo + ((res_lo >> 8) & 0xff)) * kS, 0);
    load_quarter<R>(pb_hi, lds.prof + (off_hi + ((res_hi >> 8) & 0xff)) * kS, 0);
    uint32_t cur_lo = res_lo, cur_hi = res_hi;   // residues of the current pair

    // Columns i and i + 1 from the prefetched boundary T, which is then
    // reloaded with columns i + 2 and i + 3; nx_lo / nx_hi hold the residues
    // of columns i + 2 (bits 0-7) and i + 3 (bits 8-15).
    auto column_pair = [&](int i, uint32_t nx_lo, uint32_t nx_hi,
                           TopPair& T) __attribute__((always_inline)) {
        const int16_t* ca_lo = lds.prof + (off_lo + (cur_lo & 0xff)) * kS;
        const int16_t* ca_hi = lds.prof + (off_hi + (cur_hi & 0xff)) * kS;
        const int16_t* cb_lo = lds.prof + (off_lo + ((cur_lo >> 8) & 0xff)) * kS;
        const int16_t* cb_hi = lds.prof + (off_hi + ((cur_hi >> 8) & 0xff)) * kS;
        const int16_t* na_lo = lds.prof + (off_lo + (nx_lo & 0xff)) * kS;
        const int16_t* na_hi = lds.prof + (off_hi + (nx_hi & 0xff)) * kS;
        const int16_t* nb_lo = lds.prof + (off_lo + ((nx_lo >> 8) & 0xff)) * kS;
        const int16_t* nb_hi = lds.prof + (off_hi + ((nx_hi >> 8) & 0xff)) * kS;
        cur_lo = nx_lo;
        cur_hi = nx_hi;
        if (kQuarters > 1) {
            load_quarter<R>(pa_lo, ca_lo, 1);
            load_quarter<R>(pa_hi, ca_hi, 1);
            load_quarter<R>(pb_lo, cb_lo, 1);
            load_quarter<R>(pb_hi, cb_hi, 1);
        }

        s2 fa = from_bits(T.f0 & keep), fb = from_bits(T.f1 & keep);
        s2 da = hdiag;                  // H(i-1, j0-1)
        s2 db = from_bits(T.h0 & keep); // H(i, j0-1)
        hdiag = from_bits(T.h1 & keep); // H(i+1, j0-1)
        T = load_top(tH, tF, i + 2, ncols, keep, lane);

        // column a, row 0
        {
            s2 w = row_w<R>(pa_lo, pa_hi, 0);
            s2 h = vmax(vmax(add_sat(da, w), E[0]), fa);
            da = H[0];
            Ha[0] = h;
            s2 hg = usub(h, go);
            E[0] = vmax(usub(E[0], ge), hg);
            fa = vmax(usub(fa, ge), hg);
        }
#pragma unroll
        for (int r = 1; r < R; ++r) {
            // column a, row r
            {
                s2 w = row_w<R>(pa_lo, pa_hi, r);
                s2 h = vmax(vmax(add_sat(da, w), E[r]), fa);
                da = H[r];
                Ha[r] = h;
                s2 hg = usub(h, go);
                E[r] = vmax(usub(E[r], ge), hg);
                fa = vmax(usub(fa, ge), hg);
            }
            // column b, row r - 1
            {
                const int q = r - 1;
                s2 w = row_w<R>(pb_lo, pb_hi, q);
                s2 h = vmax(vmax(add_sat(db, w), E[q]), fb);
                db = Ha[q];
                H[q] = h;
                s2 hg = usub(h, go);
                E[q] = vmax(usub(E[q], ge), hg);
                fb = vmax(usub(fb, ge), hg);
            }
            if (r % 8 == 0) {
                // rows of quarter k = r / 8 - 1 done in both columns
                const int k = r / 8 - 1;
                if (k == 0) {
                    // quarter 0 of the next pair
                    load_quarter<R>(pa_lo, na_lo, 0);
                    load_quarter<R>(pa_hi, na_hi, 0);
                    load_quarter<R>(pb_lo, nb_lo, 0);
                    load_quarter<R>(pb_hi, nb_hi, 0);
                }
                if (k + 2 < kQuarters) {
                    load_quarter<R>(pa_lo, ca_lo, k + 2);
                    load_quarter<R>(pa_hi, ca_hi, k + 2);
                    load_quarter<R>(pb_lo, cb_lo, k + 2);
                    load_quarter<R>(pb_hi, cb_hi, k + 2);
                }
            }
        }
        // column b, last row
        {
            const int q = R - 1;
            s2 w = row_w<R>(pb_lo, pb_hi, q);
            s2 h = vmax(vmax(add_sat(db, w), E[q]), fb);
            H[q] = h;
            s2 hg = usub(h, go);
            E[q] = vmax(usub(E[q], ge), hg);
            fb = vmax(usub(fb, ge), hg);
        }
        const size_t bc0 = (size_t)(i * bstep) * 64 + lane;
        bH[bc0] = bits(Ha[R - 1]);
        bF[bc0] = bits(fa);
        bH[bc0 + 64] = bits(H[R - 1]);
        bF[bc0 + 64] = bits(fb);
        // column maxima as balanced trees: independent maxes issue back to
        // back, where a running max would be one dependent chain
        const s2 cma = tree_max<R>(Ha), cmb = tree_max<R>(H);
        column_event(lds, Ha, cma, bc, i, ci_lo, ci_hi);
        column_event(lds, H, cmb, bc, i + 1, ci_lo, ci_hi);
        if (io.pub_ctr && (((i + 2) % kSyncCols) == 0 || i + 2 == ncols))
            publish(io, io.pub_base + i + 2);
    };

    // ncols is a multiple of 4 (the host pads subjects with residue 25
    // columns); res_lo / res_hi hold the residues of columns i .. i + 3
    for (int i = 0; i < ncols; i += 4) {
        if (io.wait_ctr && (i % kSyncCols) == 0 && waited < ncols) {
            // columns up to i + kAhead must be final before they are prefetched
            const int need = i + kAhead < ncols ? i + kAhead : ncols;
            if (need > waited) {
                if (!wait_for(io, io.wait_base + need)) return false;
                waited = need;
            }
        }
        const uint32_t rn_lo = load_res4(s_lo, i + 4, n_lo);
        const uint32_t rn_hi = load_res4(s_hi, i + 4, n_hi);
        column_pair(i, res_lo >> 16, res_hi >> 16, T);
        column_pair(i + 2, rn_lo, rn_hi, T);
        res_lo = rn_lo;
        res_hi = rn_hi;
    }

    // Lowest row of the saved column equal to the chunk best, then merge:
    // larger score wins; on an equal score the smaller column; on an equal
    // column the earlier chunk (already held).
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int b = half ? bc.y : bc.x;
        const int ci = half ? ci_hi : ci_lo;
        if (b <= 0) continue;
        int row = R;
#pragma unroll
        for (int r = R - 1; r >= 0; --r)
            if ((int)(int16_t)lds.saved[half][r][lane] == b) row = r;
        if (b > best.score[half] || (b == best.score[half] && ci < best.ref[half])) {
            best.score[half] = b;
            best.ref[half] = ci;
            best.read[half] = j0 + row;
            best.chunk[half] = chunk;
        }
    }
    return true;
}

// The pairs of one task as seen by one lane.
struct TaskView {
    int pair_lo, pair_hi;
    bool has_lo, has_hi;
    const int8_t* qs[kQ];    // the task's queries (wave-uniform); null past the last
    int ms[kQ];
    int m_max;
    int off_lo, off_hi;      // profile offset of each half's query
    int m_lo, m_hi;
    const int8_t* s_lo;
    const int8_t* s_hi;
    int n_lo, n_hi;
    int ncols;               // even, wave-uniform
};

// Substitution scores of rows j0 .. j0+R-1 of every query slot of the task
// for all 26 residues; rows past a query's end (and empty slots) score
// kPadScore.  Each lane fills whole rows: it loads its (slot, row)'s query
// residue once and copies that residue's column of the LDS matrix.  Written
// and read by one wavefront only.
template <int R>
__device__ __forceinline__ void build_profile(const int8_t* mat, WaveLds<R>& lds,
                                              const TaskView& t, int j0) {
    constexpr int kS = WaveLds<R>::kStride;
    const int lane = threadIdx.x % 64;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int base = 0; base < kQ * R; base += 64) {
        const int idx = base + lane;
        const int k = idx / R, r = idx % R;
        const int8_t* q = k == 0 ? t.qs[0] : k == 1 ? t.qs[1] : k == 2 ? t.qs[2] : t.qs[3];
        const int m = k == 0 ? t.ms[0] : k == 1 ? t.ms[1] : k == 2 ? t.ms[2] : t.ms[3];
        const int j = j0 + r;
        const bool real = q != nullptr && j < m;
        const int qv = real ? q[j] : 0;
        int16_t* dst = lds.prof + k * kNRes * kS + r;
#pragma unroll
        for (int s = 0; s < kPadResidue; ++s)
            dst[s * kS] = (int16_t)(real ? (int)mat[s * 25 + qv] : kPadScore);
        dst[kPadResidue * kS] = (int16_t)kPadScore;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ TaskView task_view(const SwKernelArgs& a, int task) {
    const int lane = threadIdx.x % 64;
    task = __builtin_amdgcn_readfirstlane(task);
    TaskView t;
    const int p0 = a.task_begin[task], p1 = a.task_begin[task + 1];
    const int np = p1 - p0;
    t.has_lo = lane < np;
    t.has_hi = lane + 64 < np;
    const int k_lo = p0 + (t.has_lo ? lane : 0);
    const int k_hi = p0 + (t.has_hi ? lane + 64 : 0);
    t.pair_lo = a.pair_index ? a.pair_index[k_lo] : k_lo;
    t.pair_hi = a.pair_index ? a.pair_index[k_hi] : k_hi;
    int tq[kQ];
    t.m_max = 0;
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        tq[k] = a.task_query[kQ * task + k];
        t.qs[k] = tq[k] >= 0 ? a.q_res + a.q_off[tq[k]] : nullptr;
        t.ms[k] = tq[k] >= 0 ? a.q_len[tq[k]] : 0;
        t.m_max = t.ms[k] > t.m_max ? t.ms[k] : t.m_max;
    }
    const int qid_lo = a.pair_query[t.pair_lo], qid_hi = a.pair_query[t.pair_hi];
    int slot_lo = 0, slot_hi = 0;
#pragma unroll
    for (int k = 1; k < kQ; ++k) {
        if (qid_lo == tq[k]) slot_lo = k;
        if (qid_hi == tq[k]) slot_hi = k;
    }
    t.off_lo = slot_lo * kNRes;
    t.off_hi = slot_hi * kNRes;
    t.m_lo = t.has_lo ? a.q_len[qid_lo] : 0;
    t.m_hi = t.has_hi ? a.q_len[qid_hi] : 0;
    const int sid_lo = a.pair_subject[t.pair_lo], sid_hi = a.pair_subject[t.pair_hi];
    t.s_lo = a.s_res + a.s_off[sid_lo];
    t.s_hi = a.s_res + a.s_off[sid_hi];
    t.n_lo = t.has_lo ? a.s_len[sid_lo] : 0;
    t.n_hi = t.has_hi ? a.s_len[sid_hi] : 0;
    int ncols = t.n_lo > t.n_hi ? t.n_lo : t.n_hi;
    for (int off = 32; off >= 1; off >>= 1) {
        const int o = __shfl_xor(ncols, off);
        ncols = o > ncols ? o : ncols;
    }
    ncols = __builtin_amdgcn_readfirstlane(ncols);
    t.ncols = (ncols + 3) & ~3;        // columns go in fours; pad columns score kPadScore
    return t;
}

__device__ __forceinline__ void init_best(LaneBest& best, const TaskView& t) {
    // score 0: ref_end1 -1 and read_end1 0, or -1 for an empty query
    // (ssw.c:148-149, 288-297)
    for (int h = 0; h < 2; ++h) {
        best.score[h] = 0; best.ref[h] = -1; best.chunk[h] = 0;
        best.read[h] = (h ? t.m_hi : t.m_lo) > 0 ? 0 : -1;
    }
}

__device__ __forceinline__ void write_best(const SwKernelArgs& a, const TaskView& t,
                                           const LaneBest& best) {
    if (t.has_lo) {
        a.out_score[t.pair_lo] = best.score[0] >= 32767 ? 32767 : best.score[0];
        a.out_ref[t.pair_lo] = best.ref[0];
        a.out_read[t.pair_lo] = best.read[0];
    }
    if (t.has_hi) {
        a.out_score[t.pair_hi] = best.score[1] >= 32767 ? 32767 : best.score[1];
        a.out_ref[t.pair_hi] = best.ref[1];
        a.out_read[t.pair_hi] = best.read[1];
    }
}

// LDS of one workgroup of W wavefronts.
template <int R, int W>
struct GroupLds {
    WaveLds<R> wave[W];
    int8_t mat[640];                     // the 25 x 25 substitution matrix
    int item;
    int error;
    int progress[W];
    int bscore[W][2][64], bref[W][2][64], bread[W][2][64], bchunk[W][2][64];
};

// A task worked by one wavefront: its chunks in order, boundary in place.
template <int R, int W>
__device__ void run_task_wave(const SwKernelArgs& a, WaveLds<R>& lds, uint32_t* buf, int task,
                              GroupLds<R, W>& g) {
    const TaskView t = task_view(a, task);
    LaneBest best;
    init_best(best, t);
    int chunk = 0;
    for (int j0 = 0; j0 < t.m_max; j0 += R, ++chunk) {
        build_profile(g.mat, lds, t, j0);
        ChunkIO io{};
        io.top = buf;              // in place: column c is read before it is written
        io.has_top = j0 > 0;
        io.bottom = buf;
        io.has_bottom = j0 + R < t.m_max;
        io.error = &g.error;
        io.g_error = a.error;
        run_chunk(a, lds, io, t.ncols, t.s_lo, t.n_lo, t.s_hi, t.n_hi, t.off_lo, t.off_hi, j0,
                  chunk, best);
    }
    write_best(a, t, best);
}

// A task worked by all W wavefronts of the workgroup: chunk c by wavefront
// c % W, each one trailing the wavefront of the chunk above by at least
// kAhead columns; chunk c's bottom boundary lives in buffer c % W.
template <int R, int W>
__device__ void run_task_group(const SwKernelArgs& a, GroupLds<R, W>& g, uint32_t* bufs,
                               size_t buf_words, int task) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const TaskView t = task_view(a, task);
    LaneBest best;
    init_best(best, t);
    const int n_chunks = (t.m_max + R - 1) / R;
    const int nc = t.ncols;
    for (int c = wave; c < n_chunks; c += W) {
        const int j0 = c * R;
        build_profile(g.mat, g.wave[wave], t, j0);
        ChunkIO io{};
        // chunk c writes buffer c % W, read by chunk c + 1; the wave of
        // chunk c - W + 1 (which read it last) has passed every column first
        io.bottom = bufs + (size_t)(c % W) * buf_words;
        io.has_bottom = c + 1 < n_chunks;
        io.has_top = c > 0;
        io.top = c > 0 ? bufs + (size_t)((c - 1) % W) * buf_words : io.bottom;
        if (c > 0) {
            io.wait_ctr = &g.progress[(c - 1) % W];
            io.wait_base = ((c - 1) / W) * nc;
        }
        io.pub_ctr = &g.progress[wave];
        io.pub_base = (c / W) * nc;
        io.error = &g.error;
        io.g_error = a.error;
        if (!run_chunk(a, g.wave[wave], io, nc, t.s_lo, t.n_lo, t.s_hi, t.n_hi, t.off_lo,
                       t.off_hi, j0, c, best))
            break;
        if (nc == 0) publish(io, io.pub_base);
    }
    // merge the wavefronts' bests: larger score, then smaller column, then
    // the earlier chunk
    for (int h = 0; h < 2; ++h) {
        g.bscore[wave][h][lane] = best.score[h];
        g.bref[wave][h][lane] = best.ref[h];
        g.bread[wave][h][lane] = best.read[h];
        g.bchunk[wave][h][lane] = best.chunk[h];
    }
    __syncthreads();
    if (wave == 0) {
        for (int h = 0; h < 2; ++h) {
            for (int w = 1; w < W; ++w) {
                const int sc = g.bscore[w][h][lane], rf = g.bref[w][h][lane];
                const int ck = g.bchunk[w][h][lane];
                if (sc <= 0) continue;
                if (sc > best.score[h] ||
                    (sc == best.score[h] &&
                     (rf < best.ref[h] || (rf == best.ref[h] && ck < best.chunk[h])))) {
                    best.score[h] = sc;
                    best.ref[h] = rf;
                    best.read[h] = g.bread[w][h][lane];
                    best.chunk[h] = ck;
                }
            }
        }
        if (!__hip_atomic_load(&g.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            write_best(a, t, best);
    }
}

// Work items (W entries each): item k is either one large task worked by the
// whole workgroup (item_task[W k] >= 0, item_task[W k + 1] == -2) or up to W
// small tasks, one per wavefront (-1 = none).  With a counter the grid is
// persistent and workgroup b owns workspace slot b; without one, workgroup b
// runs item b.
template <int R, int W>
__global__ void __launch_bounds__(W * 64)
sw_forward_kernel(SwKernelArgs a) {
    __shared__ GroupLds<R, W> g;
    const int wave = threadIdx.x / 64;
    const size_t buf_words = (size_t)a.ws_words_per_slot / W;
    uint32_t* bufs = a.workspace + (size_t)blockIdx.x * a.ws_words_per_slot;
    if (threadIdx.x == 0) g.error = 0;
    for (int e = threadIdx.x; e < 625; e += blockDim.x) g.mat[e] = a.matrix[e];
    for (;;) {
        if (threadIdx.x == 0) {
            g.item = a.task_counter ? atomicAdd(a.task_counter, 1) : (int)blockIdx.x;
            for (int w = 0; w < W; ++w) g.progress[w] = 0;
        }
        __syncthreads();
        const int item = g.item;
        if (item >= a.n_items || g.error) break;
        const int t0 = a.item_task[W * item];
        if (a.item_task[W * item + 1] == -2) {
            run_task_group(a, g, bufs, buf_words, t0);
        } else {
            const int task = a.item_task[W * item + wave];
            if (task >= 0) run_task_wave(a, g.wave[wave], bufs + (size_t)wave * buf_words, task, g);
        }
        __syncthreads();
        if (!a.task_counter) break;
    }
}

// Sets n words to v (graph-capturable, unlike a memset node).
__global__ void fill_i32_kernel(int32_t* p, int32_t v, int32_t n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_fill_i32(int32_t* p, int32_t v, int32_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
    return hipGetLastError();
}

namespace {
template <int R, int W>
int groups_per_cu() {
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, tpz::sw_forward_kernel<R, W>, W * 64, 0);
    return e == hipSuccess ? n : 0;
}
template <int R, int W>
hipError_t launch(const tpz::SwKernelArgs* args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((tpz::sw_forward_kernel<R, W>), dim3(grid), dim3(W * 64), 0, stream, *args);
    return hipGetLastError();
}
}  // namespace

extern "C" int tpz_sw_forward_groups_per_cu(int rows, int waves) {
    if (rows == 16 && waves == 4) return groups_per_cu<16, 4>();
    if (rows == 16 && waves == 8) return groups_per_cu<16, 8>();
    if (rows == 32 && waves == 4) return groups_per_cu<32, 4>();
    if (rows == 32 && waves == 8) return groups_per_cu<32, 8>();
    return 0;
}

extern "C" hipError_t tpz_launch_sw_forward(const tpz::SwKernelArgs* args, int grid, int rows,
                                            int waves, hipStream_t stream) {
    if (rows == 16 && waves == 4) return launch<16, 4>(args, grid, stream);
    if (rows == 16 && waves == 8) return launch<16, 8>(args, grid, stream);
    if (rows == 32 && waves == 4) return launch<32, 4>(args, grid, stream);
    if (rows == 32 && waves == 8) return launch<32, 8>(args, grid, stream);
    return hipErrorInvalidValue;
}
