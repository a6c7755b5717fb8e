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
// Mapping (inter-task): one wavefront = a task of pairs of up to four
// queries; lane l holds two runs of pairs, one in each 16-bit half of every
// VGPR, so each VOP3P instruction advances two DP cells.  A run is one pair
// or several back to back along the columns: a long pair sets the task's
// width and the other halves fill it with more pairs, instead of walking
// padding columns past a short subject's end.  Query rows are processed in
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
// best so far of the half's pair is an "event"; on an event the lane's 32 H
// values of that half are saved to LDS.  Where the pair ends (or the chunk),
// the saved column gives the lowest row, and (score, column, row) goes into
// the pair's 64-bit key with an atomic max; keys order by larger score, then
// smaller column, then smaller row, which is the reference's first-column /
// lowest-row rule, whichever chunk or wavefront finishes first.
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

// Lane runs.  A lane half works a run of pairs, one after another along the
// columns: each pair takes item_cols(n) columns (its subject, then padding
// residues), and the next pair of the run starts where they end.  Entry
// records (SwKernelArgs::entry_rec) chain a run: subject offset / 16,
// subject length, the next entry of the run (-1 at its end) and the query
// slot.  The boundaries are the same in every chunk of rows.
__device__ __forceinline__ int item_cols(int n) {
    const int c = (n + 3) & ~3;
    return c < 8 ? 8 : c;
}
constexpr int kNoBoundary = 0x7fffffff;

// The run of one lane half, as a chunk walks its columns.  The residue
// stream switches to the next pair one step of four columns before the pair
// starts (its residues are loaded a step ahead); the rest switches when it
// starts.  nrec is the record of entry en, loaded when the pair before it
// starts, so it has landed long before the boundary needs it (a pair takes
// at least 8 columns).
struct LaneRun {
    int e;          // entry of the pair being computed (-1: none)
    int c0;         // its first column
    int nb;         // column where the next pair starts (kNoBoundary: none)
    int off;        // profile offset of its query slot
    int en;         // entry of the next pair (-1: none)
    int64_t lb;     // residue stream: column c reads s_res[lb + c] while c < le
    int le;
    int4 nrec;
};

// Best (score, first column, lowest row) of a pair as one 64-bit key whose
// maximum is the reference's choice: the larger score, then the smaller
// column, then the smaller row (ssw.c:264-297).  Columns and rows below 2^24.
__device__ __forceinline__ unsigned long long best_key(int score, int col, int row) {
    return ((unsigned long long)score << 48) | ((unsigned long long)(0xffffff - col) << 24) |
           (unsigned long long)(0xffffff - row);
}

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

// Residues of columns c .. c + 3 (c a multiple of 4) of a lane's run: the
// stream's subject while c < le, padding after it (s_res + s_pad holds 16
// residues 25).  Subjects start on 16-byte boundaries and are padded with
// residue 25 to the next one.  The load is unconditional and stays in the
// global address space, and no select follows it, so the compiler keeps it
// where it is issued (a load under an exec mask feeding a phi, or a flat
// load, makes every iteration wait for all memory).
__device__ __forceinline__ uint32_t load_res4(const SwKernelArgs& a, int64_t lb, int c, int le) {
    return *reinterpret_cast<const uint32_t*>(a.s_res + (c < le ? lb + c : a.s_pad));
}

// The start of a run at column 0 of a chunk: its first pair's record.
__device__ __forceinline__ void run_start(const SwKernelArgs& a, LaneRun& L, int e0) {
    L.e = e0;
    L.c0 = 0;
    const int4 r = a.entry_rec[e0 >= 0 ? e0 : 0];
    if (e0 >= 0) {
        L.off = (r.w & 3) * kNRes;
        L.en = r.z;
        L.nb = r.z >= 0 ? item_cols(r.y) : kNoBoundary;
        L.lb = (int64_t)(uint32_t)r.x * 16;
        L.le = r.y;
    } else {
        L.off = 0; L.en = -1; L.nb = kNoBoundary; L.lb = 0; L.le = 0;
    }
    L.nrec = a.entry_rec[L.en >= 0 ? L.en : (e0 >= 0 ? e0 : 0)];
}

// One chunk of R rows [j0, j0 + R) over all columns of a task, two columns
// at a time: the second column runs one row behind the first, so every step
// advances two independent dependency chains.  Returns false if a wait on
// the producer of the top boundary gave up.
//
// The loop body covers four columns (two column pairs).  A column pair
// reads its profile columns in quarters of 8 rows, each loaded one quarter
// ahead of its rows: quarter 0 during the previous column pair (as soon as
// that pair's rows 0-7 are done, into the same registers; with the next
// run pair's query slot where a run's pair ends between them), quarter 1
// when the column pair starts, quarter k + 2 when its rows of quarter k are
// done.  Where a pair of a lane half's run ends (a multiple of four
// columns), the chunk's best of that pair goes to its key and the half's H,
// E and diagonal restart at zero.
//
// Each iteration reloads the prefetched top boundary, the residue dword and
// the next pairs' records right after their last use, into the same
// variables: no loaded value is copied at the back edge before its load was
// waited for, so no iteration waits for the loads or stores it has just
// issued.
template <int R>
__device__ __forceinline__ bool run_chunk(const SwKernelArgs& a, WaveLds<R>& lds, const ChunkIO& io,
                                          int ncols, int e0_lo, int e0_hi, int j0) {
    constexpr int kS = WaveLds<R>::kStride;
    constexpr int kQuarters = R / 8;
    const int lane = threadIdx.x % 64;
    const s2 go = s2{(short)a.gap_open, (short)a.gap_open};
    const s2 ge = s2{(short)a.gap_extend, (short)a.gap_extend};
    const s2 zero = s2{0, 0};
    s2 E[R], H[R], Ha[R];
#pragma unroll
    for (int r = 0; r < R; ++r) { E[r] = zero; H[r] = zero; }
    s2 bc = zero;                 // best of the current pair in this chunk, per half
    int ci_lo = -1, ci_hi = -1;   // the column where it was first reached
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

    LaneRun Llo, Lhi;
    run_start(a, Llo, e0_lo);
    run_start(a, Lhi, e0_hi);
    int waited = ncols < kAhead ? ncols : kAhead;       // columns known final
    if (io.wait_ctr && !wait_for(io, io.wait_base + waited)) return false;
    uint32_t res_lo = load_res4(a, Llo.lb, 0, Llo.le);
    uint32_t res_hi = load_res4(a, Lhi.lb, 0, Lhi.le);
    TopPair T = load_top(tH, tF, 0, ncols, keep, lane);
    // Let these first loads land here: a load still pending at the loop
    // entry makes the compiler wait for all vector memory at the top of every
    // iteration.
    __builtin_amdgcn_s_waitcnt(kWaitVmem);

    // The chunk's best of a half's current pair into the pair's key: the
    // lowest row of the saved column holding it.
    auto finalize = [&](int half, const LaneRun& L, int b, int ci) __attribute__((always_inline)) {
        if (L.e < 0 || b <= 0) return;
        int row = R;
#pragma unroll
        for (int r = R - 1; r >= 0; --r)
            if ((int)(int16_t)lds.saved[half][r][lane] == b) row = r;
        atomicMax(a.out_key + L.e, best_key(b, ci - L.c0, j0 + row));
    };
    // The pair after the current one starts at column i; the record of the
    // one after it is loaded now, read a step or more later.
    auto advance = [&](LaneRun& L, int i) __attribute__((always_inline)) {
        const int4 r = L.nrec;
        L.e = L.en;
        L.c0 = i;
        L.off = (r.w & 3) * kNRes;
        L.en = r.z;
        L.nb = r.z >= 0 ? i + item_cols(r.y) : kNoBoundary;
        if (r.z >= 0) L.nrec = a.entry_rec[r.z];
    };

    // profile columns of the current column pair (a = even, b = odd
    // column), starting with quarter 0 of columns 0 and 1
    ProfCol<R> pa_lo, pa_hi, pb_lo, pb_hi;
    load_quarter<R>(pa_lo, lds.prof + (Llo.off + (res_lo & 0xff)) * kS, 0);
    load_quarter<R>(pa_hi, lds.prof + (Lhi.off + (res_hi & 0xff)) * kS, 0);
    load_quarter<R>(pb_lo, lds.prof + (Llo.off + ((res_lo >> 8) & 0xff)) * kS, 0);
    load_quarter<R>(pb_hi, lds.prof + (Lhi.off + ((res_hi >> 8) & 0xff)) * kS, 0);

    // Columns i and i + 1 (residues in the low 16 bits of cur_lo / cur_hi)
    // from the prefetched boundary T, which is then reloaded with columns
    // i + 2 and i + 3; nx_lo / nx_hi hold the residues of columns i + 2 and
    // i + 3, and noff_lo / noff_hi their profile offsets.
    auto column_pair = [&](int i, uint32_t cur_lo, uint32_t cur_hi, uint32_t nx_lo, uint32_t nx_hi,
                           int noff_lo, int noff_hi, TopPair& T) __attribute__((always_inline)) {
        const int16_t* ca_lo = lds.prof + (Llo.off + (cur_lo & 0xff)) * kS;
        const int16_t* ca_hi = lds.prof + (Lhi.off + (cur_hi & 0xff)) * kS;
        const int16_t* cb_lo = lds.prof + (Llo.off + ((cur_lo >> 8) & 0xff)) * kS;
        const int16_t* cb_hi = lds.prof + (Lhi.off + ((cur_hi >> 8) & 0xff)) * kS;
        const int16_t* na_lo = lds.prof + (noff_lo + (nx_lo & 0xff)) * kS;
        const int16_t* na_hi = lds.prof + (noff_hi + (nx_hi & 0xff)) * kS;
        const int16_t* nb_lo = lds.prof + (noff_lo + ((nx_lo >> 8) & 0xff)) * kS;
        const int16_t* nb_hi = lds.prof + (noff_hi + ((nx_hi >> 8) & 0xff)) * kS;
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
                    // quarter 0 of the next column pair
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

    // ncols is a multiple of 4, as are the pairs' boundaries; res_lo / res_hi
    // hold the residues of columns i .. i + 3
    for (int i = 0; i < ncols; i += 4) {
        if (io.wait_ctr && (i % kSyncCols) == 0 && waited < ncols) {
            // columns up to i + kAhead must be final before they are prefetched
            const int need = i + kAhead < ncols ? i + kAhead : ncols;
            if (need > waited) {
                if (!wait_for(io, io.wait_base + need)) return false;
                waited = need;
            }
        }
        // a pair starts at column i in some lane half: the previous one's
        // best to its key, and the half restarts at zero
        {
            const bool st_lo = i == Llo.nb, st_hi = i == Lhi.nb;
            if (__builtin_amdgcn_ballot_w64(st_lo || st_hi)) {
                if (st_lo) finalize(0, Llo, bc.x, ci_lo);
                if (st_hi) finalize(1, Lhi, bc.y, ci_hi);
                const uint32_t m = (st_lo ? 0xffff0000u : 0xffffffffu) & (st_hi ? 0x0000ffffu : 0xffffffffu);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    E[r] = from_bits(bits(E[r]) & m);
                    H[r] = from_bits(bits(H[r]) & m);
                }
                hdiag = from_bits(bits(hdiag) & m);
                bc = from_bits(bits(bc) & m);
                if (st_lo) { advance(Llo, i); ci_lo = -1; }
                if (st_hi) { advance(Lhi, i); ci_hi = -1; }
            }
        }
        // the residue stream moves to the next pair a step before it starts
        if (i + 4 == Llo.nb) {
            Llo.lb = (int64_t)(uint32_t)Llo.nrec.x * 16 - (i + 4);
            Llo.le = i + 4 + Llo.nrec.y;
        }
        if (i + 4 == Lhi.nb) {
            Lhi.lb = (int64_t)(uint32_t)Lhi.nrec.x * 16 - (i + 4);
            Lhi.le = i + 4 + Lhi.nrec.y;
        }
        const uint32_t rn_lo = load_res4(a, Llo.lb, i + 4, Llo.le);
        const uint32_t rn_hi = load_res4(a, Lhi.lb, i + 4, Lhi.le);
        // the column pair after this step's second one starts a run's next
        // pair where that pair begins at column i + 4
        const int noff_lo = i + 4 == Llo.nb ? (Llo.nrec.w & 3) * kNRes : Llo.off;
        const int noff_hi = i + 4 == Lhi.nb ? (Lhi.nrec.w & 3) * kNRes : Lhi.off;
        column_pair(i, res_lo, res_hi, res_lo >> 16, res_hi >> 16, Llo.off, Lhi.off, T);
        column_pair(i + 2, res_lo >> 16, res_hi >> 16, rn_lo, rn_hi, noff_lo, noff_hi, T);
        res_lo = rn_lo;
        res_hi = rn_hi;
    }
    // the pairs still open at the chunk's last column
    finalize(0, Llo, bc.x, ci_lo);
    finalize(1, Lhi, bc.y, ci_hi);
    return true;
}

// A task as seen by one lane: the first entries of its two runs and the
// task's queries.
struct TaskView {
    int e_lo, e_hi;          // entries starting the lane's runs (-1: none)
    const int8_t* qs[kQ];    // the task's queries (wave-uniform); null past the last
    int ms[kQ];
    int m_max;
    int ncols;               // columns of the task, a multiple of 4, wave-uniform
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
    // the first 128 entries of a task start its runs (entry p0 + k: lane
    // k % 64, half k / 64)
    const int np = p1 - p0 < 128 ? p1 - p0 : 128;
    t.e_lo = lane < np ? p0 + lane : -1;
    t.e_hi = lane + 64 < np ? p0 + 64 + lane : -1;
    t.m_max = 0;
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        const int q = a.task_query[kQ * task + k];
        t.qs[k] = q >= 0 ? a.q_res + a.q_off[q] : nullptr;
        t.ms[k] = q >= 0 ? a.q_len[q] : 0;
        t.m_max = t.ms[k] > t.m_max ? t.ms[k] : t.m_max;
    }
    t.ncols = __builtin_amdgcn_readfirstlane(a.task_cols[task]);
    return t;
}

// LDS of one workgroup of W wavefronts.
template <int R, int W>
struct GroupLds {
    WaveLds<R> wave[W];
    int8_t mat[640];                     // the 25 x 25 substitution matrix
    int item;
    int error;
    int progress[W];
};

// A task worked by one wavefront: its chunks in order, boundary in place.
template <int R, int W>
__device__ void run_task_wave(const SwKernelArgs& a, WaveLds<R>& lds, uint32_t* buf, int task,
                              GroupLds<R, W>& g) {
    const TaskView t = task_view(a, task);
    for (int j0 = 0; j0 < t.m_max; j0 += R) {
        build_profile(g.mat, lds, t, j0);
        ChunkIO io{};
        io.top = buf;              // in place: column c is read before it is written
        io.has_top = j0 > 0;
        io.bottom = buf;
        io.has_bottom = j0 + R < t.m_max;
        io.error = &g.error;
        io.g_error = a.error;
        run_chunk(a, lds, io, t.ncols, t.e_lo, t.e_hi, j0);
    }
}

// A task worked by all W wavefronts of the workgroup: chunk c by wavefront
// c % W, each one trailing the wavefront of the chunk above by at least
// kAhead columns; chunk c's bottom boundary lives in buffer c % W.  Every
// chunk merges its bests into the pairs' keys itself.
template <int R, int W>
__device__ void run_task_group(const SwKernelArgs& a, GroupLds<R, W>& g, uint32_t* bufs,
                               size_t buf_words, int task) {
    const int wave = threadIdx.x / 64;
    const TaskView t = task_view(a, task);
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
        if (!run_chunk(a, g.wave[wave], io, nc, t.e_lo, t.e_hi, j0)) break;
        if (nc == 0) publish(io, io.pub_base);
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

// Results of every entry's pair from its key: score, ref_end1 and
// read_end1; key 0 (no cell above zero) is score 0, ref_end1 -1 and
// read_end1 0, or -1 for an empty query (ssw.c:148-149, 288-297).
__global__ void sw_decode_kernel(SwKernelArgs a, int32_t n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int p = a.pair_index ? a.pair_index[e] : e;
    const unsigned long long k = a.out_key[e];
    if (k == 0) {
        a.out_score[p] = 0;
        a.out_ref[p] = -1;
        a.out_read[p] = a.q_len[a.pair_query[p]] > 0 ? 0 : -1;
    } else {
        a.out_score[p] = (int)(k >> 48);
        a.out_ref[p] = 0xffffff - (int)((k >> 24) & 0xffffff);
        a.out_read[p] = 0xffffff - (int)(k & 0xffffff);
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

extern "C" hipError_t tpz_launch_sw_decode(const tpz::SwKernelArgs* args, int32_t n_entries,
                                           hipStream_t stream) {
    if (n_entries <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::sw_decode_kernel, dim3((n_entries + 255) / 256), dim3(256), 0, stream,
                       *args, n_entries);
    return hipGetLastError();
}

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
