// Traceback statistics of reported hits on gfx950: TOPAZ's ssw_align_stats
// (ssw.c:900-978) = a reverse Smith-Waterman pass for the start positions,
// then banded global alignment with traceback (banded_sw, ssw.c:538-784)
// for alignment length, gap openings, mismatches and identity.
//
// Only hits that are printed in BLAST tabular mode get here (at most -H per
// query).  The reverse pass, a DP over the whole prefix, takes a wavefront
// per hit; the banded kernels take one thread per hit, each with a slice of
// a workspace sized by the host.
//
// Reverse pass: the reversed query prefix q[read_end1..0] against subject
// columns ref_end1..0.  The reference runs it with 16 byte lanes when
// bias + score1 < 255, else 8 word lanes, padding the query to a multiple
// of the lane count with rows scoring 0 that take part in the column max;
// for gapO > gapE its H equals textbook local alignment (see
// sw_kernels.hip).  The scan stops after the first column whose max equals
// score1; the start is that last column of a strict max rise and the lowest
// real row there equal to the max.
//
// banded_sw: band half-width w starts at |refLen - readLen| + 1 and doubles
// until the band's best cell reaches score1 (search kernel); the passes are
// then replayed with their direction cells and the final band traced back
// from the bottom-right corner while the row is > 0 (trace kernel).
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "traceback_kernels.h"

namespace tpz {

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// ---- reverse pass: one wavefront per hit ----
//
// The DP runs as a systolic array over the wavefront: lane l holds query
// rows l*32 .. l*32+31 of a pass of 2048 rows in registers and works on
// subject column t - l at step t, taking the H and F of the row above (and
// the column max of the rows above) from lane l-1's previous step.  A
// query prefix longer than 2048 rows takes several passes; the last lane
// hands each column's bottom-row H, F and column max to the next pass
// through the workspace.  The scores of a pass's rows for all 25 residues
// sit in LDS.
//
// The scan's stopping rule needs the column maxima in scan order, and the
// start row needs the H of one column, so the kernel (a) sweeps all
// columns recording each column's max and, per lane, the first (column,
// row) where a real row's H equals score1; (b) replays the reference's scan
// over the recorded maxima; (c) only when the scan's last rise is not to
// score1 (a score1 that is not the pair's forward score), sweeps again up
// to that column to find the row.
// Rows per lane (reverse_rows_per_lane): a kernel instance per class, so a
// short prefix's launch holds a small LDS profile (25 x 64 x rows bytes)
// and more hits fit a CU.

struct RevTrack {
    int k, j;                                   // first (column, row) with H == target
};

__device__ __forceinline__ int extract_s8(const uint32_t* w, int r) {
    return (int)(int8_t)(w[r >> 2] >> (8 * (r & 3)));
}

// Columns 0..k_last of the reversed prefix DP (column k is subject residue
// se - k, row j query residue qe - j, rows >= L score 0, rows >= rows do
// not exist).  Writes each column's max to cmax; returns this lane's
// first (column, row) with H == target among real rows.
template <int kRevRows>
__device__ __forceinline__ RevTrack rev_sweep(const int8_t* q, const int8_t* s, int qe, int se,
                                              int L, int rows, const int8_t* mat_lds, int go,
                                              int ge, int8_t* prof, int32_t* bot_h,
                                              int32_t* bot_f, int32_t* cmax, int k_last,
                                              int target) {
    constexpr int kRevPassRows = 64 * kRevRows;
    const int lane = threadIdx.x;
    const int n_pass = (rows + kRevPassRows - 1) / kRevPassRows;
    RevTrack tr{INT_MAX, INT_MAX};
    for (int pass = 0; pass < n_pass; ++pass) {
        const int base = pass * kRevPassRows + lane * kRevRows;
        // this pass's profile: prof[r][x] = W(r, row pass*2048 + x)
        __threadfence_block();
        __syncthreads();
        for (int x = 0; x < kRevRows; x += 4) {
            int qr[4];
            for (int y = 0; y < 4; ++y) {
                const int j = base + x + y;
                qr[y] = j < L ? (int)q[qe - j] : -1;
            }
            for (int r = 0; r < 25; ++r) {
                uint32_t word = 0;
                for (int y = 0; y < 4; ++y) {
                    const int w = qr[y] >= 0 ? (int)mat_lds[r * 25 + qr[y]] : 0;
                    word |= (uint32_t)(uint8_t)(int8_t)w << (8 * y);
                }
                *reinterpret_cast<uint32_t*>(prof + r * kRevPassRows + lane * kRevRows + x) = word;
            }
        }
        __syncthreads();
        int32_t H[kRevRows], E[kRevRows];
        for (int r = 0; r < kRevRows; ++r) { H[r] = 0; E[r] = 0; }
        const bool last_pass = pass + 1 == n_pass;
        int prev_top = 0;                        // H of the row above at the previous column
        int out_h = 0, out_f = 0, out_cm = 0, out_sres = 0;
        int s_buf = 0, bh_buf = 0, bf_buf = 0, bc_buf = 0;
        const int steps = k_last + 1 + 63;
        for (int t = 0; t < steps; ++t) {
            // every 64 steps the wavefront fetches the next 64 columns'
            // subject residues (and the previous pass's bottom row) that
            // lane 0 feeds in, one per step
            if ((t & 63) == 0) {
                const int kk = t + lane;
                const bool ok = kk <= k_last;
                s_buf = ok ? (int)s[se - kk] : 0;
                if (pass > 0) {
                    bh_buf = ok ? bot_h[kk] : 0;
                    bf_buf = ok ? bot_f[kk] : 0;
                    bc_buf = ok ? cmax[kk] : 0;
                }
            }
            int rh = __shfl_up(out_h, 1), rf = __shfl_up(out_f, 1);
            int rc = __shfl_up(out_cm, 1), rs = __shfl_up(out_sres, 1);
            const int s0 = __shfl(s_buf, t & 63);
            int h0 = 0, f0 = 0, c0 = 0;
            if (pass > 0) {
                h0 = __shfl(bh_buf, t & 63);
                f0 = __shfl(bf_buf, t & 63);
                c0 = __shfl(bc_buf, t & 63);
            }
            const int k = t - lane;
            const bool active = k >= 0 && k <= k_last;
            if (lane == 0) {
                rh = active ? h0 : 0;
                rf = active ? f0 : 0;
                rc = active ? c0 : 0;
                rs = s0;
            }
            if (active) {
                uint32_t pw[kRevRows / 4];
                const int8_t* src = prof + rs * kRevPassRows + lane * kRevRows;
                if constexpr (kRevRows == 8) {
                    const uint2 p0 = *reinterpret_cast<const uint2*>(src);
                    pw[0] = p0.x; pw[1] = p0.y;
                } else {
#pragma unroll
                    for (int c = 0; c < kRevRows / 16; ++c) {
                        const uint4 pc = reinterpret_cast<const uint4*>(src)[c];
                        pw[4 * c] = pc.x; pw[4 * c + 1] = pc.y; pw[4 * c + 2] = pc.z; pw[4 * c + 3] = pc.w;
                    }
                }
                int diag = prev_top, f = rf, mymax = 0;
#pragma unroll
                for (int r = 0; r < kRevRows; ++r) {
                    const int w = extract_s8(pw, r);
                    int h = max(diag + w, 0);
                    h = max(h, max(E[r], f));
                    const int hg = max(h - go, 0);
                    E[r] = max(max(E[r] - ge, 0), hg);
                    f = max(max(f - ge, 0), hg);
                    diag = H[r];
                    H[r] = h;
                    const int j = base + r;
                    if (j < rows) mymax = max(mymax, h);
                    if (h == target && j < L && k < tr.k) { tr.k = k; tr.j = j; }
                }
                out_h = H[kRevRows - 1];
                out_f = f;
                out_cm = max(rc, mymax);
                out_sres = rs;
                prev_top = rh;
                if (lane == 63) {
                    cmax[k] = out_cm;
                    if (!last_pass) { bot_h[k] = out_h; bot_f[k] = out_f; }
                }
            } else {
                out_h = 0; out_f = 0; out_cm = 0; out_sres = rs;
                prev_top = k < 0 ? 0 : prev_top;
            }
        }
    }
    // the wavefront's column maxima and bottom rows are read by other lanes next
    __threadfence_block();
    __syncthreads();
    return tr;
}

__device__ __forceinline__ RevTrack wave_min(RevTrack t) {
    for (int d = 32; d >= 1; d >>= 1) {
        const int ok = __shfl_xor(t.k, d), oj = __shfl_xor(t.j, d);
        if (ok < t.k || (ok == t.k && oj < t.j)) { t.k = ok; t.j = oj; }
    }
    return t;
}

template <int kRows>
__global__ void __launch_bounds__(64)
sw_reverse_kernel(ReverseArgs a) {
    __shared__ __attribute__((aligned(16))) int8_t prof[25 * 64 * kRows];
    __shared__ int8_t mat[625];
    const int p = a.hits ? a.hits[blockIdx.x] : blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 625; i += 64) mat[i] = a.matrix[i];
    __syncthreads();
    const int8_t* q = a.q_res + a.q_off[a.pair_query[p]];
    const int8_t* s = a.s_res + a.s_off[a.pair_subject[p]];
    const int qe = a.read_end[p], se = a.ref_end[p], score1 = a.score[p];
    const int L = qe + 1;
    const bool byte = a.bias + score1 < 255;
    const int P = byte ? 16 : 8;
    const int cap = byte ? 255 - a.bias : 32767;
    const int rows = (L + P - 1) / P * P;
    int32_t* bot_h = a.workspace + (size_t)a.ws_off[p];
    int32_t* bot_f = bot_h + (se + 1);
    int32_t* cmax = bot_f + (se + 1);
    // (a) sweeps all columns; (c) when needed, sweeps again up to the rise
    // column (one call site keeps one copy of the sweep's code)
    int best = 0, end_ref = byte ? -1 : 0, rise_k = -1;
    bool overflow = false;
    int end_read = L - 1;
    int k_last = se, target = score1;
    for (int phase = 0; phase < 2; ++phase) {
        const RevTrack t = wave_min(rev_sweep<kRows>(q, s, qe, se, L, rows, mat, a.gap_open,
                                                     a.gap_extend, prof, bot_h, bot_f, cmax,
                                                     k_last, target));
        if (phase == 1) {
            if (t.k == rise_k) end_read = t.j;
            break;
        }
        // (b) the reference's scan (ssw.c:264-297 on the reversed prefixes):
        // the start is the column of the last strict rise of the running
        // max, the scan stopping at the first column whose max equals score1
        // or at an overflow of the byte range
        bool stop = false;
        for (int c0 = 0; c0 <= se && !stop; c0 += 64) {
            const int v = c0 + lane <= se ? cmax[c0 + lane] : 0;
            const int n = min(64, se + 1 - c0);
            for (int i = 0; i < n; ++i) {
                int cm = __shfl(v, i);
                if (cm >= cap) cm = cap;
                if (cm > best) {
                    best = cm;
                    if (best >= cap) { overflow = true; stop = true; break; }
                    end_ref = se - (c0 + i);
                    rise_k = c0 + i;
                }
                if (cm == score1) { stop = true; break; }
            }
        }
        // the lowest real row of the rise column equal to the max: from this
        // sweep when the max is score1, else from a second sweep
        if (overflow || rise_k < 0) break;
        if (best == score1) {
            if (t.k == rise_k) end_read = t.j;
            break;
        }
        k_last = rise_k;
        target = best;
    }
    if (lane == 0) {
        a.out_ref_begin[p] = end_ref;
        a.out_read_begin[p] = qe - end_read;
    }
}

// Band slot of column j on row i (slot 0 is the left border) and direction
// cell, for band half-width w.
__device__ __forceinline__ int band_slot(int w, int i, int j) { return j - imax(i - w, 0) + 1; }
__device__ __forceinline__ int band_dir(int w, int i, int j, int d) {
    return (j - imax(i - w, 0)) * 3 + d;
}

// One pass of the band DP with half-width w; returns the pass's best cell.
// As in the reference (ssw.c:559-660) the three band rows live across all
// passes: H is cleared on slots 1..width-2 at the start of a pass, E never,
// and slots a row does not write keep what an earlier row or pass left.
// When dirs is set the pass writes its direction cells there, row i at
// byte (2w+1) * 3 * i.
__device__ int band_pass(const int8_t* ref, const int8_t* read, int refLen, int readLen, int w,
                         const int8_t* mat, int go, int ge, int32_t* __restrict__ hb,
                         int32_t* __restrict__ eb, int32_t* __restrict__ hc,
                         int8_t* __restrict__ dirs) {
    const int width = 2 * w + 3, width_d = 2 * w + 1;
    int best = 0;
    for (int j = 1; j < width - 1; ++j) hb[j] = 0;
    for (int i = 0; i < readLen; ++i) {
        const int beg = imax(0, i - w);
        const int end = imin(refLen - 1, i + w);
        const int edge = imin(end + 1, width - 1);
        int8_t* row = dirs ? dirs + (size_t)width_d * i * 3 : nullptr;
        const int8_t* mrow = mat + (int)read[i];
        int f = 0, u = 0;
        hb[0] = eb[0] = hb[edge] = eb[edge] = hc[0] = 0;
        // the cell to the left (hc[left]) and the diagonal (hb[dg]) are the
        // previous cell's hc[u] and hb[up]: kept in registers
        int left_h = 0;
        int up = band_slot(w, i - 1, beg);
        int diag_h = hb[band_slot(w, i - 1, beg - 1)];
        {
            const int left = band_slot(w, i, beg - 1);
            left_h = hc[left];
        }
        for (int j = beg; j <= end; ++j) {
            u = band_slot(w, i, j);
            const int hb_up = hb[up], eb_up = eb[up];
            int open = (i == 0) ? -go : hb_up - go;
            int ext = (i == 0) ? -ge : eb_up - ge;
            const int e = open > ext ? open : ext;
            eb[u] = e;
            const int8_t de = open > ext ? 3 : 2;
            open = left_h - go;
            ext = f - ge;
            f = open > ext ? open : ext;
            const int8_t df = open > ext ? 5 : 4;
            const int e1 = imax(e, 0), f1 = imax(f, 0);
            const int gap = e1 > f1 ? e1 : f1;
            const int sub = diag_h + mrow[(int)ref[j] * 25];
            const int h = gap > sub ? gap : sub;
            hc[u] = h;
            best = imax(best, h);
            if (row) {
                row[band_dir(w, i, j, 0)] = de;
                row[band_dir(w, i, j, 1)] = df;
                row[band_dir(w, i, j, 2)] = gap <= sub ? 1 : (e1 > f1 ? de : df);
            }
            left_h = h;
            diag_h = hb_up;
            ++up;
        }
        for (int j = 1; j <= u; ++j) hb[j] = hc[j];
    }
    return best;
}

struct BandView {
    const int8_t* ref;
    const int8_t* read;
    int refLen, readLen, w0;
    int32_t *hb, *eb, *hc;
};

__device__ __forceinline__ BandView band_view(const BandArgs& a, int p) {
    BandView v;
    const int rb = a.ref_begin[p], qb = a.read_begin[p];
    v.refLen = a.ref_end[p] - rb + 1;
    v.readLen = a.read_end[p] - qb + 1;
    v.ref = a.s_res + a.s_off[a.pair_subject[p]] + rb;
    v.read = a.q_res + a.q_off[a.pair_query[p]] + qb;
    v.w0 = v.refLen > v.readLen ? v.refLen - v.readLen + 1 : v.readLen - v.refLen + 1;
    const int32_t words = band_row_words(v.refLen, v.readLen);
    v.hb = a.workspace + (size_t)a.ws_off[p];
    v.eb = v.hb + words;
    v.hc = v.eb + words;
    // the reference's band rows start as fresh allocations: zero here
    for (int j = 0; j < words; ++j) { v.hb[j] = 0; v.eb[j] = 0; v.hc[j] = 0; }
    return v;
}

__global__ void __launch_bounds__(64)
banded_search_kernel(BandArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.n_list) return;
    const int p = a.hits ? a.hits[x] : x;
    const BandView v = band_view(a, p);
    int w = v.w0, best = 0, status = 0;
    for (;;) {
        const int64_t cells = (int64_t)(2 * w + 1) * v.readLen * 3;
        // the reference's direction buffer stops growing at 2^31 bytes
        if (cells >= ((int64_t)1 << 30) || 2 * w + 4 > band_row_words(v.refLen, v.readLen)) { status = -1; break; }
        best = imax(best, band_pass(v.ref, v.read, v.refLen, v.readLen, w, a.matrix, a.gap_open,
                                    a.gap_extend, v.hb, v.eb, v.hc, nullptr));
        if (best >= a.score[p]) break;
        // once the band covers the whole rectangle, wider bands give the same
        // best cell and the reference doubles until its size check fails
        if (w >= v.refLen + v.readLen) { status = -1; break; }
        w *= 2;
    }
    a.out_band[p] = status < 0 ? -1 : w;
}

// Traces back from the bottom-right corner as ssw.c:663-739 does, indexing
// the direction buffer flat: a step that leaves the final band's row reads
// what the neighbouring row or an earlier pass wrote there.  A read outside
// the buffer ends as the reference's "trace back error".
__device__ void band_walk(const BandArgs& a, int p, const BandView& v, int wf, const int8_t* dirs,
                          int64_t size) {
    const int width_d = 2 * wf + 1;
    int i = v.readLen - 1, j = v.refLen - 1, state = 2;
    int length = 0, gapopen = 0, mismatch = 0, identical = 0, rc = 1;
    char prev = 'M';
    int64_t line = (int64_t)width_d * 3 * (v.readLen > 0 ? v.readLen - 1 : 0);
    while (i > 0) {
        const int64_t c = line + band_dir(wf, i, j, state);
        const int8_t code = (c >= 0 && c < size) ? dirs[c] : 0;
        char op;
        if (code == 1) { --i; --j; state = 2; line -= width_d * 3; op = 'M'; }
        else if (code == 2) { --i; state = 0; line -= width_d * 3; op = 'I'; }
        else if (code == 3) { --i; state = 2; line -= width_d * 3; op = 'I'; }
        else if (code == 4) { --j; state = 1; op = 'D'; }
        else if (code == 5) { --j; state = 2; op = 'D'; }
        else { rc = 0; break; }
        if (op == 'M') {
            if (v.read[i] != v.ref[j]) ++mismatch;
            else ++identical;
        }
        if ((op == 'I' && prev != 'I') || (op == 'D' && prev != 'D')) ++gapopen;
        ++length;
        prev = op;
    }
    a.out_rc[p] = rc;
    a.out_length[p] = length;
    a.out_gapopen[p] = gapopen;
    a.out_mismatch[p] = mismatch;
    a.out_identical[p] = rc ? identical : -2;     // -2: identity stays 0.0
}

// Replays every pass up to the final band into the hit's zeroed direction
// buffer (dir_size[p] bytes, the reference's power-of-two size), then
// traces back (band_walk).
__global__ void __launch_bounds__(64)
banded_trace_kernel(BandArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.n_list) return;
    const int p = a.hits ? a.hits[x] : x;
    const int wf = a.out_band[p];
    if (wf < 0) {
        a.out_rc[p] = 0; a.out_length[p] = -1; a.out_gapopen[p] = -1;
        a.out_mismatch[p] = -1; a.out_identical[p] = -1;
        return;
    }
    const BandView v = band_view(a, p);
    int8_t* dirs = a.dirs + a.dir_off[p];
    const int64_t size = a.dir_size[p];
    for (int w = v.w0;; w *= 2) {
        band_pass(v.ref, v.read, v.refLen, v.readLen, w, a.matrix, a.gap_open, a.gap_extend,
                  v.hb, v.eb, v.hc, dirs);
        if (w >= wf) break;
    }
    band_walk(a, p, v, wf, dirs, size);
}


// ---- banded_sw with one wavefront per hit ----
//
// The same cells, values and direction codes as band_pass, with the lanes
// across the band and the three band rows in LDS.  Within a row, E and the
// diagonal come from the row above; the horizontal gap F is a running max
// along the row.  With G the cell's value without F (max(E, 0) and the
// diagonal; G >= 0), the previous cell's value is H = max(G, F, 0), so
// F(j) = max(H(j-1) - gapO, F(j-1) - gapE)
//      = max(G(j-1) - gapO, F(j-1) - min(gapO, gapE))
// (the -gapO of a zero H is below G - gapO).  F over a chunk of 64 cells is
// then a prefix max of G(j-1) - gapO + l*c, c = min(gapO, gapE), carried
// from chunk to chunk.
struct BandLds {
    int32_t hb[kBandSlots], eb[kBandSlots], hc[kBandSlots];
};

__device__ int band_pass_wave(const int8_t* ref, const int8_t* read, int refLen, int readLen,
                              int w, const int8_t* mat, int go, int ge, BandLds& L, int8_t* dirs) {
    const int lane = threadIdx.x;
    const int width = 2 * w + 3, width_d = 2 * w + 1;
    int best = 0;
    for (int j = 1 + lane; j < width - 1; j += 64) L.hb[j] = 0;
    __syncthreads();
    for (int i = 0; i < readLen; ++i) {
        const int beg = imax(0, i - w);
        const int end = imin(refLen - 1, i + w);
        const int edge = imin(end + 1, width - 1);
        if (lane == 0) {
            L.hb[0] = 0; L.eb[0] = 0; L.hb[edge] = 0; L.eb[edge] = 0; L.hc[0] = 0;
        }
        __syncthreads();
        const int8_t* mrow = mat + (int)read[i];
        int8_t* row = dirs ? dirs + (size_t)width_d * i * 3 : nullptr;
        const int sh_prev = imax(i - 1 - w, 0), sh = imax(i - w, 0);
        // the cell left of the first: hc[0] = 0, and F = 0 before the row
        int carry_g = 0, carry_f = 0, carry_h = 0;
        for (int j0 = beg; j0 <= end; j0 += 64) {
            const int j = j0 + lane;
            const bool valid = j <= end;
            const int up = j - sh_prev + 1, u = j - sh + 1;
            int hb_up = 0, eb_up = 0, hb_dg = 0, wsc = 0;
            if (valid) {
                hb_up = L.hb[up];
                eb_up = L.eb[up];
                hb_dg = L.hb[up - 1];
                wsc = mrow[(int)ref[j] * 25];
            }
            const int open = (i == 0) ? -go : hb_up - go;
            const int ext = (i == 0) ? -ge : eb_up - ge;
            const int e = open > ext ? open : ext;
            const int8_t de = open > ext ? 3 : 2;
            const int sub = hb_dg + wsc;
            const int e1 = imax(e, 0);
            const int g = imax(e1, sub);
            int g_prev = __shfl_up(g, 1);
            if (lane == 0) g_prev = carry_g;
            const int c = imin(go, ge);
            int b = g_prev - go + lane * c;
            if (lane == 0) b = imax(b, carry_f - c);
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(b, d);
                if (lane >= d) b = imax(b, o);
            }
            const int f = b - lane * c;
            const int f1 = imax(f, 0);
            const int gap = e1 > f1 ? e1 : f1;
            const int h = gap > sub ? gap : sub;
            int h_prev = __shfl_up(h, 1), f_prev = __shfl_up(f, 1);
            if (lane == 0) { h_prev = carry_h; f_prev = carry_f; }
            const int8_t df = (h_prev - go) > (f_prev - ge) ? 5 : 4;
            if (valid) {
                L.eb[u] = e;
                L.hc[u] = h;
                best = imax(best, h);
                if (row) {
                    row[band_dir(w, i, j, 0)] = de;
                    row[band_dir(w, i, j, 1)] = df;
                    row[band_dir(w, i, j, 2)] = gap <= sub ? 1 : (e1 > f1 ? de : df);
                }
            }
            const int last = imin(63, end - j0);
            carry_g = __shfl(g, last);
            carry_f = __shfl(f, last);
            carry_h = __shfl(h, last);
        }
        __syncthreads();
        const int u_last = end >= beg ? end - sh + 1 : 0;
        for (int jj = 1 + lane; jj <= u_last; jj += 64) L.hb[jj] = L.hc[jj];
        __syncthreads();
    }
    for (int d = 32; d >= 1; d >>= 1) best = imax(best, __shfl_xor(best, d));
    return best;
}

__device__ __forceinline__ BandView band_view_wave(const BandArgs& a, int p, BandLds& L) {
    BandView v;
    const int rb = a.ref_begin[p], qb = a.read_begin[p];
    v.refLen = a.ref_end[p] - rb + 1;
    v.readLen = a.read_end[p] - qb + 1;
    v.ref = a.s_res + a.s_off[a.pair_subject[p]] + rb;
    v.read = a.q_res + a.q_off[a.pair_query[p]] + qb;
    v.w0 = v.refLen > v.readLen ? v.refLen - v.readLen + 1 : v.readLen - v.refLen + 1;
    v.hb = v.eb = v.hc = nullptr;
    // the reference's band rows start as fresh allocations: zero here
    for (int j = threadIdx.x; j < kBandSlots; j += 64) { L.hb[j] = 0; L.eb[j] = 0; L.hc[j] = 0; }
    __syncthreads();
    return v;
}

__global__ void __launch_bounds__(64)
banded_search_wave_kernel(BandArgs a) {
    __shared__ BandLds L;
    const int p = a.hits ? a.hits[blockIdx.x] : blockIdx.x;
    const BandView v = band_view_wave(a, p, L);
    int w = v.w0, best = 0, status = 0;
    for (;;) {
        const int64_t cells = (int64_t)(2 * w + 1) * v.readLen * 3;
        if (cells >= ((int64_t)1 << 30) || 2 * w + 4 > band_row_words(v.refLen, v.readLen)) { status = -1; break; }
        if (2 * w + 3 > kBandSlots) { status = -2; break; }
        best = imax(best, band_pass_wave(v.ref, v.read, v.refLen, v.readLen, w, a.matrix,
                                         a.gap_open, a.gap_extend, L, nullptr));
        if (best >= a.score[p]) break;
        if (w >= v.refLen + v.readLen) { status = -1; break; }
        w *= 2;
    }
    if (threadIdx.x == 0) a.out_band[p] = status < 0 ? status : w;
}

__global__ void __launch_bounds__(64)
banded_trace_wave_kernel(BandArgs a) {
    __shared__ BandLds L;
    const int p = a.hits ? a.hits[blockIdx.x] : blockIdx.x;
    const int wf = a.out_band[p];
    if (wf < 0) {
        if (threadIdx.x == 0) {
            a.out_rc[p] = 0; a.out_length[p] = -1; a.out_gapopen[p] = -1;
            a.out_mismatch[p] = -1; a.out_identical[p] = -1;
        }
        return;
    }
    const BandView v = band_view_wave(a, p, L);
    int8_t* dirs = a.dirs + a.dir_off[p];
    for (int w = v.w0;; w *= 2) {
        band_pass_wave(v.ref, v.read, v.refLen, v.readLen, w, a.matrix, a.gap_open, a.gap_extend,
                       L, dirs);
        if (w >= wf) break;
    }
    // the walk reads the direction cells the other lanes wrote
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0) band_walk(a, p, v, wf, dirs, a.dir_size[p]);
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_sw_reverse(const tpz::ReverseArgs* a, int rpl, hipStream_t stream) {
    const int n = a->hits ? a->n_list : a->n_pairs;
    if (n <= 0) return hipSuccess;
    if (rpl == 8)
        hipLaunchKernelGGL(tpz::sw_reverse_kernel<8>, dim3(n), dim3(64), 0, stream, *a);
    else if (rpl == 16)
        hipLaunchKernelGGL(tpz::sw_reverse_kernel<16>, dim3(n), dim3(64), 0, stream, *a);
    else if (rpl == 32)
        hipLaunchKernelGGL(tpz::sw_reverse_kernel<32>, dim3(n), dim3(64), 0, stream, *a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_band_search(const tpz::BandArgs* a, int wave, hipStream_t stream) {
    if (a->n_list <= 0) return hipSuccess;
    if (wave)
        hipLaunchKernelGGL(tpz::banded_search_wave_kernel, dim3(a->n_list), dim3(64), 0, stream, *a);
    else
        hipLaunchKernelGGL(tpz::banded_search_kernel, dim3((a->n_list + 63) / 64), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_band_trace(const tpz::BandArgs* a, int wave, hipStream_t stream) {
    if (a->n_list <= 0) return hipSuccess;
    if (wave)
        hipLaunchKernelGGL(tpz::banded_trace_wave_kernel, dim3(a->n_list), dim3(64), 0, stream, *a);
    else
        hipLaunchKernelGGL(tpz::banded_trace_kernel, dim3((a->n_list + 63) / 64), dim3(64), 0, stream, *a);
    return hipGetLastError();
}
