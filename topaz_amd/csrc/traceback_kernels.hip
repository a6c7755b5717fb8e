// Traceback statistics of reported hits on gfx950: TOPAZ's ssw_align_stats
// (ssw.c:900-978) = a reverse Smith-Waterman pass for the start positions,
// then banded global alignment with traceback (banded_sw, ssw.c:538-784)
// for alignment length, gap openings, mismatches and identity.
//
// Only hits that are printed in BLAST tabular mode get here (at most -H per
// query), so the kernels use one thread per hit; each thread owns a slice of
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
#include <stdint.h>

#include "traceback_kernels.h"

namespace tpz {

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

__global__ void __launch_bounds__(64)
sw_reverse_kernel(ReverseArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_pairs) return;
    const int8_t* q = a.q_res + a.q_off[a.pair_query[p]];
    const int8_t* s = a.s_res + a.s_off[a.pair_subject[p]];
    const int qe = a.read_end[p], se = a.ref_end[p], score1 = a.score[p];
    const int L = qe + 1;
    const bool byte = a.bias + score1 < 255;
    const int P = byte ? 16 : 8;
    const int cap = byte ? 255 - a.bias : 32767;
    const int seg = (L + P - 1) / P;
    const int rows = seg * P;
    int32_t* H = a.workspace + (size_t)a.ws_off[p];
    int32_t* E = H + rows;
    for (int j = 0; j < rows; ++j) { H[j] = 0; E[j] = 0; }
    int best = 0, end_ref = byte ? -1 : 0, end_read = L - 1;
    bool overflow = false;
    for (int i = se; i >= 0; --i) {
        const int8_t* wrow = a.matrix + (int)s[i] * 25;
        int diag = 0, f = 0, colmax = 0;
        for (int j = 0; j < rows; ++j) {
            const int w = j < L ? (int)wrow[q[qe - j]] : 0;
            int h = imax(diag + w, 0);
            h = imax(h, imax(E[j], f));
            const int hg = imax(h - a.gap_open, 0);
            E[j] = imax(imax(E[j] - a.gap_extend, 0), hg);
            f = imax(imax(f - a.gap_extend, 0), hg);
            diag = H[j];
            H[j] = h;
            colmax = imax(colmax, h);
        }
        if (colmax >= cap) colmax = cap;
        if (colmax > best) {
            best = colmax;
            if (best >= cap) { overflow = true; break; }
            end_ref = i;
            end_read = L - 1;
            for (int j = 0; j < L; ++j)
                if (H[j] == best) { end_read = j; break; }
        }
        if (colmax == score1) break;
    }
    if (overflow) end_read = L - 1;
    a.out_ref_begin[p] = end_ref;
    a.out_read_begin[p] = qe - end_read;
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
                         const int8_t* mat, int go, int ge, int32_t* hb, int32_t* eb,
                         int32_t* hc, int8_t* dirs) {
    const int width = 2 * w + 3, width_d = 2 * w + 1;
    int best = 0;
    for (int j = 1; j < width - 1; ++j) hb[j] = 0;
    for (int i = 0; i < readLen; ++i) {
        const int beg = imax(0, i - w);
        const int end = imin(refLen - 1, i + w);
        const int edge = imin(end + 1, width - 1);
        int8_t* row = dirs ? dirs + (size_t)width_d * i * 3 : nullptr;
        int f = 0, u = 0;
        hb[0] = eb[0] = hb[edge] = eb[edge] = hc[0] = 0;
        for (int j = beg; j <= end; ++j) {
            const int up = band_slot(w, i - 1, j);
            const int left = band_slot(w, i, j - 1);
            const int dg = band_slot(w, i - 1, j - 1);
            u = band_slot(w, i, j);
            int open = (i == 0) ? -go : hb[up] - go;
            int ext = (i == 0) ? -ge : eb[up] - ge;
            eb[u] = open > ext ? open : ext;
            const int8_t de = open > ext ? 3 : 2;
            open = hc[left] - go;
            ext = f - ge;
            f = open > ext ? open : ext;
            const int8_t df = open > ext ? 5 : 4;
            const int e1 = imax(eb[u], 0), f1 = imax(f, 0);
            const int gap = e1 > f1 ? e1 : f1;
            const int sub = hb[dg] + mat[(int)ref[j] * 25 + read[i]];
            hc[u] = gap > sub ? gap : sub;
            best = imax(best, hc[u]);
            if (row) {
                row[band_dir(w, i, j, 0)] = de;
                row[band_dir(w, i, j, 1)] = df;
                row[band_dir(w, i, j, 2)] = gap <= sub ? 1 : (e1 > f1 ? de : df);
            }
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
    v.hb = a.workspace + (size_t)a.ws_off[p];
    v.eb = v.hb + a.ws_band_words;
    v.hc = v.eb + a.ws_band_words;
    // the reference's band rows start as fresh allocations: zero here
    for (int j = 0; j < a.ws_band_words; ++j) { v.hb[j] = 0; v.eb[j] = 0; v.hc[j] = 0; }
    return v;
}

__global__ void __launch_bounds__(64)
banded_search_kernel(BandArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_pairs) return;
    const BandView v = band_view(a, p);
    int w = v.w0, best = 0, status = 0;
    for (;;) {
        const int64_t cells = (int64_t)(2 * w + 1) * v.readLen * 3;
        // the reference's direction buffer stops growing at 2^31 bytes
        if (cells >= ((int64_t)1 << 30) || 2 * w + 4 > a.ws_band_words) { status = -1; break; }
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

// Replays every pass up to the final band into the hit's zeroed direction
// buffer (dir_size[p] bytes, the reference's power-of-two size), then
// traces back as ssw.c:663-739 does, indexing the buffer flat: a step that
// leaves the final band's row reads what the neighbouring row or an earlier
// pass wrote there.  A read outside the buffer ends as the reference's
// "trace back error".
__global__ void __launch_bounds__(64)
banded_trace_kernel(BandArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_pairs) return;
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

}  // namespace tpz

extern "C" hipError_t tpz_launch_sw_reverse(const tpz::ReverseArgs* a, hipStream_t stream) {
    if (a->n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::sw_reverse_kernel, dim3((a->n_pairs + 63) / 64), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_band_search(const tpz::BandArgs* a, hipStream_t stream) {
    if (a->n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::banded_search_kernel, dim3((a->n_pairs + 63) / 64), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_band_trace(const tpz::BandArgs* a, hipStream_t stream) {
    if (a->n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::banded_trace_kernel, dim3((a->n_pairs + 63) / 64), dim3(64), 0, stream, *a);
    return hipGetLastError();
}
