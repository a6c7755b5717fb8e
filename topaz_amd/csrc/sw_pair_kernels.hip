// Forward Smith-Waterman of one pair per wavefront on gfx950: the kernel
// of the plain schedule, which launches once per query and gives each of
// the query's pairs a wavefront of its own (a query's ~1,000 pairs then
// fill the device without being packed with other queries' pairs).
//
// Result per pair as ssw_align (ssw.c:835-897) reports it: the best score
// (32767 where the word path overflows), ref_end1 the first subject column
// reaching it and read_end1 the lowest query row reaching it in that column
// (ssw.c:264-297), i.e. the least (column, row) among the best cells; for
// gap_open > gap_extend the lane recurrence equals textbook affine-gap local
// alignment, which this kernel computes in int32.
//
// Systolic array: lane l holds query rows l*R .. l*R+R-1 of a pass of 64R
// rows in registers and works on subject column t - l at step t, taking the
// H and F of the row above from lane l-1's previous step (and the column's
// subject residue with them).  Longer queries take several passes; the last
// lane hands each column's bottom-row H and F to the next pass through the
// workspace.  The pass's query profile, W(residue, row) for 25 residues,
// sits in LDS.  A lane's cells come in (column, row) order, so its first
// strict improvement of the best is its least (column, row) with that
// score; passes and lanes combine by score, then column, then row.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "sw_pair_kernels.h"

namespace tpz {

struct PairBest {
    int s, k, j;
};

__device__ __forceinline__ bool better(const PairBest& a, const PairBest& b) {
    return a.s != b.s ? a.s > b.s : a.k != b.k ? a.k < b.k : a.j < b.j;
}

template <int R>
__global__ void __launch_bounds__(64) sw_pair_kernel(PairArgs a) {
    __shared__ __attribute__((aligned(16))) int8_t prof[25 * 64 * R];
    __shared__ int8_t mat[640];
    const int lane = threadIdx.x;
    const int p = a.pairs ? a.pairs[blockIdx.x] : (int)blockIdx.x;
    for (int i = lane; i < 625; i += 64) mat[i] = a.matrix[i];
    const int8_t* q = a.q_res + a.q_off[a.pair_query[p]];
    const int8_t* s = a.s_res + a.s_off[a.pair_subject[p]];
    const int m = a.q_len[a.pair_query[p]], n = a.s_len[a.pair_subject[p]];
    const int go = a.gap_open, ge = a.gap_extend;
    constexpr int kPass = 64 * R;
    const int n_pass = (m + kPass - 1) / kPass;
    int32_t* bot_h = a.workspace ? a.workspace + (size_t)a.ws_off[blockIdx.x] : nullptr;
    int32_t* bot_f = bot_h ? bot_h + n : nullptr;
    // score 0: ref_end1 -1, read_end1 0 (-1 for an empty query), ssw.c:148-149, 288-297
    PairBest best{0, -1, m > 0 ? 0 : -1};
    for (int pass = 0; pass < n_pass; ++pass) {
        const int base = pass * kPass + lane * R;
        __syncthreads();
        for (int x = 0; x < R; x += 4) {
            int qr[4];
            for (int y = 0; y < 4; ++y) qr[y] = base + x + y < m ? (int)q[base + x + y] : -1;
            for (int r = 0; r < 25; ++r) {
                uint32_t word = 0;
                for (int y = 0; y < 4; ++y) {
                    const int w = qr[y] >= 0 ? (int)mat[r * 25 + qr[y]] : 0;
                    word |= (uint32_t)(uint8_t)(int8_t)w << (8 * y);
                }
                *reinterpret_cast<uint32_t*>(prof + r * kPass + lane * R + x) = word;
            }
        }
        __syncthreads();
        const int real = min(R, max(m - base, 0));       // this lane's rows that exist
        int32_t H[R], E[R];
#pragma unroll
        for (int r = 0; r < R; ++r) { H[r] = 0; E[r] = 0; }
        PairBest pb{0, INT_MAX, INT_MAX};
        const bool last_pass = pass + 1 == n_pass;
        int prev_top = 0, out_h = 0, out_f = 0, out_sres = 0;
        int s_buf = 0, bh_buf = 0, bf_buf = 0;
        const int steps = n + 63;
        for (int t = 0; t < steps; ++t) {
            // every 64 steps: the next 64 columns' residues (and the previous
            // pass's bottom row) that lane 0 feeds in, one per step
            if ((t & 63) == 0) {
                const int kk = t + lane;
                const bool ok = kk < n;
                s_buf = ok ? (int)s[kk] : 0;
                if (pass > 0) {
                    bh_buf = ok ? bot_h[kk] : 0;
                    bf_buf = ok ? bot_f[kk] : 0;
                }
            }
            int rh = __shfl_up(out_h, 1), rf = __shfl_up(out_f, 1), rs = __shfl_up(out_sres, 1);
            const int s0 = __shfl(s_buf, t & 63);
            int h0 = 0, f0 = 0;
            if (pass > 0) {
                h0 = __shfl(bh_buf, t & 63);
                f0 = __shfl(bf_buf, t & 63);
            }
            const int k = t - lane;
            const bool active = k >= 0 && k < n;
            if (lane == 0) {
                rh = active ? h0 : 0;
                rf = active ? f0 : 0;
                rs = s0;
            }
            if (active) {
                uint32_t pw[R / 4];
                const int8_t* src = prof + rs * kPass + lane * R;
                if constexpr (R == 8) {
                    const uint2 p0 = *reinterpret_cast<const uint2*>(src);
                    pw[0] = p0.x; pw[1] = p0.y;
                } else {
#pragma unroll
                    for (int c = 0; c < R / 16; ++c) {
                        const uint4 pc = reinterpret_cast<const uint4*>(src)[c];
                        pw[4 * c] = pc.x; pw[4 * c + 1] = pc.y; pw[4 * c + 2] = pc.z; pw[4 * c + 3] = pc.w;
                    }
                }
                int diag = prev_top, f = rf;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int w = (int)(int8_t)(pw[r >> 2] >> (8 * (r & 3)));
                    const int h = max(max(diag + w, 0), max(E[r], f));
                    const int hg = h - go;
                    E[r] = max(E[r] - ge, hg);
                    f = max(f - ge, hg);
                    diag = H[r];
                    H[r] = h;
                    if (r < real && h > pb.s) { pb.s = h; pb.k = k; pb.j = base + r; }
                }
                out_h = H[R - 1];
                out_f = f;
                out_sres = rs;
                prev_top = rh;
                if (lane == 63 && !last_pass) {
                    bot_h[k] = out_h;
                    bot_f[k] = out_f;
                }
            } else {
                out_h = 0;
                out_f = 0;
                out_sres = rs;
                prev_top = k < 0 ? 0 : prev_top;
            }
        }
        if (pb.s > 0 && better(pb, best)) best = pb;
        // the bottom row is read by other lanes in the next pass
        __threadfence_block();
        __syncthreads();
    }
    for (int d = 32; d >= 1; d >>= 1) {
        PairBest o{__shfl_xor(best.s, d), __shfl_xor(best.k, d), __shfl_xor(best.j, d)};
        if (better(o, best)) best = o;
    }
    if (lane == 0) {
        a.out_score[p] = best.s >= 32767 ? 32767 : best.s;
        a.out_ref[p] = best.k;
        a.out_read[p] = best.j;
    }
}

}  // namespace tpz

extern "C" int tpz_sw_pair_rows(int32_t m) {
    return m <= 512 ? 8 : m <= 1024 ? 16 : 32;
}

extern "C" hipError_t tpz_launch_sw_pair(const tpz::PairArgs* a, int rows, hipStream_t stream) {
    if (a->n_list <= 0) return hipSuccess;
    const dim3 grid(a->n_list), block(64);
    switch (rows) {
        case 8: hipLaunchKernelGGL(tpz::sw_pair_kernel<8>, grid, block, 0, stream, *a); break;
        case 16: hipLaunchKernelGGL(tpz::sw_pair_kernel<16>, grid, block, 0, stream, *a); break;
        case 32: hipLaunchKernelGGL(tpz::sw_pair_kernel<32>, grid, block, 0, stream, *a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
