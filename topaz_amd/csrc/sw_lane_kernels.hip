// The reference's striped Smith-Waterman lanes on gfx950, for the gap
// settings (gap_open <= gap_extend) where they differ from textbook
// affine-gap local alignment: sw_sse2_byte (ssw.c:127-332) and
// sw_sse2_word (ssw.c:358-536) reproduced lane for lane, with ssw_align's
// dispatch (ssw.c:835-897) and ssw_align_stats' reverse pass
// (ssw.c:900-941).
//
// The reference splits the query into P lanes (16 in the byte pass, 8 in
// the word pass) of segLen = ceil(m/P) consecutive rows, padded with rows
// that score 0.  Per subject column its main loop carries the vertical gap
// value only inside a lane (restarting it on each lane's first row), keeps
// Hp = max(H_diag + W, E, F_in, 0) and updates E from Hp; the lazy-F loop
// then carries each lane's last gap value into the next lane, raising H but
// never E.  The byte loop tests before raising and wraps around the
// segments with a further lane shift; the word loop raises and then tests,
// for at most 8 sweeps.  Values are floored at 0 (saturating arithmetic);
// the byte pass gives up once a column's maximum reaches 255 - bias (its
// saturation point, reported as 255), the word pass at 32767.
//
// Mapping: a group of 16 lanes per pair, 4 pairs per wavefront; lane L of a
// group is vector element L of the reference.  H of the previous column, H
// of this one and E live in the group's workspace, stored segment-major
// (segment s of lane L at s * P + L), so a group's accesses per segment are
// consecutive.  The lazy-F exit tests are group ballots.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_lane_kernels.h"

namespace tpz {

namespace {

struct ScanResult {
    int score, ref, read;
};

__device__ __forceinline__ int group_max(int v) {
    for (int d = 8; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 16));
    return v;
}

__device__ __forceinline__ int group_min(int v) {
    for (int d = 8; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 16));
    return v;
}

// The bits of this lane's group in a wavefront ballot.
__device__ __forceinline__ uint32_t group_ballot(bool p, int grp) {
    return (uint32_t)(__ballot(p) >> (16 * grp)) & 0xffffu;
}

// Workspace rows written by some lanes of a group and read by others (the
// group's lanes are in one wavefront, on one CU).
__device__ __forceinline__ void column_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// A lane's carry handed to the next lane; lane 0 gets 0.
__device__ __forceinline__ int shift_lanes(int v, int lane) {
    const int u = __shfl_up(v, 1, 16);
    return lane == 0 ? 0 : u;
}

// One scan (ssw.c:127-332 for P = 16, :358-536 for P = 8).  q(j): the matrix
// row of query row j < m; columns i = 0 .. n-1 (dir 0) or n-1 .. 0 (dir 1)
// of subject s.  Stops after the first column whose maximum equals
// terminate (-1: never).
template <typename Query>
__device__ ScanResult scan(const Query& q, int m, const int8_t* s, int n, int dir, int P,
                           int cap, int ovf, int terminate, const int8_t* mat, int go, int ge,
                           int32_t* hprev, int32_t* hcur, int32_t* e, int lane, int grp) {
    if (m <= 0) return ScanResult{0, P == 16 ? -1 : 0, -1};
    const int seg = (m + P - 1) / P;
    const int rows = seg * P;
    const bool on = lane < P;
    for (int x = lane; x < rows; x += 16) {
        hprev[x] = 0;
        e[x] = 0;
    }
    column_fence();
    int best = 0, end_ref = P == 16 ? -1 : 0;
    // end_read of the column kept at the last rise (0 while none: the
    // reference's kept column starts as zeros)
    int end_read = m > 0 ? 0 : -1;
    for (int k = 0; k < n; ++k) {
        const int i = dir ? n - 1 - k : k;
        const int8_t* wrow = mat + 25 * (int)s[i];
        // main loop (ssw.c:203-228, 428-452)
        int fin = 0, colmax = 0;
        for (int sg = 0; sg < seg; ++sg) {
            if (on) {
                const int j = lane * seg + sg;
                const int d = sg > 0 ? hprev[(sg - 1) * P + lane]
                                     : lane > 0 ? hprev[(seg - 1) * P + lane - 1] : 0;
                const int w = j < m ? (int)wrow[q(j)] : 0;
                const int x = sg * P + lane;
                int hp = max(d + w, 0);
                hp = max(hp, e[x]);
                hp = max(hp, fin);
                const int hgo = max(hp - go, 0);
                e[x] = max(max(e[x] - ge, 0), hgo);
                fin = max(max(fin - ge, 0), hgo);
                hcur[x] = hp;
                colmax = max(colmax, hp);
            }
        }
        // lazy-F (ssw.c:230-262, 454-466): the lane carries, shifted by one lane
        int carry = shift_lanes(on ? fin : 0, lane);
        if (!on) carry = 0;
        if (P == 16) {
            int j = 0;
            // each step lowers a carry by ge or raises H to it; a full sweep
            // shifts the carries a lane on, so every carry is 0 after P sweeps
            for (int step = 0, limit = 2 * P * seg + 2; step < limit; ++step) {
                const int hv = on ? hcur[j * P + lane] : 0;
                if (group_ballot(on && carry > max(hv - go, 0), grp) == 0) break;
                if (on) {
                    const int h2 = max(hv, carry);
                    hcur[j * P + lane] = h2;
                    colmax = max(colmax, h2);
                }
                carry = max(carry - ge, 0);
                if (++j >= seg) {
                    j = 0;
                    carry = shift_lanes(carry, lane);
                    if (!on) carry = 0;
                }
            }
        } else {
            bool done = false;
            for (int sweep = 0; sweep < 8 && !done; ++sweep) {
                if (sweep > 0) {
                    carry = shift_lanes(carry, lane);
                    if (!on) carry = 0;
                }
                for (int j = 0; j < seg; ++j) {
                    int h2 = 0;
                    if (on) {
                        h2 = max(hcur[j * P + lane], carry);
                        hcur[j * P + lane] = h2;
                        colmax = max(colmax, h2);
                    }
                    carry = max(carry - ge, 0);
                    if (group_ballot(on && carry > max(h2 - go, 0), grp) == 0) {
                        done = true;
                        break;
                    }
                }
            }
        }
        // column bookkeeping (ssw.c:264-297, 469-500)
        colmax = group_max(colmax);
        if (colmax >= cap) colmax = cap;
        if (colmax > best) {
            best = colmax;
            if (best >= cap) {
                // saturated: the reference looks the cap up in the column kept
                // at the last rise, whose values are all below it
                end_read = m - 1;
                break;
            }
            end_ref = i;
            // the lowest real row of this column holding the maximum
            int first = 0x7fffffff;
            if (on)
                for (int sg = 0; sg < seg; ++sg) {
                    const int j = lane * seg + sg;
                    if (j < m && hcur[sg * P + lane] == best) {
                        first = j;
                        break;
                    }
                }
            first = group_min(first);
            end_read = first == 0x7fffffff ? m - 1 : first;
        }
        if (colmax == terminate) break;
        int32_t* t = hprev;
        hprev = hcur;
        hcur = t;
        // the next column reads rows other lanes of the group wrote
        column_fence();
    }
    ScanResult r;
    r.score = best >= cap ? ovf : best;
    r.ref = end_ref;
    r.read = end_read;
    return r;
}

struct ForwardQuery {
    const int8_t* q;
    __device__ int operator()(int j) const { return q[j]; }
};

struct ReversedPrefix {
    const int8_t* q;
    int last;
    __device__ int operator()(int j) const { return q[last - j]; }
};

}  // namespace

__global__ void __launch_bounds__(64) sw_lane_kernel(LaneArgs a) {
    __shared__ int8_t mat[640];
    for (int x = threadIdx.x; x < 625; x += blockDim.x) mat[x] = a.matrix[x];
    __syncthreads();
    const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int64_t groups = (int64_t)gridDim.x * kLaneGroupsPerBlock;
    const int64_t g0 = (int64_t)blockIdx.x * kLaneGroupsPerBlock + grp;
    int32_t* ws = a.workspace + g0 * 3 * a.ws_rows;
    for (int64_t t = g0; t < a.n_list; t += groups) {
        const int p = a.pairs ? a.pairs[t] : (int)t;
        const int qi = a.pair_query[p], si = a.pair_subject[p];
        const int8_t* q = a.q_res + a.q_off[qi];
        const int8_t* s = a.s_res + a.s_off[si];
        if (!a.read_end) {
            // ssw_align: the byte pass, the word pass when it saturates
            const int m = a.q_len[qi], n = a.s_len[si];
            ScanResult r = scan(ForwardQuery{q}, m, s, n, 0, 16, 255 - a.bias, 255, -1, mat,
                                a.gap_open, a.gap_extend, ws, ws + a.ws_rows,
                                ws + 2 * a.ws_rows, lane, grp);
            if (r.score == 255)
                r = scan(ForwardQuery{q}, m, s, n, 0, 8, 32767, 32767, -1, mat, a.gap_open,
                         a.gap_extend, ws, ws + a.ws_rows, ws + 2 * a.ws_rows, lane, grp);
            if (m == 0) r = ScanResult{0, -1, -1};
            if (lane == 0) {
                a.out_score[p] = r.score;
                a.out_ref[p] = r.ref;
                a.out_read[p] = r.read;
            }
        } else {
            // ssw_align_stats' reverse pass
            const int re = a.read_end[p], fe = a.ref_end[p], sc = a.score[p];
            const bool byte = a.bias + sc < 255;
            const ScanResult r =
                scan(ReversedPrefix{q, re}, re + 1, s, fe + 1, 1, byte ? 16 : 8,
                     byte ? 255 - a.bias : 32767, byte ? 255 : 32767, sc, mat, a.gap_open,
                     a.gap_extend, ws, ws + a.ws_rows, ws + 2 * a.ws_rows, lane, grp);
            if (lane == 0) {
                a.out_ref[p] = r.ref;
                a.out_read[p] = re - r.read;
            }
        }
    }
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_sw_lane(const tpz::LaneArgs* a, int blocks, hipStream_t stream) {
    if (a->n_list <= 0) return hipSuccess;
    if (blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tpz::sw_lane_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}
