// Gapless HSP extension of suffix-array seed hits on gfx950 (TOPAZ's sa_hsp,
// search.c:99-190), batched over many seeds.
//
// A seed is (query q, query offset o, text position idx) -- idx = SA[p] when
// the seed names a suffix-array slot p.  With T the newline-separated
// database text and W the substitution matrix:
//   * if T[idx] != Q[o] the seed has length -1 and its score is left as is;
//   * right: over i = 0 .. min(Tsize-idx, m-o, 128)-1, stopping at a '\n'
//     in T, the running sum of W(T[idx+i], Q[o+i]) and the last i at which
//     it reaches its maximum (>= 0; 0 and 0 when never);
//   * left: the same over i = 1 .. min(idx, o, 128)-1 on T[idx-i], Q[o-i];
//   * score = best_left + best_right, length = len_left + len_right.
// Text residues map to matrix rows through aa_table2; query residues through
// aa_table2 of upper_aa_only (lexicographical.h:19-74), so any byte outside
// 'A'..'Y', '*', '\n' in the query scores as 'V' (row 23).
//
// Mapping: one thread per seed loads the seed (16 B, coalesced) and tests the
// first residue; about one seed in seventeen passes on uniform seeds.  The
// wavefront then takes its passing seeds four at a time: a group of 16 lanes
// per seed, each lane scoring 8 consecutive positions of the right window and
// 8 of the left, so the group reads its 128-byte text windows as whole
// segments.  Prefix sums, the first '\n' and the last argmax are 16-lane
// scans and reductions.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gapless_kernels.h"

namespace tpz {

constexpr int kWindow = 128;
constexpr int kGroup = 16;          // lanes per seed
constexpr int kPerLane = kWindow / kGroup;

// Inclusive prefix sum over the 16 lanes of a group.
__device__ __forceinline__ int group_scan_add(int v, int k) {
#pragma unroll
    for (int d = 1; d < kGroup; d <<= 1) {
        const int o = __shfl_up(v, d, kGroup);
        if (k >= d) v += o;
    }
    return v;
}
__device__ __forceinline__ int group_max(int v) {
#pragma unroll
    for (int d = kGroup / 2; d >= 1; d >>= 1) {
        const int o = __shfl_xor(v, d, kGroup);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ int group_min(int v) {
#pragma unroll
    for (int d = kGroup / 2; d >= 1; d >>= 1) {
        const int o = __shfl_xor(v, d, kGroup);
        v = o < v ? o : v;
    }
    return v;
}

// One side of the extension for the seed of this lane's group.  Positions
// i = first .. limit-1 (limit <= 128), lane k scoring i = 8k .. 8k+7; the
// text index of position i is idx + dir*i, the query index o + dir*i.
// Returns best << 8 | argmax (best >= 0), or 0 when no prefix is >= 0.
__device__ __forceinline__ int extend_side(const uint8_t* T, int64_t idx, const uint8_t* Q,
                                           int o, int dir, int first, int limit, int k,
                                           const int8_t* tmap, const int8_t* qmap,
                                           const int8_t* mat) {
    int sc[kPerLane];
    int nl = kWindow;                 // first '\n' position scored by this lane
    int sum = 0;
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int i = k * kPerLane + t;
        sc[t] = 0;
        if (i >= first && i < limit) {
            const uint8_t c = T[idx + (int64_t)dir * i];
            if (c == '\n') {
                nl = nl < i ? nl : i;
            } else {
                const uint8_t qc = Q[o + dir * i];
                sc[t] = mat[tmap[c & 127] * 25 + qmap[qc & 127]];
            }
        }
    }
    const int cut = group_min(nl);    // the scan stops at the first '\n'
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int i = k * kPerLane + t;
        if (i >= cut) sc[t] = 0;
        sum += sc[t];
    }
    const int before = group_scan_add(sum, k) - sum;   // exclusive prefix of lanes
    int key = 0;
    int run = before;
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int i = k * kPerLane + t;
        run += sc[t];
        if (i >= first && i < limit && i < cut && run >= 0) {
            const int kk = (run << 8) | i;
            key = kk > key ? kk : key;
        }
    }
    return group_max(key);
}

__global__ void __launch_bounds__(256)
gapless_kernel(GaplessArgs a) {
    __shared__ int8_t s_tmap[128], s_qmap[128], s_mat[625];
    for (int e = threadIdx.x; e < 128; e += blockDim.x) {
        s_tmap[e] = a.text_map[e];
        s_qmap[e] = a.query_map[e];
    }
    for (int e = threadIdx.x; e < 625; e += blockDim.x) s_mat[e] = a.matrix[e];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int64_t base = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63));
    for (int64_t w0 = base; w0 < a.n_seeds; w0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sid = w0 + lane;
        bool pass = false;
        int64_t idx = 0;
        int o = 0, qid = 0;
        if (sid < a.n_seeds) {
            const SeedRec s = a.seeds[sid];
            qid = s.query;
            o = s.offset;
            idx = a.sa ? a.sa[s.position] : s.position;
            const int m = a.q_len[qid];
            // positions outside the text, or query offsets outside the query,
            // are not seeds (the reference reads out of bounds there)
            if (idx >= 0 && idx < a.t_size && o >= 0 && o < m) {
                pass = a.text[idx] == a.q_chars[a.q_off[qid] + o];
            }
            if (!pass) a.out_length[sid] = -1;
        }
        uint64_t todo = __ballot(pass);
        const int g = lane / kGroup, k = lane % kGroup;
        while (todo) {
            // the g-th remaining passing lane is served by group g
            uint64_t rest = todo;
            int src = -1;
            for (int j = 0; j < 4 && rest; ++j) {
                const int b = __builtin_ctzll(rest);
                if (j == g) src = b;
                rest &= rest - 1;
            }
            const int64_t s_idx = __shfl(idx, src < 0 ? 0 : src);
            const int s_o = __shfl(o, src < 0 ? 0 : src);
            const int s_q = __shfl(qid, src < 0 ? 0 : src);
            if (src >= 0) {
                const uint8_t* Q = a.q_chars + a.q_off[s_q];
                const int m = a.q_len[s_q];
                int64_t rlim = a.t_size - s_idx;
                if ((int64_t)(m - s_o) < rlim) rlim = m - s_o;
                if (rlim > kWindow) rlim = kWindow;
                int64_t llim = s_idx < s_o ? s_idx : s_o;
                if (llim > kWindow) llim = kWindow;
                const int kr = extend_side(a.text, s_idx, Q, s_o, 1, 0, (int)rlim, k,
                                           s_tmap, s_qmap, s_mat);
                const int kl = extend_side(a.text, s_idx, Q, s_o, -1, 1, (int)llim, k,
                                           s_tmap, s_qmap, s_mat);
                if (k == 0) {
                    const int64_t out = w0 + src;
                    a.out_score[out] = (kr >> 8) + (kl >> 8);
                    a.out_length[out] = (kr & 255) + (kl & 255);
                }
            }
            // drop the (up to) four served lanes
            for (int j = 0; j < 4 && todo; ++j) todo &= todo - 1;
        }
    }
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_gapless(const tpz::GaplessArgs* args, int grid,
                                         hipStream_t stream) {
    hipLaunchKernelGGL(tpz::gapless_kernel, dim3(grid), dim3(256), 0, stream, *args);
    return hipGetLastError();
}
