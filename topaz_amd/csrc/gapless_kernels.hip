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
// aa_table2 of upper_aa_only (lexicographical.h:19-74).
//
// Both sides are the same scan over 128 positions.  The host keeps the text
// as letter codes, and the queries as codes twice, forward and reversed.  The
// right side reads T[idx..] and Q[o..]; the left side reads the same text
// window backwards (T[idx-1], T[idx-2], ...) and the reversed query copy
// forwards from Q[o-1].  Every bound of the rule is a terminator in the data:
// '\n' and the text's end in the text, the query's end in the forward query
// copy, Q[0] in the reversed copy, and T[0] in a small head window that
// replaces the text for seeds with idx < 128 (the left scan never reaches
// index 0: i < min(idx, o)).  A terminator scores -30000, after which no
// prefix sum is >= 0 again, so a scan simply runs its 128 positions; only
// the left side's 128th position (i = 128, outside i < 128) is dropped
// explicitly.  The two lanes of a seed read adjacent text, 260 bytes in all.
//
// Mapping: phase 1 gives one lane per seed: it loads the seed (16 B,
// coalesced) and tests the first residue; about one uniform seed in
// seventeen passes.  The passing seeds of the wavefront are compacted in LDS
// and phase 2 serves them 32 at a time, two lanes per seed: the even lane
// scans the right side, the odd lane the left.  A lane loads its 132-byte
// text and query windows with dword loads and aligns them with v_alignbyte;
// per position pair one v_perm forms (text code << 8 | doubled query code),
// taking the text bytes in reverse order on the left lane (whose text dwords
// are also taken last to first): each 16-bit half is the byte address of W in
// the LDS table.  The running maximum and its last position are one max over
// (sum << 8 | i).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gapless_kernels.h"

namespace tpz {

constexpr int kWaves = 4;
constexpr int kWindowDwords = kWindow / 4;
// A scan's sum is below this only after a terminator: 128 matrix scores sum
// to at least 128 x (-128) = -16384.
constexpr int kDoneBelow = -20000;

struct Slot {
    int64_t idx;
    int64_t sid;
    int32_t offset;
    int32_t query;
};

struct GaplessLds {
    int16_t table[kTableRows * kTableCols];
    Slot slot[kWaves][64];
};

// 33 dwords starting at the dword holding p, aligned so that w[k] holds the
// bytes p + 4k .. p + 4k + 3.  (Loading only the dwords a seed needs, the
// query's end being known beforehand, cut the fetched bytes by 6 % but made
// the kernel 2 % slower.)
__device__ __forceinline__ void load_window(const uint8_t* p, uint32_t (&w)[kWindowDwords]) {
    typedef const __attribute__((address_space(1))) uint32_t* global_u32;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const global_u32 d = (global_u32)(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3);
    uint32_t x[kWindowDwords + 1];
#pragma unroll
    for (int k = 0; k <= kWindowDwords; ++k) x[k] = d[k];
#pragma unroll
    for (int k = 0; k < kWindowDwords; ++k) w[k] = __builtin_amdgcn_alignbyte(x[k + 1], x[k], sh);
}

// One side: the last maximum of the prefix sums over positions 0..127 as
// (sum << 8 | position), or -1 when no prefix sum is >= 0.  t: the 128-byte
// text window, read forwards (right side) or backwards from its end (left
// side); q: the query codes in scan order.
__device__ __forceinline__ int scan_side(const uint8_t* t, const uint8_t* q, bool left,
                                         const GaplessLds& lds) {
    uint32_t tw[kWindowDwords], qw[kWindowDwords];
    load_window(t, tw);
    load_window(q, qw);
    const char* tab = reinterpret_cast<const char*>(lds.table);
    // perm selectors (bytes 0-3: query dword, 4-7: text dword) giving
    // q(2h), t, q(2h+1), t with the text bytes ascending or descending
    const uint32_t sel0 = left ? 0x06010700u : 0x05010400u;
    const uint32_t sel1 = left ? 0x04030502u : 0x07030602u;
    int sum = 0, key = -1;
#pragma unroll
    for (int k = 0; k < kWindowDwords; ++k) {
        const uint32_t td = left ? tw[kWindowDwords - 1 - k] : tw[k];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // A sum below kDoneBelow means a terminator has been passed,
            // and the at most 127 remaining scores (each at most 127) cannot
            // bring it back to zero, so no later position can raise the
            // maximum: the lane stops looking up scores.  Such lookups were
            // the main source of LDS bank conflicts: every lane past its
            // query's end reads the terminator column, one bank, in a
            // different row.
            if (sum > kDoneBelow) {
                const uint32_t x = __builtin_amdgcn_perm(td, qw[k], h ? sel1 : sel0);
                const int p = 4 * k + 2 * h;
                int w0 = *reinterpret_cast<const int16_t*>(tab + (x & 0xffffu));
                int w1 = *reinterpret_cast<const int16_t*>(tab + (x >> 16));
                if (p + 1 == kWindow - 1 && left) w1 = kTermScore;
                sum += w0;
                key = max(key, (int)(((uint32_t)sum << 8) | (uint32_t)p));
                sum += w1;
                key = max(key, (int)(((uint32_t)sum << 8) | (uint32_t)(p + 1)));
            }
        }
    }
    return key;
}

// Phase 2 of both kernels: the wavefront's n_pass compacted seeds (first
// residues equal), 32 at a time, two lanes per seed: the even lane scans the
// right side, the odd lane the left; write(sid, score, length) stores a
// seed's result.
template <typename Write>
__device__ __forceinline__ void extend_passing(const GaplessArgs& a, const GaplessLds& lds,
                                               const Slot* slots, int n_pass, int lane,
                                               Write&& write) {
    for (int r = 0; r < n_pass; r += 32) {
        const int j = r + lane / 2;
        const bool active = j < n_pass;
        const bool left = lane & 1;
        int key = -1;
        Slot s{};
        if (active) {
            s = slots[j];
            const int m = a.q_len[s.query];
            // left: the 128 text bytes before idx (the head window when
            // idx < 128), scanned from the end
            const uint8_t* t = !left ? a.t_fwd + s.idx
                              : s.idx >= kWindow ? a.t_fwd + (s.idx - kWindow)
                                                 : a.t_head + s.idx;
            const uint8_t* q = left ? a.q_rev + a.qr_off[s.query] + (m - s.offset)
                                    : a.q_fwd + a.qf_off[s.query] + s.offset;
            key = scan_side(t, q, left, lds);
        }
        const int other = __shfl_xor(key, 1);
        if (active && !left) {
            // right: i = position; left: i = position + 1; -1: (0, 0)
            const int sr = key < 0 ? 0 : key >> 8, lr = key < 0 ? 0 : key & 255;
            const int sl = other < 0 ? 0 : other >> 8, ll = other < 0 ? 0 : (other & 255) + 1;
            write(s.sid, sr + sl, lr + ll);
        }
    }
}

__global__ void __launch_bounds__(kWaves * 64)
gapless_kernel(GaplessArgs a) {
    __shared__ GaplessLds lds;
    for (int e = threadIdx.x; e < kTableRows * kTableCols; e += blockDim.x) lds.table[e] = a.table[e];
    __syncthreads();

    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    Slot* slots = lds.slot[wave];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t w0 = (int64_t)blockIdx.x * blockDim.x + wave * 64; w0 < a.n_seeds; w0 += stride) {
        // phase 1: one lane per seed
        const int64_t sid = w0 + lane;
        bool pass = false;
        int64_t idx = 0;
        int o = 0, qid = 0;
        if (sid < a.n_seeds) {
            const SeedRec s = a.seeds[sid];
            qid = s.query;
            o = s.offset;
            // slots or positions outside the text, and queries or query
            // offsets that do not exist, are not seeds (the reference reads
            // out of bounds there)
            const bool in_sa = s.position >= 0 && s.position < a.t_size;
            idx = a.sa ? (in_sa ? a.sa[s.position] : -1) : s.position;
            if (idx >= 0 && idx < a.t_size && qid >= 0 && qid < a.n_queries && o >= 0 &&
                o < a.q_len[qid])
                pass = 2 * (int)a.t_fwd[idx] == (int)a.q_fwd[a.qf_off[qid] + o];
            if (!pass) {
                a.out_length[sid] = -1;
                a.out_score[sid] = 0;
            }
        }
        const uint64_t mask = __ballot(pass);
        const int n_pass = __popcll(mask);
        if (pass) {
            const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
            slots[rank] = Slot{idx, sid, o, qid};
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // phase 2: two lanes per passing seed, even = right side, odd = left
        extend_passing(a, lds, slots, n_pass, lane, [&](int64_t sid, int score, int length) {
            a.out_score[sid] = score;
            a.out_length[sid] = length;
        });
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One round of walk-stream extension.  A unit is up to `width` consecutive
// slots of one request: with a chunk wider than 32 slots width is 64 and a
// wavefront iteration serves one unit (request u / spr, part u % spr, spr =
// the parts of a chunk); with a narrower chunk width is the chunk rounded up
// to a power of two and a wavefront serves 64 / width units at once, so the
// first rounds (32 slots per stream) keep every lane busy.  Lane l takes
// slot l % width of its unit: first + (have + part * 64 + l % width) * dir;
// then the same compaction and two-lane scans as gapless_kernel.  Each
// lane's slot is marked in LDS when its element is -1; each unit's first
// lane lowers its request's req_end to the unit's first marked slot.
__global__ void __launch_bounds__(kWaves * 64)
stream_round_kernel(StreamRoundArgs sa) {
    __shared__ GaplessLds lds;
    __shared__ uint8_t is_end[kWaves][64];
    const GaplessArgs& a = sa.g;
    for (int e = threadIdx.x; e < kTableRows * kTableCols; e += blockDim.x) lds.table[e] = a.table[e];
    __syncthreads();

    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    Slot* slots = lds.slot[wave];
    uint8_t* ends = is_end[wave];
    int width = 64;
    if (sa.chunk <= 32)
        for (width = 1; width < sa.chunk; width <<= 1) {}
    const int units_per_wave = 64 / width;
    const int lane_in = lane & (width - 1);
    const uint64_t unit_mask = width == 64 ? ~0ull : (1ull << width) - 1;
    const int64_t spr = width == 64 ? (sa.chunk + 63) / 64 : 1;
    const int64_t n_units = sa.n_req * spr;
    const int64_t n_iter = (n_units + units_per_wave - 1) / units_per_wave;
    const int64_t waves_total = (int64_t)gridDim.x * kWaves;
    for (int64_t g = (int64_t)blockIdx.x * kWaves + wave; g < n_iter; g += waves_total) {
        const int64_t unit = g * units_per_wave + lane / width;
        const bool live = unit < n_units;
        int64_t req = 0, out0 = 0;
        int32_t part = 0, count = 0;
        StreamDef d{};
        if (live) {
            req = unit / spr;
            part = (int32_t)(unit - req * spr);
            d = sa.defs[sa.req_stream[req]];
            count = min(width, sa.chunk - part * 64);
            out0 = req * sa.chunk + (int64_t)part * 64;
        }
        bool pass = false;
        int64_t idx = 0;
        if (lane_in < count) {
            const int64_t slot = d.first + (sa.have + (int64_t)part * 64 + lane_in) * d.dir;
            uint32_t prot = 0xffffffffu;
            if (slot >= 0 && slot < a.t_size) {
                idx = sa.sa32 ? (int64_t)sa.sa32[slot] : a.sa[slot];
                if (idx >= 0 && idx < a.t_size) {
                    // the protein holding idx: the one holding idx & ~63 or a
                    // later one (the last protein's newline is the text's end)
                    uint32_t k = sa.prot_block[idx >> 6];
                    while ((uint64_t)idx > sa.prot_end[k]) ++k;
                    prot = k;
                    if (d.query >= 0 && d.query < a.n_queries && d.offset >= 0 &&
                        d.offset < a.q_len[d.query])
                        pass = 2 * (int)a.t_fwd[idx] == (int)a.q_fwd[a.qf_off[d.query] + d.offset];
                }
            }
            sa.out_protein[out0 + lane_in] = prot;
            if (!pass) sa.out[out0 + lane_in] = -1;
        }
        ends[lane] = !pass;
        const uint64_t mask = __ballot(pass);
        const int n_pass = __popcll(mask);
        if (pass) {
            const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
            // the output element, with the lane that owns it in the low 6 bits
            slots[rank] = Slot{idx, ((out0 + lane_in) << 6) | lane, d.offset, d.query};
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int min_len = sa.min_len;
        extend_passing(a, lds, slots, n_pass, lane, [&](int64_t s_id, int score, int length) {
            const bool seed = length >= min_len;
            sa.out[s_id >> 6] = (int16_t)(seed ? score : -1);
            ends[s_id & 63] = !seed;
        });
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint64_t end_mask = __ballot(lane_in < count && ends[lane]);
        if (live && lane_in == 0) {
            const uint64_t mine = (end_mask >> lane) & unit_mask;
            if (mine) atomicMin(sa.req_end + req, part * 64 + (int32_t)__builtin_ctzll(mine));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One thread per (query, offset) record: binary search over the suffix
// array for the first suffix not below the query's suffix P (bytes compared
// as the reference's signed chars; a suffix that ends inside P sorts below
// it, one that P is a prefix of does not).  The bytes P shares with the
// suffixes on both sides of the interval are skipped (the smaller of the two
// common prefixes holds for every suffix between them).
__global__ void __launch_bounds__(256) lower_bound_kernel(LowerBoundArgs a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int32_t q = a.rec_query[k], o = a.rec_offset[k];
    const int64_t m = (int64_t)a.q_len[q] - o;
    if (m <= 0 || o < 0 || a.t_size <= 0) {
        a.out[k] = -1;
        return;
    }
    const signed char* T = reinterpret_cast<const signed char*>(a.text);
    const signed char* P = reinterpret_cast<const signed char*>(a.q_chars + a.q_off[q] + o);
    const int64_t n = a.t_size;
    int64_t lo = 0, hi = n, ml = 0, mr = 0;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        int64_t j = ml < mr ? ml : mr;
        int64_t i = (a.sa32 ? (int64_t)a.sa32[mid] : a.sa[mid]) + j;
        int r = 0;
        while (i < n && j < m) {
            r = (int)T[i] - (int)P[j];
            if (r != 0) break;
            ++i;
            ++j;
        }
        const bool below = r != 0 ? r < 0 : j < m;
        if (below) {
            lo = mid + 1;
            ml = j;
        } else {
            hi = mid;
            mr = j;
        }
    }
    a.out[k] = lo;
}

}  // namespace tpz

namespace tpz {

// search.c:17-30: compare the suffix at suf with P from *match on; bytes as
// the reference's (signed) char differences.
__device__ __forceinline__ int sa_compare(const char* T, int64_t Tsize, const char* P,
                                          int64_t Psize, int64_t suf, int64_t* match) {
    int64_t i = suf + *match, j = *match;
    int r = 0;
    for (; i < Tsize && j < Psize && (r = (int)(signed char)T[i] - (int)(signed char)P[j]) == 0;
         ++i, ++j) {}
    *match = j;
    return r == 0 ? -(j != Psize) : r;
}

// One thread per query offset, the reference's binary search step for step
// (search.c:266-344 without USE_LCP): which SA entries it compares, and the
// match lengths it carries, decide the reported index and last_position
// when the suffix does not occur.
__global__ void __launch_bounds__(64) sa_search_kernel(SaSearchArgs a) {
    extern __shared__ char qs[];
    const bool in_lds = a.q_len <= kSaSearchQueryLds;
    if (in_lds) {
        for (int64_t x = threadIdx.x; x < a.q_len; x += blockDim.x) qs[x] = a.query[x];
        __syncthreads();
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const char* T = a.text;
    const int64_t* SA = a.sa;
    const int64_t Tsize = a.t_size;
    const int64_t offset = a.offsets[k];
    const char* P = (in_lds ? qs : a.query) + offset;
    const int64_t Psize = a.q_len - offset;
    int64_t* out = a.out + 3 * (int64_t)k;
    out[0] = -1;
    if (T == nullptr || SA == nullptr || Tsize <= 0 || Psize <= 0) {
        a.status[k] = -1;
        return;
    }
    int64_t size, lsize, rsize, half;
    int64_t match, lmatch, rmatch, llmatch, lrmatch, rlmatch, rrmatch;
    int64_t i, j, kk;
    int r;
    for (i = j = kk = 0, match = lmatch = rmatch = 0, size = Tsize, half = size >> 1; 0 < size;
         size = half, half >>= 1) {
        match = lmatch < rmatch ? lmatch : rmatch;
        r = sa_compare(T, Tsize, P, Psize, SA[i + half], &match);
        if (r < 0) {
            i += half + 1;
            half -= (size & 1) ^ 1;
            lmatch = match;
        } else if (r > 0) {
            rmatch = match;
        } else {
            lsize = half, j = i, rsize = size - half - 1, kk = i + half + 1;
            for (llmatch = lmatch, lrmatch = match, half = lsize >> 1; 0 < lsize;
                 lsize = half, half >>= 1) {
                lmatch = llmatch < lrmatch ? llmatch : lrmatch;
                r = sa_compare(T, Tsize, P, match, SA[j + half], &lmatch);
                if (r < 0) {
                    j += half + 1;
                    half -= (lsize & 1) ^ 1;
                    llmatch = lmatch;
                } else {
                    lrmatch = lmatch;
                }
            }
            for (rlmatch = match, rrmatch = rmatch, half = rsize >> 1; 0 < rsize;
                 rsize = half, half >>= 1) {
                rmatch = rlmatch < rrmatch ? rlmatch : rrmatch;
                r = sa_compare(T, Tsize, P, match, SA[kk + half], &rmatch);
                if (r <= 0) {
                    kk += half + 1;
                    half -= (rsize & 1) ^ 1;
                    rlmatch = rmatch;
                } else {
                    rrmatch = rmatch;
                }
            }
            break;
        }
    }
    out[0] = (0 < (kk - j)) ? j : i;
    out[1] = kk - j;
    out[2] = match;
    a.status[k] = 0;
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_sa_search(const tpz::SaSearchArgs* args, hipStream_t stream) {
    if (args->n <= 0) return hipSuccess;
    const size_t lds = args->q_len <= tpz::kSaSearchQueryLds ? (size_t)args->q_len + 1 : 0;
    hipLaunchKernelGGL(tpz::sa_search_kernel, dim3((args->n + 63) / 64), dim3(64), lds, stream,
                       *args);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_gapless(const tpz::GaplessArgs* args, int grid,
                                         hipStream_t stream) {
    hipLaunchKernelGGL(tpz::gapless_kernel, dim3(grid), dim3(tpz::kWaves * 64), 0, stream, *args);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_lower_bound(const tpz::LowerBoundArgs* args, hipStream_t stream) {
    if (args->n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tpz::lower_bound_kernel, dim3((unsigned)((args->n + 255) / 256)), dim3(256),
                       0, stream, *args);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_stream_round(const tpz::StreamRoundArgs* args, int grid,
                                              hipStream_t stream) {
    if (args->n_req <= 0) return hipSuccess;
    if (args->chunk < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tpz::stream_round_kernel, dim3(grid), dim3(tpz::kWaves * 64), 0, stream,
                       *args);
    return hipGetLastError();
}
