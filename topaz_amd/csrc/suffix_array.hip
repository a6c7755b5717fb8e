// Suffix array of TOPAZ's text on gfx950, for `topaz index`: the order
// divsufsort64 gives (index.c:66), i.e. the suffixes of the text as
// unsigned bytes, a suffix before every longer suffix it is a prefix of.
// The suffix array of a string is unique, so any correct construction
// gives the reference's .SA byte for byte.
//
// Prefix doubling: rank[i] orders the suffixes by their first h bytes
// (bytes count as byte + 1, 0 past the end).  The first round ranks the
// two-byte prefixes from a bitmap of the byte pairs present, and in a text
// of few distinct byte pairs (protein, DNA) the second round ranks the
// four-byte prefixes the same way.  Each later round sorts the positions
// by the pair (rank[i], rank[i + h]) with a device radix sort, numbers the
// distinct pairs in sorted order with a scan, and doubles h, until every
// rank is distinct.  Each round is a few
// passes over 16-byte records in HBM; a round's sort covers only the bits
// the current ranks need (2 x bit width of the largest rank).  Rounds:
// about log2 of the longest repeated substring, plus one.
//
// Once at most a quarter of the suffixes share their group with another
// (n - groups <= n / 8), the ranks become slot numbers: a group's rank is
// 1 + the index of its first suffix in the sorted order, which is where
// the group's suffixes sit in the final suffix array.  The later rounds
// then sort only the suffixes of groups of two or more, by (rank[i],
// rank[i + h]), and split each group inside its own slots.  In a protein
// database that is the repeated sequences, a few percent of the text.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>

#include "../../include/topaz_amd.h"
#include "../../include/topaz_index_compat.h"
#include "runtime.h"

namespace tpz {
namespace {

// The first round needs no sort: the rank of a suffix by its first two
// bytes is the number of distinct byte pairs (T[i] + 1, T[i + 1] + 1, 0
// past the end) below its own, plus one.  A bitmap of the 257 x 257 pair
// codes marks the pairs present; word prefix counts of the bitmap then
// number them.
constexpr int kPairCodes = 257 * 257;
constexpr int kPairWords = (kPairCodes + 31) / 32;

__device__ inline uint32_t sa_pair_code(const uint8_t* T, int64_t n, int64_t i) {
    return ((uint32_t)T[i] + 1) * 257 + (i + 1 < n ? (uint32_t)T[i + 1] + 1 : 0);
}

__global__ void sa_pair_bits(const uint8_t* __restrict__ T, int64_t n, uint32_t* __restrict__ bits) {
    __shared__ uint32_t b[kPairWords];
    for (int w = threadIdx.x; w < kPairWords; w += blockDim.x) b[w] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = sa_pair_code(T, n, i);
        atomicOr(&b[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    for (int w = threadIdx.x; w < kPairWords; w += blockDim.x)
        if (b[w]) atomicOr(&bits[w], b[w]);
}

// prefix[w] = codes present in words 0..w-1; prefix[words] = all
__global__ void sa_bit_count(const uint32_t* __restrict__ bits, int words,
                             uint32_t* __restrict__ prefix) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t sum = 0;
    for (int w = 0; w < words; ++w) {
        prefix[w] = sum;
        sum += __popc(bits[w]);
    }
    prefix[words] = sum;
}

// Round 2 the same way when the text has few distinct byte pairs (p <=
// 511: protein and DNA text): the rank by four bytes is 1 + the number of
// distinct codes (rank[i] - 1) x (p + 1) + rank[i + 2] (0 past the end)
// below the suffix's own.  The bitmap of the p (p + 1) codes fits in
// 32 KB of LDS.
constexpr int kQuadWords = 8192;
constexpr uint32_t kQuadMaxPairs = 511;   // 511 x 512 codes <= 32 x kQuadWords

__device__ inline uint32_t sa_quad_code(const uint32_t* rank, int64_t n, int64_t i, uint32_t p) {
    return (rank[i] - 1) * (p + 1) + (i + 2 < n ? rank[i + 2] : 0);
}

__global__ void sa_quad_bits(const uint32_t* __restrict__ rank, int64_t n, uint32_t p,
                             uint32_t* __restrict__ bits) {
    __shared__ uint32_t b[kQuadWords];
    for (int w = threadIdx.x; w < kQuadWords; w += blockDim.x) b[w] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = sa_quad_code(rank, n, i, p);
        atomicOr(&b[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    for (int w = threadIdx.x; w < kQuadWords; w += blockDim.x)
        if (b[w]) atomicOr(&bits[w], b[w]);
}

__global__ void sa_quad_rank(const uint32_t* __restrict__ rank, int64_t n, uint32_t p,
                             const uint32_t* __restrict__ bits, const uint32_t* __restrict__ prefix,
                             uint32_t* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = sa_quad_code(rank, n, i, p);
        out[i] = prefix[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1)) + 1;
    }
}

__global__ void sa_pair_rank(const uint8_t* __restrict__ T, int64_t n,
                             const uint32_t* __restrict__ bits, const uint32_t* __restrict__ prefix,
                             uint32_t* __restrict__ rank) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = sa_pair_code(T, n, i);
        rank[i] = prefix[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1)) + 1;
    }
}

// the pair (rank[i], rank[i + h]) packed in 2 * bits bits, and position i
__global__ void sa_pairs(const uint32_t* __restrict__ rank, int64_t n, int64_t h, int bits,
                         uint64_t* __restrict__ key, uint32_t* __restrict__ pos) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t second = i + h < n ? rank[i + h] : 0;
        key[i] = ((uint64_t)rank[i] << bits) | second;
        pos[i] = (uint32_t)i;
    }
}

// 1 where a sorted pair differs from the one before
__global__ void sa_heads(const uint64_t* __restrict__ key, int64_t n, uint32_t* __restrict__ head) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        head[k] = (k == 0 || key[k] != key[k - 1]) ? 1u : 0u;
}

__global__ void sa_rerank(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ group,
                          int64_t n, uint32_t* __restrict__ rank) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        rank[pos[k]] = group[k];
}

// Slot ranks over the whole sorted array: mark[k] = k where the sorted
// pair differs from the one before, else 0; a max-scan then gives the
// first index of each group.
__global__ void sa_marks_whole(const uint64_t* __restrict__ key, int64_t n,
                               uint32_t* __restrict__ mark) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        mark[k] = (k == 0 || key[k] != key[k - 1]) ? (uint32_t)k : 0u;
}

// 1 for a suffix whose sorted pair equals a neighbour's (its group has two
// or more suffixes)
__device__ inline uint8_t sa_shared(const uint64_t* key, int64_t k, int64_t m) {
    const bool first = k == 0 || key[k] != key[k - 1];
    const bool last = k + 1 == m || key[k + 1] != key[k];
    return (first && last) ? 0 : 1;
}

__global__ void sa_place_whole(const uint64_t* __restrict__ key, const uint32_t* __restrict__ pos,
                               const uint32_t* __restrict__ start, int64_t n,
                               uint32_t* __restrict__ rank, uint8_t* __restrict__ shared) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        rank[pos[k]] = start[k] + 1;
        shared[k] = sa_shared(key, k, n);
    }
}

// pairs of the unresolved suffixes list[0..m): ranks are slot numbers of
// at most `bits` bits
__global__ void sa_list_pairs(const uint32_t* __restrict__ list, int64_t m,
                              const uint32_t* __restrict__ rank, int64_t n, int64_t h, int bits,
                              uint64_t* __restrict__ key, uint32_t* __restrict__ pos) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = list[j];
        const uint64_t second = i + h < n ? rank[i + h] : 0;
        key[j] = ((uint64_t)rank[i] << bits) | second;
        pos[j] = (uint32_t)i;
    }
}

// Two marks per sorted entry j, packed high | low: j where the group rank
// (key >> bits) changes, j where the whole pair changes; 0 elsewhere.
__global__ void sa_marks(const uint64_t* __restrict__ key, int64_t m, int bits,
                         uint64_t* __restrict__ mark) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const bool group = j == 0 || (key[j] >> bits) != (key[j - 1] >> bits);
        const bool pair = j == 0 || key[j] != key[j - 1];
        mark[j] = ((group ? (uint64_t)j : 0) << 32) | (pair ? (uint64_t)j : 0);
    }
}

// componentwise max of the two 32-bit halves
struct HalvesMax {
    __host__ __device__ uint64_t operator()(uint64_t a, uint64_t b) const {
        const uint64_t hi = (a >> 32) > (b >> 32) ? (a >> 32) : (b >> 32);
        const uint64_t lo = (uint32_t)a > (uint32_t)b ? (uint32_t)a : (uint32_t)b;
        return (hi << 32) | lo;
    }
};

// Entry j of a group whose slots begin at g = rank - 1 and whose entries
// begin at r in the sorted list goes to slot g + j - r; its new rank is
// 1 + the slot of the first entry with the same pair.
__global__ void sa_place(const uint64_t* __restrict__ key, const uint32_t* __restrict__ pos,
                         const uint64_t* __restrict__ start, int64_t m, int bits,
                         uint32_t* __restrict__ rank, uint32_t* __restrict__ sa,
                         uint8_t* __restrict__ shared) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t g = (key[j] >> bits) - 1;
        const uint64_t r = start[j] >> 32, first = (uint32_t)start[j];
        rank[pos[j]] = (uint32_t)(g + (first - r) + 1);
        sa[g + (j - r)] = pos[j];
        shared[j] = sa_shared(key, j, m);
    }
}

__global__ void sa_widen(const uint32_t* __restrict__ pos, int64_t n, int64_t* __restrict__ sa) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        sa[k] = pos[k];
}

int bit_width(uint64_t v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

struct SaBuffers {
    DeviceBuffer text, rank, key0, key1, pos0, pos1, head, tmp, count, pairs;
};

double seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Download of a large suffix array: the 32-bit positions cross PCIe into
// pinned staging buffers (half the bytes of the 64-bit entries), and host
// threads widen them into the caller's array.  Each thread owns a stream,
// two 4 MB halves of the staging memory and a contiguous range of the
// array, so the copy of one half overlaps the widening of the other.  The
// staging memory, streams and events are made on the first large call and
// kept until the process ends; calls from several threads take turns.
constexpr int kReadThreads = 8;
constexpr int64_t kReadChunk = int64_t(1) << 20;          // entries per half
constexpr int64_t kReadStaged = int64_t(4) << 20;         // from this many entries on

struct Staging {
    std::mutex mu;
    uint32_t* host = nullptr;
    hipStream_t st[kReadThreads] = {};
    hipEvent_t ev[kReadThreads][2] = {};
};

Staging& staging() {
    static Staging* s = new Staging;   // never destroyed: its memory lives with the process
    return *s;
}

int staging_init(Staging& s) {
    if (s.host) return 0;
    void* p = nullptr;
    if (check(hipHostMalloc(&p, kReadThreads * 2 * kReadChunk * 4, hipHostMallocDefault),
              "hipHostMalloc"))
        return -1;
    for (int t = 0; t < kReadThreads; ++t)
        if (check(hipStreamCreateWithFlags(&s.st[t], hipStreamNonBlocking), "hipStreamCreate") ||
            check(hipEventCreateWithFlags(&s.ev[t][0], hipEventDisableTiming), "hipEventCreate") ||
            check(hipEventCreateWithFlags(&s.ev[t][1], hipEventDisableTiming), "hipEventCreate"))
            return -1;
    s.host = (uint32_t*)p;
    return 0;
}

hipError_t read_range(Staging& s, int t, const uint32_t* dev, int64_t lo, int64_t hi,
                      int64_t* sa) {
    uint32_t* half[2] = {s.host + (2 * t) * kReadChunk, s.host + (2 * t + 1) * kReadChunk};
    const int64_t chunks = (hi - lo + kReadChunk - 1) / kReadChunk;
    auto issue = [&](int64_t c) {
        const int64_t a = lo + c * kReadChunk, m = std::min(kReadChunk, hi - a);
        hipError_t e = hipMemcpyAsync(half[c & 1], dev + a, m * 4, hipMemcpyDeviceToHost, s.st[t]);
        return e == hipSuccess ? hipEventRecord(s.ev[t][c & 1], s.st[t]) : e;
    };
    // on an error, let the copies already issued land before the staging
    // memory is used again
    auto drain = [&](hipError_t e) {
        (void)hipStreamSynchronize(s.st[t]);
        return e;
    };
    for (int64_t c = 0; c < std::min<int64_t>(2, chunks); ++c)
        if (hipError_t e = issue(c)) return drain(e);
    for (int64_t c = 0; c < chunks; ++c) {
        if (hipError_t e = hipEventSynchronize(s.ev[t][c & 1])) return drain(e);
        const int64_t a = lo + c * kReadChunk, m = std::min(kReadChunk, hi - a);
        const uint32_t* src = half[c & 1];
        int64_t* dst = sa + a;
        for (int64_t i = 0; i < m; ++i) dst[i] = src[i];
        if (c + 2 < chunks)
            if (hipError_t e = issue(c + 2)) return drain(e);
    }
    return hipSuccess;
}

// sa[0..n) = pos[0..n) widened; pos is complete on the device
int sa_download(const uint32_t* pos, int64_t n, int64_t* sa) {
    Staging& s = staging();
    std::lock_guard<std::mutex> lk(s.mu);
    if (staging_init(s)) return -1;
    hipError_t err[kReadThreads] = {};
    std::thread th[kReadThreads];
    for (int t = 0; t < kReadThreads; ++t)
        th[t] = std::thread([&, t] {
            err[t] = read_range(s, t, pos, n * t / kReadThreads, n * (t + 1) / kReadThreads, sa);
        });
    for (auto& x : th) x.join();
    for (int t = 0; t < kReadThreads; ++t)
        if (check(err[t], "suffix array download")) return -1;
    return 0;
}

}  // namespace
}  // namespace tpz

using namespace tpz;

extern "C" {

int tpz_suffix_array(const uint8_t* text, int64_t n, int64_t* sa, tpz_sa_stats* stats) {
    if (stats) *stats = tpz_sa_stats{};
    if (n < 0 || (n > 0 && (!text || !sa))) return fail("tpz_suffix_array: bad arguments");
    if (n == 0) return 0;
    if (n > (int64_t)UINT32_MAX - 1)
        return fail("tpz_suffix_array: %lld bytes; positions and ranks are 32-bit (at most "
                    "4,294,967,294 bytes)", (long long)n);
    if (ensure_init()) return -1;
    const double t0 = seconds();
    hipStream_t st = stream();
    SaBuffers b;
    const size_t N = (size_t)n;
    if (b.text.upload(text, N) || b.rank.alloc(N * 4) || b.key0.alloc(N * 8) ||
        b.key1.alloc(N * 8) || b.pos0.alloc(N * 4) || b.pos1.alloc(N * 4) ||
        b.head.alloc(N * 4) || b.count.alloc(8) || b.pairs.alloc((2 * kQuadWords + 1) * 4))
        return -1;
    // scratch of the sorts, scans and selection, the largest of them
    size_t sort_bytes = 0, scan_bytes = 0, max_bytes = 0, pair_bytes = 0, select_bytes = 0;
    {
        rocprim::double_buffer<uint64_t> kb((uint64_t*)b.key0.ptr, (uint64_t*)b.key1.ptr);
        rocprim::double_buffer<uint32_t> vb((uint32_t*)b.pos0.ptr, (uint32_t*)b.pos1.ptr);
        uint32_t* u = (uint32_t*)b.head.ptr;
        uint64_t* w = (uint64_t*)b.key1.ptr;
        if (check(rocprim::radix_sort_pairs(nullptr, sort_bytes, kb, vb, N, 0, 64, st),
                  "radix sort size") ||
            check(rocprim::inclusive_scan(nullptr, scan_bytes, u, u, N, rocprim::plus<uint32_t>(),
                                          st), "scan size") ||
            check(rocprim::inclusive_scan(nullptr, max_bytes, u, u, N,
                                          rocprim::maximum<uint32_t>(), st), "scan size") ||
            check(rocprim::inclusive_scan(nullptr, pair_bytes, w, w, N / 4 + 1, HalvesMax(), st),
                  "scan size") ||
            check(rocprim::select(nullptr, select_bytes, u, (const uint8_t*)w, u,
                                  (uint32_t*)b.count.ptr, N, st), "select size"))
            return -1;
    }
    const size_t tmp_bytes = std::max({sort_bytes, scan_bytes, max_bytes, pair_bytes, select_bytes});
    if (b.tmp.alloc(tmp_bytes + 256)) return -1;
    const int block = 256;
    auto grid_for = [&](int64_t m) {
        return (int)std::max<int64_t>(1, std::min<int64_t>((m + block - 1) / block,
                                                           (int64_t)num_cu() * 64));
    };
    const int grid = grid_for(n);
    if (check(hipStreamSynchronize(st), "upload")) return -1;
    const double t1 = seconds();
    // round 1: ranks by the first two bytes
    uint32_t* pair_bits = (uint32_t*)b.pairs.ptr;
    uint32_t* pair_prefix = pair_bits + kPairWords;
    uint32_t pairs = 0;
    if (check(hipMemsetAsync(pair_bits, 0, kPairWords * 4, st), "clear"))
        return -1;
    hipLaunchKernelGGL(sa_pair_bits, dim3(std::min(grid, num_cu() * 4)), dim3(block), 0, st,
                       (const uint8_t*)b.text.ptr, n, pair_bits);
    hipLaunchKernelGGL(sa_bit_count, dim3(1), dim3(64), 0, st, (const uint32_t*)pair_bits,
                       kPairWords, pair_prefix);
    hipLaunchKernelGGL(sa_pair_rank, dim3(grid), dim3(block), 0, st, (const uint8_t*)b.text.ptr,
                       n, (const uint32_t*)pair_bits, (const uint32_t*)pair_prefix,
                       (uint32_t*)b.rank.ptr);
    if (check(hipGetLastError(), "pair ranks") ||
        check(hipMemcpyAsync(&pairs, pair_prefix + kPairWords, 4, hipMemcpyDeviceToHost, st),
              "readback") ||
        check(hipStreamSynchronize(st), "pair ranks"))
        return -1;
    uint64_t max_rank = pairs;
    int rounds = 1;
    int64_t h = 2;
    if (pairs <= kQuadMaxPairs && (int64_t)pairs < n) {
        // round 2: ranks by four bytes, into the head buffer, which then
        // becomes the rank buffer
        uint32_t* quad_bits = (uint32_t*)b.pairs.ptr;
        uint32_t* quad_prefix = quad_bits + kQuadWords;
        uint32_t* out = (uint32_t*)b.head.ptr;
        uint32_t quads = 0;
        if (check(hipMemsetAsync(quad_bits, 0, kQuadWords * 4, st), "clear")) return -1;
        hipLaunchKernelGGL(sa_quad_bits, dim3(std::min(grid, num_cu() * 2)), dim3(block), 0, st,
                           (const uint32_t*)b.rank.ptr, n, pairs, quad_bits);
        hipLaunchKernelGGL(sa_bit_count, dim3(1), dim3(64), 0, st, (const uint32_t*)quad_bits,
                           kQuadWords, quad_prefix);
        hipLaunchKernelGGL(sa_quad_rank, dim3(grid), dim3(block), 0, st,
                           (const uint32_t*)b.rank.ptr, n, pairs, (const uint32_t*)quad_bits,
                           (const uint32_t*)quad_prefix, out);
        if (check(hipGetLastError(), "four-byte ranks") ||
            check(hipMemcpyAsync(&quads, quad_prefix + kQuadWords, 4, hipMemcpyDeviceToHost,
                                 st), "readback") ||
            check(hipStreamSynchronize(st), "four-byte ranks"))
            return -1;
        std::swap(b.rank.ptr, b.head.ptr);
        max_rank = quads;
        rounds = 2;
        h = 4;
    }
    uint32_t* pos_sorted = nullptr;
    bool compact = false;
    rocprim::double_buffer<uint64_t> kb((uint64_t*)b.key0.ptr, (uint64_t*)b.key1.ptr);
    rocprim::double_buffer<uint32_t> vb((uint32_t*)b.pos0.ptr, (uint32_t*)b.pos1.ptr);
    for (;; h *= 2) {
        const int bits = bit_width(max_rank);
        kb = rocprim::double_buffer<uint64_t>((uint64_t*)b.key0.ptr, (uint64_t*)b.key1.ptr);
        vb = rocprim::double_buffer<uint32_t>((uint32_t*)b.pos0.ptr, (uint32_t*)b.pos1.ptr);
        hipLaunchKernelGGL(sa_pairs, dim3(grid), dim3(block), 0, st, (const uint32_t*)b.rank.ptr,
                           n, h, bits, kb.current(), vb.current());
        size_t tb = tmp_bytes;
        if (check(hipGetLastError(), "sa_pairs") ||
            check(rocprim::radix_sort_pairs(b.tmp.ptr, tb, kb, vb, N, 0, 2 * bits, st),
                  "radix sort"))
            return -1;
        hipLaunchKernelGGL(sa_heads, dim3(grid), dim3(block), 0, st,
                           (const uint64_t*)kb.current(), n, (uint32_t*)b.head.ptr);
        // group numbers (1-based, in sorted order) into the free key buffer
        uint32_t* group = reinterpret_cast<uint32_t*>(kb.alternate());
        size_t sb = tmp_bytes;
        uint32_t groups = 0;
        if (check(hipGetLastError(), "sa_heads") ||
            check(rocprim::inclusive_scan(b.tmp.ptr, sb, (uint32_t*)b.head.ptr, group, N,
                                          rocprim::plus<uint32_t>(), st),
                  "scan") ||
            check(hipMemcpyAsync(&groups, group + (n - 1), 4, hipMemcpyDeviceToHost, st),
                  "readback") ||
            check(hipStreamSynchronize(st), "suffix array round"))
            return -1;
        ++rounds;
        pos_sorted = vb.current();
        if ((int64_t)groups == n || h >= n) break;
        if ((n - (int64_t)groups) * 8 <= n) {
            compact = true;
            break;
        }
        hipLaunchKernelGGL(sa_rerank, dim3(grid), dim3(block), 0, st,
                           (const uint32_t*)vb.current(), (const uint32_t*)group, n,
                           (uint32_t*)b.rank.ptr);
        if (check(hipGetLastError(), "sa_rerank")) return -1;
        max_rank = groups;
    }
    if (compact) {
        // Slot ranks for all suffixes, and the list of those still sharing a
        // group (at most 2 (n - groups) <= n / 4 of them).  Buffers: the
        // suffix array in slot order stays in pos_sorted; the sorted keys'
        // buffer holds the list rounds' keys (16 m bytes), marks and starts
        // (8 m each); the other key buffer holds the flags (n bytes) and the
        // list (4 m bytes from byte n rounded up to 16); the other position
        // buffer and the head buffer hold the list rounds' positions.
        uint64_t* K = kb.current();
        uint8_t* shared = reinterpret_cast<uint8_t*>(kb.alternate());
        uint32_t* list = reinterpret_cast<uint32_t*>(shared + ((N + 15) & ~(size_t)15));
        uint32_t* mark32 = (uint32_t*)b.head.ptr;
        uint32_t* start32 = vb.alternate();
        uint32_t* sa32 = pos_sorted;
        uint32_t* posA = vb.alternate();
        uint32_t* posB = (uint32_t*)b.head.ptr;
        hipLaunchKernelGGL(sa_marks_whole, dim3(grid), dim3(block), 0, st, (const uint64_t*)K, n,
                           mark32);
        size_t sb = tmp_bytes;
        if (check(hipGetLastError(), "sa_marks_whole") ||
            check(rocprim::inclusive_scan(b.tmp.ptr, sb, mark32, start32, N,
                                          rocprim::maximum<uint32_t>(), st), "max scan"))
            return -1;
        hipLaunchKernelGGL(sa_place_whole, dim3(grid), dim3(block), 0, st, (const uint64_t*)K,
                           (const uint32_t*)sa32, (const uint32_t*)start32, n,
                           (uint32_t*)b.rank.ptr, shared);
        uint32_t m = 0;
        sb = tmp_bytes;
        if (check(hipGetLastError(), "sa_place_whole") ||
            check(rocprim::select(b.tmp.ptr, sb, sa32, (const uint8_t*)shared, list,
                                  (uint32_t*)b.count.ptr, N, st), "select") ||
            check(hipMemcpyAsync(&m, b.count.ptr, 4, hipMemcpyDeviceToHost, st), "readback") ||
            check(hipStreamSynchronize(st), "suffix array round"))
            return -1;
        if ((int64_t)m * 4 > n)
            return fail("tpz_suffix_array: %u unresolved suffixes of %lld", m, (long long)n);
        const int bits = bit_width((uint64_t)n);
        while (m > 0) {
            h *= 2;
            if (h >= 2 * n) return fail("tpz_suffix_array: rounds did not converge");
            const int64_t M = m;
            const int g = grid_for(M);
            rocprim::double_buffer<uint64_t> lk(K, K + M);
            rocprim::double_buffer<uint32_t> lv(posA, posB);
            uint64_t* marks = K + 2 * M;
            uint64_t* starts = K + 3 * M;
            hipLaunchKernelGGL(sa_list_pairs, dim3(g), dim3(block), 0, st, (const uint32_t*)list,
                               M, (const uint32_t*)b.rank.ptr, n, h, bits, lk.current(),
                               lv.current());
            size_t tb = tmp_bytes;
            if (check(hipGetLastError(), "sa_list_pairs") ||
                check(rocprim::radix_sort_pairs(b.tmp.ptr, tb, lk, lv, (size_t)M, 0, 2 * bits,
                                                st), "radix sort"))
                return -1;
            hipLaunchKernelGGL(sa_marks, dim3(g), dim3(block), 0, st,
                               (const uint64_t*)lk.current(), M, bits, marks);
            sb = tmp_bytes;
            if (check(hipGetLastError(), "sa_marks") ||
                check(rocprim::inclusive_scan(b.tmp.ptr, sb, marks, starts, (size_t)M,
                                              HalvesMax(), st), "pair scan"))
                return -1;
            hipLaunchKernelGGL(sa_place, dim3(g), dim3(block), 0, st,
                               (const uint64_t*)lk.current(), (const uint32_t*)lv.current(),
                               (const uint64_t*)starts, M, bits, (uint32_t*)b.rank.ptr, sa32,
                               shared);
            sb = tmp_bytes;
            if (check(hipGetLastError(), "sa_place") ||
                check(rocprim::select(b.tmp.ptr, sb, lv.current(), (const uint8_t*)shared, list,
                                      (uint32_t*)b.count.ptr, (size_t)M, st), "select") ||
                check(hipMemcpyAsync(&m, b.count.ptr, 4, hipMemcpyDeviceToHost, st),
                      "readback") ||
                check(hipStreamSynchronize(st), "suffix array round"))
                return -1;
            ++rounds;
        }
        pos_sorted = sa32;
    }
    const double t2 = seconds();
    if (n >= kReadStaged) {
        if (sa_download(pos_sorted, n, sa)) return -1;
    } else {
        // positions as the reference's 64-bit entries, into a key buffer
        // that no longer holds anything needed
        int64_t* out = (int64_t*)b.key0.ptr;
        hipLaunchKernelGGL(sa_widen, dim3(grid), dim3(block), 0, st, (const uint32_t*)pos_sorted,
                           n, out);
        if (check(hipGetLastError(), "sa_widen") ||
            check(hipMemcpyAsync(sa, out, N * 8, hipMemcpyDeviceToHost, st), "readback") ||
            check(hipStreamSynchronize(st), "suffix array"))
            return -1;
    }
    if (stats) {
        stats->rounds = rounds;
        stats->device_bytes = (int64_t)(N * 33 + 8 + tmp_bytes);
        stats->alloc_s = t1 - t0;
        stats->sort_s = t2 - t1;
        stats->readback_s = seconds() - t2;
    }
    return 0;
}

int32_t divsufsort64(const uint8_t* T, int64_t* SA, int64_t n) {
    if (!T || !SA || n < 0) return -1;                  // divsufsort.c:338
    if (tpz_suffix_array(T, n, SA, nullptr)) {
        fprintf(stderr, "topaz_amd: divsufsort64: %s\n", tpz_last_error());
        return -2;
    }
    return 0;
}

}  // extern "C"
