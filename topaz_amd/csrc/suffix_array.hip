// Suffix array of TOPAZ's text on gfx950, for `topaz index`: the order
// divsufsort64 gives (index.c:66), i.e. the suffixes of the text as
// unsigned bytes, a suffix before every longer suffix it is a prefix of.
// The suffix array of a string is unique, so any correct construction
// gives the reference's .SA byte for byte.
//
// Prefix doubling: rank[i] orders the suffixes by their first h bytes
// (the byte + 1 to begin with, 0 past the end); each round sorts the
// positions by the pair (rank[i], rank[i + h]) with a device radix sort,
// numbers the distinct pairs in sorted order with a scan, and doubles h,
// until every rank is distinct.  Each round is a few passes over 16-byte
// records in HBM; a round's sort covers only the bits the current ranks
// need (2 x bit width of the largest rank), 18 bits in the first.  Rounds:
// about log2 of the longest repeated substring, plus one.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "../../include/topaz_amd.h"
#include "../../include/topaz_index_compat.h"
#include "runtime.h"

namespace tpz {
namespace {

__global__ void sa_first_rank(const uint8_t* __restrict__ T, int64_t n, uint32_t* __restrict__ rank) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        rank[i] = (uint32_t)T[i] + 1;
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
    DeviceBuffer text, rank, key0, key1, pos0, pos1, head, tmp;
};

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
    hipStream_t st = stream();
    SaBuffers b;
    const size_t N = (size_t)n;
    if (b.text.upload(text, N) || b.rank.alloc(N * 4) || b.key0.alloc(N * 8) ||
        b.key1.alloc(N * 8) || b.pos0.alloc(N * 4) || b.pos1.alloc(N * 4) || b.head.alloc(N * 4))
        return -1;
    // scratch of the sort and the scan, the larger of the two
    size_t sort_bytes = 0, scan_bytes = 0;
    {
        rocprim::double_buffer<uint64_t> kb((uint64_t*)b.key0.ptr, (uint64_t*)b.key1.ptr);
        rocprim::double_buffer<uint32_t> vb((uint32_t*)b.pos0.ptr, (uint32_t*)b.pos1.ptr);
        if (check(rocprim::radix_sort_pairs(nullptr, sort_bytes, kb, vb, N, 0, 64, st),
                  "radix sort size") ||
            check(rocprim::inclusive_scan(nullptr, scan_bytes, (uint32_t*)b.head.ptr,
                                          (uint32_t*)b.key1.ptr, N, rocprim::plus<uint32_t>(), st),
                  "scan size"))
            return -1;
    }
    if (b.tmp.alloc(std::max(sort_bytes, scan_bytes) + 256)) return -1;
    const int block = 256;
    const int grid = (int)std::min<int64_t>((n + block - 1) / block, (int64_t)num_cu() * 64);
    hipLaunchKernelGGL(sa_first_rank, dim3(grid), dim3(block), 0, st, (const uint8_t*)b.text.ptr,
                       n, (uint32_t*)b.rank.ptr);
    if (check(hipGetLastError(), "sa_first_rank")) return -1;
    uint64_t max_rank = 256;
    int rounds = 0;
    uint32_t* pos_sorted = nullptr;
    for (int64_t h = 1;; h *= 2) {
        const int bits = bit_width(max_rank);
        rocprim::double_buffer<uint64_t> kb((uint64_t*)b.key0.ptr, (uint64_t*)b.key1.ptr);
        rocprim::double_buffer<uint32_t> vb((uint32_t*)b.pos0.ptr, (uint32_t*)b.pos1.ptr);
        hipLaunchKernelGGL(sa_pairs, dim3(grid), dim3(block), 0, st, (const uint32_t*)b.rank.ptr,
                           n, h, bits, kb.current(), vb.current());
        size_t tb = sort_bytes;
        if (check(hipGetLastError(), "sa_pairs") ||
            check(rocprim::radix_sort_pairs(b.tmp.ptr, tb, kb, vb, N, 0, 2 * bits, st),
                  "radix sort"))
            return -1;
        hipLaunchKernelGGL(sa_heads, dim3(grid), dim3(block), 0, st,
                           (const uint64_t*)kb.current(), n, (uint32_t*)b.head.ptr);
        // group numbers (1-based, in sorted order) into the free key buffer
        uint32_t* group = reinterpret_cast<uint32_t*>(kb.alternate());
        size_t sb = scan_bytes;
        if (check(hipGetLastError(), "sa_heads") ||
            check(rocprim::inclusive_scan(b.tmp.ptr, sb, (uint32_t*)b.head.ptr, group, N,
                                          rocprim::plus<uint32_t>(), st),
                  "scan"))
            return -1;
        hipLaunchKernelGGL(sa_rerank, dim3(grid), dim3(block), 0, st,
                           (const uint32_t*)vb.current(), (const uint32_t*)group, n,
                           (uint32_t*)b.rank.ptr);
        uint32_t groups = 0;
        if (check(hipGetLastError(), "sa_rerank") ||
            check(hipMemcpyAsync(&groups, group + (n - 1), 4, hipMemcpyDeviceToHost, st),
                  "readback") ||
            check(hipStreamSynchronize(st), "suffix array round"))
            return -1;
        ++rounds;
        pos_sorted = vb.current();
        if ((int64_t)groups == n || h >= n) break;
        max_rank = groups;
    }
    // positions as the reference's 64-bit entries, into a key buffer
    int64_t* out = (int64_t*)b.key0.ptr;
    hipLaunchKernelGGL(sa_widen, dim3(grid), dim3(block), 0, st, (const uint32_t*)pos_sorted, n, out);
    if (check(hipGetLastError(), "sa_widen") ||
        check(hipMemcpyAsync(sa, out, N * 8, hipMemcpyDeviceToHost, st), "readback") ||
        check(hipStreamSynchronize(st), "suffix array"))
        return -1;
    if (stats) {
        stats->rounds = rounds;
        stats->device_bytes = (int64_t)(N * 33 + std::max(sort_bytes, scan_bytes));
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
