// Stage G (sa_hsp over a batch of seeds) for callers that cannot allocate:
// tpz_gapless_run_noalloc.  Inputs and outputs are the caller's device
// arrays; everything else the gapless kernel reads -- the score table, the
// forward and reversed query code copies, their offsets and checked lengths
// -- is built by small kernels in a fixed device array of this file
// (`scratch`, kScratchBytes per device, held until the process ends).
//
// Nothing here calls hipMalloc, hipFree, hipHostMalloc or a copy from host
// memory: the text's letter codes, the alphabet maps and the matrix travel
// as kernel arguments.  Every call on a device shares that device's array,
// so calls that overlap on one device (on different streams, or from
// different threads without ordering) race; calls queued on one stream are
// ordered and safe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/topaz_amd.h"
#include "gapless_kernels.h"
#include "runtime.h"

namespace tpz {
namespace {

constexpr size_t kScratchBytes = (size_t)32 << 20;
// The device array.  In host code `scratch` is only a stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices];

// The address of this device's array (looked up on the first call on it,
// with the device current).
int scratch_address(int dev, uint8_t** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail("device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return -1;
        g_scratch_ptr[dev] = p;
    }
    *out = (uint8_t*)g_scratch_ptr[dev];
    return 0;
}

// Alphabet, codes and matrix, passed by value (1.2 KB of kernel arguments).
struct Tables {
    uint8_t code[256];       // text byte -> text code, 0xff: not in the text
    uint8_t byte_of[32];     // text code -> byte
    uint8_t used[32];        // text codes of letters of the text
    int8_t trow[128];        // byte -> matrix row as a text residue
    int8_t qrow[128];        // byte -> matrix row as a query residue
    int8_t matrix[625];
};

// Scratch layout: the table, then per query its copy offset and checked
// length, then the forward and reversed code copies.
struct Layout {
    size_t table, base, len, fwd, rev, end;
};

size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

Layout layout(int64_t n_queries, int64_t q_bytes) {
    Layout L;
    const size_t copy = round256((size_t)q_bytes + (size_t)n_queries + kTextPad + 8);
    L.table = 0;
    L.base = round256(L.table + sizeof(int16_t) * kTableRows * kTableCols);
    L.len = round256(L.base + 8 * (size_t)n_queries);
    L.fwd = round256(L.len + 4 * (size_t)n_queries);
    L.rev = L.fwd + copy;
    L.end = L.rev + copy;
    return L;
}

__device__ __forceinline__ uint8_t query_code(const Tables& t, uint8_t c) {
    return t.code[c] != 0xff ? t.code[c] : (uint8_t)(kForeignCode + t.qrow[c & 127]);
}

// W(text code, query code) as gapless_table builds it on the host, and
// terminators over both code copies (so that gaps between the caller's
// queries hold terminators too).
__global__ void noalloc_fill_kernel(Tables t, int16_t* table, uint8_t* fwd, uint8_t* rev,
                                    int64_t copy_bytes) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t e = i0; e < kTableRows * kTableCols; e += stride) {
        const int tc = (int)(e / kTableCols), q = (int)(e % kTableCols);
        int w = kTermScore;
        if (tc < kTextTerm && t.used[tc] && q != kQueryTerm) {
            const int tr = t.trow[t.byte_of[tc] & 127];
            int qr = -1;
            if ((q < kTextTerm && t.used[q]) || q == kTextTerm) qr = t.qrow[t.byte_of[q] & 127];
            else if (q >= kForeignCode && q < kForeignCode + 25) qr = q - kForeignCode;
            if (qr >= 0) w = t.matrix[tr * 25 + qr];
        }
        table[e] = (int16_t)w;
    }
    for (int64_t e = i0; e < copy_bytes; e += stride) {
        fwd[e] = 2 * kQueryTerm;
        rev[e] = 2 * kQueryTerm;
    }
}

// One workgroup per query: its copies go to offset q_off[k] + k (each
// earlier query adds one terminator), forward q[0..m-1] then a terminator,
// reversed q[m-1..1] then two terminators.  A query whose bytes lie outside
// [0, q_bytes) or overlap the next query's is given length 0 (no seed of it
// passes) and sets *status.
__global__ void noalloc_encode_kernel(Tables t, const char* q_chars, int64_t q_bytes,
                                      const int64_t* q_off, const int32_t* q_len,
                                      int32_t n_queries, int64_t* base, int32_t* len,
                                      uint8_t* fwd, uint8_t* rev, int32_t* status) {
    const int k = blockIdx.x;
    if (k >= n_queries) return;
    const int64_t o = q_off[k];
    const int32_t m = q_len[k];
    bool ok = o >= 0 && m >= 0 && o + m <= q_bytes;
    if (ok && k + 1 < n_queries) ok = q_off[k + 1] >= o + m;
    if (threadIdx.x == 0) {
        base[k] = ok ? o + k : 0;
        len[k] = ok ? m : 0;
        if (!ok && status) atomicOr(status, 1);
    }
    if (!ok) return;
    const int64_t b = o + k;
    const uint8_t* q = reinterpret_cast<const uint8_t*>(q_chars) + o;
    for (int32_t j = threadIdx.x; j < m; j += blockDim.x) {
        const uint8_t c = (uint8_t)(2 * query_code(t, q[j]));
        fwd[b + j] = c;
        if (j >= 1) rev[b + (m - 1 - j)] = c;
    }
}

__global__ void noalloc_status_kernel(int32_t* status) {
    if (threadIdx.x == 0) *status = 0;
}

}  // namespace
}  // namespace tpz

using namespace tpz;

extern "C" int tpz_gapless_run_noalloc(const tpz_text* text, int32_t n_queries,
                                       const char* q_chars, int64_t q_bytes,
                                       const int64_t* q_off, const int32_t* q_len,
                                       int64_t n_seeds, const tpz_seed* seeds,
                                       const int8_t* matrix, int32_t* score, int32_t* length,
                                       int32_t* status, void* stream) {
    if (!text) return fail("tpz_gapless_run_noalloc: null text");
    if (!matrix) return fail("tpz_gapless_run_noalloc: null matrix");
    if (n_queries < 0 || n_seeds < 0 || q_bytes < 0)
        return fail("tpz_gapless_run_noalloc: negative counts");
    if (n_queries > 0 && (!q_off || !q_len || (q_bytes > 0 && !q_chars)))
        return fail("tpz_gapless_run_noalloc: null query arrays");
    if (n_seeds > 0 && (!seeds || !score || !length))
        return fail("tpz_gapless_run_noalloc: null seed or result arrays");
    const Layout L = layout(n_queries, q_bytes);
    if (L.end > kScratchBytes)
        return fail("tpz_gapless_run_noalloc: %d queries of %lld bytes need %zu bytes of "
                    "scratch, more than the %zu this entry point has (use "
                    "tpz_gapless_batch_create)", n_queries, (long long)q_bytes, L.end,
                    kScratchBytes);
    if (ensure_init()) return -1;
    int dev = -1;
    if (check(hipGetDevice(&dev), "hipGetDevice")) return -1;
    if (dev != device_id())
        return fail("tpz_gapless_run_noalloc: the calling thread's device %d is not the "
                    "library's device %d", dev, device_id());
    uint8_t* base = nullptr;
    if (scratch_address(dev, &base)) return -1;
    hipStream_t st = stream ? (hipStream_t)stream : tpz::stream();

    Tables t;
    memcpy(t.code, text->codes.code, 256);
    memcpy(t.byte_of, text->codes.byte_of, 32);
    for (int c = 0; c < 32; ++c) t.used[c] = text->codes.used[c] ? 1 : 0;
    alphabet_rows(t.trow, t.qrow);
    memcpy(t.matrix, matrix, 625);

    int16_t* table = (int16_t*)(base + L.table);
    int64_t* q_base = (int64_t*)(base + L.base);
    int32_t* q_len_ok = (int32_t*)(base + L.len);
    uint8_t* fwd = base + L.fwd;
    uint8_t* rev = base + L.rev;
    const int64_t copy = (int64_t)(L.rev - L.fwd);
    if (status) {
        hipLaunchKernelGGL(noalloc_status_kernel, dim3(1), dim3(64), 0, st, status);
        if (check(hipGetLastError(), "noalloc status")) return -1;
    }
    const int fill_grid = (int)std::min<int64_t>(1024, (copy + 255) / 256 + 16);
    hipLaunchKernelGGL(noalloc_fill_kernel, dim3(fill_grid), dim3(256), 0, st, t, table, fwd,
                       rev, copy);
    if (check(hipGetLastError(), "noalloc fill")) return -1;
    if (n_queries > 0) {
        hipLaunchKernelGGL(noalloc_encode_kernel, dim3(n_queries), dim3(256), 0, st, t, q_chars,
                           q_bytes, q_off, q_len, n_queries, q_base, q_len_ok, fwd, rev,
                           status);
        if (check(hipGetLastError(), "noalloc encode")) return -1;
    }
    if (n_seeds == 0) return 0;
    GaplessArgs a{};
    a.t_fwd = (const uint8_t*)text->d_fwd.ptr;
    a.t_head = (const uint8_t*)text->d_head.ptr;
    a.t_size = (int64_t)text->t_size;
    a.sa = text->has_sa ? (const int64_t*)text->d_sa.ptr : nullptr;
    a.q_fwd = fwd;
    a.qf_off = q_base;
    a.q_rev = rev;
    a.qr_off = q_base;
    a.q_len = q_len_ok;
    a.n_queries = n_queries;
    a.seeds = reinterpret_cast<const SeedRec*>(seeds);
    a.n_seeds = n_seeds;
    a.table = table;
    a.out_score = score;
    a.out_length = length;
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((n_seeds + 255) / 256, (int64_t)std::max(num_cu(), 1) * 32));
    return check(tpz_launch_gapless(&a, grid, st), "gapless (noalloc)");
}
