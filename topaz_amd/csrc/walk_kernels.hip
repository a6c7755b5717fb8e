// The GPU seed walk: walk_core.h's walk of one query per wavefront, and the
// copy that gathers each stream's rounds into one contiguous run of
// elements before it.
//
// The walk is a chain of dependent steps (run queue, hit list), so a query gets
// one wavefront whose 64 lanes all execute the same steps on the same data:
// their loads and stores hit the same addresses and coalesce, and every
// lane sees its own stores.  Only the scan of a run's elements is split
// over the lanes, 64 elements per step: the first element that ends the run
// or that the hit list must see is found with a ballot.  Walk state lives
// in global memory (one WalkMem block per query), written -1 where the walk
// expects it before the walk starts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_core.h"
#include "walk_kernels.h"

namespace tpz {

struct WaveScan {
    int lane;
    __device__ int32_t operator()(const int16_t* v, int32_t i, int32_t len, int32_t start,
                                  int32_t leader, bool all_slow, int32_t worst) const {
        if (all_slow) return i;
        for (int32_t b = i;; b += 64) {
            const int32_t j = b + lane;
            bool hit = true;
            if (j < len) {
                const int32_t x = v[j];
                hit = (j > start && (x == kStreamEnd || x < leader)) || x > worst;
            }
            const uint64_t m = __ballot(hit);
            if (m) return b + (int32_t)__builtin_ctzll(m);
        }
    }
};

// The run queue's stack tops (one int per leader score, read and written at
// every pop and push) in LDS when they fit kWalkTopLds bytes: dynamic shared
// memory of the launch, addressed as LDS so that its loads wait only for
// LDS.
constexpr int kWalkTopLds = 32768;
typedef __attribute__((address_space(3))) int32_t lds_i32;
struct WaveMem : WalkMem {
    lds_i32* top;
};

__global__ void __launch_bounds__(64) walk_kernel(WalkArgs a) {
    extern __shared__ int32_t lds_top[];
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const int32_t s0 = a.q_stream[q], n_streams = a.q_stream[q + 1] - s0;
    const int32_t capacity = a.q_capacity[q];
    const WalkMemLayout L = walk_mem_layout(n_streams, capacity, a.max_score);
    char* base = a.mem + a.q_mem[q];
    uint4* init = reinterpret_cast<uint4*>(base);
    const uint4 ones = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    for (size_t w = lane; w < walk_mem_init_bytes(L) / 16; w += 64) init[w] = ones;
    const bool top_in_lds = 4 * (a.max_score + 1) <= kWalkTopLds;
    if (top_in_lds)
        for (int s = lane; s <= a.max_score; s += 64) lds_top[s] = -1;
    __syncthreads();
    const WalkMem m = walk_mem_at(base, L);
    const WalkStreams st{n_streams, a.s_base + s0, a.s_len + s0, a.s_open + s0, a.vals, a.prot};
    WalkResult r;
    if (top_in_lds) {
        WaveMem wm;
        static_cast<WalkMem&>(wm) = m;
        wm.top = (lds_i32*)lds_top;
        r = walk_query(st, wm, capacity, a.limit, a.max_score, WaveScan{lane},
                       a.out_index + a.q_out[q], a.out_score + a.q_out[q]);
    } else {
        r = walk_query(st, m, capacity, a.limit, a.max_score, WaveScan{lane},
                       a.out_index + a.q_out[q], a.out_score + a.q_out[q]);
    }
    if (lane == 0) {
        a.out_n[q] = r.n_hits;
        a.out_status[q] = r.status;
        a.out_stats[3 * (int64_t)q] = r.pops;
        a.out_stats[3 * (int64_t)q + 1] = r.runs;
        a.out_stats[3 * (int64_t)q + 2] = r.counted;
    }
}

// One job per wavefront iteration, 64 elements per step.
__global__ void __launch_bounds__(256) compact_kernel(CompactArgs a) {
    const int lane = threadIdx.x % 64;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t j = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; j < a.n_jobs;
         j += waves) {
        const CompactJob job = a.jobs[j];
        const int16_t* sv = a.round_vals[job.round] + job.src;
        const uint32_t* sp = a.round_prot[job.round] + job.src;
        for (int32_t e = lane; e < job.n; e += 64) {
            a.vals[job.dst + e] = sv[e];
            a.prot[job.dst + e] = sp[e];
        }
    }
}

}  // namespace tpz

extern "C" hipError_t tpz_launch_compact(const tpz::CompactArgs* a, hipStream_t stream) {
    if (a->n_jobs <= 0) return hipSuccess;
    const int64_t blocks = (a->n_jobs + 3) / 4;
    hipLaunchKernelGGL(tpz::compact_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)),
                       dim3(256), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t tpz_launch_walk(const tpz::WalkArgs* a, hipStream_t stream) {
    if (a->n_queries <= 0) return hipSuccess;
    const size_t top_bytes = 4 * (size_t)(a->max_score + 1);
    hipLaunchKernelGGL(tpz::walk_kernel, dim3((unsigned)a->n_queries), dim3(64),
                       top_bytes <= (size_t)tpz::kWalkTopLds ? top_bytes : 0, stream, *a);
    return hipGetLastError();
}
