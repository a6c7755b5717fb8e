// Roofline calibration for the gapless extension: the time to read, per
// seed, the two text windows a seed's scans read (132 bytes on each side of
// a random position, 264 contiguous bytes from a dword boundary) from a
// device-resident text, with no compute beyond folding the words into one
// result per seed.  The gapless kernel cannot move its windows faster than
// this, whatever its arithmetic; the ratio of the two times is how close the
// kernel is to the memory system's rate for random window reads.
//
//     probe_window_gather [text_bytes] [seeds]
//
// Prints one JSON line: milliseconds for two windows per seed (the gapless
// pattern) and for one.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 1;                                                             \
        }                                                                         \
    } while (0)

namespace {

constexpr int kWindowDwords = 33;     // a 128-byte window from a dword boundary, and its tail

// Seed i reads [pos[i] - 128, pos[i] + 136) (two windows) or [pos[i], pos[i]
// + 136) (one); the caller keeps every window inside the text.
__global__ void __launch_bounds__(256) gather_kernel(const uint8_t* text, const int64_t* pos,
                                                     int64_t n, int sides, uint32_t* out) {
    typedef const __attribute__((address_space(1))) uint32_t* global_u32;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t p = pos[i];
        uint32_t acc = 0;
        for (int s = 0; s < sides; ++s) {
            const uintptr_t a = reinterpret_cast<uintptr_t>(text + (s ? p - 128 : p));
            const global_u32 d = (global_u32)(a & ~(uintptr_t)3);
            uint32_t x[kWindowDwords];
#pragma unroll
            for (int k = 0; k < kWindowDwords; ++k) x[k] = d[k];
#pragma unroll
            for (int k = 0; k < kWindowDwords; ++k) acc ^= x[k] * (uint32_t)(k + 1);
        }
        out[i] = acc;
    }
}

}  // namespace

int main(int argc, char** argv) {
    const int64_t t_size = argc > 1 ? atoll(argv[1]) : 1000000000LL;
    const int64_t n = argc > 2 ? atoll(argv[2]) : 50000000LL;
    if (t_size < 4096 || n < 1) {
        fprintf(stderr, "need text_bytes >= 4096 and seeds >= 1\n");
        return 2;
    }
    // positions in [256, t_size - 512): both windows and their tails inside
    std::vector<int64_t> pos((size_t)n);
    std::mt19937_64 rng(2026);
    for (auto& p : pos) p = 256 + (int64_t)(rng() % (uint64_t)(t_size - 768));
    uint8_t* d_text = nullptr;
    int64_t* d_pos = nullptr;
    uint32_t* d_out = nullptr;
    CHECK(hipMalloc(&d_text, (size_t)t_size));
    CHECK(hipMalloc(&d_pos, (size_t)n * 8));
    CHECK(hipMalloc(&d_out, (size_t)n * 4));
    CHECK(hipMemset(d_text, 7, (size_t)t_size));
    CHECK(hipMemcpy(d_pos, pos.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    float best[3] = {1e30f, 1e30f, 1e30f};
    for (int sides = 1; sides <= 2; ++sides) {
        for (int rep = 0; rep < 5; ++rep) {
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(gather_kernel, dim3(8192), dim3(256), 0, 0, d_text, d_pos, n, sides,
                               d_out);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms = 0;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < best[sides]) best[sides] = ms;     // the first run warms up
        }
    }
    printf("{\"text_bytes\": %lld, \"seeds\": %lld, \"ms_two_windows\": %.4f, "
           "\"ms_one_window\": %.4f}\n",
           (long long)t_size, (long long)n, best[2], best[1]);
    (void)hipFree(d_text);
    (void)hipFree(d_pos);
    (void)hipFree(d_out);
    return 0;
}
