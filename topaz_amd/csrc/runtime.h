// Internal host-side helpers shared by the library's translation units.
#ifndef TOPAZ_AMD_RUNTIME_H
#define TOPAZ_AMD_RUNTIME_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace tpz {

int fail(const char* fmt, ...);          // set the error string, return -1
int check(hipError_t e, const char* what);
int ensure_init();                       // tpz_init(0) on first use
hipStream_t stream();
int num_cu();

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    void reset();
    int alloc(size_t n);
    int upload(const void* src, size_t n);
};

// Letter codes of a text for the gapless kernel (see gapless_kernels.h).
struct TextCodes {
    uint8_t code[256];        // text byte -> code; 0xff: the byte is not in the text
    uint8_t byte_of[32];      // code -> byte
    int n_codes = 0;          // codes 0 .. n_codes-1 (kTextTerm is '\n')
};

// Query copies of a gapless batch in the kernel's layout: forward and
// reversed doubled codes with terminators, and their offsets.
struct QueryCodes {
    std::vector<uint8_t> fwd, rev;
    std::vector<int64_t> fwd_off, rev_off;
};
int gapless_query_codes(const TextCodes& tc, int32_t n_queries, const char* q_chars,
                        const int64_t* q_off, const int32_t* q_len, QueryCodes& out);
// W(text code, query code) for the kernel, kTableRows x kTableCols int16.
void gapless_table(const TextCodes& tc, const int8_t* matrix, int16_t* table);

}  // namespace tpz

// Device-resident text of the gapless kernel.
struct tpz_text {
    uint64_t t_size = 0;
    bool has_sa = false;
    tpz::TextCodes codes;
    tpz::DeviceBuffer d_fwd, d_rev, d_sa;
};

#endif
