// Internal host-side helpers shared by the library's translation units.
#ifndef TOPAZ_AMD_RUNTIME_H
#define TOPAZ_AMD_RUNTIME_H

#include <hip/hip_runtime.h>
#include <stddef.h>

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

}  // namespace tpz

#endif
