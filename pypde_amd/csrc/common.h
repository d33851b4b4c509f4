// Shared helpers for the pypde_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdarg>
#include <string>

namespace pypde {

// Thread-local last-error message returned by pypde_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

enum Status : int {
    OK = 0,
    ERR_ARG = -1,     // invalid argument (shape, pointer, size)
    ERR_HIP = -2,     // HIP runtime error
    ERR_UNSUP = -3,   // unsupported configuration
};

inline int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return ERR_HIP;
    }
    return OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;

// Workgroups are dispatched to the device's 8 XCDs in turn (workgroup i to
// XCD i mod 8), and each XCD has its own L2.  Renumbering them so that XCD x
// runs a contiguous range of work items keeps the cache lines that
// neighbouring items share (GEMM operand panels, the 128-byte lines that
// neighbouring columns of a row-major array share) in one L2 instead of
// fetching them into all eight.  The map is a bijection on [0, total):
// results do not change, only where the work runs.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_remap(int pid, int total)
{
    const int q = total / kXcd, r = total - q * kXcd;
    const int x = pid % kXcd, local = pid / kXcd;
    return x * q + (x < r ? x : r) + local;
}

// Bytes of the device tables of the multi launches (one per translation unit).
long long gemm_multi_bytes();
long long recurrence_multi_bytes();
long long products_multi_bytes();

}  // namespace pypde

#define PYPDE_REQUIRE(cond, ...)                 \
    do {                                         \
        if (!(cond)) {                           \
            ::pypde::set_error(__VA_ARGS__);     \
            return ::pypde::ERR_ARG;             \
        }                                        \
    } while (0)

#define PYPDE_HIP(call)                                                        \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) {                                                \
            ::pypde::set_error("%s: %s", #call, hipGetErrorString(_e));        \
            return ::pypde::ERR_HIP;                                           \
        }                                                                      \
    } while (0)
