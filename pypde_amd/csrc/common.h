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
