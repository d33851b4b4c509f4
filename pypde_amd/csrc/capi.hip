// Library-level entry points: error reporting, version, small pointwise ops.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <cstring>

namespace pypde {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

__global__ void axpby_kernel(long long n, double alpha, const double* __restrict__ x, double beta,
                             double* __restrict__ y)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = beta == 0.0 ? alpha * x[i] : alpha * x[i] + beta * y[i];   // beta 0: y not read
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_version(void) { return 1; }

extern "C" const char* pypde_last_error(void) { return last_error(); }

extern "C" int pypde_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int pypde_axpby(long long n, double alpha, const double* x, double beta, double* y,
                           void* stream)
{
    PYPDE_REQUIRE(n >= 0, "negative length");
    if (n == 0) return OK;
    PYPDE_REQUIRE(x && y, "null pointer");
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), n,
                       alpha, x, beta, y);
    return check_launch("pypde_axpby");
}
