// Line operators along one axis of a batch of row-major 2-D float64 arrays.
//
// Every operator is a sequential recurrence (or a short banded stencil) along
// a line, exactly as the reference's Fortran kernels run it.  A block stages a
// panel of W lines into LDS with coalesced global accesses (rows of 64
// adjacent columns for axis 0, whole contiguous rows for axis 1), one lane
// then runs the operator on its own line in LDS, and the panel is written back
// coalesced.  The arithmetic of each operator is the reference's, statement by
// statement, with FMA contraction disabled, so results are bit-identical to
// the Fortran/C restatement of the reference.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <algorithm>

namespace pypde {

struct LineGeom {
    const double* src;
    double* dst;
    long long src_ld, dst_ld, src_bs, dst_bs;
    int len_in, len_out, nlines, batch, axis;
    int W;      // lines per block (power of two)
    int pitch;  // LDS doubles per line (odd)
};

// One block handles W lines of one batch entry.
template <class Op>
__global__ void __launch_bounds__(256) line_panel_kernel(LineGeom g, Op op)
{
    extern __shared__ double lds[];
    const int bpe = (g.nlines + g.W - 1) / g.W;
    const int b = blockIdx.x / bpe;
    const int l0 = (blockIdx.x - b * bpe) * g.W;
    const int nl = min(g.W, g.nlines - l0);
    const double* __restrict__ src = g.src + (long long)b * g.src_bs;
    double* __restrict__ dst = g.dst + (long long)b * g.dst_bs;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;

    if (g.axis == 0) {
        // rows of the panel: element (k, c) of the input lives at src[k*ld + l0 + c]
        const int rows_per_pass = nt / g.W;  // W divides nt or W >= nt
        if (rows_per_pass >= 1) {
            const int c = tid & (g.W - 1);
            for (int k = tid / g.W; k < g.len_in; k += rows_per_pass)
                if (c < nl) lds[c * g.pitch + k] = src[(long long)k * g.src_ld + l0 + c];
        } else {
            for (int k = 0; k < g.len_in; ++k)
                for (int c = tid; c < nl; c += nt) lds[c * g.pitch + k] = src[(long long)k * g.src_ld + l0 + c];
        }
    } else {
        for (int c = 0; c < nl; ++c) {
            const double* row = src + (long long)(l0 + c) * g.src_ld;
            for (int k = tid; k < g.len_in; k += nt) lds[c * g.pitch + k] = row[k];
        }
    }
    __syncthreads();
    for (int c = tid; c < nl; c += nt) op(lds + c * g.pitch, b, l0 + c);
    __syncthreads();
    if (g.axis == 0) {
        const int rows_per_pass = nt / g.W;
        if (rows_per_pass >= 1) {
            const int c = tid & (g.W - 1);
            for (int k = tid / g.W; k < g.len_out; k += rows_per_pass)
                if (c < nl) dst[(long long)k * g.dst_ld + l0 + c] = lds[c * g.pitch + k];
        } else {
            for (int k = 0; k < g.len_out; ++k)
                for (int c = tid; c < nl; c += nt) dst[(long long)k * g.dst_ld + l0 + c] = lds[c * g.pitch + k];
        }
    } else {
        for (int c = 0; c < nl; ++c) {
            double* row = dst + (long long)(l0 + c) * g.dst_ld;
            for (int k = tid; k < g.len_out; k += nt) row[k] = lds[c * g.pitch + k];
        }
    }
}

// ------------------------------------------------------------------------
// Operators.  Each works in place on one LDS line x[0 .. max(len_in,len_out)).
// ------------------------------------------------------------------------
#pragma clang fp contract(off)

// u = S v  (chebyshev.py:287-293): u_i = s_{i-2} v_{i-2} + v_i, summed in the
// column order of the csc stencil; i >= M keeps only the s-term.
struct StencilOp {
    const double* s;
    int M;
    __device__ __forceinline__ void apply(double* x) const
    {
        const int N = M + 2;
        for (int i = N - 1; i >= 0; --i) {
            double y;
            if (i >= 2) {
                const double t = s[i - 2] * x[i - 2];
                y = (i < M) ? t + x[i] : t;
            } else {
                y = x[i];
            }
            x[i] = y;
        }
    }
    __device__ void operator()(double* x, int, int) const { apply(x); }
};

// Chebyshev derivative recurrence (differentiate_cheby.f90:28-53) applied
// `order` times, then divided by scale_pow (field_operations.py:40-45).
__device__ __forceinline__ void cheb_deriv_inplace(double* x, int n)
{
    double cn = x[n - 1];       // c_{j+1} (input value) for the next j
    double dp2 = 0.0;           // dc_{j+2}
    double dp1 = 0.0;           // dc_{j+1}
    x[n - 1] = 0.0;
    {
        const double cj = x[n - 2];
        const double dj = (double)(2 * (n - 1)) * cn;
        x[n - 2] = dj;
        dp2 = 0.0;
        dp1 = dj;
        cn = cj;
    }
    for (int j = n - 3; j >= 1; --j) {
        const double cj = x[j];
        const double dj = dp2 + (double)(2 * (j + 1)) * cn;
        x[j] = dj;
        dp2 = dp1;
        dp1 = dj;
        cn = cj;
    }
    // dc_0 = dc_2 / 2 + c_1   (dp2 holds dc_2 here, cn holds c_1)
    x[0] = dp2 / 2.0 + cn;
}

struct DerivOp {
    const double* s;  // stencil (nullptr for Chebyshev input)
    int len_in;       // M for Galerkin input, N for Chebyshev
    int n;            // Chebyshev length
    int order;
    double scale_pow;
    __device__ void operator()(double* x, int, int) const
    {
        if (s) StencilOp{s, len_in}.apply(x);
        for (int o = 0; o < order; ++o) cheb_deriv_inplace(x, n);
        for (int i = 0; i < n; ++i) x[i] = x[i] / scale_pow;
    }
};

// v = (S^T S)^-1 S^T u with the offset-2 tridiagonal solve (tdma.f90:56-107).
struct StencilInvOp {
    const double *s, *a, *b, *w, *den;
    int M;
    __device__ void operator()(double* x, int, int) const
    {
        // y_k = u_k + s_k u_{k+2}  (csc column order of S^T)
        for (int k = 0; k < M; ++k) x[k] = x[k] + s[k] * x[k + 2];
        // forward sweep g
        double g2 = x[0] / b[0];
        double g1 = x[1] / b[1];
        x[0] = g2;
        x[1] = g1;
        for (int i = 2; i < M; ++i) {
            const double gi = (x[i] - a[i - 2] * g2) / den[i];
            x[i] = gi;
            g2 = g1;
            g1 = gi;
        }
        // back substitution  x_j = g_j - w_j x_{j+2}
        double xp2 = x[M - 1];
        double xp1 = x[M - 2];
        for (int j = M - 3; j >= 0; --j) {
            const double xj = x[j] - w[j] * xp2;
            x[j] = xj;
            xp2 = xp1;
            xp1 = xj;
        }
    }
};

// Pre-factored four-diagonal solve (fdma.f90:1-38), tables per batch entry.
__device__ __forceinline__ void fdma_inplace(double* x, int n, const double* __restrict__ l,
                                             const double* __restrict__ d,
                                             const double* __restrict__ u1,
                                             const double* __restrict__ u2)
{
    double a2 = x[0], a1 = x[1];
    for (int i = 2; i < n; ++i) {
        const double v = x[i] - l[i - 2] * a2;
        x[i] = v;
        a2 = a1;
        a1 = v;
    }
    double p1 = x[n - 1] / d[n - 1];
    double p2 = x[n - 2] / d[n - 2];
    x[n - 1] = p1;
    x[n - 2] = p2;
    double p3 = (x[n - 3] - u1[n - 3] * p1) / d[n - 3];
    x[n - 3] = p3;
    double p4 = (x[n - 4] - u1[n - 4] * p2) / d[n - 4];
    x[n - 4] = p4;
    // rolling window: x_{i+1}=p4? keep x_{i+2}, x_{i+4}
    double xi1 = p4, xi2 = p3, xi3 = p2, xi4 = p1;  // x_{i+1..i+4} for i = n-5
    for (int i = n - 5; i >= 0; --i) {
        const double v = (x[i] - u1[i] * xi2 - u2[i] * xi4) / d[i];
        x[i] = v;
        xi4 = xi3;
        xi3 = xi2;
        xi2 = xi1;
        xi1 = v;
    }
}

// Transposed solve A^T x = b with A = L U from the same tables.
__device__ __forceinline__ void fdma_t_inplace(double* x, int n, const double* __restrict__ l,
                                               const double* __restrict__ d,
                                               const double* __restrict__ u1,
                                               const double* __restrict__ u2)
{
    // U^T z = b : z_i = (b_i - u1_{i-2} z_{i-2} - u2_{i-4} z_{i-4}) / d_i
    double z1 = 0.0, z2 = 0.0, z3 = 0.0, z4 = 0.0;  // z_{i-1..i-4}
    for (int i = 0; i < n; ++i) {
        double v = x[i];
        if (i >= 2) v = v - u1[i - 2] * z2;
        if (i >= 4) v = v - u2[i - 4] * z4;
        v = v / d[i];
        x[i] = v;
        z4 = z3;
        z3 = z2;
        z2 = z1;
        z1 = v;
    }
    // L^T x = z : x_j = z_j - l_j x_{j+2}
    double xp2 = x[n - 1], xp1 = x[n - 2];
    for (int j = n - 3; j >= 0; --j) {
        const double v = x[j] - l[j] * xp2;
        x[j] = v;
        xp2 = xp1;
        xp1 = v;
    }
}

struct FdmaOp {
    const double *l, *d, *u1, *u2;
    long long tbs;
    int n;
    int transpose;
    __device__ void operator()(double* x, int b, int) const
    {
        const long long o = (long long)b * tbs;
        if (transpose) fdma_t_inplace(x, n, l + o, d + o, u1 + o, u2 + o);
        else fdma_inplace(x, n, l + o, d + o, u1 + o, u2 + o);
    }
};

// Poisson line solve with per-line tables (d | u1 | u2 | l) of length n each.
struct PoissonOp {
    const double* tab;
    const int* sing;
    int n;
    int transpose;
    __device__ void operator()(double* x, int, int line) const
    {
        const double* t = tab + (long long)line * 4 * n;
        const double *d = t, *u1 = t + n, *u2 = t + 2 * n, *l = t + 3 * n;
        if (sing[line]) {
            // reduced system on x[1:], tables stored from index 0 for length n-1
            if (transpose) fdma_t_inplace(x + 1, n - 1, l, d, u1, u2);
            else fdma_inplace(x + 1, n - 1, l, d, u1, u2);
            x[0] = 0.0;
        } else {
            if (transpose) fdma_t_inplace(x, n, l, d, u1, u2);
            else fdma_inplace(x, n, l, d, u1, u2);
        }
    }
};

// Upper two-band back substitution (twodma.f90:2-25).
struct TwodmaOp {
    const double *d, *u;
    int n;
    __device__ void operator()(double* x, int, int) const
    {
        double p1 = x[n - 1] / d[n - 1];
        double p2 = x[n - 2] / d[n - 2];
        x[n - 1] = p1;
        x[n - 2] = p2;
        for (int i = n - 3; i >= 0; --i) {
            const double v = (x[i] - u[i] * p1) / d[i];
            x[i] = v;
            p1 = p2;
            p2 = v;
        }
    }
};

// PDMA_Solve (pdma.py:74-90) with PDMA_LU tables.
struct PdmaOp {
    const double *a, *b, *d, *e, *f;
    int n;
    __device__ void operator()(double* x, int, int) const
    {
        x[2] -= b[0] * x[0];
        x[3] -= b[1] * x[1];
        for (int k = 4; k < n; ++k) x[k] -= b[k - 2] * x[k - 2] + a[k - 4] * x[k - 4];
        x[n - 1] /= d[n - 1];
        x[n - 2] /= d[n - 2];
        x[n - 3] = (x[n - 3] - e[n - 3] * x[n - 1]) / d[n - 3];
        x[n - 4] = (x[n - 4] - e[n - 4] * x[n - 2]) / d[n - 4];
        for (int k = n - 5; k >= 0; --k) x[k] = (x[k] - e[k] * x[k + 2] - f[k] * x[k + 4]) / d[k];
    }
};

// Banded y = A x, diagonals with offsets in [-4, 4], ascending column order.
constexpr int kMaxDiag = 9;
struct BandedOp {
    const double* vals;
    int off[kMaxDiag];
    int nd;
    int len_in, len_out;
    __device__ void operator()(double* x, int, int) const
    {
        double old1 = 0.0, old2 = 0.0, old3 = 0.0, old4 = 0.0;  // input x_{r-1..r-4}
        for (int r = 0; r < len_out; ++r) {
            const double xr = (r < len_in) ? x[r] : 0.0;
            double acc = 0.0;
            bool first = true;
            for (int q = 0; q < nd; ++q) {
                const int c = r + off[q];
                if (c < 0 || c >= len_in) continue;
                const double a = vals[(long long)q * len_out + r];
                if (a == 0.0) continue;
                double xc;
                const int o = off[q];
                if (o == 0) xc = xr;
                else if (o == -1) xc = old1;
                else if (o == -2) xc = old2;
                else if (o == -3) xc = old3;
                else if (o == -4) xc = old4;
                else xc = x[c];
                const double t = a * xc;
                acc = first ? t : acc + t;
                first = false;
            }
            x[r] = acc;
            old4 = old3;
            old3 = old2;
            old2 = old1;
            old1 = xr;
        }
    }
};

#pragma clang fp contract(on)

// ------------------------------------------------------------------------
// Launch helpers
// ------------------------------------------------------------------------
static int pick_lines_per_block(int len, int hint, int nlines)
{
    if (hint > 0) {
        int w = 1;
        while (w * 2 <= hint) w *= 2;
        return w;
    }
    const int pitch = (len | 1);
    const int budget = 64 * 1024;  // bytes of LDS per block
    int w = 1;
    while (w * 2 <= 64 && (w * 2) * pitch * 8 <= budget) w *= 2;
    // do not stage far more lines than exist
    while (w > 1 && w / 2 >= nlines) w /= 2;
    return w;
}

static int make_geom(const pypde_lines* in, LineGeom& g)
{
    PYPDE_REQUIRE(in != nullptr, "null pypde_lines");
    PYPDE_REQUIRE(in->src && in->dst, "null src/dst pointer");
    PYPDE_REQUIRE(in->axis == 0 || in->axis == 1, "axis must be 0 or 1 (got %d)", in->axis);
    PYPDE_REQUIRE(in->len_in >= 1 && in->len_out >= 1 && in->nlines >= 0 && in->batch >= 0,
                  "invalid line geometry len_in=%d len_out=%d nlines=%d batch=%d",
                  in->len_in, in->len_out, in->nlines, in->batch);
    g.src = in->src;
    g.dst = in->dst;
    g.src_ld = in->src_ld;
    g.dst_ld = in->dst_ld;
    g.src_bs = in->src_bs;
    g.dst_bs = in->dst_bs;
    g.len_in = in->len_in;
    g.len_out = in->len_out;
    g.nlines = in->nlines;
    g.batch = in->batch;
    g.axis = in->axis;
    const int len = std::max(in->len_in, in->len_out);
    g.pitch = len | 1;
    g.W = pick_lines_per_block(len, in->lines_per_block, in->nlines);
    PYPDE_REQUIRE((long long)g.W * g.pitch * 8 <= 160 * 1024,
                  "line length %d too long for an LDS panel", len);
    return OK;
}

template <class Op>
static int launch_lines(const LineGeom& g, const Op& op, void* stream, const char* name)
{
    if (g.nlines == 0 || g.batch == 0) return OK;
    const int bpe = (g.nlines + g.W - 1) / g.W;
    const long long blocks = (long long)bpe * g.batch;
    PYPDE_REQUIRE(blocks < (1LL << 31), "%s: grid too large", name);
    const int threads = std::min(256, std::max(64, g.W));
    const size_t shmem = (size_t)g.W * g.pitch * sizeof(double);
    hipLaunchKernelGGL(line_panel_kernel<Op>, dim3((unsigned)blocks), dim3(threads), shmem,
                       as_stream(stream), g, op);
    return check_launch(name);
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_stencil_apply(const pypde_lines* in, const double* s, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(s, "null stencil");
    PYPDE_REQUIRE(g.len_out == g.len_in + 2, "stencil apply needs len_out == len_in + 2");
    return launch_lines(g, StencilOp{s, g.len_in}, stream, "pypde_stencil_apply");
}

extern "C" int pypde_stencil_inverse(const pypde_lines* in, const double* s, const double* a,
                                     const double* b, const double* w, const double* den,
                                     void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(s && a && b && w && den, "null table pointer");
    PYPDE_REQUIRE(g.len_in == g.len_out + 2 && g.len_out >= 3,
                  "stencil inverse needs len_in == len_out + 2 >= 5");
    return launch_lines(g, StencilInvOp{s, a, b, w, den, g.len_out}, stream,
                        "pypde_stencil_inverse");
}

extern "C" int pypde_cheb_deriv(const pypde_lines* in, const double* s, int order,
                                double scale_pow, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(order >= 0, "negative derivative order");
    const int n = s ? g.len_in + 2 : g.len_in;
    PYPDE_REQUIRE(g.len_out == n, "derivative output length must be %d", n);
    PYPDE_REQUIRE(n >= 3 || order == 0, "derivative needs n >= 3");
    return launch_lines(g, DerivOp{s, g.len_in, n, order, scale_pow}, stream, "pypde_cheb_deriv");
}

extern "C" int pypde_banded_apply(const pypde_lines* in, int ndiag, const int* offsets_host,
                                  const double* vals, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(ndiag >= 1 && ndiag <= kMaxDiag, "ndiag must be 1..%d", kMaxDiag);
    PYPDE_REQUIRE(vals && offsets_host, "null diagonal pointer");
    BandedOp op{};
    op.vals = vals;
    op.nd = ndiag;
    op.len_in = g.len_in;
    op.len_out = g.len_out;
    for (int q = 0; q < ndiag; ++q) {
        PYPDE_REQUIRE(offsets_host[q] >= -4 && offsets_host[q] <= 4, "offset out of [-4,4]");
        if (q) PYPDE_REQUIRE(offsets_host[q] > offsets_host[q - 1], "offsets must ascend");
        op.off[q] = offsets_host[q];
    }
    return launch_lines(g, op, stream, "pypde_banded_apply");
}

static int fdma_common(const pypde_lines* in, const double* l, const double* d, const double* u1,
                       const double* u2, long long tbs, int transpose, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(l && d && u1 && u2, "null fdma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 5, "fdma needs len_in == len_out >= 5");
    return launch_lines(g, FdmaOp{l, d, u1, u2, tbs, g.len_in, transpose}, stream,
                        transpose ? "pypde_fdma_solve_t" : "pypde_fdma_solve");
}

extern "C" int pypde_fdma_solve(const pypde_lines* in, const double* l, const double* d,
                                const double* u1, const double* u2, long long table_bs,
                                void* stream)
{
    return fdma_common(in, l, d, u1, u2, table_bs, 0, stream);
}

extern "C" int pypde_fdma_solve_t(const pypde_lines* in, const double* l, const double* d,
                                  const double* u1, const double* u2, long long table_bs,
                                  void* stream)
{
    return fdma_common(in, l, d, u1, u2, table_bs, 1, stream);
}

extern "C" int pypde_poisson_solve(const pypde_lines* in, const double* tab, const int* sing,
                                   int transpose, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(tab && sing, "null poisson table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 6, "poisson needs len_in == len_out >= 6");
    return launch_lines(g, PoissonOp{tab, sing, g.len_in, transpose}, stream, "pypde_poisson_solve");
}

extern "C" int pypde_twodma_solve(const pypde_lines* in, const double* d, const double* u,
                                  void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(d && u, "null twodma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 3, "twodma needs len_in == len_out >= 3");
    return launch_lines(g, TwodmaOp{d, u, g.len_in}, stream, "pypde_twodma_solve");
}

extern "C" int pypde_pdma_solve(const pypde_lines* in, const double* a, const double* b,
                                const double* d, const double* e, const double* f, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(a && b && d && e && f, "null pdma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 6, "pdma needs len_in == len_out >= 6");
    return launch_lines(g, PdmaOp{a, b, d, e, f, g.len_in}, stream, "pypde_pdma_solve");
}
