// Line operators along one axis of a batch of row-major 2-D float64 arrays.
//
// Two kernel families:
//
// * Banded applies (stencil S, pseudo-inverse B, B S, ...): no recurrence, so
//   every output element is a sum of at most nine inputs of its line.  A
//   thread computes one output, or a register window of R consecutive outputs
//   of its line, or (along rows) one entry of many rows with that entry's
//   coefficients held in registers.  Neighbouring threads own neighbouring
//   columns (axis 0) or neighbouring entries of a row (axis 1), so accesses
//   coalesce.
//
// * Recurrences (four-diagonal LU sweeps, the offset-2 tridiagonal solve, the
//   Poisson per-line solves, two-/five-band solves, the derivative recurrence):
//   a block stages a panel of W lines into LDS with coalesced accesses (rows of
//   W adjacent columns for axis 0, whole rows for axis 1) together with the
//   operator tables.  Every operator here couples only indices of equal
//   parity (all band offsets are even), so each line is swept by two lanes, one
//   for the even and one for the odd entries, which halves the serial chain.
//   The sweeps load their inputs a few elements ahead of the chain.
//
// The arithmetic of every operator is the reference's, statement by statement,
// with FMA contraction disabled, so results are bit-identical to the
// Fortran/C restatement of the reference.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <algorithm>

namespace pypde {

#pragma clang fp contract(off)

struct LineGeom {
    const double* src;
    double* dst;
    long long src_ld, dst_ld, src_bs, dst_bs;
    int len_in, len_out, nlines, batch, axis;
    int W;      // lines per block
    int pitch;  // LDS doubles per line (odd)
};

// ---------------------------------------------------------------------------
// Banded applies: y_r = sum_q vals[q][r] * x_{r + off_q}  (ascending columns)
// ---------------------------------------------------------------------------
constexpr int kMaxDiag = 9;
struct BandSpec {
    const double* vals;  // [nd][len_out]
    int off[kMaxDiag];
    int nd;
    int stencil;         // 1: stencil form y_i = s_{i-2} x_{i-2} + x_i (vals = s)
    int qmap[kMaxDiag];  // diagonal q at offset off[0] + d (-1: none), for the window forms
};

// Output element r of a line from its input window (ascending columns, zero
// entries skipped as the reference's csr storage does).
template <class Get>
__device__ __forceinline__ double band_row(const BandSpec& bs, int r, int len_in, int len_out,
                                           const Get& X)
{
    if (bs.stencil) {
        if (r >= 2) {
            const double tt = bs.vals[r - 2] * X(-2);
            return (r < len_in) ? tt + X(0) : tt;
        }
        return X(0);
    }
    double acc = 0.0;
    bool first = true;
    for (int q = 0; q < bs.nd; ++q) {
        const int k = r + bs.off[q];
        if (k < 0 || k >= len_in) continue;
        const double a = bs.vals[(long long)q * len_out + r];
        if (a == 0.0) continue;
        const double tt = a * X(bs.off[q]);
        acc = first ? tt : acc + tt;
        first = false;
    }
    return acc;
}

// Axis 0: one thread per column and `rows` consecutive output rows (rows =
// g.W, 1..8, the lines_per_block hint; MULTI = false is the one-row form);
// grid.x over columns (neighbouring threads read neighbouring columns), grid.y
// over row groups, grid.z over batch entries.  Rows shared by neighbouring
// outputs are served from cache.
constexpr int kMaxRowsPerThread = 8;

template <bool MULTI>
__global__ void __launch_bounds__(256) banded_axis0_kernel(LineGeom g, BandSpec bs)
{
    const int rows = MULTI ? g.W : 1;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * rows;
    const int b = blockIdx.z;
    if (c >= g.nlines) return;
    const double* __restrict__ col = g.src + (long long)b * g.src_bs + c;
    double* __restrict__ out = g.dst + (long long)b * g.dst_bs + c;
    const long long ls = g.src_ld;
    if (!MULTI) {
        const double y = band_row(bs, r0, g.len_in, g.len_out,
                                  [&](int o) { return col[(long long)(r0 + o) * ls]; });
        out[(long long)r0 * g.dst_ld] = y;
        return;
    }
    for (int j = 0; j < rows; ++j) {
        const int r = r0 + j;
        if (r >= g.len_out) break;
        const double y = band_row(bs, r, g.len_in, g.len_out,
                                  [&](int o) { return col[(long long)(r + o) * ls]; });
        out[(long long)r * g.dst_ld] = y;
    }
}

// Axis 1: one thread per output element of a row, `rows` rows per block
// (rows = g.W, 1..8; MULTI = false is the one-row form); grid.x over the row,
// grid.y over row groups, grid.z over batch entries.
template <bool MULTI>
__global__ void __launch_bounds__(256) banded_axis1_kernel(LineGeom g, BandSpec bs)
{
    const int rows = MULTI ? g.W : 1;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.z;
    if (r >= g.len_out) return;
    if (!MULTI) {
        const int c = blockIdx.y;
        const double* __restrict__ row = g.src + (long long)b * g.src_bs + (long long)c * g.src_ld;
        const double y = band_row(bs, r, g.len_in, g.len_out, [&](int o) { return row[r + o]; });
        g.dst[(long long)b * g.dst_bs + (long long)c * g.dst_ld + r] = y;
        return;
    }
    for (int j = 0; j < rows; ++j) {
        const int c = blockIdx.y * rows + j;
        if (c >= g.nlines) break;
        const double* __restrict__ row = g.src + (long long)b * g.src_bs + (long long)c * g.src_ld;
        const double y = band_row(bs, r, g.len_in, g.len_out, [&](int o) { return row[r + o]; });
        g.dst[(long long)b * g.dst_bs + (long long)c * g.dst_ld + r] = y;
    }
}

// ---------------------------------------------------------------------------
// Recurrence panel kernel
// ---------------------------------------------------------------------------
// Op interface:
//   static constexpr int kTables;             doubles of shared table per entry (x n)
//   __device__ void load_tables(double* t, int b, int tid, int nt) const;
//   __device__ void operator()(double* x, const double* t, int parity, int b, int line) const;
// `t` points at the LDS copy of the tables of entry b.
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
    double* tab = lds + g.W * g.pitch;
    op.load_tables(tab, b, tid, nt);

    if (g.axis == 0) {
        const int rows_per_pass = nt / g.W;
        const int c = tid % g.W;
        for (int k = tid / g.W; k < g.len_in; k += rows_per_pass)
            if (c < nl) lds[c * g.pitch + k] = src[(long long)k * g.src_ld + l0 + c];
    } else {
        for (int c = 0; c < nl; ++c) {
            const double* row = src + (long long)(l0 + c) * g.src_ld;
            for (int k = tid; k < g.len_in; k += nt) lds[c * g.pitch + k] = row[k];
        }
    }
    __syncthreads();
    for (int u = tid; u < 2 * nl; u += nt) op(lds + (u >> 1) * g.pitch, tab, u & 1, b, l0 + (u >> 1));
    __syncthreads();
    if (g.axis == 0) {
        const int rows_per_pass = nt / g.W;
        const int c = tid % g.W;
        for (int k = tid / g.W; k < g.len_out; k += rows_per_pass)
            if (c < nl) dst[(long long)k * g.dst_ld + l0 + c] = lds[c * g.pitch + k];
    } else {
        for (int c = 0; c < nl; ++c) {
            double* row = dst + (long long)(l0 + c) * g.dst_ld;
            for (int k = tid; k < g.len_out; k += nt) row[k] = lds[c * g.pitch + k];
        }
    }
}

__device__ __forceinline__ void copy_tab(double* dst, const double* src, int n, int tid, int nt)
{
    for (int i = tid; i < n; i += nt) dst[i] = src[i];
}

// Largest index <= top with the given parity.
__device__ __forceinline__ int top_of(int top, int parity) { return ((top - parity) & 1) ? top - 1 : top; }

// ---- four-diagonal sweeps (fdma.f90:1-38) on the entries of one parity ----
// Forward: x_i -= l_{i-2} x_{i-2}; backward as in the reference.
__device__ __forceinline__ void fdma_par(double* x, int n, int p, const double* l, const double* d,
                                         const double* u1, const double* u2)
{
    // forward elimination along i = p+2, p+4, ...
    {
        double prev = x[p];
        int i = p + 2;
        for (; i + 6 < n; i += 8) {
            const double x0 = x[i], x1 = x[i + 2], x2 = x[i + 4], x3 = x[i + 6];
            const double l0 = l[i - 2], l1 = l[i], l2 = l[i + 2], l3 = l[i + 4];
            const double y0 = x0 - l0 * prev;
            const double y1 = x1 - l1 * y0;
            const double y2 = x2 - l2 * y1;
            const double y3 = x3 - l3 * y2;
            x[i] = y0; x[i + 2] = y1; x[i + 4] = y2; x[i + 6] = y3;
            prev = y3;
        }
        for (; i < n; i += 2) {
            const double y = x[i] - l[i - 2] * prev;
            x[i] = y;
            prev = y;
        }
    }
    // backward substitution
    const int t = top_of(n - 1, p);
    double a1 = x[t] / d[t];                                   // x_{i+2}
    x[t] = a1;
    double a0 = (x[t - 2] - u1[t - 2] * a1) / d[t - 2];        // x_{i+2} for next, a1 -> x_{i+4}
    x[t - 2] = a0;
    double x2 = a0, x4 = a1;
    int i = t - 4;
    for (; i - 6 >= 0; i -= 8) {
        const double b0 = x[i], b1 = x[i - 2], b2 = x[i - 4], b3 = x[i - 6];
        const double c0 = u1[i], c1 = u1[i - 2], c2 = u1[i - 4], c3 = u1[i - 6];
        const double e0 = u2[i], e1 = u2[i - 2], e2 = u2[i - 4], e3 = u2[i - 6];
        const double d0 = d[i], d1 = d[i - 2], d2 = d[i - 4], d3 = d[i - 6];
        const double y0 = (b0 - c0 * x2 - e0 * x4) / d0;
        const double y1 = (b1 - c1 * y0 - e1 * x2) / d1;
        const double y2 = (b2 - c2 * y1 - e2 * y0) / d2;
        const double y3 = (b3 - c3 * y2 - e3 * y1) / d3;
        x[i] = y0; x[i - 2] = y1; x[i - 4] = y2; x[i - 6] = y3;
        x4 = y2;
        x2 = y3;
    }
    for (; i >= 0; i -= 2) {
        const double y = (x[i] - u1[i] * x2 - u2[i] * x4) / d[i];
        x[i] = y;
        x4 = x2;
        x2 = y;
    }
}

// Transposed: U^T z = b ascending, then L^T x = z descending.
__device__ __forceinline__ void fdma_t_par(double* x, int n, int p, const double* l, const double* d,
                                           const double* u1, const double* u2)
{
    double z2 = 0.0, z4 = 0.0;  // z_{i-2}, z_{i-4}
    for (int i = p; i < n; i += 2) {
        double v = x[i];
        if (i >= 2) v = v - u1[i - 2] * z2;
        if (i >= 4) v = v - u2[i - 4] * z4;
        v = v / d[i];
        x[i] = v;
        z4 = z2;
        z2 = v;
    }
    const int t = top_of(n - 1, p);
    double xp = x[t];
    for (int j = t - 2; j >= 0; j -= 2) {
        const double v = x[j] - l[j] * xp;
        x[j] = v;
        xp = v;
    }
}

struct FdmaOp {
    const double *l, *d, *u1, *u2;
    long long tbs;
    int n;
    int transpose;
    __device__ void load_tables(double* t, int b, int tid, int nt) const
    {
        const long long o = (long long)b * tbs;
        copy_tab(t, l + o, n - 2, tid, nt);
        copy_tab(t + n, d + o, n, tid, nt);
        copy_tab(t + 2 * n, u1 + o, n - 2, tid, nt);
        copy_tab(t + 3 * n, u2 + o, n - 4, tid, nt);
    }
    __device__ void operator()(double* x, const double* t, int p, int, int) const
    {
        if (transpose) fdma_t_par(x, n, p, t, t + n, t + 2 * n, t + 3 * n);
        else fdma_par(x, n, p, t, t + n, t + 2 * n, t + 3 * n);
    }
};

// Poisson per-line tables (d | u1 | u2 | l), read from global memory.
struct PoissonOp {
    const double* tab;
    const int* sing;
    int n;
    int transpose;
    __device__ void load_tables(double*, int, int, int) const {}
    __device__ void operator()(double* x, const double*, int p, int, int line) const
    {
        const double* t = tab + (long long)line * 4 * n;
        const double *d = t, *u1 = t + n, *u2 = t + 2 * n, *l = t + 3 * n;
        if (sing[line]) {
            if (transpose) fdma_t_par(x + 1, n - 1, p, l, d, u1, u2);
            else fdma_par(x + 1, n - 1, p, l, d, u1, u2);
            if (p == 1) x[0] = 0.0;
        } else {
            if (transpose) fdma_t_par(x, n, p, l, d, u1, u2);
            else fdma_par(x, n, p, l, d, u1, u2);
        }
    }
};

// v = (S^T S)^-1 S^T u, offset-2 tridiagonal solve (tdma.f90:56-107).
struct StencilInvOp {
    const double *s, *a, *b, *w, *den;
    int M;
    __device__ void load_tables(double* t, int, int tid, int nt) const
    {
        copy_tab(t, s, M, tid, nt);
        copy_tab(t + M, a, M - 2, tid, nt);
        copy_tab(t + 2 * M, b, M, tid, nt);
        copy_tab(t + 3 * M, w, M - 2, tid, nt);
        copy_tab(t + 4 * M, den, M, tid, nt);
    }
    __device__ void operator()(double* x, const double* t, int p, int, int) const
    {
        const double *s_ = t, *a_ = t + M, *b_ = t + 2 * M, *w_ = t + 3 * M, *den_ = t + 4 * M;
        // y_k = u_k + s_k u_{k+2}   (csc column order of S^T)
        for (int k = p; k < M; k += 2) x[k] = x[k] + s_[k] * x[k + 2];
        // forward sweep g
        double g = x[p] / b_[p];
        x[p] = g;
        for (int i = p + 2; i < M; i += 2) {
            g = (x[i] - a_[i - 2] * g) / den_[i];
            x[i] = g;
        }
        // back substitution x_j = g_j - w_j x_{j+2}
        const int t0 = top_of(M - 1, p);
        double xp = x[t0];
        for (int j = t0 - 2; j >= 0; j -= 2) {
            xp = x[j] - w_[j] * xp;
            x[j] = xp;
        }
    }
};

// Chebyshev derivative recurrence (differentiate_cheby.f90:28-53), `order`
// times, optionally after the stencil, then / scale_pow.  The recurrence reads
// the other parity's input, so one lane does the whole line.
__device__ __forceinline__ void cheb_deriv_inplace(double* x, int n)
{
    double cn = x[n - 1];
    x[n - 1] = 0.0;
    double dp2 = 0.0, dp1;
    {
        const double cj = x[n - 2];
        const double dj = (double)(2 * (n - 1)) * cn;
        x[n - 2] = dj;
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
    x[0] = dp2 / 2.0 + cn;
}

struct DerivOp {
    const double* s;
    int len_in, n, order;
    double scale_pow;
    __device__ void load_tables(double* t, int, int tid, int nt) const
    {
        if (s) copy_tab(t, s, len_in, tid, nt);
    }
    __device__ void operator()(double* x, const double* t, int p, int, int) const
    {
        if (p) return;
        if (s) {
            const int M = len_in, N = M + 2;
            for (int i = N - 1; i >= 0; --i) {
                double y;
                if (i >= 2) {
                    const double tt = t[i - 2] * x[i - 2];
                    y = (i < M) ? tt + x[i] : tt;
                } else {
                    y = x[i];
                }
                x[i] = y;
            }
        }
        for (int o = 0; o < order; ++o) cheb_deriv_inplace(x, n);
        for (int i = 0; i < n; ++i) x[i] = x[i] / scale_pow;
    }
};

// Upper two-band back substitution (twodma.f90:2-25).
struct TwodmaOp {
    const double *d, *u;
    int n;
    __device__ void load_tables(double* t, int, int tid, int nt) const
    {
        copy_tab(t, d, n, tid, nt);
        copy_tab(t + n, u, n - 2, tid, nt);
    }
    __device__ void operator()(double* x, const double* t, int p, int, int) const
    {
        const double *d_ = t, *u_ = t + n;
        const int t0 = top_of(n - 1, p);
        double xp = x[t0] / d_[t0];
        x[t0] = xp;
        for (int i = t0 - 2; i >= 0; i -= 2) {
            xp = (x[i] - u_[i] * xp) / d_[i];
            x[i] = xp;
        }
    }
};

// PDMA_Solve (pdma.py:74-90) with PDMA_LU tables.
struct PdmaOp {
    const double *a, *b, *d, *e, *f;
    int n;
    __device__ void load_tables(double* t, int, int tid, int nt) const
    {
        copy_tab(t, a, n - 4, tid, nt);
        copy_tab(t + n, b, n - 2, tid, nt);
        copy_tab(t + 2 * n, d, n, tid, nt);
        copy_tab(t + 3 * n, e, n - 2, tid, nt);
        copy_tab(t + 4 * n, f, n - 4, tid, nt);
    }
    __device__ void operator()(double* x, const double* t, int p, int, int) const
    {
        const double *a_ = t, *b_ = t + n, *d_ = t + 2 * n, *e_ = t + 3 * n, *f_ = t + 4 * n;
        x[2 + p] -= b_[p] * x[p];
        for (int k = 4 + p; k < n; k += 2) x[k] -= b_[k - 2] * x[k - 2] + a_[k - 4] * x[k - 4];
        const int t0 = top_of(n - 1, p);
        x[t0] /= d_[t0];
        x[t0 - 2] = (x[t0 - 2] - e_[t0 - 2] * x[t0]) / d_[t0 - 2];
        for (int k = t0 - 4; k >= 0; k -= 2) x[k] = (x[k] - e_[k] * x[k + 2] - f_[k] * x[k + 4]) / d_[k];
    }
};

// ------------------------------------------------------------------------
// Launch helpers
// ------------------------------------------------------------------------
constexpr int kLdsBudget = 64 * 1024;

static int pick_lines_per_block(int pitch, int tab_doubles, int hint, int nlines)
{
    int w = 1;
    if (hint > 0) {
        while (w * 2 <= hint && w * 2 <= 128) w *= 2;
    } else {
        while (w * 2 <= 64 && ((long long)(w * 2) * pitch + tab_doubles) * 8 <= kLdsBudget) w *= 2;
    }
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
    g.pitch = std::max(in->len_in, in->len_out) | 1;
    g.W = 1;
    return OK;
}

template <class Op>
static int launch_lines(LineGeom g, const Op& op, int ntab, int hint, void* stream, const char* name)
{
    if (g.nlines == 0 || g.batch == 0) return OK;
    const int tab_doubles = ntab;
    g.W = pick_lines_per_block(g.pitch, tab_doubles, hint, g.nlines);
    const size_t shmem = ((size_t)g.W * g.pitch + tab_doubles) * sizeof(double);
    PYPDE_REQUIRE(shmem <= (size_t)kLdsBudget, "%s: line length %d too long for an LDS panel",
                  name, g.pitch);
    const int bpe = (g.nlines + g.W - 1) / g.W;
    const long long blocks = (long long)bpe * g.batch;
    PYPDE_REQUIRE(blocks < (1LL << 31), "%s: grid too large", name);
    const int threads = std::min(256, std::max(64, 2 * g.W));
    hipLaunchKernelGGL(line_panel_kernel<Op>, dim3((unsigned)blocks), dim3(threads), shmem,
                       as_stream(stream), g, op);
    return check_launch(name);
}

// rows per thread (axis 0) / per block (axis 1): the lines_per_block hint, 1..8
// (0 = 1), chosen per shape by the host-side autotuner.
// Register-window forms: a thread owns R consecutive outputs of its line and
// loads the R + span inputs they share once (span = last - first offset <= 8),
// where the forms above re-read them for every output.  The sum runs over the
// same terms in the same order (ascending offsets; terms outside the input
// and zero coefficients skipped; the stencil's s x_{i-2} + x_i), so results
// are bit-identical.  Axis 0: a thread per column and R rows; axis 1: a thread
// per R consecutive entries of a row.
template <int R, bool STENCIL>
__device__ __forceinline__ void window_rows(const BandSpec& bs, int r0, int len_in, int len_out,
                                            const double (&w)[R + 8], double (&y)[R])
{
    const int o0 = STENCIL ? -2 : bs.off[0];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = r0 + i;
        double acc = 0.0;
        bool first = true;
#pragma unroll
        for (int d = 0; d < kMaxDiag; ++d) {
            double a;
            if (STENCIL) {
                if (d != 0 && d != 2) continue;
                if (d == 0) {
                    if (r < 2) continue;
                    a = bs.vals[r - 2];
                } else {
                    a = 1.0;
                }
            } else {
                const int q = bs.qmap[d];
                if (q < 0 || r >= len_out) continue;
                a = bs.vals[(long long)q * len_out + r];
            }
            const int k = r + o0 + d;
            if (k < 0 || k >= len_in) continue;
            if (!STENCIL && a == 0.0) continue;
            const double tt = (STENCIL && d == 2) ? w[i + d] : a * w[i + d];
            acc = first ? tt : acc + tt;
            first = false;
        }
        y[i] = acc;
    }
}

template <int R, bool STENCIL>
__global__ void __launch_bounds__(256) banded_axis0_win_kernel(LineGeom g, BandSpec bs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * R;
    const int b = blockIdx.z;
    if (c >= g.nlines) return;
    const double* __restrict__ col = g.src + (long long)b * g.src_bs + c;
    double* __restrict__ out = g.dst + (long long)b * g.dst_bs + c;
    const int o0 = STENCIL ? -2 : bs.off[0];
    const int span = STENCIL ? 2 : bs.off[bs.nd - 1] - o0;
    double w[R + 8];
#pragma unroll
    for (int j = 0; j < R + 8; ++j) {
        const int k = r0 + o0 + j;
        w[j] = (j <= R - 1 + span && k >= 0 && k < g.len_in) ? col[(long long)k * g.src_ld] : 0.0;
    }
    double y[R];
    window_rows<R, STENCIL>(bs, r0, g.len_in, g.len_out, w, y);
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (r0 + i < g.len_out) out[(long long)(r0 + i) * g.dst_ld] = y[i];
}

template <int R, bool STENCIL>
__global__ void __launch_bounds__(256) banded_axis1_win_kernel(LineGeom g, BandSpec bs)
{
    // threads over (row, chunk of R entries), chunks of a row adjacent
    const int nch = (g.len_out + R - 1) / R;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nch * g.nlines) return;
    const int c = (int)(t / nch);
    const int r0 = (int)(t - (long long)c * nch) * R;
    const int b = blockIdx.z;
    const double* __restrict__ row = g.src + (long long)b * g.src_bs + (long long)c * g.src_ld;
    double* __restrict__ out = g.dst + (long long)b * g.dst_bs + (long long)c * g.dst_ld;
    const int o0 = STENCIL ? -2 : bs.off[0];
    const int span = STENCIL ? 2 : bs.off[bs.nd - 1] - o0;
    double w[R + 8];
#pragma unroll
    for (int j = 0; j < R + 8; ++j) {
        const int k = r0 + o0 + j;
        w[j] = (j <= R - 1 + span && k >= 0 && k < g.len_in) ? row[k] : 0.0;
    }
    double y[R];
    window_rows<R, STENCIL>(bs, r0, g.len_in, g.len_out, w, y);
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (r0 + i < g.len_out) out[r0 + i] = y[i];
}

// Axis 1, coefficient-resident form: a thread owns entry r of the output rows,
// keeps the coefficients of output r (and which terms exist) in registers, and
// runs over g.W rows, U at a time with their loads issued together.  Each
// load of a wave is 64 consecutive entries of a row shifted by one band
// offset, so all accesses coalesce and the shifted copies hit in cache.  The
// sum is band_row's, term for term.  NDT: the number of diagonals when it is
// known at compile time (3), else kMaxDiag.  Every load is unconditional: a
// term that does not exist for this output (its column outside the input)
// reads a clamped column and is left out of the sum, so the loads are not
// placed behind per-lane branches that would wait for each in turn.
template <int U, bool STENCIL, int NDT>
__global__ void __launch_bounds__(256) banded_axis1_res_kernel(LineGeom g, BandSpec bs)
{
    constexpr int ND = STENCIL ? 2 : NDT;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.z;
    if (r >= g.len_out) return;
    double a[ND];
    int off[ND];  // column offset of the term's load, clamped into the input row
    bool use[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        int k;
        if (STENCIL) {
            // term 0: s_{r-2} x_{r-2} (r >= 2); term 1: x_r (r < len_in)
            k = r + (q == 0 ? -2 : 0);
            use[q] = q == 0 ? r >= 2 : r < g.len_in;
            a[q] = (q == 0 && use[q]) ? bs.vals[r - 2] : 1.0;
        } else {
            const bool present = NDT < kMaxDiag || q < bs.nd;
            k = r + (present ? bs.off[q] : 0);
            use[q] = present && k >= 0 && k < g.len_in;
            a[q] = use[q] ? bs.vals[(long long)q * g.len_out + r] : 0.0;
            use[q] = use[q] && a[q] != 0.0;
        }
        off[q] = min(max(k, 0), g.len_in - 1) - r;
    }
    const int c0 = blockIdx.y * g.W;
    const int c1 = min(c0 + g.W, g.nlines);
    const double* __restrict__ src = g.src + (long long)b * g.src_bs + r;
    double* __restrict__ dst = g.dst + (long long)b * g.dst_bs + r;
    for (int c = c0; c < c1; c += U) {
        double x[U][ND];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double* row = src + (long long)min(c + u, c1 - 1) * g.src_ld;
#pragma unroll
            for (int q = 0; q < ND; ++q) x[u][q] = row[off[q]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (c + u >= c1) break;
            double acc = 0.0;
            bool first = true;
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                if (!use[q]) continue;
                const double tt = (STENCIL && q == 1) ? x[u][q] : a[q] * x[u][q];
                acc = first ? tt : acc + tt;
                first = false;
            }
            dst[(long long)(c + u) * g.dst_ld] = acc;
        }
    }
}

template <int R>
static void launch_window(const LineGeom& g, const BandSpec& bs, void* stream)
{
    if (g.axis == 0) {
        dim3 grid((g.nlines + 255) / 256, (g.len_out + R - 1) / R, g.batch);
        if (bs.stencil) hipLaunchKernelGGL((banded_axis0_win_kernel<R, true>), grid, dim3(256), 0, as_stream(stream), g, bs);
        else hipLaunchKernelGGL((banded_axis0_win_kernel<R, false>), grid, dim3(256), 0, as_stream(stream), g, bs);
    } else {
        const long long threads = (long long)((g.len_out + R - 1) / R) * g.nlines;
        dim3 grid((unsigned)((threads + 255) / 256), 1, g.batch);
        if (bs.stencil) hipLaunchKernelGGL((banded_axis1_win_kernel<R, true>), grid, dim3(256), 0, as_stream(stream), g, bs);
        else hipLaunchKernelGGL((banded_axis1_win_kernel<R, false>), grid, dim3(256), 0, as_stream(stream), g, bs);
    }
}

static int launch_banded(LineGeom g, const BandSpec& bs, int hint, void* stream, const char* name)
{
    if ((long long)g.len_out * g.nlines * g.batch == 0) return OK;
    if (hint >= 33 && hint <= 40) {
        // coefficient-resident form along rows: 33..36 four rows in flight and
        // 4, 8, 16, 32 rows per block; 37..40 eight rows and 8, 16, 32, 64
        PYPDE_REQUIRE(g.axis == 1, "%s: row hints 33..40 are for axis 1", name);
        const bool u8 = hint >= 37;
        g.W = (u8 ? 8 : 4) << ((hint - 33) & 3);
        PYPDE_REQUIRE(g.batch <= 65535 && (g.nlines + g.W - 1) / g.W <= 65535,
                      "%s: grid too large", name);
        dim3 grid((g.len_out + 255) / 256, (g.nlines + g.W - 1) / g.W, g.batch);
        hipStream_t st = as_stream(stream);
        if (bs.stencil) {
            if (u8) hipLaunchKernelGGL((banded_axis1_res_kernel<8, true, 2>), grid, dim3(256), 0, st, g, bs);
            else hipLaunchKernelGGL((banded_axis1_res_kernel<4, true, 2>), grid, dim3(256), 0, st, g, bs);
        } else if (bs.nd == 3) {
            if (u8) hipLaunchKernelGGL((banded_axis1_res_kernel<8, false, 3>), grid, dim3(256), 0, st, g, bs);
            else hipLaunchKernelGGL((banded_axis1_res_kernel<4, false, 3>), grid, dim3(256), 0, st, g, bs);
        } else {
            if (u8) hipLaunchKernelGGL((banded_axis1_res_kernel<8, false, kMaxDiag>), grid, dim3(256), 0, st, g, bs);
            else hipLaunchKernelGGL((banded_axis1_res_kernel<4, false, kMaxDiag>), grid, dim3(256), 0, st, g, bs);
        }
        return check_launch(name);
    }
    if (hint == 18 || hint == 20 || hint == 24) {
        // register-window forms, R = hint - 16 outputs per thread
        PYPDE_REQUIRE(g.batch <= 65535 && g.len_out <= 65535 * 256 && g.nlines <= 65535 * 256 &&
                          (long long)g.len_out * g.nlines < (1LL << 38),
                      "%s: grid too large", name);
        if (hint == 18) launch_window<2>(g, bs, stream);
        else if (hint == 20) launch_window<4>(g, bs, stream);
        else launch_window<8>(g, bs, stream);
        return check_launch(name);
    }
    g.W = std::min(kMaxRowsPerThread, std::max(1, hint));
    const int kRowsPerThread = g.W, kRowsPerBlock1 = g.W;
    if (g.axis == 0) {
        PYPDE_REQUIRE(g.batch <= 65535 && (g.len_out + kRowsPerThread - 1) / kRowsPerThread <= 65535,
                      "%s: grid too large", name);
        dim3 grid((g.nlines + 255) / 256, (g.len_out + kRowsPerThread - 1) / kRowsPerThread, g.batch);
        if (g.W == 1) hipLaunchKernelGGL(banded_axis0_kernel<false>, grid, dim3(256), 0, as_stream(stream), g, bs);
        else hipLaunchKernelGGL(banded_axis0_kernel<true>, grid, dim3(256), 0, as_stream(stream), g, bs);
    } else {
        PYPDE_REQUIRE(g.batch <= 65535 && (g.nlines + kRowsPerBlock1 - 1) / kRowsPerBlock1 <= 65535,
                      "%s: grid too large", name);
        dim3 grid((g.len_out + 255) / 256, (g.nlines + kRowsPerBlock1 - 1) / kRowsPerBlock1, g.batch);
        if (g.W == 1) hipLaunchKernelGGL(banded_axis1_kernel<false>, grid, dim3(256), 0, as_stream(stream), g, bs);
        else hipLaunchKernelGGL(banded_axis1_kernel<true>, grid, dim3(256), 0, as_stream(stream), g, bs);
    }
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
    PYPDE_REQUIRE(g.src != g.dst, "stencil apply cannot run in place");
    BandSpec bs{};
    bs.vals = s;
    bs.stencil = 1;
    return launch_banded(g, bs, in->lines_per_block, stream, "pypde_stencil_apply");
}

extern "C" int pypde_banded_apply(const pypde_lines* in, int ndiag, const int* offsets_host,
                                  const double* vals, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(ndiag >= 1 && ndiag <= kMaxDiag, "ndiag must be 1..%d", kMaxDiag);
    PYPDE_REQUIRE(vals && offsets_host, "null diagonal pointer");
    PYPDE_REQUIRE(g.src != g.dst, "banded apply cannot run in place");
    BandSpec bs{};
    bs.vals = vals;
    bs.nd = ndiag;
    for (int d = 0; d < kMaxDiag; ++d) bs.qmap[d] = -1;
    for (int q = 0; q < ndiag; ++q) {
        PYPDE_REQUIRE(offsets_host[q] >= -4 && offsets_host[q] <= 4, "offset out of [-4,4]");
        if (q) PYPDE_REQUIRE(offsets_host[q] > offsets_host[q - 1], "offsets must ascend");
        bs.off[q] = offsets_host[q];
        bs.qmap[offsets_host[q] - offsets_host[0]] = q;
    }
    return launch_banded(g, bs, in->lines_per_block, stream, "pypde_banded_apply");
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
    const int M = g.len_out;
    return launch_lines(g, StencilInvOp{s, a, b, w, den, M}, 5 * M, in->lines_per_block, stream,
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
    return launch_lines(g, DerivOp{s, g.len_in, n, order, scale_pow}, s ? g.len_in : 0,
                        in->lines_per_block, stream, "pypde_cheb_deriv");
}

static int fdma_common(const pypde_lines* in, const double* l, const double* d, const double* u1,
                       const double* u2, long long tbs, int transpose, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(l && d && u1 && u2, "null fdma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 6, "fdma needs len_in == len_out >= 6");
    const int n = g.len_in;
    return launch_lines(g, FdmaOp{l, d, u1, u2, tbs, n, transpose}, 4 * n, in->lines_per_block,
                        stream, transpose ? "pypde_fdma_solve_t" : "pypde_fdma_solve");
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
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 7, "poisson needs len_in == len_out >= 7");
    return launch_lines(g, PoissonOp{tab, sing, g.len_in, transpose}, 0, in->lines_per_block, stream,
                        "pypde_poisson_solve");
}

extern "C" int pypde_twodma_solve(const pypde_lines* in, const double* d, const double* u,
                                  void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(d && u, "null twodma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 3, "twodma needs len_in == len_out >= 3");
    return launch_lines(g, TwodmaOp{d, u, g.len_in}, 2 * g.len_in, in->lines_per_block, stream,
                        "pypde_twodma_solve");
}

extern "C" int pypde_pdma_solve(const pypde_lines* in, const double* a, const double* b,
                                const double* d, const double* e, const double* f, void* stream)
{
    LineGeom g;
    int rc = make_geom(in, g);
    if (rc) return rc;
    PYPDE_REQUIRE(a && b && d && e && f, "null pdma table");
    PYPDE_REQUIRE(g.len_in == g.len_out && g.len_in >= 6, "pdma needs len_in == len_out >= 6");
    return launch_lines(g, PdmaOp{a, b, d, e, f, g.len_in}, 5 * g.len_in, in->lines_per_block,
                        stream, "pypde_pdma_solve");
}
