/*
 * ORACLE (test infrastructure only): C restatement of the reference's four
 * f2py Fortran kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library; the product path never does.
 *
 * Arrays are C-contiguous numpy arrays of shape (n, m); "axis 0" means the
 * first numpy index, exactly as the f2py wrappers expose the Fortran x(n,m).
 * Arithmetic follows the Fortran statement by statement (same operation order,
 * no FMA contraction: build with -ffp-contract=off).
 *
 *   diff_2d          pypde/bases/fortran/differentiate_cheby.f90:28-53
 *   solve_tdma_2d    pypde/bases/linalg/fortran/tdma.f90:56-107
 *   solve_fdma_1d    pypde/solver/linalg/fortran/fdma.f90:1-38
 *   solve_fdma_2d    pypde/solver/linalg/fortran/fdma.f90:40-98
 *   init_fdma        pypde/solver/linalg/fortran/fdma.f90:102-143
 *   solve_fdma_type2 pypde/solver/linalg/fortran/fdma.f90:146-195
 *   solve_twodma_2d  pypde/solver/linalg/fortran/twodma.f90:28-62
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define X(i, j) x[(size_t)(i) * (size_t)m + (size_t)(j)]

/* differentiate_cheby.f90:28-53 -- derivative recurrence along axis 0.
 * dc is fully written (the f2py intent(out) array is zero-filled, so the last
 * row stays 0). */
void ora_diff_2d(const double* c, double* dc, int n, int m)
{
    memset(dc, 0, sizeof(double) * (size_t)n * (size_t)m);
    if (n < 2) return;
    for (int j = 0; j < m; ++j)
        dc[(size_t)(n - 2) * m + j] = (double)(2 * (n - 1)) * c[(size_t)(n - 1) * m + j];
    /* Fortran i = n-2 .. 2 (1-based): dc(i) = dc(i+2) + 2*i*c(i+1) */
    for (int i = n - 2; i >= 2; --i) {
        const int r = i - 1;
        for (int j = 0; j < m; ++j)
            dc[(size_t)r * m + j] = dc[(size_t)(r + 2) * m + j] + (double)(2 * i) * c[(size_t)(r + 1) * m + j];
    }
    if (n >= 3) {
        for (int j = 0; j < m; ++j)
            dc[j] = dc[2 * (size_t)m + j] / 2.0 + c[(size_t)m + j];
    } else {
        for (int j = 0; j < m; ++j) dc[j] = c[(size_t)m + j];
    }
}

/* tdma.f90:56-107 -- offset-k tridiagonal solve along axis 0. */
void ora_tdma_2d(const double* a, const double* b, const double* c, const double* d,
                 int k, double* x, int n, int m)
{
    double* w = (double*)malloc(sizeof(double) * (size_t)(n > k ? n - k : 1));
    double* g = (double*)malloc(sizeof(double) * (size_t)n * (size_t)m);
    /* Fortran i = 1..n-k */
    for (int i = 1; i <= n - k; ++i) {
        if (i < k + 1) w[i - 1] = c[i - 1] / b[i - 1];
        else           w[i - 1] = c[i - 1] / (b[i - 1] - a[i - k - 1] * w[i - k - 1]);
    }
    for (int i = 1; i <= n; ++i) {
        for (int j = 0; j < m; ++j) {
            if (i < k + 1) g[(size_t)(i - 1) * m + j] = d[(size_t)(i - 1) * m + j] / b[i - 1];
            else g[(size_t)(i - 1) * m + j] =
                (d[(size_t)(i - 1) * m + j] - a[i - k - 1] * g[(size_t)(i - k - 1) * m + j]) /
                (b[i - 1] - a[i - k - 1] * w[i - k - 1]);
        }
    }
    /* x(n-k:n) = g(n-k:n) */
    for (int i = n - k; i <= n; ++i)
        for (int j = 0; j < m; ++j) X(i - 1, j) = g[(size_t)(i - 1) * m + j];
    /* do i = n-k+1, 2, -1: x(i-1) = g(i-1) - w(i-1)*x(i+k-1) */
    for (int i = n - k + 1; i >= 2; --i)
        for (int j = 0; j < m; ++j)
            X(i - 2, j) = g[(size_t)(i - 2) * m + j] - w[i - 2] * X(i + k - 2, j);
    free(w);
    free(g);
}

/* fdma.f90:1-38 on a strided line (stride in elements). */
static void fdma_line(const double* l, const double* d, const double* u1, const double* u2,
                      double* x, int n, size_t s)
{
#define L1(i) x[(size_t)((i) - 1) * s]
    for (int i = 3; i <= n; ++i) L1(i) = L1(i) - l[i - 3] * L1(i - 2);
    L1(n) = L1(n) / d[n - 1];
    L1(n - 1) = L1(n - 1) / d[n - 2];
    L1(n - 2) = (L1(n - 2) - u1[n - 3] * L1(n)) / d[n - 3];
    L1(n - 3) = (L1(n - 3) - u1[n - 4] * L1(n - 1)) / d[n - 4];
    for (int i = n - 4; i >= 1; --i)
        L1(i) = (L1(i) - u1[i - 1] * L1(i + 2) - u2[i - 1] * L1(i + 4)) / d[i - 1];
#undef L1
}

void ora_fdma_1d(const double* l, const double* d, const double* u1, const double* u2,
                 double* x, int n)
{
    fdma_line(l, d, u1, u2, x, n, 1);
}

/* fdma.f90:40-98 -- pre-factored four-diagonal solve along axis 0 or 1. */
void ora_fdma_2d(const double* l, const double* d, const double* u1, const double* u2,
                 double* x, int axis, int n, int m)
{
    if (axis == 0) {
        for (int j = 0; j < m; ++j) fdma_line(l, d, u1, u2, x + j, n, (size_t)m);
    } else {
        for (int i = 0; i < n; ++i) fdma_line(l, d, u1, u2, x + (size_t)i * m, m, 1);
    }
}

/* fdma.f90:102-143 -- extract diagonals -2,0,2,4 of a row-major n x n matrix
 * and run the no-pivot LU sweep.  A is indexed A(i,j) = A[(i-1)*lda + (j-1)]. */
static void init_fdma_ld(const double* A, size_t lda, double* d, double* u1, double* u2,
                         double* l, int n)
{
    for (int i = 0; i < n; ++i) d[i] = 0.0;
    for (int i = 0; i < n - 2; ++i) { u1[i] = 0.0; l[i] = 0.0; }
    for (int i = 0; i < n - 4; ++i) u2[i] = 0.0;
#define AA(i, j) A[(size_t)((i) - 1) * lda + (size_t)((j) - 1)]
    for (int i = 1; i <= n; ++i) {
        d[i - 1] = AA(i, i);
        if (i > 2) l[i - 3] = AA(i, i - 2);
        if (i < n - 1) u1[i - 1] = AA(i, i + 2);
        if (i < n - 3) u2[i - 1] = AA(i, i + 4);
    }
#undef AA
    for (int i = 3; i <= n; ++i) {
        l[i - 3] = l[i - 3] / d[i - 3];
        d[i - 1] = d[i - 1] - l[i - 3] * u1[i - 3];
        if (i < n - 1) u1[i - 1] = u1[i - 1] - l[i - 3] * u2[i - 3];
    }
}

void ora_init_fdma(const double* A, double* d, double* u1, double* u2, double* l, int n)
{
    init_fdma_ld(A, (size_t)n, d, u1, u2, l, n);
}

/* fdma.f90:146-195 -- for every line i form the dense (A + C*lam(i)), factor it
 * and solve; the singular branch (|lam|<1e-10) solves the trailing (n-1) block
 * and sets x(1,i)=0.  A and C are row-major N x N with N the solve length. */
void ora_fdma_type2(const double* A, const double* C, const double* lam, double* x,
                    int axis, int singular, int n, int m)
{
    const int len = (axis == 0) ? n : m;
    const int nlines = (axis == 0) ? m : n;
    const size_t s = (axis == 0) ? (size_t)m : 1;
    double* T = (double*)malloc(sizeof(double) * (size_t)len * (size_t)len);
    double* d = (double*)malloc(sizeof(double) * (size_t)len);
    double* u1 = (double*)malloc(sizeof(double) * (size_t)len);
    double* u2 = (double*)malloc(sizeof(double) * (size_t)len);
    double* l = (double*)malloc(sizeof(double) * (size_t)len);
    for (int i = 0; i < nlines; ++i) {
        double* line = (axis == 0) ? x + i : x + (size_t)i * m;
        /* the singular branch exists only for axis 0 in the reference */
        if (axis == 0 && singular && fabs(lam[i]) < 1e-10) {
            const int nn = len - 1;
            for (int r = 0; r < nn; ++r)
                for (int c = 0; c < nn; ++c) {
                    const size_t src = (size_t)(r + 1) * len + (size_t)(c + 1);
                    T[(size_t)r * nn + c] = A[src] + C[src] * lam[i];
                }
            init_fdma_ld(T, (size_t)nn, d, u1, u2, l, nn);
            fdma_line(l, d, u1, u2, line + s, nn, s);
            line[0] = 0.0;
        } else {
            for (size_t e = 0; e < (size_t)len * len; ++e) T[e] = A[e] + C[e] * lam[i];
            init_fdma_ld(T, (size_t)len, d, u1, u2, l, len);
            fdma_line(l, d, u1, u2, line, len, s);
        }
    }
    free(T); free(d); free(u1); free(u2); free(l);
}

/* twodma.f90:28-62 -- upper (0,+2) back substitution along axis 0 or 1. */
static void twodma_line(const double* d, const double* u, double* x, int n, size_t s)
{
#define L1(i) x[(size_t)((i) - 1) * s]
    L1(n) = L1(n) / d[n - 1];
    L1(n - 1) = L1(n - 1) / d[n - 2];
    for (int i = n - 2; i >= 1; --i) L1(i) = (L1(i) - u[i - 1] * L1(i + 2)) / d[i - 1];
#undef L1
}

void ora_twodma_2d(const double* d, const double* u, double* x, int axis, int n, int m)
{
    if (axis == 0) {
        for (int j = 0; j < m; ++j) twodma_line(d, u, x + j, n, (size_t)m);
    } else {
        for (int i = 0; i < n; ++i) twodma_line(d, u, x + (size_t)i * m, m, 1);
    }
}
