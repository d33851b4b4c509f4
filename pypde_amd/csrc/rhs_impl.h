// Banded right-hand side of the fused RBC2D step: the per-output device
// function shared by the standalone kernel (rhs.hip) and the launch that
// merges it with the convective products (pointwise.hip).  Both translation
// units compile these sums without contraction into FMAs, so every launch form
// gives the same bits.
#pragma once
#include "common.h"

namespace pypde {

#pragma clang fp contract(off)

// Diagonal tables (diag[d][i] = A(i, i + off_d), zero where the column falls
// outside the operator): x tables M0 long, y tables M1 long.
enum { TX_BD = 0, TX_B = 3, TX_BS = 6, TX_BST = 10, TX_BD2S = 14, TX_BD2ST = 16, TX_COUNT = 18 };
enum { TY_B = 0, TY_BD = 3, TY_BS = 6, TY_BD2S = 10, TY_COUNT = 12 };

// Every load below is unconditional: an element outside the array is read at
// a clamped, valid index and its value replaced by zero (its term dropped)
// with a select afterwards.  A load under a branch makes the wave wait for it
// before it issues the next; the selects keep all of an output's loads in
// flight together and give the same sums, term for term.

// element c of a row of ncol, zero outside the row
__device__ __forceinline__ double row_at(const double* __restrict__ row, int c, int ncol)
{
    const double v = row[min(max(c, 0), ncol - 1)];
    return ((c >= 0) & (c < ncol)) ? v : 0.0;
}

// sum over the y diagonals (offsets oy[0..ny)) of row r of the input
template <int NY>
__device__ __forceinline__ double ysum(const double* __restrict__ row, int j, int ncol,
                                       const int (&oy)[NY], const double* __restrict__ ty, int M1)
{
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NY; ++b) s = s + row_at(row, j + oy[b], ncol) * ty[b * M1 + j];
    return s;
}

// ysum with the NY coefficients of column j already in registers
template <int NY>
__device__ __forceinline__ double ysum_c(const double* __restrict__ row, int j, int ncol,
                                         const int (&oy)[NY], const double (&coef)[NY])
{
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NY; ++b) s = s + row_at(row, j + oy[b], ncol) * coef[b];
    return s;
}

// ysum_c for a row that is read once per tile: the loads of the row's
// columns inside the row only (no clamped addresses to hold), which keeps the
// tile form's registers at three waves per SIMD; the same terms and order
template <int NY>
__device__ __forceinline__ double ysum_c_inside(const double* __restrict__ row, int j, int ncol,
                                                const int (&oy)[NY], const double (&coef)[NY])
{
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NY; ++b) {
        const int c = j + oy[b];
        const double v = (c >= 0 && c < ncol) ? row[c] : 0.0;
        s = s + v * coef[b];
    }
    return s;
}

// sum_a tx[a][i] * ysum(row i + ox[a]) over the rows inside the input
template <int NX, int NY>
__device__ __forceinline__ double band2(const double* __restrict__ in, long long ld, int nrow, int ncol,
                                        int i, int j, const int (&ox)[NX], const double* __restrict__ tx,
                                        int M0, const int (&oy)[NY], const double* __restrict__ ty,
                                        int M1)
{
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < NX; ++a) {
        const int r = i + ox[a];
        const int rc = min(max(r, 0), nrow - 1);
        const double s = ysum<NY>(in + (long long)rc * ld, j, ncol, oy, ty, M1);
        const double t = acc + tx[a * M0 + i] * s;
        acc = ((r >= 0) & (r < nrow)) ? t : acc;
    }
    return acc;
}

__device__ __forceinline__ void rhs_banded_at(const double* __restrict__ X, const double* __restrict__ pres,
                                              const double* __restrict__ cst, double* __restrict__ R,
                                              const double* __restrict__ mdta,
                                              const double* __restrict__ dta,
                                              const double* __restrict__ fdiff,
                                              const double* __restrict__ tx,
                                              const double* __restrict__ ty, int M0, int M1, int N0,
                                              int N1, long long t)
{
    const long long per = (long long)M0 * M1;
    const int o135[3] = {1, 3, 5};
    const int o024[3] = {0, 2, 4};
    const int o4[4] = {-2, 0, 2, 4};
    const int o02[2] = {0, 2};
    {
        const long long e = t / per;  // (instance, field) entry
        const long long r = t - e * per;
        const int b = (int)(e / 3), f = (int)(e - 3LL * b);
        const int i = (int)(r / M1), j = (int)(r - (long long)i * M1);
        const double* Xf = X + e * per;
        const double* tbs = tx + (long long)(f == 2 ? TX_BST : TX_BS) * M0;
        // old value (B S)_x X_f (B S)_y^T
        double acc = band2<4, 4>(Xf, M1, M0, M1, i, j, o4, tbs, M0, o4, ty + TY_BS * M1, M1);
        if (f < 2) {
            const double* pb = pres + (long long)b * N0 * N1;
            const double g = f == 0
                ? band2<3, 3>(pb, N1, N0, N1, i, j, o135, tx + TX_BD * M0, M0, o024, ty + TY_B * M1, M1)
                : band2<3, 3>(pb, N1, N0, N1, i, j, o024, tx + TX_B * M0, M0, o135, ty + TY_BD * M1, M1);
            acc = acc + mdta[b] * g;
        }
        if (f == 1) {
            // buoyancy from the temperature of the same instance
            const double* XT = X + (e + 1) * per;
            const double bt = band2<4, 4>(XT, M1, M0, M1, i, j, o4, tx + TX_BST * M0, M0, o4,
                                          ty + TY_BS * M1, M1);
            acc = acc + dta[b] * bt;
        }
        if (fdiff) {
            const double* td2 = tx + (long long)(f == 2 ? TX_BD2ST : TX_BD2S) * M0;
            const double dxx = band2<2, 4>(Xf, M1, M0, M1, i, j, o02, td2, M0, o4, ty + TY_BS * M1, M1);
            const double dyy = band2<4, 2>(Xf, M1, M0, M1, i, j, o4, tbs, M0, o02, ty + TY_BD2S * M1, M1);
            acc = acc + fdiff[e] * (dxx + dyy);
        }
        R[t] = cst ? cst[t] + acc : acc;
    }
}

}  // namespace pypde
