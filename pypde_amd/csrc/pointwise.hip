// Pointwise kernels of the fused RBC2D step.
#include "common.h"
#include "../../include/pypde_amd.h"
#include "rhs_impl.h"

#include <cstring>

namespace pypde {

#pragma clang fp contract(off)

// Convective products on the dealiased grid (field_operations.py:124-129,160-164):
//   PU = dUdx*ux + dUdy*uz,  PV = dVdx*ux + dVdy*uz,
//   PT = dTdx*ux + dTdy*uz + uz*dTbcdz1.
// Y holds, per instance and field f in (U, V, T), a [2*L0, 2*L1] array whose
// blocks are value (0,0), d/dy (0,1) and d/dx (1,0) on the L0 x L1 grid.
__global__ void rbc_products_kernel(const double* __restrict__ Y, const double* __restrict__ dtbc,
                                    double* __restrict__ P, int L0, int L1, int B)
{
    const long long per = (long long)L0 * L1;
    const long long ld = 2LL * L1;
    const long long fs = 4LL * per;  // field stride in Y
    const long long total = per * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long b = t / per;
        const long long r = t - b * per;
        const int i = (int)(r / L1), j = (int)(r - (long long)i * L1);
        const double* Yb = Y + b * 3 * fs;
        const long long v = (long long)i * ld + j;        // value block
        const long long dy = v + L1;                      // d/dy block
        const long long dx = (long long)(i + L0) * ld + j;  // d/dx block
        const double ux = Yb[v];
        const double uz = Yb[fs + v];
        const double pu = Yb[dx] * ux + Yb[dy] * uz;
        const double pv = Yb[fs + dx] * ux + Yb[fs + dy] * uz;
        const double pt = (Yb[2 * fs + dx] * ux + Yb[2 * fs + dy] * uz) + uz * dtbc[r];
        double* Pb = P + b * 3 * per;
        Pb[r] = pu;
        Pb[per + r] = pv;
        Pb[2 * per + r] = pt;
    }
}

// Convective products from y-folded fields.  Per instance and field f in
// (U, V, T), Y holds six [L0, H] blocks (H = L1 / 2) over the first half of the
// y grid: q = 0, 1 value (even / odd y-coefficient part), 2, 3 d/dy, 4, 5 d/dx.
// The field at grid column l is E + O and at its mirror L1-1-l it is s (E - O),
// s = sv for values and d/dx, sd for d/dy.  The products at both columns are
// written folded for the forward transform along y: per field, block 0 holds
// P(l) + P(L1-1-l) and block 1 P(l) - P(L1-1-l).  With Yold (the previous RK3
// stage, same layout), ratio times the products of the current gradients with
// the old velocities is added (the c-term of rbc2d.py:261-283).
__device__ __forceinline__ void products_yfold_at(const double* __restrict__ Y,
                                                  const double* __restrict__ Yold,
                                                  const double* __restrict__ dtbc,
                                                  double* __restrict__ P, int L0, int L1, double sv,
                                                  double sd, double ratio, long long t)
{
    const int H = L1 / 2;
    const long long per = (long long)L0 * H;
    const long long fs = 6LL * per;
    {
        const long long b = t / per;
        const long long r = t - b * per;
        const int i = (int)(r / H), l = (int)(r - (long long)i * H);
        const double* Yb = Y + b * 3 * fs + r;
        double v1[3], v2[3], y1[3], y2[3], x1[3], x2[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const double* q = Yb + f * fs;
            const double vE = q[0], vO = q[per];
            const double dE = q[2 * per], dO = q[3 * per];
            const double xE = q[4 * per], xO = q[5 * per];
            v1[f] = vE + vO;
            v2[f] = sv * (vE - vO);
            y1[f] = dE + dO;
            y2[f] = sd * (dE - dO);
            x1[f] = xE + xO;
            x2[f] = sv * (xE - xO);
        }
        const double* tb = dtbc + (long long)i * L1;
        const double tb1 = tb[l], tb2 = tb[L1 - 1 - l];
        double pu1 = x1[0] * v1[0] + y1[0] * v1[1];
        double pv1 = x1[1] * v1[0] + y1[1] * v1[1];
        double pt1 = (x1[2] * v1[0] + y1[2] * v1[1]) + v1[1] * tb1;
        double pu2 = x2[0] * v2[0] + y2[0] * v2[1];
        double pv2 = x2[1] * v2[0] + y2[1] * v2[1];
        double pt2 = (x2[2] * v2[0] + y2[2] * v2[1]) + v2[1] * tb2;
        if (Yold) {
            const double* Ob = Yold + b * 3 * fs + r;
            const double ux1 = Ob[0] + Ob[per], ux2 = sv * (Ob[0] - Ob[per]);
            const double uz1 = Ob[fs] + Ob[fs + per], uz2 = sv * (Ob[fs] - Ob[fs + per]);
            pu1 = pu1 + ratio * (x1[0] * ux1 + y1[0] * uz1);
            pv1 = pv1 + ratio * (x1[1] * ux1 + y1[1] * uz1);
            pt1 = pt1 + ratio * ((x1[2] * ux1 + y1[2] * uz1) + uz1 * tb1);
            pu2 = pu2 + ratio * (x2[0] * ux2 + y2[0] * uz2);
            pv2 = pv2 + ratio * (x2[1] * ux2 + y2[1] * uz2);
            pt2 = pt2 + ratio * ((x2[2] * ux2 + y2[2] * uz2) + uz2 * tb2);
        }
        double* Pb = P + b * 3 * 2 * per + r;
        Pb[0] = pu1 + pu2;
        Pb[per] = pu1 - pu2;
        Pb[2 * per] = pv1 + pv2;
        Pb[3 * per] = pv1 - pv2;
        Pb[4 * per] = pt1 + pt2;
        Pb[5 * per] = pt1 - pt2;
    }
}

__global__ void rbc_products_yfold_kernel(const double* __restrict__ Y, const double* __restrict__ Yold,
                                          const double* __restrict__ dtbc, double* __restrict__ P,
                                          int L0, int L1, double sv, double sd, double ratio, int B)
{
    const long long total = (long long)L0 * (L1 / 2) * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x)
        products_yfold_at(Y, Yold, dtbc, P, L0, L1, sv, sd, ratio, t);
}

// The y-folded products of many members (size classes) in one launch: block
// range [start[i], start[i+1]) serves member i, one element per thread.
constexpr int kMaxProdMulti = 64;
struct ProdMember {
    const double* Y;
    const double* Yold;
    const double* dtbc;
    double* P;
    int L0, L1, batch;
    double sv, sd, ratio;
};
struct ProdMulti {
    int count, total;
    int start[kMaxProdMulti + 1];
    ProdMember m[kMaxProdMulti];
};

long long products_multi_bytes() { return (long long)sizeof(ProdMulti); }

__global__ void rbc_products_yfold_multi(const ProdMulti* __restrict__ mm)
{
    const int blk = blockIdx.x;
    int lo = 0, hi = mm->count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mm->start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const ProdMember q = mm->m[lo];
    const long long t = (long long)(blk - mm->start[lo]) * blockDim.x + threadIdx.x;
    if (t < (long long)q.L0 * (q.L1 / 2) * q.batch)
        products_yfold_at(q.Y, q.Yold, q.dtbc, q.P, q.L0, q.L1, q.sv, q.sd, q.ratio, t);
}

// Adjoint of rbc_products_kernel: given the cotangents Pb of the three
// products, the cotangents of the value / d/dy / d/dx blocks of Y (the
// d2/dxdy blocks and the temperature value block receive zero).
__global__ void rbc_products_vjp_kernel(const double* __restrict__ Y, const double* __restrict__ dtbc,
                                        const double* __restrict__ Pb, double* __restrict__ Yb,
                                        int L0, int L1, int B)
{
    const long long per = (long long)L0 * L1;
    const long long ld = 2LL * L1;
    const long long fs = 4LL * per;
    const long long total = per * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long b = t / per;
        const long long r = t - b * per;
        const int i = (int)(r / L1), j = (int)(r - (long long)i * L1);
        const double* Yb_in = Y + b * 3 * fs;
        double* Yo = Yb + b * 3 * fs;
        const long long v = (long long)i * ld + j;
        const long long dy = v + L1;
        const long long dx = (long long)(i + L0) * ld + j;
        const long long dxy = dx + L1;
        const double ux = Yb_in[v], uz = Yb_in[fs + v];
        const double* Pbb = Pb + b * 3 * per;
        const double pu = Pbb[r], pv = Pbb[per + r], pt = Pbb[2 * per + r];
        Yo[dx] = pu * ux;
        Yo[dy] = pu * uz;
        Yo[fs + dx] = pv * ux;
        Yo[fs + dy] = pv * uz;
        Yo[2 * fs + dx] = pt * ux;
        Yo[2 * fs + dy] = pt * uz;
        Yo[v] = pu * Yb_in[dx] + pv * Yb_in[fs + dx] + pt * Yb_in[2 * fs + dx];
        Yo[fs + v] = pu * Yb_in[dy] + pv * Yb_in[fs + dy] + pt * (Yb_in[2 * fs + dy] + dtbc[r]);
        Yo[2 * fs + v] = 0.0;
        Yo[dxy] = 0.0;
        Yo[fs + dxy] = 0.0;
        Yo[2 * fs + dxy] = 0.0;
    }
}

// Adjoint of the products with the three-stage scheme's c-term,
//   P_f = gx_f (ux + r ux_o) + gy_f (uz + r uz_o)   (+ (uz + r uz_o) dTbc/dz for T),
// the current gradients times the current velocities plus r = c / b times the
// previous stage's velocities (Yold).  Writes the cotangents of the current
// blocks into Yb (as rbc_products_vjp_kernel, velocities now u + r u_o in the
// gradient terms) and those of the previous stage's value blocks into Yob (U
// and V values r * the velocity cotangent; every other block of Yob zero).
__global__ void rbc_products_vjp_c_kernel(const double* __restrict__ Y, const double* __restrict__ Yold,
                                          const double* __restrict__ dtbc,
                                          const double* __restrict__ Pb, double* __restrict__ Yb,
                                          double* __restrict__ Yob, double ratio, int L0, int L1,
                                          int B)
{
    const long long per = (long long)L0 * L1;
    const long long ld = 2LL * L1;
    const long long fs = 4LL * per;
    const long long total = per * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long b = t / per;
        const long long r = t - b * per;
        const int i = (int)(r / L1), j = (int)(r - (long long)i * L1);
        const double* Yi = Y + b * 3 * fs;
        const double* Yo_in = Yold + b * 3 * fs;
        double* Yo = Yb + b * 3 * fs;
        double* Yp = Yob + b * 3 * fs;
        const long long v = (long long)i * ld + j;
        const long long dy = v + L1;
        const long long dx = (long long)(i + L0) * ld + j;
        const long long dxy = dx + L1;
        const double ux = Yi[v] + ratio * Yo_in[v];
        const double uz = Yi[fs + v] + ratio * Yo_in[fs + v];
        const double* Pbb = Pb + b * 3 * per;
        const double pu = Pbb[r], pv = Pbb[per + r], pt = Pbb[2 * per + r];
        Yo[dx] = pu * ux;
        Yo[dy] = pu * uz;
        Yo[fs + dx] = pv * ux;
        Yo[fs + dy] = pv * uz;
        Yo[2 * fs + dx] = pt * ux;
        Yo[2 * fs + dy] = pt * uz;
        const double gux = pu * Yi[dx] + pv * Yi[fs + dx] + pt * Yi[2 * fs + dx];
        const double guz = pu * Yi[dy] + pv * Yi[fs + dy] + pt * (Yi[2 * fs + dy] + dtbc[r]);
        Yo[v] = gux;
        Yo[fs + v] = guz;
        Yo[2 * fs + v] = 0.0;
        Yo[dxy] = 0.0;
        Yo[fs + dxy] = 0.0;
        Yo[2 * fs + dxy] = 0.0;
        Yp[v] = ratio * gux;
        Yp[fs + v] = ratio * guz;
        Yp[2 * fs + v] = 0.0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            Yp[f * fs + dy] = 0.0;
            Yp[f * fs + dx] = 0.0;
            Yp[f * fs + dxy] = 0.0;
        }
    }
}

// y[b] = y[b] - (coef[b] * x[b]) * c2   (pres -= nu*div*beta, rbc2d.py:391)
__global__ void scaled_sub_kernel(long long per, int B, const double* __restrict__ coef, double c2,
                                  const double* __restrict__ x, double* __restrict__ y)
{
    const long long total = per * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const double c = coef[t / per];
        y[t] = y[t] - (c * x[t]) * c2;
    }
}

__global__ void set_entries_kernel(double* p, long long stride, int n, double value)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[(long long)e * stride] = value;
}

// Sum of K-split partial products onto an addend, in a fixed order:
//   C[e] = (D[e] + (W[e S] + W[e S + 1] + ... + W[e S + S-1])) + X[e / xm]  (X only for e % xm == xr)
// over `entries` matrices of mn elements each (all contiguous).  The sum runs
// in partial order whatever the launch order, so results are reproducible.
__global__ void splitk_reduce_kernel(const double* __restrict__ W, int S, const double* __restrict__ D,
                                     const double* __restrict__ X, int xm, int xr,
                                     double* __restrict__ C, long long mn, long long total)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = t / mn, r = t - e * mn;
        const double* w = W + e * S * mn + r;
        double acc = w[0];
        for (int k = 1; k < S; ++k) acc = acc + w[k * mn];
        double v = D ? D[t] + acc : acc;
        if (X && (int)(e % xm) == xr) v = v + X[(e / xm) * mn + r];
        C[t] = v;
    }
}

// dst[b][j][i] = src[b][i][j] over i < rows, j < cols: 32 x 32 tiles through
// LDS (one padding column, so the column reads hit distinct banks); loads and
// stores each run along rows of 32 consecutive doubles.
constexpr int kTrTile = 32;

__global__ void __launch_bounds__(256) transpose_kernel(const double* __restrict__ src,
                                                        long long src_ld, long long src_bs,
                                                        double* __restrict__ dst, long long dst_ld,
                                                        long long dst_bs, int rows, int cols)
{
    __shared__ double t[kTrTile][kTrTile + 1];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * kTrTile, j0 = blockIdx.x * kTrTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    const double* s = src + (long long)b * src_bs;
    double* d = dst + (long long)b * dst_bs;
#pragma unroll
    for (int r = ty; r < kTrTile; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < rows && j < cols) t[r][tx] = s[(long long)i * src_ld + j];
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < kTrTile; r += 8) {
        const int j = j0 + r, i = i0 + tx;
        if (i < rows && j < cols) d[(long long)j * dst_ld + i] = t[tx][r];
    }
}

static unsigned grid_for(long long n)
{
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// The banded right-hand side (blocks [0, nbr), one output per thread) and the
// y-folded products (blocks [nbr, nbr + nbp)) in one launch: the two read
// disjoint inputs (the state; the dealiased fields) and write disjoint outputs,
// so they overlap on the device instead of running one after the other.  The
// right-hand side, the longer of the two, gets the first blocks.  Each output
// is computed exactly as by its own kernel.
__global__ void __launch_bounds__(256) rbc_products_rhs_kernel(ProdMember q, pypde_rbc_rhs a,
                                                               unsigned nbr, unsigned nbp)
{
    if (blockIdx.x < nbr) {
        const long long total = 3LL * a.M0 * a.M1 * a.batch;
        for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < total;
             t += (long long)nbr * 256)
            rhs_banded_at(a.X, a.pres, a.cst, a.R, a.mdta, a.dta, a.fdiff, a.tx, a.ty, a.M0,
                          a.M1, a.N0, a.N1, t);
        return;
    }
    const long long total = (long long)q.L0 * (q.L1 / 2) * q.batch;
    for (long long t = (long long)(blockIdx.x - nbr) * 256 + threadIdx.x; t < total;
         t += (long long)nbp * 256)
        products_yfold_at(q.Y, q.Yold, q.dtbc, q.P, q.L0, q.L1, q.sv, q.sd, q.ratio, t);
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_rbc_products(const double* Y, const double* dtbc, double* P, int L0, int L1,
                                  int batch, void* stream)
{
    PYPDE_REQUIRE(Y && dtbc && P, "null pointer");
    PYPDE_REQUIRE(L0 > 0 && L1 > 0 && batch >= 0, "invalid product geometry");
    if (batch == 0) return OK;
    hipLaunchKernelGGL(rbc_products_kernel, dim3(grid_for((long long)L0 * L1 * batch)), dim3(256), 0,
                       as_stream(stream), Y, dtbc, P, L0, L1, batch);
    return check_launch("pypde_rbc_products");
}

extern "C" int pypde_rbc_products_yfold(const double* Y, const double* Yold, const double* dtbc,
                                        double* P, int L0, int L1, double sign_value,
                                        double sign_dy, double ratio, int batch, void* stream)
{
    PYPDE_REQUIRE(Y && dtbc && P, "null pointer");
    PYPDE_REQUIRE(L0 > 0 && L1 > 0 && L1 % 2 == 0 && batch >= 0,
                  "invalid product geometry (L1 must be even)");
    if (batch == 0) return OK;
    hipLaunchKernelGGL(rbc_products_yfold_kernel, dim3(grid_for((long long)L0 * (L1 / 2) * batch)), dim3(256), 0,
                       as_stream(stream), Y, Yold, dtbc, P, L0, L1, sign_value, sign_dy, ratio, batch);
    return check_launch("pypde_rbc_products_yfold");
}

extern "C" int pypde_rbc_products_rhs(const double* Y, const double* Yold, const double* dtbc,
                                      double* P, int L0, int L1, double sign_value,
                                      double sign_dy, double ratio, int batch,
                                      const pypde_rbc_rhs* rhs, void* stream)
{
    PYPDE_REQUIRE(Y && dtbc && P && rhs, "null pointer");
    PYPDE_REQUIRE(L0 > 0 && L1 > 0 && L1 % 2 == 0 && batch >= 0,
                  "invalid geometry (L1 must be even: the y-folded layout)");
    PYPDE_REQUIRE(rhs->X && rhs->pres && rhs->R && rhs->mdta && rhs->dta && rhs->tx && rhs->ty,
                  "null right-hand-side pointer (X, pres, R, mdta, dta, tx and ty are required)");
    PYPDE_REQUIRE(rhs->M0 >= 1 && rhs->M1 >= 1 && rhs->N0 == rhs->M0 + 2 &&
                      rhs->N1 == rhs->M1 + 2 && rhs->batch >= 0,
                  "invalid right-hand-side geometry: M0 %d M1 %d N0 %d N1 %d batch %d", rhs->M0,
                  rhs->M1, rhs->N0, rhs->N1, rhs->batch);
    PYPDE_REQUIRE(rhs->R != rhs->X, "R must not alias X (neighbouring entries are read)");
    PYPDE_REQUIRE(rhs->rows_per_thread == 1,
                  "the merged launch computes one right-hand-side output per thread "
                  "(rows_per_thread must be 1, got %d)", rhs->rows_per_thread);
    const long long tr = 3LL * rhs->M0 * rhs->M1 * rhs->batch;
    const long long tp = (long long)L0 * (L1 / 2) * batch;
    if (tr + tp == 0) return OK;
    long long nbr = (tr + 255) / 256;
    if (nbr > 16384) nbr = 16384;
    const unsigned nbp = tp ? grid_for(tp) : 0;
    const ProdMember q{Y, Yold, dtbc, P, L0, L1, batch, sign_value, sign_dy, ratio};
    hipLaunchKernelGGL(rbc_products_rhs_kernel, dim3((unsigned)nbr + nbp), dim3(256), 0,
                       as_stream(stream), q, *rhs, (unsigned)nbr, nbp);
    return check_launch("pypde_rbc_products_rhs");
}

extern "C" int pypde_products_multi_pack(const double* const* Y, const double* const* Yold,
                                         const double* const* dtbc, double* const* P,
                                         const int* L0, const int* L1, const double* sign_value,
                                         const double* sign_dy, const double* ratio,
                                         const int* batch, int count, void* table,
                                         long long table_bytes, pypde_multi_info* info)
{
    PYPDE_REQUIRE(Y && Yold && dtbc && P && L0 && L1 && sign_value && sign_dy && ratio && batch &&
                      table && info,
                  "null argument");
    PYPDE_REQUIRE(table_bytes >= (long long)sizeof(ProdMulti), "products table needs %zu bytes",
                  sizeof(ProdMulti));
    PYPDE_REQUIRE(count >= 1 && count <= kMaxProdMulti, "multi products: count %d not in 1..%d",
                  count, kMaxProdMulti);
    ProdMulti* mm = static_cast<ProdMulti*>(table);
    memset(mm, 0, sizeof(ProdMulti));
    mm->count = count;
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        PYPDE_REQUIRE(Y[i] && dtbc[i] && P[i], "member %d: null pointer", i);
        PYPDE_REQUIRE(L0[i] > 0 && L1[i] > 0 && L1[i] % 2 == 0 && batch[i] >= 0,
                      "member %d: invalid product geometry (L1 must be even)", i);
        mm->m[i] = ProdMember{Y[i], Yold[i], dtbc[i], P[i], L0[i], L1[i], batch[i],
                              sign_value[i], sign_dy[i], ratio[i]};
        mm->start[i] = (int)total;
        total += ((long long)L0[i] * (L1[i] / 2) * batch[i] + 255) / 256;
        PYPDE_REQUIRE(total < (1LL << 31), "multi products too large");
    }
    mm->start[count] = (int)total;
    mm->total = (int)total;
    info->kind = PYPDE_MULTI_PRODUCTS;
    info->count = count;
    info->variant = 0;
    info->flags = 0;
    info->total = total;
    return OK;
}

extern "C" int pypde_products_multi(const void* table, const pypde_multi_info* info, void* stream)
{
    PYPDE_REQUIRE(table && info, "null argument");
    PYPDE_REQUIRE(info->kind == PYPDE_MULTI_PRODUCTS, "not a products table");
    if (info->total == 0) return OK;
    hipLaunchKernelGGL(rbc_products_yfold_multi, dim3((unsigned)info->total), dim3(256), 0,
                       as_stream(stream), static_cast<const ProdMulti*>(table));
    return check_launch("pypde_products_multi");
}

extern "C" int pypde_rbc_products_vjp(const double* Y, const double* dtbc, const double* Pb,
                                      double* Yb, int L0, int L1, int batch, void* stream)
{
    PYPDE_REQUIRE(Y && dtbc && Pb && Yb, "null pointer");
    PYPDE_REQUIRE(L0 > 0 && L1 > 0 && batch >= 0, "invalid product geometry");
    if (batch == 0) return OK;
    hipLaunchKernelGGL(rbc_products_vjp_kernel, dim3(grid_for((long long)L0 * L1 * batch)), dim3(256),
                       0, as_stream(stream), Y, dtbc, Pb, Yb, L0, L1, batch);
    return check_launch("pypde_rbc_products_vjp");
}

extern "C" int pypde_rbc_products_vjp_c(const double* Y, const double* Yold, const double* dtbc,
                                        const double* Pb, double* Yb, double* Yob, double ratio,
                                        int L0, int L1, int batch, void* stream)
{
    PYPDE_REQUIRE(Y && Yold && dtbc && Pb && Yb && Yob, "null pointer");
    PYPDE_REQUIRE(L0 > 0 && L1 > 0 && batch >= 0, "invalid product geometry");
    if (batch == 0) return OK;
    hipLaunchKernelGGL(rbc_products_vjp_c_kernel, dim3(grid_for((long long)L0 * L1 * batch)),
                       dim3(256), 0, as_stream(stream), Y, Yold, dtbc, Pb, Yb, Yob, ratio, L0, L1,
                       batch);
    return check_launch("pypde_rbc_products_vjp_c");
}

extern "C" int pypde_scaled_sub(long long per, int batch, const double* coef, double c2,
                                const double* x, double* y, void* stream)
{
    PYPDE_REQUIRE(coef && x && y, "null pointer");
    PYPDE_REQUIRE(per >= 0 && batch >= 0, "negative size");
    if (per == 0 || batch == 0) return OK;
    hipLaunchKernelGGL(scaled_sub_kernel, dim3(grid_for(per * batch)), dim3(256), 0,
                       as_stream(stream), per, batch, coef, c2, x, y);
    return check_launch("pypde_scaled_sub");
}

extern "C" int pypde_transpose(const double* src, long long src_ld, long long src_bs, double* dst,
                               long long dst_ld, long long dst_bs, int rows, int cols, int batch,
                               void* stream)
{
    PYPDE_REQUIRE(src && dst, "null pointer");
    PYPDE_REQUIRE(src != dst, "the transpose cannot run in place");
    PYPDE_REQUIRE(rows >= 0 && cols >= 0 && batch >= 0, "negative size");
    if ((long long)rows * cols * batch == 0) return OK;
    PYPDE_REQUIRE(src_ld >= cols && dst_ld >= rows,
                  "leading dimensions too small (src_ld %lld < %d or dst_ld %lld < %d)", src_ld,
                  cols, dst_ld, rows);
    PYPDE_REQUIRE(batch < 65536, "batch must be below 65536 (got %d)", batch);
    const dim3 grid((cols + kTrTile - 1) / kTrTile, (rows + kTrTile - 1) / kTrTile, batch);
    PYPDE_REQUIRE(grid.y < 65536, "too many rows (%d)", rows);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, as_stream(stream), src, src_ld, src_bs,
                       dst, dst_ld, dst_bs, rows, cols);
    return check_launch("pypde_transpose");
}

extern "C" int pypde_set_entries(double* p, long long stride, int n, double value, void* stream)
{
    PYPDE_REQUIRE(p, "null pointer");
    if (n <= 0) return OK;
    hipLaunchKernelGGL(set_entries_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), p,
                       stride, n, value);
    return check_launch("pypde_set_entries");
}

extern "C" int pypde_splitk_reduce(const double* W, int S, const double* D, const double* X,
                                   int xm, int xr, double* C, long long mn, int entries,
                                   void* stream)
{
    PYPDE_REQUIRE(W && C, "null pointer");
    PYPDE_REQUIRE(S >= 1 && S <= 64, "splits must be 1..64 (got %d)", S);
    PYPDE_REQUIRE(mn >= 0 && entries >= 0, "negative size");
    PYPDE_REQUIRE(!X || (xm >= 1 && xr >= 0 && xr < xm), "extra addend needs 0 <= xr < xm");
    if (mn == 0 || entries == 0) return OK;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for(mn * entries)), dim3(256), 0,
                       as_stream(stream), W, S, D, X, xm < 1 ? 1 : xm, xr, C, mn, mn * entries);
    return check_launch("pypde_splitk_reduce");
}
