// Banded part of the Helmholtz right-hand side of the fused RBC2D step.
//
// Per instance b and field f (U, V, T) the right-hand side of the implicit
// solve is  BF_x Z_f + R_f, where Z_f carries the convective term and R_f the
// terms whose x and y operators are banded (navier/rbc2d.py:213-283 with the
// plan right-hand sides of templates/hholtz.py:70-93):
//   pressure gradient  U: -dt a (B D / s)_x pres B_y^T   (offsets 1,3,5 x 0,2,4)
//                      V: -dt a B_x pres (B D / s)_y^T   (offsets 0,2,4 x 1,3,5)
//   old value          (B S)_x X_f (B S)_y^T             (offsets -2..4 x -2..4)
//   buoyancy (V)       dt a (B S_T)_x X_T (B S)_y^T
//   explicit diffusion f_f [(B D2 S)_x X_f (B S)_y^T + (B S)_x X_f (B D2 S)_y^T]
//   constant           const_f (the lifted boundary terms)
// Each output sums its neighbourhood: for every x diagonal, the y-direction
// sum of the input row first, then that times the x coefficient, in ascending
// diagonal order.  One thread per output; the waves of a block run along
// contiguous entries of the output rows, so the row loads coalesce and the
// neighbouring rows they share come from the caches.
#include "common.h"
#include "../../include/pypde_amd.h"
#include "rhs_impl.h"

#include <cstring>

namespace pypde {

#pragma clang fp contract(off)

// The same outputs for RS consecutive rows i0 .. i0+RS-1 of one column j per
// thread: the y-direction sums of each input row are formed once and shared
// by the RS outputs that use that row (about 10 loads per output instead of
// 16 for the old-value term).  Each output sums the same terms in the same
// order as rhs_banded_at, so the results are bit-identical.
template <int RS>
__device__ __forceinline__ void rhs_banded_strip(const double* __restrict__ X,
                                                 const double* __restrict__ pres,
                                                 const double* __restrict__ cst,
                                                 double* __restrict__ R,
                                                 const double* __restrict__ mdta,
                                                 const double* __restrict__ dta,
                                                 const double* __restrict__ fdiff,
                                                 const double* __restrict__ tx,
                                                 const double* __restrict__ ty, int M0, int M1,
                                                 int N0, int N1, long long t)
{
    const int o135[3] = {1, 3, 5};
    const int o024[3] = {0, 2, 4};
    const int o4[4] = {-2, 0, 2, 4};
    const int o02[2] = {0, 2};
    const int nstrip = (M0 + RS - 1) / RS;
    const long long per = (long long)nstrip * M1;
    const long long e = t / per;
    const long long rr = t - e * per;
    const int s = (int)(rr / M1), j = (int)(rr - (long long)s * M1);
    const int b = (int)(e / 3), f = (int)(e - 3LL * b);
    const int i0 = s * RS;
    const long long mm = (long long)M0 * M1;
    const double* Xf = X + e * mm;
    const double* tbs = tx + (long long)(f == 2 ? TX_BST : TX_BS) * M0;
    // the column's (B S)_y coefficients, once for all the strip's rows
    double cbs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) cbs[c] = ty[(TY_BS + c) * M1 + j];
    // y sums with (B S)_y of the rows i0-2 .. i0+RS+3 of X_f
    double ys[RS + 6];
#pragma unroll
    for (int q = 0; q < RS + 6; ++q) {
        const int r = i0 - 2 + q;
        const double v = ysum_c<4>(Xf + (long long)min(max(r, 0), M0 - 1) * M1, j, M1, o4, cbs);
        ys[q] = ((r >= 0) & (r < M0)) ? v : 0.0;
    }
    double acc[RS];
#pragma unroll
    for (int k = 0; k < RS; ++k) {
        const int i = i0 + k;
        double a = 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int r = i + o4[d];
            const double sum = a + tbs[d * M0 + min(i, M0 - 1)] * ys[k + d * 2];
            a = ((i < M0) & (r >= 0) & (r < M0)) ? sum : a;
        }
        acc[k] = a;
    }
    if (f < 2) {
        // pressure rows i0+1 .. i0+RS+4 (U, offsets 1,3,5) or i0 .. i0+RS+3 (V, 0,2,4)
        const double* pb = pres + (long long)b * N0 * N1;
        const int first = f == 0 ? 1 : 0;
        double cp[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) cp[c] = ty[((f == 0 ? TY_B : TY_BD) + c) * M1 + j];
        double yp[RS + 4];
#pragma unroll
        for (int q = 0; q < RS + 4; ++q) {
            const int r = i0 + first + q;
            const double* prow = pb + (long long)min(r, N0 - 1) * N1;
            const double v = f == 0 ? ysum_c<3>(prow, j, N1, o024, cp) : ysum_c<3>(prow, j, N1, o135, cp);
            yp[q] = r < N0 ? v : 0.0;
        }
        const double* tp = tx + (long long)(f == 0 ? TX_BD : TX_B) * M0;
        const double m = mdta[b];
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            const int i = i0 + k;
            double g = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int r = i + (f == 0 ? o135[d] : o024[d]);
                const double sum = g + tp[d * M0 + min(i, M0 - 1)] * yp[k + 2 * d];
            g = ((i < M0) & (r < N0)) ? sum : g;
            }
            acc[k] = acc[k] + m * g;
        }
    }
    if (f == 1) {
        // buoyancy from the temperature rows i0-2 .. i0+RS+3
        const double* XT = X + (e + 1) * mm;
        double yt[RS + 6];
#pragma unroll
        for (int q = 0; q < RS + 6; ++q) {
            const int r = i0 - 2 + q;
            const double v = ysum_c<4>(XT + (long long)min(max(r, 0), M0 - 1) * M1, j, M1, o4, cbs);
            yt[q] = ((r >= 0) & (r < M0)) ? v : 0.0;
        }
        const double* tt = tx + (long long)TX_BST * M0;
        const double c = dta[b];
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            const int i = i0 + k;
            double bt = 0.0;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int r = i + o4[d];
                const double sum = bt + tt[d * M0 + min(i, M0 - 1)] * yt[k + d * 2];
            bt = ((i < M0) & (r >= 0) & (r < M0)) ? sum : bt;
            }
            acc[k] = acc[k] + c * bt;
        }
    }
    if (fdiff) {
        double yd[RS + 6];
#pragma unroll
        for (int q = 0; q < RS + 6; ++q) {
            const int r = i0 - 2 + q;
            const double v = ysum<2>(Xf + (long long)min(max(r, 0), M0 - 1) * M1, j, M1, o02,
                                     ty + TY_BD2S * M1, M1);
            yd[q] = ((r >= 0) & (r < M0)) ? v : 0.0;
        }
        const double* td2 = tx + (long long)(f == 2 ? TX_BD2ST : TX_BD2S) * M0;
        const double fe = fdiff[e];
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            const int i = i0 + k;
            double dxx = 0.0, dyy = 0.0;
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int r = i + o02[d];
                const double sum = dxx + td2[d * M0 + min(i, M0 - 1)] * ys[k + 2 + d * 2];
            dxx = ((i < M0) & (r < M0)) ? sum : dxx;
            }
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int r = i + o4[d];
                const double sum = dyy + tbs[d * M0 + min(i, M0 - 1)] * yd[k + d * 2];
            dyy = ((i < M0) & (r >= 0) & (r < M0)) ? sum : dyy;
            }
            acc[k] = acc[k] + fe * (dxx + dyy);
        }
    }
#pragma unroll
    for (int k = 0; k < RS; ++k) {
        const int i = i0 + k;
        if (i < M0) {
            const long long o = e * mm + (long long)i * M1 + j;
            R[o] = cst ? cst[o] + acc[k] : acc[k];
        }
    }
}

// Tile form: a block of 256 threads serves kRhsTI rows x 64 columns of one
// (instance, field) entry.  First the y-direction sums of every input row the
// tile needs -- rows i0-2 .. i0+kRhsTI+3 of X_f, and of X_T and of the pressure
// where the field has those terms -- go to LDS, one column per lane, each formed
// once (with the column's y coefficients held in registers); then every thread
// sums its outputs over the x diagonals from LDS.  About 6-17 global loads per
// output instead of 25-95.  The same terms in the same order as rhs_banded_at:
// bit-identical.
constexpr int kRhsTI = 16, kRhsTJ = 64, kRhsRows = kRhsTI + 6;

__device__ __forceinline__ void rhs_banded_tile(const pypde_rbc_rhs& a, long long e, int ti, int tj)
{
    // LDS of the launch (rhs_tile_lds_bytes): three arrays, four with the
    // explicit diffusion
    extern __shared__ double rhs_lds[];
    double (*ys)[kRhsTJ] = reinterpret_cast<double (*)[kRhsTJ]>(rhs_lds);   // (B S)_y sums of X_f
    double (*yt)[kRhsTJ] = ys + kRhsRows;      // (B S)_y sums of X_T (buoyancy, V)
    double (*yp)[kRhsTJ] = yt + kRhsRows;      // pressure sums (U: B_y, V: (B D)_y)
    double (*yd)[kRhsTJ] = yp + kRhsRows;      // (B D2 S)_y sums of X_f (explicit diffusion)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int M0 = a.M0, M1 = a.M1, N0 = a.N0, N1 = a.N1;
    const int b = (int)(e / 3), f = (int)(e - 3LL * b);
    const int i0 = ti * kRhsTI, j = tj * kRhsTJ + lane;
    const bool jin = j < M1;
    const long long mm = (long long)M0 * M1;
    const double* Xf = a.X + e * mm;
    const double* pb = a.pres + (long long)b * N0 * N1;
    const int o4[4] = {-2, 0, 2, 4};
    const int o02[2] = {0, 2};
    const int o024[3] = {0, 2, 4};
    const int o135[3] = {1, 3, 5};
    const int pfirst = f == 0 ? 1 : 0;   // first pressure row: U offsets 1,3,5; V 0,2,4
    // this column's y coefficients, loaded once for all the tile's rows
    double cbs[4], cd2[2], cp[3];
#pragma unroll
    for (int c = 0; c < 4; ++c) cbs[c] = jin ? a.ty[(TY_BS + c) * M1 + j] : 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) cd2[c] = (jin && a.fdiff) ? a.ty[(TY_BD2S + c) * M1 + j] : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        cp[c] = (jin && f < 2) ? a.ty[((f == 0 ? TY_B : TY_BD) + c) * M1 + j] : 0.0;
    // y sums (ysum's terms and order, the coefficients from registers): LDS row
    // q holds X row i0 - 2 + q and pressure row i0 + pfirst + q
    // the constant terms and x coefficients of this thread's outputs (rows
    // i0 + 4 w + k), fetched while the y sums form
    const double* tbs = a.tx + (long long)(f == 2 ? TX_BST : TX_BS) * M0;
    const double* tp = a.tx + (long long)(f == 0 ? TX_BD : TX_B) * M0;
    const double* tt = a.tx + (long long)TX_BST * M0;
    const double* td2 = a.tx + (long long)(f == 2 ? TX_BD2ST : TX_BD2S) * M0;
    constexpr int KR = kRhsTI / 4;
    double cv[KR], xbs[KR][4], xp[KR][3], xt[KR][4], xd[KR][2];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int i = i0 + 4 * w + k;
        const bool ok = jin && i < M0;
        cv[k] = (a.cst && ok) ? a.cst[e * mm + (long long)i * M1 + j] : 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) xbs[k][d] = ok ? tbs[d * M0 + i] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) xp[k][d] = (ok && f < 2) ? tp[d * M0 + i] : 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) xt[k][d] = (ok && f == 1) ? tt[d * M0 + i] : 0.0;
#pragma unroll
        for (int d = 0; d < 2; ++d) xd[k][d] = (ok && a.fdiff) ? td2[d * M0 + i] : 0.0;
    }
    const double mb = f < 2 ? a.mdta[b] : 0.0;
    const double db = f == 1 ? a.dta[b] : 0.0;
    const double fe = a.fdiff ? a.fdiff[e] : 0.0;
    // one input row per iteration, its loads together: with every row's loads
    // in flight at once the tile takes the registers of two more waves and
    // runs slower (265 against 208 us at 64 x 256x256)
#pragma unroll
    for (int t = 0; t < (kRhsRows + 3) / 4; ++t) {
        const int q = w + 4 * t;
        if (q >= kRhsRows || !jin) continue;
        const int r = i0 - 2 + q;
        const bool rin = r >= 0 && r < M0;
        ys[q][lane] = rin ? ysum_c_inside<4>(Xf + (long long)r * M1, j, M1, o4, cbs) : 0.0;
        if (f == 1)
            yt[q][lane] = rin ? ysum_c_inside<4>(a.X + (e + 1) * mm + (long long)r * M1, j, M1, o4, cbs)
                              : 0.0;
        if (f < 2) {
            const int rp = i0 + pfirst + q;
            yp[q][lane] = rp < N0 ? (f == 0 ? ysum_c_inside<3>(pb + (long long)rp * N1, j, N1, o024, cp)
                                            : ysum_c_inside<3>(pb + (long long)rp * N1, j, N1, o135, cp))
                                  : 0.0;
        }
        if (a.fdiff)
            yd[q][lane] = rin ? ysum_c_inside<2>(Xf + (long long)r * M1, j, M1, o02, cd2) : 0.0;
    }
    __syncthreads();
    if (!jin) return;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int i = i0 + 4 * w + k;
        if (i >= M0) break;
        const int qi = i - i0 + 2;   // LDS row of X row i
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int r = i + o4[d];
            if (r >= 0 && r < M0) acc = acc + xbs[k][d] * ys[qi + o4[d]][lane];
        }
        if (f < 2) {
            double g = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int off = f == 0 ? o135[d] : o024[d];
                if (i + off < N0) g = g + xp[k][d] * yp[i - i0 + off - pfirst][lane];
            }
            acc = acc + mb * g;
        }
        if (f == 1) {
            double bt = 0.0;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int r = i + o4[d];
                if (r >= 0 && r < M0) bt = bt + xt[k][d] * yt[qi + o4[d]][lane];
            }
            acc = acc + db * bt;
        }
        if (a.fdiff) {
            double dxx = 0.0, dyy = 0.0;
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int r = i + o02[d];
                if (r < M0) dxx = dxx + xd[k][d] * ys[qi + o02[d]][lane];
            }
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int r = i + o4[d];
                if (r >= 0 && r < M0) dyy = dyy + xbs[k][d] * yd[qi + o4[d]][lane];
            }
            acc = acc + fe * (dxx + dyy);
        }
        const long long o = e * mm + (long long)i * M1 + j;
        a.R[o] = a.cst ? cv[k] + acc : acc;
    }
}

__device__ __forceinline__ long long rhs_tiles_per_entry(const pypde_rbc_rhs& a)
{
    return (long long)((a.M0 + kRhsTI - 1) / kRhsTI) * ((a.M1 + kRhsTJ - 1) / kRhsTJ);
}

__device__ __forceinline__ void rhs_tile_at(const pypde_rbc_rhs& a, long long t)
{
    const int ntj = (a.M1 + kRhsTJ - 1) / kRhsTJ;
    const long long per = rhs_tiles_per_entry(a);
    const long long e = t / per;
    const int r = (int)(t - e * per);
    rhs_banded_tile(a, e, r / ntj, r % ntj);
}

// LDS bytes of a tile launch: the y-sum arrays it uses (the diffusion one only
// when some member has the explicit diffusion)
static size_t rhs_tile_lds_bytes(bool diffusion)
{
    return sizeof(double) * kRhsRows * kRhsTJ * (diffusion ? 4 : 3);
}

__global__ void __launch_bounds__(256) rbc_rhs_tile_kernel(pypde_rbc_rhs a)
{
    rhs_tile_at(a, blockIdx.x);
}

template <int RS>
__global__ void __launch_bounds__(64) rbc_rhs_strip_kernel(pypde_rbc_rhs a)
{
    const long long total = 3LL * a.batch * ((a.M0 + RS - 1) / RS) * a.M1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total)
        rhs_banded_strip<RS>(a.X, a.pres, a.cst, a.R, a.mdta, a.dta, a.fdiff, a.tx, a.ty, a.M0,
                             a.M1, a.N0, a.N1, t);
}

__global__ void rbc_rhs_banded_kernel(const double* __restrict__ X, const double* __restrict__ pres,
                                      const double* __restrict__ cst, double* __restrict__ R,
                                      const double* __restrict__ mdta, const double* __restrict__ dta,
                                      const double* __restrict__ fdiff, const double* __restrict__ tx,
                                      const double* __restrict__ ty, int M0, int M1, int N0, int N1,
                                      int B)
{
    const long long total = 3LL * M0 * M1 * B;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x)
        rhs_banded_at(X, pres, cst, R, mdta, dta, fdiff, tx, ty, M0, M1, N0, N1, t);
}

// The banded right-hand sides of many members (size classes) in one launch:
// block range [start[i], start[i+1]) serves member i, one output per thread.
constexpr int kMaxRhsMulti = 64;
struct RhsMulti {
    int count, total;
    int start[kMaxRhsMulti + 1];
    pypde_rbc_rhs m[kMaxRhsMulti];
};

// RS = 1: one output per thread (rhs_banded_at); 4, 8: a strip of RS rows of
// one column per thread (rhs_banded_strip); the same bits either way.
template <int RS>
__global__ void __launch_bounds__(256) rbc_rhs_banded_multi(const RhsMulti* __restrict__ mm)
{
    const int blk = blockIdx.x;
    int lo = 0, hi = mm->count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mm->start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const pypde_rbc_rhs a = mm->m[lo];
    if (RS == 16) {
        rhs_tile_at(a, blk - mm->start[lo]);
        return;
    }
    const long long t = (long long)(blk - mm->start[lo]) * blockDim.x + threadIdx.x;
    if (RS == 1) {
        if (t < 3LL * a.M0 * a.M1 * a.batch)
            rhs_banded_at(a.X, a.pres, a.cst, a.R, a.mdta, a.dta, a.fdiff, a.tx, a.ty, a.M0,
                          a.M1, a.N0, a.N1, t);
    } else if (t < 3LL * a.batch * ((a.M0 + RS - 1) / RS) * a.M1) {
        rhs_banded_strip<RS>(a.X, a.pres, a.cst, a.R, a.mdta, a.dta, a.fdiff, a.tx, a.ty, a.M0,
                             a.M1, a.N0, a.N1, t);
    }
}

}  // namespace pypde

using namespace pypde;

static int check_rhs(const pypde_rbc_rhs* a)
{
    PYPDE_REQUIRE(a, "null argument");
    PYPDE_REQUIRE(a->X && a->pres && a->R && a->mdta && a->dta && a->tx && a->ty,
                  "null pointer (X, pres, R, mdta, dta, tx and ty are required)");
    PYPDE_REQUIRE(a->M0 >= 1 && a->M1 >= 1 && a->N0 == a->M0 + 2 && a->N1 == a->M1 + 2 &&
                      a->batch >= 0,
                  "invalid geometry: M0 %d M1 %d N0 %d N1 %d batch %d", a->M0, a->M1, a->N0,
                  a->N1, a->batch);
    PYPDE_REQUIRE(a->R != a->X, "R must not alias X (neighbouring entries are read)");
    return OK;
}

extern "C" int pypde_rhs_multi_pack(const pypde_rbc_rhs* args, int count, void* table,
                                    long long table_bytes, pypde_multi_info* info)
{
    PYPDE_REQUIRE(args && table && info, "null argument");
    PYPDE_REQUIRE(table_bytes >= (long long)sizeof(RhsMulti), "right-hand-side table needs %zu bytes",
                  sizeof(RhsMulti));
    PYPDE_REQUIRE(count >= 1 && count <= kMaxRhsMulti, "multi right-hand side: count %d not in 1..%d",
                  count, kMaxRhsMulti);
    // one kernel form for all members: rows_per_thread 0 or 1 one output per
    // thread, 4 or 8 a strip of that many rows
    const int rs = args[0].rows_per_thread <= 1 ? 1 : args[0].rows_per_thread;
    PYPDE_REQUIRE(rs == 1 || rs == 4 || rs == 8 || rs == 16,
                  "rows_per_thread must be 0, 1, 4, 8 or 16 (got %d)", args[0].rows_per_thread);
    RhsMulti* mm = static_cast<RhsMulti*>(table);
    memset(mm, 0, sizeof(RhsMulti));
    mm->count = count;
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        const int rc = check_rhs(&args[i]);
        if (rc != OK) return rc;
        const int ri = args[i].rows_per_thread <= 1 ? 1 : args[i].rows_per_thread;
        PYPDE_REQUIRE(ri == rs, "member %d: every member needs the same rows_per_thread", i);
        mm->m[i] = args[i];
        mm->start[i] = (int)total;
        if (rs == 16)
            total += 3LL * args[i].batch * ((args[i].M0 + 15) / 16) * ((args[i].M1 + 63) / 64);
        else
            total += (3LL * args[i].batch * ((args[i].M0 + rs - 1) / rs) * args[i].M1 + 255) / 256;
        PYPDE_REQUIRE(total < (1LL << 31), "multi right-hand side too large");
    }
    mm->start[count] = (int)total;
    mm->total = (int)total;
    bool diffusion = false;
    for (int i = 0; i < count; ++i) diffusion = diffusion || args[i].fdiff != nullptr;
    info->kind = PYPDE_MULTI_RHS;
    info->count = count;
    info->variant = rs;
    info->flags = diffusion ? 1 : 0;   // bit 0: some member has the explicit diffusion
    info->total = total;
    return OK;
}

extern "C" int pypde_rhs_multi(const void* table, const pypde_multi_info* info, void* stream)
{
    PYPDE_REQUIRE(table && info, "null argument");
    PYPDE_REQUIRE(info->kind == PYPDE_MULTI_RHS, "not a right-hand-side table");
    if (info->total == 0) return OK;
    const RhsMulti* mm = static_cast<const RhsMulti*>(table);
    switch (info->variant) {
    case 0:
    case 1:
        hipLaunchKernelGGL(rbc_rhs_banded_multi<1>, dim3((unsigned)info->total), dim3(256), 0,
                           as_stream(stream), mm);
        break;
    case 4:
        hipLaunchKernelGGL(rbc_rhs_banded_multi<4>, dim3((unsigned)info->total), dim3(256), 0,
                           as_stream(stream), mm);
        break;
    case 8:
        hipLaunchKernelGGL(rbc_rhs_banded_multi<8>, dim3((unsigned)info->total), dim3(256), 0,
                           as_stream(stream), mm);
        break;
    case 16:
        hipLaunchKernelGGL(rbc_rhs_banded_multi<16>, dim3((unsigned)info->total), dim3(256),
                           rhs_tile_lds_bytes(info->flags & 1),
                           as_stream(stream), mm);
        break;
    default:
        set_error("pypde_rhs_multi: unknown kernel form %d (the table's rows per thread)",
                  info->variant);
        return ERR_ARG;
    }
    return check_launch("pypde_rhs_multi");
}

extern "C" long long pypde_multi_table_bytes(int kind)
{
    switch (kind) {
    case PYPDE_MULTI_GEMM: return gemm_multi_bytes();
    case PYPDE_MULTI_RECURRENCE: return recurrence_multi_bytes();
    case PYPDE_MULTI_PRODUCTS: return products_multi_bytes();
    case PYPDE_MULTI_RHS: return (long long)sizeof(RhsMulti);
    default: return -1;
    }
}

extern "C" int pypde_rbc_rhs_banded(const pypde_rbc_rhs* a, void* stream)
{
    const int rc = check_rhs(a);
    if (rc != OK) return rc;
    if (a->batch == 0) return OK;
    const int rs = a->rows_per_thread == 0 ? 4 : a->rows_per_thread;
    PYPDE_REQUIRE(rs == 1 || rs == 4 || rs == 8 || rs == 16,
                  "rows_per_thread must be 0, 1, 4, 8 or 16 (got %d)", a->rows_per_thread);
    if (rs == 16) {
        const long long blocks =
            3LL * a->batch * ((a->M0 + 15) / 16) * ((a->M1 + 63) / 64);
        PYPDE_REQUIRE(blocks < (1LL << 31), "right-hand side too large");
        hipLaunchKernelGGL(rbc_rhs_tile_kernel, dim3((unsigned)blocks), dim3(256),
                           rhs_tile_lds_bytes(a->fdiff != nullptr),
                           as_stream(stream), *a);
        return check_launch("pypde_rbc_rhs_banded");
    }
    if (rs == 1) {
        const long long total = 3LL * a->M0 * a->M1 * a->batch;
        long long blocks = (total + 255) / 256;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(rbc_rhs_banded_kernel, dim3((unsigned)blocks), dim3(256), 0,
                           as_stream(stream), a->X, a->pres, a->cst, a->R, a->mdta, a->dta,
                           a->fdiff, a->tx, a->ty, a->M0, a->M1, a->N0, a->N1, a->batch);
        return check_launch("pypde_rbc_rhs_banded");
    }
    const long long total = 3LL * a->batch * ((a->M0 + rs - 1) / rs) * a->M1;
    const long long blocks = (total + 63) / 64;
    PYPDE_REQUIRE(blocks < (1LL << 31), "right-hand side too large");
    if (rs == 4)
        hipLaunchKernelGGL(rbc_rhs_strip_kernel<4>, dim3((unsigned)blocks), dim3(64), 0,
                           as_stream(stream), *a);
    else
        hipLaunchKernelGGL(rbc_rhs_strip_kernel<8>, dim3((unsigned)blocks), dim3(64), 0,
                           as_stream(stream), *a);
    return check_launch("pypde_rbc_rhs_banded");
}
