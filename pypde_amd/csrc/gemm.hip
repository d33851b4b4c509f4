// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
//   C[b] = alpha * op(A[b]) @ op(B[b]) + beta * D[b]      (row-major)
//
// This carries every dense operator of the spectral step: the Chebyshev
// transforms (as contractions against the collocation matrices, with stencil,
// derivative, padding and truncation folded into the matrix), the Poisson
// eigen-mixing Hy / Qy, and the pseudo-inverse right-hand sides.
//
// Block tile BM x BN (64x64 or 32x32), BK = 16, 256 threads = 4 wave64s in a
// 2x2 arrangement; each wave owns a (BM/2)x(BN/2) sub-tile made of 16x16 MFMA
// tiles.  Tiles of A and B are staged through LDS (k-major, one double of
// padding per row), the next K-slab is prefetched into registers while the
// current one feeds the MFMAs.
#include "common.h"
#include "../../include/pypde_amd.h"

namespace pypde {

typedef double v4d __attribute__((ext_vector_type(4)));

// Batch index of one operand: entry e is at  (e/d1 % m1)*s1 + (e/d2 % m2)*s2
// (m = 0 means no modulus).  A plain strided batch is {s, 0, 1, 0, 1, 0}.
struct BIdx {
    long long s1, s2;
    int d1, m1, d2, m2;
};

__device__ __forceinline__ long long bofs(int e, const BIdx& x)
{
    int i1 = e / x.d1;
    if (x.m1) i1 %= x.m1;
    int i2 = e / x.d2;
    if (x.m2) i2 %= x.m2;
    return (long long)i1 * x.s1 + (long long)i2 * x.s2;
}

struct GemmArgs {
    const double* A;
    const double* B;
    const double* D;
    double* C;
    long long lda, ldb, ldd, ldc;
    BIdx iA, iB, iD, iC;
    int m, n, k;
    double alpha, beta;
    const double* alpha_vec;  // alpha * alpha_vec[e / alpha_div] when given
    int alpha_div;
};

template <int BM, int BN, bool TA, bool TB>
__global__ void __launch_bounds__(256) dgemm_mfma_kernel(GemmArgs g)
{
    constexpr int BK = 16;
    constexpr int PA = BM + 1;
    constexpr int PB = BN + 1;
    constexpr int WM = BM / 2;  // wave tile rows
    constexpr int WN = BN / 2;
    constexpr int TM = WM / 16; // MFMA tiles per wave (rows)
    constexpr int TN = WN / 16;
    constexpr int LA = BM * BK / 256;  // A elements loaded per thread per slab
    constexpr int LB = BN * BK / 256;
    __shared__ double As[BK * PA];
    __shared__ double Bs[BK * PB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int bz = blockIdx.z;
    const int m0 = blockIdx.y * BM;
    const int n0 = blockIdx.x * BN;

    const double* __restrict__ A = g.A + bofs(bz, g.iA);
    const double* __restrict__ B = g.B + bofs(bz, g.iB);

    v4d acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[LA], rb[LB];

    auto load_slab = [&](int k0) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;  // r: row of op(A) within tile (m), c: k within slab
            if (!TA) { c = tid % BK; r = tid / BK + q * (256 / BK); }
            else     { r = tid % BM; c = tid / BM + q * (256 / BM); }
            const int gm = m0 + r, gk = k0 + c;
            double v = 0.0;
            if (gm < g.m && gk < g.k) v = TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + gk];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;  // r: k within slab, c: column of op(B) within tile (n)
            if (!TB) { c = tid % BN; r = tid / BN + q * (256 / BN); }
            else     { r = tid % BK; c = tid / BK + q * (256 / BK); }
            const int gk = k0 + r, gn = n0 + c;
            double v = 0.0;
            if (gk < g.k && gn < g.n) v = TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn];
            rb[q] = v;
        }
    };
    auto store_slab = [&]() {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            if (!TA) { c = tid % BK; r = tid / BK + q * (256 / BK); }
            else     { r = tid % BM; c = tid / BM + q * (256 / BM); }
            As[c * PA + r] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            if (!TB) { c = tid % BN; r = tid / BN + q * (256 / BN); }
            else     { r = tid % BK; c = tid / BK + q * (256 / BK); }
            Bs[r * PB + c] = rb[q];
        }
    };

    load_slab(0);
    for (int k0 = 0; k0 < g.k; k0 += BK) {
        __syncthreads();
        store_slab();
        __syncthreads();
        if (k0 + BK < g.k) load_slab(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = As[kr * PA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = Bs[kr * PB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }

    // Epilogue: for the f64 16x16x4 MFMA, accumulator item r of a lane holds
    // row (lane/16) + 4*r, column lane%16 of its 16x16 tile.
    double* __restrict__ C = g.C + bofs(bz, g.iC);
    const double* __restrict__ Dm = g.D ? g.D + bofs(bz, g.iD) : nullptr;
    const double alpha = g.alpha_vec ? g.alpha * g.alpha_vec[bz / g.alpha_div] : g.alpha;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
                const int gn = n0 + wn * WN + j * 16 + (lane & 15);
                if (gm < g.m && gn < g.n) {
                    double v = alpha * acc[i][j][r];
                    if (Dm && g.beta != 0.0) v += g.beta * Dm[(long long)gm * g.ldd + gn];
                    C[(long long)gm * g.ldc + gn] = v;
                }
            }
}

template <int BM, int BN>
static void launch_gemm(bool ta, bool tb, const GemmArgs& g, int batch, hipStream_t s)
{
    dim3 grid((g.n + BN - 1) / BN, (g.m + BM - 1) / BM, batch);
    dim3 block(256);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, false, false>), grid, block, 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, false, true>), grid, block, 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, true, false>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, true, true>), grid, block, 0, s, g);
}

static BIdx bidx(const pypde_bidx& b)
{
    BIdx r{b.s1, b.s2, b.d1 < 1 ? 1 : b.d1, b.m1, b.d2 < 1 ? 1 : b.d2, b.m2};
    return r;
}

static BIdx strided(long long s)
{
    BIdx r{s, 0, 1, 0, 1, 0};
    return r;
}

static int launch(int transA, int transB, GemmArgs& g, int batch, hipStream_t stream)
{
    PYPDE_REQUIRE(g.m >= 0 && g.n >= 0 && g.k >= 0 && batch >= 0, "negative GEMM size");
    if (g.m == 0 || g.n == 0 || batch == 0) return OK;
    PYPDE_REQUIRE(g.A && g.B && g.C, "null GEMM operand");
    PYPDE_REQUIRE(batch <= 65535, "GEMM batch %d exceeds 65535", batch);
    // Small problems use 32x32 tiles so that more CUs get work.
    const long long tiles64 = (long long)((g.m + 63) / 64) * ((g.n + 63) / 64) * batch;
    if (tiles64 >= 512) launch_gemm<64, 64>(transA, transB, g, batch, stream);
    else launch_gemm<32, 32>(transA, transB, g, batch, stream);
    return check_launch("pypde_dgemm");
}

int dgemm(int transA, int transB, int m, int n, int k, double alpha, const double* A,
          long long lda, long long strideA, const double* B, long long ldb, long long strideB,
          double beta, const double* D, long long ldd, long long strideD, double* C,
          long long ldc, long long strideC, int batch, hipStream_t stream)
{
    GemmArgs g{A, B, D, C, lda, ldb, ldd, ldc, strided(strideA), strided(strideB),
               strided(strideD), strided(strideC), m, n, k, alpha, beta, nullptr, 1};
    return launch(transA, transB, g, batch, stream);
}

int dgemm_batched(const pypde_gemm_desc* d, hipStream_t stream)
{
    PYPDE_REQUIRE(d, "null GEMM descriptor");
    PYPDE_REQUIRE(!d->alpha_vec || d->alpha_div >= 1, "alpha_div must be >= 1");
    GemmArgs g{d->A, d->B, d->D, d->C, d->lda, d->ldb, d->ldd, d->ldc, bidx(d->iA), bidx(d->iB),
               bidx(d->iD), bidx(d->iC), d->m, d->n, d->k, d->alpha, d->beta, d->alpha_vec,
               d->alpha_div < 1 ? 1 : d->alpha_div};
    return launch(d->transA, d->transB, g, d->batch, stream);
}

}  // namespace pypde

extern "C" int pypde_dgemm(int transA, int transB, int m, int n, int k, double alpha,
                           const double* A, long long lda, long long strideA, const double* B,
                           long long ldb, long long strideB, double beta, const double* D,
                           long long ldd, long long strideD, double* C, long long ldc,
                           long long strideC, int batch, void* stream)
{
    return pypde::dgemm(transA, transB, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, D,
                        ldd, strideD, C, ldc, strideC, batch, pypde::as_stream(stream));
}

extern "C" int pypde_dgemm_batched(const pypde_gemm_desc* d, void* stream)
{
    return pypde::dgemm_batched(d, pypde::as_stream(stream));
}
