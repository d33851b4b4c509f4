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

struct GemmArgs {
    const double* A;
    const double* B;
    const double* D;
    double* C;
    long long lda, ldb, ldd, ldc;
    long long sA, sB, sD, sC;
    int m, n, k;
    double alpha, beta;
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

    const double* __restrict__ A = g.A + (long long)bz * g.sA;
    const double* __restrict__ B = g.B + (long long)bz * g.sB;

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

    // Epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each 16x16 tile.
    double* __restrict__ C = g.C + (long long)bz * g.sC;
    const double* __restrict__ Dm = g.D ? g.D + (long long)bz * g.sD : nullptr;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn * WN + j * 16 + (lane & 15);
                if (gm < g.m && gn < g.n) {
                    double v = g.alpha * acc[i][j][r];
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

int dgemm(int transA, int transB, int m, int n, int k, double alpha, const double* A,
          long long lda, long long strideA, const double* B, long long ldb, long long strideB,
          double beta, const double* D, long long ldd, long long strideD, double* C,
          long long ldc, long long strideC, int batch, hipStream_t stream)
{
    PYPDE_REQUIRE(m >= 0 && n >= 0 && k >= 0 && batch >= 0, "negative GEMM size");
    if (m == 0 || n == 0 || batch == 0) return OK;
    PYPDE_REQUIRE(A && B && C, "null GEMM operand");
    PYPDE_REQUIRE(batch <= 65535, "GEMM batch %d exceeds 65535", batch);
    GemmArgs g{A, B, D, C, lda, ldb, ldd, ldc, strideA, strideB, strideD, strideC, m, n, k, alpha, beta};
    // Small problems use 32x32 tiles so that more CUs get work.
    const long long tiles64 = (long long)((m + 63) / 64) * ((n + 63) / 64) * batch;
    if (tiles64 >= 512) launch_gemm<64, 64>(transA, transB, g, batch, stream);
    else launch_gemm<32, 32>(transA, transB, g, batch, stream);
    return check_launch("pypde_dgemm");
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
