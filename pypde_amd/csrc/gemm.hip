// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
//   C[b] = alpha * op(A[b]) @ op(B[b]) + beta * D[b]      (row-major)
//
// This carries every dense operator of the spectral step: the Chebyshev
// transforms (as contractions against the collocation matrices, with stencil,
// derivative, padding and truncation folded into the matrix), the Poisson
// eigen-mixing Hy / Qy, and the pseudo-inverse right-hand sides.
//
// Two kernel families:
//
//  * LDS-tiled (throughput): block tile BM x BN (128x128 over 8 wave64s in a
//    4x2 arrangement, or 64x64 / 32x32 over 4 waves in 2x2), BK = 16 (32 or
//    64 in variants 33..38); each
//    wave owns a (BM/WGM)x(BN/WGN) sub-tile made of 16x16 MFMA tiles.  Tiles
//    of A and B are staged through LDS (k-major, one double of padding per
//    row), the next K-slab is prefetched into registers while the current one
//    feeds the MFMAs; the double-buffered form stores it into the other LDS
//    buffer and needs one barrier per slab.  128x128 tiles reach about 70% of
//    the FP64 MFMA peak on batched 256-point transforms, 64x64 about 55%.
//
//  * Split-K, LDS-staged (latency): KW waves share one output tile, each
//    staging its own K range through a private LDS slab with coalesced loads;
//    partial tiles are summed through LDS.
//
//  * Split-K, register-direct (latency): one workgroup of KW waves owns a
//    (16 TM) x (16 TN) output tile.  The K range is cut into 4 KW contiguous
//    chunks, one per (wave, lane group of 16): the f64 MFMA sums over its four
//    lane groups, so giving each lane group its own chunk makes every MFMA step
//    consume four chunks at once.  Operands are loaded straight into registers
//    (no LDS staging, no barriers in the K loop); the KW partial tiles are
//    reduced through LDS in the epilogue.  Single-instance 256x256 operators
//    give 64-200 32x32 tiles -- too few for 256 CUs -- and this variant puts
//    4-16x more waves on the same problem with a K loop 4 KW times shorter.
//
// The variant is chosen per call (pypde_gemm_desc.variant, filled by the
// host-side autotuner) or by a size heuristic when the variant is 0.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <cstring>

namespace pypde {

typedef double v4d __attribute__((ext_vector_type(4)));

// Orders LDS traffic within one wavefront (other waves are not synchronised).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
    const double* E;          // optional second addend
    long long lde;
    BIdx iE;
    const double* gamma_vec;  // gamma * gamma_vec[e / gamma_div] scales E when given
    double gamma;
    int gamma_div;
    int zero_first;           // set element (0,0) of every C[e] to 0 (pressure gauge)
    long long csn;            // column stride of C, D and E (1: contiguous rows)
    long long ska;            // element stride of A along k when A is not transposed
};

// Output update shared by every kernel family:
//   C = (beta D - (g_e E) gamma) + alpha_e acc      (E term and D term optional)
// in the order of the reference's pressure update (pres -= nu div beta, then
// pres += ...), with the gauge C[0,0] = 0 applied last when requested.  The
// arithmetic is spelled out (one FMA for the E term, every other product and
// sum rounded on its own) so that each kernel family rounds it the same way.
struct Epilogue {
    double* C;
    const double* D;
    const double* E;
    long long ldc, ldd, lde, csn;
    double alpha, beta, ge;
    bool zero_first;
    __device__ __forceinline__ Epilogue(const GemmArgs& g, int bz)
    {
        C = g.C + bofs(bz, g.iC);
        D = (g.D && g.beta != 0.0) ? g.D + bofs(bz, g.iD) : nullptr;
        E = g.E ? g.E + bofs(bz, g.iE) : nullptr;
        ldc = g.ldc;
        ldd = g.ldd;
        lde = g.lde;
        csn = g.csn;
        alpha = g.alpha_vec ? g.alpha * g.alpha_vec[bz / g.alpha_div] : g.alpha;
        beta = g.beta;
        ge = g.gamma_vec ? g.gamma_vec[bz / g.gamma_div] : 1.0;
        beta_e_gamma = g.gamma;
        zero_first = g.zero_first != 0;
    }
    // beta D - (g_e E) gamma at (gm, gn), for the terms present (HD, HE)
    template <bool HD, bool HE>
    __device__ __forceinline__ double addend(int gm, int gn) const
    {
#pragma clang fp contract(off)
        const long long cn = (long long)gn * csn;
        double d = HD ? beta * D[(long long)gm * ldd + cn] : 0.0;
        if (HE) d = __builtin_fma(-(ge * E[(long long)gm * lde + cn]), beta_e_gamma, d);
        return d;
    }
    __device__ __forceinline__ void put(int gm, int gn, double acc, double add, bool has_add) const
    {
#pragma clang fp contract(off)
        double v = alpha * acc;
        if (has_add) v = add + v;
        if (zero_first && gm == 0 && gn == 0) v = 0.0;
        C[(long long)gm * ldc + (long long)gn * csn] = v;
    }
    double beta_e_gamma;
};

// Stores the NE output elements a thread owns; elem(e, gm, gn, ok, v) gives
// element e's position, whether it lies inside C, and its accumulated value.
// With a D or E term the elements go in groups of four: the group's D / E
// loads are issued first, at a clamped address (element (0, 0)) for elements
// outside C, so that the wave waits once per group instead of once per element
// behind a branch; groups of four keep the registers held by the loads few.
constexpr int kEpiGroup = 4;  // elements per group of epilogue loads
template <bool HD, bool HE, int NE, class Elem>
__device__ __forceinline__ void epilogue_groups(const Epilogue& ep, Elem elem)
{
#pragma unroll
    for (int c = 0; c < NE; c += kEpiGroup) {
        double add[kEpiGroup];
#pragma unroll
        for (int u = 0; u < kEpiGroup; ++u) {
            if (c + u >= NE) break;
            int gm, gn;
            bool ok;
            double v;
            elem(c + u, gm, gn, ok, v);
            add[u] = ep.addend<HD, HE>(ok ? gm : 0, ok ? gn : 0);
        }
#pragma unroll
        for (int u = 0; u < kEpiGroup; ++u) {
            if (c + u >= NE) break;
            int gm, gn;
            bool ok;
            double v;
            elem(c + u, gm, gn, ok, v);
            if (ok) ep.put(gm, gn, v, add[u], true);
        }
    }
}

template <int NE, class Elem>
__device__ __forceinline__ void epilogue_store(const Epilogue& ep, Elem elem)
{
    if (ep.D && ep.E) epilogue_groups<true, true, NE>(ep, elem);
    else if (ep.D) epilogue_groups<true, false, NE>(ep, elem);
    else if (ep.E) epilogue_groups<false, true, NE>(ep, elem);
    else {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int gm, gn;
            bool ok;
            double v;
            elem(e, gm, gn, ok, v);
            if (ok) ep.put(gm, gn, v, 0.0, false);
        }
    }
}

// GEMM tiles go through xcd_remap (common.h): the tiles that one XCD runs are
// a contiguous range, so the operand panels they share stay in one L2.

// Tile (bx, by, bz) of a 3-D grid after the XCD renumbering.
__device__ __forceinline__ void grid_tile(int& bx, int& by, int& bz)
{
    const int tx = gridDim.x, ty = gridDim.y;
    const int total = tx * ty * gridDim.z;
    const int t = xcd_remap(blockIdx.x + tx * (blockIdx.y + ty * blockIdx.z), total);
    bx = t % tx;
    const int r = t / tx;
    by = r % ty;
    bz = r / ty;
}

// Several independent GEMMs of one kernel form in one launch (grouped
// launch): the grid is the concatenation of the members' tile grids, block t
// runs tile t - start[i] of member i.  Members share the kernel template, so
// each computes exactly what its own launch would.
constexpr int kMaxGroup = 4;
struct GemmGroup {
    GemmArgs a[kMaxGroup];
    int tx[kMaxGroup], ty[kMaxGroup];
    int start[kMaxGroup + 1];
    int count;
};

__device__ __forceinline__ int group_tile(const GemmGroup& gg, int& bx, int& by, int& bz)
{
    const int t = xcd_remap(blockIdx.x, gg.start[gg.count]);
    int i = 0;
    while (i + 1 < gg.count && t >= gg.start[i + 1]) ++i;
    const int local = t - gg.start[i];
    bx = local % gg.tx[i];
    const int r = local / gg.tx[i];
    by = r % gg.ty[i];
    bz = r / gg.ty[i];
    return i;
}

// PIPE 0: one LDS buffer, the next slab's loads in flight during a slab; 1:
// two LDS buffers (one barrier per slab); 2: one LDS buffer, two slabs of
// loads in flight (registers for two).  Every form runs the same slabs in the
// same order, so results are bit-identical.
template <int BM, int BN, int WGM, int WGN, int PIPE, bool TA, bool TB, int BK, int PAD>
__device__ __forceinline__ void dgemm_mfma_body(const GemmArgs& g, int bx, int by, int bz)
{
    constexpr int NT = 64 * WGM * WGN;  // threads
    constexpr int PA = BM + PAD;  // LDS row pitch (doubles)
    constexpr int PB = BN + PAD;
    constexpr int WM = BM / WGM;  // wave tile rows
    constexpr int WN = BN / WGN;
    constexpr int TM = WM / 16;   // MFMA tiles per wave (rows)
    constexpr int TN = WN / 16;
    constexpr int LA = BM * BK / NT;  // A elements loaded per thread per slab
    constexpr int LB = BN * BK / NT;
    constexpr int NBUF = PIPE == 1 ? 2 : 1;
    constexpr int DEPTH = PIPE == 2 ? 2 : 1;
    static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");
    static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "tile / thread mismatch");
    static_assert(NT % BK == 0 && NT % BM == 0 && NT % BN == 0, "load mapping");
    __shared__ double As[NBUF][BK * PA];
    __shared__ double Bs[NBUF][BK * PB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WGN, wn = wave % WGN;
    const int m0 = by * BM;
    const int n0 = bx * BN;

    const double* __restrict__ A = g.A + bofs(bz, g.iA);
    const double* __restrict__ B = g.B + bofs(bz, g.iB);

    v4d acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[DEPTH][LA], rb[DEPTH][LB];

    // (r, c) of the q-th element a thread moves: consecutive threads walk the
    // contiguous dimension of the operand in memory.
    auto a_rc = [&](int q, int& r, int& c) {  // r: row of op(A) in tile, c: k in slab
        if (!TA) { c = tid % BK; r = tid / BK + q * (NT / BK); }
        else     { r = tid % BM; c = tid / BM + q * (NT / BM); }
    };
    auto b_rc = [&](int q, int& r, int& c) {  // r: k in slab, c: column of op(B) in tile
        if (!TB) { c = tid % BN; r = tid / BN + q * (NT / BN); }
        else     { r = tid % BK; c = tid / BK + q * (NT / BK); }
    };
    auto load_slab = [&](int st, int k0) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            const int gm = m0 + r, gk = k0 + c;
            double v = 0.0;
            if (gm < g.m && gk < g.k) v = TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + (long long)gk * g.ska];
            ra[st][q] = v;
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            const int gk = k0 + r, gn = n0 + c;
            double v = 0.0;
            if (gk < g.k && gn < g.n) v = TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn];
            rb[st][q] = v;
        }
    };
    auto store_slab = [&](int buf, int st) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            As[buf][c * PA + r] = ra[st][q];
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            Bs[buf][r * PB + c] = rb[st][q];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = As[buf][kr * PA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = Bs[buf][kr * PB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    };

    if (PIPE == 1) {
        // one barrier per slab: the next slab is stored into the other buffer
        load_slab(0, 0);
        store_slab(0, 0);
        __syncthreads();
        int cur = 0;
        for (int k0 = 0; k0 < g.k; k0 += BK) {
            const bool more = k0 + BK < g.k;
            if (more) load_slab(0, k0 + BK);
            compute(cur);
            if (more) store_slab(cur ^ 1, 0);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        // slab (i * DEPTH + st) * BK lives in register stage st
#pragma unroll
        for (int st = 0; st < DEPTH; ++st)
            if (st * BK < g.k) load_slab(st, st * BK);
        for (int k0 = 0; k0 < g.k; k0 += DEPTH * BK) {
#pragma unroll
            for (int st = 0; st < DEPTH; ++st) {
                const int ks = k0 + st * BK;
                if (ks >= g.k) break;  // uniform across the block
                __syncthreads();
                store_slab(0, st);
                __syncthreads();
                if (ks + DEPTH * BK < g.k) load_slab(st, ks + DEPTH * BK);
                compute(0);
            }
        }
    }

    // Epilogue: for the f64 16x16x4 MFMA, accumulator item r of a lane holds
    // row (lane/16) + 4*r, column lane%16 of its 16x16 tile.
    const Epilogue ep(g, bz);
    epilogue_store<TM * TN * 4>(ep, [&](int e, int& gm, int& gn, bool& ok, double& v) {
        const int i = e / (TN * 4), j = (e / 4) % TN, r = e % 4;
        gm = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
        gn = n0 + wn * WN + j * 16 + (lane & 15);
        ok = gm < g.m && gn < g.n;
        v = acc[i][j][r];
    });
}

template <int BM, int BN, int WGM, int WGN, int PIPE, bool TA, bool TB, int BK, int PAD>
__global__ void __launch_bounds__(64 * WGM * WGN) dgemm_mfma_kernel(GemmArgs g)
{
    int bx, by, bz;
    grid_tile(bx, by, bz);
    dgemm_mfma_body<BM, BN, WGM, WGN, PIPE, TA, TB, BK, PAD>(g, bx, by, bz);
}

template <int BM, int BN, int WGM, int WGN, int PIPE, bool TA, bool TB, int BK, int PAD>
__global__ void __launch_bounds__(64 * WGM * WGN) dgemm_mfma_group(GemmGroup gg)
{
    int bx, by, bz;
    const int i = group_tile(gg, bx, by, bz);
    dgemm_mfma_body<BM, BN, WGM, WGN, PIPE, TA, TB, BK, PAD>(gg.a[i], bx, by, bz);
}

// Many independent GEMMs of one LDS-tiled kernel form in one launch, their
// arguments in a table in device memory (up to kMaxMulti members, beyond what
// kernel arguments hold): a heterogeneous batch runs each operator of the step
// for every size class in one launch.  Block t runs tile t - start[i] of the
// member i found by binary search; members share the kernel template, so each
// computes exactly what its own launch of that variant computes.
constexpr int kMaxMulti = 64;
struct GemmMulti {
    int count, total;
    int start[kMaxMulti + 1];
    int tx[kMaxMulti], ty[kMaxMulti];
    GemmArgs a[kMaxMulti];
};

__device__ __forceinline__ int multi_tile(const GemmMulti* __restrict__ mm, int& bx, int& by, int& bz)
{
    const int t = xcd_remap(blockIdx.x, mm->total);
    int lo = 0, hi = mm->count - 1;
    while (lo < hi) {  // the last member whose range starts at or before t
        const int mid = (lo + hi + 1) >> 1;
        if (mm->start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const int local = t - mm->start[lo];
    bx = local % mm->tx[lo];
    const int r = local / mm->tx[lo];
    by = r % mm->ty[lo];
    bz = r / mm->ty[lo];
    return lo;
}

template <int BM, int BN, int WGM, int WGN, int PIPE, bool TA, bool TB, int BK, int PAD>
__global__ void __launch_bounds__(64 * WGM * WGN) dgemm_mfma_multi(const GemmMulti* __restrict__ mm)
{
    int bx, by, bz;
    const int i = multi_tile(mm, bx, by, bz);
    const GemmArgs g = mm->a[i];
    dgemm_mfma_body<BM, BN, WGM, WGN, PIPE, TA, TB, BK, PAD>(g, bx, by, bz);
}

template <int BM, int BN, int WGM, int WGN, int PIPE, int BK = 16, int PAD = 1>
static void launch_gemm_multi(bool ta, bool tb, const GemmMulti* mm, int total, hipStream_t s)
{
    dim3 grid(total);
    dim3 block(64 * WGM * WGN);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_mfma_multi<BM, BN, WGM, WGN, PIPE, false, false, BK, PAD>), grid, block, 0, s, mm);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_mfma_multi<BM, BN, WGM, WGN, PIPE, false, true, BK, PAD>), grid, block, 0, s, mm);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_mfma_multi<BM, BN, WGM, WGN, PIPE, true, false, BK, PAD>), grid, block, 0, s, mm);
    else hipLaunchKernelGGL((dgemm_mfma_multi<BM, BN, WGM, WGN, PIPE, true, true, BK, PAD>), grid, block, 0, s, mm);
}

// Block tile (rows, columns) of the LDS-tiled variants; 0 for the others.
static void multi_tile_shape(int v, int& bm, int& bn)
{
    bm = bn = 0;
    switch (v) {
    case 1: case 25: case 37: bm = 64; bn = 64; break;
    case 3: case 4: case 26: bm = 128; bn = 128; break;
    case 5: case 7: case 38: bm = 128; bn = 64; break;
    case 6: case 24: bm = 64; bn = 128; break;
    case 2: case 27: case 28: case 29: case 30: case 31: case 32: case 33: case 34: case 35:
    case 36: case 39: case 40: case 41: case 42: bm = 32; bn = 32; break;
    default: break;
    }
}

template <int BM, int BN, int WGM, int WGN, int PIPE, int BK = 16, int PAD = 1>
static void launch_gemm(bool ta, bool tb, const GemmArgs& g, int batch, hipStream_t s)
{
    dim3 grid((g.n + BN - 1) / BN, (g.m + BM - 1) / BM, batch);
    dim3 block(64 * WGM * WGN);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, WGM, WGN, PIPE, false, false, BK, PAD>), grid, block, 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, WGM, WGN, PIPE, false, true, BK, PAD>), grid, block, 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, WGM, WGN, PIPE, true, false, BK, PAD>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, WGM, WGN, PIPE, true, true, BK, PAD>), grid, block, 0, s, g);
}

static void group_grid(GemmGroup& gg, const int* batch, int bm, int bn)
{
    gg.start[0] = 0;
    for (int i = 0; i < gg.count; ++i) {
        gg.tx[i] = (gg.a[i].n + bn - 1) / bn;
        gg.ty[i] = (gg.a[i].m + bm - 1) / bm;
        const long long t = (gg.a[i].m > 0 && gg.a[i].n > 0) ? (long long)gg.tx[i] * gg.ty[i] * batch[i] : 0;
        gg.start[i + 1] = gg.start[i] + (int)t;
    }
}

template <int BM, int BN, int WGM, int WGN, int PIPE, int BK = 16, int PAD = 1>
static void launch_gemm_group(bool ta, bool tb, GemmGroup& gg, const int* batch, hipStream_t s)
{
    group_grid(gg, batch, BM, BN);
    if (gg.start[gg.count] == 0) return;
    dim3 grid(gg.start[gg.count]);
    dim3 block(64 * WGM * WGN);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_mfma_group<BM, BN, WGM, WGN, PIPE, false, false, BK, PAD>), grid, block, 0, s, gg);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_mfma_group<BM, BN, WGM, WGN, PIPE, false, true, BK, PAD>), grid, block, 0, s, gg);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_mfma_group<BM, BN, WGM, WGN, PIPE, true, false, BK, PAD>), grid, block, 0, s, gg);
    else hipLaunchKernelGGL((dgemm_mfma_group<BM, BN, WGM, WGN, PIPE, true, true, BK, PAD>), grid, block, 0, s, gg);
}

// ---------------------------------------------------------------------------
// Split-K register-direct kernel.
template <int TM, int TN, int KW, bool TA, bool TB>
__device__ __forceinline__ void dgemm_splitk_body(const GemmArgs& g, int bx, int by, int bz)
{
    constexpr int U = 4;  // MFMA steps per register block
    __shared__ double red[KW * TM * TN * 4 * 64];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r16 = lane & 15;
    const int q = lane >> 4;
    const int m0 = by * (16 * TM);
    const int n0 = bx * (16 * TN);

    const double* __restrict__ A = g.A + bofs(bz, g.iA);
    const double* __restrict__ B = g.B + bofs(bz, g.iB);

    // this lane group's K chunk [kb, ke)
    const int P = 4 * KW;
    const int c = (g.k + P - 1) / P;
    const int kb = (wave * 4 + q) * c;
    const int ke = min(kb + c, g.k);

    // row / column of op(A) / op(B) each lane feeds, with clamped addresses
    const double* pa[TM];
    bool va[TM];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int gm = m0 + i * 16 + r16;
        va[i] = gm < g.m;
        const int gmc = va[i] ? gm : 0;
        pa[i] = TA ? A + gmc : A + (long long)gmc * g.lda;
    }
    const double* pb[TN];
    bool vb[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int gn = n0 + j * 16 + r16;
        vb[j] = gn < g.n;
        const int gnc = vb[j] ? gn : 0;
        pb[j] = TB ? B + (long long)gnc * g.ldb : B + gnc;
    }
    const long long sa = TA ? g.lda : g.ska;  // stride of one k step
    const long long sb = TB ? 1 : g.ldb;

    v4d acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[2][U][TM], rb[2][U][TN];
    auto load = [&](int buf, int k0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k0 + u;
            const bool ok = kk < ke;
#pragma unroll
            for (int i = 0; i < TM; ++i) ra[buf][u][i] = (ok && va[i]) ? pa[i][kk * sa] : 0.0;
#pragma unroll
            for (int j = 0; j < TN; ++j) rb[buf][u][j] = (ok && vb[j]) ? pb[j][kk * sb] : 0.0;
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[buf][u][i], rb[buf][u][j],
                                                                      acc[i][j], 0, 0, 0);
    };
    // The trip count is uniform over the workgroup (c steps per lane group);
    // lanes past the end of their chunk feed zeros.
    const int nblk = (c + U - 1) / U;
    if (nblk > 0) load(0, kb);
    int blk = 0;
    for (; blk + 2 <= nblk; blk += 2) {
        load(1, kb + (blk + 1) * U);
        compute(0);
        if (blk + 2 < nblk) load(0, kb + (blk + 2) * U);
        compute(1);
    }
    if (blk < nblk) compute(0);

    // reduce the KW partial tiles
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                red[((wave * TM * TN + i * TN + j) * 4 + r) * 64 + lane] = acc[i][j][r];
    __syncthreads();

    const Epilogue ep(g, bz);
    constexpr int E = TM * TN * 4 * 64;
    // the tile's E items over 64 KW threads: item tid + q * 64 KW is a thread's q-th
    epilogue_store<(E + 64 * KW - 1) / (64 * KW)>(ep, [&](int q, int& gm, int& gn, bool& ok, double& v) {
        const int e = tid + q * 64 * KW;
        const int ec = e < E ? e : 0;
        v = 0.0;
#pragma unroll
        for (int w = 0; w < KW; ++w) v += red[w * E + ec];
        const int l = ec & 63;
        const int r = (ec >> 6) & 3;
        const int t = ec >> 8;
        gm = m0 + (t / TN) * 16 + (l >> 4) + 4 * r;
        gn = n0 + (t % TN) * 16 + (l & 15);
        ok = e < E && gm < g.m && gn < g.n;
    });
}

template <int TM, int TN, int KW, bool TA, bool TB>
__global__ void __launch_bounds__(64 * KW) dgemm_splitk_kernel(GemmArgs g)
{
    int bx, by, bz;
    grid_tile(bx, by, bz);
    dgemm_splitk_body<TM, TN, KW, TA, TB>(g, bx, by, bz);
}

template <int TM, int TN, int KW, bool TA, bool TB>
__global__ void __launch_bounds__(64 * KW) dgemm_splitk_group(GemmGroup gg)
{
    int bx, by, bz;
    const int i = group_tile(gg, bx, by, bz);
    dgemm_splitk_body<TM, TN, KW, TA, TB>(gg.a[i], bx, by, bz);
}

template <int TM, int TN, int KW>
static void launch_splitk(bool ta, bool tb, const GemmArgs& g, int batch, hipStream_t s)
{
    dim3 grid((g.n + 16 * TN - 1) / (16 * TN), (g.m + 16 * TM - 1) / (16 * TM), batch);
    dim3 block(64 * KW);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_splitk_kernel<TM, TN, KW, false, false>), grid, block, 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_splitk_kernel<TM, TN, KW, false, true>), grid, block, 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_splitk_kernel<TM, TN, KW, true, false>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((dgemm_splitk_kernel<TM, TN, KW, true, true>), grid, block, 0, s, g);
}

template <int TM, int TN, int KW>
static void launch_splitk_group(bool ta, bool tb, GemmGroup& gg, const int* batch, hipStream_t s)
{
    group_grid(gg, batch, 16 * TM, 16 * TN);
    if (gg.start[gg.count] == 0) return;
    dim3 grid(gg.start[gg.count]);
    dim3 block(64 * KW);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_splitk_group<TM, TN, KW, false, false>), grid, block, 0, s, gg);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_splitk_group<TM, TN, KW, false, true>), grid, block, 0, s, gg);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_splitk_group<TM, TN, KW, true, false>), grid, block, 0, s, gg);
    else hipLaunchKernelGGL((dgemm_splitk_group<TM, TN, KW, true, true>), grid, block, 0, s, gg);
}

// ---------------------------------------------------------------------------
// Split-K, LDS-staged kernel (latency): KW waves share one BM x BN output tile;
// wave w owns a contiguous K range and stages its own 16-deep slabs of A and B
// through a private LDS region with coalesced loads (consecutive lanes walk the
// contiguous dimension of each operand), so the K loop has no block barriers.
// DEPTH slabs of loads are in flight per wave (registers of DEPTH slabs; the
// slab order and so the sums are the same for every DEPTH).
// The KW partial tiles are summed through LDS at the end.
template <int BM, int BN, int KW, bool TA, bool TB, int DEPTH>
__device__ __forceinline__ void dgemm_wsplit_body(const GemmArgs& g, int bx, int by, int bz)
{
    constexpr int BK = 16;
    constexpr int PA = BM + 1, PB = BN + 1;
    constexpr int SLAB = BK * PA + BK * PB;  // doubles per wave
    constexpr int TM = BM / 16, TN = BN / 16;
    constexpr int LA = BM * BK / 64, LB = BN * BK / 64;
    constexpr int E = BM * BN;
    constexpr int AREA = (KW * SLAB > KW * E) ? KW * SLAB : KW * E;
    static_assert(64 % BK == 0 && 64 % BM == 0 && 64 % BN == 0, "load mapping: BM, BN divide 64");
    __shared__ double lds[AREA];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int m0 = by * BM;
    const int n0 = bx * BN;
    const double* __restrict__ A = g.A + bofs(bz, g.iA);
    const double* __restrict__ B = g.B + bofs(bz, g.iB);
    double* As = lds + wave * SLAB;
    double* Bs = As + BK * PA;

    // K range of this wave, in whole slabs
    const int kc = ((g.k + KW - 1) / KW + BK - 1) / BK * BK;
    const int kb = wave * kc;
    const int ke = min(kb + kc, g.k);

    v4d acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[DEPTH][LA], rb[DEPTH][LB];
    auto a_rc = [&](int q, int& r, int& c) {
        if (!TA) { c = lane % BK; r = lane / BK + q * (64 / BK); }
        else     { r = lane % BM; c = lane / BM + q * (64 / BM); }
    };
    auto b_rc = [&](int q, int& r, int& c) {
        if (!TB) { c = lane % BN; r = lane / BN + q * (64 / BN); }
        else     { r = lane % BK; c = lane / BK + q * (64 / BK); }
    };
    auto load = [&](int st, int k0) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            const int gm = m0 + r, gk = k0 + c;
            ra[st][q] = (gm < g.m && gk < ke) ? (TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + (long long)gk * g.ska]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            const int gk = k0 + r, gn = n0 + c;
            rb[st][q] = (gk < ke && gn < g.n) ? (TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn]) : 0.0;
        }
    };
    auto store = [&](int st) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            As[c * PA + r] = ra[st][q];
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            Bs[r * PB + c] = rb[st][q];
        }
    };

    // slab kb + (i * DEPTH + st) * BK lives in register stage st
#pragma unroll
    for (int st = 0; st < DEPTH; ++st)
        if (kb + st * BK < ke) load(st, kb + st * BK);
    for (int k0 = kb; k0 < ke; k0 += DEPTH * BK) {
#pragma unroll
        for (int st = 0; st < DEPTH; ++st) {
            const int ks = k0 + st * BK;
            if (ks >= ke) break;  // uniform across the wave
            wave_lds_sync();  // previous slab fully read
            store(st);
            wave_lds_sync();
            if (ks + DEPTH * BK < ke) load(st, ks + DEPTH * BK);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                const int kr = kk + (lane >> 4);
                double af[TM], bf[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) af[i] = As[kr * PA + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < TN; ++j) bf[j] = Bs[kr * PB + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    __syncthreads();  // every wave is done with its slabs; the area becomes the reduction buffer
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                lds[wave * E + ((i * TN + j) * 4 + r) * 64 + lane] = acc[i][j][r];
    __syncthreads();

    const Epilogue ep(g, bz);
    // the tile's E items over 64 KW threads: item tid + q * 64 KW is a thread's q-th
    epilogue_store<(E + 64 * KW - 1) / (64 * KW)>(ep, [&](int q, int& gm, int& gn, bool& ok, double& v) {
        const int e = tid + q * 64 * KW;
        const int ec = e < E ? e : 0;
        v = 0.0;
#pragma unroll
        for (int w = 0; w < KW; ++w) v += lds[w * E + ec];
        const int l = ec & 63;
        const int r = (ec >> 6) & 3;
        const int t = ec >> 8;
        gm = m0 + (t / TN) * 16 + (l >> 4) + 4 * r;
        gn = n0 + (t % TN) * 16 + (l & 15);
        ok = e < E && gm < g.m && gn < g.n;
    });
}

template <int BM, int BN, int KW, bool TA, bool TB, int DEPTH>
__global__ void __launch_bounds__(64 * KW) dgemm_wsplit_kernel(GemmArgs g)
{
    int bx, by, bz;
    grid_tile(bx, by, bz);
    dgemm_wsplit_body<BM, BN, KW, TA, TB, DEPTH>(g, bx, by, bz);
}

template <int BM, int BN, int KW, bool TA, bool TB, int DEPTH>
__global__ void __launch_bounds__(64 * KW) dgemm_wsplit_group(GemmGroup gg)
{
    int bx, by, bz;
    const int i = group_tile(gg, bx, by, bz);
    dgemm_wsplit_body<BM, BN, KW, TA, TB, DEPTH>(gg.a[i], bx, by, bz);
}

template <int BM, int BN, int KW, int DEPTH = 1>
static void launch_wsplit(bool ta, bool tb, const GemmArgs& g, int batch, hipStream_t s)
{
    dim3 grid((g.n + BN - 1) / BN, (g.m + BM - 1) / BM, batch);
    dim3 block(64 * KW);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_wsplit_kernel<BM, BN, KW, false, false, DEPTH>), grid, block, 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_wsplit_kernel<BM, BN, KW, false, true, DEPTH>), grid, block, 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_wsplit_kernel<BM, BN, KW, true, false, DEPTH>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((dgemm_wsplit_kernel<BM, BN, KW, true, true, DEPTH>), grid, block, 0, s, g);
}

template <int BM, int BN, int KW, int DEPTH = 1>
static void launch_wsplit_group(bool ta, bool tb, GemmGroup& gg, const int* batch, hipStream_t s)
{
    group_grid(gg, batch, BM, BN);
    if (gg.start[gg.count] == 0) return;
    dim3 grid(gg.start[gg.count]);
    dim3 block(64 * KW);
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_wsplit_group<BM, BN, KW, false, false, DEPTH>), grid, block, 0, s, gg);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_wsplit_group<BM, BN, KW, false, true, DEPTH>), grid, block, 0, s, gg);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_wsplit_group<BM, BN, KW, true, false, DEPTH>), grid, block, 0, s, gg);
    else hipLaunchKernelGGL((dgemm_wsplit_group<BM, BN, KW, true, true, DEPTH>), grid, block, 0, s, gg);
}

// ---------------------------------------------------------------------------
// Parity-folded GEMMs (LDS-tiled, BK = 16, single LDS buffer).
//
// The collocation operators of the spectral step are mirror-symmetric: for a
// backward operator G (grid point j, coefficient k) G[L-1-j, k] =
// s (-1)^k G[j, k], and for a forward operator F, F[k, L-1-j] = s (-1)^k F[k, j].
// Each half of the work is then a GEMM of half the size:
//   MODE 1 / 2 (output fold, rows / columns of C): with E = A0 B0 (even k) and
//     O = A1 B1 (odd k) over the first half of each length-Lb block of output
//     rows (columns), C[j] = E + O and C[Lb-1-j] = sign_b (E - O).  One
//     workgroup runs both products (interleaved slab by slab) into two sets of
//     accumulators and writes both mirrored outputs.
//   MODE 3 / 4 (input fold of A's / B's k): the operand element k of a GEMM
//     over the first half of a length-Kf range is  X[k] + s_e X[Kf-1-k]; the
//     sign s_e alternates with the output parity, which the caller maps to
//     batch entries (even / odd output rows or columns, written with row or
//     column stride 2).
// Both halve the MFMA work of a transform.
struct FoldArgs {
    const double* A1;  // second product (modes 1, 2)
    const double* B1;
    long long lda1, ldb1;
    int k1;
    long long ska, skb;   // k-stride of A (not transposed) and of B (transposed)
    int fold_len;         // modes 1, 2: block length Lb; modes 3, 4: Kf
    unsigned fold_sign;   // modes 1, 2: bit b set: sign_b = -1
    double mirror_sign;   // modes 3, 4: s_e = mirror_sign * (-1)^((e / mirror_div) % 2)
    int mirror_div;
};

// Folded GEMMs of one mode and tile form for many members in one launch (the
// folded x-pass of every size class of a heterogeneous batch): the plain multi
// table layout plus each member's fold arguments.
struct FoldMulti {
    int count, total;
    int start[kMaxMulti + 1];
    int tx[kMaxMulti], ty[kMaxMulti];
    GemmArgs a[kMaxMulti];
    FoldArgs f[kMaxMulti];
};

// MULTI = false: one folded GEMM, its tile from the launch grid (mm unused).
// MULTI = true: the member and tile of this block from the table mm (g and f
// unused).  One kernel body for both, so that the single form keeps its own
// register allocation.
template <int BM, int BN, int WGM, int WGN, int MODE, bool TA, bool TB, bool MULTI>
__global__ void __launch_bounds__(64 * WGM * WGN) dgemm_fold_kernel(GemmArgs g0, FoldArgs f0,
                                                                   const FoldMulti* __restrict__ mm)
{
    constexpr int BK = 16;
    constexpr int NT = 64 * WGM * WGN;
    constexpr int PA = BM + 1;
    constexpr int PB = BN + 1;
    constexpr int WM = BM / WGM;
    constexpr int WN = BN / WGN;
    constexpr int TM = WM / 16;
    constexpr int TN = WN / 16;
    constexpr int LA = BM * BK / NT;
    constexpr int LB = BN * BK / NT;
    constexpr bool DUAL = (MODE == 1 || MODE == 2);
    constexpr int NACC = DUAL ? 2 : 1;
    static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");
    static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "tile / thread mismatch");
    static_assert(NT % BK == 0 && NT % BM == 0 && NT % BN == 0, "load mapping");
    __shared__ double As[BK * PA];
    __shared__ double Bs[BK * PB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WGN, wn = wave % WGN;
    int bx, by, bz;
    GemmArgs g;
    FoldArgs f;
    if constexpr (MULTI) {
        const int t = xcd_remap(blockIdx.x, mm->total);
        int lo = 0, hi = mm->count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (mm->start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int local = t - mm->start[lo];
        bx = local % mm->tx[lo];
        const int r = local / mm->tx[lo];
        by = r % mm->ty[lo];
        bz = r / mm->ty[lo];
        g = mm->a[lo];
        f = mm->f[lo];
    } else {
        grid_tile(bx, by, bz);
        g = g0;
        f = f0;
    }
    const int m0 = by * BM;
    const int n0 = bx * BN;
    const long long oA = bofs(bz, g.iA), oB = bofs(bz, g.iB);
    const double se = (MODE >= 3) ? (((bz / f.mirror_div) & 1) ? -f.mirror_sign : f.mirror_sign) : 0.0;
    const int kf1 = f.fold_len - 1;

    v4d acc[NACC][TM][TN];
#pragma unroll
    for (int a = 0; a < NACC; ++a)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[a][i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[LA], rb[LB];
    // modes 3 / 4: the mirrored element of each A / B element and whether the
    // element lies inside the operand
    double ra2[LA] = {}, rb2[LB] = {};
    bool ina[LA] = {}, inb[LB] = {};
    auto a_rc = [&](int q, int& r, int& c) {
        if (!TA) { c = tid % BK; r = tid / BK + q * (NT / BK); }
        else     { r = tid % BM; c = tid / BM + q * (NT / BM); }
    };
    auto b_rc = [&](int q, int& r, int& c) {
        if (!TB) { c = tid % BN; r = tid / BN + q * (NT / BN); }
        else     { r = tid % BK; c = tid / BK + q * (NT / BK); }
    };
    // slab v: product p = v & 1 (dual modes; else 0), k0 = (v >> DUAL) * BK
    auto load_slab = [&](int v) {
        const int p = DUAL ? (v & 1) : 0;
        const int k0 = (DUAL ? (v >> 1) : v) * BK;
        const double* __restrict__ A = (p ? f.A1 : g.A) + oA;
        const double* __restrict__ B = (p ? f.B1 : g.B) + oB;
        const long long lda = p ? f.lda1 : g.lda, ldb = p ? f.ldb1 : g.ldb;
        const int K = p ? f.k1 : g.k;
        // The mirrored operand (modes 3, 4) needs two loads per element: both
        // are issued unconditionally at clamped, valid indices, and the pair is
        // summed (and zeroed outside the operand) only when the slab is stored
        // to LDS, so that no use of a load in this phase makes the wave wait
        // for it while the current slab's MFMAs could run.
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            const int gm = m0 + r, gk = k0 + c;
            double v = 0.0;
            if (MODE == 3) {
                ina[q] = (gm < g.m) & (gk < K);
                const int gmc = min(gm, g.m - 1), gkc = min(gk, K - 1), mk = max(kf1 - gkc, 0);
                v = TA ? A[(long long)gkc * lda + gmc] : A[(long long)gmc * lda + (long long)gkc * f.ska];
                ra2[q] = TA ? A[(long long)mk * lda + gmc] : A[(long long)gmc * lda + (long long)mk * f.ska];
            } else if (gm < g.m && gk < K) {
                v = TA ? A[(long long)gk * lda + gm] : A[(long long)gm * lda + (long long)gk * f.ska];
            }
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            const int gk = k0 + r, gn = n0 + c;
            double v = 0.0;
            if (MODE == 4) {
                inb[q] = (gk < K) & (gn < g.n);
                const int gnc = min(gn, g.n - 1), gkc = min(gk, K - 1), mk = max(kf1 - gkc, 0);
                v = TB ? B[(long long)gnc * ldb + (long long)gkc * f.skb] : B[(long long)gkc * ldb + gnc];
                rb2[q] = TB ? B[(long long)gnc * ldb + (long long)mk * f.skb] : B[(long long)mk * ldb + gnc];
            } else if (gk < K && gn < g.n) {
                v = TB ? B[(long long)gn * ldb + (long long)gk * f.skb] : B[(long long)gk * ldb + gn];
            }
            rb[q] = v;
        }
    };
    auto store_slab = [&]() {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            const double sum = ra[q] + se * ra2[q];
            As[c * PA + r] = MODE == 3 ? (ina[q] ? sum : 0.0) : ra[q];
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            const double sum = rb[q] + se * rb2[q];
            Bs[r * PB + c] = MODE == 4 ? (inb[q] ? sum : 0.0) : rb[q];
        }
    };
    auto compute = [&](int a) {
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
                    acc[a][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[a][i][j], 0, 0, 0);
        }
    };

    const int kmax = DUAL ? max(g.k, f.k1) : g.k;
    const int nslab = (kmax + BK - 1) / BK;
    const int nv = DUAL ? 2 * nslab : nslab;
    if (nv > 0) load_slab(0);
    for (int v = 0; v < nv; ++v) {
        __syncthreads();
        store_slab();
        __syncthreads();
        if (v + 1 < nv) load_slab(v + 1);
        if (DUAL && (v & 1)) compute(NACC - 1);
        else compute(0);
    }

    const Epilogue ep(g, bz);
    // output elements: one per accumulator item, or (DUAL) the mirrored pair
    // x1 = even + odd and x2 = sign (even - odd), x2 only when it differs
    epilogue_store<TM * TN * 4 * (DUAL ? 2 : 1)>(ep, [&](int e, int& gm, int& gn, bool& ok, double& v) {
        const int q = DUAL ? e >> 1 : e;
        const int i = q / (TN * 4), j = (q / 4) % TN, r = q % 4;
        gm = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
        gn = n0 + wn * WN + j * 16 + (lane & 15);
        ok = gm < g.m && gn < g.n;
        if (!DUAL) {
            v = acc[0][i][j][r];
            return;
        }
        const double ev = acc[0][i][j][r], od = acc[NACC - 1][i][j][r];
        const int Lb = f.fold_len, Lh = (Lb + 1) / 2;
        const int x = ok ? ((MODE == 1) ? gm : gn) : 0;  // elements outside C are not stored
        const int b = x / Lh, jj = x - b * Lh;
        const int x1 = b * Lb + jj, x2 = b * Lb + Lb - 1 - jj;
        const double sg = ((f.fold_sign >> b) & 1u) ? -1.0 : 1.0;
        const int xo = (e & 1) ? x2 : x1;
        if (MODE == 1) gm = xo;
        else gn = xo;
        ok = ok && (!(e & 1) || x2 != x1);
        v = (e & 1) ? sg * (ev - od) : ev + od;
    });
}


// Output folds (modes 1, 2) split over waves: waves [0, KW) run the first
// product (even k) and waves [KW, 2 KW) the second (odd k), each wave a
// contiguous K range of its product staged through its own LDS slabs (no
// block barriers in the K loop), so a wave's serial chain is a 1 / (2 KW) share
// of the single-block form's.  The partial tiles meet in LDS; E and O are each
// summed in slice order and written through the fold epilogue.
template <int BM, int BN, int KW, int MODE, bool TA, bool TB>
__global__ void __launch_bounds__(128 * KW) dgemm_fold_wsplit_kernel(GemmArgs g, FoldArgs f)
{
    static_assert(MODE == 1 || MODE == 2, "output folds only");
    constexpr int BK = 16;
    constexpr int PA = BM + 1, PB = BN + 1;
    constexpr int SLAB = BK * PA + BK * PB;
    constexpr int TM = BM / 16, TN = BN / 16;
    constexpr int LA = BM * BK / 64, LB = BN * BK / 64;
    constexpr int NW = 2 * KW;
    constexpr int E = BM * BN;
    constexpr int AREA = (NW * SLAB > NW * E) ? NW * SLAB : NW * E;
    static_assert(64 % BK == 0 && 64 % BM == 0 && 64 % BN == 0, "load mapping: BM, BN divide 64");
    __shared__ double lds[AREA];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int prod = wave / KW, slice = wave - prod * KW;
    int bx, by, bz;
    grid_tile(bx, by, bz);
    const int m0 = by * BM;
    const int n0 = bx * BN;
    const long long oA = bofs(bz, g.iA), oB = bofs(bz, g.iB);
    const double* __restrict__ A = (prod ? f.A1 : g.A) + oA;
    const double* __restrict__ B = (prod ? f.B1 : g.B) + oB;
    const long long lda = prod ? f.lda1 : g.lda, ldb = prod ? f.ldb1 : g.ldb;
    const int K = prod ? f.k1 : g.k;
    const int kc = ((K + KW - 1) / KW + BK - 1) / BK * BK;
    const int kb = slice * kc;
    const int ke = min(kb + kc, K);
    double* As = lds + wave * SLAB;
    double* Bs = As + BK * PA;

    v4d acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[LA], rb[LB];
    auto a_rc = [&](int q, int& r, int& c) {
        if (!TA) { c = lane % BK; r = lane / BK + q * (64 / BK); }
        else     { r = lane % BM; c = lane / BM + q * (64 / BM); }
    };
    auto b_rc = [&](int q, int& r, int& c) {
        if (!TB) { c = lane % BN; r = lane / BN + q * (64 / BN); }
        else     { r = lane % BK; c = lane / BK + q * (64 / BK); }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            const int gm = m0 + r, gk = k0 + c;
            ra[q] = (gm < g.m && gk < ke) ? (TA ? A[(long long)gk * lda + gm] : A[(long long)gm * lda + (long long)gk * f.ska]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            const int gk = k0 + r, gn = n0 + c;
            rb[q] = (gk < ke && gn < g.n) ? (TB ? B[(long long)gn * ldb + (long long)gk * f.skb] : B[(long long)gk * ldb + gn]) : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            int r, c;
            a_rc(q, r, c);
            As[c * PA + r] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            int r, c;
            b_rc(q, r, c);
            Bs[r * PB + c] = rb[q];
        }
    };

    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
        wave_lds_sync();  // previous slab fully read
        store();
        wave_lds_sync();
        if (k0 + BK < ke) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = As[kr * PA + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = Bs[kr * PB + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }

    __syncthreads();  // every wave is done with its slabs; the area becomes the reduction buffer
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                lds[wave * E + ((i * TN + j) * 4 + r) * 64 + lane] = acc[i][j][r];
    __syncthreads();

    const Epilogue ep(g, bz);
    const int Lb = f.fold_len, Lh = (Lb + 1) / 2;
    // tile item tid + q * 64 NW of the E items gives the mirrored pair of
    // output elements 2 q (x1) and 2 q + 1 (x2, only when it differs)
    epilogue_store<2 * ((E + 64 * NW - 1) / (64 * NW))>(ep, [&](int o, int& gm, int& gn, bool& ok, double& v) {
        const int q = o >> 1;
        const int e = tid + q * 64 * NW;
        const int ec = e < E ? e : 0;
        double ev = 0.0, od = 0.0;
#pragma unroll
        for (int w = 0; w < KW; ++w) {
            ev += lds[w * E + ec];
            od += lds[(KW + w) * E + ec];
        }
        const int l = ec & 63;
        const int r = (ec >> 6) & 3;
        const int t = ec >> 8;
        gm = m0 + (t / TN) * 16 + (l >> 4) + 4 * r;
        gn = n0 + (t % TN) * 16 + (l & 15);
        ok = e < E && gm < g.m && gn < g.n;
        const int x = ok ? ((MODE == 1) ? gm : gn) : 0;  // elements outside C are not stored
        const int b = x / Lh, jj = x - b * Lh;
        const int x1 = b * Lb + jj, x2 = b * Lb + Lb - 1 - jj;
        const double sg = ((f.fold_sign >> b) & 1u) ? -1.0 : 1.0;
        const int xo = (o & 1) ? x2 : x1;
        if (MODE == 1) gm = xo;
        else gn = xo;
        ok = ok && (!(o & 1) || x2 != x1);
        v = (o & 1) ? sg * (ev - od) : ev + od;
    });
}

template <int BM, int BN, int KW>
static void launch_fold_wsplit(int mode, bool ta, bool tb, const GemmArgs& g, const FoldArgs& f,
                               int batch, hipStream_t s)
{
    dim3 grid((g.n + BN - 1) / BN, (g.m + BM - 1) / BM, batch);
    dim3 block(128 * KW);
#define PYPDE_FW(M, TA_, TB_) hipLaunchKernelGGL((dgemm_fold_wsplit_kernel<BM, BN, KW, M, TA_, TB_>), grid, block, 0, s, g, f)
    if (mode == 1) {
        if (!ta && !tb) PYPDE_FW(1, false, false);
        else if (!ta && tb) PYPDE_FW(1, false, true);
        else if (ta && !tb) PYPDE_FW(1, true, false);
        else PYPDE_FW(1, true, true);
    } else {
        if (!ta && !tb) PYPDE_FW(2, false, false);
        else if (!ta && tb) PYPDE_FW(2, false, true);
        else if (ta && !tb) PYPDE_FW(2, true, false);
        else PYPDE_FW(2, true, true);
    }
#undef PYPDE_FW
}

template <int BM, int BN, int WGM, int WGN, int MODE>
static void launch_fold_m(bool ta, bool tb, const GemmArgs& g, const FoldArgs& f, int batch, hipStream_t s)
{
    dim3 grid((g.n + BN - 1) / BN, (g.m + BM - 1) / BM, batch);
    dim3 block(64 * WGM * WGN);
    const FoldMulti* none = nullptr;
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, false, false, false>), grid, block, 0, s, g, f, none);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, false, true, false>), grid, block, 0, s, g, f, none);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, true, false, false>), grid, block, 0, s, g, f, none);
    else hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, true, true, false>), grid, block, 0, s, g, f, none);
}

template <int BM, int BN, int WGM, int WGN>
static void launch_fold(int mode, bool ta, bool tb, const GemmArgs& g, const FoldArgs& f, int batch,
                        hipStream_t s)
{
    switch (mode) {
    case 1: launch_fold_m<BM, BN, WGM, WGN, 1>(ta, tb, g, f, batch, s); break;
    case 2: launch_fold_m<BM, BN, WGM, WGN, 2>(ta, tb, g, f, batch, s); break;
    case 3: launch_fold_m<BM, BN, WGM, WGN, 3>(ta, tb, g, f, batch, s); break;
    default: launch_fold_m<BM, BN, WGM, WGN, 4>(ta, tb, g, f, batch, s); break;
    }
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

// Variant codes (pypde_gemm_desc.variant):
//   0 heuristic, 1 LDS 64x64 (4 waves), 2 LDS 32x32 (4 waves),
//   3 LDS 128x128 (8 waves), 4 LDS 128x128 (8 waves, double-buffered LDS),
//   5 LDS 128x64 (8 waves), 6 LDS 64x128 (8 waves),
//   7, 24, 25, 26, 27 LDS 128x64, 64x128, 64x64, 128x128, 32x32 with two slabs
//   of loads in flight, 28 LDS 32x32 double-buffered, 29 and 30 LDS 32x32 over
//   2 waves (16x32 each) and 1 wave (four 16x16 accumulators), 31 and 32 those
//   two with two slabs of loads in flight, 33 and 34 variants 2 and 27 with
//   32-deep K slabs instead of 16, 35 and 36 the same with 64-deep slabs,
//   37 and 38 variants 1 and 5 with 32-deep slabs, 39..42 variants 2, 27,
//   33, 34 with LDS rows padded by two doubles instead of one (consecutive k
//   of one tile row then lie two 8-byte banks apart and the next row fills
//   the banks between, so a slab store by 32 lanes, 16 k x 2 rows, touches 32
//   distinct banks),
//   8 split-K LDS 32x32 over 4 waves, 9 the same over 8 waves, 10 64x32 over
//   4 waves, 11 32x64 over 4 waves, 12..15 those four with two slabs in flight,
//   16 + c split-K with TM = 1 + (c & 1), TN = 1 + ((c >> 1) & 1), KW = (c & 4) ? 8 : 4.
static int heuristic_variant(const GemmArgs& g, int batch)
{
    const long long tiles128 = (long long)((g.m + 127) / 128) * ((g.n + 127) / 128) * batch;
    if (tiles128 >= 512) return 3;
    const long long tiles64 = (long long)((g.m + 63) / 64) * ((g.n + 63) / 64) * batch;
    if (tiles64 >= 512) return 1;
    const long long tiles32 = (long long)((g.m + 31) / 32) * ((g.n + 31) / 32) * batch;
    if (tiles32 >= 1024) return 2;
    // latency-bound: 32x32 split-K tiles when they still give >= 256 workgroups
    const int big = (tiles32 >= 256) ? 3 : 0;
    return 16 + big + (g.k > 512 ? 4 : 0);
}

static int launch(int transA, int transB, GemmArgs& g, int batch, int variant, hipStream_t stream)
{
    PYPDE_REQUIRE(g.m >= 0 && g.n >= 0 && g.k >= 0 && batch >= 0, "negative GEMM size");
    if (g.m == 0 || g.n == 0 || batch == 0) return OK;
    PYPDE_REQUIRE(g.A && g.B && g.C, "null GEMM operand");
    PYPDE_REQUIRE(batch <= 65535, "GEMM batch %d exceeds 65535", batch);
    PYPDE_REQUIRE((long long)((g.m + 15) / 16) * ((g.n + 15) / 16) * batch < (1LL << 31),
                  "GEMM grid too large");
    PYPDE_REQUIRE(variant >= 0 && variant <= 42, "unknown GEMM variant %d", variant);
    if (variant == 0) variant = heuristic_variant(g, batch);
    switch (variant) {
    case 1: launch_gemm<64, 64, 2, 2, false>(transA, transB, g, batch, stream); break;
    case 2: launch_gemm<32, 32, 2, 2, false>(transA, transB, g, batch, stream); break;
    case 3: launch_gemm<128, 128, 4, 2, false>(transA, transB, g, batch, stream); break;
    case 4: launch_gemm<128, 128, 4, 2, true>(transA, transB, g, batch, stream); break;
    case 5: launch_gemm<128, 64, 4, 2, false>(transA, transB, g, batch, stream); break;
    case 6: launch_gemm<64, 128, 2, 4, false>(transA, transB, g, batch, stream); break;
    case 7: launch_gemm<128, 64, 4, 2, 2>(transA, transB, g, batch, stream); break;
    case 24: launch_gemm<64, 128, 2, 4, 2>(transA, transB, g, batch, stream); break;
    case 25: launch_gemm<64, 64, 2, 2, 2>(transA, transB, g, batch, stream); break;
    case 26: launch_gemm<128, 128, 4, 2, 2>(transA, transB, g, batch, stream); break;
    case 27: launch_gemm<32, 32, 2, 2, 2>(transA, transB, g, batch, stream); break;
    case 28: launch_gemm<32, 32, 2, 2, 1>(transA, transB, g, batch, stream); break;
    case 29: launch_gemm<32, 32, 2, 1, 0>(transA, transB, g, batch, stream); break;
    case 30: launch_gemm<32, 32, 1, 1, 0>(transA, transB, g, batch, stream); break;
    case 31: launch_gemm<32, 32, 2, 1, 2>(transA, transB, g, batch, stream); break;
    case 32: launch_gemm<32, 32, 1, 1, 2>(transA, transB, g, batch, stream); break;
    case 33: launch_gemm<32, 32, 2, 2, 0, 32>(transA, transB, g, batch, stream); break;
    case 34: launch_gemm<32, 32, 2, 2, 2, 32>(transA, transB, g, batch, stream); break;
    case 35: launch_gemm<32, 32, 2, 2, 0, 64>(transA, transB, g, batch, stream); break;
    case 36: launch_gemm<32, 32, 2, 2, 2, 64>(transA, transB, g, batch, stream); break;
    case 37: launch_gemm<64, 64, 2, 2, 0, 32>(transA, transB, g, batch, stream); break;
    case 38: launch_gemm<128, 64, 4, 2, 0, 32>(transA, transB, g, batch, stream); break;
    case 39: launch_gemm<32, 32, 2, 2, 0, 16, 2>(transA, transB, g, batch, stream); break;
    case 40: launch_gemm<32, 32, 2, 2, 2, 16, 2>(transA, transB, g, batch, stream); break;
    case 41: launch_gemm<32, 32, 2, 2, 0, 32, 2>(transA, transB, g, batch, stream); break;
    case 42: launch_gemm<32, 32, 2, 2, 2, 32, 2>(transA, transB, g, batch, stream); break;
    case 8: launch_wsplit<32, 32, 4>(transA, transB, g, batch, stream); break;
    case 9: launch_wsplit<32, 32, 8>(transA, transB, g, batch, stream); break;
    case 10: launch_wsplit<64, 32, 4>(transA, transB, g, batch, stream); break;
    case 11: launch_wsplit<32, 64, 4>(transA, transB, g, batch, stream); break;
    case 12: launch_wsplit<32, 32, 4, 2>(transA, transB, g, batch, stream); break;
    case 13: launch_wsplit<32, 32, 8, 2>(transA, transB, g, batch, stream); break;
    case 14: launch_wsplit<64, 32, 4, 2>(transA, transB, g, batch, stream); break;
    case 15: launch_wsplit<32, 64, 4, 2>(transA, transB, g, batch, stream); break;
    case 16: launch_splitk<1, 1, 4>(transA, transB, g, batch, stream); break;
    case 17: launch_splitk<2, 1, 4>(transA, transB, g, batch, stream); break;
    case 18: launch_splitk<1, 2, 4>(transA, transB, g, batch, stream); break;
    case 19: launch_splitk<2, 2, 4>(transA, transB, g, batch, stream); break;
    case 20: launch_splitk<1, 1, 8>(transA, transB, g, batch, stream); break;
    case 21: launch_splitk<2, 1, 8>(transA, transB, g, batch, stream); break;
    case 22: launch_splitk<1, 2, 8>(transA, transB, g, batch, stream); break;
    default: launch_splitk<2, 2, 8>(transA, transB, g, batch, stream); break;
    }
    return check_launch("pypde_dgemm");
}

int dgemm(int transA, int transB, int m, int n, int k, double alpha, const double* A,
          long long lda, long long strideA, const double* B, long long ldb, long long strideB,
          double beta, const double* D, long long ldd, long long strideD, double* C,
          long long ldc, long long strideC, int batch, hipStream_t stream)
{
    GemmArgs g{A, B, D, C, lda, ldb, ldd, ldc, strided(strideA), strided(strideB),
               strided(strideD), strided(strideC), m, n, k, alpha, beta, nullptr, 1,
               nullptr, 0, strided(0), nullptr, 0.0, 1, 0, 1, 1};
    return launch(transA, transB, g, batch, 0, stream);
}

int dgemm_batched(const pypde_gemm_desc* d, hipStream_t stream)
{
    PYPDE_REQUIRE(d, "null GEMM descriptor");
    PYPDE_REQUIRE(!d->alpha_vec || d->alpha_div >= 1, "alpha_div must be >= 1");
    PYPDE_REQUIRE(!d->gamma_vec || d->gamma_div >= 1, "gamma_div must be >= 1");
    GemmArgs g{d->A, d->B, d->D, d->C, d->lda, d->ldb, d->ldd, d->ldc, bidx(d->iA), bidx(d->iB),
               bidx(d->iD), bidx(d->iC), d->m, d->n, d->k, d->alpha, d->beta, d->alpha_vec,
               d->alpha_div < 1 ? 1 : d->alpha_div, d->E, d->lde, bidx(d->iE), d->gamma_vec,
               d->gamma, d->gamma_div < 1 ? 1 : d->gamma_div, d->zero_first,
               d->csn < 1 ? 1 : d->csn, d->ska < 1 ? 1 : d->ska};
    if (d->fold == 0) return launch(d->transA, d->transB, g, d->batch, d->variant, stream);

    // parity-folded forms
    PYPDE_REQUIRE(d->fold >= 1 && d->fold <= 4, "unknown fold mode %d", d->fold);
    PYPDE_REQUIRE(g.m >= 0 && g.n >= 0 && g.k >= 0 && d->batch >= 0, "negative GEMM size");
    if (g.m == 0 || g.n == 0 || d->batch == 0) return OK;
    PYPDE_REQUIRE(g.A && g.B && g.C, "null GEMM operand");
    PYPDE_REQUIRE(d->batch <= 65535, "GEMM batch %d exceeds 65535", d->batch);
    PYPDE_REQUIRE(d->fold_len >= 1, "fold_len must be >= 1");
    PYPDE_REQUIRE((long long)((g.m + 15) / 16) * ((g.n + 15) / 16) * d->batch < (1LL << 31),
                  "GEMM grid too large");
    FoldArgs f{d->A1, d->B1, d->lda1, d->ldb1, d->k1, d->ska < 1 ? 1 : d->ska,
               d->skb < 1 ? 1 : d->skb, d->fold_len, d->fold_sign, d->mirror_sign,
               d->mirror_div < 1 ? 1 : d->mirror_div};
    if (d->fold <= 2) {
        PYPDE_REQUIRE(d->A1 && d->B1 && d->k1 >= 0, "output fold needs the second product (A1, B1, k1)");
        const long long half = (d->fold_len + 1) / 2;
        const long long dim = d->fold == 1 ? g.m : g.n;
        PYPDE_REQUIRE(dim % half == 0, "output fold: the folded extent %lld is not a multiple of %lld",
                      dim, half);
    } else {
        PYPDE_REQUIRE(d->fold_len % 2 == 0 && 2 * (long long)g.k == d->fold_len,
                      "input fold: k (%d) must be half of an even fold_len (%d)", g.k, d->fold_len);
    }
    int variant = d->variant;
    if (variant == 0) {
        const long long tiles128 = (long long)((g.m + 127) / 128) * ((g.n + 127) / 128) * d->batch;
        const long long tiles64 = (long long)((g.m + 63) / 64) * ((g.n + 63) / 64) * d->batch;
        variant = tiles128 >= 512 ? 3 : (tiles64 >= 512 ? 1 : 2);
    }
    PYPDE_REQUIRE(variant >= 1 && variant <= 7, "folded GEMMs support variants 1-7 (got %d)", variant);
    PYPDE_REQUIRE(variant <= 5 || d->fold <= 2, "fold variants 6, 7 are for the output folds (1, 2)");
    const bool ta = d->transA, tb = d->transB;
    // smaller wave tiles than the plain kernels: the output folds carry two
    // accumulator sets (32x32 per wave at 128x128 over 16 waves)
    switch (variant) {
    case 1: launch_fold<64, 64, 4, 2>(d->fold, ta, tb, g, f, d->batch, stream); break;
    case 2: launch_fold<32, 32, 2, 2>(d->fold, ta, tb, g, f, d->batch, stream); break;
    case 3: launch_fold<128, 128, 4, 4>(d->fold, ta, tb, g, f, d->batch, stream); break;
    case 4: launch_fold<128, 64, 4, 2>(d->fold, ta, tb, g, f, d->batch, stream); break;
    case 5: launch_fold<64, 128, 2, 4>(d->fold, ta, tb, g, f, d->batch, stream); break;
    // each product over its own waves, K split KW ways (32 x 32 tiles)
    case 6: launch_fold_wsplit<32, 32, 2>(d->fold, ta, tb, g, f, d->batch, stream); break;
    default: launch_fold_wsplit<32, 32, 4>(d->fold, ta, tb, g, f, d->batch, stream); break;
    }
    return check_launch("pypde_dgemm_batched (folded)");
}

static GemmArgs args_of(const pypde_gemm_desc* d)
{
    GemmArgs g{d->A, d->B, d->D, d->C, d->lda, d->ldb, d->ldd, d->ldc, bidx(d->iA), bidx(d->iB),
               bidx(d->iD), bidx(d->iC), d->m, d->n, d->k, d->alpha, d->beta, d->alpha_vec,
               d->alpha_div < 1 ? 1 : d->alpha_div, d->E, d->lde, bidx(d->iE), d->gamma_vec,
               d->gamma, d->gamma_div < 1 ? 1 : d->gamma_div, d->zero_first,
               d->csn < 1 ? 1 : d->csn, d->ska < 1 ? 1 : d->ska};
    return g;
}

int dgemm_grouped(const pypde_gemm_desc* d, int count, hipStream_t stream)
{
    PYPDE_REQUIRE(d, "null GEMM descriptor array");
    PYPDE_REQUIRE(count >= 1 && count <= kMaxGroup, "grouped GEMM: count %d not in 1..%d", count, kMaxGroup);
    GemmGroup gg;
    int batch[kMaxGroup];
    gg.count = count;
    const int ta = d[0].transA, tb = d[0].transB;
    int variant = d[0].variant;
    long long tiles = 0;
    for (int i = 0; i < count; ++i) {
        const pypde_gemm_desc& e = d[i];
        PYPDE_REQUIRE(e.fold == 0, "grouped GEMM: member %d is folded (not supported)", i);
        PYPDE_REQUIRE(e.transA == ta && e.transB == tb, "grouped GEMM: member %d has other transposes", i);
        PYPDE_REQUIRE(e.variant == variant, "grouped GEMM: member %d has another variant", i);
        PYPDE_REQUIRE(e.m >= 0 && e.n >= 0 && e.k >= 0 && e.batch >= 0, "negative GEMM size (member %d)", i);
        PYPDE_REQUIRE(!e.alpha_vec || e.alpha_div >= 1, "alpha_div must be >= 1");
        PYPDE_REQUIRE(!e.gamma_vec || e.gamma_div >= 1, "gamma_div must be >= 1");
        if (e.m > 0 && e.n > 0 && e.batch > 0)
            PYPDE_REQUIRE(e.A && e.B && e.C, "null GEMM operand (member %d)", i);
        gg.a[i] = args_of(&e);
        batch[i] = e.batch;
        tiles += (long long)((e.m + 15) / 16) * ((e.n + 15) / 16) * e.batch;
    }
    PYPDE_REQUIRE(tiles < (1LL << 31), "grouped GEMM too large");
    PYPDE_REQUIRE(variant >= 0 && variant <= 42, "unknown GEMM variant %d", variant);
    if (variant == 0) variant = heuristic_variant(gg.a[0], batch[0]);
    switch (variant) {
    case 1: launch_gemm_group<64, 64, 2, 2, false>(ta, tb, gg, batch, stream); break;
    case 2: launch_gemm_group<32, 32, 2, 2, false>(ta, tb, gg, batch, stream); break;
    case 3: launch_gemm_group<128, 128, 4, 2, false>(ta, tb, gg, batch, stream); break;
    case 4: launch_gemm_group<128, 128, 4, 2, true>(ta, tb, gg, batch, stream); break;
    case 5: launch_gemm_group<128, 64, 4, 2, false>(ta, tb, gg, batch, stream); break;
    case 6: launch_gemm_group<64, 128, 2, 4, false>(ta, tb, gg, batch, stream); break;
    case 7: launch_gemm_group<128, 64, 4, 2, 2>(ta, tb, gg, batch, stream); break;
    case 24: launch_gemm_group<64, 128, 2, 4, 2>(ta, tb, gg, batch, stream); break;
    case 25: launch_gemm_group<64, 64, 2, 2, 2>(ta, tb, gg, batch, stream); break;
    case 26: launch_gemm_group<128, 128, 4, 2, 2>(ta, tb, gg, batch, stream); break;
    case 27: launch_gemm_group<32, 32, 2, 2, 2>(ta, tb, gg, batch, stream); break;
    case 28: launch_gemm_group<32, 32, 2, 2, 1>(ta, tb, gg, batch, stream); break;
    case 29: launch_gemm_group<32, 32, 2, 1, 0>(ta, tb, gg, batch, stream); break;
    case 30: launch_gemm_group<32, 32, 1, 1, 0>(ta, tb, gg, batch, stream); break;
    case 31: launch_gemm_group<32, 32, 2, 1, 2>(ta, tb, gg, batch, stream); break;
    case 32: launch_gemm_group<32, 32, 1, 1, 2>(ta, tb, gg, batch, stream); break;
    case 33: launch_gemm_group<32, 32, 2, 2, 0, 32>(ta, tb, gg, batch, stream); break;
    case 34: launch_gemm_group<32, 32, 2, 2, 2, 32>(ta, tb, gg, batch, stream); break;
    case 35: launch_gemm_group<32, 32, 2, 2, 0, 64>(ta, tb, gg, batch, stream); break;
    case 36: launch_gemm_group<32, 32, 2, 2, 2, 64>(ta, tb, gg, batch, stream); break;
    case 37: launch_gemm_group<64, 64, 2, 2, 0, 32>(ta, tb, gg, batch, stream); break;
    case 38: launch_gemm_group<128, 64, 4, 2, 0, 32>(ta, tb, gg, batch, stream); break;
    case 39: launch_gemm_group<32, 32, 2, 2, 0, 16, 2>(ta, tb, gg, batch, stream); break;
    case 40: launch_gemm_group<32, 32, 2, 2, 2, 16, 2>(ta, tb, gg, batch, stream); break;
    case 41: launch_gemm_group<32, 32, 2, 2, 0, 32, 2>(ta, tb, gg, batch, stream); break;
    case 42: launch_gemm_group<32, 32, 2, 2, 2, 32, 2>(ta, tb, gg, batch, stream); break;
    case 8: launch_wsplit_group<32, 32, 4>(ta, tb, gg, batch, stream); break;
    case 9: launch_wsplit_group<32, 32, 8>(ta, tb, gg, batch, stream); break;
    case 10: launch_wsplit_group<64, 32, 4>(ta, tb, gg, batch, stream); break;
    case 11: launch_wsplit_group<32, 64, 4>(ta, tb, gg, batch, stream); break;
    case 12: launch_wsplit_group<32, 32, 4, 2>(ta, tb, gg, batch, stream); break;
    case 13: launch_wsplit_group<32, 32, 8, 2>(ta, tb, gg, batch, stream); break;
    case 14: launch_wsplit_group<64, 32, 4, 2>(ta, tb, gg, batch, stream); break;
    case 15: launch_wsplit_group<32, 64, 4, 2>(ta, tb, gg, batch, stream); break;
    case 16: launch_splitk_group<1, 1, 4>(ta, tb, gg, batch, stream); break;
    case 17: launch_splitk_group<2, 1, 4>(ta, tb, gg, batch, stream); break;
    case 18: launch_splitk_group<1, 2, 4>(ta, tb, gg, batch, stream); break;
    case 19: launch_splitk_group<2, 2, 4>(ta, tb, gg, batch, stream); break;
    case 20: launch_splitk_group<1, 1, 8>(ta, tb, gg, batch, stream); break;
    case 21: launch_splitk_group<2, 1, 8>(ta, tb, gg, batch, stream); break;
    case 22: launch_splitk_group<1, 2, 8>(ta, tb, gg, batch, stream); break;
    default: launch_splitk_group<2, 2, 8>(ta, tb, gg, batch, stream); break;
    }
    return check_launch("pypde_dgemm_grouped");
}

long long gemm_multi_bytes()
{
    return (long long)(sizeof(GemmMulti) > sizeof(FoldMulti) ? sizeof(GemmMulti) : sizeof(FoldMulti));
}

// Block tile of the folded variants 1-5 (0 for others).
static void fold_tile_shape(int v, int& bm, int& bn)
{
    bm = bn = 0;
    switch (v) {
    case 1: bm = 64; bn = 64; break;
    case 2: bm = 32; bn = 32; break;
    case 3: bm = 128; bn = 128; break;
    case 4: bm = 128; bn = 64; break;
    case 5: bm = 64; bn = 128; break;
    default: break;
    }
}

static FoldArgs fold_args_of(const pypde_gemm_desc* d)
{
    FoldArgs f{d->A1, d->B1, d->lda1, d->ldb1, d->k1, d->ska < 1 ? 1 : d->ska,
               d->skb < 1 ? 1 : d->skb, d->fold_len, d->fold_sign, d->mirror_sign,
               d->mirror_div < 1 ? 1 : d->mirror_div};
    return f;
}

// Folded members: one fold mode and transposes for all, variants 1-5.
static int fold_multi_pack(const pypde_gemm_desc* d, int count, int variant, void* table,
                           pypde_multi_info* info)
{
    int bm, bn;
    fold_tile_shape(variant, bm, bn);
    PYPDE_REQUIRE(bm > 0, "multi GEMM: folded members run variants 1-5 (got %d)", variant);
    FoldMulti* mm = static_cast<FoldMulti*>(table);
    memset(mm, 0, sizeof(FoldMulti));
    mm->count = count;
    const int ta = d[0].transA, tb = d[0].transB, mode = d[0].fold;
    PYPDE_REQUIRE(mode >= 1 && mode <= 4, "multi GEMM: fold mode %d", mode);
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        const pypde_gemm_desc& e = d[i];
        PYPDE_REQUIRE(e.fold == mode, "multi GEMM: member %d has another fold mode", i);
        PYPDE_REQUIRE(e.transA == ta && e.transB == tb, "multi GEMM: member %d has other transposes", i);
        PYPDE_REQUIRE(e.m >= 0 && e.n >= 0 && e.k >= 0 && e.batch >= 0, "negative GEMM size (member %d)", i);
        PYPDE_REQUIRE(!e.alpha_vec || e.alpha_div >= 1, "alpha_div must be >= 1");
        PYPDE_REQUIRE(!e.gamma_vec || e.gamma_div >= 1, "gamma_div must be >= 1");
        PYPDE_REQUIRE(e.fold_len >= 1, "member %d: fold_len must be >= 1", i);
        const bool live = e.m > 0 && e.n > 0 && e.batch > 0;
        if (live) PYPDE_REQUIRE(e.A && e.B && e.C, "null GEMM operand (member %d)", i);
        if (mode <= 2) {
            PYPDE_REQUIRE(!live || (e.A1 && e.B1 && e.k1 >= 0),
                          "member %d: output fold needs the second product (A1, B1, k1)", i);
            const long long half = (e.fold_len + 1) / 2;
            PYPDE_REQUIRE((mode == 1 ? e.m : e.n) % half == 0,
                          "member %d: the folded extent is not a multiple of %lld", i, half);
        } else {
            PYPDE_REQUIRE(e.fold_len % 2 == 0 && 2LL * e.k == e.fold_len,
                          "member %d: k must be half of an even fold_len", i);
        }
        mm->a[i] = args_of(&e);
        mm->f[i] = fold_args_of(&e);
        mm->tx[i] = (e.n + bn - 1) / bn;
        mm->ty[i] = (e.m + bm - 1) / bm;
        mm->start[i] = (int)total;
        if (live) total += (long long)mm->tx[i] * mm->ty[i] * e.batch;
        PYPDE_REQUIRE(total < (1LL << 31), "multi GEMM too large");
    }
    mm->start[count] = (int)total;
    mm->total = (int)total;
    info->kind = PYPDE_MULTI_GEMM;
    info->count = count;
    info->variant = variant;
    info->flags = (ta ? 1 : 0) | (tb ? 2 : 0) | (mode << 2);
    info->total = total;
    return OK;
}

template <int BM, int BN, int WGM, int WGN, int MODE>
static void launch_fold_multi_m(bool ta, bool tb, const FoldMulti* mm, int total, hipStream_t s)
{
    dim3 grid(total), block(64 * WGM * WGN);
    const GemmArgs g{};
    const FoldArgs f{};
    if (!ta && !tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, false, false, true>), grid, block, 0, s, g, f, mm);
    else if (!ta && tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, false, true, true>), grid, block, 0, s, g, f, mm);
    else if (ta && !tb) hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, true, false, true>), grid, block, 0, s, g, f, mm);
    else hipLaunchKernelGGL((dgemm_fold_kernel<BM, BN, WGM, WGN, MODE, true, true, true>), grid, block, 0, s, g, f, mm);
}

template <int BM, int BN, int WGM, int WGN>
static void launch_fold_multi(int mode, bool ta, bool tb, const FoldMulti* mm, int total, hipStream_t s)
{
    switch (mode) {
    case 1: launch_fold_multi_m<BM, BN, WGM, WGN, 1>(ta, tb, mm, total, s); break;
    case 2: launch_fold_multi_m<BM, BN, WGM, WGN, 2>(ta, tb, mm, total, s); break;
    case 3: launch_fold_multi_m<BM, BN, WGM, WGN, 3>(ta, tb, mm, total, s); break;
    default: launch_fold_multi_m<BM, BN, WGM, WGN, 4>(ta, tb, mm, total, s); break;
    }
}

int gemm_multi_pack(const pypde_gemm_desc* d, int count, int variant, void* table,
                    long long table_bytes, pypde_multi_info* info)
{
    PYPDE_REQUIRE(d && table && info, "null argument");
    PYPDE_REQUIRE(table_bytes >= gemm_multi_bytes(), "GEMM table needs %lld bytes",
                  gemm_multi_bytes());
    PYPDE_REQUIRE(count >= 1 && count <= kMaxMulti, "multi GEMM: count %d not in 1..%d", count,
                  kMaxMulti);
    if (d[0].fold != 0) return fold_multi_pack(d, count, variant, table, info);
    int bm, bn;
    multi_tile_shape(variant, bm, bn);
    PYPDE_REQUIRE(bm > 0, "multi GEMM: variant %d is not an LDS-tiled variant", variant);
    GemmMulti* mm = static_cast<GemmMulti*>(table);
    memset(mm, 0, sizeof(GemmMulti));
    mm->count = count;
    const int ta = d[0].transA, tb = d[0].transB;
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        const pypde_gemm_desc& e = d[i];
        PYPDE_REQUIRE(e.fold == 0, "multi GEMM: member %d is folded (not supported)", i);
        PYPDE_REQUIRE(e.transA == ta && e.transB == tb, "multi GEMM: member %d has other transposes", i);
        PYPDE_REQUIRE(e.m >= 0 && e.n >= 0 && e.k >= 0 && e.batch >= 0, "negative GEMM size (member %d)", i);
        PYPDE_REQUIRE(!e.alpha_vec || e.alpha_div >= 1, "alpha_div must be >= 1");
        PYPDE_REQUIRE(!e.gamma_vec || e.gamma_div >= 1, "gamma_div must be >= 1");
        const bool live = e.m > 0 && e.n > 0 && e.batch > 0;
        if (live) PYPDE_REQUIRE(e.A && e.B && e.C, "null GEMM operand (member %d)", i);
        mm->a[i] = args_of(&e);
        mm->tx[i] = (e.n + bn - 1) / bn;
        mm->ty[i] = (e.m + bm - 1) / bm;
        mm->start[i] = (int)total;
        if (live) total += (long long)mm->tx[i] * mm->ty[i] * e.batch;
        PYPDE_REQUIRE(total < (1LL << 31), "multi GEMM too large");
    }
    mm->start[count] = (int)total;
    mm->total = (int)total;
    info->kind = PYPDE_MULTI_GEMM;
    info->count = count;
    info->variant = variant;
    info->flags = (ta ? 1 : 0) | (tb ? 2 : 0);
    info->total = total;
    return OK;
}

int dgemm_multi(const void* table, const pypde_multi_info* info, hipStream_t stream)
{
    PYPDE_REQUIRE(table && info, "null argument");
    PYPDE_REQUIRE(info->kind == PYPDE_MULTI_GEMM, "not a GEMM table");
    if (info->total == 0) return OK;
    const GemmMulti* mm = static_cast<const GemmMulti*>(table);
    const bool ta = info->flags & 1, tb = info->flags & 2;
    const int total = (int)info->total;
    const int mode = (info->flags >> 2) & 7;
    if (mode) {
        const FoldMulti* fm = static_cast<const FoldMulti*>(table);
        PYPDE_REQUIRE(mode <= 4, "multi GEMM: fold mode %d", mode);
        switch (info->variant) {
        case 1: launch_fold_multi<64, 64, 4, 2>(mode, ta, tb, fm, total, stream); break;
        case 2: launch_fold_multi<32, 32, 2, 2>(mode, ta, tb, fm, total, stream); break;
        case 3: launch_fold_multi<128, 128, 4, 4>(mode, ta, tb, fm, total, stream); break;
        case 4: launch_fold_multi<128, 64, 4, 2>(mode, ta, tb, fm, total, stream); break;
        case 5: launch_fold_multi<64, 128, 2, 4>(mode, ta, tb, fm, total, stream); break;
        default:
            set_error("multi GEMM: folded members run variants 1-5 (got %d)", info->variant);
            return ERR_ARG;
        }
        return check_launch("pypde_dgemm_multi (folded)");
    }
    switch (info->variant) {
    case 1: launch_gemm_multi<64, 64, 2, 2, false>(ta, tb, mm, total, stream); break;
    case 2: launch_gemm_multi<32, 32, 2, 2, false>(ta, tb, mm, total, stream); break;
    case 3: launch_gemm_multi<128, 128, 4, 2, false>(ta, tb, mm, total, stream); break;
    case 4: launch_gemm_multi<128, 128, 4, 2, true>(ta, tb, mm, total, stream); break;
    case 5: launch_gemm_multi<128, 64, 4, 2, false>(ta, tb, mm, total, stream); break;
    case 6: launch_gemm_multi<64, 128, 2, 4, false>(ta, tb, mm, total, stream); break;
    case 7: launch_gemm_multi<128, 64, 4, 2, 2>(ta, tb, mm, total, stream); break;
    case 24: launch_gemm_multi<64, 128, 2, 4, 2>(ta, tb, mm, total, stream); break;
    case 25: launch_gemm_multi<64, 64, 2, 2, 2>(ta, tb, mm, total, stream); break;
    case 26: launch_gemm_multi<128, 128, 4, 2, 2>(ta, tb, mm, total, stream); break;
    case 27: launch_gemm_multi<32, 32, 2, 2, 2>(ta, tb, mm, total, stream); break;
    case 28: launch_gemm_multi<32, 32, 2, 2, 1>(ta, tb, mm, total, stream); break;
    case 29: launch_gemm_multi<32, 32, 2, 1, 0>(ta, tb, mm, total, stream); break;
    case 30: launch_gemm_multi<32, 32, 1, 1, 0>(ta, tb, mm, total, stream); break;
    case 31: launch_gemm_multi<32, 32, 2, 1, 2>(ta, tb, mm, total, stream); break;
    case 32: launch_gemm_multi<32, 32, 1, 1, 2>(ta, tb, mm, total, stream); break;
    case 33: launch_gemm_multi<32, 32, 2, 2, 0, 32>(ta, tb, mm, total, stream); break;
    case 34: launch_gemm_multi<32, 32, 2, 2, 2, 32>(ta, tb, mm, total, stream); break;
    case 35: launch_gemm_multi<32, 32, 2, 2, 0, 64>(ta, tb, mm, total, stream); break;
    case 36: launch_gemm_multi<32, 32, 2, 2, 2, 64>(ta, tb, mm, total, stream); break;
    case 37: launch_gemm_multi<64, 64, 2, 2, 0, 32>(ta, tb, mm, total, stream); break;
    case 38: launch_gemm_multi<128, 64, 4, 2, 0, 32>(ta, tb, mm, total, stream); break;
    case 39: launch_gemm_multi<32, 32, 2, 2, 0, 16, 2>(ta, tb, mm, total, stream); break;
    case 40: launch_gemm_multi<32, 32, 2, 2, 2, 16, 2>(ta, tb, mm, total, stream); break;
    case 41: launch_gemm_multi<32, 32, 2, 2, 0, 32, 2>(ta, tb, mm, total, stream); break;
    case 42: launch_gemm_multi<32, 32, 2, 2, 2, 32, 2>(ta, tb, mm, total, stream); break;
    default:
        set_error("multi GEMM: variant %d is not an LDS-tiled variant", info->variant);
        return ERR_ARG;
    }
    return check_launch("pypde_dgemm_multi");
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

extern "C" int pypde_dgemm_grouped(const pypde_gemm_desc* descs, int count, void* stream)
{
    return pypde::dgemm_grouped(descs, count, pypde::as_stream(stream));
}

extern "C" int pypde_gemm_multi_pack(const pypde_gemm_desc* descs, int count, int variant,
                                     void* table, long long table_bytes, pypde_multi_info* info)
{
    return pypde::gemm_multi_pack(descs, count, variant, table, table_bytes, info);
}

extern "C" int pypde_dgemm_multi(const void* table, const pypde_multi_info* info, void* stream)
{
    return pypde::dgemm_multi(table, info, pypde::as_stream(stream));
}
