// Register-resident banded line solves along one axis (the hot recurrences of
// the time step: four-diagonal Helmholtz sweeps, eigen-split Poisson line
// solves, the Galerkin projection (S^T S)^-1 S^T, two-band solves).
//
// Every operator is written in one normalised form over positions i of a line
// (all offsets are even, so even and odd positions form independent chains):
//
//   pre-stencil (optional)  y_i = u_i + pre_i u_{i+2}
//   forward                 y_i = F_i y_i + LM_i y_{i-2}            (ascending)
//   backward                z_i = R_i y_i + MU_i z_{i+2} + NU_i z_{i+4}  (descending)
//
// with zero-padded tables that encode chain starts and ends, and reciprocals
// of the pivots precomputed at plan time (no division in the kernel).
//
// Block form: a block owns 64 lines (one lane each) and P = ceil(n / CH)
// wave64s (CH = 16 or 32, below); wave w holds positions [w*CH, (w+1)*CH) of its lane's
// line in registers.  Each sweep runs twice per chunk: once to obtain the chunk's
// affine transfer (end value and carry coefficient), then, after the carries
// are combined across the P chunks through LDS, once more from the exact
// incoming values.  Loads and stores are coalesced: along columns (axis 0)
// neighbouring lanes read neighbouring columns; along rows (axis 1) each wave
// transposes 64 x 16 tiles (128-byte row segments) through its own LDS tile,
// synchronising only within the wave.  The 32-line variant of the block form
// gives the two half-waves of a wave neighbouring chunks of 32 lines (32 x 32
// transpose tiles along rows), and the lane-chunk form further below gives
// each line one wave.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <cstring>

#include <algorithm>

namespace pypde {

// Positions per chunk (per lane): 16 halves the serial chain of each wave and
// the registers it holds, and is faster for the four-diagonal and per-line
// Poisson forms up to n = 256 (16 waves); the pre-stencil form in 64-line
// blocks and longer lines use 32.
constexpr int kMaxWaves = 16;
constexpr int TW = 16;    // row-tile width for axis-1 transposes
constexpr int TP = TW + 1;
constexpr int TP32 = 33;  // pitch of the 32 x 32 tiles of the 32-line form

struct RecArgs {
    const double* src;
    double* dst;
    long long src_ld, dst_ld, src_bs, dst_bs;
    int n_in, n, nlines, batch, axis;
    const double *F, *LM, *R, *MU, *NU, *pre;
    long long tbs;   // table stride between batch entries (shared mode)
    int per_line;    // tables laid out [n][nlines] per table
};

// Orders LDS traffic of one wavefront (a wave-local tile written by some lanes
// and read by others); waves of the block are not synchronised.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Table access: per-line tables from global memory ([n][nlines], coalesced
// across lanes); shared tables from the LDS copy of the block's batch entry,
// which stores only the present tables in the order F, LM, R, MU, NU, pre (the
// slot of each is fixed by the template flags).
template <bool PERLINE, bool HASF, bool FWD, bool HASR, bool HASNU>
struct Tab {
    const double* g;   // per-line: this lane's column of slot 1 (LM); shared: LDS base
    int nlines, line, n;
    __device__ __forceinline__ double operator()(int which, int pos) const
    {
        // per-line: the offset is uniform across the wave, the base per lane
        if (PERLINE) return g[(unsigned)(((which - 1) * n + pos) * nlines)];
        const int slot = (which > 0 && HASF) + (which > 1 && FWD) + (which > 2 && HASR) + (which > 3) +
                         (which > 4 && HASNU);
        return g[slot * n + pos];
    }
};

// Sweeps over one chunk held in registers.  GUARD: skip positions >= n (only
// the last chunk of a line needs it).
// P chunks per line, w = this lane's chunk, lane = its line within the block;
// the exchange area holds LS lines per slot.
template <bool GUARD, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE, int LS, int NX>
__device__ __forceinline__ void sweeps(double (&v)[NX], const Tab<PERLINE, HASF, FWD, HASR, HASNU>& T,
                                       int base, int n,
                                       int P, int w, int lane, double* xch)
{
    constexpr int CH = PRE ? NX - 2 : NX;
    // table slots: 0 F, 1 LM, 2 R, 3 MU, 4 NU, 5 pre
    if (PRE) {
#pragma unroll
        for (int k = 0; k < CH; ++k)
            if (!GUARD || base + k < n) v[k] = v[k] + T(5, base + k) * v[k + 2];
    }
    if (FWD) {
        double ye[2] = {0.0, 0.0}, pe[2] = {1.0, 1.0};
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int p = k & 1;
            if (!GUARD || base + k < n) {
                const double f = HASF ? T(0, base + k) : 1.0;
                const double m = T(1, base + k);
                ye[p] = f * v[k] + m * ye[p];
                pe[p] = m * pe[p];
            }
            if (PERLINE && (k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
        xch[(w * 4 + 0) * LS + lane] = ye[0];
        xch[(w * 4 + 1) * LS + lane] = ye[1];
        xch[(w * 4 + 2) * LS + lane] = pe[0];
        xch[(w * 4 + 3) * LS + lane] = pe[1];
        __syncthreads();
        double y[2] = {0.0, 0.0};
        for (int u = 0; u < w; ++u) {
#pragma unroll
            for (int p = 0; p < 2; ++p)
                y[p] = xch[(u * 4 + p) * LS + lane] + xch[(u * 4 + 2 + p) * LS + lane] * y[p];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int p = k & 1;
            if (!GUARD || base + k < n) {
                const double f = HASF ? T(0, base + k) : 1.0;
                y[p] = f * v[k] + T(1, base + k) * y[p];
                v[k] = y[p];
            }
            if (PERLINE && (k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
    }
    if constexpr (!HASNU) {
        // backward with one super-diagonal: the chunk's lowest value of each
        // parity as s + a Z, Z = z at the next chunk's lowest position of that
        // parity (a product by a constant zero NU would not be folded away, so
        // the two-term state below is not reused with NU = 0)
        double s[2] = {0, 0}, am[2] = {1, 1};
#pragma unroll
        for (int k = CH - 1; k >= 0; --k) {
            const int p = k & 1;
            if (!GUARD || base + k < n) {
                const double r = HASR ? T(2, base + k) : 1.0;
                const double mu = T(3, base + k);
                s[p] = r * v[k] + mu * s[p];
                am[p] = mu * am[p];
            }
            // at most eight positions' table reads hoisted ahead of their use
            if ((k & 7) == 0) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            xch[(w * 4 + p) * LS + lane] = s[p];
            xch[(w * 4 + 2 + p) * LS + lane] = am[p];
        }
        __syncthreads();
        double Z[2] = {0.0, 0.0};
        for (int u = P - 1; u > w; --u) {
#pragma unroll
            for (int p = 0; p < 2; ++p) Z[p] = xch[(u * 4 + p) * LS + lane] + xch[(u * 4 + 2 + p) * LS + lane] * Z[p];
        }
        __syncthreads();
#pragma unroll
        for (int k = CH - 1; k >= 0; --k) {
            const int p = k & 1;
            if (!GUARD || base + k < n) {
                const double r = HASR ? T(2, base + k) : 1.0;
                Z[p] = r * v[k] + T(3, base + k) * Z[p];
                v[k] = Z[p];
            }
            if ((k & 7) == 0) __builtin_amdgcn_sched_barrier(0);
        }
        return;
    }
    // backward: lowest two same-parity values as affine functions of the
    // incoming (Z1, Z2) = z at the next chunk's lowest positions
    double s1[2] = {0, 0}, a1[2] = {1, 1}, b1[2] = {0, 0};
    double s2[2] = {0, 0}, a2[2] = {0, 0}, b2[2] = {1, 1};
#pragma unroll
    for (int k = CH - 1; k >= 0; --k) {
        const int p = k & 1;
        if (!GUARD || base + k < n) {
            const double r = HASR ? T(2, base + k) : 1.0;
            const double mu = T(3, base + k);
            const double nu = HASNU ? T(4, base + k) : 0.0;
            const double sv = r * v[k] + mu * s1[p] + nu * s2[p];
            const double al = mu * a1[p] + nu * a2[p];
            const double be = mu * b1[p] + nu * b2[p];
            s2[p] = s1[p]; a2[p] = a1[p]; b2[p] = b1[p];
            s1[p] = sv; a1[p] = al; b1[p] = be;
        }
        if (PERLINE && (k & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double* o = xch + ((w * 2 + p) * 6) * LS + lane;
        o[0] = s1[p]; o[LS] = a1[p]; o[2 * LS] = b1[p];
        o[3 * LS] = s2[p]; o[4 * LS] = a2[p]; o[5 * LS] = b2[p];
    }
    __syncthreads();
    double Z1[2] = {0.0, 0.0}, Z2[2] = {0.0, 0.0};
    for (int u = P - 1; u > w; --u) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double* o = xch + ((u * 2 + p) * 6) * LS + lane;
            const double z1 = o[0] + o[LS] * Z1[p] + o[2 * LS] * Z2[p];
            const double z2 = o[3 * LS] + o[4 * LS] * Z1[p] + o[5 * LS] * Z2[p];
            Z1[p] = z1;
            Z2[p] = z2;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = CH - 1; k >= 0; --k) {
        const int p = k & 1;
        if (!GUARD || base + k < n) {
            const double r = HASR ? T(2, base + k) : 1.0;
            const double z = r * v[k] + T(3, base + k) * Z1[p] + (HASNU ? T(4, base + k) * Z2[p] : 0.0);
            v[k] = z;
            Z2[p] = Z1[p];
            Z1[p] = z;
        }
        if (PERLINE && (k & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// Occupancy hint (measured at n = 254, 512 lines of 256): the pre-stencil
// (tri-diagonal projection) form with 16 positions per chunk runs within 128
// VGPRs (4 waves per SIMD); with 32 positions it runs best capped at 168 (3
// waves per SIMD: under 128 it spills, and along rows the spills cost more than
// the extra wave gains); the four-diagonal and Poisson forms spill under 128
// and run best uncapped, except the 32-line, 16-position shared-table form.
//
// LPB = lines per block.  64: lane = line, wave = chunk (P waves).  32: the two
// half-waves hold two neighbouring chunks of the same 32 lines (ceil(P / 2)
// waves), which halves the block's LDS and threads so that two blocks share a
// CU and one block's loads overlap another's sweeps; accesses along columns
// are two 256-byte row segments per load, and along rows each wave transposes
// 32 x 32 tiles.
template <int CH, bool PRE, bool PERLINE, int LPB>
constexpr int rec_min_waves()
{
    return (PRE && CH == 32) ? 3 : (PRE || (LPB == 32 && CH == 16 && !PERLINE)) ? 4 : 1;
}

template <int AXIS, int CH, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE, int MAXT,
          int LPB>
__global__ void __launch_bounds__(MAXT, (rec_min_waves<CH, PRE, PERLINE, LPB>())) recurrence_kernel(RecArgs a)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int li = LPB == 64 ? lane : (lane & 31);            // line within the block
    const int w = LPB == 64 ? wv : 2 * wv + (lane >> 5);       // chunk
    const int P = LPB == 64 ? (int)(blockDim.x >> 6) : 2 * (int)(blockDim.x >> 6);
    const int bpe = (a.nlines + LPB - 1) / LPB;
    // along columns (axis 0) neighbouring blocks share the cache lines at the
    // edges of their row segments: they run on one XCD (one L2)
    const int blk = AXIS == 0 ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    const int b = blk / bpe;
    const int l0 = (blk - b * bpe) * LPB;
    const int line = l0 + li;
    const bool live = line < a.nlines;
    const int base = w * CH;
    const double* __restrict__ src = a.src + (long long)b * a.src_bs;
    double* __restrict__ dst = a.dst + (long long)b * a.dst_bs;
    const int n = a.n;

    // LDS: [present tables, n each (shared mode)] [transpose tiles (axis 1), reused
    // as the carry-exchange area P*12*LPB during the sweeps]
    //
    // Every global load below is unconditional: an element past the end of
    // the input line or past the last line reads a clamped, valid address.
    // Such values are never used (the sweeps skip positions >= n, and n_in >=
    // n; lanes past the last line store nothing), so they are not even
    // replaced by zeros: a load whose value is selected under a condition is
    // moved by the compiler into a branch that waits for it before the next
    // load issues, while unconditional loads are all in flight together
    // (tables first, then the data, so that the table copy into LDS waits
    // only for the tables).
    double* ltab = lds;
    constexpr bool kPresent[6] = {HASF, FWD, HASR, true, HASNU, PRE};  // F, LM, R, MU, NU, pre
    constexpr int ntab = HASF + FWD + HASR + 1 + HASNU + PRE;
    double* tiles = lds + (PERLINE ? 0 : ntab * n);
    double* xch = tiles;
    // one table entry per thread (the block has at least n threads)
    double tv[6];
    if (!PERLINE) {
        const long long eo = (long long)b * a.tbs;
        const double* src_t[6] = {a.F, a.LM, a.R, a.MU, a.NU, a.pre};
        const int i = min((int)threadIdx.x, n - 1);
#pragma unroll
        for (int t = 0; t < 6; ++t)
            if (kPresent[t]) tv[t] = src_t[t][eo + i];
    }
    Tab<PERLINE, HASF, FWD, HASR, HASNU> T;
    T.nlines = a.nlines;
    T.line = live ? line : 0;
    T.n = n;
    T.g = PERLINE ? nullptr : ltab;
    constexpr int NX = PRE ? CH + 2 : CH;
    double v[NX];
    const int lc = min(line, a.nlines - 1);  // clamped line of this lane

    if constexpr (AXIS == 0) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int pos = base + k;
            v[k] = src[(long long)min(pos, a.n_in - 1) * a.src_ld + lc];
        }
    } else {
        // positions of the lane's transposed tiles, loaded before any tile is
        // written: LPB 64 -- CH / TW rounds of 64 lines x TW positions;
        // LPB 32 -- CH / 16 rounds of 32 lines x 32 positions
        constexpr int NR = LPB == 64 ? CH / TW : CH / 16;
        constexpr int NQ = LPB == 64 ? TW : 16;
        double tl[NR][NQ];
#pragma unroll
        for (int rr = 0; rr < NR; ++rr)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int r = LPB == 64 ? (lane / TW) + q * (64 / TW) : (lane >> 5) + 2 * q;
                const int pos = LPB == 64 ? base + rr * TW + lane % TW : wv * 2 * CH + rr * 32 + (lane & 31);
                const int ln = l0 + r;
                tl[rr][q] = src[(long long)min(ln, a.nlines - 1) * a.src_ld + min(pos, a.n_in - 1)];
            }
        // the pre-stencil's two look-ahead positions: direct loads
#pragma unroll
        for (int k = CH; k < NX; ++k) {
            const int pos = base + k;
            v[k] = src[(long long)lc * a.src_ld + min(pos, a.n_in - 1)];
        }
        if (!PERLINE) {
            const int i = (int)threadIdx.x;
            int slot = 0;
#pragma unroll
            for (int t = 0; t < 6; ++t)
                if (kPresent[t]) {
                    if (i < n) ltab[slot * n + i] = tv[t];
                    ++slot;
                }
        }
        if constexpr (LPB == 64) {
            double* tile = tiles + wv * 64 * TP;
#pragma unroll
            for (int rr = 0; rr < NR; ++rr) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) tile[((lane / TW) + q * (64 / TW)) * TP + lane % TW] = tl[rr][q];
                wave_lds_sync();
#pragma unroll
                for (int c = 0; c < TW; ++c) v[rr * TW + c] = tile[lane * TP + c];
                wave_lds_sync();
            }
        } else {
            // CH = 16: one round holds both half-waves' chunks, CH = 32: round
            // rr holds half-wave rr's chunk
            double* tile = tiles + wv * 32 * TP32;
#pragma unroll
            for (int rr = 0; rr < NR; ++rr) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) tile[((lane >> 5) + 2 * q) * TP32 + (lane & 31)] = tl[rr][q];
                wave_lds_sync();
                if (CH == 16) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) v[k] = tile[li * TP32 + (lane >> 5) * 16 + k];
                } else if ((lane >> 5) == rr) {
#pragma unroll
                    for (int k = 0; k < CH && k < 32; ++k) v[k] = tile[li * TP32 + k];
                }
                wave_lds_sync();
            }
        }
    }
    if (AXIS == 0 && !PERLINE) {
        const int i = (int)threadIdx.x;
        int slot = 0;
#pragma unroll
        for (int t = 0; t < 6; ++t)
            if (kPresent[t]) {
                if (i < n) ltab[slot * n + i] = tv[t];
                ++slot;
            }
    }
    // table copy visible; every wave is done with its tile before the area is
    // reused for the carry exchange
    __syncthreads();

    // unguarded sweeps when every chunk of the wave lies inside the line (a
    // wave-uniform test: the sweeps contain block barriers)
    const bool inside = LPB == 64 ? base + CH <= n : (2 * wv + 2) * CH <= n;
    if constexpr (PERLINE) {
        // per-line mode only ever uses LM, R, MU, NU (Poisson): address them
        // through a 4-slot layout [slot][n][nlines] starting at LM
        Tab<true, HASF, FWD, HASR, HASNU> TL;
        TL.nlines = a.nlines;
        TL.line = live ? line : 0;
        TL.n = n;
        TL.g = a.LM + (live ? line : 0);  // slot s of this lane at g + (s-1)*n*nlines
        if (inside) sweeps<false, PRE, FWD, HASF, HASR, HASNU, true, LPB>(v, TL, base, n, P, w, li, xch);
        else sweeps<true, PRE, FWD, HASF, HASR, HASNU, true, LPB>(v, TL, base, n, P, w, li, xch);
    } else {
        if (inside) sweeps<false, PRE, FWD, HASF, HASR, HASNU, false, LPB>(v, T, base, n, P, w, li, xch);
        else sweeps<true, PRE, FWD, HASF, HASR, HASNU, false, LPB>(v, T, base, n, P, w, li, xch);
    }

    if constexpr (AXIS == 0) {
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int pos = base + k;
            if (live && pos < a.n) dst[(long long)pos * a.dst_ld + line] = v[k];
        }
    } else if constexpr (LPB == 64) {
        double* tile = tiles + wv * 64 * TP;
#pragma unroll
        for (int t0 = 0; t0 < CH; t0 += TW) {
#pragma unroll
            for (int c = 0; c < TW; ++c) tile[lane * TP + c] = v[t0 + c];
            wave_lds_sync();
#pragma unroll
            for (int q = 0; q < TW; ++q) {
                const int r = (lane / TW) + q * (64 / TW);
                const int c = lane % TW;
                const int pos = base + t0 + c;
                const int ln = l0 + r;
                if (ln < a.nlines && pos < a.n) dst[(long long)ln * a.dst_ld + pos] = tile[r * TP + c];
            }
            wave_lds_sync();
        }
    } else {
        double* tile = tiles + wv * 32 * TP32;
        const int wbase = wv * 2 * CH;
#pragma unroll
        for (int rr = 0; rr < CH / 16; ++rr) {
            if (CH == 16) {
#pragma unroll
                for (int k = 0; k < 16; ++k) tile[li * TP32 + (lane >> 5) * 16 + k] = v[k];
            } else if ((lane >> 5) == rr) {
#pragma unroll
                for (int k = 0; k < CH && k < 32; ++k) tile[li * TP32 + k] = v[k];
            }
            wave_lds_sync();
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int r = (lane >> 5) + 2 * q;
                const int c = lane & 31;
                const int pos = wbase + rr * 32 + c;
                const int ln = l0 + r;
                if (ln < a.nlines && pos < a.n) dst[(long long)ln * a.dst_ld + pos] = tile[r * TP32 + c];
            }
            wave_lds_sync();
        }
    }
}

template <int AXIS, int CH, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE, int MAXT,
          int LPB>
static int launch_rec_t(const RecArgs& a, hipStream_t s)
{
    const int P = (a.n + CH - 1) / CH;
    const int waves = LPB == 64 ? P : (P + 1) / 2;
    const int chunks = LPB == 64 ? P : 2 * waves;
    const int bpe = (a.nlines + LPB - 1) / LPB;
    const long long blocks = (long long)bpe * a.batch;
    if (blocks == 0) return OK;
    PYPDE_REQUIRE(waves * 64 <= MAXT, "recurrence: %d waves exceed the launch bound", waves);
    PYPDE_REQUIRE(waves * 64 >= a.n, "recurrence: %d threads cannot copy %d table entries", waves * 64, a.n);
    PYPDE_REQUIRE(blocks <= 0x7fffffffLL, "recurrence: %lld blocks exceed the grid", blocks);
    const int ntab = (a.F != nullptr) + (a.LM != nullptr) + (a.R != nullptr) + (a.MU != nullptr) +
                     (a.NU != nullptr) + (a.pre != nullptr);
    size_t tabs = PERLINE ? 0 : (size_t)ntab * a.n;
    size_t tiles = AXIS == 1 ? (size_t)waves * (LPB == 64 ? 64 * TP : 32 * TP32) : 0;
    // carry exchange: 4 slots per chunk for the forward and one-super-diagonal
    // backward sweeps, 12 for the two-super-diagonal backward sweep
    size_t xch = (size_t)chunks * (HASNU ? 12 : 4) * LPB;
    const size_t shmem = (tabs + std::max(tiles, xch)) * sizeof(double);
    auto fn = recurrence_kernel<AXIS, CH, PRE, FWD, HASF, HASR, HASNU, PERLINE, MAXT, LPB>;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    PYPDE_REQUIRE(shmem <= 160 * 1024, "recurrence: LDS %zu too large", shmem);
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(waves * 64), shmem, s, a);
    return check_launch("pypde_recurrence_solve");
}

// Block size (waves * 64 threads); the launch bound is the smallest that fits
// so the compiler may give each lane enough registers for its chunk.
template <int AXIS, int CH, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE, int LPB>
static int launch_rec_ch(const RecArgs& a, hipStream_t s)
{
    const int P = (a.n + CH - 1) / CH;
    const int waves = LPB == 64 ? P : (P + 1) / 2;
    if (waves <= 4) return launch_rec_t<AXIS, CH, PRE, FWD, HASF, HASR, HASNU, PERLINE, 256, LPB>(a, s);
    if (waves <= 8) return launch_rec_t<AXIS, CH, PRE, FWD, HASF, HASR, HASNU, PERLINE, 512, LPB>(a, s);
    return launch_rec_t<AXIS, CH, PRE, FWD, HASF, HASR, HASNU, PERLINE, 1024, LPB>(a, s);
}

template <bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE, int LPB>
static int launch_rec_lpb(const RecArgs& a, hipStream_t s, bool long32)
{
    PYPDE_REQUIRE(a.n >= 1 && a.n <= kMaxWaves * 32, "recurrence: line length %d exceeds %d", a.n,
                  kMaxWaves * 32);
    // 16 positions per chunk up to n = 256, except for the pre-stencil form in
    // 64-line blocks (16 waves of it do not fit two blocks per CU) and when 32
    // are asked for (long32: half the chunks, so half the carry chain)
    const bool short16 = !long32 && (!PRE || LPB == 32) && a.n <= kMaxWaves * 16;
    if (a.axis == 0) {
        if (short16) return launch_rec_ch<0, 16, PRE, FWD, HASF, HASR, HASNU, PERLINE, LPB>(a, s);
        return launch_rec_ch<0, 32, PRE, FWD, HASF, HASR, HASNU, PERLINE, LPB>(a, s);
    }
    if (short16) return launch_rec_ch<1, 16, PRE, FWD, HASF, HASR, HASNU, PERLINE, LPB>(a, s);
    return launch_rec_ch<1, 32, PRE, FWD, HASF, HASR, HASNU, PERLINE, LPB>(a, s);
}

// Block form with 64 (half = false) or 32 (half = true) lines per block;
// long32: 32 positions per chunk at every length.
template <bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool PERLINE>
static int launch_rec(const RecArgs& a, hipStream_t s, bool half, bool long32)
{
    if (half) return launch_rec_lpb<PRE, FWD, HASF, HASR, HASNU, PERLINE, 32>(a, s, long32);
    return launch_rec_lpb<PRE, FWD, HASF, HASR, HASNU, PERLINE, 64>(a, s, long32);
}

// ---------------------------------------------------------------------------
// Lane-chunk form, for launches with few lines (one instance: a few hundred
// lines, which the block form above spreads over only a handful of CUs).
// One wave per line; lane l holds positions [l*C, l*C + C) (C = 2, 4 or 8, so
// n <= 64*C; C even keeps each lane's first position even, so the parity of
// a position is the parity of its index in the chunk).  The chunk transfers of
// the two sweeps are combined across the 64 lanes by scans: the forward
// chain's affine maps y -> A y + B upward, the backward chain's 2-D maps
// (Z1, Z2) -> s + M (Z1, Z2) downward, each as four DPP shift steps within the
// rows of 16 lanes and one step across the rows with the rows' totals read
// from single lanes (no LDS traffic, unlike shuffles).  Tables come straight
// from global memory (shared tables are a few KB and stay in cache; per-line
// tables are read line-major, [nlines][n], so a wave reads its line's
// positions contiguously).  No LDS and no block barrier: waves are independent.

// Cross-lane moves for the lane-chunk scans.  dpp_f64<CTRL>: a DPP move of a
// double (two 32-bit moves); lanes whose source lies outside the pattern keep
// `old`.  Controls used: row_shr:d (0x110 + d) and row_shl:d (0x100 + d) move
// within rows of 16 lanes, wave_shr:1 (0x138) and wave_shl:1 (0x130) across the
// wave.  readlane_f64: one lane's value, for every lane.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double old, double x)
{
    const unsigned long long ov = __builtin_bit_cast(unsigned long long, old);
    const unsigned long long xv = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)ov, (int)(unsigned)xv, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(ov >> 32), (int)(unsigned)(xv >> 32), CTRL,
                                               0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const unsigned long long xv = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)xv, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(xv >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Forward chain maps y -> A y + B; (A, B) after (Ap, Bp): A Ap, A Bp + B.
// One step of the inclusive scan within rows of 16 lanes: lane l composes the
// map of lane l - D of its row (when there is one) before its own.
template <int D>
__device__ __forceinline__ void fwd_row_step(double& A, double& B, int lane)
{
    const double Au = dpp_f64<0x110 + D>(1.0, A);
    const double Bu = dpp_f64<0x110 + D>(0.0, B);
    const double nb = A * Bu + B, na = A * Au;
    const bool in = (lane & 15) >= D;
    B = in ? nb : B;
    A = in ? na : A;
}

// Backward chain maps (Z1, Z2) -> s + M (Z1, Z2), M = [a1 b1; a2 b2]; a lane's
// map applied after the map of the lane above it (its input is that lane's
// output): own.s1 + own.a1 t.s1 + own.b1 t.s2, ...
struct BwdMap {
    double s1, s2, a1, b1, a2, b2;
};

__device__ __forceinline__ BwdMap bwd_compose(const BwdMap& x, const BwdMap& t)
{
    return BwdMap{x.s1 + x.a1 * t.s1 + x.b1 * t.s2, x.s2 + x.a2 * t.s1 + x.b2 * t.s2,
                  x.a1 * t.a1 + x.b1 * t.a2,         x.a1 * t.b1 + x.b1 * t.b2,
                  x.a2 * t.a1 + x.b2 * t.a2,         x.a2 * t.b1 + x.b2 * t.b2};
}

template <int D>
__device__ __forceinline__ void bwd_row_step(BwdMap& m, int lane)
{
    const BwdMap t{dpp_f64<0x100 + D>(0.0, m.s1), dpp_f64<0x100 + D>(0.0, m.s2),
                   dpp_f64<0x100 + D>(1.0, m.a1), dpp_f64<0x100 + D>(0.0, m.b1),
                   dpp_f64<0x100 + D>(0.0, m.a2), dpp_f64<0x100 + D>(1.0, m.b2)};
    const BwdMap c = bwd_compose(m, t);
    const bool in = (lane & 15) + D < 16;
    m.s1 = in ? c.s1 : m.s1; m.s2 = in ? c.s2 : m.s2;
    m.a1 = in ? c.a1 : m.a1; m.b1 = in ? c.b1 : m.b1;
    m.a2 = in ? c.a2 : m.a2; m.b2 = in ? c.b2 : m.b2;
}

__device__ __forceinline__ BwdMap bwd_readlane(const BwdMap& m, int l)
{
    return BwdMap{readlane_f64(m.s1, l), readlane_f64(m.s2, l), readlane_f64(m.a1, l),
                  readlane_f64(m.b1, l), readlane_f64(m.a2, l), readlane_f64(m.b2, l)};
}

// The lane's table entries and inputs are fetched into registers before any
// of them is used: positions past the line's end read a clamped, valid entry
// (the last position of the line) whose value is then discarded, so every
// load is unconditional and the wave keeps all of them in flight at once
// instead of waiting for each behind a branch.
template <int C, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU>
struct LaneTab {
    double t[6][C];  // F, LM, R, MU, NU, pre at the lane's positions (present tables only)
    __device__ __forceinline__ double operator()(int which, int k) const { return t[which][k]; }
};

template <int AXIS, int C, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool LINEMAJOR>
__device__ __forceinline__ void lane_chunk_line(const RecArgs& a, long long gl)
{
    constexpr int NX = PRE ? C + 2 : C;
    const int lane = threadIdx.x;
    const int b = (int)(gl / a.nlines);
    const int line = (int)(gl - (long long)b * a.nlines);
    const int n = a.n;
    const int base = lane * C;
    const double* __restrict__ src = a.src + (long long)b * a.src_bs;
    double* __restrict__ dst = a.dst + (long long)b * a.dst_bs;

    LaneTab<C, PRE, FWD, HASF, HASR, HASNU> T;
    {
        const double* tabs[6] = {a.F, a.LM, a.R, a.MU, a.NU, a.pre};
        const bool present[6] = {HASF, FWD, HASR, true, HASNU, PRE};
        const long long off = (long long)b * a.tbs + (LINEMAJOR ? (long long)line * n : 0);
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            if (!present[t]) continue;
            const double* __restrict__ tp = tabs[t] + off;
#pragma unroll
            for (int k = 0; k < C; ++k) T.t[t][k] = tp[min(base + k, n - 1)];
        }
    }

    double v[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) {
        const int pos = base + k;
        const int pc = min(pos, a.n_in - 1);
        const double x = AXIS == 0 ? src[(long long)pc * a.src_ld + line] : src[(long long)line * a.src_ld + pc];
        v[k] = pos < a.n_in ? x : 0.0;
    }

    if (PRE) {
#pragma unroll
        for (int k = 0; k < C; ++k)
            if (base + k < n) v[k] = v[k] + T(5, k) * v[k + 2];
    }

    if (FWD) {
        // chunk transfer per parity: y_out = A y_in + B
        double A[2] = {1.0, 1.0}, B[2] = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const int p = k & 1;
            if (base + k < n) {
                const double f = HASF ? T(0, k) : 1.0;
                const double m = T(1, k);
                B[p] = f * v[k] + m * B[p];
                A[p] = m * A[p];
            }
        }
        // inclusive scan upward (lane l composes its map after lanes < l):
        // within rows of 16 lanes by DPP shifts, then across the four rows
        // with the rows' totals read from their last lanes (15, 31, 47)
        double y[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            fwd_row_step<1>(A[p], B[p], lane);
            fwd_row_step<2>(A[p], B[p], lane);
            fwd_row_step<4>(A[p], B[p], lane);
            fwd_row_step<8>(A[p], B[p], lane);
            const double A0 = readlane_f64(A[p], 15), B0 = readlane_f64(B[p], 15);
            const double A1 = readlane_f64(A[p], 31), B1 = readlane_f64(B[p], 31);
            const double A2 = readlane_f64(A[p], 47), B2 = readlane_f64(B[p], 47);
            const double PA2 = A1 * A0, PB2 = A1 * B0 + B1;     // rows 0-1
            const double PA3 = A2 * PA2, PB3 = A2 * PB2 + B2;   // rows 0-2
            const int row = lane >> 4;
            const double PA = row == 1 ? A0 : (row == 2 ? PA2 : PA3);
            const double PB = row == 1 ? B0 : (row == 2 ? PB2 : PB3);
            const double nb = A[p] * PB + B[p], na = A[p] * PA;
            B[p] = row > 0 ? nb : B[p];
            A[p] = row > 0 ? na : A[p];
            // the chunk's incoming value: the lane below's outgoing one
            y[p] = dpp_f64<0x138>(0.0, B[p]);
        }
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const int p = k & 1;
            if (base + k < n) {
                const double f = HASF ? T(0, k) : 1.0;
                y[p] = f * v[k] + T(1, k) * y[p];
                v[k] = y[p];
            }
        }
    }

    // backward: the chunk's lowest two same-parity values as an affine map of
    // the next chunk's lowest two, (Z1, Z2) -> s + M (Z1, Z2)
    double s1[2] = {0, 0}, a1[2] = {1, 1}, b1[2] = {0, 0};
    double s2[2] = {0, 0}, a2[2] = {0, 0}, b2[2] = {1, 1};
#pragma unroll
    for (int k = C - 1; k >= 0; --k) {
        const int p = k & 1;
        if (base + k < n) {
            const double r = HASR ? T(2, k) : 1.0;
            const double mu = T(3, k);
            const double nu = HASNU ? T(4, k) : 0.0;
            const double sv = r * v[k] + mu * s1[p] + nu * s2[p];
            const double al = mu * a1[p] + nu * a2[p];
            const double be = mu * b1[p] + nu * b2[p];
            s2[p] = s1[p]; a2[p] = a1[p]; b2[p] = b1[p];
            s1[p] = sv; a1[p] = al; b1[p] = be;
        }
    }
    // inclusive scan downward (lane l's map applied after the maps of lanes >
    // l): within rows of 16 lanes by DPP shifts, then across the rows with the
    // rows' totals read from their first lanes (16, 32, 48)
    double Z1[2], Z2[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        BwdMap m{s1[p], s2[p], a1[p], b1[p], a2[p], b2[p]};
        bwd_row_step<1>(m, lane);
        bwd_row_step<2>(m, lane);
        bwd_row_step<4>(m, lane);
        bwd_row_step<8>(m, lane);
        const BwdMap R1 = bwd_readlane(m, 16), R2 = bwd_readlane(m, 32), R3 = bwd_readlane(m, 48);
        const BwdMap Q1 = bwd_compose(R2, R3);   // rows 2-3
        const BwdMap Q0 = bwd_compose(R1, Q1);   // rows 1-3
        const int row = lane >> 4;
        const BwdMap& Q = row == 0 ? Q0 : (row == 1 ? Q1 : R3);
        const BwdMap c = bwd_compose(m, Q);
        const bool up = row < 3;
        // the chunk's incoming values: the lane above's outgoing ones
        Z1[p] = dpp_f64<0x130>(0.0, up ? c.s1 : m.s1);
        Z2[p] = dpp_f64<0x130>(0.0, up ? c.s2 : m.s2);
    }
#pragma unroll
    for (int k = C - 1; k >= 0; --k) {
        const int p = k & 1;
        if (base + k < n) {
            const double r = HASR ? T(2, k) : 1.0;
            const double z = r * v[k] + T(3, k) * Z1[p] + (HASNU ? T(4, k) * Z2[p] : 0.0);
            v[k] = z;
            Z2[p] = Z1[p];
            Z1[p] = z;
        }
    }

#pragma unroll
    for (int k = 0; k < C; ++k) {
        const int pos = base + k;
        if (pos < n) {
            if (AXIS == 0) dst[(long long)pos * a.dst_ld + line] = v[k];
            else dst[(long long)line * a.dst_ld + pos] = v[k];
        }
    }
}

template <int AXIS, int C, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool LINEMAJOR>
__global__ void __launch_bounds__(64) lane_chunk_kernel(RecArgs a)
{
    // (batch entry, line); along axis 0 neighbouring lines share every cache
    // line they touch, so they run on one XCD
    const long long gl = AXIS == 0 ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : blockIdx.x;
    lane_chunk_line<AXIS, C, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, gl);
}

// Lane-chunk solves of many members (the same table form, each its own line
// geometry and tables) in one launch, the arguments in a device table: block
// t runs line t - start[i] of member i.  Lines shorter than 64 C leave their
// upper lanes idle, as in a single launch with the same C.
constexpr int kMaxRecMulti = 64;
struct RecMulti {
    int count, total;
    int start[kMaxRecMulti + 1];
    RecArgs a[kMaxRecMulti];
};

template <int AXIS, int C, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool LINEMAJOR>
__global__ void __launch_bounds__(64) lane_chunk_multi(const RecMulti* __restrict__ mm)
{
    const int t = AXIS == 0 ? xcd_remap((int)blockIdx.x, mm->total) : (int)blockIdx.x;
    int lo = 0, hi = mm->count - 1;
    while (lo < hi) {  // the last member whose range starts at or before t
        const int mid = (lo + hi + 1) >> 1;
        if (mm->start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const RecArgs a = mm->a[lo];
    lane_chunk_line<AXIS, C, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, t - mm->start[lo]);
}

template <int AXIS, int C, bool LINEMAJOR>
static void launch_lane_multi_t(const RecMulti* mm, int total, hipStream_t s)
{
    hipLaunchKernelGGL((lane_chunk_multi<AXIS, C, false, true, false, true, true, LINEMAJOR>),
                       dim3((unsigned)total), dim3(64), 0, s, mm);
}

template <int AXIS, int C, bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool LINEMAJOR>
static int launch_lane_t(const RecArgs& a, hipStream_t s)
{
    const long long blocks = (long long)a.nlines * a.batch;
    if (blocks == 0) return OK;
    PYPDE_REQUIRE(blocks <= 0x7fffffffLL, "recurrence: %lld lines exceed the grid", blocks);
    hipLaunchKernelGGL((lane_chunk_kernel<AXIS, C, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>),
                       dim3((unsigned)blocks), dim3(64), 0, s, a);
    return check_launch("pypde_recurrence_solve");
}

template <bool PRE, bool FWD, bool HASF, bool HASR, bool HASNU, bool LINEMAJOR>
static int launch_lane(const RecArgs& a, hipStream_t s)
{
    PYPDE_REQUIRE(a.n >= 1 && a.n <= 64 * 8, "recurrence: line length %d exceeds 512", a.n);
    if (a.axis == 0) {
        if (a.n <= 128) return launch_lane_t<0, 2, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
        if (a.n <= 256) return launch_lane_t<0, 4, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
        return launch_lane_t<0, 8, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
    }
    if (a.n <= 128) return launch_lane_t<1, 2, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
    if (a.n <= 256) return launch_lane_t<1, 4, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
    return launch_lane_t<1, 8, PRE, FWD, HASF, HASR, HASNU, LINEMAJOR>(a, s);
}

long long recurrence_multi_bytes() { return (long long)sizeof(RecMulti); }

// flags of a recurrence multi table: bit 0 axis, bit 1 line-major per-line
// tables, bits 2.. log2 of the chunk C (2, 4 or 8)
int recurrence_multi_pack(const pypde_lines* g, const pypde_recurrence* r, int count, void* table,
                          long long table_bytes, pypde_multi_info* info)
{
    PYPDE_REQUIRE(g && r && table && info, "null argument");
    PYPDE_REQUIRE(table_bytes >= (long long)sizeof(RecMulti), "recurrence table needs %zu bytes",
                  sizeof(RecMulti));
    PYPDE_REQUIRE(count >= 1 && count <= kMaxRecMulti, "multi recurrence: count %d not in 1..%d",
                  count, kMaxRecMulti);
    RecMulti* mm = static_cast<RecMulti*>(table);
    memset(mm, 0, sizeof(RecMulti));
    mm->count = count;
    const int axis = g[0].axis;
    const int linemajor = r[0].per_line == 2;
    int nmax = 0;
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        const pypde_lines& gi = g[i];
        const pypde_recurrence& ri = r[i];
        PYPDE_REQUIRE(gi.src && gi.dst, "member %d: null src/dst pointer", i);
        PYPDE_REQUIRE(gi.axis == axis, "member %d: another axis", i);
        PYPDE_REQUIRE((ri.per_line == 2) == (linemajor != 0) && ri.per_line != 1,
                      "member %d: tables must be shared/per entry (per_line 0) or line-major "
                      "(per_line 2) like member 0", i);
        PYPDE_REQUIRE(!ri.pre && ri.LM && !ri.F && ri.R && ri.MU && ri.NU,
                      "member %d: multi launches support the (LM, R, MU, NU) form only", i);
        PYPDE_REQUIRE(gi.len_out >= 3 && gi.len_in == gi.len_out && gi.nlines >= 0 && gi.batch >= 0,
                      "member %d: invalid line geometry", i);
        PYPDE_REQUIRE(gi.len_out <= 512, "member %d: line length %d exceeds 512", i, gi.len_out);
        mm->a[i] = RecArgs{gi.src, gi.dst, gi.src_ld, gi.dst_ld, gi.src_bs, gi.dst_bs, gi.len_in,
                           gi.len_out, gi.nlines, gi.batch, gi.axis, ri.F, ri.LM, ri.R, ri.MU,
                           ri.NU, ri.pre, ri.entry_stride, ri.per_line};
        mm->start[i] = (int)total;
        total += (long long)gi.nlines * gi.batch;
        PYPDE_REQUIRE(total < (1LL << 31), "multi recurrence too large");
        if (gi.len_out > nmax) nmax = gi.len_out;
    }
    mm->start[count] = (int)total;
    mm->total = (int)total;
    const int lc = nmax <= 128 ? 1 : (nmax <= 256 ? 2 : 3);
    info->kind = PYPDE_MULTI_RECURRENCE;
    info->count = count;
    info->variant = 2;
    info->flags = axis | (linemajor << 1) | (lc << 2);
    info->total = total;
    return OK;
}

int recurrence_multi(const void* table, const pypde_multi_info* info, hipStream_t s)
{
    PYPDE_REQUIRE(table && info, "null argument");
    PYPDE_REQUIRE(info->kind == PYPDE_MULTI_RECURRENCE, "not a recurrence table");
    if (info->total == 0) return OK;
    const RecMulti* mm = static_cast<const RecMulti*>(table);
    const int axis = info->flags & 1, lm = (info->flags >> 1) & 1, lc = info->flags >> 2;
    const int total = (int)info->total;
    PYPDE_REQUIRE(lc >= 1 && lc <= 3, "recurrence table: bad chunk code %d", lc);
#define PYPDE_LANE_MULTI(AX, CC)                                                   \
    (lm ? launch_lane_multi_t<AX, CC, true>(mm, total, s)                         \
        : launch_lane_multi_t<AX, CC, false>(mm, total, s))
    if (axis == 0) {
        if (lc == 1) PYPDE_LANE_MULTI(0, 2);
        else if (lc == 2) PYPDE_LANE_MULTI(0, 4);
        else PYPDE_LANE_MULTI(0, 8);
    } else {
        if (lc == 1) PYPDE_LANE_MULTI(1, 2);
        else if (lc == 2) PYPDE_LANE_MULTI(1, 4);
        else PYPDE_LANE_MULTI(1, 8);
    }
#undef PYPDE_LANE_MULTI
    return check_launch("pypde_recurrence_multi");
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_recurrence_multi_pack(const pypde_lines* g, const pypde_recurrence* r,
                                           int count, void* table, long long table_bytes,
                                           pypde_multi_info* info)
{
    return recurrence_multi_pack(g, r, count, table, table_bytes, info);
}

extern "C" int pypde_recurrence_multi(const void* table, const pypde_multi_info* info, void* stream)
{
    return recurrence_multi(table, info, as_stream(stream));
}

extern "C" int pypde_recurrence_solve(const pypde_lines* g, const pypde_recurrence* r, void* stream)
{
    PYPDE_REQUIRE(g && r, "null argument");
    PYPDE_REQUIRE(g->src && g->dst, "null src/dst pointer");
    PYPDE_REQUIRE(g->axis == 0 || g->axis == 1, "axis must be 0 or 1 (got %d)", g->axis);
    PYPDE_REQUIRE(r->MU, "recurrence needs the backward multiplier table MU");
    const bool pre = r->pre != nullptr;
    PYPDE_REQUIRE(g->len_out >= 3 && g->nlines >= 0 && g->batch >= 0, "invalid line geometry");
    PYPDE_REQUIRE(g->len_in == g->len_out + (pre ? 2 : 0),
                  "recurrence: len_in must be len_out%s", pre ? " + 2" : "");
    PYPDE_REQUIRE(!(g->src == g->dst && pre), "pre-stencil cannot run in place");
    PYPDE_REQUIRE(r->per_line >= 0 && r->per_line <= 2, "per_line must be 0, 1 or 2");
    PYPDE_REQUIRE(r->variant >= 0 && r->variant <= 4, "variant must be 0, 1, 2, 3 or 4");
    RecArgs a{g->src, g->dst, g->src_ld, g->dst_ld, g->src_bs, g->dst_bs, g->len_in, g->len_out,
              g->nlines, g->batch, g->axis, r->F, r->LM, r->R, r->MU, r->NU, r->pre,
              r->entry_stride, r->per_line};
    hipStream_t s = as_stream(stream);
    const bool fwd = r->LM != nullptr, hf = r->F != nullptr, hr = r->R != nullptr,
               hn = r->NU != nullptr;
    // lane-chunk form: asked for, required by line-major tables, or (variant 0)
    // chosen when the launch has few lines
    const long long lines = (long long)g->nlines * g->batch;
    PYPDE_REQUIRE(!(r->variant >= 3 && r->per_line == 2),
                  "the block forms read per-line tables [n][nlines] (per_line = 1)");
    const bool lane = r->variant == 2 || r->per_line == 2 ||
                      (r->variant == 0 && r->per_line == 0 && lines <= 2048);
    if (lane) {
        PYPDE_REQUIRE(r->per_line != 1, "the lane-chunk form reads per-line tables line-major (per_line = 2)");
        if (r->per_line == 2) {
            PYPDE_REQUIRE(!pre && fwd && !hf && hr && hn,
                          "per-line tables support the (LM, R, MU, NU) form only");
            return launch_lane<false, true, false, true, true, true>(a, s);
        }
        if (!pre && fwd && !hf && hr && hn) return launch_lane<false, true, false, true, true, false>(a, s);
        if (pre && fwd && hf && !hr && !hn) return launch_lane<true, true, true, false, false, false>(a, s);
        if (!pre && !fwd && !hf && hr && !hn) return launch_lane<false, false, false, true, false, false>(a, s);
        if (!pre && fwd && hf && hr && hn) return launch_lane<false, true, true, true, true, false>(a, s);
        set_error("pypde_recurrence_solve: unsupported table combination (pre=%d F=%d LM=%d R=%d NU=%d)",
                  (int)pre, (int)hf, (int)fwd, (int)hr, (int)hn);
        return ERR_UNSUP;
    }
    // block form: 64 lines per block (variant 0 / 1) or 32 (variant 3; variant
    // 4: 32 lines and 32 positions per chunk)
    const bool half = r->variant >= 3, long32 = r->variant == 4;
    // the combinations used by the plans
    if (r->per_line) {
        PYPDE_REQUIRE(!pre && fwd && !hf && hr && hn,
                      "per-line tables support the (LM, R, MU, NU) form only");
        PYPDE_REQUIRE(r->R == r->LM + (long long)g->len_out * g->nlines &&
                      r->MU == r->R + (long long)g->len_out * g->nlines &&
                      r->NU == r->MU + (long long)g->len_out * g->nlines,
                      "per-line tables must be contiguous slots LM | R | MU | NU of [n][nlines]");
        return launch_rec<false, true, false, true, true, true>(a, s, half, long32);
    }
    if (!pre && fwd && !hf && hr && hn) return launch_rec<false, true, false, true, true, false>(a, s, half, long32);
    if (pre && fwd && hf && !hr && !hn) return launch_rec<true, true, true, false, false, false>(a, s, half, long32);
    if (!pre && !fwd && !hf && hr && !hn) return launch_rec<false, false, false, true, false, false>(a, s, half, long32);
    if (!pre && fwd && hf && hr && hn) return launch_rec<false, true, true, true, true, false>(a, s, half, long32);
    set_error("pypde_recurrence_solve: unsupported table combination (pre=%d F=%d LM=%d R=%d NU=%d)",
              (int)pre, (int)hf, (int)fwd, (int)hr, (int)hn);
    return ERR_UNSUP;
}
