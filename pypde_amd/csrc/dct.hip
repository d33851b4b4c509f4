// Chebyshev transforms as LDS-staged cosine sums on the FP64 vector units
// (the alternative to the dense MFMA contraction; the tuner picks one per
// shape).
//
//   y[o] = a_o * sum_i cos(pi o i / P) * (b_i x[i]),   o < n_out = P + 1, i < n_in
//
// covers the backward transform (a = 1, b_i = (-1)^i, P = grid points - 1,
// chebyshev.py:78-83) and the square forward transform with its mass weights
// (a_o = w_o (-1)^o / (2P), b_i = 1 at the ends and 2 inside,
// chebyshev.py:67-76,143-149).  The outputs come in mirrored pairs:
// cos(pi (P - o) i / P) = (-1)^i cos(pi o i / P), so y[o] = a_o (E + O) and
// y[P - o] = a_(P-o) (E - O) from the even-i and odd-i partial sums E, O of one
// pass (half the multiply-adds of the dense product).
//
// Layout: a block serves 64 * LC lines (lane l holds lines l, l + 64, ...) and
// a range of output pairs over its waves, RO pairs per wave.  The input lines
// are staged through LDS in chunks of KC positions, scaled by b on the way
// in (coalesced loads along the other axis for axis-0 lines; rows of a tile
// for axis-1 lines).  The cosine of every multiple of pi / P comes from a
// table of 2P values in LDS, read with wave-uniform addresses (broadcast); a
// pair's table index advances by o per input position, modulo 2P.
#include "common.h"
#include "../../include/pypde_amd.h"

namespace pypde {

constexpr int kDctKC = 32;   // input positions per LDS chunk

struct DctArgs {
    const double* src;
    double* dst;
    long long src_ld, dst_ld, src_bs, dst_bs;
    int n_in, n_out, nlines, batch, P;
    const double* a;     // [n_out]
    const double* b;     // [n_in]
    const double* ctab;  // [2P]: cos(pi m / P)
};

template <int AXIS, int RO, int LC, int W>
__global__ void __launch_bounds__(64 * W) dct_fold_kernel(DctArgs g)
{
    constexpr int NL = 64 * LC;           // lines per block
    __shared__ double xs[kDctKC][NL + 1];
    extern __shared__ double ctab_s[];    // 2P values
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int P = g.P, twoP = 2 * P;
    const int npairs = P / 2 + 1;         // o = 0 .. floor(P/2)
    // lines of this block: global line index = entry * nlines + line
    const long long total_lines = (long long)g.nlines * g.batch;
    const long long line0 = (long long)blockIdx.x * NL;
    const int o0 = (blockIdx.y * W + wave) * RO;   // first pair of this wave

    for (int m = tid; m < twoP; m += 64 * W) ctab_s[m] = g.ctab[m];

    // destination of each line this lane holds
    double* dp[LC];
    bool live[LC];
#pragma unroll
    for (int c = 0; c < LC; ++c) {
        const long long gl = line0 + lane + 64 * c;
        live[c] = gl < total_lines;
        const long long e = live[c] ? gl / g.nlines : 0;
        const long long l = live[c] ? gl - e * g.nlines : 0;
        dp[c] = g.dst + e * g.dst_bs + (AXIS == 0 ? l : l * g.dst_ld);
    }

    double E[RO][LC], O[RO][LC];
#pragma unroll
    for (int r = 0; r < RO; ++r)
#pragma unroll
        for (int c = 0; c < LC; ++c) E[r][c] = O[r][c] = 0.0;

    for (int i0 = 0; i0 < g.n_in; i0 += kDctKC) {
        __syncthreads();
        // stage x[i0 .. i0+KC) of the block's lines, scaled by b
        if (AXIS == 0) {
            // row k of the chunk: NL consecutive lines (columns) of one entry run
            for (int q = tid; q < kDctKC * NL; q += 64 * W) {
                const int k = q / NL, c = q - k * NL;
                const long long gl = line0 + c;
                const int i = i0 + k;
                double v = 0.0;
                if (i < g.n_in && gl < total_lines) {
                    const long long e = gl / g.nlines, l = gl - e * g.nlines;
                    v = g.b[i] * g.src[e * g.src_bs + (long long)i * g.src_ld + l];
                }
                xs[k][c] = v;
            }
        } else {
            // a line is a row: consecutive threads read consecutive positions of it
            for (int q = tid; q < kDctKC * NL; q += 64 * W) {
                const int c = q / kDctKC, k = q - c * kDctKC;
                const long long gl = line0 + c;
                const int i = i0 + k;
                double v = 0.0;
                if (i < g.n_in && gl < total_lines) {
                    const long long e = gl / g.nlines, l = gl - e * g.nlines;
                    v = g.b[i] * g.src[e * g.src_bs + l * g.src_ld + i];
                }
                xs[k][c] = v;
            }
        }
        __syncthreads();
        if (o0 >= npairs) continue;
        int idx[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r) idx[r] = (int)(((long long)(o0 + r) * i0) % twoP);
        const int kmax = min(kDctKC, g.n_in - i0);
        for (int k = 0; k < kmax; k += 2) {
            double x0[LC], x1[LC];
#pragma unroll
            for (int c = 0; c < LC; ++c) {
                x0[c] = xs[k][lane + 64 * c];
                x1[c] = xs[k + 1][lane + 64 * c];   // zero beyond n_in (staged as 0)
            }
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const double c0 = ctab_s[idx[r]];
                int j = idx[r] + o0 + r;
                j = j >= twoP ? j - twoP : j;
                const double c1 = ctab_s[j];
                j += o0 + r;
                idx[r] = j >= twoP ? j - twoP : j;
#pragma unroll
                for (int c = 0; c < LC; ++c) {
                    E[r][c] = fma(c0, x0[c], E[r][c]);   // i0 + k even (i0, k even)
                    O[r][c] = fma(c1, x1[c], O[r][c]);
                }
            }
        }
    }
    // outputs: y[o] = a_o (E + O), y[P - o] = a_(P-o) (E - O)
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        const int o = o0 + r;
        if (o >= npairs) break;
        const int om = P - o;
        const double ao = g.a[o], am = g.a[om];
#pragma unroll
        for (int c = 0; c < LC; ++c) {
            if (!live[c]) continue;
            const double s = E[r][c] + O[r][c], d = E[r][c] - O[r][c];
            if (AXIS == 0) {
                dp[c][(long long)o * g.dst_ld] = ao * s;
                if (om != o) dp[c][(long long)om * g.dst_ld] = am * d;
            } else {
                dp[c][o] = ao * s;
                if (om != o) dp[c][om] = am * d;
            }
        }
    }
}

template <int AXIS, int RO, int LC, int W>
static int launch_dct_w(const DctArgs& g, hipStream_t s)
{
    const long long lines = (long long)g.nlines * g.batch;
    const long long bx = (lines + 64 * LC - 1) / (64 * LC);
    const int npairs = g.P / 2 + 1;
    const int by = (npairs + RO * W - 1) / (RO * W);
    PYPDE_REQUIRE(bx < (1LL << 31) && by < 65536, "transform grid too large");
    const size_t shmem = sizeof(double) * 2 * (size_t)g.P;
    hipLaunchKernelGGL((dct_fold_kernel<AXIS, RO, LC, W>), dim3((unsigned)bx, (unsigned)by),
                       dim3(64 * W), shmem, s, g);
    return check_launch("pypde_cheb_dct");
}

// as many waves per block as the output pairs need (up to 16), so that a
// block stages each input line once
template <int AXIS, int RO, int LC>
static int launch_dct(const DctArgs& g, hipStream_t s)
{
    const int waves = (g.P / 2 + 1 + RO - 1) / RO;
    if (waves <= 4) return launch_dct_w<AXIS, RO, LC, 4>(g, s);
    if (waves <= 8) return launch_dct_w<AXIS, RO, LC, 8>(g, s);
    return launch_dct_w<AXIS, RO, LC, 16>(g, s);
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_cheb_dct(const pypde_lines* g, int P, const double* a, const double* b,
                              const double* ctab, int variant, void* stream)
{
    PYPDE_REQUIRE(g && a && b && ctab, "null argument");
    PYPDE_REQUIRE(g->src && g->dst, "null src/dst pointer");
    PYPDE_REQUIRE(g->src != g->dst, "the transform cannot run in place");
    PYPDE_REQUIRE(g->axis == 0 || g->axis == 1, "axis must be 0 or 1 (got %d)", g->axis);
    PYPDE_REQUIRE(P >= 1 && P <= 2047, "period P must be 1..2047 (got %d)", P);
    PYPDE_REQUIRE(g->len_out == P + 1, "the folded transform writes P + 1 = %d outputs (got %d)",
                  P + 1, g->len_out);
    PYPDE_REQUIRE(g->len_in >= 1 && g->len_in <= P + 1, "len_in must be 1..P+1 (got %d)",
                  g->len_in);
    PYPDE_REQUIRE(g->nlines >= 0 && g->batch >= 0, "invalid line geometry");
    if ((long long)g->nlines * g->batch == 0) return OK;
    DctArgs d{g->src, g->dst, g->src_ld, g->dst_ld, g->src_bs, g->dst_bs, g->len_in, g->len_out,
              g->nlines, g->batch, P, a, b, ctab};
    const hipStream_t s = as_stream(stream);
    // variant: 1 = 8 pairs per wave, one line per lane; 2 = 8 pairs, two lines per
    // lane; 3 = 4 pairs, two lines per lane (0: 2)
    const int v = variant == 0 ? 2 : variant;
    PYPDE_REQUIRE(v >= 1 && v <= 3, "variant must be 0..3 (got %d)", variant);
    if (g->axis == 0) {
        if (v == 1) return launch_dct<0, 8, 1>(d, s);
        if (v == 2) return launch_dct<0, 8, 2>(d, s);
        return launch_dct<0, 4, 2>(d, s);
    }
    if (v == 1) return launch_dct<1, 8, 1>(d, s);
    if (v == 2) return launch_dct<1, 8, 2>(d, s);
    return launch_dct<1, 4, 2>(d, s);
}
