/*
 * pypde_amd C-ABI: MI355X (gfx950) kernels behind pypde's spectral-Galerkin
 * operator surface.  All entry points are extern "C", take plain pointers to
 * device memory (row-major float64 arrays owned by the caller), run
 * asynchronously on the given HIP stream (NULL = default stream), and return
 * 0 on success or a negative code; pypde_last_error() then holds the message.
 *
 * Array geometry used by every along-axis entry point (pypde_lines):
 *   a batch of 2-D row-major arrays, entry b at  ptr + b*bs,  element (i,j) at
 *   ptr + b*bs + i*ld + j.  "axis 0" lines are columns (element stride ld),
 *   "axis 1" lines are rows (element stride 1).  len_in / len_out are the line
 *   lengths along the axis before / after the operation, nlines the extent of
 *   the other axis.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * preiter93/pypde).
 */
#ifndef PYPDE_AMD_H
#define PYPDE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pypde_lines {
    const double* src;
    double* dst;
    long long src_ld, dst_ld; /* elements between rows of one 2-D array */
    long long src_bs, dst_bs; /* elements between batch entries          */
    int len_in, len_out;      /* line length before / after the op        */
    int nlines;               /* number of lines per batch entry          */
    int batch;                /* number of batch entries                  */
    int axis;                 /* 0: columns, 1: rows                      */
    int lines_per_block;      /* 0 = library default (autotuner hint); for
                                 the banded / stencil applies 1..8 outputs
                                 per thread (axis 0) or rows per block
                                 (axis 1), 18 / 20 / 24 the register-window
                                 form with 2 / 4 / 8 outputs per thread,
                                 33..40 (axis 1 only) the coefficient-
                                 resident form: 33..36 four rows in
                                 flight and 4 / 8 / 16 / 32 rows per
                                 block, 37..40 eight rows in flight and
                                 8 / 16 / 32 / 64 rows per block */
} pypde_lines;

/* ---- library ---------------------------------------------------------- */
int pypde_version(void);
const char* pypde_last_error(void);
/* Number of visible HIP devices (0 when none); never throws. */
int pypde_device_count(void);

/* ---- dense operators (MFMA FP64 GEMM) ---------------------------------- */
/* C[b] = alpha * op(A[b]) @ op(B[b]) + beta * D[b]   (row-major, op = T if trans)
 * Replaces the dense / csr products of MatrixBase.dot (pypde/solver/matrix.py:41-56)
 * and carries the Chebyshev transforms as dense contractions against the
 * collocation matrices (pypde/bases/chebyshev.py:67-93 forward_fft/backward_fft/dctn). */
int pypde_dgemm(int transA, int transB, int m, int n, int k, double alpha,
                const double* A, long long lda, long long strideA,
                const double* B, long long ldb, long long strideB,
                double beta, const double* D, long long ldd, long long strideD,
                double* C, long long ldc, long long strideC, int batch, void* stream);

/* Batch index of one GEMM operand: entry e starts at
 *   (e/d1 % m1)*s1 + (e/d2 % m2)*s2   elements   (m = 0: no modulus, d >= 1). */
typedef struct pypde_bidx {
    long long s1, s2;
    int d1, m1, d2, m2;
} pypde_bidx;

typedef struct pypde_gemm_desc {
    int transA, transB, m, n, k, batch;
    double alpha, beta;
    const double* A;
    const double* B;
    const double* D; /* may be NULL (beta ignored) or alias C */
    double* C;
    long long lda, ldb, ldd, ldc;
    pypde_bidx iA, iB, iD, iC;
    const double* alpha_vec; /* NULL, or per-entry scale alpha_vec[e / alpha_div] */
    int alpha_div;
    int variant; /* kernel choice: 0 size heuristic; 1 LDS-tiled 64x64; 2 LDS-tiled
                    32x32; 3 LDS-tiled 128x128; 4 the same, double-buffered; 5 LDS-tiled
                    128x64; 6 LDS-tiled 64x128; 7, 24, 25, 26, 27 LDS-tiled 128x64,
                    64x128, 64x64, 128x128, 32x32 with two slabs of loads in flight;
                    28 LDS-tiled 32x32, double-buffered; 29, 30 LDS-tiled 32x32 over
                    2 waves / 1 wave; 31, 32 the same with two slabs in flight;
                    33, 34 variants 2 and 27 with 32-deep K slabs; 35, 36 the same
                    with 64-deep slabs; 37, 38 variants 1 and 5 with 32-deep slabs;
                    39..42 variants 2, 27, 33, 34 with LDS rows padded by 2 doubles;
                    8..11 split-K LDS-staged (32x32 over 4 / 8 waves, 64x32, 32x64);
                    12..15 the same with two slabs of loads in flight per wave;
                    16..23 split-K register-direct, 16 + (TM-1) + 2 (TN-1)
                    + 4 (KW == 8) for a (16 TM) x (16 TN) tile over KW waves */
    /* optional second addend and gauge (fused pressure update of rbc2d.py:384-394):
     *   C[e] = (beta D[e] - (gamma_vec[e / gamma_div] E[e]) gamma) + alpha_e op(A) op(B)
     * E NULL: no E term (gamma_vec NULL: factor 1); zero_first = 1 sets C[e](0,0) = 0 last. */
    const double* E;
    long long lde;
    pypde_bidx iE;
    const double* gamma_vec;
    double gamma;
    int gamma_div;
    int zero_first;
    long long csn; /* column stride of C, D and E (0 or 1: contiguous) */
    /* parity folds of mirror-symmetric operators (halve the MFMA work), variants 1-5
     * (64x64, 32x32, 128x128, 128x64, 64x128 tiles); for the output folds also 6, 7
     * (32x32 tiles, each product over its own waves with K split 2 / 4 ways):
     *   fold 1 / 2: output rows / columns of C come in blocks of fold_len mirrored
     *     pairs; m (fold 1) or n (fold 2) counts the first halves of the blocks;
     *     E = A B over k (first product), O = A1 B1 over k1 (second product) and
     *     C[j] = E + O, C[fold_len-1-j] = sign_b (E - O), sign_b = -1 where bit b
     *     of fold_sign is set (b = block index);
     *   fold 3 / 4: the k-th element of A (3) or B (4) is X[k] + s_e X[fold_len-1-k]
     *     with k < fold_len / 2 = k and s_e = mirror_sign * (-1)^((e / mirror_div) % 2).
     * ska / skb: element stride along k of A (not transposed) / B (transposed);
     * ska applies to every variant, skb to the folded ones. */
    int fold;
    const double* A1;
    const double* B1;
    long long lda1, ldb1;
    int k1;
    long long ska, skb;
    int fold_len;
    unsigned fold_sign;
    double mirror_sign;
    int mirror_div;
} pypde_gemm_desc;

/* Batched GEMM  C[e] = alpha_e * op(A[e]) @ op(B[e]) + beta * D[e]  with the
 * operand batch indexing above: one launch applies an operator stack to the
 * U, V, T fields of many RBC instances (per-instance time steps via
 * alpha_vec).  Same reference interfaces as pypde_dgemm. */
int pypde_dgemm_batched(const pypde_gemm_desc* d, void* stream);

/* Grouped launch: `count` (1..4) independent plain descriptors with the same
 * transposes and the same kernel variant in one kernel launch (the tile grids
 * are concatenated); every member computes exactly what pypde_dgemm_batched
 * computes for it.  For steps whose GEMMs are too small to fill the device
 * alone.  Folded descriptors are rejected.  Same reference interfaces as
 * pypde_dgemm. */
int pypde_dgemm_grouped(const pypde_gemm_desc* descs, int count, void* stream);

/* Sum of K-split GEMM partials onto an addend (the second half of a split-K
 * product; the first half is a batched GEMM whose batch runs over K slices):
 *   C[e] = (D[e] + sum_{s<S} W[e*S + s]) + X[e / xm]   (X term only when e % xm == xr)
 * for `entries` contiguous matrices of mn elements; D and X may be NULL.  The
 * partials are summed in slice order, so the result does not depend on the
 * launch order.  Serves the fused step's right-hand-side assembly (the RHS of
 * update_U / update_V / update_T, navier/rbc2d.py:213-260, with the buoyancy
 * term of update_V as the X addend of the V entries). */
int pypde_splitk_reduce(const double* W, int S, const double* D, const double* X, int xm, int xr,
                        double* C, long long mn, int entries, void* stream);

/* ---- line operators (one lane per line, LDS-staged panels) ------------ */
/* Galerkin -> Chebyshev stencil u = S v, S(i,i)=1, S(i+2,i)=s[i]; len_out = len_in+2.
 * Replaces GalerkinChebyshev.to_chebyshev (pypde/bases/chebyshev.py:287-293). */
int pypde_stencil_apply(const pypde_lines* g, const double* s, void* stream);

/* Chebyshev -> Galerkin v = (S^T S)^-1 S^T u via the offset-2 tridiagonal solve;
 * a,b,c are the -2,0,+2 diagonals of S^T S, w/den the plan-time sweep factors.
 * Replaces GalerkinChebyshev.from_chebyshev/_solve_stencil_inv
 * (pypde/bases/chebyshev.py:295-337) -> tdma.solve_tdma_2d (tdma.f90:56-107). */
int pypde_stencil_inverse(const pypde_lines* g, const double* s, const double* a,
                          const double* b, const double* w, const double* den,
                          void* stream);

/* Chebyshev-coefficient derivative of the given order, optionally preceded by
 * the stencil (s != NULL, Galerkin input), result divided by scale_pow.
 * Replaces Chebyshev.derivative -> differentiate_cheby.diff_2d
 * (pypde/bases/chebyshev.py:117-131, differentiate_cheby.f90:28-53) and the
 * rescale in grad (pypde/field_operations.py:40-45). */
int pypde_cheb_deriv(const pypde_lines* g, const double* s, int order, double scale_pow,
                     void* stream);

/* Banded matrix-vector product along an axis, y = A x with A given by ndiag
 * diagonals (offsets ascending, vals[d*len_out + r] = A(r, r+off[d])), summed in
 * ascending column order.  Replaces PlanRHS with a sparse operator
 * (pypde/solver/plans.py:54-74, matrix.py:41-56). */
int pypde_banded_apply(const pypde_lines* g, int ndiag, const int* offsets_host,
                       const double* vals, void* stream);

/* Pre-factored four-diagonal (-2,0,2,4) solve, in place (src == dst allowed).
 * Tables l,d,u1,u2 per batch entry at offset b*table_bs (0 = shared).
 * Replaces Plan_fdma.solve -> fdma.solve_fdma_2d (plans.py:184-224, fdma.f90:40-98). */
int pypde_fdma_solve(const pypde_lines* g, const double* l, const double* d,
                     const double* u1, const double* u2, long long table_bs,
                     void* stream);

/* Transposed four-diagonal solve x = A^-T b with the same LU tables (adjoint).
 * No reference counterpart: used by the discrete adjoint of the Helmholtz step. */
int pypde_fdma_solve_t(const pypde_lines* g, const double* l, const double* d,
                       const double* u1, const double* u2, long long table_bs,
                       void* stream);

/* Poisson eigen-split solve: line j solves (A + lam_j C) x_j = b_j with the
 * factor tables of line j at  tab + j*4*len  (d | u1 | u2 | l, each len long);
 * lines with sing[j] != 0 solve the trailing (len-1) block and set x_0 = 0.
 * Replaces Plan_Poisson.solve_fortran -> fdma.solve_fdma_type2
 * (plans.py:310-314, fdma.f90:146-195); the per-call dense formation and
 * factorisation of the reference is done once at plan time. */
int pypde_poisson_solve(const pypde_lines* g, const double* tab, const int* sing,
                        int transpose, void* stream);

/* Upper two-band (0,+2) back substitution, in place.
 * Replaces Plan_twodma.solve -> twodma.solve_twodma_2d (plans.py:137-181, twodma.f90:28-62). */
int pypde_twodma_solve(const pypde_lines* g, const double* d, const double* u, void* stream);

/* Pentadiagonal (-4,-2,0,2,4) solve with PDMA_LU tables a,b,d,e,f, in place.
 * Replaces PDMA.solve (pypde/solver/linalg/pdma.py:74-93). */
int pypde_pdma_solve(const pypde_lines* g, const double* a, const double* b, const double* d,
                     const double* e, const double* f, void* stream);

/* Chebyshev transform along an axis as an LDS-staged cosine sum on the FP64
 * vector units, the alternative to the dense MFMA contraction (pypde_dgemm with
 * the transform matrix; the autotuner picks one per shape):
 *   dst[o] = a[o] * sum_i cos(pi o i / P) * b[i] * src[i],  o <= P (len_out = P + 1),
 *   i < len_in <= P + 1,
 * with ctab[m] = cos(pi m / P) for m < 2P.  a = 1, b[i] = (-1)^i is the backward
 * transform of chebyshev.py:78-83; a[o] = w_o (-1)^o / (2P), b = (1, 2, .., 2, 1)
 * the forward transform with the mass weights of chebyshev.py:67-76,143-149.
 * Output pairs (o, P - o) come from one pass over even and odd i.  Not in place.
 * variant: 1, 2, 3 (0 = 2) choose the register tiling (output pairs per wave,
 * lines per lane). */
int pypde_cheb_dct(const pypde_lines* g, int P, const double* a, const double* b,
                   const double* ctab, int variant, void* stream);

/* ---- register-resident recurrences (the hot path) ---------------------- */
/* Banded line solve in one normalised form over positions i of each line:
 *   pre-stencil (pre != NULL)  y_i = u_i + pre_i u_{i+2}     (len_in = len_out + 2)
 *   forward     (LM  != NULL)  y_i = F_i y_i + LM_i y_{i-2}   ascending  (F NULL: 1)
 *   backward                   z_i = R_i y_i + MU_i z_{i+2} + NU_i z_{i+4}  descending
 *                              (R NULL: 1, NU NULL: 0)
 * Tables are zero-padded so that chain starts/ends are encoded (LM_i = 0 for
 * i < 2, MU_i = 0 for i >= n-2, NU_i = 0 for i >= n-4) and hold reciprocal
 * pivots, so the kernel does not divide.  Tables are per batch entry at
 * entry_stride (0 = shared) or per line: per_line = 1 lays them out
 * [n][nlines] (contiguous slots LM | R | MU | NU), per_line = 2 line-major,
 * [nlines][n] per table.
 * variant: 1 = block form (64 lines per block, the line split into chunks over
 * up to 16 waves; for many lines), 2 = lane-chunk form (one wave per line, a
 * chunk of n/64 positions per lane, carries combined by wave scans; for few
 * lines), 3 = block form with 32 lines per block (half-waves hold neighbouring
 * chunks; half the LDS, so two blocks share a CU), 4 = variant 3 with 32
 * positions per chunk at every length (half the chunks, so a shorter carry
 * chain), 0 = chosen by the number of lines (lane-chunk up to 2048 lines,
 * else variant 1).
 * per_line = 2 implies the lane-chunk form, per_line = 1 a block form.
 * Carries: the four-diagonal Helmholtz sweeps (Plan_fdma.solve, plans.py:184-224
 * -> fdma.f90:40-98), the Poisson line solves (Plan_Poisson.solve_fortran,
 * plans.py:310-314 -> fdma.f90:146-195, factors formed once at plan time), the
 * Galerkin projection (chebyshev.py:295-337 -> tdma.f90:56-107) and the
 * two-band solve (plans.py:137-181 -> twodma.f90:28-62).  Results agree with
 * the reference to rounding (about 1e-15 relative); the bit-exact
 * restatements above remain available. */
typedef struct pypde_recurrence {
    const double* F;
    const double* LM;
    const double* R;
    const double* MU;
    const double* NU;
    const double* pre;
    long long entry_stride;
    int per_line;
    int variant;
} pypde_recurrence;

int pypde_recurrence_solve(const pypde_lines* g, const pypde_recurrence* r, void* stream);

/* ---- pointwise --------------------------------------------------------- */
/* y = alpha*x + beta*y over n elements (beta == 0: y is only written, as in
 * BLAS, so it may hold anything on entry). */
int pypde_axpby(long long n, double alpha, const double* x, double beta, double* y,
                void* stream);

/* dst[b][j*dst_ld + i] = src[b][i*src_ld + j] for i < rows, j < cols, b < batch
 * (batch strides src_bs, dst_bs in elements; not in place).  The reference
 * transposes by numpy's swapaxes (pypde/solver/plans.py:40-51); here it lets
 * the solves along rows run as solves along columns. */
int pypde_transpose(const double* src, long long src_ld, long long src_bs, double* dst,
                    long long dst_ld, long long dst_bs, int rows, int cols, int batch,
                    void* stream);

/* ---- fused RBC2D step helpers ------------------------------------------ */
/* Dealiased convective products of the U, V, T fields, with the lifted
 * boundary term uz*dTbcdz1 added to the temperature product.  Y is
 * [batch, 3, 2*L0, 2*L1] (value, d/dy, d/dx blocks per field), P is
 * [batch, 3, L0, L1].  Replaces the products of conv_term / convective_term
 * (pypde/field_operations.py:124-129,160-164) and rbc2d.py:342. */
int pypde_rbc_products(const double* Y, const double* dtbc, double* P, int L0, int L1,
                       int batch, void* stream);

/* The same products from y-folded blocks (the half-size y-pass output): per
 * instance and field, Y holds six [L0, L1/2] blocks -- value even / odd
 * y-coefficient parts, d/dy even / odd, d/dx even / odd -- over the first half of
 * the y grid; the field at column l is E + O and at L1-1-l it is s (E - O), with
 * s = sign_value for values and d/dx and sign_dy for d/dy.  P receives per field
 * two [L0, L1/2] blocks, P(l) + P(L1-1-l) and P(l) - P(L1-1-l), the input of the
 * half-size forward transform along y.  L1 must be even.  Yold (may be NULL): the
 * previous RK3 stage in the same layout; ratio * (products of the current
 * gradients with its velocities) is added, the c-term of the three-stage scheme
 * (rbc2d.py:261-283, ratio = c / b).  Same reference interfaces as
 * pypde_rbc_products. */
int pypde_rbc_products_yfold(const double* Y, const double* Yold, const double* dtbc, double* P,
                             int L0, int L1, double sign_value, double sign_dy, double ratio,
                             int batch, void* stream);

/* Discrete adjoint of pypde_rbc_products: from the cotangents Pb
 * [batch, 3, L0, L1] of the three products and the forward blocks Y, write the
 * cotangents Yb [batch, 3, 2*L0, 2*L1] of the value, d/dy and d/dx blocks.
 * No reference counterpart: the adjoint of the convective step used by the
 * gradient of the time step (the reference's navier/rbc2d_adj.py:162-187 forms
 * a continuous adjoint instead). */
int pypde_rbc_products_vjp(const double* Y, const double* dtbc, const double* Pb, double* Yb,
                           int L0, int L1, int batch, void* stream);

/* Banded part of the Helmholtz right-hand side of the fused step, per instance
 * b and field f in (U, V, T), all on Galerkin coefficients [M0, M1]:
 *   R_f = cst_f + (B S)_x X_f (B S)_y^T                          old value
 *         + mdta_b (B D/s)_x pres B_y^T            (f = U)        pressure gradient
 *         + mdta_b B_x pres (B D/s)_y^T            (f = V)
 *         + dta_b (B S_T)_x X_T (B S)_y^T          (f = V)        buoyancy
 *         + fdiff_{b,f} [(B D2 S/s^2)_x X_f (B S)_y^T + (B S)_x X_f (B D2 S/s^2)_y^T]
 * (S_T, the temperature's x stencil, replaces S in the x operators of f = T).
 * tx holds the x operators as diagonals, diag[d][i] = A(i, i + off_d) with
 * zeros where the column falls outside, M0 entries each, in this order:
 *   B D/s (offsets 1,3,5), B (0,2,4), B S (-2,0,2,4), B S_T (-2,0,2,4),
 *   B D2 S/s^2 (0,2), B D2 S_T/s^2 (0,2)                        18 diagonals;
 * ty the y operators, M1 entries each: B (0,2,4), B D/s (1,3,5), B S (-2,0,2,4),
 *   B D2 S/s^2 (0,2)                                              12 diagonals.
 * cst (may be NULL) and R are [batch, 3, M0, M1]; X [batch, 3, M0, M1]; pres
 * [batch, N0, N1]; mdta, dta per instance; fdiff [batch, 3] or NULL (beta = 1).
 * Replaces the sparse PlanRHS applies of the U, V and T updates (the solve_rhs /
 * solve_old plans of templates/hholtz.py:70-93 applied in navier/rbc2d.py:213-283,
 * through pypde/solver/plans.py:54-74 and matrix.py:41-56). */
typedef struct pypde_rbc_rhs {
    const double* X;
    const double* pres;
    const double* cst;
    double* R;
    const double* mdta;
    const double* dta;
    const double* fdiff;
    const double* tx;
    const double* ty;
    int M0, M1, N0, N1, batch;
    int rows_per_thread; /* kernel form: 1 one output per thread; 4 or 8 (0: 4) a strip
                            of rows per thread sharing the y sums of its input rows;
                            16 tiles of 16 rows x 64 columns per block with the y sums
                            in LDS (the same sums in the same order: bit-identical) */
} pypde_rbc_rhs;

int pypde_rbc_rhs_banded(const pypde_rbc_rhs* args, void* stream);

/* pypde_rbc_products_yfold and pypde_rbc_rhs_banded (one output per thread:
 * rhs->rows_per_thread must be 1) in one launch.  The two are independent --
 * the products read the dealiased fields, the right-hand side the state -- so
 * merged they overlap on the device; each output is computed exactly as by
 * its own entry point.  Same reference interfaces as those two. */
int pypde_rbc_products_rhs(const double* Y, const double* Yold, const double* dtbc, double* P,
                           int L0, int L1, double sign_value, double sign_dy, double ratio,
                           int batch, const pypde_rbc_rhs* rhs, void* stream);

/* Discrete adjoint of the products with the c-term of the three-stage scheme:
 * the forward products are P_f = grad_f . (u + ratio u_old) (plus the lifted
 * boundary term with uz + ratio uz_old for T), the gradients of the current
 * stage times its velocities plus ratio = c / b times the previous stage's
 * velocities (rbc2d.py:261-283).  Y and Yold are the natural-layout blocks
 * [batch, 3, 2*L0, 2*L1] of the current and previous stage; Yb receives the
 * cotangents of Y's blocks and Yob those of Yold's (velocity values only).
 * No reference counterpart: adjoint of the three-stage step. */
int pypde_rbc_products_vjp_c(const double* Y, const double* Yold, const double* dtbc,
                             const double* Pb, double* Yb, double* Yob, double ratio, int L0,
                             int L1, int batch, void* stream);

/* y[b] -= (coef[b] * x[b]) * c2 over per elements of each batch entry:
 * the pressure update pres -= nu*div*beta of rbc2d.py:390-394, per instance. */
int pypde_scaled_sub(long long per, int batch, const double* coef, double c2,
                     const double* x, double* y, void* stream);

/* p[e*stride] = value for e < n: the gauge P[0,0] = 0 of rbc2d.py:384-385. */
int pypde_set_entries(double* p, long long stride, int n, double value, void* stream);

/* ---- multi launches: one launch for many members (heterogeneous batches) -- */
/* A heterogeneous batch runs each operator of the step once per size class.  A
 * multi launch serves every class in one launch: the members' arguments are
 * packed on the host into a table (pypde_*_multi_pack, into table_bytes of
 * host memory, at least pypde_multi_table_bytes(kind)), the caller copies the
 * table to device memory once, and each launch passes the device table with
 * the info the pack filled.  Every member computes exactly what its own
 * launch of the same kernel form computes.  No reference counterpart: the
 * reference steps one NavierStokes object at a time (navier/rbc2d.py:396-434). */
#define PYPDE_MULTI_GEMM 1
#define PYPDE_MULTI_RECURRENCE 2
#define PYPDE_MULTI_PRODUCTS 3
#define PYPDE_MULTI_RHS 4
#define PYPDE_MULTI_MAX 64 /* members per table */
typedef struct pypde_multi_info {
    int kind;       /* PYPDE_MULTI_* */
    int count;      /* members */
    int variant;    /* kernel form */
    int flags;      /* GEMM: bit 0 transA, bit 1 transB; recurrence: the table form */
    long long total;/* work items of the launch (tiles, lines, blocks) */
} pypde_multi_info;

long long pypde_multi_table_bytes(int kind);
/* GEMM members: plain descriptors with the same transposes, run with one
 * LDS-tiled variant (1..7, 24..42); or folded descriptors with one fold mode
 * and the same transposes, run with one folded variant (1..5). */
int pypde_gemm_multi_pack(const pypde_gemm_desc* descs, int count, int variant, void* table,
                          long long table_bytes, pypde_multi_info* info);
int pypde_dgemm_multi(const void* table, const pypde_multi_info* info, void* stream);
/* Lane-chunk recurrences (pypde_recurrence_solve, variant 2), the (LM, R, MU, NU)
 * form, one axis and one table layout (per_line 0 or 2) for all members; each
 * member has its own line geometry and tables. */
int pypde_recurrence_multi_pack(const pypde_lines* g, const pypde_recurrence* r, int count,
                                void* table, long long table_bytes, pypde_multi_info* info);
int pypde_recurrence_multi(const void* table, const pypde_multi_info* info, void* stream);
/* y-folded convective products (pypde_rbc_products_yfold), member i with the
 * arguments at index i of each array. */
int pypde_products_multi_pack(const double* const* Y, const double* const* Yold,
                              const double* const* dtbc, double* const* P, const int* L0,
                              const int* L1, const double* sign_value, const double* sign_dy,
                              const double* ratio, const int* batch, int count, void* table,
                              long long table_bytes, pypde_multi_info* info);
int pypde_products_multi(const void* table, const pypde_multi_info* info, void* stream);
/* Banded right-hand sides (pypde_rbc_rhs_banded), one argument struct per member;
 * one kernel form for all of them: every member's rows_per_thread must be the
 * same (0 or 1: one output per thread; 4 or 8: row strips; 16: LDS tiles). */
int pypde_rhs_multi_pack(const pypde_rbc_rhs* args, int count, void* table, long long table_bytes,
                         pypde_multi_info* info);
int pypde_rhs_multi(const void* table, const pypde_multi_info* info, void* stream);

/* ---- step plans: a whole time step behind one call ------------------- */
/* A step plan is the ordered list of the launches above that make up one fused
 * time step, recorded once (descriptors copied into host memory) and replayed
 * by one call.  Pointer arguments are recorded relative to `regions`: address
 * ranges declared at creation (state arrays, workspace).  A run may move each
 * region to another base -- a pointer inside region r is shifted by the move
 * of r -- so one plan steps any buffers of the declared sizes; pointers outside
 * every region (the plan-owned operator tables) are used as recorded.  Running
 * a plan allocates nothing; runs are asynchronous on the given stream, and
 * runs that overlap on the same buffers race.  Plans are immutable after
 * recording and may be run from several host threads.
 * No reference counterpart as an object: a plan carries the operation
 * sequence of NavierStokes.update (navier/rbc2d.py:396-434), which the
 * reference executes as Python calls. */
#define PYPDE_PLAN_MAX_REGIONS 64
typedef struct pypde_plan pypde_plan;

int pypde_plan_create(int nregions, const void* const* bases, const long long* bytes,
                      pypde_plan** out);
int pypde_plan_destroy(pypde_plan* p);
/* Recording: the arguments of pypde_dgemm_batched, pypde_dgemm_grouped,
 * pypde_recurrence_solve, pypde_rbc_products, pypde_rbc_products_yfold,
 * pypde_splitk_reduce, pypde_rbc_rhs_banded, pypde_rbc_products_rhs, pypde_fdma_solve
 * and pypde_poisson_solve, without the stream.  Arguments are checked when run. */
int pypde_plan_add_gemm(pypde_plan* p, const pypde_gemm_desc* d);
int pypde_plan_add_grouped(pypde_plan* p, const pypde_gemm_desc* descs, int count);
int pypde_plan_add_recurrence(pypde_plan* p, const pypde_lines* g, const pypde_recurrence* r);
int pypde_plan_add_products(pypde_plan* p, const double* Y, const double* dtbc, double* P,
                            int L0, int L1, int batch);
int pypde_plan_add_products_yfold(pypde_plan* p, const double* Y, const double* Yold,
                                  const double* dtbc, double* P, int L0, int L1,
                                  double sign_value, double sign_dy, double ratio, int batch);
int pypde_plan_add_splitk_reduce(pypde_plan* p, const double* W, int S, const double* D,
                                 const double* X, int xm, int xr, double* C, long long mn,
                                 int entries);
int pypde_plan_add_rhs_banded(pypde_plan* p, const pypde_rbc_rhs* args);
int pypde_plan_add_products_rhs(pypde_plan* p, const double* Y, const double* Yold,
                                const double* dtbc, double* P, int L0, int L1,
                                double sign_value, double sign_dy, double ratio, int batch,
                                const pypde_rbc_rhs* rhs);
int pypde_plan_add_fdma(pypde_plan* p, const pypde_lines* g, const double* l, const double* d,
                        const double* u1, const double* u2, long long table_bs);
int pypde_plan_add_poisson(pypde_plan* p, const pypde_lines* g, const double* tab,
                           const int* sing, int transpose);
/* A multi launch (the device table stays where it is; its members' pointers are
 * not moved by a run). */
int pypde_plan_add_multi(pypde_plan* p, const void* table, const pypde_multi_info* info);
/* The pointwise launches of the step's adjoint: the arguments of pypde_axpby,
 * pypde_transpose, pypde_scaled_sub, pypde_set_entries, pypde_rbc_products_vjp
 * and pypde_rbc_products_vjp_c, without the stream. */
int pypde_plan_add_axpby(pypde_plan* p, long long n, double alpha, const double* x, double beta,
                         double* y);
int pypde_plan_add_transpose(pypde_plan* p, const double* src, long long src_ld, long long src_bs,
                             double* dst, long long dst_ld, long long dst_bs, int rows, int cols,
                             int batch);
int pypde_plan_add_scaled_sub(pypde_plan* p, long long per, int batch, const double* coef,
                              double c2, const double* x, double* y);
int pypde_plan_add_set_entries(pypde_plan* p, double* ptr, long long stride, int n, double value);
int pypde_plan_add_products_vjp(pypde_plan* p, const double* Y, const double* dtbc,
                                const double* Pb, double* Yb, int L0, int L1, int batch);
int pypde_plan_add_products_vjp_c(pypde_plan* p, const double* Y, const double* Yold,
                                  const double* dtbc, const double* Pb, double* Yb, double* Yob,
                                  double ratio, int L0, int L1, int batch);
/* Number of recorded launches (-1 for a null plan); bytes of a region (-1 if none). */
int pypde_plan_size(const pypde_plan* p);
long long pypde_plan_region_bytes(const pypde_plan* p, int region);
/* Run the recorded launches n_steps times; bases[r] (NULL entry or NULL array:
 * as declared) is region r's base for this run. */
int pypde_plan_run(const pypde_plan* p, const void* const* bases, int n_steps, void* stream);

/* n_steps IMEX time steps of B RBC instances of one shape (one instance: the
 * drop-in NavierStokes.update; B instances with per-instance Ra, Pr, dt: the
 * batched form), by a step plan whose regions are, in order,
 *   0  X     [B, 3, M0, M1]  Galerkin coefficients of U, V, T
 *   1  pres  [B, N0, N1]     Chebyshev coefficients of the pressure
 *   2  P     [B, M0, M1]     pressure-correction potential
 *   3  workspace of pypde_plan_region_bytes(plan, 3) bytes (any contents).
 * In place on X, pres and P.  Replaces NavierStokes.update / iterate
 * (navier/rbc2d.py:396-434, pypde/solver/integrator.py:14-41). */
int pypde_rbc_step(const pypde_plan* plan, double* X, double* pres, double* P, void* workspace,
                   int n_steps, void* stream);

/* n_steps of a heterogeneous batch (instances of several shapes, each size
 * class with its own state and workspace) by a plan of multi launches, one per
 * operator of the step for all classes, recorded on the classes' buffers:
 * runs them in place there.  Same reference interface as pypde_rbc_step. */
int pypde_rbc_step_batched(const pypde_plan* plan, int n_steps, void* stream);

/* The vector-Jacobian product of one time step (the discrete adjoint: the
 * step's launches reversed, with transposed operators), by a plan whose
 * regions are, in order,
 *   0  X     [B, 3, M0, M1]  the step's input state (read only)
 *   1  pres  [B, N0, N1]     its pressure (read only)
 *   2  Xb    [B, 3, M0, M1]  in: cotangent of the step's output X; out: of its input X
 *   3  pb    [B, N0, N1]     in: cotangent of the output pres; out: of the input pres
 *   4 ... workspace regions, carved one after the other from `workspace`
 *         (each start rounded up to 256 bytes; pypde_rbc_step_vjp_workspace_bytes
 *         gives the total; any contents).
 * Euler and the three-stage scheme (whose stage states the plan recomputes in
 * the workspace), any implicit fraction.  No reference counterpart: the
 * reference differentiates nothing (navier/rbc2d_adj.py:321-356 is an
 * adjoint-descent iteration of a continuous adjoint); this is the gradient
 * of objectives of NavierStokes.update trajectories (navier/rbc2d.py:396-434). */
long long pypde_rbc_step_vjp_workspace_bytes(const pypde_plan* plan);
int pypde_rbc_step_vjp(const pypde_plan* plan, const double* X, const double* pres, double* Xb,
                       double* pb, void* workspace, void* stream);

/* pypde_rbc_step without a caller workspace: the plan's workspace region is
 * moved to a device array the library reserves, PYPDE_SCRATCH_BYTES (48 MiB)
 * on every GPU it runs on, held until the process ends.  Calls no allocator
 * (no hipMalloc / hipFree, host or device).  Rejects a plan whose workspace
 * is larger (pypde_plan_region_bytes(plan, 3) > pypde_scratch_bytes()): a
 * single 256 x 256 instance fits.  All calls on one GPU share its array, so
 * calls that overlap in time on one GPU (two streams, two host threads) race;
 * calls in one stream do not.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * allocation). */
#define PYPDE_SCRATCH_BYTES (48ull << 20)
long long pypde_scratch_bytes(void);
int pypde_rbc_step_noalloc(const pypde_plan* plan, double* X, double* pres, double* P,
                           int n_steps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYPDE_AMD_H */
