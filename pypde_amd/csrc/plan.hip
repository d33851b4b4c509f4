// Step plans: the launches of one fused time step as a replayable list.
//
// A plan records the library's own launches (batched and grouped GEMMs,
// recurrences, the convective-product and split-K reduce kernels) in order,
// with their descriptors copied into host memory at recording time.  Every
// pointer argument is classified against the plan's regions -- address ranges
// declared at creation (the state arrays, the workspace) -- so that a run can
// move each region to another base: a pointer inside region r is shifted by
// (new base of r - declared base of r); pointers outside every region
// (operator tables) are used as recorded.  Running a plan allocates nothing.
#include "common.h"
#include "../../include/pypde_amd.h"

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

namespace pypde {

enum PlanOpKind { OP_GEMM, OP_GROUPED, OP_RECURRENCE, OP_PRODUCTS, OP_PRODUCTS_YFOLD, OP_REDUCE,
                  OP_RHS_BANDED, OP_FDMA, OP_POISSON, OP_MULTI, OP_AXPBY, OP_TRANSPOSE,
                  OP_SCALED_SUB, OP_SET_ENTRIES, OP_PRODUCTS_VJP, OP_PRODUCTS_RHS };

struct PlanOp {
    PlanOpKind kind;
    std::vector<pypde_gemm_desc> descs;  // OP_GEMM (one) / OP_GROUPED (count)
    pypde_lines lines;                   // OP_RECURRENCE
    pypde_recurrence rec;
    pypde_rbc_rhs rhs;                   // OP_RHS_BANDED
    const double* tab[4];                // OP_FDMA: l, d, u1, u2; OP_POISSON: tab, (sing)
    const int* sing;
    long long table_bs;
    int transpose;
    const void* multi_table;             // OP_MULTI: device argument table and its info
    pypde_multi_info multi;
    // products: Y, Yold (yfold only), dtbc, P; reduce: W, D, X, C;
    // products VJP: Y, Yold, dtbc, Pb, Yb, Yob; axpby / scaled_sub: x, y (coef in p2);
    // transpose: src, dst; set_entries: p3
    const double* p0;
    const double* p1;
    const double* p2;
    double* p3;
    double* p4;
    double* p5;
    int L0, L1, batch;
    double sign_value, sign_dy, ratio;
    int S, xm, xr, entries;
    long long mn;
    long long ld0, ld1, bs0, bs1;        // transpose strides
    double alpha, beta;                  // axpby; scaled_sub c2 in alpha; set_entries value
};

}  // namespace pypde

struct pypde_plan {
    std::vector<uintptr_t> base;
    std::vector<long long> bytes;
    std::vector<pypde::PlanOp> ops;
};

namespace pypde {

// Shift p by the move of the region that contains it (none: unchanged).
template <typename T>
static inline T* relocate(const pypde_plan* pl, const intptr_t* delta, T* p)
{
    if (!p) return p;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    for (size_t r = 0; r < pl->base.size(); ++r) {
        if (a >= pl->base[r] && a < pl->base[r] + (uintptr_t)pl->bytes[r])
            return reinterpret_cast<T*>(a + delta[r]);
    }
    return p;
}

static void relocate_desc(const pypde_plan* pl, const intptr_t* dl, pypde_gemm_desc& d)
{
    d.A = relocate(pl, dl, d.A);
    d.B = relocate(pl, dl, d.B);
    d.D = relocate(pl, dl, d.D);
    d.C = relocate(pl, dl, d.C);
    d.alpha_vec = relocate(pl, dl, d.alpha_vec);
    d.E = relocate(pl, dl, d.E);
    d.gamma_vec = relocate(pl, dl, d.gamma_vec);
    d.A1 = relocate(pl, dl, d.A1);
    d.B1 = relocate(pl, dl, d.B1);
}

static int run_op(const pypde_plan* pl, const intptr_t* dl, bool moved,
                  const PlanOp& op, void* stream)
{
    switch (op.kind) {
    case OP_GEMM:
    case OP_GROUPED: {
        if (!moved) {
            return op.kind == OP_GEMM ? pypde_dgemm_batched(op.descs.data(), stream)
                                      : pypde_dgemm_grouped(op.descs.data(), (int)op.descs.size(),
                                                            stream);
        }
        pypde_gemm_desc d[4];
        const int n = (int)op.descs.size();
        for (int i = 0; i < n; ++i) {
            d[i] = op.descs[i];
            relocate_desc(pl, dl, d[i]);
        }
        return op.kind == OP_GEMM ? pypde_dgemm_batched(d, stream)
                                  : pypde_dgemm_grouped(d, n, stream);
    }
    case OP_RECURRENCE: {
        if (!moved) return pypde_recurrence_solve(&op.lines, &op.rec, stream);
        pypde_lines g = op.lines;
        pypde_recurrence r = op.rec;
        g.src = relocate(pl, dl, g.src);
        g.dst = relocate(pl, dl, g.dst);
        r.F = relocate(pl, dl, r.F);
        r.LM = relocate(pl, dl, r.LM);
        r.R = relocate(pl, dl, r.R);
        r.MU = relocate(pl, dl, r.MU);
        r.NU = relocate(pl, dl, r.NU);
        r.pre = relocate(pl, dl, r.pre);
        return pypde_recurrence_solve(&g, &r, stream);
    }
    case OP_PRODUCTS:
        return pypde_rbc_products(relocate(pl, dl, op.p0), relocate(pl, dl, op.p2),
                                  relocate(pl, dl, op.p3), op.L0, op.L1, op.batch, stream);
    case OP_PRODUCTS_YFOLD:
        return pypde_rbc_products_yfold(relocate(pl, dl, op.p0), relocate(pl, dl, op.p1),
                                        relocate(pl, dl, op.p2), relocate(pl, dl, op.p3), op.L0,
                                        op.L1, op.sign_value, op.sign_dy, op.ratio, op.batch,
                                        stream);
    case OP_RHS_BANDED:
    case OP_PRODUCTS_RHS: {
        pypde_rbc_rhs a = op.rhs;
        a.X = relocate(pl, dl, a.X);
        a.pres = relocate(pl, dl, a.pres);
        a.cst = relocate(pl, dl, a.cst);
        a.R = relocate(pl, dl, a.R);
        a.mdta = relocate(pl, dl, a.mdta);
        a.dta = relocate(pl, dl, a.dta);
        a.fdiff = relocate(pl, dl, a.fdiff);
        a.tx = relocate(pl, dl, a.tx);
        a.ty = relocate(pl, dl, a.ty);
        if (op.kind == OP_RHS_BANDED) return pypde_rbc_rhs_banded(&a, stream);
        return pypde_rbc_products_rhs(relocate(pl, dl, op.p0), relocate(pl, dl, op.p1),
                                      relocate(pl, dl, op.p2), relocate(pl, dl, op.p3), op.L0,
                                      op.L1, op.sign_value, op.sign_dy, op.ratio, op.batch, &a,
                                      stream);
    }
    case OP_FDMA: {
        pypde_lines g = op.lines;
        g.src = relocate(pl, dl, g.src);
        g.dst = relocate(pl, dl, g.dst);
        return pypde_fdma_solve(&g, relocate(pl, dl, op.tab[0]), relocate(pl, dl, op.tab[1]),
                                relocate(pl, dl, op.tab[2]), relocate(pl, dl, op.tab[3]),
                                op.table_bs, stream);
    }
    case OP_POISSON: {
        pypde_lines g = op.lines;
        g.src = relocate(pl, dl, g.src);
        g.dst = relocate(pl, dl, g.dst);
        return pypde_poisson_solve(&g, relocate(pl, dl, op.tab[0]), relocate(pl, dl, op.sing),
                                   op.transpose, stream);
    }
    case OP_MULTI:
        // the members' pointers live in the device table: used as recorded
        switch (op.multi.kind) {
        case PYPDE_MULTI_GEMM: return pypde_dgemm_multi(op.multi_table, &op.multi, stream);
        case PYPDE_MULTI_RECURRENCE: return pypde_recurrence_multi(op.multi_table, &op.multi, stream);
        case PYPDE_MULTI_PRODUCTS: return pypde_products_multi(op.multi_table, &op.multi, stream);
        case PYPDE_MULTI_RHS: return pypde_rhs_multi(op.multi_table, &op.multi, stream);
        default: break;
        }
        set_error("step plan: unknown multi launch kind %d", op.multi.kind);
        return ERR_ARG;
    case OP_AXPBY:
        return pypde_axpby(op.mn, op.alpha, relocate(pl, dl, op.p0), op.beta,
                           relocate(pl, dl, op.p3), stream);
    case OP_TRANSPOSE:
        return pypde_transpose(relocate(pl, dl, op.p0), op.ld0, op.bs0, relocate(pl, dl, op.p3),
                               op.ld1, op.bs1, op.L0, op.L1, op.batch, stream);
    case OP_SCALED_SUB:
        return pypde_scaled_sub(op.mn, op.batch, relocate(pl, dl, op.p2), op.alpha,
                                relocate(pl, dl, op.p0), relocate(pl, dl, op.p3), stream);
    case OP_SET_ENTRIES:
        return pypde_set_entries(relocate(pl, dl, op.p3), op.mn, op.entries, op.alpha, stream);
    case OP_PRODUCTS_VJP:
        if (!op.p1)
            return pypde_rbc_products_vjp(relocate(pl, dl, op.p0), relocate(pl, dl, op.p2),
                                          relocate(pl, dl, op.p3), relocate(pl, dl, op.p4), op.L0,
                                          op.L1, op.batch, stream);
        return pypde_rbc_products_vjp_c(relocate(pl, dl, op.p0), relocate(pl, dl, op.p1),
                                        relocate(pl, dl, op.p2), relocate(pl, dl, op.p3),
                                        relocate(pl, dl, op.p4), relocate(pl, dl, op.p5), op.ratio,
                                        op.L0, op.L1, op.batch, stream);
    case OP_REDUCE:
        return pypde_splitk_reduce(relocate(pl, dl, op.p0), op.S, relocate(pl, dl, op.p1),
                                   relocate(pl, dl, op.p2), op.xm, op.xr, relocate(pl, dl, op.p3),
                                   op.mn, op.entries, stream);
    }
    set_error("step plan: unknown operation");
    return ERR_ARG;
}

static int add_op(pypde_plan* p, PlanOp&& op)
{
    try {
        p->ops.push_back(std::move(op));
    } catch (const std::bad_alloc&) {
        set_error("step plan: out of host memory");
        return ERR_ARG;
    }
    return OK;
}

}  // namespace pypde

using namespace pypde;

extern "C" int pypde_plan_create(int nregions, const void* const* bases, const long long* bytes,
                                 pypde_plan** out)
{
    PYPDE_REQUIRE(out, "null output handle");
    *out = nullptr;
    PYPDE_REQUIRE(nregions >= 0 && nregions <= PYPDE_PLAN_MAX_REGIONS,
                  "regions must be 0..%d (got %d)", PYPDE_PLAN_MAX_REGIONS, nregions);
    PYPDE_REQUIRE(nregions == 0 || (bases && bytes), "null region arrays");
    for (int r = 0; r < nregions; ++r) {
        PYPDE_REQUIRE(bases[r] && bytes[r] > 0, "region %d: null base or empty range", r);
        for (int q = 0; q < r; ++q) {
            const uintptr_t a0 = (uintptr_t)bases[q], a1 = a0 + (uintptr_t)bytes[q];
            const uintptr_t b0 = (uintptr_t)bases[r], b1 = b0 + (uintptr_t)bytes[r];
            PYPDE_REQUIRE(b1 <= a0 || a1 <= b0, "regions %d and %d overlap", q, r);
        }
    }
    pypde_plan* p = new (std::nothrow) pypde_plan();
    PYPDE_REQUIRE(p, "step plan: out of host memory");
    for (int r = 0; r < nregions; ++r) {
        p->base.push_back((uintptr_t)bases[r]);
        p->bytes.push_back(bytes[r]);
    }
    *out = p;
    return OK;
}

extern "C" int pypde_plan_destroy(pypde_plan* p)
{
    delete p;
    return OK;
}

extern "C" int pypde_plan_add_gemm(pypde_plan* p, const pypde_gemm_desc* d)
{
    PYPDE_REQUIRE(p && d, "null argument");
    PlanOp op{};
    op.kind = OP_GEMM;
    op.descs.assign(d, d + 1);
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_grouped(pypde_plan* p, const pypde_gemm_desc* descs, int count)
{
    PYPDE_REQUIRE(p && descs, "null argument");
    PYPDE_REQUIRE(count >= 1 && count <= 4, "grouped launch of 1..4 descriptors (got %d)", count);
    PlanOp op{};
    op.kind = OP_GROUPED;
    op.descs.assign(descs, descs + count);
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_recurrence(pypde_plan* p, const pypde_lines* g,
                                         const pypde_recurrence* r)
{
    PYPDE_REQUIRE(p && g && r, "null argument");
    PlanOp op{};
    op.kind = OP_RECURRENCE;
    op.lines = *g;
    op.rec = *r;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_products(pypde_plan* p, const double* Y, const double* dtbc, double* P,
                                       int L0, int L1, int batch)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_PRODUCTS;
    op.p0 = Y;
    op.p2 = dtbc;
    op.p3 = P;
    op.L0 = L0;
    op.L1 = L1;
    op.batch = batch;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_products_yfold(pypde_plan* p, const double* Y, const double* Yold,
                                             const double* dtbc, double* P, int L0, int L1,
                                             double sign_value, double sign_dy, double ratio,
                                             int batch)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_PRODUCTS_YFOLD;
    op.p0 = Y;
    op.p1 = Yold;
    op.p2 = dtbc;
    op.p3 = P;
    op.L0 = L0;
    op.L1 = L1;
    op.sign_value = sign_value;
    op.sign_dy = sign_dy;
    op.ratio = ratio;
    op.batch = batch;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_splitk_reduce(pypde_plan* p, const double* W, int S, const double* D,
                                            const double* X, int xm, int xr, double* C,
                                            long long mn, int entries)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_REDUCE;
    op.p0 = W;
    op.p1 = D;
    op.p2 = X;
    op.p3 = C;
    op.S = S;
    op.xm = xm;
    op.xr = xr;
    op.mn = mn;
    op.entries = entries;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_rhs_banded(pypde_plan* p, const pypde_rbc_rhs* args)
{
    PYPDE_REQUIRE(p && args, "null argument");
    PlanOp op{};
    op.kind = OP_RHS_BANDED;
    op.rhs = *args;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_products_rhs(pypde_plan* p, const double* Y, const double* Yold,
                                           const double* dtbc, double* P, int L0, int L1,
                                           double sign_value, double sign_dy, double ratio,
                                           int batch, const pypde_rbc_rhs* rhs)
{
    PYPDE_REQUIRE(p && rhs, "null argument");
    PlanOp op{};
    op.kind = OP_PRODUCTS_RHS;
    op.p0 = Y;
    op.p1 = Yold;
    op.p2 = dtbc;
    op.p3 = P;
    op.L0 = L0;
    op.L1 = L1;
    op.sign_value = sign_value;
    op.sign_dy = sign_dy;
    op.ratio = ratio;
    op.batch = batch;
    op.rhs = *rhs;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_fdma(pypde_plan* p, const pypde_lines* g, const double* l,
                                   const double* d, const double* u1, const double* u2,
                                   long long table_bs)
{
    PYPDE_REQUIRE(p && g, "null argument");
    PlanOp op{};
    op.kind = OP_FDMA;
    op.lines = *g;
    op.tab[0] = l;
    op.tab[1] = d;
    op.tab[2] = u1;
    op.tab[3] = u2;
    op.table_bs = table_bs;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_poisson(pypde_plan* p, const pypde_lines* g, const double* tab,
                                      const int* sing, int transpose)
{
    PYPDE_REQUIRE(p && g, "null argument");
    PlanOp op{};
    op.kind = OP_POISSON;
    op.lines = *g;
    op.tab[0] = tab;
    op.sing = sing;
    op.transpose = transpose;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_multi(pypde_plan* p, const void* table, const pypde_multi_info* info)
{
    PYPDE_REQUIRE(p && table && info, "null argument");
    PlanOp op{};
    op.kind = OP_MULTI;
    op.multi_table = table;
    op.multi = *info;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_axpby(pypde_plan* p, long long n, double alpha, const double* x,
                                    double beta, double* y)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_AXPBY;
    op.mn = n;
    op.alpha = alpha;
    op.beta = beta;
    op.p0 = x;
    op.p3 = y;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_transpose(pypde_plan* p, const double* src, long long src_ld,
                                        long long src_bs, double* dst, long long dst_ld,
                                        long long dst_bs, int rows, int cols, int batch)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_TRANSPOSE;
    op.p0 = src;
    op.p3 = dst;
    op.ld0 = src_ld;
    op.bs0 = src_bs;
    op.ld1 = dst_ld;
    op.bs1 = dst_bs;
    op.L0 = rows;
    op.L1 = cols;
    op.batch = batch;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_scaled_sub(pypde_plan* p, long long per, int batch,
                                         const double* coef, double c2, const double* x, double* y)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_SCALED_SUB;
    op.mn = per;
    op.batch = batch;
    op.p2 = coef;
    op.alpha = c2;
    op.p0 = x;
    op.p3 = y;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_set_entries(pypde_plan* p, double* ptr, long long stride, int n,
                                          double value)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_SET_ENTRIES;
    op.p3 = ptr;
    op.mn = stride;
    op.entries = n;
    op.alpha = value;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_products_vjp(pypde_plan* p, const double* Y, const double* dtbc,
                                           const double* Pb, double* Yb, int L0, int L1, int batch)
{
    PYPDE_REQUIRE(p, "null plan");
    PlanOp op{};
    op.kind = OP_PRODUCTS_VJP;
    op.p0 = Y;
    op.p1 = nullptr;
    op.p2 = dtbc;
    op.p3 = const_cast<double*>(Pb);
    op.p4 = Yb;
    op.L0 = L0;
    op.L1 = L1;
    op.batch = batch;
    return add_op(p, std::move(op));
}

extern "C" int pypde_plan_add_products_vjp_c(pypde_plan* p, const double* Y, const double* Yold,
                                             const double* dtbc, const double* Pb, double* Yb,
                                             double* Yob, double ratio, int L0, int L1, int batch)
{
    PYPDE_REQUIRE(p, "null plan");
    PYPDE_REQUIRE(Yold, "the c-term form needs the previous stage's fields");
    PlanOp op{};
    op.kind = OP_PRODUCTS_VJP;
    op.p0 = Y;
    op.p1 = Yold;
    op.p2 = dtbc;
    op.p3 = const_cast<double*>(Pb);
    op.p4 = Yb;
    op.p5 = Yob;
    op.ratio = ratio;
    op.L0 = L0;
    op.L1 = L1;
    op.batch = batch;
    return add_op(p, std::move(op));
}

extern "C" int pypde_rbc_step_batched(const pypde_plan* p, int n_steps, void* stream)
{
    PYPDE_REQUIRE(p, "null plan");
    return pypde_plan_run(p, nullptr, n_steps, stream);
}

extern "C" int pypde_plan_size(const pypde_plan* p)
{
    return p ? (int)p->ops.size() : -1;
}

extern "C" long long pypde_plan_region_bytes(const pypde_plan* p, int region)
{
    if (!p || region < 0 || region >= (int)p->base.size()) return -1;
    return p->bytes[region];
}

extern "C" int pypde_plan_run(const pypde_plan* p, const void* const* bases, int n_steps,
                              void* stream)
{
    PYPDE_REQUIRE(p, "null plan");
    PYPDE_REQUIRE(n_steps >= 0, "negative step count");
    // per-region moves on the stack: a run allocates nothing, host or device
    const size_t nr = p->base.size();
    intptr_t delta[PYPDE_PLAN_MAX_REGIONS] = {};
    bool moved = false;
    if (bases) {
        for (size_t r = 0; r < nr; ++r) {
            if (!bases[r]) continue;
            delta[r] = (intptr_t)((uintptr_t)bases[r] - p->base[r]);
            moved = moved || delta[r] != 0;
        }
    }
    for (int s = 0; s < n_steps; ++s) {
        for (size_t i = 0; i < p->ops.size(); ++i) {
            const int rc = run_op(p, delta, moved, p->ops[i], stream);
            if (rc != OK) {
                // keep the failing launch's message, name its place in the plan
                char msg[768];
                snprintf(msg, sizeof(msg), "%s", last_error());
                set_error("step plan, step %d, launch %zu: %s", s, i, msg);
                return rc;
            }
        }
    }
    return OK;
}

// ---- the step without a caller workspace ---------------------------------
// Device memory that the library reserves for callers that cannot allocate:
// one array per GPU, held until the process ends.  Host code never uses its
// name as an address: the device address comes from hipGetSymbolAddress, once
// per device, with that device current.
static __device__ __align__(256) unsigned char scratch[PYPDE_SCRATCH_BYTES];

namespace pypde {
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::atomic<int> g_scratch_ready[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The current device's address of `scratch` (resolved on the first call on
// that device; later calls read the kept pointer).
static int device_scratch(void** out)
{
    int dev = -1;
    PYPDE_HIP(hipGetDevice(&dev));
    PYPDE_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d: at most %d devices are supported",
                  dev, kMaxDevices);
    if (!g_scratch_ready[dev].load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ready[dev].load(std::memory_order_relaxed)) {
            void* ptr = nullptr;
            PYPDE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            PYPDE_REQUIRE(ptr, "hipGetSymbolAddress returned no address for the scratch array");
            g_scratch_ptr[dev] = ptr;
            g_scratch_ready[dev].store(1, std::memory_order_release);
        }
    }
    *out = g_scratch_ptr[dev];
    return OK;
}
}  // namespace pypde

extern "C" long long pypde_scratch_bytes(void) { return (long long)PYPDE_SCRATCH_BYTES; }

extern "C" int pypde_rbc_step_noalloc(const pypde_plan* p, double* X, double* pres, double* P,
                                      int n_steps, void* stream)
{
    PYPDE_REQUIRE(p, "null plan");
    PYPDE_REQUIRE(p->base.size() == 4,
                  "an RBC step plan has the regions X, pres, P and workspace (got %zu regions)",
                  p->base.size());
    PYPDE_REQUIRE(p->bytes[3] <= (long long)PYPDE_SCRATCH_BYTES,
                  "the plan's workspace needs %lld bytes; the library's scratch array holds %lld "
                  "(use pypde_rbc_step with a caller workspace)",
                  p->bytes[3], (long long)PYPDE_SCRATCH_BYTES);
    PYPDE_REQUIRE(X && pres && P, "null state pointer");
    void* ws = nullptr;
    const int rc = device_scratch(&ws);
    if (rc != OK) return rc;
    // every workspace entry the step reads is written earlier in the same
    // step, so the contents left by an earlier call need no clearing
    const void* bases[4] = {X, pres, P, ws};
    return pypde_plan_run(p, bases, n_steps, stream);
}

namespace pypde {
constexpr int kVjpFixedRegions = 4;   // X, pres, Xb, pb
static long long align256(long long b) { return (b + 255) / 256 * 256; }
}  // namespace pypde

extern "C" long long pypde_rbc_step_vjp_workspace_bytes(const pypde_plan* p)
{
    if (!p || (int)p->base.size() <= kVjpFixedRegions) return -1;
    long long total = 0;
    for (size_t r = kVjpFixedRegions; r < p->base.size(); ++r) total += align256(p->bytes[r]);
    return total;
}

extern "C" int pypde_rbc_step_vjp(const pypde_plan* p, const double* X, const double* pres,
                                  double* Xb, double* pb, void* workspace, void* stream)
{
    PYPDE_REQUIRE(p, "null plan");
    PYPDE_REQUIRE((int)p->base.size() > kVjpFixedRegions,
                  "a step VJP plan has the regions X, pres, Xb, pb and at least one workspace "
                  "region (got %zu regions)", p->base.size());
    PYPDE_REQUIRE(X && pres && Xb && pb && workspace, "null state, cotangent or workspace pointer");
    const void* bases[PYPDE_PLAN_MAX_REGIONS] = {X, pres, Xb, pb};
    uintptr_t at = (uintptr_t)workspace;
    PYPDE_REQUIRE(at % 256 == 0, "the workspace must be 256-byte aligned");
    for (size_t r = kVjpFixedRegions; r < p->base.size(); ++r) {
        bases[r] = (const void*)at;
        at += (uintptr_t)align256(p->bytes[r]);
    }
    return pypde_plan_run(p, bases, 1, stream);
}

extern "C" int pypde_rbc_step(const pypde_plan* p, double* X, double* pres, double* P,
                              void* workspace, int n_steps, void* stream)
{
    PYPDE_REQUIRE(p, "null plan");
    PYPDE_REQUIRE(p->base.size() == 4,
                  "an RBC step plan has the regions X, pres, P and workspace (got %zu regions)",
                  p->base.size());
    PYPDE_REQUIRE(X && pres && P && workspace, "null state or workspace pointer");
    const void* bases[4] = {X, pres, P, workspace};
    return pypde_plan_run(p, bases, n_steps, stream);
}
