// C ABI of the GPU time loop: operators, media, runs, and the drop-in
// replacements of csr_matvec / abc_mpml / seismic_source.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "seisfem_gpu.h"
#include "seisfem_host.h"
#include "sfg_internal.h"

using namespace sfg;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;
static thread_local int g_device = 0;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(SFG_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));                \
    } while (0)

extern "C" const char *sfg_last_error(void) { return g_err.c_str(); }

extern "C" int sfg_set_device(int device)
{
    HIPCHK(hipSetDevice(device));
    g_device = device;
    return SFG_OK;
}

template <class T>
static int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(SFG_E_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    return SFG_OK;
}

template <class T>
static void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// drop-ins

extern "C" double seismic_source(double f0, double t0, double magnitide, double t)
{
    // Ricker wavelet (source_receiver/seismic_source.c:37-39); same operand order
    const double pi = 3.141592653589793238462643;
    const double a = (pi * f0 * (t - t0)) * (pi * f0 * (t - t0));
    return magnitide * (1 - 2 * a) * exp(-a);
}

namespace {
// Scratch device buffers for the host-pointer drop-ins, grown on demand.
struct DropinScratch {
    std::mutex mu;
    void *buf = nullptr;
    size_t cap = 0;
    int get(size_t bytes, void **out)
    {
        if (bytes > cap) {
            if (buf) (void)hipFree(buf);
            buf = nullptr;
            cap = 0;
            hipError_t e = hipMalloc(&buf, bytes);
            if (e != hipSuccess) return fail(SFG_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
            cap = bytes;
        }
        *out = buf;
        return SFG_OK;
    }
};
DropinScratch g_scratch;

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
} // namespace

static int spmv_host(int Bp_size, const int *Bp, const int *Bj, const double *Bx, const double *x,
                     double *y)
{
    const int n = Bp_size - 1;
    if (n <= 0 || !Bp || !Bj || !Bx || !x || !y) return fail(SFG_E_ARG, "csr_matvec: bad arguments");
    const int nnz = Bp[n];
    if (Bp[0] != 0 || nnz < 0) return fail(SFG_E_ARG, "csr_matvec: row pointers must start at 0");
    for (int i = 0; i < n; ++i)
        if (Bp[i + 1] < Bp[i]) return fail(SFG_E_ARG, "csr_matvec: row pointers decrease at row %d", i);
    for (int k = 0; k < nnz; ++k)
        if (Bj[k] < 0 || Bj[k] >= n) return fail(SFG_E_ARG, "csr_matvec: column %d out of range", Bj[k]);
    std::lock_guard<std::mutex> lk(g_scratch.mu);
    const size_t o_rp = 0, o_cj = align256(o_rp + sizeof(int) * (n + 1));
    const size_t o_a = align256(o_cj + sizeof(int) * std::max(nnz, 1));
    const size_t o_x = align256(o_a + sizeof(double) * std::max(nnz, 1));
    const size_t o_y = align256(o_x + sizeof(double) * n);
    const size_t total = o_y + sizeof(double) * n;
    void *base;
    int rc = g_scratch.get(total, &base);
    if (rc) return rc;
    char *b = (char *)base;
    HIPCHK(hipMemcpy(b + o_rp, Bp, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (nnz) {
        HIPCHK(hipMemcpy(b + o_cj, Bj, sizeof(int) * nnz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b + o_a, Bx, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(b + o_x, x, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(launch_spmv64(n, (const int *)(b + o_rp), (const int *)(b + o_cj), (const double *)(b + o_a),
                         (const double *)(b + o_x), (double *)(b + o_y), nullptr));
    HIPCHK(hipMemcpy(y, b + o_y, sizeof(double) * n, hipMemcpyDeviceToHost));
    return SFG_OK;
}

extern "C" void csr_matvec(int Bp_size, int *Bp, int *Bj, double *Bx, double *x, double *y)
{
    g_err.clear();
    (void)spmv_host(Bp_size, Bp, Bj, Bx, x, y);
}

static int mpml_host(int node_num, int element_num, int element_order, const int *element_node,
                     const double *X, const double *Y, int mpml_nx, int mpml_ny, double h,
                     double xmin, double xmax, double ymin, double ymax, double vp_max, int uxn,
                     int uxx, int uyn, int uyx, double *const out[6])
{
    (void)element_num;
    (void)element_order;
    (void)element_node;
    if (node_num <= 0 || !X || !Y) return fail(SFG_E_ARG, "abc_mpml: bad arguments");
    for (int k = 0; k < 6; ++k)
        if (!out[k]) return fail(SFG_E_ARG, "abc_mpml: null output %d", k);
    // scalars of pml/abc_mpml.c:50-61, evaluated on the host as the reference does
    MpmlGeo g;
    g.pxy = 0.15;
    g.pyx = 0.15;
    const double Rcoef = 0.0001;
    g.tx = h * mpml_nx;
    g.ty = h * mpml_ny;
    g.x0 = xmin + g.tx;
    g.x1 = xmax - g.tx;
    g.y0 = ymin + g.ty;
    g.y1 = ymax - g.ty;
    g.d0x = 3.0 * vp_max * log10(1.0 / Rcoef) / (2.0 * g.tx);
    g.d0y = 3.0 * vp_max * log10(1.0 / Rcoef) / (2.0 * g.ty);
    g.use_xmin = uxn;
    g.use_xmax = uxx;
    g.use_ymin = uyn;
    g.use_ymax = uyx;
    std::lock_guard<std::mutex> lk(g_scratch.mu);
    const size_t n = (size_t)node_num;
    void *base;
    int rc = g_scratch.get(sizeof(double) * n * 8, &base);
    if (rc) return rc;
    double *dX = (double *)base, *dY = dX + n, *dO = dY + n;
    HIPCHK(hipMemcpy(dX, X, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dY, Y, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(launch_mpml(node_num, dX, dY, g, dO, nullptr));
    for (int k = 0; k < 6; ++k)
        HIPCHK(hipMemcpy(out[k], dO + k * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    return SFG_OK;
}

extern "C" void abc_mpml(int node_num, int element_num, int element_order, int *element_node,
                         double **node_xy, int mpml_nx, int mpml_ny, double h, double xmin,
                         double xmax, double ymin, double ymax, double vp_max, int use_mpml_xmin,
                         int use_mpml_xmax, int use_mpml_ymin, int use_mpml_ymax, double *mpml_dx,
                         double *mpml_dy, double *mpml_dxx, double *mpml_dyy, double *mpml_dxx_pyx,
                         double *mpml_dyy_pxy)
{
    g_err.clear();
    double *const out[6] = {mpml_dx, mpml_dy, mpml_dxx, mpml_dyy, mpml_dxx_pyx, mpml_dyy_pxy};
    if (!node_xy) {
        fail(SFG_E_ARG, "abc_mpml: node_xy is null");
        return;
    }
    (void)mpml_host(node_num, element_num, element_order, element_node, node_xy[0], node_xy[1],
                    mpml_nx, mpml_ny, h, xmin, xmax, ymin, ymax, vp_max, use_mpml_xmin,
                    use_mpml_xmax, use_mpml_ymin, use_mpml_ymax, out);
}

// ---------------------------------------------------------------------------
// operators

struct sfg_operator {
    int n = 0, nnz = 0, device = 0;
    int *rp = nullptr, *ci = nullptr;
    std::vector<int> hrp, hci;  // host copy of the union pattern (block halos)
    double *v64 = nullptr; // [7][nnz]
    float *v32 = nullptr;  // [7][nnz]
    double *p64 = nullptr; // [nnz][8]: M, S1..S6, 0 (one wide load per nonzero)
    float *p32 = nullptr;  // [nnz][8]
};

static int check_csr(int n, const int *p, const int *j, int m)
{
    if (!p || !j) return fail(SFG_E_ARG, "matrix %d: null CSR array", m);
    if (p[0] != 0) return fail(SFG_E_ARG, "matrix %d: row pointers must start at 0", m);
    for (int i = 0; i < n; ++i) {
        if (p[i + 1] < p[i]) return fail(SFG_E_ARG, "matrix %d: row pointers decrease at row %d", m, i);
        for (int k = p[i]; k < p[i + 1]; ++k) {
            if (j[k] < 0 || j[k] >= n) return fail(SFG_E_ARG, "matrix %d: column out of range in row %d", m, i);
            if (k > p[i] && j[k] <= j[k - 1])
                return fail(SFG_E_ARG, "matrix %d: columns not strictly ascending in row %d", m, i);
        }
    }
    return SFG_OK;
}

extern "C" int sfg_operator_create(sfg_operator **out, int n, const int *const csr_p[7],
                                   const int *const csr_j[7], const double *const csr_x[7])
{
    g_err.clear();
    if (!out || n <= 0 || !csr_p || !csr_j || !csr_x) return fail(SFG_E_ARG, "operator: bad arguments");
    *out = nullptr;
    long long sum = 0;
    for (int m = 0; m < 7; ++m) {
        if (!csr_x[m]) return fail(SFG_E_ARG, "matrix %d: null values", m);
        int rc = check_csr(n, csr_p[m], csr_j[m], m);
        if (rc) return rc;
        sum += csr_p[m][n];
    }
    const int cap = (int)std::min<long long>(sum, 0x7fffffffLL);
    std::vector<int> up(n + 1), uj(cap);
    std::vector<double> ux((size_t)7 * cap);
    double *uxp[7];
    for (int m = 0; m < 7; ++m) uxp[m] = ux.data() + (size_t)m * cap;
    const int nnz = sfh_csr_union(n, 7, csr_p, csr_j, csr_x, cap, up.data(), uj.data(), uxp);
    if (nnz < 0) return fail(SFG_E_ARG, "operator: union pattern failed");
    sfg_operator *op = new sfg_operator;
    op->n = n;
    op->nnz = nnz;
    op->device = g_device;
    op->hrp = up;
    op->hci.assign(uj.begin(), uj.begin() + nnz);
    int rc = SFG_OK;
    if ((rc = dalloc(&op->rp, n + 1)) || (rc = dalloc(&op->ci, nnz)) ||
        (rc = dalloc(&op->v64, (size_t)7 * nnz)) || (rc = dalloc(&op->v32, (size_t)7 * nnz)) ||
        (rc = dalloc(&op->p64, (size_t)8 * nnz)) || (rc = dalloc(&op->p32, (size_t)8 * nnz))) {
        sfg_operator_destroy(op);
        return rc;
    }
    hipError_t e = hipMemcpy(op->rp, up.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(op->ci, uj.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
    for (int m = 0; m < 7 && e == hipSuccess; ++m)
        e = hipMemcpy(op->v64 + (size_t)m * nnz, uxp[m], sizeof(double) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_to_f32(op->v64, op->v32, (size_t)7 * nnz, nullptr);
    if (e == hipSuccess) {
        std::vector<double> pk((size_t)8 * nnz, 0.0);
        for (int k = 0; k < nnz; ++k)
            for (int m = 0; m < 7; ++m) pk[(size_t)8 * k + m] = uxp[m][k];
        e = hipMemcpy(op->p64, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_to_f32(op->p64, op->p32, (size_t)8 * nnz, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        sfg_operator_destroy(op);
        return fail(SFG_E_HIP, "operator upload: %s", hipGetErrorString(e));
    }
    *out = op;
    return SFG_OK;
}

extern "C" int sfg_operator_nnz(const sfg_operator *op) { return op ? op->nnz : 0; }

extern "C" void sfg_operator_destroy(sfg_operator *op)
{
    if (!op) return;
    dfree(op->rp);
    dfree(op->ci);
    dfree(op->v64);
    dfree(op->v32);
    dfree(op->p64);
    dfree(op->p32);
    delete op;
}

// ---------------------------------------------------------------------------
// media

struct sfg_medium {
    int n = 0;
    double *coef64 = nullptr; // [NCOEF][n], MASS = mass_lump
    float *coef32 = nullptr;  // [NCOEF][n], MASS = 1/mass_lump, 0 where clamped
    int *flag = nullptr;      // [n]
    int zero_mass = 0;
};

extern "C" int sfg_medium_create(sfg_medium **out, const sfg_operator *op, const double *const c[4],
                                 const double *mass_lump, const int *dirichlet,
                                 const double *const mpml[6])
{
    g_err.clear();
    if (!out || !op || !c || !mass_lump || !dirichlet || !mpml) return fail(SFG_E_ARG, "medium: bad arguments");
    *out = nullptr;
    const int n = op->n;
    for (int k = 0; k < 4; ++k)
        if (!c[k]) return fail(SFG_E_ARG, "medium: null c[%d]", k);
    for (int k = 0; k < 6; ++k)
        if (!mpml[k]) return fail(SFG_E_ARG, "medium: null mpml[%d]", k);
    std::vector<double> h((size_t)NCOEF * n);
    int zero = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) h[(size_t)k * n + i] = c[k][i];
        for (int k = 0; k < 6; ++k) h[(size_t)(DX + k) * n + i] = mpml[k][i];
        h[(size_t)MASS * n + i] = mass_lump[i];
        h[(size_t)MROW * n + i] = 1.0;
        if (mass_lump[i] == 0.0) ++zero;
    }
    sfg_medium *md = new sfg_medium;
    md->n = n;
    md->zero_mass = zero;
    int rc = SFG_OK;
    if ((rc = dalloc(&md->coef64, (size_t)NCOEF * n)) || (rc = dalloc(&md->coef32, (size_t)NCOEF * n)) ||
        (rc = dalloc(&md->flag, n))) {
        sfg_medium_destroy(md);
        return rc;
    }
    hipError_t e = hipMemcpy(md->coef64, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(md->flag, dirichlet, sizeof(int) * n, hipMemcpyHostToDevice);
    // FP32 copy on the GPU: 1/mass_lump in FP64 with the clamp folded in
    if (e == hipSuccess) e = launch_medium_f32(n, md->coef64, md->flag, md->coef32, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        sfg_medium_destroy(md);
        return fail(SFG_E_HIP, "medium upload: %s", hipGetErrorString(e));
    }
    *out = md;
    return SFG_OK;
}

extern "C" int sfg_medium_set_row_scale(sfg_medium *md, const double *row_scale)
{
    g_err.clear();
    if (!md || !row_scale) return fail(SFG_E_ARG, "row scale: bad arguments");
    const size_t n = (size_t)md->n;
    HIPCHK(hipMemcpy(md->coef64 + MROW * n, row_scale, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(launch_medium_f32(md->n, md->coef64, md->flag, md->coef32, nullptr));
    HIPCHK(hipDeviceSynchronize());
    return SFG_OK;
}

extern "C" void sfg_medium_destroy(sfg_medium *md)
{
    if (!md) return;
    dfree(md->coef64);
    dfree(md->coef32);
    dfree(md->flag);
    delete md;
}

// ---------------------------------------------------------------------------
// runs

// 4- or 8-byte words per (node, shot) of one Kahan-term buffer
static size_t comp_fields(int prec) { return prec == 3 ? 3 : prec == 2 ? 6 : 0; }

struct sfg_run {
    const sfg_operator *op = nullptr;
    const sfg_medium *md = nullptr;
    int prec = 0, nshots = 0, S = 1, SP = 1, block = 256, nbx = 1, nby = 1;
    int max_steps = 0, rec_num = 0, guard_every = 50, nslots = 1;
    bool use_graph = true;
    bool cps = false;           // coefficients per (node, shot): one medium per shot
    void *coefp = nullptr;      // [NCOEF][n][S] packed coefficients (cps)
    int *flagp = nullptr;       // [n][S] packed Dirichlet flags (cps)
    std::vector<int> perm;      // internal shot -> user shot
    std::vector<int> steps_int; // executed steps per internal shot
    void *state = nullptr, *comp = nullptr;
    void *uw = nullptr;         // FP32 forms: [2][2][n][S] U and W sums of the state
    int *pml = nullptr;         // FP32 forms: rows with a nonzero M-PML profile
    Ctrl *ctrl = nullptr;
    int *active = nullptr, *src = nullptr, *rec_ptr = nullptr, *rec_idx = nullptr;
    int *shot_steps = nullptr;  // [S] executed steps per internal shot (0 for padding)
    int *order = nullptr;       // row_order on the device (or null)
    int *hptr = nullptr, *hnode = nullptr, *kslot = nullptr, *rslot = nullptr; // LDS halos
    int *sell_row = nullptr, *sell_off = nullptr, *sell_ci = nullptr; // sliced layout (variants 8, 9)
    int *sell_own = nullptr;    // variant 9: LDS slot of each lane's own row
    void *sell_val = nullptr;
    size_t lds_bytes = 0;       // dynamic LDS of the LDS-staged kernel
    int *src_shot = nullptr;    // src = per-node entry pointers, then shot and weights per entry
    double *src_wt = nullptr;
    double *wavelet = nullptr, *seis_u = nullptr, *seis_w = nullptr, *energy = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // guard_every steps + guard, by parity of the first step
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int executed = 0;
    double last_ms = 0.0;
    bool in_multi = false;      // stepped by an sfg_multi (shares its step counter)
    StepArgs<double> a64;
    StepArgs<float> a32;
};

static void run_free(sfg_run *r)
{
    for (hipGraphExec_t g : r->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    dfree(r->state);
    dfree(r->comp);
    dfree(r->uw);
    dfree(r->pml);
    dfree(r->ctrl);
    dfree(r->active);
    dfree(r->shot_steps);
    dfree(r->order);
    dfree(r->hptr);
    dfree(r->hnode);
    dfree(r->kslot);
    dfree(r->rslot);
    dfree(r->sell_row);
    dfree(r->sell_off);
    dfree(r->sell_ci);
    dfree(r->sell_own);
    dfree(r->sell_val);
    dfree(r->src);
    dfree(r->src_shot);
    dfree(r->src_wt);
    dfree(r->rec_ptr);
    dfree(r->rec_idx);
    dfree(r->wavelet);
    dfree(r->seis_u);
    dfree(r->seis_w);
    dfree(r->energy);
    dfree(r->coefp);
    dfree(r->flagp);
    delete r;
}

extern "C" void sfg_run_destroy(sfg_run *r)
{
    if (!r) return;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    run_free(r);
}

// slot = parity of the step this launch executes
static int launch_one_step(sfg_run *r, int slot)
{
    const void *args = (r->prec == 0) ? (const void *)&r->a64 : (const void *)&r->a32;
    HIPCHK(launch_step(r->prec, r->SP, r->cps, args, r->nbx, r->nby, r->block, slot, r->lds_bytes, r->stream));
    return SFG_OK;
}

// Layout of the run's 26-field state (the step kernel derives the same from SP).
static StateLayout state_layout(const sfg_run *r)
{
    StateLayout L;
    L.gs = node_major(r->SP) ? (r->SP >= 64 ? 64 : r->SP) : 0;
    L.ng = L.gs ? r->S / L.gs : 0;
    L.nf = NFIELD;
    L.S = r->S;
    L.F = (size_t)r->op->n * r->S;
    return L;
}

// ctrl / stream default to the run's own (a multi-run passes its shared ones)
static int launch_guard_now(sfg_run *r, const Ctrl *ctrl = nullptr, hipStream_t st = nullptr)
{
    HIPCHK(launch_guard(r->prec, r->state, r->uw, r->op->n, state_layout(r), ctrl ? ctrl : r->ctrl, r->shot_steps,
                        r->energy, r->nslots, r->guard_every, st ? st : r->stream));
    return SFG_OK;
}

// Captures what `launches` enqueues on `st` into an instantiated graph.
template <class F>
static int capture_graph(hipStream_t st, hipGraphExec_t *out, F launches)
{
    hipGraph_t g = nullptr;
    int rc = SFG_OK;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = launches();
        hipError_t e2 = hipStreamEndCapture(st, &g);
        if (rc == SFG_OK && e2 != hipSuccess) rc = fail(SFG_E_HIP, "graph capture: %s", hipGetErrorString(e2));
    } else {
        rc = fail(SFG_E_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    if (rc == SFG_OK) {
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(SFG_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

static int run_create_impl(sfg_run **out, const sfg_operator *op, const sfg_medium *const *media, int nmedia,
                           const sfg_shot_desc *d)
{
    g_err.clear();
    if (!out || !op || !media || nmedia <= 0 || !d) return fail(SFG_E_ARG, "run: bad arguments");
    *out = nullptr;
    if (d->precision < 0 || d->precision > 3) return fail(SFG_E_ARG, "run: unknown precision %d", d->precision);
    if (d->kernel_variant < 0 || d->kernel_variant > 9)
        return fail(SFG_E_ARG, "run: unknown kernel_variant %d", d->kernel_variant);
    if (d->xcd_remap < -1 || d->xcd_remap > (1 << 20))
        return fail(SFG_E_ARG, "run: xcd_remap %d (0, -1 or a stripe of 1 .. 2^20 row blocks)", d->xcd_remap);
    if (d->nshots <= 0 || !d->src_node) return fail(SFG_E_ARG, "run: need >= 1 shot with a source node");
    if (nmedia != 1 && nmedia != d->nshots) return fail(SFG_E_ARG, "run: need one medium or one per shot");
    for (int k = 0; k < nmedia; ++k) {
        if (!media[k] || media[k]->n != op->n) return fail(SFG_E_ARG, "run: medium %d does not match the operator", k);
        if (d->precision != 0 && media[k]->zero_mass)
            return fail(SFG_E_ZERO_MASS, "run: medium %d has %d zero lumped-mass entries (FP32 paths need mass > 0)",
                        k, media[k]->zero_mass);
    }
    const sfg_medium *md = media[0];
    if (d->max_steps <= 0) return fail(SFG_E_ARG, "run: max_steps must be > 0");
    if (d->rec_num < 0 || (d->rec_num > 0 && !d->rec_node)) return fail(SFG_E_ARG, "run: bad receivers");
    const int n = op->n;
    for (int s = 0; s < d->nshots; ++s)
        if (d->src_node[s] < -1 || d->src_node[s] >= n) return fail(SFG_E_ARG, "run: source node out of range");
    if (d->src_count) {
        long long tot = 0;
        for (int s = 0; s < d->nshots; ++s) {
            if (d->src_count[s] < 0) return fail(SFG_E_ARG, "run: negative src_count");
            tot += d->src_count[s];
        }
        if (tot > 0 && (!d->src_nodes || !d->src_weight)) return fail(SFG_E_ARG, "run: source entries missing");
        for (long long q = 0; q < tot; ++q)
            if (d->src_nodes[q] < -1 || d->src_nodes[q] >= n) return fail(SFG_E_ARG, "run: source entry node out of range");
    }
    for (int q = 0; q < d->rec_num; ++q)
        if (d->rec_node[q] < 0 || d->rec_node[q] >= n)
            return fail(SFG_E_ARG, "run: receiver %d node %d out of range (the reference reads U[-1])", q, d->rec_node[q]);
    std::vector<int> steps(d->nshots);
    for (int s = 0; s < d->nshots; ++s) {
        steps[s] = d->nsteps ? d->nsteps[s] : d->max_steps;
        if (steps[s] < 0 || steps[s] > d->max_steps) return fail(SFG_E_ARG, "run: nsteps[%d] out of range", s);
    }

    sfg_run *r = new sfg_run;
    r->op = op;
    r->md = md;
    r->cps = nmedia > 1;
    r->prec = d->precision;
    r->nshots = d->nshots;
    r->max_steps = d->max_steps;
    r->rec_num = d->rec_num;
    r->guard_every = d->guard_every > 0 ? d->guard_every : 50;
    r->use_graph = d->use_graph != 0;
    // lane grouping: shots per lane group and the padded shot stride
    int sp = d->shots_per_lane_group;
    if (sp == 0) sp = d->nshots <= 1 ? 1 : d->nshots <= 4 ? 4 : d->nshots <= 16 ? 16 : 64;
    if (sp != 1 && sp != 4 && sp != 16 && sp != 32 && sp != 64) {
        run_free(r);
        return fail(SFG_E_ARG, "run: shots_per_lane_group must be 1, 4, 16, 32 or 64");
    }
    // shots are padded to whole lane groups; each group of sp shots is one
    // column of tiles (blockIdx.y)
    r->SP = sp;
    r->S = sp * ((d->nshots + sp - 1) / sp);
    r->nby = r->S / sp;
    if ((size_t)n * r->S >= (size_t)1 << 31) {
        run_free(r);
        return fail(SFG_E_ARG, "run: node_num * padded shots must stay below 2^31");
    }
    r->block = d->block > 0 ? d->block : (sp == 1 && n < 256 * 1024 ? 64 : 256);
    if (r->block % 64 != 0 || r->block > 256) {
        run_free(r);
        return fail(SFG_E_ARG, "run: block must be 64, 128, 192 or 256 threads (kernel launch bound)");
    }
    const int rows_per_block = (r->block / 64) * (sp >= 64 ? 1 : 64 / sp);
    r->nbx = (n + rows_per_block - 1) / rows_per_block;

    // shots sorted by decreasing step count, so the active shots are a prefix
    r->perm.resize(d->nshots);
    std::iota(r->perm.begin(), r->perm.end(), 0);
    std::stable_sort(r->perm.begin(), r->perm.end(), [&](int x, int y) { return steps[x] > steps[y]; });
    r->steps_int.resize(d->nshots);
    for (int s = 0; s < d->nshots; ++s) r->steps_int[s] = steps[r->perm[s]];
    // sources grouped by node: (internal shot, u weight, w weight); a point force
    // (the reference's only source) has weights sin/cos(angle) with the
    // truncated pi of elastic_wave.c:67
    const double pi_ref = 3.1415926535898;
    const double su = sin(d->angle_deg * pi_ref / 180.0), sw = cos(d->angle_deg * pi_ref / 180.0);
    std::vector<int> first(d->nshots + 1, 0);
    for (int u = 0; u < d->nshots; ++u) first[u + 1] = first[u] + (d->src_count ? d->src_count[u] : 1);
    struct SrcEntry { int node, shot; double wu, ww; };
    std::vector<SrcEntry> ent;
    for (int s = 0; s < d->nshots; ++s) {
        const int u = r->perm[s];
        for (int q = first[u]; q < first[u + 1]; ++q) {
            const int node = d->src_count ? d->src_nodes[q] : d->src_node[u];
            const double wu = d->src_count ? d->src_weight[2 * q] : su;
            const double ww = d->src_count ? d->src_weight[2 * q + 1] : sw;
            if (node < 0) continue; // no source for this shot
            ent.push_back({node, s, wu, ww});
        }
    }
    // entries grouped by (node, shot); repeated (node, shot) pairs are summed in input order
    std::stable_sort(ent.begin(), ent.end(), [](const SrcEntry &x, const SrcEntry &y) {
        return x.node != y.node ? x.node < y.node : x.shot < y.shot;
    });
    {
        size_t m = 0;
        for (size_t q = 0; q < ent.size(); ++q) {
            if (m > 0 && ent[m - 1].node == ent[q].node && ent[m - 1].shot == ent[q].shot) {
                ent[m - 1].wu += ent[q].wu;
                ent[m - 1].ww += ent[q].ww;
            } else {
                ent[m++] = ent[q];
            }
        }
        ent.resize(m);
    }
    std::vector<int> sptr(n + 1, 0), sshot(std::max<size_t>(ent.size(), 1), 0);
    std::vector<double> swt(2 * std::max<size_t>(ent.size(), 1), 0.0);
    for (size_t q = 0; q < ent.size(); ++q) {
        sptr[ent[q].node + 1]++;
        sshot[q] = ent[q].shot;
        swt[2 * q] = ent[q].wu;
        swt[2 * q + 1] = ent[q].ww;
    }
    for (int i = 0; i < n; ++i) sptr[i + 1] += sptr[i];
    std::vector<int> act(d->max_steps);
    for (int t = 0; t < d->max_steps; ++t) {
        int c = 0;
        while (c < d->nshots && r->steps_int[c] > t) ++c;
        act[t] = c;
    }
    // Ricker table: step t of the run is the reference's it = t + 2, time (it+1)*dt
    std::vector<double> wav(d->max_steps);
    for (int t = 0; t < d->max_steps; ++t) wav[t] = seismic_source(d->f0, d->t0, d->amplitude, (t + 3) * d->dt);
    // receivers grouped by node
    std::vector<int> rptr(n + 1, 0), ridx(std::max(d->rec_num, 1));
    for (int q = 0; q < d->rec_num; ++q) rptr[d->rec_node[q] + 1]++;
    for (int i = 0; i < n; ++i) rptr[i + 1] += rptr[i];
    {
        std::vector<int> fillp(rptr.begin(), rptr.end() - 1);
        for (int q = 0; q < d->rec_num; ++q) ridx[fillp[d->rec_node[q]]++] = q;
    }
    r->nslots = (d->max_steps + r->guard_every - 1) / r->guard_every + 1;
    if (d->row_order) {
        std::vector<char> seen(n, 0);
        for (int i = 0; i < n; ++i) {
            const int v = d->row_order[i];
            if (v < 0 || v >= n || seen[v]) {
                run_free(r);
                return fail(SFG_E_ARG, "run: row_order is not a permutation of the %d nodes", n);
            }
            seen[v] = 1;
        }
    }

    const size_t F = (size_t)n * r->S;
    const size_t w = r->prec == 0 ? 8 : 4;
    const size_t seis_n = (size_t)r->S * d->max_steps * std::max(d->rec_num, 1);

    // LDS-staged gathers: the halo (distinct columns) of each block of rows
    std::vector<int> hptr, hnode, kslot, rslot;
    int lds_stride = 0;
    if (d->kernel_variant == 7) {
        if (!node_major(sp)) {
            run_free(r);
            return fail(SFG_E_ARG, "run: kernel_variant 7 (LDS-staged gathers) needs 16, 32 or 64 shots per lane group");
        }
        const int R = rows_per_block, gs = sp >= 64 ? 64 : sp;
        hptr.assign(r->nbx + 1, 0);
        kslot.assign(op->nnz, 0);
        rslot.assign(n, 0);
        std::vector<int> stamp(n, -1), slot_of(n, 0), blk;
        size_t maxh = 0;
        for (int b = 0; b < r->nbx; ++b) {
            blk.clear();
            auto add = [&](int x) { if (stamp[x] != b) { stamp[x] = b; blk.push_back(x); } };
            const int l1 = std::min(n, (b + 1) * R);
            for (int l = b * R; l < l1; ++l) {
                const int i = d->row_order ? d->row_order[l] : l;
                add(i);
                for (int k = op->hrp[i]; k < op->hrp[i + 1]; ++k) add(op->hci[k]);
            }
            std::sort(blk.begin(), blk.end()); // ascending node ids: staging reads walk memory upwards
            for (size_t h = 0; h < blk.size(); ++h) slot_of[blk[h]] = (int)h;
            for (int l = b * R; l < l1; ++l) {
                const int i = d->row_order ? d->row_order[l] : l;
                rslot[i] = slot_of[i];
                for (int k = op->hrp[i]; k < op->hrp[i + 1]; ++k) kslot[k] = slot_of[op->hci[k]];
            }
            hnode.insert(hnode.end(), blk.begin(), blk.end());
            hptr[b + 1] = (int)hnode.size();
            maxh = std::max(maxh, blk.size());
        }
        int lim = 0;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g_device) != hipSuccess) lim = 65536;
        // pad each slot so the lane groups of one wave start on different LDS banks
        lds_stride = 20 * gs + (gs < 64 ? gs : 0);
        if (maxh * lds_stride * w > (size_t)lim) lds_stride = 20 * gs;
        r->lds_bytes = maxh * lds_stride * w;
        if (r->lds_bytes > (size_t)lim) {
            run_free(r);
            return fail(SFG_E_ARG, "run: a block's halo of %zu nodes needs %zu bytes of LDS (limit %d); use a "
                        "smaller block or a locality row order", maxh, r->lds_bytes, lim);
        }
        hipError_t ea = allow_step_lds(r->prec, sp, r->cps, r->lds_bytes);
        if (ea != hipSuccess) {
            run_free(r);
            return fail(SFG_E_HIP, "run: LDS size %zu: %s", r->lds_bytes, hipGetErrorString(ea));
        }
    }
    // sliced layout (kernel variants 8 and 9): the row order cut into windows,
    // each window's rows sorted by decreasing length (stable), then slices of 64
    // consecutive rows; a slice's slots run to its longest row.  Variant 9 also
    // stages each block's halo (the distinct columns of its rows) in LDS: its
    // windows lie inside blocks (default: one window per block), and the slots
    // hold the columns' positions in the block's sorted halo list.
    std::vector<int> srow, soff, sci, ssrc, sown;
    if (d->kernel_variant == 8 || d->kernel_variant == 9) {
        const bool staged = d->kernel_variant == 9;
        const int R = r->block;                                  // rows per block (one per lane)
        const int win = d->sell_sigma > 0 ? d->sell_sigma : 64;
        if (sp != 1 || win % 64 || (staged && R % win)) {
            run_free(r);
            return fail(SFG_E_ARG, "run: kernel_variant %d (sliced layout) needs one shot per lane group and a "
                        "sorting window of a multiple of 64 rows%s (got %d shots per group, window %d)",
                        d->kernel_variant, staged ? " dividing the block" : "", sp, win);
        }
        const size_t nslices = (size_t)r->nbx * (R / 64);
        std::vector<int> logical(n);
        for (int i = 0; i < n; ++i) logical[i] = d->row_order ? d->row_order[i] : i;
        auto len = [&](int i) { return op->hrp[i + 1] - op->hrp[i]; };
        for (int w0 = 0; w0 < n; w0 += win)
            std::stable_sort(logical.begin() + w0, logical.begin() + std::min(n, w0 + win),
                             [&](int x, int y) { return len(x) > len(y); });
        srow.assign(nslices * 64, -1);
        std::copy(logical.begin(), logical.end(), srow.begin());
        soff.assign(nslices + 1, 0);
        for (size_t sl = 0; sl < nslices; ++sl) {
            int m = 0;
            for (int l = 0; l < 64; ++l)
                if (srow[sl * 64 + l] >= 0) m = std::max(m, len(srow[sl * 64 + l]));
            if ((long long)soff[sl] + m > 0x7fffffffLL / 64) {
                run_free(r);
                return fail(SFG_E_ARG, "run: sliced layout beyond 2^31 slots");
            }
            soff[sl + 1] = soff[sl] + m;
        }
        // halo of each block (variant 9): its rows' columns, sorted, and their slots
        std::vector<int> slot_of;
        size_t maxh = 0;
        if (staged) {
            hptr.assign(r->nbx + 1, 0);
            std::vector<int> stamp(n, -1), blk;
            slot_of.assign(n, 0);
            sown.assign(nslices * 64, 0);
            for (int b = 0; b < r->nbx; ++b) {
                blk.clear();
                for (int l = b * R; l < (b + 1) * R; ++l) {
                    const int i = srow[l];
                    if (i < 0) continue;
                    for (int k = op->hrp[i]; k < op->hrp[i + 1]; ++k)
                        if (stamp[op->hci[k]] != b) { stamp[op->hci[k]] = b; blk.push_back(op->hci[k]); }
                    if (stamp[i] != b) { stamp[i] = b; blk.push_back(i); }
                }
                std::sort(blk.begin(), blk.end());
                for (size_t h = 0; h < blk.size(); ++h) slot_of[blk[h]] = (int)h;
                for (int l = b * R; l < (b + 1) * R; ++l)
                    if (srow[l] >= 0) sown[l] = slot_of[srow[l]];
                hnode.insert(hnode.end(), blk.begin(), blk.end());
                hptr[b + 1] = (int)hnode.size();
                maxh = std::max(maxh, blk.size());
                // slots of this block's slices, while slot_of holds its halo
                for (int sl = b * (R / 64); sl < (b + 1) * (R / 64); ++sl)
                    for (int l = 0; l < 64; ++l) {
                        const int row = srow[(size_t)sl * 64 + l];
                        if (row < 0) continue;
                        for (int k = 0; k < soff[sl + 1] - soff[sl]; ++k) {
                            const size_t q = ((size_t)soff[sl] + k) * 64 + l;
                            if (sci.size() <= q) {
                                sci.resize(std::max<size_t>((size_t)soff[nslices] * 64, 64), 0);
                                ssrc.resize(sci.size(), -1);
                            }
                            if (k < len(row)) {
                                sci[q] = slot_of[op->hci[op->hrp[row] + k]];
                                ssrc[q] = op->hrp[row] + k;
                            } else {
                                sci[q] = slot_of[row];  // the row itself, value 0
                            }
                        }
                    }
            }
            int lim = 0;
            if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g_device) != hipSuccess)
                lim = 65536;
            lds_stride = (int)maxh;
            r->lds_bytes = 20 * maxh * w;
            if (r->lds_bytes > (size_t)lim) {
                run_free(r);
                return fail(SFG_E_ARG, "run: a block's halo of %zu nodes needs %zu bytes of LDS (limit %d); use a "
                            "smaller block or a row order that keeps blocks compact (a patch order)", maxh,
                            r->lds_bytes, lim);
            }
            hipError_t ea = allow_step_lds(r->prec, 1, r->cps, r->lds_bytes);
            if (ea != hipSuccess) {
                run_free(r);
                return fail(SFG_E_HIP, "run: LDS size %zu: %s", r->lds_bytes, hipGetErrorString(ea));
            }
        }
        const size_t slots = (size_t)soff[nslices] * 64;
        sci.resize(std::max<size_t>(slots, 64), 0);
        ssrc.resize(sci.size(), -1);
        if (!staged)
            for (size_t sl = 0; sl < nslices; ++sl)
                for (int l = 0; l < 64; ++l) {
                    const int row = srow[sl * 64 + l];
                    for (int k = 0; k < soff[sl + 1] - soff[sl]; ++k) {
                        const size_t q = ((size_t)soff[sl] + k) * 64 + l;
                        if (row >= 0 && k < len(row)) {
                            sci[q] = op->hci[op->hrp[row] + k];
                            ssrc[q] = op->hrp[row] + k;
                        } else {
                            sci[q] = row >= 0 ? row : 0;   // column = the row itself, value 0
                        }
                    }
                }
    }
    int rc = SFG_OK;
    hipError_t e = hipSuccess;
    if ((rc = dalloc((char **)&r->state, 2 * NFIELD * F * w)) ||
        (r->prec >= 2 && (rc = dalloc((char **)&r->comp, 2 * comp_fields(r->prec) * F * w))) ||
        (r->prec != 0 && ((rc = dalloc((char **)&r->uw, 2 * 2 * F * w)) || (rc = dalloc(&r->pml, n)))) ||
        (rc = dalloc(&r->ctrl, 1)) ||
        (rc = dalloc(&r->active, d->max_steps)) || (rc = dalloc(&r->shot_steps, r->S)) ||
        (rc = dalloc(&r->src, n + 1)) ||
        (rc = dalloc(&r->src_shot, sshot.size())) || (rc = dalloc(&r->src_wt, swt.size())) ||
        (rc = dalloc(&r->rec_ptr, n + 1)) || (rc = dalloc(&r->rec_idx, ridx.size())) ||
        (rc = dalloc(&r->wavelet, d->max_steps)) || (rc = dalloc(&r->seis_u, seis_n)) ||
        (rc = dalloc(&r->seis_w, seis_n)) || (rc = dalloc(&r->energy, (size_t)r->nslots * r->S * 2)) ||
        (d->row_order && (rc = dalloc(&r->order, n))) ||
        (!hptr.empty() && ((rc = dalloc(&r->hptr, hptr.size())) || (rc = dalloc(&r->hnode, hnode.size())) ||
                           (rc = dalloc(&r->kslot, kslot.size())) || (rc = dalloc(&r->rslot, rslot.size()))))) {
        run_free(r);
        return rc;
    }
    e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&r->ev0);
    if (e == hipSuccess) e = hipEventCreate(&r->ev1);
    if (e == hipSuccess) e = launch_fill_zero(r->state, 2 * NFIELD * F * w, r->stream);
    if (e == hipSuccess && r->comp) e = launch_fill_zero(r->comp, 2 * comp_fields(r->prec) * F * w, r->stream);
    if (e == hipSuccess && r->uw) e = launch_fill_zero(r->uw, 2 * 2 * F * w, r->stream);
    if (e == hipSuccess) e = launch_fill_zero(r->ctrl, sizeof(Ctrl), r->stream);
    if (e == hipSuccess) e = launch_fill_zero(r->seis_u, seis_n * 8, r->stream);
    if (e == hipSuccess) e = launch_fill_zero(r->seis_w, seis_n * 8, r->stream);
    if (e == hipSuccess) e = launch_fill_zero(r->energy, (size_t)r->nslots * r->S * 16, r->stream);
    std::vector<int> shot_steps(r->S, 0);
    std::copy(r->steps_int.begin(), r->steps_int.end(), shot_steps.begin());
    if (e == hipSuccess) e = hipMemcpyAsync(r->active, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->shot_steps, shot_steps.data(), sizeof(int) * r->S, hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && r->order)
        e = hipMemcpyAsync(r->order, d->row_order, sizeof(int) * n, hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && r->hptr) e = hipMemcpyAsync(r->hptr, hptr.data(), sizeof(int) * hptr.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && r->hnode) e = hipMemcpyAsync(r->hnode, hnode.data(), sizeof(int) * hnode.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && r->kslot && !kslot.empty())
        e = hipMemcpyAsync(r->kslot, kslot.data(), sizeof(int) * kslot.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && r->rslot && !rslot.empty())
        e = hipMemcpyAsync(r->rslot, rslot.data(), sizeof(int) * rslot.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess && !srow.empty()) {
        int *dsrc = nullptr;
        if ((rc = dalloc(&r->sell_row, srow.size())) || (rc = dalloc(&r->sell_off, soff.size())) ||
            (rc = dalloc(&r->sell_ci, sci.size())) || (rc = dalloc((char **)&r->sell_val, sci.size() * 7 * w)) ||
            (rc = dalloc(&dsrc, ssrc.size()))) {
            dfree(dsrc);
            run_free(r);
            return rc;
        }
        if (!sown.empty()) {
            if ((rc = dalloc(&r->sell_own, sown.size()))) {
                dfree(dsrc);
                run_free(r);
                return rc;
            }
            e = hipMemcpyAsync(r->sell_own, sown.data(), sizeof(int) * sown.size(), hipMemcpyHostToDevice, r->stream);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->sell_row, srow.data(), sizeof(int) * srow.size(), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->sell_off, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(r->sell_ci, sci.data(), sizeof(int) * sci.size(), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dsrc, ssrc.data(), sizeof(int) * ssrc.size(), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess)
            e = launch_sell_values(r->prec == 0 ? (const void *)op->v64 : (const void *)op->v32, (size_t)op->nnz, dsrc,
                                   sci.size(), (int)w, r->sell_val, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        dfree(dsrc);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(r->src, sptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->src_shot, sshot.data(), sizeof(int) * sshot.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->src_wt, swt.data(), sizeof(double) * swt.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->rec_ptr, rptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->rec_idx, ridx.data(), sizeof(int) * ridx.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->wavelet, wav.data(), sizeof(double) * wav.size(), hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) {
        run_free(r);
        return fail(SFG_E_HIP, "run setup: %s", hipGetErrorString(e));
    }
    if (r->cps) {
        // one medium per shot: coefficients and flags become [.][node][shot] arrays
        // in the internal (step-sorted) shot order
        std::vector<const void *> csrc(d->nshots), fsrc(d->nshots);
        for (int s = 0; s < d->nshots; ++s) {
            const sfg_medium *m = media[r->perm[s]];
            csrc[s] = r->prec == 0 ? (const void *)m->coef64 : (const void *)m->coef32;
            fsrc[s] = m->flag;
        }
        const void **dsrc = nullptr;
        if ((rc = dalloc((char **)&r->coefp, (size_t)NCOEF * F * w)) || (rc = dalloc(&r->flagp, F)) ||
            (rc = dalloc((char **)&dsrc, sizeof(void *) * 2 * d->nshots))) {
            dfree(dsrc);
            run_free(r);
            return rc;
        }
        e = hipMemcpy(dsrc, csrc.data(), sizeof(void *) * d->nshots, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(dsrc + d->nshots, fsrc.data(), sizeof(void *) * d->nshots, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_pack_per_shot(dsrc, d->nshots, NCOEF, n, r->S, (int)w, r->coefp, r->stream);
        if (e == hipSuccess) e = launch_pack_per_shot(dsrc + d->nshots, d->nshots, 1, n, r->S, 4, r->flagp, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        dfree(dsrc);
        if (e != hipSuccess) {
            run_free(r);
            return fail(SFG_E_HIP, "run setup (per-shot media): %s", hipGetErrorString(e));
        }
    }

    if (r->pml) {
        // rows outside every damping layer (for all shots) take the short gather
        e = launch_pml_rows(r->prec, r->cps ? r->coefp : (const void *)md->coef32, n, r->cps ? r->S : 1, op->rp,
                            op->ci, r->pml, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) {
            run_free(r);
            return fail(SFG_E_HIP, "run setup (M-PML rows): %s", hipGetErrorString(e));
        }
    }

    // kernel arguments (identical for every step: the step index lives on the device)
    const double alpha = NEWMARK_ALPHA, delta = NEWMARK_DELTA;
    auto fill = [&](auto &a, auto *state, auto *comp, auto *const vals, auto *coef, auto cast) {
        a.n = n;
        a.S = r->S;
        a.rows_per_block = rows_per_block;
        a.nsteps_max = d->max_steps;
        a.xcd_remap = d->xcd_remap == 0 ? 1 : 0;
        a.xcd_stripe = d->xcd_remap > 0 ? d->xcd_remap : 0;
        a.sp = r->SP;
        a.cps = r->cps ? 1 : 0;
        a.nbx = r->nbx;
        a.rp = op->rp;
        a.order = r->order;
        a.hptr = r->hptr;
        a.hnode = r->hnode;
        a.kslot = r->kslot;
        a.rslot = r->rslot;
        a.lds_stride = lds_stride;
        a.sell_row = r->sell_row;
        a.sell_off = r->sell_off;
        a.sell_ci = r->sell_ci;
        a.sell_own = r->sell_own;
        a.ci = op->ci;
        for (int m = 0; m < 7; ++m) a.A[m] = vals + (size_t)m * op->nnz;
        // descriptor variant -> kernel variant (see step_tile)
        static const int kv[10] = {1, 1, 0, 2, 3, 4, 5, 6, 7, 8};
        a.variant = kv[d->kernel_variant];
        a.coef = r->cps ? (decltype(coef))r->coefp : coef;
        a.flag = r->cps ? r->flagp : md->flag;
        a.state = state;
        a.comp = comp;
        a.uw = (decltype(comp))r->uw;
        a.pml = r->pml;
        // the LDS-staged walk reads every neighbour's split fields, so it keeps them on all
        // rows; one row per lane (lane groups of 1 or 4 shots) mixes row kinds in a wave
        // and measured slower with summed rows
        a.agg = (r->prec != 0 && d->kernel_variant != 7 && node_major(r->SP)) ? 1 : 0;
        a.fstride = F;
        a.active = r->active;
        a.src_ptr = r->src;
        a.src_shot = r->src_shot;
        a.src_wt = r->src_wt;
        a.wavelet = r->wavelet;
        a.src_u = su;
        a.src_w = sw;
        a.dt = cast(d->dt);
        a.ca = cast(0.5 - alpha);
        a.alpha = cast(alpha);
        a.cd = cast(1 - delta);
        a.delta = cast(delta);
        a.rec_ptr = r->rec_ptr;
        a.rec_idx = r->rec_idx;
        a.rec_num = d->rec_num;
        a.seis_u = r->seis_u;
        a.seis_w = r->seis_w;
        a.ctrl = r->ctrl;
    };
    if (r->prec == 0)
        fill(r->a64, (double *)r->state, (double *)r->comp, op->v64, md->coef64, [](double v) { return v; }),
            r->a64.Ap = op->p64, r->a64.sell_val = (const double *)r->sell_val;
    else
        fill(r->a32, (float *)r->state, (float *)r->comp, op->v32, md->coef32, [](double v) { return (float)v; }),
            r->a32.Ap = op->p32, r->a32.sell_val = (const float *)r->sell_val;

    // captured graphs of guard_every steps followed by the energy guard: one
    // per parity of the first step (a second one only when guard_every is odd)
    if (r->use_graph) {
        for (int p = 0; p < (r->guard_every % 2 ? 2 : 1) && rc == SFG_OK; ++p)
            rc = capture_graph(r->stream, &r->graph[p], [&]() {
                int c = SFG_OK;
                for (int k = 0; k < r->guard_every && c == SFG_OK; ++k) c = launch_one_step(r, (p + k) & 1);
                return c == SFG_OK ? launch_guard_now(r) : c;
            });
        if (rc != SFG_OK) {
            run_free(r);
            return rc;
        }
    }
    *out = r;
    return SFG_OK;
}

extern "C" int sfg_run_create(sfg_run **out, const sfg_operator *op, const sfg_medium *md,
                              const sfg_shot_desc *d)
{
    return run_create_impl(out, op, &md, 1, d);
}

extern "C" int sfg_run_create_media(sfg_run **out, const sfg_operator *op, const sfg_medium *const *media,
                                    const sfg_shot_desc *d)
{
    if (!d) return fail(SFG_E_ARG, "run: bad arguments");
    return run_create_impl(out, op, media, d->nshots, d);
}

static int check_guard(sfg_run *r, int upto_step, const Ctrl *ctrl = nullptr)
{
    const int nchk = std::min(r->nslots, (upto_step + r->guard_every - 1) / r->guard_every + 1);
    std::vector<double> en((size_t)nchk * r->S * 2);
    HIPCHK(hipMemcpy(en.data(), r->energy, sizeof(double) * en.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < nchk; ++c)
        for (int s = 0; s < r->nshots; ++s) {
            const double eu = en[((size_t)c * r->S + s) * 2], ew = en[((size_t)c * r->S + s) * 2 + 1];
            if (!(eu <= 10e6) || !(ew <= 10e6))
                return fail(SFG_E_ENERGY, "energy guard: shot %d sum(U^2)=%g sum(W^2)=%g at check %d (step <= %d)",
                            r->perm[s], eu, ew, c, c * r->guard_every);
        }
    Ctrl h;
    HIPCHK(hipMemcpy(&h, ctrl ? ctrl : r->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (h.err & 1) return fail(SFG_E_ZERO_MASS, "zero lumped mass met (outputs of those nodes left unchanged)");
    const int dev_steps = std::max(h.step[0], h.step[1]);
    if (dev_steps != r->executed) return fail(SFG_E_HIP, "device step counter %d != host %d", dev_steps, r->executed);
    return SFG_OK;
}

extern "C" int sfg_run_advance(sfg_run *r, int nsteps)
{
    g_err.clear();
    if (!r || nsteps < 0) return fail(SFG_E_ARG, "advance: bad arguments");
    if (r->in_multi) return fail(SFG_E_ARG, "advance: this run is stepped by a multi-run");
    int todo = std::min(nsteps, r->max_steps - r->executed);
    const int steps = todo;
    HIPCHK(hipEventRecord(r->ev0, r->stream));
    while (todo > 0) {
        hipGraphExec_t g = r->graph[r->executed & 1];
        if (g && todo >= r->guard_every && r->executed % r->guard_every == 0) {
            HIPCHK(hipGraphLaunch(g, r->stream));
            r->executed += r->guard_every;
            todo -= r->guard_every;
        } else {
            int rc = launch_one_step(r, r->executed & 1);
            if (rc) return rc;
            r->executed += 1;
            todo -= 1;
            if (r->executed % r->guard_every == 0 || r->executed == r->max_steps) {
                rc = launch_guard_now(r);
                if (rc) return rc;
            }
        }
    }
    HIPCHK(hipEventRecord(r->ev1, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    r->last_ms = steps > 0 ? (double)ms / steps : 0.0;
    int rc = check_guard(r, r->executed);
    if (rc) return rc;
    return r->executed;
}

extern "C" int sfg_run_reset(sfg_run *r)
{
    g_err.clear();
    if (!r) return fail(SFG_E_ARG, "reset: null run");
    if (r->in_multi) return fail(SFG_E_ARG, "reset: this run is stepped by a multi-run");
    const size_t F = (size_t)r->op->n * r->S;
    const size_t w = r->prec == 0 ? 8 : 4;
    const size_t seis_n = (size_t)r->S * r->max_steps * std::max(r->rec_num, 1);
    // the same zero fills as at creation (elastic_wave.c:466-502), on the run's stream
    HIPCHK(launch_fill_zero(r->state, 2 * NFIELD * F * w, r->stream));
    if (r->comp) HIPCHK(launch_fill_zero(r->comp, 2 * comp_fields(r->prec) * F * w, r->stream));
    if (r->uw) HIPCHK(launch_fill_zero(r->uw, 2 * 2 * F * w, r->stream));
    HIPCHK(launch_fill_zero(r->ctrl, sizeof(Ctrl), r->stream));
    HIPCHK(launch_fill_zero(r->seis_u, seis_n * 8, r->stream));
    HIPCHK(launch_fill_zero(r->seis_w, seis_n * 8, r->stream));
    HIPCHK(launch_fill_zero(r->energy, (size_t)r->nslots * r->S * 16, r->stream));
    r->executed = 0;
    r->last_ms = 0.0;
    return SFG_OK;
}

extern "C" int sfg_run_sync(sfg_run *r)
{
    if (!r) return fail(SFG_E_ARG, "sync: null run");
    HIPCHK(hipStreamSynchronize(r->stream));
    return r->executed;
}

extern "C" double sfg_run_last_ms_per_step(const sfg_run *r) { return r ? r->last_ms : 0.0; }

extern "C" double sfg_run_step_bytes(const sfg_run *r, int active)
{
    if (!r) return 0.0;
    const double N = r->op->n, nnz = r->op->nnz, w = r->prec == 0 ? 8.0 : 4.0;
    // Kahan terms read and written: 6 words, or 6 bfloat16 (3 words) in precision 3
    const double K = r->prec >= 2 ? 2.0 * comp_fields(r->prec) * w : 0.0;
    // per-shot media read 12 coefficients per (node, shot) instead of 11 per node
    const double coef = r->cps ? (double)active * N * 12.0 * w : 11.0 * N * w;
    return 4.0 * (N + 1) + nnz * (4.0 + 7.0 * w) + coef + (double)active * N * (52.0 * w + K);
}

extern "C" int sfg_run_seismograms(sfg_run *r, double *u, double *w)
{
    g_err.clear();
    if (!r || !u || !w) return fail(SFG_E_ARG, "seismograms: bad arguments");
    HIPCHK(hipStreamSynchronize(r->stream));
    const size_t per = (size_t)r->max_steps * r->rec_num;
    if (per == 0) return SFG_OK;
    std::vector<double> hu(per * r->S), hw(per * r->S);
    HIPCHK(hipMemcpy(hu.data(), r->seis_u, sizeof(double) * hu.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hw.data(), r->seis_w, sizeof(double) * hw.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < r->nshots; ++s) {
        std::memcpy(u + per * r->perm[s], hu.data() + per * s, sizeof(double) * per);
        std::memcpy(w + per * r->perm[s], hw.data() + per * s, sizeof(double) * per);
    }
    return SFG_OK;
}

extern "C" int sfg_run_snapshot(sfg_run *r, int shot, double *U, double *W)
{
    g_err.clear();
    if (!r || !U || !W || shot < 0 || shot >= r->nshots) return fail(SFG_E_ARG, "snapshot: bad arguments");
    int s = 0;
    while (r->perm[s] != shot) ++s;
    // a shot that stopped early holds its state in the buffer of its own last step
    const int done = std::min(r->executed, r->steps_int[s]);
    const int n = r->op->n;
    double *dU, *dW;
    int rc = dalloc(&dU, 2 * (size_t)n);
    if (rc) return rc;
    dW = dU + n;
    hipError_t e = launch_gather_uw(r->prec, r->state, r->uw, n, state_layout(r), s, done & 1, dU, dW,
                                    r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e == hipSuccess) e = hipMemcpy(U, dU, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(W, dW, sizeof(double) * n, hipMemcpyDeviceToHost);
    dfree(dU);
    if (e != hipSuccess) return fail(SFG_E_HIP, "snapshot: %s", hipGetErrorString(e));
    return SFG_OK;
}

// ---------------------------------------------------------------------------
// allocation-free snapshots

// Staging vector of sfg_run_snapshot_noalloc: U then W of one shot, one copy
// per GPU (module-scope device memory, held until the process ends).  The
// array's name is a host-side stand-in; host code uses only the address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[2 * (size_t)SFG_NOALLOC_MAX_NODES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;               // serialises the calls of this process
double *g_noalloc_ptr[kMaxDevices] = {}; // device address of `scratch`, per device
} // namespace

extern "C" int sfg_run_snapshot_noalloc(sfg_run *r, int shot, double *U, double *W)
{
    g_err.clear();
    if (!r || !U || !W || shot < 0 || shot >= r->nshots) return fail(SFG_E_ARG, "snapshot_noalloc: bad arguments");
    const int n = r->op->n;
    if (n > SFG_NOALLOC_MAX_NODES)
        return fail(SFG_E_ARG, "snapshot_noalloc: %d nodes exceed the staging limit of %d (use sfg_run_snapshot)", n,
                    SFG_NOALLOC_MAX_NODES);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SFG_E_ARG, "snapshot_noalloc: device %d out of range", dev);
    if (dev != r->op->device)
        return fail(SFG_E_ARG, "snapshot_noalloc: run lives on device %d, current device is %d", r->op->device, dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        // first call on this device: resolve this GPU's copy of the array
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(SFG_E_HIP, "snapshot_noalloc: no device address for the staging array");
        g_noalloc_ptr[dev] = (double *)p;
    }
    double *dU = g_noalloc_ptr[dev], *dW = dU + n;
    int s = 0;
    while (r->perm[s] != shot) ++s;
    // a shot that stopped early holds its state in the buffer of its own last step
    const int done = std::min(r->executed, r->steps_int[s]);
    // every one of the 2n staged values is written by the gather: nothing to clear
    HIPCHK(launch_gather_uw(r->prec, r->state, r->uw, n, state_layout(r), s, done & 1, dU, dW, r->stream));
    HIPCHK(hipMemcpyAsync(U, dU, sizeof(double) * n, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipMemcpyAsync(W, dW, sizeof(double) * n, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return SFG_OK;
}

extern "C" int sfg_run_energy(sfg_run *r, double *energy, int max_checks)
{
    g_err.clear();
    if (!r || !energy || max_checks <= 0) return fail(SFG_E_ARG, "energy: bad arguments");
    HIPCHK(hipStreamSynchronize(r->stream));
    const int nchk = std::min(max_checks, r->nslots);
    std::vector<double> en((size_t)nchk * r->S * 2);
    HIPCHK(hipMemcpy(en.data(), r->energy, sizeof(double) * en.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < nchk; ++c)
        for (int s = 0; s < r->nshots; ++s)
            for (int k = 0; k < 2; ++k)
                energy[((size_t)c * r->nshots + r->perm[s]) * 2 + k] = en[((size_t)c * r->S + s) * 2 + k];
    return nchk;
}

// ---------------------------------------------------------------------------
// dt estimate (no counterpart in the reference: dt is user input, seisfem.c:144)

extern "C" int sfg_estimate_dt(const sfg_operator *op, const sfg_medium *md, int max_iter, double tol,
                               double *lambda_max, double *dt_max)
{
    g_err.clear();
    if (!op || !md || md->n != op->n || max_iter <= 0 || !lambda_max || !dt_max)
        return fail(SFG_E_ARG, "estimate_dt: bad arguments");
    const int n = op->n;
    const size_t n2 = 2 * (size_t)n;
    // deterministic start vector with every component non-zero
    std::vector<double> x0(n2);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    double nrm = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        x0[i] = 0.5 + (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
        nrm += x0[i] * x0[i];
    }
    for (auto &v : x0) v /= std::sqrt(nrm);
    double *x = nullptr, *y = nullptr, *acc = nullptr;
    int rc = dalloc(&x, n2);
    if (!rc) rc = dalloc(&y, n2);
    if (!rc) rc = dalloc(&acc, 1);
    if (rc) { dfree(x); dfree(y); dfree(acc); return rc; }
    hipError_t e = hipMemcpy(x, x0.data(), sizeof(double) * n2, hipMemcpyHostToDevice);
    double lam = 0.0, prev = 0.0;
    int stable = 0, it = 0;
    for (; it < max_iter && e == hipSuccess; ++it) {
        e = launch_fill_zero(acc, sizeof(double), nullptr);
        if (e == hipSuccess)
            e = launch_elastic_apply(n, op->rp, op->ci, op->v64, (size_t)op->nnz, md->coef64, md->flag, x, y, acc, nullptr);
        double ss = 0.0;
        if (e == hipSuccess) e = hipMemcpy(&ss, acc, sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) break;
        lam = std::sqrt(ss); // |A x| with |x| = 1
        if (!(lam > 0.0)) break;
        e = launch_scale(n2, y, acc, nullptr);
        std::swap(x, y);
        if (std::fabs(lam - prev) <= tol * lam) { if (++stable >= 5) break; } else stable = 0;
        prev = lam;
    }
    dfree(x); dfree(y); dfree(acc);
    if (e != hipSuccess) return fail(SFG_E_HIP, "estimate_dt: %s", hipGetErrorString(e));
    if (!(lam > 0.0)) return fail(SFG_E_ARG, "estimate_dt: operator is zero (all rows clamped?)");
    *lambda_max = lam;
    // the reference's Newmark variant (alpha = 1, delta = 1.5) is stable for omega*dt <= 2
    *dt_max = 2.0 / std::sqrt(lam);
    return it;
}

// ---------------------------------------------------------------------------
// multi-runs: several runs stepped by one kernel launch per step

// One launch of a multi-run step: the runs it steps, their argument table and
// the first tile of each (runs in order, tiles of a run = its row blocks x shot
// groups, row blocks fastest).
struct MultiGroup {
    std::vector<sfg_run *> runs;
    int ntiles = 0;
    int xcd_remap = 1;          // the group's tile map: 0 round-robin, 1 contiguous, Q + 1 stripes of Q
    size_t lds_bytes = 0;       // dynamic LDS of the launch (LDS-staged group)
    void *args = nullptr;       // device StepArgs<T>[runs]
    int *tile_start = nullptr;  // device [runs + 1]
};

struct sfg_multi {
    std::vector<sfg_run *> runs;
    int prec = 0, block = 256, max_steps = 0, guard_every = 50, executed = 0;
    Ctrl *ctrl = nullptr;       // shared step counter
    // Per step: each batched run (several shots per lane group) that fills the
    // chip on its own takes its own step launch (the single-run kernel, with its
    // compaction of finished shot groups), reading the shared counter; then one
    // launch of the smaller batched runs (grp[0]), one of the single-shot runs in
    // the per-row or sliced layouts (grp[1], a kernel with only those forms) and
    // one of the single-shot runs with LDS-staged halos (grp[2]), each only when
    // it has runs.
    std::vector<sfg_run *> solo;
    std::vector<StepArgs<double>> solo64;   // their arguments with the shared counter
    std::vector<StepArgs<float>> solo32;
    MultiGroup grp[3];
    // persistent schedule (sfg_multi_set_persistent): resident blocks of one
    // cooperative launch per guard interval, or 0 for the per-step launches
    int persist_blocks = 0;
    int persist_group = -1;
    int *queue = nullptr;       // device int[16]: tile queues by step parity and XCD
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // by parity of the first step
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

static void multi_free(sfg_multi *m)
{
    for (sfg_run *r : m->runs) r->in_multi = false;
    for (hipGraphExec_t g : m->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    dfree(m->ctrl);
    dfree(m->queue);
    for (MultiGroup &g : m->grp) {
        dfree(g.args);
        dfree(g.tile_start);
    }
    delete m;
}

static int multi_launch_step(sfg_multi *m, int slot)
{
    for (size_t k = 0; k < m->solo.size(); ++k) {
        const sfg_run *r = m->solo[k];
        const void *args = m->prec == 0 ? (const void *)&m->solo64[k] : (const void *)&m->solo32[k];
        HIPCHK(launch_step(r->prec, r->SP, r->cps, args, r->nbx, r->nby, r->block, slot, r->lds_bytes, m->stream));
    }
    for (int gi = 0; gi < 2; ++gi) {
        const MultiGroup &g = m->grp[gi];
        if (!g.runs.empty())
            HIPCHK(launch_multi_step(m->prec, g.args, g.tile_start, (int)g.runs.size(), g.ntiles, m->block,
                                     g.xcd_remap, slot, gi == 1, m->stream));
    }
    const MultiGroup &g2 = m->grp[2];
    if (!g2.runs.empty())
        HIPCHK(launch_multi_step_lds(m->prec, g2.args, g2.tile_start, (int)g2.runs.size(), g2.ntiles, m->block,
                                     g2.xcd_remap, slot, g2.lds_bytes, m->stream));
    return SFG_OK;
}

static int multi_launch_guards(sfg_multi *m)
{
    for (sfg_run *r : m->runs) {
        int rc = launch_guard_now(r, m->ctrl, m->stream);
        if (rc) return rc;
    }
    return SFG_OK;
}

extern "C" int sfg_multi_create(sfg_multi **out, sfg_run *const *runs, int nruns)
{
    g_err.clear();
    if (!out || !runs || nruns <= 0) return fail(SFG_E_ARG, "multi: bad arguments");
    *out = nullptr;
    const sfg_run *r0 = runs[0];
    for (int k = 0; k < nruns; ++k) {
        const sfg_run *r = runs[k];
        if (!r || r->in_multi) return fail(SFG_E_ARG, "multi: run %d is null or already in a multi-run", k);
        if (r->executed != 0) return fail(SFG_E_ARG, "multi: run %d has already stepped", k);
        if (r->prec != r0->prec || r->max_steps != r0->max_steps || r->block != r0->block ||
            r->guard_every != r0->guard_every)
            return fail(SFG_E_ARG, "multi: runs must share precision, max_steps, block and guard interval");
    }
    sfg_multi *m = new sfg_multi;
    m->prec = r0->prec;
    m->block = r0->block;
    m->max_steps = r0->max_steps;
    m->guard_every = r0->guard_every;
    // which launch steps each run (see sfg_multi)
    auto staged = [](const sfg_run *r) { return (r->prec == 0 ? r->a64.variant : r->a32.variant) == 8; };
    for (int k = 0; k < nruns; ++k) {
        sfg_run *r = runs[k];
        m->runs.push_back(r);
        if (r->SP == 1) m->grp[staged(r) ? 2 : 1].runs.push_back(r);
        else if ((r->SP == 32 || r->SP == 64) && r->nbx * r->nby < 256 && r->lds_bytes == 0) m->grp[0].runs.push_back(r);
        else m->solo.push_back(r);
    }
    int rc = SFG_OK;
    if ((rc = dalloc(&m->ctrl, 1))) {
        multi_free(m);
        return rc;
    }
    for (sfg_run *r : m->solo) {
        if (m->prec == 0) {
            m->solo64.push_back(r->a64);
            m->solo64.back().ctrl = m->ctrl;
        } else {
            m->solo32.push_back(r->a32);
            m->solo32.back().ctrl = m->ctrl;
        }
    }
    const size_t asz = m->prec == 0 ? sizeof(StepArgs<double>) : sizeof(StepArgs<float>);
    std::vector<std::vector<char>> tabs(3);
    std::vector<std::vector<int>> tss(3);
    for (int gi = 0; gi < 3; ++gi) {
        MultiGroup &g = m->grp[gi];
        const int n = (int)g.runs.size();
        if (n == 0) continue;
        {
            // the group's first run sets its tile map
            const sfg_run *f = g.runs[0];
            const int remap = f->prec == 0 ? f->a64.xcd_remap : f->a32.xcd_remap;
            const int stripe = f->prec == 0 ? f->a64.xcd_stripe : f->a32.xcd_stripe;
            g.xcd_remap = stripe > 0 ? stripe + 1 : remap;
        }
        std::vector<int> &ts = tss[gi];
        ts.assign(n + 1, 0);
        for (int k = 0; k < n; ++k) {
            ts[k + 1] = ts[k] + g.runs[k]->nbx * g.runs[k]->nby;
            g.lds_bytes = std::max(g.lds_bytes, g.runs[k]->lds_bytes);
        }
        g.ntiles = ts[n];
        if ((rc = dalloc((char **)&g.args, asz * n)) || (rc = dalloc(&g.tile_start, n + 1))) {
            multi_free(m);
            return rc;
        }
        // argument table: each run's own arguments with the shared step counter
        std::vector<char> &tab = tabs[gi];
        tab.resize(asz * n);
        for (int k = 0; k < n; ++k) {
            if (m->prec == 0) {
                StepArgs<double> a = g.runs[k]->a64;
                a.ctrl = m->ctrl;
                std::memcpy(tab.data() + asz * k, &a, asz);
            } else {
                StepArgs<float> a = g.runs[k]->a32;
                a.ctrl = m->ctrl;
                std::memcpy(tab.data() + asz * k, &a, asz);
            }
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess && !m->grp[2].runs.empty()) e = allow_multi_step_lds(m->prec, m->grp[2].lds_bytes);
    if (e == hipSuccess) e = hipEventCreate(&m->ev0);
    if (e == hipSuccess) e = hipEventCreate(&m->ev1);
    if (e == hipSuccess) e = launch_fill_zero(m->ctrl, sizeof(Ctrl), m->stream);
    for (int gi = 0; gi < 3; ++gi) {
        if (m->grp[gi].runs.empty()) continue;
        if (e == hipSuccess)
            e = hipMemcpyAsync(m->grp[gi].args, tabs[gi].data(), tabs[gi].size(), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(m->grp[gi].tile_start, tss[gi].data(), sizeof(int) * tss[gi].size(),
                               hipMemcpyHostToDevice, m->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        multi_free(m);
        return fail(SFG_E_HIP, "multi setup: %s", hipGetErrorString(e));
    }
    // graphs: guard_every launches, then every run's energy guard (per start parity)
    for (int p = 0; p < (m->guard_every % 2 ? 2 : 1) && rc == SFG_OK; ++p)
        rc = capture_graph(m->stream, &m->graph[p], [&]() {
            int c = SFG_OK;
            for (int k = 0; k < m->guard_every && c == SFG_OK; ++k) c = multi_launch_step(m, (p + k) & 1);
            return c == SFG_OK ? multi_launch_guards(m) : c;
        });
    if (rc != SFG_OK) {
        multi_free(m);
        return rc;
    }
    for (sfg_run *r : m->runs) r->in_multi = true;
    *out = m;
    return SFG_OK;
}

extern "C" int sfg_multi_advance(sfg_multi *m, int nsteps)
{
    g_err.clear();
    if (!m || nsteps < 0) return fail(SFG_E_ARG, "multi advance: bad arguments");
    int todo = std::min(nsteps, m->max_steps - m->executed);
    const int steps = todo;
    HIPCHK(hipEventRecord(m->ev0, m->stream));
    while (todo > 0 && m->persist_blocks > 0) {
        // one cooperative launch up to the next guard
        const MultiGroup &g = m->grp[m->persist_group];
        const int seg = std::min(todo, m->guard_every - m->executed % m->guard_every);
        HIPCHK(launch_multi_persistent(m->prec, g.args, g.tile_start, (int)g.runs.size(), g.ntiles, m->block,
                                       m->persist_blocks, m->executed & 1, seg, m->queue, m->persist_group == 1,
                                       m->stream));
        m->executed += seg;
        todo -= seg;
        if (m->executed % m->guard_every == 0 || m->executed == m->max_steps) {
            int rc = multi_launch_guards(m);
            if (rc) return rc;
        }
    }
    while (todo > 0) {
        hipGraphExec_t g = m->graph[m->executed & 1];
        if (g && todo >= m->guard_every && m->executed % m->guard_every == 0) {
            HIPCHK(hipGraphLaunch(g, m->stream));
            m->executed += m->guard_every;
            todo -= m->guard_every;
        } else {
            int rc = multi_launch_step(m, m->executed & 1);
            if (rc) return rc;
            m->executed += 1;
            todo -= 1;
            if (m->executed % m->guard_every == 0 || m->executed == m->max_steps) {
                rc = multi_launch_guards(m);
                if (rc) return rc;
            }
        }
    }
    HIPCHK(hipEventRecord(m->ev1, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_ms = steps > 0 ? (double)ms / steps : 0.0;
    for (sfg_run *r : m->runs) {
        r->executed = m->executed;
        int rc = check_guard(r, m->executed, m->ctrl);
        if (rc) return rc;
    }
    return m->executed;
}

extern "C" int sfg_multi_set_persistent(sfg_multi *m, int blocks_per_cu)
{
    g_err.clear();
    if (!m || blocks_per_cu < 0) return fail(SFG_E_ARG, "multi persistent: bad arguments");
    if (blocks_per_cu == 0) {
        m->persist_blocks = 0;
        return SFG_OK;
    }
    int gi = -1, groups = 0;
    for (int k = 0; k < 3; ++k)
        if (!m->grp[k].runs.empty()) {
            gi = k;
            ++groups;
        }
    if (!m->solo.empty() || groups != 1 || gi == 2)
        return fail(SFG_E_ARG, "multi persistent: needs every run in one shared launch (single-shot runs in the "
                               "per-row or sliced layouts, or batched runs too small to fill the GPU alone)");
    int per_cu = 0, dev = 0, cus = 0;
    HIPCHK(multi_persistent_blocks_per_cu(m->prec, gi == 1, m->block, &per_cu));
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (blocks_per_cu > per_cu)
        return fail(SFG_E_ARG, "multi persistent: %d blocks per CU requested, %d fit", blocks_per_cu, per_cu);
    if (!m->queue) {
        int rc = dalloc(&m->queue, 16);
        if (rc) return rc;
        HIPCHK(launch_fill_zero(m->queue, 16 * sizeof(int), m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    }
    m->persist_group = gi;
    m->persist_blocks = blocks_per_cu * cus;
    return SFG_OK;
}

extern "C" double sfg_multi_last_ms_per_step(const sfg_multi *m) { return m ? m->last_ms : 0.0; }

extern "C" double sfg_multi_step_bytes(const sfg_multi *m, int step)
{
    if (!m) return 0.0;
    double b = 0.0;
    for (const sfg_run *r : m->runs) {
        int act = 0;
        for (int s = 0; s < r->nshots; ++s) act += r->steps_int[s] > step ? 1 : 0;
        if (act) b += sfg_run_step_bytes(r, act);
    }
    return b;
}

extern "C" void sfg_multi_destroy(sfg_multi *m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    multi_free(m);
}

// ---------------------------------------------------------------------------
// HBM yardstick (no counterpart in the reference)

extern "C" int sfg_stream_copy(size_t bytes, int reps, double *gbs, double *ms_per_copy)
{
    g_err.clear();
    if (bytes < 16 || bytes % 16 || reps <= 0 || !gbs || !ms_per_copy)
        return fail(SFG_E_ARG, "stream_copy: bytes must be a positive multiple of 16, reps > 0");
    char *a = nullptr, *b = nullptr;
    int rc = dalloc(&a, bytes);
    if (!rc) rc = dalloc(&b, bytes);
    if (rc) { dfree(a); dfree(b); return rc; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_fill_zero(a, bytes, nullptr);
    // two untimed copies, then `reps` timed ones
    for (int w = 0; w < 2 && e == hipSuccess; ++w) e = launch_stream_copy(a, b, bytes / 16, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stream_copy(a, b, bytes / 16, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    if (e != hipSuccess) return fail(SFG_E_HIP, "stream_copy: %s", hipGetErrorString(e));
    *ms_per_copy = (double)ms / reps;
    *gbs = 2.0 * (double)bytes / (*ms_per_copy * 1e-3) / 1e9;
    return SFG_OK;
}
