// Host side of libscp3g: the C ABI declared in include/scp3g.h.
//
// It owns the device copies of the system (x0, omega, MUa, Sobol direction
// integers, basis row descriptors) and the running FP64 sum, and drives the
// per-tile kernel chain that replaces SCP3_a's block loop
// (SCP3.cpp:187-232): sobol -> map -> pes -> weight -> basis -> syrk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstring>
#include <new>

#include "scp3g_internal.h"

namespace scp3g {

hipError_t launch_qtip4pf(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                          int fp32, double *E);
hipError_t launch_ttm3f(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                        int fp32, double *E);
void release_tiles(scp3g_ctx *ctx);

int fail(scp3g_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int check_hip(scp3g_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    return fail(ctx, -100 - (int)e, std::string(what) + ": " + hipGetErrorString(e));
}

static hipEvent_t pool_event(scp3g_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t ev = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    hipEventCreate(&ev);
    return ev;
}

void timer_begin(scp3g_ctx *ctx, const char *, hipEvent_t *ev0)
{
    *ev0 = nullptr;
    if (!ctx->timing) return;
    *ev0 = pool_event(ctx);
    hipEventRecord(*ev0, ctx->stream);
}

static void timer_drain(scp3g_ctx *ctx, KernelTimer &t)
{
    for (auto &pr : t.pending) {
        float ms = 0.f;
        hipEventSynchronize(pr.second);
        hipEventElapsedTime(&ms, pr.first, pr.second);
        t.total_ms += ms;
        t.launches += 1;
        ctx->event_pool.push_back(pr.first);
        ctx->event_pool.push_back(pr.second);
    }
    t.pending.clear();
}

void timer_end(scp3g_ctx *ctx, const char *name, hipEvent_t ev0)
{
    if (!ctx->timing || !ev0) return;
    hipEvent_t ev1 = pool_event(ctx);
    hipEventRecord(ev1, ctx->stream);
    KernelTimer &t = ctx->timers[name];
    t.pending.emplace_back(ev0, ev1);
    if (t.pending.size() > 256) timer_drain(ctx, t);
}

}  // namespace scp3g

using namespace scp3g;

// ---- helpers ------------------------------------------------------------

static int64_t basis_size(int nsel, int n2, int n3)
{
    return 1 + (int64_t)nsel + (int64_t)n2 * (n2 + 1) / 2 +
           ((int64_t)n3 * (1 + n3) * (2 + n3)) / 6;
}

// get_bra's loop nest (ScaledMatrixElements.cpp:870-930) as row descriptors
static std::vector<RowDesc> build_desc(int nsel, int n2, int n3)
{
    std::vector<RowDesc> d;
    auto add = [&](int m0, int g0, int m1, int g1, int m2, int g2) {
        RowDesc r;
        r.f[0] = pack_factor(m0, g0);
        r.f[1] = pack_factor(m1, g1);
        r.f[2] = pack_factor(m2, g2);
        d.push_back(r);
    };
    add(0, 0, 0, 0, 0, 0);                                          // 1
    for (int m = 0; m < nsel; ++m) add(m, 1, 0, 0, 0, 0);            // h1(q_m)
    for (int m = 0; m < n2; ++m) add(m, 2, 0, 0, 0, 0);              // h2(q_m)
    for (int m = 0; m < n2; ++m)
        for (int n = m + 1; n < n2; ++n) add(m, 1, n, 1, 0, 0);      // h1 h1
    for (int m = 0; m < n3; ++m) add(m, 3, 0, 0, 0, 0);              // h3(q_m)
    for (int m = 0; m < n3; ++m)
        for (int n = m + 1; n < n3; ++n) {
            add(m, 2, n, 1, 0, 0);                                    // h2 h1
            add(m, 1, n, 2, 0, 0);                                    // h1 h2
        }
    for (int m = 0; m < n3; ++m)
        for (int n = m + 1; n < n3; ++n)
            for (int p = n + 1; p < n3; ++p) add(m, 1, n, 1, p, 1);   // h1 h1 h1
    return d;
}

// HO diagonal (addHODiagonal, ScaledMatrixElements.cpp:735-784) plus the
// zero-point energy of unselected modes (SCP3.cpp:238,270).
std::vector<double> scp3g::ho_diagonal(const scp3g_ctx *ctx)
{
    const int nsel = (int)ctx->modes.size();
    std::vector<double> om(nsel);
    for (int k = 0; k < nsel; ++k) om[k] = ctx->omega[ctx->modes[k]];
    double E0 = 0.0;
    for (double o : om) E0 += o;
    E0 *= 0.5;
    double sum_all = 0.0, sum_sel = 0.0;
    for (double o : ctx->omega) sum_all += o;
    for (double o : om) sum_sel += o;
    const double zpe_rest = 0.5 * (sum_all - sum_sel);
    std::vector<double> dg;
    dg.reserve((size_t)ctx->nbasis);
    dg.push_back(E0);
    for (int k = 0; k < nsel; ++k) dg.push_back(E0 + om[k]);
    for (int k = 0; k < ctx->nmodes2; ++k) dg.push_back(E0 + 2 * om[k]);
    for (int k = 0; k < ctx->nmodes2; ++k)
        for (int l = k + 1; l < ctx->nmodes2; ++l) dg.push_back(E0 + om[k] + om[l]);
    for (int k = 0; k < ctx->nmodes3; ++k) dg.push_back(E0 + 3.0 * om[k]);
    for (int k = 0; k < ctx->nmodes3; ++k)
        for (int l = k + 1; l < ctx->nmodes3; ++l) {
            dg.push_back(E0 + 2.0 * om[k] + om[l]);
            dg.push_back(E0 + om[k] + 2.0 * om[l]);
        }
    for (int k = 0; k < ctx->nmodes3; ++k)
        for (int l = k + 1; l < ctx->nmodes3; ++l)
            for (int j = l + 1; j < ctx->nmodes3; ++j) dg.push_back(E0 + om[k] + om[l] + om[j]);
    for (double &v : dg) v += zpe_rest;
    return dg;
}

template <typename T>
static void dfree(T *&p)
{
    if (p) hipFree((void *)p);
    p = nullptr;
}

static void free_work(scp3g_ctx *ctx)
{
    dfree(ctx->d_y);
    dfree(ctx->d_crd);
    dfree(ctx->d_E);
    dfree(ctx->d_w);
    dfree(ctx->d_B);
    dfree(ctx->d_G);
    dfree(ctx->d_Vq);
    dfree(ctx->d_wp);
    dfree(ctx->d_wm);
    dfree(ctx->d_sc);
    ctx->work_tile = 0;
    ctx->work_n = 0;
    ctx->work_precision = -1;
    ctx->work_estimator = -1;
}

static void free_sums(scp3g_ctx *ctx)
{
    dfree(ctx->d_H);
    dfree(ctx->d_Hc);
    dfree(ctx->d_F);
    dfree(ctx->d_Heff);
    dfree(ctx->d_mout);
    dfree(ctx->d_diag);
    dfree(ctx->d_offset);
    ctx->h_n = 0;
    ctx->points_done = 0;
    ctx->norm_sum = 0.0;
}

// Points per tile when the caller sets none: per-point kernels need >= 64k
// points to fill 256 CUs and the split-K SYRK partials amortise over the
// tile, so take the largest power of two <= 2^19 whose FP64 basis table
// (rows_pad128 x T doubles) fits in 8 GiB (HBM is 288 GB), but at least
// 4096.  Both forms use the same tile; the mixed form allocates only its
// two BF16 planes.  (1225 states, 2^20 points: 116, 124, 128, 128 M
// points/s at 2^17, 2^18, 2^19, 2^20; 20825 states: the mixed SYRK is 1.8%
// slower per point at 65536 than at 32768.)
static int64_t auto_tile(int64_t n)
{
    const double rows = (double)rows_pad128(n) + 32;
    int64_t T = 1 << 19;
    while (T > 4096 && 8.0 * rows * (double)T > 8.0 * (1 << 30)) T >>= 1;
    return T;
}

static int ensure_work(scp3g_ctx *ctx)
{
    const int64_t n = ctx->nbasis;
    const int64_t Tw = ctx->tile > 0 ? ctx->tile : auto_tile(n);
    const int grad = ctx->estimator != SCP3G_EST_ENERGY;
    if (ctx->work_tile == Tw && ctx->work_n == n && ctx->work_precision == ctx->precision &&
        (ctx->work_estimator != SCP3G_EST_ENERGY) == grad && ctx->work_estimator >= 0 && ctx->d_B)
        return 0;
    free_work(ctx);
    const size_t T = (size_t)Tw;
    const size_t rows = (size_t)rows_pad128(n);  // FP64 table rows (zero padded)
    int rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_y, sizeof(double) * ctx->nmodes0 * T), "alloc y")))
        return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_crd, sizeof(double) * 9 * ctx->nw * T), "alloc crd")))
        return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_E, sizeof(double) * T), "alloc E"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_w, sizeof(double) * T), "alloc w"))) return rc;
    // the FP64 table (rows x T doubles) in the FP64 form; in the mixed form
    // the two BF16 planes of X (rows padded to 256) and the sign words, which
    // also hold the direct-load kernels' FP32 table (rows x T floats)
    const size_t rows_f = (size_t)fast_rows_pad(n);
    // (the gradient estimator's mixed form holds the phi + s Psi and phi - s Psi
    // planes side by side: twice the points)
    const size_t bbytes = ctx->precision == SCP3G_FP64 ? sizeof(double) * rows * T
                          : grad ? 8 * rows_f * T + T / 4 + 256
                                 : 4 * rows_f * T + T / 8 + 256;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_B, bbytes), "alloc B"))) return rc;
    if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_B, 0, bbytes, ctx->stream), "clear B"))) return rc;
    if (grad) {
        // gradient estimator: dE/dx, Vq over the selected modes, weights, scale
        const size_t nsel = ctx->modes.size();
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_G, sizeof(double) * 9 * ctx->nw * T), "alloc G")))
            return rc;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_Vq, sizeof(double) * nsel * T), "alloc Vq"))) return rc;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_wp, sizeof(double) * T), "alloc wp"))) return rc;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_wm, sizeof(double) * T), "alloc wm"))) return rc;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_sc, sizeof(double) * T), "alloc sc"))) return rc;
    }
    ctx->work_estimator = ctx->estimator;
    ctx->work_precision = ctx->precision;
    ctx->work_tile = (int)Tw;
    ctx->work_n = n;
    return 0;
}

static int ensure_sums(scp3g_ctx *ctx)
{
    const int64_t n = ctx->nbasis;
    const bool grad = ctx->estimator != SCP3G_EST_ENERGY;
    int rc;
    if (grad && (rc = grad_setup(ctx))) return rc;
    if (ctx->h_n == n && ctx->d_H && ctx->d_Hc && (!grad || (ctx->d_F && ctx->d_Heff))) return 0;
    free_sums(ctx);
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_H, bytes), "alloc H"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_Hc, bytes), "alloc Hc"))) return rc;
    if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_H, 0, bytes, ctx->stream), "clear H"))) return rc;
    if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_Hc, 0, bytes, ctx->stream), "clear Hc")))
        return rc;
    if (grad) {
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_Heff, bytes), "alloc Heff"))) return rc;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_F, sizeof(double) * ctx->ngfix), "alloc F"))) return rc;
        if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_F, 0, sizeof(double) * ctx->ngfix, ctx->stream),
                            "clear F")))
            return rc;
    }
    ctx->h_n = n;
    ctx->points_done = 0;
    ctx->norm_sum = 0.0;
    return 0;
}

static int ready(scp3g_ctx *ctx)
{
    if (!ctx) return -1;
    if (ctx->nw <= 0 || ctx->nmodes0 <= 0) return fail(ctx, -2, "system not set (scp3g_set_system)");
    if (ctx->nbasis <= 0) return fail(ctx, -2, "basis not set (scp3g_set_basis)");
    return 0;
}

// ---- C ABI ----------------------------------------------------------------

extern "C" {

int scp3g_create(scp3g_ctx **out, int device)
{
    if (!out) return -1;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -3;
    if (device < 0 || device >= ndev) return -3;
    scp3g_ctx *ctx = new (std::nothrow) scp3g_ctx();
    if (!ctx) return -2;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return -3;
    }
    // the static arrays of scp3g_sobol_normals_noalloc, once per device
    if (sobol_static_setup(device) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return -3;
    }
    *out = ctx;
    return 0;
}

void scp3g_destroy(scp3g_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->timers)
        for (auto &pr : kv.second.pending) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
    for (hipEvent_t ev : ctx->event_pool) hipEventDestroy(ev);
    free_work(ctx);
    free_sums(ctx);
    grad_release(ctx);
    scp1_release(ctx);
    dfree(ctx->d_part);
    dfree(ctx->d_mono);
    dfree(ctx->d_sobol);
    dfree(ctx->d_x0);
    dfree(ctx->d_omega);
    dfree(ctx->d_MUa);
    dfree(ctx->d_modes);
    dfree(ctx->d_desc);
    release_tiles(ctx);
    mlp_release(ctx);
    release_fast_tiles(ctx);
    release_tiles128(ctx);
    release_davidson(ctx);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *scp3g_last_error(const scp3g_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int scp3g_sync(scp3g_ctx *ctx)
{
    if (!ctx) return -1;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "stream sync");
}

int scp3g_set_system(scp3g_ctx *ctx, int nw, const double *x0, int nmodes0,
                     const double *omega, const double *MUa)
{
    if (!ctx) return -1;
    if (nw <= 0 || nw > 64) return fail(ctx, -4, "nw must be in 1..64");
    if (nmodes0 <= 0 || nmodes0 > 9 * nw) return fail(ctx, -4, "nmodes0 must be in 1..9nw");
    if (nmodes0 > 1111) return fail(ctx, -4, "at most 1111 Sobol dimensions (sobol.cpp:14292)");
    if (!x0 || !omega || !MUa) return fail(ctx, -4, "null input");
    hipSetDevice(ctx->device);
    const int n3 = 9 * nw;
    ctx->nw = nw;
    ctx->nmodes0 = nmodes0;
    ctx->x0.assign(x0, x0 + n3);
    ctx->omega.assign(omega, omega + nmodes0);
    ctx->MUa.assign(MUa, MUa + (size_t)n3 * nmodes0);
    std::vector<uint64_t> dirs((size_t)nmodes0 * kSobolCols + nmodes0, 0);
    sobol_directions(nmodes0, dirs.data());
    if (ctx->scramble_on)
        sobol_scramble(nmodes0, ctx->scramble, dirs.data(), dirs.data() + (size_t)nmodes0 * kSobolCols);
    dfree(ctx->d_sobol);
    dfree(ctx->d_x0);
    dfree(ctx->d_omega);
    dfree(ctx->d_MUa);
    int rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_sobol, dirs.size() * 8), "alloc sobol"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_x0, n3 * 8), "alloc x0"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_omega, nmodes0 * 8), "alloc omega"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_MUa, ctx->MUa.size() * 8), "alloc MUa"))) return rc;
    hipMemcpy(ctx->d_sobol, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice);
    hipMemcpy(ctx->d_x0, x0, n3 * 8, hipMemcpyHostToDevice);
    hipMemcpy(ctx->d_omega, omega, nmodes0 * 8, hipMemcpyHostToDevice);
    rc = check_hip(ctx, hipMemcpy(ctx->d_MUa, MUa, ctx->MUa.size() * 8, hipMemcpyHostToDevice),
                   "upload system");
    if (rc) return rc;
    // the work buffers depend on nw / nmodes0, the running sums on the
    // system; the basis stays only if every selected mode still exists
    // (a mode list keeps the caller's order, so check each entry)
    free_work(ctx);
    free_sums(ctx);
    grad_release(ctx);
    bool keep = !ctx->modes.empty();
    for (int m : ctx->modes) keep = keep && m >= 0 && m < nmodes0;
    if (!keep) {
        ctx->modes.clear();
        ctx->nbasis = 0;
    }
    return 0;
}

int scp3g_set_basis(scp3g_ctx *ctx, int nsel, const int *modes, int nmodes2, int nmodes3)
{
    if (!ctx) return -1;
    if (ctx->nmodes0 <= 0) return fail(ctx, -2, "system not set (scp3g_set_system)");
    if (nsel <= 0 || !modes) return fail(ctx, -4, "no modes selected");
    // SCP3.cpp:156-160
    if (nmodes2 < 0 || nmodes3 < 0 || nmodes2 > nsel || nmodes3 > nsel)
        return fail(ctx, -4, "Number of double or triple excitations exceed number of modes.");
    for (int k = 0; k < nsel; ++k)
        if (modes[k] < 0 || modes[k] >= ctx->nmodes0)
            return fail(ctx, -4, "selected mode out of range");
    if (nsel > 0xffff) return fail(ctx, -4, "too many selected modes");
    hipSetDevice(ctx->device);
    ctx->modes.assign(modes, modes + nsel);
    ctx->nmodes2 = nmodes2;
    ctx->nmodes3 = nmodes3;
    ctx->nbasis = basis_size(nsel, nmodes2, nmodes3);
    ctx->desc = build_desc(nsel, nmodes2, nmodes3);
    if ((int64_t)ctx->desc.size() != ctx->nbasis) return fail(ctx, -5, "R != basis size");
    dfree(ctx->d_modes);
    dfree(ctx->d_desc);
    int rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_modes, nsel * sizeof(int)), "alloc modes"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_desc, ctx->desc.size() * sizeof(RowDesc)), "alloc desc")))
        return rc;
    hipMemcpy(ctx->d_modes, modes, nsel * sizeof(int), hipMemcpyHostToDevice);
    rc = check_hip(ctx, hipMemcpy(ctx->d_desc, ctx->desc.data(), ctx->desc.size() * sizeof(RowDesc),
                                  hipMemcpyHostToDevice),
                   "upload basis");
    if (rc) return rc;
    free_work(ctx);
    free_sums(ctx);
    grad_release(ctx);
    return 0;
}

int scp3g_basis_size(const scp3g_ctx *ctx, int64_t *n)
{
    if (!ctx || !n) return -1;
    *n = ctx->nbasis;
    return 0;
}

int scp3g_sub_basis_size(const scp3g_ctx *ctx, int level, int64_t *n)
{
    if (!ctx || !n) return -1;
    const int64_t nsel = (int64_t)ctx->modes.size();
    const int64_t n2 = ctx->nmodes2, n3 = ctx->nmodes3;
    int64_t s = 0;
    if (level >= 3) s += (n3 * (1 + n3) * (2 + n3)) / 6;
    if (level >= 2) s += n2 * (n2 + 1) / 2;
    if (level >= 1) s += nsel;
    if (level >= 0) s += 1;
    *n = s;
    return 0;
}

int scp3g_set_pes(scp3g_ctx *ctx, const char *name)
{
    if (!ctx) return -1;
    if (!name) return fail(ctx, -4, "null PES name");
    if (!strcmp(name, "qtip4pf")) ctx->pes = SCP3G_PES_QTIP4PF;
    else if (!strcmp(name, "ttm3f")) ctx->pes = SCP3G_PES_TTM3F;
    else return fail(ctx, -4, std::string("Unknown H2O potential: ") + name);
    return 0;
}

int scp3g_set_precision(scp3g_ctx *ctx, int mode)
{
    if (!ctx) return -1;
    if (mode != SCP3G_FP64 && mode != SCP3G_MIXED) return fail(ctx, -4, "bad precision mode");
    if (mode != ctx->precision) {
        ctx->precision = mode;
        free_sums(ctx);
    }
    return 0;
}

int scp3g_set_mixed_kernel(scp3g_ctx *ctx, int fast)
{
    if (!ctx) return -1;
    if (fast < 0 || fast > 1) return fail(ctx, -4, "mixed kernel must be 0 or 1");
    ctx->mixed_fast = fast != 0;
    return 0;
}

int scp3g_set_cv_terms(scp3g_ctx *ctx, int terms)
{
    if (!ctx) return -1;
    if (terms != 1 && terms != 3) return fail(ctx, -4, "terms must be 1 or 3");
    ctx->cv_one_term = terms == 1;
    return 0;
}

int scp3g_set_estimator(scp3g_ctx *ctx, int estimator)
{
    if (!ctx) return -1;
    if (estimator != SCP3G_EST_ENERGY && estimator != SCP3G_EST_GRADIENT &&
        estimator != SCP3G_EST_GRADIENT_STEIN)
        return fail(ctx, -4, "bad estimator");
    if (estimator != ctx->estimator) {
        hipSetDevice(ctx->device);
        ctx->estimator = estimator;
        free_sums(ctx);
        grad_release(ctx);
    }
    return 0;
}

int scp3g_set_scramble(scp3g_ctx *ctx, int scrambled, uint64_t seed)
{
    if (!ctx) return -1;
    if (ctx->nmodes0 <= 0) return fail(ctx, -2, "system not set (scp3g_set_system)");
    hipSetDevice(ctx->device);
    const int M = ctx->nmodes0;
    std::vector<uint64_t> dirs((size_t)M * kSobolCols + M, 0);
    sobol_directions(M, dirs.data());
    if (scrambled) sobol_scramble(M, seed, dirs.data(), dirs.data() + (size_t)M * kSobolCols);
    int rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!rc)
        rc = check_hip(ctx, hipMemcpy(ctx->d_sobol, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice),
                       "upload scrambled directions");
    if (!rc) {
        ctx->scramble_on = scrambled != 0;
        ctx->scramble = seed;
    }
    return rc;
}

int scp3g_set_tile(scp3g_ctx *ctx, int tile)
{
    if (!ctx) return -1;
    if (tile < 0 || tile % 256) return fail(ctx, -4, "tile must be 0 or a positive multiple of 256");
    ctx->tile = tile;
    free_work(ctx);
    return 0;
}

// The running sum, normalised, becomes a fixed term of every later Mout and
// level solve, and the running sum restarts at zero (a control variate's
// correction term C / P_t under a surrogate term grown checkpoint by
// checkpoint).  Energy estimator only.
int scp3g_hold_as_offset(scp3g_ctx *ctx)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (ctx->estimator != SCP3G_EST_ENERGY) return fail(ctx, -4, "a fixed term needs the energy (-G) estimator");
    if (!ctx->d_H || ctx->points_done <= 0 || !(ctx->norm_sum > 0.0)) return fail(ctx, -6, "no points accumulated");
    hipSetDevice(ctx->device);
    const int64_t n = ctx->nbasis;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    // this sum / its normaliser + the earlier fixed term (if any), upper
    // triangle, into the Mout staging buffer, which then becomes the fixed
    // term (and the earlier one the staging buffer)
    if (!ctx->d_mout && (rc = check_hip(ctx, hipMalloc(&ctx->d_mout, bytes), "alloc Mout"))) return rc;
    if ((rc = check_hip(ctx, launch_finalize(ctx, ctx->d_H, n, ctx->norm_sum, nullptr, ctx->d_mout, ctx->d_offset),
                        "offset")))
        return rc;
    std::swap(ctx->d_offset, ctx->d_mout);
    return scp3g_reset(ctx);
}

int scp3g_clear_offset(scp3g_ctx *ctx)
{
    if (!ctx) return -1;
    hipSetDevice(ctx->device);
    dfree(ctx->d_offset);
    return 0;
}

int scp3g_reset(scp3g_ctx *ctx)
{
    int rc = ready(ctx);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    const size_t bytes = sizeof(double) * (size_t)ctx->nbasis * (size_t)ctx->nbasis;
    if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_H, 0, bytes, ctx->stream), "clear H"))) return rc;
    if ((rc = check_hip(ctx, hipMemsetAsync(ctx->d_Hc, 0, bytes, ctx->stream), "clear Hc")))
        return rc;
    if (ctx->d_F &&
        (rc = check_hip(ctx, hipMemsetAsync(ctx->d_F, 0, sizeof(double) * ctx->ngfix, ctx->stream), "clear F")))
        return rc;
    ctx->points_done = 0;
    ctx->norm_sum = 0.0;
    return 0;
}

// one tile: y already in ctx->d_y (nmodes0 x tile, halved), np points.
// wmode 0: plain weights bV (SCP3.cpp:219-223); 1: surrogate weights w_nn;
// 2: control-variate correction bV - w_nn; 3: plain weights from energies
// the caller put in ctx->d_E; 5: the correction bV - w_nn as a sum of its
// own.  Modes 1, 2 and 5 multiply the weights by `scale`; the normaliser
// grows by np for modes 0, 3 and 5, np*scale for 1.  When the surrogate
// term takes one BF16 product (cv_one_term), the correction is bV with
// three products less w_nn with that same one product, so that the
// surrogate's rounding cancels in expectation and the estimator keeps the
// three-product form's precision.
}  // extern "C"

// The gradient estimator's tile (SCP3.cpp:205-231 with use_gradient):
// map, energy and gradient of the device potential (or the caller's, given
// = true), bV, Vq, the SYRK pair of phi +- s Psi (two FP64 launches, or one
// split-BF16 launch over both halves in the mixed form), and the block
// corrections and fixed elements.
static int run_tile_grad(scp3g_ctx *ctx, int64_t np, bool given)
{
    const int64_t T = ctx->work_tile;
    hipEvent_t ev;
    int rc;
    if (given) goto weights;  // the caller's energies and gradients are in d_E, d_G
    timer_begin(ctx, "map", &ev);
    rc = check_hip(ctx, launch_map(ctx, ctx->d_y, T, np, ctx->d_crd, T, 0), "map");
    timer_end(ctx, "map", ev);
    if (rc) return rc;
    timer_begin(ctx, "pes", &ev);
    if (ctx->pes == SCP3G_PES_TTM3F)
        rc = check_hip(ctx, launch_ttm3f_grad(ctx, ctx->nw, np, (const double *)ctx->d_crd, T, ctx->d_E,
                                              ctx->d_G, T),
                       "pes gradient");
    else
        rc = check_hip(ctx, launch_qtip4pf_grad(ctx, ctx->nw, np, (const double *)ctx->d_crd, T, ctx->d_E,
                                                ctx->d_G, T),
                       "pes gradient");
    timer_end(ctx, "pes", ev);
    if (rc) return rc;
weights:
    timer_begin(ctx, "weight", &ev);
    rc = check_hip(ctx, launch_weight(ctx, ctx->d_E, ctx->d_y, T, np, ctx->d_w), "weight");
    if (!rc)
        rc = check_hip(ctx, launch_vq(ctx, ctx->d_G, T, ctx->d_y, T, ctx->d_w, np, ctx->d_Vq, T, ctx->d_wp,
                                      ctx->d_wm, ctx->d_sc),
                       "vq");
    timer_end(ctx, "weight", ev);
    if (rc) return rc;
    if (ctx->precision == SCP3G_MIXED) {
        // the pair phi +- s Psi as one signed split-BF16 SYRK over 2 Kh slots
        const int64_t Kh = (np + 63) / 64 * 64;
        const int64_t rows_pad = fast_rows_pad(ctx->nbasis);
        unsigned short *Xh = reinterpret_cast<unsigned short *>(ctx->d_B);
        unsigned short *Xl = Xh + (size_t)rows_pad * 2 * T;
        unsigned *sgn = reinterpret_cast<unsigned *>(Xl + (size_t)rows_pad * 2 * T);
        timer_begin(ctx, "basis", &ev);
        rc = check_hip(ctx, launch_grad_split(ctx, ctx->d_y, T, np, Kh, ctx->d_w, ctx->d_Vq, T, ctx->d_sc,
                                              ctx->d_wp, Xh, Xl, sgn),
                       "grad split");
        timer_end(ctx, "basis", ev);
        if (rc) return rc;
        timer_begin(ctx, "syrk", &ev);
        rc = check_hip(ctx, launch_syrk_x3(ctx, ctx->nbasis, Xh, Xl, sgn, 2 * Kh, ctx->d_H, ctx->d_Hc, 3),
                       "syrk");
        timer_end(ctx, "syrk", ev);
        if (rc) return rc;
        goto corrections;
    }
    {
    double *B = (double *)ctx->d_B;
    for (int pass = 0; pass < 2; ++pass) {
        timer_begin(ctx, "basis", &ev);
        rc = check_hip(ctx, launch_grad_table(ctx, ctx->d_y, T, np, ctx->d_w, ctx->d_Vq, T, ctx->d_sc,
                                              pass ? -1.0 : 1.0, B, T),
                       "grad table");
        timer_end(ctx, "basis", ev);
        if (rc) return rc;
        timer_begin(ctx, "syrk", &ev);
        rc = check_hip(ctx, launch_syrk_f64_lds(ctx, ctx->nbasis, B, T, pass ? ctx->d_wm : ctx->d_wp, np,
                                                ctx->d_H),
                       "syrk");
        timer_end(ctx, "syrk", ev);
        if (rc) return rc;
    }
    }
corrections:
    timer_begin(ctx, "grad_blocks", &ev);
    rc = check_hip(ctx, launch_grad_corrections(ctx, ctx->d_y, T, np, ctx->d_w, ctx->d_Vq, T, ctx->d_H,
                                                ctx->d_F),
                   "grad corrections");
    timer_end(ctx, "grad_blocks", ev);
    if (rc) return rc;
    ctx->points_done += np;
    ctx->norm_sum += (double)np;
    return 0;
}

int scp3g::read_sums(scp3g_ctx *ctx, const double **H)
{
    if (ctx->estimator == SCP3G_EST_ENERGY) {
        *H = ctx->d_H;
        return 0;
    }
    int rc = ensure_sums(ctx);
    if (rc) return rc;
    rc = check_hip(ctx, launch_grad_effective(ctx, ctx->d_H, ctx->d_F, ctx->d_Heff), "grad effective");
    *H = ctx->d_Heff;
    return rc;
}

int scp3g::run_tile(scp3g_ctx *ctx, int64_t np, int wmode, double scale)
{
    const int64_t T = ctx->work_tile;
    const int fp32 = ctx->precision == SCP3G_MIXED;
    // a correction term matched to a one-product surrogate term
    const bool matched = (wmode == 2 || wmode == 5) && fp32 && ctx->mixed_fast && ctx->cv_one_term;
    hipEvent_t ev;
    int rc;
    if (ctx->estimator != SCP3G_EST_ENERGY) {
        if (wmode != 0 && wmode != 4)
            return fail(ctx, -4, "the gradient estimator has no surrogate or energy-input form "
                                 "(scp3g_accumulate_gradients takes energies with gradients)");
        if (ctx->precision == SCP3G_MIXED && !ctx->mixed_fast)
            return fail(ctx, -4, "the gradient estimator's mixed form needs the LDS-tiled SYRK (mixed kernel 1)");
        if (ctx->nmodes3 > ctx->nmodes2) return fail(ctx, -4, "the gradient estimator needs nmodes3 <= nmodes2");
        return run_tile_grad(ctx, np, wmode == 4);
    }
    if (wmode == 1) {
        timer_begin(ctx, "mlp", &ev);
        rc = check_hip(ctx, launch_mlp_weights(ctx, ctx->d_y, T, np, nullptr, 0, scale, ctx->d_w), "mlp");
        timer_end(ctx, "mlp", ev);
        if (rc) return rc;
        goto basis;
    }
    if (wmode == 3) goto weights;  // the caller's energies are in d_E
    timer_begin(ctx, "map", &ev);
    rc = check_hip(ctx, launch_map(ctx, ctx->d_y, T, np, ctx->d_crd, T, fp32), "map");
    timer_end(ctx, "map", ev);
    if (rc) return rc;
    timer_begin(ctx, "pes", &ev);
    if (ctx->pes == SCP3G_PES_TTM3F)
        rc = check_hip(ctx, launch_ttm3f(ctx, ctx->nw, np, ctx->d_crd, T, fp32, ctx->d_E), "pes");
    else
        rc = check_hip(ctx, launch_qtip4pf(ctx, ctx->nw, np, ctx->d_crd, T, fp32, ctx->d_E), "pes");
    timer_end(ctx, "pes", ev);
    if (rc) return rc;
weights:
    if (matched) {
        // first the true weights alone, scaled
        timer_begin(ctx, "weight", &ev);
        rc = check_hip(ctx, launch_weight(ctx, ctx->d_E, ctx->d_y, T, np, ctx->d_w, scale), "weight");
        timer_end(ctx, "weight", ev);
    } else if (wmode == 2 || wmode == 5) {
        timer_begin(ctx, "mlp", &ev);
        rc = check_hip(ctx, launch_mlp_weights(ctx, ctx->d_y, T, np, ctx->d_E, 1, scale, ctx->d_w), "mlp");
        timer_end(ctx, "mlp", ev);
    } else {
        timer_begin(ctx, "weight", &ev);
        rc = check_hip(ctx, launch_weight(ctx, ctx->d_E, ctx->d_y, T, np, ctx->d_w), "weight");
        timer_end(ctx, "weight", ev);
    }
    if (rc) return rc;
basis:
    if (fp32 && ctx->mixed_fast) {
        // signed split-BF16 operand X = sqrt|w| B in the k-blocked layout
        const int64_t K = (np + 63) / 64 * 64;
        const int64_t rows_pad = fast_rows_pad(ctx->nbasis);
        unsigned short *Xh = reinterpret_cast<unsigned short *>(ctx->d_B);
        unsigned short *Xl = Xh + (size_t)rows_pad * T;
        unsigned *sgn = reinterpret_cast<unsigned *>(Xl + (size_t)rows_pad * T);
        for (int pass = 0; pass < (matched ? 2 : 1); ++pass) {
            if (pass == 1) {
                // then -w_nn, scaled, with the surrogate term's one product
                timer_begin(ctx, "mlp", &ev);
                rc = check_hip(ctx, launch_mlp_weights(ctx, ctx->d_y, T, np, nullptr, 0, -scale, ctx->d_w),
                               "mlp");
                timer_end(ctx, "mlp", ev);
                if (rc) return rc;
            }
            // a one-product SYRK reads the hi plane only
            const int terms = ((wmode == 1 && ctx->cv_one_term) || pass == 1) ? 1 : 3;
            timer_begin(ctx, "basis", &ev);
            rc = check_hip(ctx, launch_basis_split(ctx, ctx->d_y, T, np, K, ctx->d_w, Xh, terms == 1 ? nullptr : Xl, sgn),
                           "basis");
            timer_end(ctx, "basis", ev);
            if (rc) return rc;
            timer_begin(ctx, "syrk", &ev);
            rc = check_hip(ctx, launch_syrk_x3(ctx, ctx->nbasis, Xh, Xl, sgn, K, ctx->d_H, ctx->d_Hc, terms),
                           "syrk");
            timer_end(ctx, "syrk", ev);
            if (rc) return rc;
        }
    } else {
        timer_begin(ctx, "basis", &ev);
        rc = check_hip(ctx, launch_basis(ctx, ctx->d_y, T, np, ctx->d_B, T, fp32), "basis");
        timer_end(ctx, "basis", ev);
        if (rc) return rc;
        timer_begin(ctx, "syrk", &ev);
        rc = check_hip(ctx, launch_syrk(ctx, ctx->precision, ctx->nbasis, ctx->d_B, T, ctx->d_w, np,
                                        ctx->d_H, ctx->d_Hc),
                       "syrk");
        timer_end(ctx, "syrk", ev);
        if (rc) return rc;
    }
    ctx->points_done += np;
    if (wmode == 0 || wmode == 3 || wmode == 5) ctx->norm_sum += (double)np;
    else if (wmode == 1) ctx->norm_sum += (double)np * scale;
    return 0;
}

extern "C" {

int scp3g_accumulate(scp3g_ctx *ctx, int64_t skip, int64_t npts)
{
    return scp3g_accumulate_mode(ctx, skip, npts, 0, 1.0);
}

// Sobol seeds skip..skip+npts-1 with weight mode wmode (see run_tile).
int scp3g_accumulate_mode(scp3g_ctx *ctx, int64_t skip, int64_t npts, int wmode, double scale)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (npts < 0 || skip < 0) return fail(ctx, -4, "negative skip or point count");
    // i8_sobol's column limit (sobol.cpp:27841-27852) with the 64-lane stride
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    if ((rc = ensure_work(ctx))) return rc;
    const int64_t T = ctx->work_tile;
    for (int64_t done = 0; done < npts; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        hipEvent_t ev;
        timer_begin(ctx, "sobol", &ev);
        rc = check_hip(ctx, launch_sobol(ctx, ctx->d_sobol, ctx->nmodes0, skip + done, np, 1,
                                         ctx->d_y, T, nullptr),
                       "sobol");
        timer_end(ctx, "sobol", ev);
        if (rc) return rc;
        if ((rc = run_tile(ctx, np, wmode, scale))) return rc;
    }
    return 0;
}

int scp3g_accumulate_normals(scp3g_ctx *ctx, const double *normals, int64_t npts)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (npts < 0 || (!normals && npts > 0)) return fail(ctx, -4, "bad normals");
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    if ((rc = ensure_work(ctx))) return rc;
    const int64_t T = ctx->work_tile;
    const int M = ctx->nmodes0;
    std::vector<double> yt((size_t)M * T);
    for (int64_t done = 0; done < npts; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        // point-major -> mode-major, divided by sqrt(2) (SCP3.cpp:201)
        for (int64_t i = 0; i < np; ++i)
            for (int k = 0; k < M; ++k)
                yt[(size_t)k * T + i] = normals[(size_t)(done + i) * M + k] / std::sqrt(2.0);
        if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) return rc;
        if ((rc = check_hip(ctx, hipMemcpy(ctx->d_y, yt.data(), yt.size() * 8, hipMemcpyHostToDevice),
                            "upload normals")))
            return rc;
        if ((rc = run_tile(ctx, np, 0, 1.0))) return rc;
    }
    return 0;
}

int scp3g_points_coords(scp3g_ctx *ctx, int64_t skip, int64_t npts, double *crd_host)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (npts < 0 || skip < 0 || (!crd_host && npts > 0)) return fail(ctx, -4, "bad arguments");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    hipSetDevice(ctx->device);
    if ((rc = ensure_work(ctx))) return rc;
    const int64_t T = ctx->work_tile;
    const int n3 = 9 * ctx->nw;
    std::vector<double> ct((size_t)n3 * T);
    for (int64_t done = 0; done < npts; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        rc = check_hip(ctx, launch_sobol(ctx, ctx->d_sobol, ctx->nmodes0, skip + done, np, 1, ctx->d_y, T, nullptr),
                       "sobol");
        if (!rc) rc = check_hip(ctx, launch_map(ctx, ctx->d_y, T, np, ctx->d_crd, T, 0), "map");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (!rc)
            rc = check_hip(ctx, hipMemcpy(ct.data(), ctx->d_crd, ct.size() * 8, hipMemcpyDeviceToHost), "coords");
        if (rc) return rc;
        // [coordinate][point] -> point-major, 9nw per point (OHH, A)
        for (int64_t i = 0; i < np; ++i)
            for (int c = 0; c < n3; ++c) crd_host[(size_t)(done + i) * n3 + c] = ct[(size_t)c * T + i];
    }
    return 0;
}

int scp3g_accumulate_energies(scp3g_ctx *ctx, int64_t skip, int64_t npts, const double *E_host)
{
    int rc = ready(ctx);
    if (rc) return rc;
    SyncOnError guard(ctx, rc);
    if (npts < 0 || skip < 0 || (!E_host && npts > 0)) return fail(ctx, -4, "bad arguments");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    if ((rc = ensure_work(ctx))) return rc;
    const int64_t T = ctx->work_tile;
    for (int64_t done = 0; done < npts; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        rc = check_hip(ctx, launch_sobol(ctx, ctx->d_sobol, ctx->nmodes0, skip + done, np, 1, ctx->d_y, T, nullptr),
                       "sobol");
        if (!rc)
            rc = check_hip(ctx, hipMemcpyAsync(ctx->d_E, E_host + done, (size_t)np * 8, hipMemcpyHostToDevice,
                                               ctx->stream),
                           "upload energies");
        if (!rc) rc = run_tile(ctx, np, 3, 1.0);
        // the host buffer is read asynchronously: finish before the next tile
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (rc) return rc;
    }
    return 0;
}

int scp3g_accumulate_gradients(scp3g_ctx *ctx, int64_t skip, int64_t npts, const double *E_host,
                               const double *G_host)
{
    int rc = ready(ctx);
    if (rc) return rc;
    SyncOnError guard(ctx, rc);
    if (npts < 0 || skip < 0 || ((!E_host || !G_host) && npts > 0)) return fail(ctx, -4, "bad arguments");
    if (ctx->estimator == SCP3G_EST_ENERGY)
        return fail(ctx, -4, "scp3g_accumulate_gradients needs a gradient estimator (scp3g_set_estimator)");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    if ((rc = ensure_work(ctx))) return rc;
    const int64_t T = ctx->work_tile;
    const int n3 = 9 * ctx->nw;
    std::vector<double> gt((size_t)n3 * T);
    for (int64_t done = 0; done < npts; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        // point-major (9nw per point) -> [coordinate][point]
        for (int64_t i = 0; i < np; ++i)
            for (int c = 0; c < n3; ++c) gt[(size_t)c * T + i] = G_host[(size_t)(done + i) * n3 + c];
        rc = check_hip(ctx, launch_sobol(ctx, ctx->d_sobol, ctx->nmodes0, skip + done, np, 1, ctx->d_y, T, nullptr),
                       "sobol");
        if (!rc)
            rc = check_hip(ctx, hipMemcpyAsync(ctx->d_E, E_host + done, (size_t)np * 8, hipMemcpyHostToDevice,
                                               ctx->stream),
                           "upload energies");
        if (!rc)
            rc = check_hip(ctx, hipMemcpyAsync(ctx->d_G, gt.data(), gt.size() * 8, hipMemcpyHostToDevice, ctx->stream),
                           "upload gradients");
        if (!rc) rc = run_tile(ctx, np, 4, 1.0);
        // the host buffers are read asynchronously: finish before the next tile
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (rc) return rc;
    }
    return 0;
}

int scp3g_points_done(const scp3g_ctx *ctx, int64_t *npts)
{
    if (!ctx || !npts) return -1;
    *npts = ctx->points_done;
    return 0;
}

int scp3g_h_download(scp3g_ctx *ctx, double *M_host, int raw)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (!M_host) return fail(ctx, -4, "null output");
    hipSetDevice(ctx->device);
    if ((rc = ensure_sums(ctx))) return rc;
    const int64_t n = ctx->nbasis;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
    const double *Hs = nullptr;
    if ((rc = read_sums(ctx, &Hs))) return rc;
    if (raw) {
        if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) return rc;
        return check_hip(ctx, hipMemcpy(M_host, Hs, bytes, hipMemcpyDeviceToHost), "download H");
    }
    if (ctx->points_done <= 0) return fail(ctx, -6, "no points accumulated");
    std::vector<double> dg = ho_diagonal(ctx);
    // the staging buffers stay with the running sums (freed with them)
    if (!ctx->d_diag && (rc = check_hip(ctx, hipMalloc(&ctx->d_diag, n * 8), "alloc diag"))) return rc;
    if (!ctx->d_mout && (rc = check_hip(ctx, hipMalloc(&ctx->d_mout, bytes), "alloc Mout"))) return rc;
    if ((rc = check_hip(ctx, hipMemcpyAsync(ctx->d_diag, dg.data(), n * 8, hipMemcpyHostToDevice, ctx->stream),
                        "upload diag")))
        return rc;
    hipEvent_t ev;
    timer_begin(ctx, "finalize", &ev);
    rc = check_hip(ctx, launch_finalize(ctx, Hs, n, ctx->norm_sum, ctx->d_diag, ctx->d_mout, ctx->d_offset),
                   "finalize");
    timer_end(ctx, "finalize", ev);
    // dg is read by the queued upload: wait before returning either way
    const int rs = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!rc) rc = rs;
    if (!rc) rc = check_hip(ctx, hipMemcpy(M_host, ctx->d_mout, bytes, hipMemcpyDeviceToHost), "download Mout");
    return rc;
}

// ---- per-stage entry points ------------------------------------------------

int scp3g_sobol_normals(scp3g_ctx *ctx, int dims, int64_t skip, int64_t npts, int halve,
                        double *y_dev, int64_t ldy, uint64_t *ints_dev)
{
    if (!ctx) return -1;
    if (dims < 1 || dims > 1111) return fail(ctx, -4, "dims must be in 1..1111");
    if (npts < 0 || skip < 0 || ldy < npts) return fail(ctx, -4, "bad sizes");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    if (!y_dev) return fail(ctx, -4, "null output");
    hipSetDevice(ctx->device);
    std::vector<uint64_t> dirs((size_t)dims * kSobolCols);
    sobol_directions(dims, dirs.data());
    uint64_t *d_dirs = nullptr;
    int rc;
    if ((rc = check_hip(ctx, hipMalloc(&d_dirs, dirs.size() * 8), "alloc dirs"))) return rc;
    hipMemcpyAsync(d_dirs, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice, ctx->stream);
    hipEvent_t ev;
    timer_begin(ctx, "sobol", &ev);
    rc = check_hip(ctx, launch_sobol(ctx, d_dirs, dims, skip, npts, halve, y_dev, ldy, ints_dev),
                   "sobol");
    timer_end(ctx, "sobol", ev);
    hipStreamSynchronize(ctx->stream);
    hipFree(d_dirs);
    return rc;
}

int scp3g_map_coords(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts,
                     void *crd_dev, int64_t ldc, int fp32)
{
    if (!ctx) return -1;
    if (ctx->nw <= 0) return fail(ctx, -2, "system not set (scp3g_set_system)");
    if (npts < 0 || ldy < npts || ldc < npts || !y_dev || !crd_dev) return fail(ctx, -4, "bad arguments");
    hipSetDevice(ctx->device);
    hipEvent_t ev;
    timer_begin(ctx, "map", &ev);
    int rc = check_hip(ctx, launch_map(ctx, y_dev, ldy, npts, crd_dev, ldc, fp32), "map");
    timer_end(ctx, "map", ev);
    return rc;
}

int scp3g_pes_energy(scp3g_ctx *ctx, int pes, int nw, int64_t npts, const void *crd_dev,
                     int64_t ldc, int fp32, double *E_dev)
{
    if (!ctx) return -1;
    if (nw <= 0 || nw > 64 || npts < 0 || ldc < npts || !crd_dev || !E_dev)
        return fail(ctx, -4, "bad arguments");
    hipSetDevice(ctx->device);
    hipEvent_t ev;
    int rc;
    timer_begin(ctx, "pes", &ev);
    if (pes == SCP3G_PES_QTIP4PF)
        rc = check_hip(ctx, launch_qtip4pf(ctx, nw, npts, crd_dev, ldc, fp32, E_dev), "qtip4pf");
    else if (pes == SCP3G_PES_TTM3F)
        rc = check_hip(ctx, launch_ttm3f(ctx, nw, npts, crd_dev, ldc, fp32, E_dev), "ttm3f");
    else
        rc = fail(ctx, -4, "unknown potential id");
    timer_end(ctx, "pes", ev);
    return rc;
}

int scp3g_pes_gradient(scp3g_ctx *ctx, int pes, int nw, int64_t npts, const double *crd_dev,
                       int64_t ldc, double *E_dev, double *G_dev, int64_t ldg)
{
    if (!ctx) return -1;
    if (nw <= 0 || nw > 64 || npts < 0 || ldc < npts || ldg < npts || !crd_dev || !E_dev || !G_dev)
        return fail(ctx, -4, "bad arguments");
    if (pes != SCP3G_PES_QTIP4PF && pes != SCP3G_PES_TTM3F) return fail(ctx, -4, "unknown potential id");
    hipSetDevice(ctx->device);
    hipEvent_t ev;
    timer_begin(ctx, "pes", &ev);
    int rc = pes == SCP3G_PES_TTM3F
                 ? check_hip(ctx, launch_ttm3f_grad(ctx, nw, npts, crd_dev, ldc, E_dev, G_dev, ldg), "ttm3f")
                 : check_hip(ctx, launch_qtip4pf_grad(ctx, nw, npts, crd_dev, ldc, E_dev, G_dev, ldg), "qtip4pf");
    timer_end(ctx, "pes", ev);
    return rc;
}

int scp3g_basis_table(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts,
                      void *B_dev, int64_t ldb, int fp32)
{
    int rc = ready(ctx);
    if (rc) return rc;
    if (npts < 0 || ldy < npts || ldb < npts || !y_dev || !B_dev) return fail(ctx, -4, "bad arguments");
    hipSetDevice(ctx->device);
    hipEvent_t ev;
    timer_begin(ctx, "basis", &ev);
    rc = check_hip(ctx, launch_basis(ctx, y_dev, ldy, npts, B_dev, ldb, fp32), "basis");
    timer_end(ctx, "basis", ev);
    return rc;
}

int scp3g_h_accumulate(scp3g_ctx *ctx, int mode, int64_t n, const void *B_dev, int64_t ldb,
                       const double *w_dev, int64_t npts, double *H_dev, double *Hc_dev)
{
    if (!ctx) return -1;
    if (mode != SCP3G_FP64 && mode != SCP3G_MIXED) return fail(ctx, -4, "bad mode");
    if (n <= 0 || npts < 0 || !B_dev || !w_dev || !H_dev) return fail(ctx, -4, "bad arguments");
    if (mode == SCP3G_MIXED && !Hc_dev) return fail(ctx, -4, "mixed mode needs Hc");
    // the kernels read 128-row tiles and 16/32-point chunks: the caller's B must
    // hold rows up to round_up(n, 128) and points up to round_up(npts, 32),
    // zero beyond n / npts, and w zero beyond npts
    if (ldb % 32 || ldb < (npts + 31) / 32 * 32) return fail(ctx, -4, "ldb must be a multiple of 32 covering npts");
    hipSetDevice(ctx->device);
    hipEvent_t ev;
    timer_begin(ctx, "syrk", &ev);
    int rc = check_hip(ctx, launch_syrk(ctx, mode, n, B_dev, ldb, w_dev, npts, H_dev, Hc_dev), "syrk");
    timer_end(ctx, "syrk", ev);
    return rc;
}

int scp3g_set_timing(scp3g_ctx *ctx, int on)
{
    if (!ctx) return -1;
    ctx->timing = on != 0;
    if (on) {
        for (auto &kv : ctx->timers) {
            timer_drain(ctx, kv.second);
            kv.second.total_ms = 0.0;
            kv.second.launches = 0;
        }
    }
    return 0;
}

int scp3g_kernel_time(scp3g_ctx *ctx, const char *name, double *total_ms, int64_t *launches)
{
    if (!ctx || !name || !total_ms || !launches) return -1;
    auto it = ctx->timers.find(name);
    if (it == ctx->timers.end()) {
        *total_ms = 0.0;
        *launches = 0;
        return 0;
    }
    timer_drain(ctx, it->second);
    *total_ms = it->second.total_ms;
    *launches = it->second.launches;
    return 0;
}

}  // extern "C"
