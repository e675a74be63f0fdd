// Internal declarations shared by the HIP translation units of libscp3g.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/scp3g.h"

namespace scp3g {

// Constants.cpp:12-18, Constants.h:21
constexpr double kBohr = 0.52917721092;
constexpr double kAutokcalpmol = 627.509474277194;
constexpr int kSobolCols = 62;

struct Error {
    int code;
    std::string msg;
};

// One basis row: value = (f0 * f1) * f2 with f = ho_deg(q[mode]); deg 0
// factors are 1.0.  Packed as mode (bits 0-15) | deg (bits 16-17) per
// factor, three factors per row.
struct RowDesc {
    uint32_t f[3];
};
static inline uint32_t pack_factor(int mode, int deg)
{
    return (uint32_t)mode | ((uint32_t)deg << 16);
}

// Per-kernel HIP-event timing on the context's stream.
struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    int64_t launches = 0;
};

}  // namespace scp3g

struct scp3g_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // system
    int nw = 0, nmodes0 = 0;
    std::vector<double> x0, omega, MUa;
    // basis
    std::vector<int> modes;
    int nmodes2 = 0, nmodes3 = 0;
    int64_t nbasis = 0;
    std::vector<scp3g::RowDesc> desc;
    int pes = SCP3G_PES_QTIP4PF;
    int precision = SCP3G_FP64;
    int tile = 0;  // points per tile; 0 = automatic (auto_tile)
    // LDS-tiled SYRK kernels (syrk_fast.hip: signed 256 x 256 split-BF16;
    // syrk64_fast.hip: 128 x 128 FP64) instead of the 64 x 64 direct-load
    // kernels (syrk.hip, syrk_mixed.hip)
    bool mixed_fast = true;
    // control variate, mixed form: the surrogate term with hi.hi products only
    bool cv_one_term = false;
    // SCP3G_EST_ENERGY (the -G estimator) or one of the gradient forms
    int estimator = SCP3G_EST_ENERGY;
    // Sobol digit scrambling: on/off (off = the reference's sequence) and
    // the seed of the random matrices and shifts (0 is a seed like any other)
    bool scramble_on = false;
    uint64_t scramble = 0;

    // device copies of the set-up
    uint64_t *d_sobol = nullptr;  // nmodes0 x 62, then nmodes0 shifts (scrambled)
    double *d_x0 = nullptr, *d_omega = nullptr, *d_MUa = nullptr;
    int *d_modes = nullptr;
    scp3g::RowDesc *d_desc = nullptr;

    // per-tile work buffers (tile points)
    int work_tile = 0;
    int work_precision = -1;   // precision the work buffers were sized for
    int64_t work_n = 0;
    double *d_y = nullptr;     // nmodes0 x tile
    void *d_crd = nullptr;     // 9nw x tile (double)
    double *d_E = nullptr;     // tile
    double *d_w = nullptr;     // tile
    void *d_B = nullptr;       // n x tile (double or float)
    // gradient estimator, per tile: dE/dx (9nw x tile), Vq (nsel x tile),
    // the +-SYRK weights and the per-point scale
    int work_estimator = -1;
    double *d_G = nullptr, *d_Vq = nullptr, *d_wp = nullptr, *d_wm = nullptr, *d_sc = nullptr;

    // running sums
    double *d_H = nullptr;     // n x n column-major, upper
    double *d_Hc = nullptr;    // compensation (mixed mode)
    double *d_part = nullptr;  // split-K partials
    size_t part_bytes = 0;
    // gradient estimator: running sums of the fixed elements (ngfix), the
    // A_d row lists, correction work items and fix list (grad.hip)
    double *d_F = nullptr;
    double *d_Heff = nullptr;  // the effective running sum, built when read (n x n)
    int *d_grows = nullptr;
    int4 *d_gwork = nullptr, *d_gfix = nullptr;
    double *d_gpart = nullptr;
    size_t gpart_bytes = 0;
    // mixed form: per-mode block tables and the block's FP64 sum
    void *d_gcx = nullptr;
    size_t gcx_bytes = 0;
    std::vector<int> gblock_off, gblock_na;  // A_d offset and size per mode
    int ngwork = 0, ngfix = 0, grad_form = -1, grad_tile = 64;
    int64_t grad_n = 0;
    int64_t h_n = 0;
    // Mout staging of scp3g_h_download (n x n) and the diagonal it adds,
    // kept with the running sums
    double *d_mout = nullptr, *d_diag = nullptr;
    // fixed term added to every finalised Mout (scp3g_hold_as_offset): an
    // earlier normalised sum, upper triangle, leading dimension n; freed
    // with the running sums
    double *d_offset = nullptr;
    int64_t points_done = 0;
    // normaliser of the running sum: the number of points for the plain
    // estimator, the sum of per-point scales for the control-variate one
    double norm_sum = 0.0;
    // MLP surrogate state (mlp.hip), owned by the context
    struct MlpState *mlp = nullptr;
    // Davidson work space, kept between solves (eigen.hip): device buffers
    // and pinned host staging (hipHostMalloc)
    double *d_dav = nullptr;
    size_t dav_bytes = 0;
    double *h_dav = nullptr;
    size_t hdav_bytes = 0;

    // timing
    bool timing = false;
    std::map<std::string, scp3g::KernelTimer> timers;
    std::vector<hipEvent_t> event_pool;
    // TTM3-F gradient: per-(configuration, molecule) monomer terms
    double *d_mono = nullptr;
    size_t mono_bytes = 0;
};

// Sobol seeds skip..skip+npts-1 through the tile pipeline with weight mode
// wmode (0 plain, 1 surrogate, 2 control-variate correction) and scale.
extern "C" int scp3g_accumulate_mode(scp3g_ctx *ctx, int64_t skip, int64_t npts, int wmode,
                                     double scale);

namespace scp3g {

int fail(scp3g_ctx *ctx, int code, const std::string &msg);
int check_hip(scp3g_ctx *ctx, hipError_t e, const char *what);

// Waits for the context's stream when a call returns an error: host staging
// buffers handed to asynchronous uploads (pageable vectors, the caller's
// arrays) must not be released while a copy from them may still be queued.
struct SyncOnError {
    scp3g_ctx *ctx;
    const int &rc;
    SyncOnError(scp3g_ctx *c, const int &r) : ctx(c), rc(r) {}
    ~SyncOnError()
    {
        if (rc) hipStreamSynchronize(ctx->stream);
    }
};

// timing helpers: bracket a launch on ctx->stream
void timer_begin(scp3g_ctx *ctx, const char *name, hipEvent_t *ev0);
void timer_end(scp3g_ctx *ctx, const char *name, hipEvent_t ev0);

// kernel launchers (defined in the .hip files); all enqueue on ctx->stream
hipError_t launch_sobol(scp3g_ctx *ctx, const uint64_t *dirs, int dims, int64_t skip,
                        int64_t npts, int halve, double *y, int64_t ldy, uint64_t *ints);
hipError_t launch_map(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                      void *crd, int64_t ldc, int fp32);
hipError_t launch_pes(scp3g_ctx *ctx, int pes, int nw, int64_t npts, const void *crd,
                      int64_t ldc, int fp32, double *E);
hipError_t launch_weight(scp3g_ctx *ctx, const double *E, const double *y, int64_t ldy,
                         int64_t npts, double *w, double scale = 1.0);
hipError_t launch_basis(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, void *B,
                        int64_t ldb, int fp32);
hipError_t launch_syrk(scp3g_ctx *ctx, int mode, int64_t n, const void *B, int64_t ldb,
                       const double *w, int64_t npts, double *H, double *Hc);
hipError_t launch_finalize(scp3g_ctx *ctx, const double *H, int64_t n, double scale,
                           const double *diag_add, double *out, const double *offset = nullptr);

hipError_t launch_mlp_weights(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                              const double *E, int mode, double scale, double *w);
int run_tile(scp3g_ctx *ctx, int64_t np, int wmode, double scale);
hipError_t launch_qtip4pf(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                          int fp32, double *E);
hipError_t launch_ttm3f(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                        int fp32, double *E);
int64_t fast_rows_pad(int64_t n);
hipError_t launch_basis_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t K,
                              const double *w, unsigned short *Xh, unsigned short *Xl, unsigned *sgn);
hipError_t launch_syrk_x3(scp3g_ctx *ctx, int64_t n, const unsigned short *Xh,
                          const unsigned short *Xl, const unsigned *sgn, int64_t K, double *H,
                          double *Hc, int terms);
void release_fast_tiles(scp3g_ctx *ctx);
int64_t rows_pad128(int64_t n);
hipError_t launch_syrk_f64_lds(scp3g_ctx *ctx, int64_t n, const double *B, int64_t ldb,
                               const double *w, int64_t npts, double *H);
void release_tiles128(scp3g_ctx *ctx);
double qtip4pf_energy_host(int nw, const double *crd);
hipError_t launch_qtip4pf_grad(scp3g_ctx *ctx, int nw, int64_t npts, const double *crd, int64_t ldc,
                               double *E, double *G, int64_t ldg);
hipError_t launch_ttm3f_grad(scp3g_ctx *ctx, int nw, int64_t npts, const double *crd, int64_t ldc,
                             double *E, double *G, int64_t ldg);
// gradient estimator (grad.hip)
int grad_setup(scp3g_ctx *ctx);
void grad_release(scp3g_ctx *ctx);
// SCP1 averages (scp1.hip)
void scp1_release(scp3g_ctx *ctx);
hipError_t launch_vq(scp3g_ctx *ctx, const double *G, int64_t ldg, const double *y, int64_t ldy,
                     const double *w, int64_t npts, double *Vq, int64_t ldv, double *wp, double *wm,
                     double *sc);
hipError_t launch_grad_table(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                             const double *w, const double *Vq, int64_t ldv, const double *sc,
                             double sgn, double *B, int64_t ldb);
hipError_t launch_grad_corrections(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                                   const double *w, const double *Vq, int64_t ldv, double *H,
                                   double *F);
hipError_t launch_grad_effective(scp3g_ctx *ctx, const double *S, const double *F, double *dst);
hipError_t launch_corr_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t Kh, int d,
                             const int *rows, int off, int na, const double *w, const double *Vq, int64_t ldv,
                             unsigned short *Xh, unsigned short *Xl, unsigned *sgn);
hipError_t launch_grad_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t Kh,
                             const double *w, const double *Vq, int64_t ldv, const double *sc,
                             const double *wp, unsigned short *Xh, unsigned short *Xl, unsigned *sgn);
// the running sum as the reference's M (before division by the points):
// ctx->d_H for the energy estimator, built into ctx->d_Hc for the gradient one
int read_sums(scp3g_ctx *ctx, const double **H);
double ttm3f_energy_host(int nw, const double *crd, int *iters);
void mlp_release(scp3g_ctx *ctx);
void release_davidson(scp3g_ctx *ctx);

// host helpers
std::vector<double> ho_diagonal(const scp3g_ctx *ctx);  // HO diagonal + unselected ZPE
void sobol_directions(int dims, uint64_t *v);  // dims x 62, scaled
// device addresses of the allocation-free sampler's arrays, once per device
hipError_t sobol_static_setup(int device);
// affine digit scrambling of the direction integers in place, shifts dims
void sobol_scramble(int dims, uint64_t seed, uint64_t *v, uint64_t *shift);
// lowest nvec eigenpairs of a small dense symmetric matrix (column-major,
// both triangles, overwritten); w ascending, Z m x nvec (small_eigh.cpp)
int small_eigh(int m, std::vector<double> &A, int nvec, std::vector<double> &w, std::vector<double> &Z);
int jacobi_eigh(int n, std::vector<double> &A, std::vector<double> &w,
                std::vector<double> &V);

}  // namespace scp3g
