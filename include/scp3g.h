/*
 * scp3g.h - C ABI of the MI355X Hamiltonian-assembly library (libscp3g.so).
 *
 * The library replaces the hot loop of SCP3's -G (no-gradient) estimator,
 * SCP3.cpp:187-232:
 *
 *   for each block of points:
 *     y  = sobol::std_normal(...) / sqrt(2)           SCP3.cpp:196-201
 *     r  = bohr * (x0 + MUa * y)                       SCP3.cpp:208
 *     V  = pot(nw, r) / autokcalpmol                   SCP3.cpp:219
 *     bV = V - 0.5 * y . (omega o y)                   SCP3.cpp:222-223
 *     M += bra diag(bV) bra^T   (upper triangle)       SCP3.cpp:231,
 *                                                      ScaledMatrixElements.cpp:951-971
 *   M / P + HO diagonal                                SCP3.cpp:237-239,269-271
 *
 * and the eigensolver's lowest-K levels (eigensolver.cpp:20-109).
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    scp3g_last_error() returns the message.  Nothing in the library
 *    exits or throws across this boundary (the reference exits inside
 *    ScaledMatrixElements.cpp:119-122 and h2o.h:42-45).
 *  - Matrices follow the reference: column-major, M(r, c) at c*n + r, only
 *    the upper triangle (r <= c) is accumulated.
 *  - Pointers named *_host are host memory borrowed for the call; pointers
 *    named *_dev are device memory on the context's device, and the call
 *    is ordered on the context's stream (scp3g_sync waits for it).
 *  - One context per host thread; a context is not re-entrant (as an
 *    h2o::PES instance is not, qtip4pf.cpp:137-148).
 */
#ifndef SCP3G_H
#define SCP3G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scp3g_ctx scp3g_ctx;

/* precision modes of the H accumulation */
#define SCP3G_FP64 0   /* FP64 potential, FP64 basis, FP64 MFMA SYRK          */
#define SCP3G_MIXED 1  /* FP32 potential, FP32 basis, 3-term split-BF16 MFMA  */
                       /* SYRK accumulated in FP32 per tile, folded into an   */
                       /* FP64 running H by compensated summation             */

/* estimators of the potential-energy matrix M (SCP3.cpp:203-231) */
#define SCP3G_EST_ENERGY 0          /* -G: M += B diag(bV) B^T (addEpot,       */
                                    /* ScaledMatrixElements.cpp:951-971)       */
#define SCP3G_EST_GRADIENT 1        /* the default without -G: the Stein-       */
                                    /* identity elements from bV and its        */
                                    /* gradient, placed as the reference does   */
                                    /* (ScaledMatrixElements.cpp:100-725)       */
#define SCP3G_EST_GRADIENT_STEIN 2  /* the same estimator with every element    */
                                    /* from the Stein identity (differs from 1  */
                                    /* in five element families, see grad.hip)  */

/* potentials selectable on the GPU (h2o.h:20-48 names, plus ttm3f) */
#define SCP3G_PES_QTIP4PF 0
#define SCP3G_PES_TTM3F 1

/* ---- context ------------------------------------------------------- */

/* Create a context on HIP device `device` (its own non-blocking stream). */
int scp3g_create(scp3g_ctx **ctx, int device);
void scp3g_destroy(scp3g_ctx *ctx);
const char *scp3g_last_error(const scp3g_ctx *ctx);
int scp3g_sync(scp3g_ctx *ctx);

/*
 * Molecular system: nw waters, x0_host[9nw] equilibrium Cartesians in bohr
 * in O H H order (after OHHOHH, DiskIO.cpp:142-195), omega_host[nmodes0]
 * harmonic frequencies in hartree, MUa_host[9nw x nmodes0] column-major
 * mass-and-frequency-scaled normal modes (SCP3.cpp:307-314).
 * Replaces the arguments of SCP3_a (SCP3.cpp:154).
 */
int scp3g_set_system(scp3g_ctx *ctx, int nw, const double *x0_host, int nmodes0,
                     const double *omega_host, const double *MUa_host);

/*
 * Basis: selected modes (indices into 0..nmodes0-1, ascending frequency),
 * singles over all selected modes, doubles over the first nmodes2 and
 * triples over the first nmodes3 of them.  Replaces the constructor of
 * ScaledMatrixElements (ScaledMatrixElements.h:73-83).
 */
int scp3g_set_basis(scp3g_ctx *ctx, int nsel, const int *modes_host, int nmodes2,
                    int nmodes3);
/* n = 1 + nsel + nmodes2(nmodes2+1)/2 + nmodes3(nmodes3+1)(nmodes3+2)/6,
 * ScaledMatrixElements.cpp:73-78 */
int scp3g_basis_size(const scp3g_ctx *ctx, int64_t *n);
/* ScaledMatrixElements.cpp:81-99 */
int scp3g_sub_basis_size(const scp3g_ctx *ctx, int level, int64_t *n);

/* Potential by name: "qtip4pf" or "ttm3f" (PESFromString, h2o.h:20-48). */
int scp3g_set_pes(scp3g_ctx *ctx, const char *name);
/* SCP3G_FP64 or SCP3G_MIXED */
int scp3g_set_precision(scp3g_ctx *ctx, int mode);
/* SYRK kernels: 1 (default) = signed 256 x 256-tile split-BF16 kernel on
 * X = sqrt|w| B, 8 waves on 16x16x32 MFMA, operands moved to LDS by DMA;
 * 0 = 64 x 64-tile direct-load kernels on B and w.  1 also selects the
 * LDS-tiled FP64 kernel in the FP64 form. */
int scp3g_set_mixed_kernel(scp3g_ctx *ctx, int fast);
/*
 * SCP3G_EST_ENERGY (default), SCP3G_EST_GRADIENT or SCP3G_EST_GRADIENT_STEIN
 * (SCP3.cpp:63-151 option -G selects the first; use_gradient the others).
 * Clears the running sum.  The gradient forms evaluate q-TIP4P/F or TTM3-F
 * with gradients in FP64; in SCP3G_MIXED their rank-2 term and per-mode
 * block corrections run as split-BF16 SYRKs (LDS-tiled kernel only).  They
 * need nmodes3 <= nmodes2; accumulate calls fail otherwise.
 */
int scp3g_set_estimator(scp3g_ctx *ctx, int estimator);
/*
 * Scrambled Sobol points: scrambled != 0 applies a random affine matrix
 * scramble of the binary digits to every dimension (Matousek 1998: a
 * lower-triangular binary matrix with unit diagonal and a digital shift,
 * drawn from `seed`, any value including 0; the class of
 * scramble(p, 'MatousekAffineOwen') behind the reference's -r files,
 * SCP3.cpp:188-194); scrambled == 0 (the default) is the reference's
 * unscrambled sequence (sobol.cpp:14198), seed ignored.  Needs
 * scp3g_set_system first (which keeps the setting); applies to the
 * context's accumulate calls, not to scp3g_sobol_normals.
 */
int scp3g_set_scramble(scp3g_ctx *ctx, int scrambled, uint64_t seed);
/* points per device tile (multiple of 256), or 0 (default): the largest power
 * of two <= 524288 whose FP64 basis table fits in 8 GiB, at least 4096 */
int scp3g_set_tile(scp3g_ctx *ctx, int tile);

/* ---- whole-path entry points -------------------------------------- */

/* Zero the running sum M and the point counter. */
int scp3g_reset(scp3g_ctx *ctx);

/*
 * Accumulate Sobol points with seeds skip .. skip+npts-1 (the reference's
 * sobol_skip, SCP3.cpp:48,196-199) into the running FP64 sum M.
 */
int scp3g_accumulate(scp3g_ctx *ctx, int64_t skip, int64_t npts);

/*
 * Accumulate caller-supplied standard normals (the -r rng-file branch,
 * SCP3.cpp:188-194): normals_host[npts x nmodes0] point-major; they are
 * divided by sqrt(2) as SCP3.cpp:201 does.
 */
int scp3g_accumulate_normals(scp3g_ctx *ctx, const double *normals_host, int64_t npts);

/*
 * Potentials on the CPU (WHBB, HBB2-pol or any h2o::PES, PES.h:13-23): the
 * Cartesians the reference hands to the potential for Sobol seeds
 * skip .. skip+npts-1 (SCP3.cpp:196-208: y = BSM(Sobol) / sqrt(2),
 * r = bohr (x0 + MUa y)), crd_host[npts x 9nw] point-major, O H H, A ...
 */
int scp3g_points_coords(scp3g_ctx *ctx, int64_t skip, int64_t npts, double *crd_host);
/*
 * ... and their energies E_host[npts] (kcal/mol, from the caller's
 * potential) accumulated into the running sum: bV, basis rows and the SYRK
 * on the GPU (SCP3.cpp:213-231 with the PES call made by the caller).
 */
int scp3g_accumulate_energies(scp3g_ctx *ctx, int64_t skip, int64_t npts, const double *E_host);
/*
 * The same for the gradient estimators (scp3g_set_estimator): the caller's
 * energies E_host[npts] (kcal/mol) and gradients G_host[npts x 9nw]
 * (kcal/mol/A, point-major, the gradient operator() of h2o::PES) for the
 * Cartesians of scp3g_points_coords; bV, bVy (SCP3.cpp:210-223) and the
 * estimator's SYRKs run on the GPU.
 */
int scp3g_accumulate_gradients(scp3g_ctx *ctx, int64_t skip, int64_t npts, const double *E_host,
                               const double *G_host);

/* Number of points accumulated since the last reset. */
int scp3g_points_done(const scp3g_ctx *ctx, int64_t *npts);

/*
 * Mout = M / points_done + HO diagonal + 0.5*(sum(omega) - sum(omega_sel))
 * (SCP3.cpp:237-239,269-271; ScaledMatrixElements.cpp:735-784), written to
 * M_host[n*n] column-major with the upper triangle filled and the strict
 * lower triangle zero - the layout of the reference's M and of the sM_*.h5
 * dataset.  raw != 0 returns the plain running sum M instead.  For the
 * gradient estimators M is the sum of the per-point element estimators
 * (ScaledMatrixElements.cpp:100-725), assembled from the device's running
 * sums when read.
 */
int scp3g_h_download(scp3g_ctx *ctx, double *M_host, int raw);

/*
 * Lowest k eigenvalues (hartree, ascending) of Mout, computed on the GPU in
 * FP64 (replaces eigensolver.cpp:20-80 dsyevr RANGE='I').  evecs_host may be
 * NULL; otherwise n x k column-major.  tol is the residual tolerance in
 * hartree (<= 0 selects 1e-9); k <= 28.
 */
int scp3g_lowest_levels(scp3g_ctx *ctx, int k, double tol, double *evals_host,
                        double *evecs_host);

/*
 * Lowest k eigenvalues of the leading nsub x nsub block of Mout (the E0
 * monitor of SCP3.cpp:241-247 uses the singles and singles+doubles blocks).
 */
int scp3g_lowest_levels_sub(scp3g_ctx *ctx, int64_t nsub, int k, double tol, double *evals_host);

/*
 * Lowest k eigenpairs of a caller-supplied symmetric n x n matrix
 * (column-major; the upper triangle is read unless M(0, n-1) == 0, then the
 * lower one, as eigensolver.cpp:27-30 chooses UPLO), on the context's GPU.
 * Replaces eigensolver.cpp's dsyevr on the matrix read from sM_*.h5.
 */
int scp3g_matrix_lowest_levels(scp3g_ctx *ctx, int64_t n, const double *M_host, int k, double tol,
                               double *evals_host, double *evecs_host);

/* ---- MLP surrogate and control variate --------------------------------
 * No counterpart in the reference.  A network 3 x (width `hidden`, SiLU) on
 * the normal coordinates (x = sqrt(2) y, padded to a multiple of 64) models
 * V - V0 - 1/2 sum omega y^2 (normalised); BF16 weights and activations on
 * MFMA with FP32 accumulation, FP32 master weights and Adam, all on the GPU.
 * Parameter vector layout (FP32): [W0 (h x dpad) | b0 | W1 (h x h) | b1 |
 * W2 (h x h) | b2 | w3 (h) | b3], W row-major [out][in].  norm[3] holds
 * V0 (hartree), the target mean and the target scale.
 */
int scp3g_mlp_init(scp3g_ctx *ctx, int hidden, uint64_t seed);
int scp3g_mlp_param_count(const scp3g_ctx *ctx, int64_t *n);
int scp3g_mlp_get_params(scp3g_ctx *ctx, float *P_host, double *norm_host);
int scp3g_mlp_set_params(scp3g_ctx *ctx, const float *P_host, const double *norm_host);
/* training set from host data: y_host[npts x nmodes0] (halved normals,
 * point-major), t_host[npts] normalised targets; npts a multiple of 64 */
int scp3g_mlp_load_data(scp3g_ctx *ctx, const double *y_host, const float *t_host, int64_t npts);
/* training set sampled by the path itself: Sobol seeds skip.., FP64 potential */
int scp3g_mlp_sample_data(scp3g_ctx *ctx, int64_t skip, int64_t npts);
/* `steps` Adam steps on consecutive batches; loss_host[steps] (nullable) */
int scp3g_mlp_train(scp3g_ctx *ctx, int steps, int batch, double lr, float *loss_host);
/* loss gradient over the first `batch` training points, no update */
int scp3g_mlp_grad(scp3g_ctx *ctx, int batch, float *grad_host, float *loss_host);
/* V_nn (hartree) of npts points, y_dev mode-major with leading dimension ldy */
int scp3g_mlp_eval(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts, double *V_dev);
/*
 * Control-variate accumulation: (1/P_nn) sum B B^T w_nn over seeds skip_nn..
 * plus (1/P_t) sum B B^T (bV - w_nn) over seeds skip_t.. (true potential);
 * unbiased for the same H as scp3g_accumulate.  After a reset, the running
 * sum is normalised by the accumulated weight (here 1).
 */
int scp3g_accumulate_cv(scp3g_ctx *ctx, int64_t skip_nn, int64_t P_nn, int64_t skip_t,
                        int64_t P_t);
/* mixed form: split-BF16 terms used for the surrogate term of the control
 * variate, 3 (default, hi.hi + hi.lo + lo.hi) or 1 (hi.hi).  With 1 the
 * correction term takes bV with three products less w_nn with the same one
 * product, so the one-product rounding of the surrogate term cancels in
 * expectation (no reference counterpart) */
int scp3g_set_cv_terms(scp3g_ctx *ctx, int terms);
/* the surrogate term alone: H += sum B B^T w_nn over Sobol seeds
 * skip..skip+npts-1, normaliser += npts (QMC with V_nn in place of V, the
 * first term of scp3g_accumulate_cv); no reference counterpart */
int scp3g_accumulate_surrogate(scp3g_ctx *ctx, int64_t skip, int64_t npts);
/* the correction term alone: H += sum B B^T (bV - w_nn) over Sobol seeds
 * skip..skip+npts-1 with the true potential, normaliser += npts (QMC of the
 * residual V - V_nn, the second term of scp3g_accumulate_cv); with
 * scp3g_accumulate_surrogate in another context it lets a caller grow the
 * two terms separately; no reference counterpart */
int scp3g_accumulate_correction(scp3g_ctx *ctx, int64_t skip, int64_t npts);
/* the running sum, normalised (sum / normaliser, no HO diagonal), is added
 * to an earlier fixed term (none at first), the total becomes the fixed
 * term of every later Mout (scp3g_h_download with raw = 0) and level solve
 * (scp3g_lowest_levels, _sub), and the running sum restarts at zero: after
 * scp3g_accumulate_correction, a surrogate term grown by
 * scp3g_accumulate_surrogate then gives the control variate's H = S / P_nn
 * + C / P_t at every checkpoint.  Energy estimator only; a basis, system or
 * estimator change drops the term; no reference counterpart */
int scp3g_hold_as_offset(scp3g_ctx *ctx);
/* drop the fixed term of scp3g_hold_as_offset */
int scp3g_clear_offset(scp3g_ctx *ctx);

/* ---- per-stage entry points (device pointers; used by parity tests) ---- */

/*
 * y_dev[d*ldy + i] = BSM(sobol(skip+i)[d]) / (halve ? sqrt(2) : 1) for
 * d < dims, i < npts (sobol.cpp:14198-27867, beasley_springer_moro.cpp:14-46,
 * SCP3.cpp:201).  ints_dev (nullable) receives the 62-bit Sobol integers
 * (lastq of i8_sobol) in the same layout.
 */
int scp3g_sobol_normals(scp3g_ctx *ctx, int dims, int64_t skip, int64_t npts, int halve,
                        double *y_dev, int64_t ldy, uint64_t *ints_dev);
/*
 * The same points as scp3g_sobol_normals, from an entry point that
 * allocates nothing: no counterpart in the reference, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops).  The direction integers are staged in a fixed device array of the
 * library (1111 x 62 words, 551 KiB per GPU, held until the process ends)
 * instead of a buffer allocated per call; dims above 1111 are rejected.
 * Stream-ordered on the context's stream and does not wait for it.  Calls
 * that overlap on one GPU (several contexts or threads) race on that array.
 */
int scp3g_sobol_normals_noalloc(scp3g_ctx *ctx, int dims, int64_t skip, int64_t npts, int halve,
                                double *y_dev, int64_t ldy, uint64_t *ints_dev);

/*
 * crd_dev[c*ldc + i] = bohr*(x0[c] + sum_k MUa[c,k] y[k*ldy + i]) for
 * c < 9nw (SCP3.cpp:208), in Angstrom; fp32 != 0 writes float.
 */
int scp3g_map_coords(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts,
                     void *crd_dev, int64_t ldc, int fp32);

/*
 * Energies of npts configurations crd_dev[c*ldc + i] (Angstrom, O H H
 * order) in kcal/mol: E_dev[i] = pot(nw, crd_i) (PES.h:13-23 energy-only
 * operator()).  fp32 != 0 reads float coordinates and evaluates in FP32.
 */
int scp3g_pes_energy(scp3g_ctx *ctx, int pes, int nw, int64_t npts, const void *crd_dev,
                     int64_t ldc, int fp32, double *E_dev);

/*
 * Energies and gradients of npts configurations (FP64): E_dev[i] as above and
 * G_dev[c*ldg + i] = dE/dx_c in kcal/mol/A, the gradient operator()
 * (PES.h:13-23; qtip4pf.cpp:252-397, ttm3f.cpp:325-716).
 */
int scp3g_pes_gradient(scp3g_ctx *ctx, int pes, int nw, int64_t npts, const double *crd_dev,
                       int64_t ldc, double *E_dev, double *G_dev, int64_t ldg);

/*
 * SCP1's Gaussian averages (SCP1::average, SCP1.cpp:33-86) over the Sobol
 * points skip .. skip + npts - 1 in 9nw dimensions, on this device (the
 * reference splits the range over MPI ranks and Allreduces the three sums).
 * Point i: z = std_normal(9nw, i), dq = isqrtM_U (z o sqrt(d)), r = bohr (q +
 * dq) in Angstrom, V and g = grad V of potential pes (SCP3G_PES_*) at r.
 * Host arrays: q (9nw, bohr, O H H order), d (9nw, >= 0), isqrtM_U (9nw x
 * 9nw column-major).  Outputs (host): *Upot = <V>/autokcalpmol,
 * UX = <g> bohr/autokcalpmol (9nw), UX_q = <(g bohr/autokcalpmol) dq^T>
 * (9nw x 9nw column-major), each divided by npts.  Up to 21 waters (the
 * potential kernels' limit).
 */
int scp3g_scp1_average(scp3g_ctx *ctx, int pes, int nw, const double *q, const double *d,
                       const double *isqrtM_U, int64_t skip, int64_t npts, double *Upot,
                       double *UX, double *UX_q);

/*
 * Cartesian Hessian of potential pes at q (9nw, bohr) by central differences
 * of the gradient with step s bohr, all 2 x 9nw displaced configurations in
 * one batch (getHessian, Hessian.cpp:13-43).  H (host, 9nw x 9nw
 * column-major) in Hartree/bohr^2: the gradient is converted to Hartree/bohr
 * before the difference (the reference divides kcal/mol/A by the step in
 * bohr).  As in the reference only the upper triangle is symmetrised.
 */
int scp3g_hessian(scp3g_ctx *ctx, int pes, int nw, const double *q, double s, double *H);

/*
 * Basis table B_dev[a*ldb + i] = bra_a(q_i) for the context's basis, with
 * q_i[m] = y_dev[modes[m]*ldy + i] (get_bra, ScaledMatrixElements.cpp:870-930).
 * fp32 != 0 writes float.
 */
int scp3g_basis_table(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts,
                      void *B_dev, int64_t ldb, int fp32);

/*
 * H_dev (n x n column-major, upper) += B diag(w) B^T over npts points, for
 * B_dev[a*ldb + i] (n rows).  mode SCP3G_FP64 reads double B; SCP3G_MIXED
 * reads float B and folds an FP32 tile sum into H_dev with the
 * compensation array Hc_dev (same shape, may not be NULL in that mode).
 * Replaces addEpot's dsyr2k (ScaledMatrixElements.cpp:951-971).
 */
int scp3g_h_accumulate(scp3g_ctx *ctx, int mode, int64_t n, const void *B_dev, int64_t ldb,
                       const double *w_dev, int64_t npts, double *H_dev, double *Hc_dev);

/* ---- host-side set-up (SCP3.cpp:292-318) --------------------------- */

/*
 * Normal modes from a mass-weighted Hessian (3N x 3N, atomic units):
 * eigen-decomposition (ascending), omega = sqrt(ev[6:]),
 * MUa[:, k] = U[:, 6+k] / (sqrt(mass) * sqrt(omega_k)).  omega_all_host
 * (nullable) receives all 3N sqrt(|ev|) (omega0.dat, SCP3.cpp:300-302).
 */
int scp3g_normal_modes(int n3, const double *mass_host, const double *hess_host,
                       double *omega_host, double *MUa_host, double *omega_all_host);

/*
 * CPU FP64 energy of one configuration by potential name ("qtip4pf",
 * "ttm3f"), kcal/mol, with a central-difference gradient when grad_host is
 * not NULL (PES.h interface; used for input preparation).
 */
int scp3g_pes_cpu(const char *name, int nw, const double *crd_host, double *E_host,
                  double *grad_host);

/* ---- timing ------------------------------------------------------- */

/* Enable/disable per-kernel HIP-event timing on the context's stream. */
int scp3g_set_timing(scp3g_ctx *ctx, int on);
/*
 * Total milliseconds and launch count of kernel `name` since timing was
 * enabled ("sobol", "map", "pes", "basis", "syrk", "fold", "finalize",
 * "symv").  Synchronises the stream.
 */
int scp3g_kernel_time(scp3g_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* SCP3G_H */
