/*
 * Oracle (test infrastructure only, never the product path): SCP1's
 * Gaussian averages, SCP1::average (SCP1.cpp:33-86), restated on one process
 * in the reference's per-point order.  The reference's MPI split of the
 * index range (:53-54) and its three Allreduce sums (:71-78) become one
 * loop over the whole range.  Parity pins: tests/test_scp1.py (Sobol
 * normals and potential gradients pinned against the reference build; the
 * small-width limit <UX> -> grad V(q), <UX q^T> -> Hess V(q) Cov(dq) against
 * finite differences of the reference build's gradient; the harmonic
 * closed form of the iteration).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_sobol_normals(int dims, int64_t seed0, int64_t npts, int halve, double *y);
double orc_qtip4pf_grad(int64_t nw, const double *crd, double *grad);
double orc_ttm3f_grad(int64_t nw, const double *crd, double *grad, double tol, int *iters)
    __attribute__((weak));

static const double S1_BOHR = 0.52917721092;            /* Constants.cpp:12 */
static const double S1_AUTOKCALPMOL = 627.509474277194; /* Constants.h:21 */

/*
 * q (n3, bohr), d (n3), U = isqrtM_U (n3 x n3, column-major); pes 0 =
 * q-TIP4P/F, 1 = TTM3-F.  Outputs divided by npts (SCP1.cpp:81-83).
 */
int orc_scp1_average(int nw, int pes, const double *q, const double *d, const double *U,
                     int64_t skip, int64_t npts, double *Upot, double *UX, double *UXq)
{
    if (pes == 1 && !orc_ttm3f_grad) return -4;
    const int n3 = 9 * nw;
    double *z = (double *)malloc(sizeof(double) * (size_t)n3 * (size_t)npts);
    double *y = (double *)malloc(sizeof(double) * (size_t)n3 * 4);
    if (!z || !y) { free(z); free(y); return -2; }
    double *dq = y + n3, *r = dq + n3, *g = r + n3;
    /* SCP1.cpp:55-57: point i is std_normal at index sobol_skip + i */
    int rc = orc_sobol_normals(n3, skip, npts, 0, z);
    if (rc) { free(z); free(y); return rc; }
    double U0 = 0.0;
    memset(UX, 0, sizeof(double) * n3);
    memset(UXq, 0, sizeof(double) * (size_t)n3 * n3);
    for (int64_t i = 0; i < npts; i++) {
        for (int k = 0; k < n3; k++) y[k] = z[i * n3 + k] * sqrt(d[k]);   /* :60 */
        for (int k = 0; k < n3; k++) dq[k] = 0.0;                         /* :62 */
        for (int c = 0; c < n3; c++)
            for (int k = 0; k < n3; k++) dq[k] += U[(size_t)c * n3 + k] * y[c];
        for (int k = 0; k < n3; k++) r[k] = S1_BOHR * (q[k] + dq[k]);   /* :63 */
        const double V = pes == 1 ? orc_ttm3f_grad(nw, r, g, 1e-15, 0) : orc_qtip4pf_grad(nw, r, g);
        U0 += V / S1_AUTOKCALPMOL;                                       /* :64 */
        for (int k = 0; k < n3; k++) g[k] *= S1_BOHR / S1_AUTOKCALPMOL;  /* :65 */
        for (int k = 0; k < n3; k++) UX[k] += g[k];                      /* :67 */
        for (int c = 0; c < n3; c++)                                     /* :68 */
            for (int k = 0; k < n3; k++) UXq[(size_t)c * n3 + k] += g[k] * dq[c];
    }
    *Upot = U0 / (double)npts;
    for (int k = 0; k < n3; k++) UX[k] /= (double)npts;
    for (size_t e = 0; e < (size_t)n3 * n3; e++) UXq[e] /= (double)npts;
    free(z);
    free(y);
    return 0;
}
