/*
 * CPU FP64 oracle for SCP3's gradient (Stein-identity) estimator of the
 * vibrational Hamiltonian: SCP3_a with use_gradient (SCP3.cpp:203-231) and
 * ScaledMatrixElements::addEpot(q, V, Vq, M) (ScaledMatrixElements.cpp:
 * 100-725 with the per-element formulas of TripleExcitations.h).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, never by the product library.
 *
 * Every formula of TripleExcitations.h is one instance of a single rule.
 * For a matrix element <a|V|b> with product polynomial P(q) = phi_a phi_b,
 * D = the modes excited in a or b, and q ~ exp(-q^2) per mode,
 *
 *   est(a,b) = 1/|D| sum_{d in D} [ V (P|_{q_d=0} + d/dq_d g_d) + V_d g_d ],
 *   g_d = (P - P|_{q_d=0}) / (2 q_d),
 *
 * which is unbiased by Stein's identity E[V_d g] = E[V (2 q_d g - d_d g)].
 * The oracle evaluates the rule directly with 1-D polynomial algebra per
 * mode (no factorisation), so it is independent of the GPU formulation.
 *
 * form 1 (SCP3G_EST_GRADIENT_STEIN): every element from the rule; (0,0) = sum V
 * (ScaledMatrixElements.cpp:717).
 * form 0 (SCP3G_EST_GRADIENT): the reference's element placement, which
 * differs from the rule in five element families:
 *   column h1(k)h1(l), row h1(q_k): the rule for row h2(q_k)
 *       (v[k_index(m)] = m2_f_m1_l1(k, l), ScaledMatrixElements.cpp:195-199);
 *   column h1(k)h1(l), row h2(q_k): m2_f_k1_l1(m, k, l) at m = k, i.e. the
 *       disjoint-mode formula with mode m's values taken from mode k
 *       (ScaledMatrixElements.cpp:192-194, never overwritten);
 *   column h3(k), row h1(q_k): the rule for row h2(q_k) (m2_f_m3,
 *       ScaledMatrixElements.cpp:264-265);
 *   column h3(k), row h2(q_k): m2_f_k3(m, k) at m = k (:261-263);
 *   diagonal h3(k): m3_f_m3 (TripleExcitations.h:763-773), which has
 *       V (q^2 - 3) where the rule gives V (10 q^2 - 3) inside
 *       (2q^2 - 3)(...)/6, i.e. rule + V (4.5 q^2 - 3 q^4).
 * Same-mode rows of the other loops follow the rule.  Both forms need
 * nmodes3 <= nmodes2 (the reference indexes h2 rows of triple modes).
 *
 * Parity pins (tests/test_grad.py): known-answer values of reference
 * formulas (TripleExcitations.h, retyped with line citations) at random
 * points; Gauss-Hermite exactness of form 1 against the -G estimator for
 * polynomial potentials; form 0 equal to form 1 outside the five families;
 * q-TIP4P/F gradients against the reference build and finite differences.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int64_t orc_basis_size(int nmodes, int nmodes2, int nmodes3);
int orc_sobol_normals(int dims, int64_t seed0, int64_t npts, int halve, double *y);

static const double G_BOHR = 0.52917721092;
static const double G_AUTOKCALPMOL = 627.509474277194;

/* Hermite h_0..h_3 as coefficients of x^0..x^3 (HarmonicOscillatorBasis.cpp:16-46) */
static void hpoly(int i, double c[4])
{
    const double s2 = sqrt(2.0), s3 = sqrt(3.0);
    c[0] = c[1] = c[2] = c[3] = 0.0;
    switch (i) {
    case 0: c[0] = 1.0; break;
    case 1: c[1] = s2; break;
    case 2: c[2] = 2.0 / s2; c[0] = -1.0 / s2; break;
    case 3: c[3] = 2.0 / s3; c[1] = -3.0 / s3; break;
    }
}

static double peval(const double *c, int deg, double x)
{
    double s = 0.0;
    for (int k = deg; k >= 0; k--) s = s * x + c[k];
    return s;
}

/*
 * The rule at one point for a product element over nslot factor slots.
 * Slot s carries bra degree da[s], ket degree db[s], coordinate q[s] and
 * gradient component vd[s]; slots with da = db = 0 are ignored.  V is the
 * weight bV of the point.  Slots normally hold distinct modes; giving two
 * slots the same (q, vd) evaluates a disjoint-mode formula at coinciding
 * modes, as the reference's aliased placements do.
 */
double orc_grad_rule(int nslot, const int *da, const int *db, const double *q, const double *vd,
                     double V)
{
    double C[6], c0[6], G[6], Gp[6];
    int live[6], nd = 0;
    for (int s = 0; s < nslot; s++) {
        live[s] = da[s] + db[s] > 0;
        if (!live[s]) { C[s] = 1.0; continue; }
        nd++;
        double a[4], b[4], c[7] = {0, 0, 0, 0, 0, 0, 0};
        hpoly(da[s], a);
        hpoly(db[s], b);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) c[i + j] += a[i] * b[j];
        C[s] = peval(c, 6, q[s]);
        c0[s] = c[0];
        /* G(x) = (c(x) - c(0)) / (2x) = sum_{k>=1} c_k x^{k-1} / 2 */
        double g[6], gp[5];
        for (int k = 1; k <= 6; k++) g[k - 1] = 0.5 * c[k];
        for (int k = 1; k <= 5; k++) gp[k - 1] = k * g[k];
        G[s] = peval(g, 5, q[s]);
        Gp[s] = peval(gp, 4, q[s]);
    }
    if (nd == 0) return V;
    double est = 0.0;
    for (int d = 0; d < nslot; d++) {
        if (!live[d]) continue;
        double rest = 1.0;
        for (int s = 0; s < nslot; s++)
            if (s != d) rest *= C[s];
        est += (V * (c0[d] + Gp[d]) + vd[d] * G[d]) * rest;
    }
    return est / nd;
}

/* basis function r as (mode, degree) x 3 in get_bra's order (ScaledMatrixElements.cpp:870-930) */
static void basis_desc(int nsel, int n2, int n3, int *md /* n x 3 modes */, int *dg /* n x 3 */)
{
    int64_t R = 0;
#define PUT(m0, d0, m1, d1, m2, d2)                                            \
    do {                                                                       \
        md[3 * R] = m0; md[3 * R + 1] = m1; md[3 * R + 2] = m2;                \
        dg[3 * R] = d0; dg[3 * R + 1] = d1; dg[3 * R + 2] = d2;                \
        R++;                                                                   \
    } while (0)
    PUT(0, 0, 0, 0, 0, 0);
    for (int m = 0; m < nsel; m++) PUT(m, 1, 0, 0, 0, 0);
    for (int m = 0; m < n2; m++) PUT(m, 2, 0, 0, 0, 0);
    for (int m = 0; m < n2; m++)
        for (int n = m + 1; n < n2; n++) PUT(m, 1, n, 1, 0, 0);
    for (int m = 0; m < n3; m++) PUT(m, 3, 0, 0, 0, 0);
    for (int m = 0; m < n3; m++)
        for (int n = m + 1; n < n3; n++) {
            PUT(m, 2, n, 1, 0, 0);
            PUT(m, 1, n, 2, 0, 0);
        }
    for (int m = 0; m < n3; m++)
        for (int n = m + 1; n < n3; n++)
            for (int p = n + 1; p < n3; p++) PUT(m, 1, n, 1, p, 1);
#undef PUT
}

/* the rule for basis elements (ra, rb) at one point; q, vq over the selected modes */
static double elem_rule(const int *md, const int *dg, int64_t ra, int64_t rb, const double *q,
                        const double *vq, double V)
{
    int mode[6], da[6], db[6], ns = 0;
    double qs[6], vs[6];
    for (int side = 0; side < 2; side++) {
        const int64_t r = side ? rb : ra;
        for (int t = 0; t < 3; t++) {
            const int d = dg[3 * r + t];
            if (!d) continue;
            const int m = md[3 * r + t];
            int s = 0;
            while (s < ns && mode[s] != m) s++;
            if (s == ns) {
                mode[ns] = m; da[ns] = db[ns] = 0; qs[ns] = q[m]; vs[ns] = vq[m];
                ns++;
            }
            if (side) db[s] = d; else da[s] = d;
        }
    }
    return orc_grad_rule(ns, da, db, qs, vs, V);
}

/* index of basis functions (get_bra order) */
static int64_t idx_h1(int k) { return 1 + k; }

/*
 * M (n x n column-major, upper) += sum over points of the per-point element
 * estimators.  q and vq: npts x nsel point-major (selected modes), V: the
 * weights bV.  form 0 = the reference's placement, 1 = the rule throughout.
 */
int orc_grad_accumulate(int nsel, int n2, int n3, int form, int64_t npts, const double *q,
                        const double *V, const double *vq, double *M)
{
    if (n2 > nsel || n3 > nsel || n3 > n2) return -4;
    const int64_t n = orc_basis_size(nsel, n2, n3);
    int *md = (int *)malloc(sizeof(int) * 3 * (size_t)n);
    int *dg = (int *)malloc(sizeof(int) * 3 * (size_t)n);
    if (!md || !dg) { free(md); free(dg); return -2; }
    basis_desc(nsel, n2, n3, md, dg);
    const int64_t h2_0 = 1 + nsel, kl_0 = h2_0 + n2;
    const int64_t h3_0 = kl_0 + (int64_t)n2 * (n2 - 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < n; c++) {
        for (int64_t r = 0; r <= c; r++) {
            double s = 0.0;
            if (r == 0 && c == 0) {
                for (int64_t p = 0; p < npts; p++) s += V[p];
                M[c * n + r] += s;
                continue;
            }
            /* reference placements (form 0) */
            int kind = 0, k = -1, l = -1;
            if (form == 0) {
                if (c >= kl_0 && c < h3_0 && dg[3 * c] == 1 && dg[3 * c + 1] == 1) {
                    const int ck = md[3 * c], cl = md[3 * c + 1];
                    if (r == idx_h1(ck)) { kind = 1; k = ck; l = cl; }
                    else if (r == idx_h1(cl)) { kind = 1; k = cl; l = ck; }
                    else if (r == h2_0 + ck) { kind = 2; k = ck; l = cl; }
                    else if (r == h2_0 + cl) { kind = 2; k = cl; l = ck; }
                } else if (c >= h3_0 && c < h3_0 + n3) {
                    const int ck = md[3 * c];
                    if (r == idx_h1(ck)) { kind = 3; k = ck; }
                    else if (r == h2_0 + ck) { kind = 4; k = ck; }
                    else if (r == c) { kind = 5; k = ck; }
                }
            }
            for (int64_t p = 0; p < npts; p++) {
                const double *qp = q + p * nsel, *vp = vq + p * nsel;
                double e;
                switch (kind) {
                case 1: e = elem_rule(md, dg, h2_0 + k, c, qp, vp, V[p]); break;
                case 2: {
                    /* m2_f_k1_l1(m, k, l) with m = k: bra h2(q_m), ket h1 h1 */
                    const int da[3] = {2, 0, 0}, db[3] = {0, 1, 1};
                    const double qs[3] = {qp[k], qp[k], qp[l]}, vs[3] = {vp[k], vp[k], vp[l]};
                    e = orc_grad_rule(3, da, db, qs, vs, V[p]);
                    break;
                }
                case 3: e = elem_rule(md, dg, h2_0 + k, c, qp, vp, V[p]); break;
                case 4: {
                    /* m2_f_k3(m, k) with m = k: bra h2(q_m), ket h3(q_k) */
                    const int da[2] = {2, 0}, db[2] = {0, 3};
                    const double qs[2] = {qp[k], qp[k]}, vs[2] = {vp[k], vp[k]};
                    e = orc_grad_rule(2, da, db, qs, vs, V[p]);
                    break;
                }
                case 5: {
                    const double x2 = qp[k] * qp[k];
                    e = elem_rule(md, dg, r, c, qp, vp, V[p]) + V[p] * (4.5 * x2 - 3.0 * x2 * x2);
                    break;
                }
                default: e = elem_rule(md, dg, r, c, qp, vp, V[p]);
                }
                s += e;
            }
            M[c * n + r] += s;
        }
    }
    free(md);
    free(dg);
    return 0;
}

/* ------------------------------------------------------------------ */
/* q-TIP4P/F energy and gradient, qtip4pf.cpp:57-107,252-397            */
/* ------------------------------------------------------------------ */

static const double QG_CHARGECON = 18.22234397655801030455;
#define QG_qM (-1.1128 * QG_CHARGECON)
#define QG_qH (-QG_qM / 2)
static const double QG_rOHeq = 0.9419, QG_aHOHeq = 107.4;
static const double QG_epsilon = 0.1852, QG_sigma = 3.1589;
static const double QG_alpha_r = 2.287, QG_D_r = 116.09, QG_k_theta = 87.85;
static const double QG_gammaM = 0.73612;
#define QG_gamma1 ((1.0 - QG_gammaM) / 2)

/* monomer energy and gradient (qtip4pf.cpp:57-107); g receives 9 values */
static double qg_monomer(const double *w, double *g)
{
    double d1[3], d2[3], r1 = 0, r2 = 0, d12 = 0;
    for (int i = 0; i < 3; i++) {
        d1[i] = w[i + 3] - w[i];
        r1 += d1[i] * d1[i];
        d2[i] = w[i + 6] - w[i];
        r2 += d2[i] * d2[i];
        d12 += d1[i] * d2[i];
    }
    r1 = sqrt(r1);
    r2 = sqrt(r2);
    const double x1 = QG_alpha_r * (r1 - QG_rOHeq);
    const double x2 = QG_alpha_r * (r2 - QG_rOHeq);
    const double V1 = x1 * x1 * (1.0 - x1 * (1.0 - (7.0 / 12.0) * x1)) * QG_D_r;
    const double V2 = x2 * x2 * (1.0 - x2 * (1.0 - (7.0 / 12.0) * x2)) * QG_D_r;
    const double r12 = 1.0 / r1 / r2;
    const double xi = d12 * r12;
    const double theta = acos(xi);
    const double d_theta = theta - QG_aHOHeq * M_PI / 180.0;
    /* dV/dr of the Morse-quartic terms and d(theta)/d(xi) = -1/sqrt(1-xi^2) */
    const double dV1 = QG_alpha_r * QG_D_r * x1 * (2.0 - x1 * (3.0 - (7.0 / 3.0) * x1)) / r1;
    const double dV2 = QG_alpha_r * QG_D_r * x2 * (2.0 - x2 * (3.0 - (7.0 / 3.0) * x2)) / r2;
    const double f = -QG_k_theta * d_theta / sqrt(1.0 - xi * xi);
    const double t0 = f * r12;
    const double t1 = dV1 - f * xi / r1 / r1;
    const double t2 = dV2 - f * xi / r2 / r2;
    for (int i = 0; i < 3; i++) {
        const double a1 = t0 * d2[i] + t1 * d1[i];
        const double a2 = t0 * d1[i] + t2 * d2[i];
        g[i] = -(a1 + a2);
        g[i + 3] = a1;
        g[i + 6] = a2;
    }
    return V1 + V2 + 0.5 * QG_k_theta * d_theta * d_theta;
}

/* qtip4pf::operator()(nw, crd, grad), qtip4pf.cpp:252-397; crd in A, O H H */
double orc_qtip4pf_grad(int64_t nw, const double *crd, double *grad)
{
    double *ms = (double *)calloc(6 * (size_t)(nw > 0 ? nw : 1), sizeof(double));
    double *dM = ms + 3 * nw;
    double E = 0.0;
    for (int64_t i = 0; i < nw; i++) {
        const double *O = crd + 9 * i;
        for (int k = 0; k < 3; k++) ms[3 * i + k] = QG_gammaM * O[k] + QG_gamma1 * (O[3 + k] + O[6 + k]);
        E += qg_monomer(O, grad + 9 * i);
    }
    double Evdw = 0, Eel = 0;
    for (int64_t i = 0; i < nw; i++) {
        for (int64_t j = 0; j < nw; j++) { /* H(j) - M(i), :283-310 */
            if (i == j) continue;
            for (int l = 1; l < 3; l++) {
                const double *H = crd + 9 * j + 3 * l;
                double dR[3], R2 = 0;
                for (int k = 0; k < 3; k++) { dR[k] = ms[3 * i + k] - H[k]; R2 += dR[k] * dR[k]; }
                const double r1 = 1.0 / sqrt(R2), E1 = QG_qH * QG_qM * r1, g1 = E1 * r1 * r1;
                Eel += E1;
                for (int k = 0; k < 3; k++) {
                    dM[3 * i + k] -= g1 * dR[k];
                    grad[9 * j + 3 * l + k] += g1 * dR[k];
                }
            }
        }
        for (int64_t j = i + 1; j < nw; j++) {
            for (int a = 1; a < 3; a++) /* H-H, :317-340 */
                for (int b = 1; b < 3; b++) {
                    double dR[3], R2 = 0;
                    for (int k = 0; k < 3; k++) {
                        dR[k] = crd[9 * i + 3 * a + k] - crd[9 * j + 3 * b + k];
                        R2 += dR[k] * dR[k];
                    }
                    const double r1 = 1.0 / sqrt(R2), E1 = QG_qH * QG_qH * r1, g1 = E1 * r1 * r1;
                    Eel += E1;
                    for (int k = 0; k < 3; k++) {
                        grad[9 * i + 3 * a + k] -= g1 * dR[k];
                        grad[9 * j + 3 * b + k] += g1 * dR[k];
                    }
                }
            double dR[3], R2 = 0; /* O-O, :343-357 */
            for (int k = 0; k < 3; k++) { dR[k] = crd[9 * i + k] - crd[9 * j + k]; R2 += dR[k] * dR[k]; }
            const double s2 = QG_sigma * QG_sigma / R2, d6 = s2 * s2 * s2;
            Evdw += 4 * QG_epsilon * d6 * (d6 - 1.0);
            const double gv = 24 * QG_epsilon * d6 * (2 * d6 - 1.0) / R2;
            for (int k = 0; k < 3; k++) {
                grad[9 * i + k] -= gv * dR[k];
                grad[9 * j + k] += gv * dR[k];
            }
            R2 = 0; /* M-M, :360-374 */
            for (int k = 0; k < 3; k++) { dR[k] = ms[3 * i + k] - ms[3 * j + k]; R2 += dR[k] * dR[k]; }
            const double r1 = 1.0 / sqrt(R2), E1 = QG_qM * QG_qM * r1, g1 = E1 * r1 * r1;
            Eel += E1;
            for (int k = 0; k < 3; k++) {
                dM[3 * i + k] -= g1 * dR[k];
                dM[3 * j + k] += g1 * dR[k];
            }
        }
    }
    for (int64_t i = 0; i < nw; i++) /* M-site chain rule, :379-390 */
        for (int k = 0; k < 3; k++) {
            grad[9 * i + k] += QG_gammaM * dM[3 * i + k];
            grad[9 * i + 3 + k] += QG_gamma1 * dM[3 * i + k];
            grad[9 * i + 6 + k] += QG_gamma1 * dM[3 * i + k];
        }
    free(ms);
    return E + Evdw + Eel;
}

/*
 * The SCP3_a block loop with use_gradient (SCP3.cpp:187-232), pes 0 =
 * q-TIP4P/F, 1 = TTM3-F: y (seeds seed0.. or the caller's normals, halved),
 * r = bohr (x0 + MUa y), V and grad V, bV = V/autokcal - y.(omega o y)/2,
 * bVy = MUa^T grad V bohr/autokcal - omega o y, then the estimator over the
 * selected modes.  M (n x n, upper) gains the raw sums; V_out/bV_out/Vq_out
 * (npts x nsel) on request.
 */
double orc_ttm3f_grad(int64_t nw, const double *crd, double *grad, double tol, int *iters)
    __attribute__((weak));

int orc_scp3_accumulate_grad(int nw, int nmodes0, const double *x0, const double *omega,
                             const double *MUa, int nsel, const int *modes, int nmodes2,
                             int nmodes3, int form, int pes, int64_t seed0, int64_t npts,
                             const double *normals, double *M, double *bV_out, double *Vq_out)
{
    if (pes == 1 && !orc_ttm3f_grad) return -4;
    const int n3 = 9 * nw;
    double *y = (double *)malloc(sizeof(double) * (size_t)nmodes0 * (size_t)npts);
    double *q = (double *)malloc(sizeof(double) * (size_t)(nsel ? nsel : 1) * (size_t)npts);
    double *vq = (double *)malloc(sizeof(double) * (size_t)(nsel ? nsel : 1) * (size_t)npts);
    double *bV = (double *)malloc(sizeof(double) * (size_t)npts);
    if (!y || !q || !vq || !bV) { free(y); free(q); free(vq); free(bV); return -2; }
    if (normals) {
        for (size_t k = 0; k < (size_t)nmodes0 * (size_t)npts; k++) y[k] = normals[k] / sqrt(2.0);
    } else {
        int rc = orc_sobol_normals(nmodes0, seed0, npts, 1, y);
        if (rc) { free(y); free(q); free(vq); free(bV); return rc; }
    }
#pragma omp parallel
    {
        double *r = (double *)malloc(sizeof(double) * 2 * n3);
        double *g = r + n3;
#pragma omp for schedule(static)
        for (int64_t j = 0; j < npts; j++) {
            const double *yj = y + j * nmodes0;
            for (int c = 0; c < n3; c++) {
                double s = 0.0;
                for (int k = 0; k < nmodes0; k++) s += MUa[(size_t)k * n3 + c] * yj[k];
                r[c] = G_BOHR * (x0[c] + s);
            }
            const double V = (pes == 1 ? orc_ttm3f_grad(nw, r, g, 1e-15, 0) : orc_qtip4pf_grad(nw, r, g)) /
                             G_AUTOKCALPMOL;
            double h = 0.0;
            for (int k = 0; k < nmodes0; k++) h += yj[k] * (omega[k] * yj[k]);
            bV[j] = V - 0.5 * h;
            for (int m = 0; m < nsel; m++) {
                const int k = modes[m];
                double s = 0.0;
                for (int c = 0; c < n3; c++) s += MUa[(size_t)k * n3 + c] * (g[c] * (G_BOHR / G_AUTOKCALPMOL));
                q[j * nsel + m] = yj[k];
                vq[j * nsel + m] = s - omega[k] * yj[k];
            }
        }
        free(r);
    }
    int rc = 0;
    if (M) rc = orc_grad_accumulate(nsel, nmodes2, nmodes3, form, npts, q, bV, vq, M);
    if (bV_out) memcpy(bV_out, bV, sizeof(double) * (size_t)npts);
    if (Vq_out) memcpy(Vq_out, vq, sizeof(double) * (size_t)nsel * (size_t)npts);
    free(y);
    free(q);
    free(vq);
    free(bV);
    return rc;
}
