/*
 * CPU FP64 oracle for the TTM3-F energy.  TEST INFRASTRUCTURE ONLY (see
 * scp3_oracle.c): restates ttm3f::operator()(nw, crd) (ttm3f.cpp:92-321)
 * with the Partridge-Schwenke monomer potential ps::pot_nasa (ps.cpp:486-603,
 * energy branch), the dipole surface ps::dms_nasa (ps.cpp:605-770, charge
 * branch with the TTM3 modification) and the smeared Coulomb kernels
 * (ttm3f-bits-smear.cpp:17-171).  Operation order follows the reference.
 *
 * One deliberate difference: the reference never clears its field-from-
 * dipoles buffer Efd before the first iteration (ttm3f.cpp:112 zeroes only
 * Efq, phi and part of ddt), so the first iterate starts from whatever the
 * buffer held.  The iteration is a contraction with a unique fixed point, so
 * the converged dipoles (tolerance 1e-15) do not depend on it; here Efd starts
 * at zero, which makes the function pure.
 *
 * Pins (tests/test_ttm3f.py): energies of the reference build oracle/_ref on
 * the same configurations; gammq23 against scipy's regularised incomplete
 * gamma; induced dipoles against a direct solve of the linear system the
 * iteration converges to.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../scp3_amd/csrc/ps_coeffs.inc"

/* ps.cpp:11-18,472-477 */
static const double PS_reoh = 0.958649, PS_b1 = 2.0, PS_roh = 0.9519607159623009;
static const double PS_alphaoh = 2.587949757553683, PS_deohA = 42290.92019288289;
static const double PS_phh1A = 16.94879431193463, PS_phh2 = 12.66426998162947;
static const double PS_f5z = 0.999677885;
static const double PS_b1D = 1.0, PS_a = 0.2999e0, PS_b = -0.6932e0;
static const double PS_c0 = 1.0099e0, PS_c1 = -0.1801e0, PS_c2 = 0.0892e0;

/* ps::pot_nasa energy (ps.cpp:486-575), kcal/mol */
double orc_pot_nasa(const double *rr)
{
    double ROH1[3], ROH2[3], RHH[3], dROH1 = 0, dROH2 = 0, dRHH = 0;
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = rr[3 + i] - rr[i];
        ROH2[i] = rr[6 + i] - rr[i];
        RHH[i] = rr[3 + i] - rr[6 + i];
        dROH1 += ROH1[i] * ROH1[i];
        dROH2 += ROH2[i] * ROH2[i];
        dRHH += RHH[i] * RHH[i];
    }
    dROH1 = sqrt(dROH1);
    dROH2 = sqrt(dROH2);
    dRHH = sqrt(dRHH);
    const double costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (dROH1 * dROH2);
    const double deoh = PS_f5z * PS_deohA;
    const double phh1 = PS_f5z * PS_phh1A * exp(PS_phh2);
    const double costhe = -.24780227221366464506;
    const double exp1 = exp(-PS_alphaoh * (dROH1 - PS_roh));
    const double exp2 = exp(-PS_alphaoh * (dROH2 - PS_roh));
    const double Va = deoh * (exp1 * (exp1 - 2.0) + exp2 * (exp2 - 2.0));
    const double Vb = phh1 * exp(-PS_phh2 * dRHH);
    const double x1 = (dROH1 - PS_reoh) / PS_reoh;
    const double x2 = (dROH2 - PS_reoh) / PS_reoh;
    const double x3 = costh - costhe;
    double fmat[3][16];
    for (int i = 0; i < 3; ++i) {
        fmat[i][0] = 0.0;
        fmat[i][1] = 1.0;
    }
    for (int j = 2; j < 16; ++j) {
        fmat[0][j] = fmat[0][j - 1] * x1;
        fmat[1][j] = fmat[1][j - 1] * x2;
        fmat[2][j] = fmat[2][j - 1] * x3;
    }
    const double efac = exp(-PS_b1 * (pow((dROH1 - PS_reoh), 2) + pow((dROH2 - PS_reoh), 2)));
    double sum0 = 0;
    for (int j = 1; j < 245; ++j) {
        const int inI = scp3_ps_idx1[j], inJ = scp3_ps_idx2[j], inK = scp3_ps_idx3[j];
        sum0 += scp3_ps_c5z[j] * (fmat[0][inI] * fmat[1][inJ] + fmat[0][inJ] * fmat[1][inI]) * fmat[2][inK];
    }
    const double Vc = 2 * scp3_ps_c5z[0] + efac * sum0;
    double e1 = Va + Vb + Vc;
    e1 += 0.44739574026257;
    e1 *= 0.00285914375100642899;
    return e1;
}

/* ps::dms_nasa charges (ps.cpp:605-770) with the TTM3 modification */
void orc_dms_nasa(double p1_, double p2_, double p3_, const double *rr, double *q3)
{
    const double ath0 = 1.82400520401572996557;
    const double costhe = -0.24780227221366464506;
    double ROH1[3], ROH2[3], RHH[3], dROH1 = 0, dROH2 = 0, dRHH = 0;
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = rr[3 + i] - rr[i];
        ROH2[i] = rr[6 + i] - rr[i];
        RHH[i] = rr[3 + i] - rr[6 + i];
        dROH1 += ROH1[i] * ROH1[i];
        dROH2 += ROH2[i] * ROH2[i];
        dRHH += RHH[i] * RHH[i];
    }
    dROH1 = sqrt(dROH1);
    dROH2 = sqrt(dROH2);
    dRHH = sqrt(dRHH);
    const double costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (dROH1 * dROH2);
    const double efac = exp(-PS_b1D * (pow((dROH1 - PS_reoh), 2) + pow((dROH2 - PS_reoh), 2)));
    const double x1 = (dROH1 - PS_reoh) / PS_reoh;
    const double x2 = (dROH2 - PS_reoh) / PS_reoh;
    const double x3 = costh - costhe;
    double fmat[3][16];
    for (int i = 0; i < 3; ++i) {
        fmat[i][0] = 0.0;
        fmat[i][1] = 1.0;
    }
    for (int j = 2; j < 16; ++j) {
        fmat[0][j] = fmat[0][j - 1] * x1;
        fmat[1][j] = fmat[1][j - 1] * x2;
        fmat[2][j] = fmat[2][j - 1] * x3;
    }
    double p1 = 0, p2 = 0;
    const double pl1 = costh;
    const double pl2 = 0.5 * (3 * pl1 * pl1 - 1.0);
    for (int j = 1; j < 84; ++j) {
        const int inI = scp3_ps_idxD0[j], inJ = scp3_ps_idxD1[j], inK = scp3_ps_idxD2[j];
        p1 += scp3_ps_coefD[j] * fmat[0][inI] * fmat[1][inJ] * fmat[2][inK];
        p2 += scp3_ps_coefD[j] * fmat[0][inJ] * fmat[1][inI] * fmat[2][inK];
    }
    const double xx = 0.529177249;
    const double pc0 = PS_a * (pow(dROH1, PS_b) + pow(dROH2, PS_b)) * (PS_c0 + pl1 * PS_c1 + pl2 * PS_c2);
    p1 = scp3_ps_coefD[0] + p1 * efac + pc0 * xx;
    p2 = scp3_ps_coefD[0] + p2 * efac + pc0 * xx;
    q3[0] = -(p1 + p2);
    q3[1] = p1;
    q3[2] = p2;
    /* TTM3 modification, ps.cpp:732-759 */
    double AxB[3];
    AxB[0] = ROH1[1] * ROH2[2] - ROH1[2] * ROH2[1];
    AxB[1] = -ROH1[0] * ROH2[2] + ROH1[2] * ROH2[0];
    AxB[2] = ROH1[0] * ROH2[1] - ROH1[1] * ROH2[0];
    double s0 = 0;
    for (int i = 0; i < 3; ++i) s0 += pow((AxB[i]), 2);
    const double dAxB = sqrt(s0);
    const double sinth = dAxB / (dROH1 * dROH2);
    const double ang = atan2(sinth, costh);
    p1 = p1_ * (dROH1 - p2_) + p3_ * (ang - ath0);
    p2 = p1_ * (dROH2 - p2_) + p3_ * (ang - ath0);
    q3[0] = q3[0] - (p1 + p2);
    q3[1] = q3[1] + p1;
    q3[2] = q3[2] + p2;
}

/* gammq(2/3, x) * Gamma(2/3), ttm3f-bits-smear.cpp:17-80 */
static double orc_gser23(double x)
{
    const double a = 2.0 / 3.0;
    double ap = a, sum = 1.0 / a, del = sum;
    for (;;) {
        ++ap;
        del *= x / ap;
        sum += del;
        if (fabs(del) < fabs(sum) * DBL_EPSILON) return sum * exp(-x + a * log(x));
    }
}

static double orc_gcf23(double x)
{
    const double a = 2.0 / 3.0;
    const double FPMIN = DBL_MIN / DBL_EPSILON;
    double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
    for (int i = 1;; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < FPMIN) d = FPMIN;
        c = b + an / c;
        if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) <= DBL_EPSILON) break;
    }
    return exp(-x + a * log(x)) * h;
}

double orc_gammq23(double x)
{
    const double exp_gammln_23 = 1.354117939426400e+00;
    return (x < 5.0 / 3.0 ? exp_gammln_23 - orc_gser23(x) : orc_gcf23(x));
}

static const double SM_aXX = 0.175;

/* smear0 / smear01 / smear012, ttm3f-bits-smear.cpp:92-171 */
static void orc_smear(double r12, double AA, int want, double *ts0, double *ts1, double *ts2)
{
    const double a13 = pow(SM_aXX, 1.0 / 3.0);
    const double dri = 1.0 / r12;
    const double drsqi = dri * dri;
    const double rA = r12 / AA;
    const double rA3 = rA * rA * rA;
    const double exp1 = exp(-SM_aXX * rA3);
    *ts0 = (1.0 - exp1 + a13 * rA * orc_gammq23(SM_aXX * rA3)) * dri;
    if (want >= 1) *ts1 = (1.0 - exp1) * dri * drsqi;
    if (want >= 2) *ts2 = (*ts1 - exp1 * SM_aXX / (AA * AA * AA)) * drsqi;
}

void orc_smear012(double r, double AA, double *ts)
{
    orc_smear(r, AA, 2, ts, ts + 1, ts + 2);
}

/* ttm3f.cpp:16-33, ttm3f-bits.h:10-26 */
static const double T_polarM = 1.444, T_polfacH = 0.496, T_polfacM = 0.837;
static const double T_vdwC = -0.72298855E+03, T_vdwD = 0.10211829E+06, T_vdwE = 0.37170376E+01;
static const double T_gammaM = 0.46;
#define T_gamma1 (1.0 - T_gammaM)
#define T_gamma2 (T_gammaM / 2)
static const double T_CHARGECON = 18.22234397655801030455;
static const double T_DEBYE = 4.8033324;

/*
 * ttm3f::operator()(nw, crd), energy branch, ttm3f.cpp:92-321.  iters
 * (nullable) receives the number of dipole iterations; dip_out (nullable,
 * 3nw) the converged induced dipoles; Efq_out (3nw) and ddt_out (3nw x 3nw)
 * the charge field and dipole tensor of the linear system the iteration
 * solves, dip = polarM * (Efq + ddt dip).
 */
double orc_ttm3f_full(int64_t nw, const double *crd, double tol, int *iters, double *dip_out,
                      double *Efq_out, double *ddt_out)
{
    const double AA_HH = pow(T_polfacH * T_polfacH, 1.0 / 6.0);
    const double AA_HM = pow(T_polfacH * T_polfacM, 1.0 / 6.0);
    const double AA_MM = pow(T_polfacM * T_polfacM, 1.0 / 6.0);
    const int64_t n3 = 3 * nw;
    double *mem = (double *)calloc((size_t)(7 * n3 + n3 * n3), sizeof(double));
    double *msite = mem, *charge = msite + n3, *dip = charge + n3, *dip_prev = dip + n3;
    double *Efd = dip_prev + n3, *Efq = Efd + n3, *phi = Efq + n3, *ddt = phi + n3;
    double Eint = 0;
    for (int64_t i = 0; i < nw; ++i) {
        const double *w = crd + 9 * i;
        for (int k = 0; k < 3; ++k) msite[3 * i + k] = T_gamma1 * w[k] + T_gamma2 * (w[3 + k] + w[6 + k]);
        Eint += orc_pot_nasa(w);
        double q3[3];
        orc_dms_nasa(0.5, 0.9578, 0.012, w, q3);
        const double tmp = T_gamma2 / T_gamma1;
        charge[3 * i + 0] = T_CHARGECON * q3[0] / T_gamma1;
        charge[3 * i + 1] = T_CHARGECON * (q3[1] + tmp * (q3[1] + q3[2]));
        charge[3 * i + 2] = T_CHARGECON * (q3[2] + tmp * (q3[1] + q3[2]));
    }
    double Evdw = 0;
    for (int64_t i = 0; i < nw; ++i) {
        const int64_t i3 = 3 * i, i9 = 9 * i;
        for (int64_t j = 0; j < nw; ++j) {                       /* ttm3f.cpp:148-172 */
            if (i == j) continue;
            const int64_t j3 = 3 * j;
            for (int l = 1; l < 3; ++l) {
                const int64_t j3l = j3 + l, jh = 3 * j3l;
                double Rij[3], Rsq = 0;
                for (int k = 0; k < 3; ++k) {
                    Rij[k] = msite[i3 + k] - crd[jh + k];
                    Rsq += Rij[k] * Rij[k];
                }
                double ts0, ts1, ts2;
                orc_smear(sqrt(Rsq), AA_HM, 1, &ts0, &ts1, &ts2);
                phi[i3] += ts0 * charge[j3l];
                phi[j3l] += ts0 * charge[i3];
                for (int k = 0; k < 3; ++k) Efq[i3 + k] += charge[j3l] * ts1 * Rij[k];
            }
        }
        for (int64_t j = i + 1; j < nw; ++j) {                   /* ttm3f.cpp:176-262 */
            const int64_t j3 = 3 * j, j9 = 9 * j;
            for (int a = 1; a < 3; ++a) {
                const int64_t i3a = i3 + a;
                for (int b = 1; b < 3; ++b) {
                    const int64_t j3b = j3 + b;
                    double Rsq = 0;
                    for (int k = 0; k < 3; ++k) {
                        const double dx = crd[3 * i3a + k] - crd[3 * j3b + k];
                        Rsq += dx * dx;
                    }
                    double ts0, t1, t2;
                    orc_smear(sqrt(Rsq), AA_HH, 0, &ts0, &t1, &t2);
                    phi[i3a] += charge[j3b] * ts0;
                    phi[j3b] += charge[i3a] * ts0;
                }
            }
            double Rij[3], Rsq = 0;
            for (int k = 0; k < 3; ++k) {
                Rij[k] = crd[i9 + k] - crd[j9 + k];
                Rsq += Rij[k] * Rij[k];
            }
            const double dRij = sqrt(Rsq);
            const double dR6 = Rsq * Rsq * Rsq;
            const double expon = T_vdwD * exp(-T_vdwE * dRij);
            Evdw += T_vdwC / dR6 + expon;
            Rsq = 0.0;
            for (int k = 0; k < 3; ++k) {
                Rij[k] = msite[i3 + k] - msite[j3 + k];
                Rsq += Rij[k] * Rij[k];
            }
            double ts0, ts1, ts2;
            orc_smear(sqrt(Rsq), AA_MM, 2, &ts0, &ts1, &ts2);
            phi[i3] += ts0 * charge[j3];
            phi[j3] += ts0 * charge[i3];
            for (int k = 0; k < 3; ++k) {
                Efq[i3 + k] += charge[j3] * ts1 * Rij[k];
                Efq[j3 + k] -= charge[i3] * ts1 * Rij[k];
            }
            double dd3[3][3];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) dd3[k][l] = 3.0 * ts2 * Rij[k] * Rij[l] - (k == l ? ts1 : 0.0);
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    ddt[n3 * (i3 + k) + j3 + l] = dd3[k][l];
                    ddt[n3 * (j3 + k) + i3 + l] = dd3[k][l];
                }
        }
    }
    /* induced dipoles, ttm3f.cpp:268-307 */
    for (int64_t i = 0; i < n3; ++i) {
        dip[i] = T_polarM * Efq[i];
        dip_prev[i] = dip[i];
    }
    const double dmix = 0.7;
    const double stath = T_DEBYE / T_CHARGECON / (sqrt((double)nw));
    int it = 0;
    for (; it < 500; ++it) {
        memset(Efd, 0, sizeof(double) * (size_t)n3);
        for (int64_t k = 0; k < n3; ++k)
            for (int64_t l = 0; l < n3; ++l) Efd[k] += ddt[n3 * k + l] * dip[l];
        double deltadip = 0.0;
        for (int64_t i = 0; i < n3; ++i) {
            dip[i] = T_polarM * (Efq[i] + Efd[i]);
            dip[i] = dmix * dip[i] + (1.0 - dmix) * dip_prev[i];
            const double delta = dip[i] - dip_prev[i];
            deltadip += delta * delta;
        }
        deltadip = sqrt(deltadip) * stath;
        if (deltadip < tol) break;
        memcpy(dip_prev, dip, sizeof(double) * (size_t)n3);
    }
    if (iters) *iters = it + 1;
    double Eelec = 0, Eind = 0;
    for (int64_t i = 0; i < n3; ++i) {
        Eelec += phi[i] * charge[i];
        Eind -= dip[i] * Efq[i];
    }
    Eelec /= 2;
    Eind /= 2;
    if (dip_out) memcpy(dip_out, dip, sizeof(double) * (size_t)n3);
    if (Efq_out) memcpy(Efq_out, Efq, sizeof(double) * (size_t)n3);
    if (ddt_out) memcpy(ddt_out, ddt, sizeof(double) * (size_t)(n3 * n3));
    free(mem);
    return Eint + Evdw + Eelec + Eind;
}

double orc_ttm3f(int64_t nw, const double *crd)
{
    return orc_ttm3f_full(nw, crd, 1.0e-15, NULL, NULL, NULL, NULL);
}
