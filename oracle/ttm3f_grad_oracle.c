/*
 * CPU FP64 oracle for TTM3-F energies with gradients.  TEST INFRASTRUCTURE
 * ONLY (see scp3_oracle.c): restates ttm3f::operator()(nw, crd, grad)
 * (ttm3f.cpp:325-716) with ps::pot_nasa's gradient branch (ps.cpp:486-603),
 * ps::dms_nasa with charge derivatives and the TTM3 modification
 * (ps.cpp:605-836) and the smeared kernels smear01/012/12/123
 * (ttm3f-bits-smear.cpp:116-190).  Operation order follows the reference.
 *
 * Pins (tests/test_ttm3f.py): energies and gradients of the reference build
 * oracle/_ref on the same configurations; central finite differences of
 * the energy oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../scp3_amd/csrc/ps_coeffs.inc"

double orc_gammq23(double x);

/* ps.cpp:11-18,472-477 */
static const double G_reoh = 0.958649, G_b1 = 2.0, G_roh = 0.9519607159623009;
static const double G_alphaoh = 2.587949757553683, G_deohA = 42290.92019288289;
static const double G_phh1A = 16.94879431193463, G_phh2 = 12.66426998162947;
static const double G_f5z = 0.999677885;
static const double G_b1D = 1.0, G_a = 0.2999e0, G_b = -0.6932e0;
static const double G_c0 = 1.0099e0, G_c1 = -0.1801e0, G_c2 = 0.0892e0;
static const double G_costhe = -.24780227221366464506;
static const double G_cm2kcal = 0.00285914375100642899;

/* ps::pot_nasa(rr, dr), ps.cpp:486-603: energy and gradient (kcal/mol) */
double orc_pot_nasa_grad(const double *rr, double *dr)
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
    const double deoh = G_f5z * G_deohA;
    const double phh1 = G_f5z * G_phh1A * exp(G_phh2);
    const double exp1 = exp(-G_alphaoh * (dROH1 - G_roh));
    const double exp2 = exp(-G_alphaoh * (dROH2 - G_roh));
    const double Va = deoh * (exp1 * (exp1 - 2.0) + exp2 * (exp2 - 2.0));
    const double Vb = phh1 * exp(-G_phh2 * dRHH);
    const double dVa1 = 2.0 * G_alphaoh * deoh * exp1 * (1.0 - exp1) / dROH1;
    const double dVa2 = 2.0 * G_alphaoh * deoh * exp2 * (1.0 - exp2) / dROH2;
    const double dVb = -G_phh2 * Vb / dRHH;
    const double x1 = (dROH1 - G_reoh) / G_reoh;
    const double x2 = (dROH2 - G_reoh) / G_reoh;
    const double x3 = costh - G_costhe;
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
    const double efac = exp(-G_b1 * (pow((dROH1 - G_reoh), 2) + pow((dROH2 - G_reoh), 2)));
    double sum0 = 0, sum1 = 0, sum2 = 0, sum3 = 0;
    for (int j = 1; j < 245; ++j) {
        const int inI = scp3_ps_idx1[j], inJ = scp3_ps_idx2[j], inK = scp3_ps_idx3[j];
        sum0 += scp3_ps_c5z[j] * (fmat[0][inI] * fmat[1][inJ] + fmat[0][inJ] * fmat[1][inI]) * fmat[2][inK];
        sum1 += scp3_ps_c5z[j] *
                ((inI - 1) * fmat[0][inI - 1] * fmat[1][inJ] + (inJ - 1) * fmat[0][inJ - 1] * fmat[1][inI]) *
                fmat[2][inK];
        sum2 += scp3_ps_c5z[j] *
                ((inJ - 1) * fmat[0][inI] * fmat[1][inJ - 1] + (inI - 1) * fmat[0][inJ] * fmat[1][inI - 1]) *
                fmat[2][inK];
        sum3 += scp3_ps_c5z[j] * (fmat[0][inI] * fmat[1][inJ] + fmat[0][inJ] * fmat[1][inI]) * (inK - 1) *
                fmat[2][inK - 1];
    }
    const double Vc = 2 * scp3_ps_c5z[0] + efac * sum0;
    double e1 = Va + Vb + Vc;
    e1 += 0.44739574026257;
    e1 *= G_cm2kcal;
    const double dVcdr1 = (-2 * G_b1 * efac * (dROH1 - G_reoh) * sum0 + efac * sum1 / G_reoh) / dROH1;
    const double dVcdr2 = (-2 * G_b1 * efac * (dROH2 - G_reoh) * sum0 + efac * sum2 / G_reoh) / dROH2;
    const double dVcdcth = efac * sum3;
    for (int i = 0; i < 3; ++i) {
        dr[3 + i] = dVa1 * ROH1[i] + dVb * RHH[i] + dVcdr1 * ROH1[i] +
                    dVcdcth * (ROH2[i] / (dROH1 * dROH2) - costh * ROH1[i] / (dROH1 * dROH1));
        dr[6 + i] = dVa2 * ROH2[i] - dVb * RHH[i] + dVcdr2 * ROH2[i] +
                    dVcdcth * (ROH1[i] / (dROH1 * dROH2) - costh * ROH2[i] / (dROH2 * dROH2));
        dr[0 + i] = -(dr[3 + i] + dr[6 + i]);
    }
    for (int i = 0; i < 9; ++i) dr[i] *= G_cm2kcal;
    return e1;
}

/* ps::dms_nasa(p1, p2, p3, rr, q3, dq3, true), ps.cpp:605-836; dq3[k + 3(j + 3i)]:
 * derivative along k of charge j (H1, H2, O) w.r.t. atom i (H1, H2, O) */
void orc_dms_nasa_grad(double p1_, double p2_, double p3_, const double *rr, double *q3, double *dq3)
{
    const double ath0 = 1.82400520401572996557;
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
    const double efac = exp(-G_b1D * (pow((dROH1 - G_reoh), 2) + pow((dROH2 - G_reoh), 2)));
    const double x1 = (dROH1 - G_reoh) / G_reoh;
    const double x2 = (dROH2 - G_reoh) / G_reoh;
    const double x3 = costh - G_costhe;
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
    const double pl1 = costh, pl2 = 0.5 * (3 * pl1 * pl1 - 1.0);
    double dp1dr1 = 0, dp1dr2 = 0, dp1dcabc = 0, dp2dr1 = 0, dp2dr2 = 0, dp2dcabc = 0;
    for (int j = 1; j < 84; ++j) {
        const int inI = scp3_ps_idxD0[j], inJ = scp3_ps_idxD1[j], inK = scp3_ps_idxD2[j];
        const double c = scp3_ps_coefD[j];
        p1 += c * fmat[0][inI] * fmat[1][inJ] * fmat[2][inK];
        p2 += c * fmat[0][inJ] * fmat[1][inI] * fmat[2][inK];
        dp1dr1 += c * (inI - 1) * fmat[0][inI - 1] * fmat[1][inJ] * fmat[2][inK];
        dp1dr2 += c * (inJ - 1) * fmat[0][inI] * fmat[1][inJ - 1] * fmat[2][inK];
        dp1dcabc += c * (inK - 1) * fmat[0][inI] * fmat[1][inJ] * fmat[2][inK - 1];
        dp2dr1 += c * (inJ - 1) * fmat[0][inJ - 1] * fmat[1][inI] * fmat[2][inK];
        dp2dr2 += c * (inI - 1) * fmat[0][inJ] * fmat[1][inI - 1] * fmat[2][inK];
        dp2dcabc += c * (inK - 1) * fmat[0][inJ] * fmat[1][inI] * fmat[2][inK - 1];
    }
    const double xx = 0.529177249, xx2 = xx * xx;
    dp1dr1 /= G_reoh / xx;
    dp1dr2 /= G_reoh / xx;
    dp2dr1 /= G_reoh / xx;
    dp2dr2 /= G_reoh / xx;
    const double pc0 = G_a * (pow(dROH1, G_b) + pow(dROH2, G_b)) * (G_c0 + pl1 * G_c1 + pl2 * G_c2);
    const double dpc0dr1 = G_a * G_b * pow(dROH1, G_b - 1) * (G_c0 + pl1 * G_c1 + pl2 * G_c2) * xx2;
    const double dpc0dr2 = G_a * G_b * pow(dROH2, G_b - 1) * (G_c0 + pl1 * G_c1 + pl2 * G_c2) * xx2;
    const double dpc0dcabc = G_a * (pow(dROH1, G_b) + pow(dROH2, G_b)) * (G_c1 + 0.5 * (6.0 * pl1) * G_c2) * xx;
    const double defacdr1 = -2.0 * G_b1D * (dROH1 - G_reoh) * efac * xx;
    const double defacdr2 = -2.0 * G_b1D * (dROH2 - G_reoh) * efac * xx;
    dp1dr1 = dp1dr1 * efac + p1 * defacdr1 + dpc0dr1;
    dp1dr2 = dp1dr2 * efac + p1 * defacdr2 + dpc0dr2;
    dp1dcabc = dp1dcabc * efac + dpc0dcabc;
    dp2dr1 = dp2dr1 * efac + p2 * defacdr1 + dpc0dr1;
    dp2dr2 = dp2dr2 * efac + p2 * defacdr2 + dpc0dr2;
    dp2dcabc = dp2dcabc * efac + dpc0dcabc;
    p1 = scp3_ps_coefD[0] + p1 * efac + pc0 * xx;
    p2 = scp3_ps_coefD[0] + p2 * efac + pc0 * xx;
    q3[0] = -(p1 + p2);
    q3[1] = p1;
    q3[2] = p2;
    dp1dr1 /= xx;
    dp1dr2 /= xx;
    dp2dr1 /= xx;
    dp2dr2 /= xx;
    /* TTM3 modification, ps.cpp:732-769 */
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
    dp1dr1 += p1_;
    dp2dr2 += p1_;
    dp1dcabc -= p3_ / sinth;
    dp2dcabc -= p3_ / sinth;
    /* ps.cpp:773-826 */
    const double f1q1r13 = (dp1dr1 - (dp1dcabc * costh / dROH1)) / dROH1;
    const double f1q1r23 = dp1dcabc / (dROH1 * dROH2);
    const double f2q1r23 = (dp1dr2 - (dp1dcabc * costh / dROH2)) / dROH2;
    const double f2q1r13 = dp1dcabc / (dROH2 * dROH1);
    const double f1q2r13 = (dp2dr1 - (dp2dcabc * costh / dROH1)) / dROH1;
    const double f1q2r23 = dp2dcabc / (dROH1 * dROH2);
    const double f2q2r23 = (dp2dr2 - (dp2dcabc * costh / dROH2)) / dROH2;
    const double f2q2r13 = dp2dcabc / (dROH2 * dROH1);
#define GQ(i, j, k) dq3[(k) + 3 * ((j) + 3 * (i))]
    for (int k = 0; k < 3; ++k) {
        GQ(0, 0, k) = f1q1r13 * ROH1[k] + f1q1r23 * ROH2[k];
        GQ(1, 0, k) = f2q1r13 * ROH1[k] + f2q1r23 * ROH2[k];
        GQ(2, 0, k) = -(GQ(0, 0, k) + GQ(1, 0, k));
        GQ(0, 1, k) = f1q2r13 * ROH1[k] + f1q2r23 * ROH2[k];
        GQ(1, 1, k) = f2q2r13 * ROH1[k] + f2q2r23 * ROH2[k];
        GQ(2, 1, k) = -(GQ(0, 1, k) + GQ(1, 1, k));
        GQ(0, 2, k) = -(GQ(0, 0, k) + GQ(0, 1, k));
        GQ(1, 2, k) = -(GQ(1, 0, k) + GQ(1, 1, k));
        GQ(2, 2, k) = -(GQ(2, 0, k) + GQ(2, 1, k));
    }
#undef GQ
}

/* ttm3f.cpp:16-33, ttm3f-bits.h:10-26 */
static const double TG_polarM = 1.444, TG_polfacH = 0.496, TG_polfacM = 0.837;
static const double TG_vdwC = -0.72298855E+03, TG_vdwD = 0.10211829E+06, TG_vdwE = 0.37170376E+01;
static const double TG_dms1 = 0.5, TG_dms2 = 0.9578, TG_dms3 = 0.012;
static const double TG_gammaM = 0.46;
#define TG_gamma1 (1.0 - TG_gammaM)
#define TG_gamma2 (TG_gammaM / 2)
static const double TG_CHARGECON = 18.22234397655801030455;
static const double TG_DEBYE = 4.8033324;
static const double TG_aXX = 0.175;

/* smeared kernels, ttm3f-bits-smear.cpp:116-190; want: 1 = 01, 2 = 012, 12, 123 */
static void tg_smear(double r12, double AA, double *ts0, double *ts1, double *ts2, double *ts3)
{
    const double dri = 1.0 / r12;
    const double drsqi = dri * dri;
    const double rA = r12 / AA;
    const double rA3 = rA * rA * rA;
    const double exp1 = exp(-TG_aXX * rA3);
    const double AA3 = AA * AA * AA;
    if (ts0) *ts0 = (1.0 - exp1 + pow(TG_aXX, 1.0 / 3.0) * rA * orc_gammq23(TG_aXX * rA3)) * dri;
    *ts1 = (1.0 - exp1) * dri * drsqi;
    if (ts2) *ts2 = (*ts1 - exp1 * TG_aXX / AA3) * drsqi;
    if (ts3) *ts3 = (*ts2 - 0.6 * exp1 * r12 * TG_aXX * TG_aXX / (AA3 * AA3)) * drsqi;
}

/* ttm3f::operator()(nw, crd, grad), ttm3f.cpp:325-716; crd in A, O H H */
double orc_ttm3f_grad(int64_t nw, const double *crd, double *grad, double tol, int *iters)
{
    const double AA_HH = pow(TG_polfacH * TG_polfacH, 1.0 / 6.0);
    const double AA_HM = pow(TG_polfacH * TG_polfacM, 1.0 / 6.0);
    const double AA_MM = pow(TG_polfacM * TG_polfacM, 1.0 / 6.0);
    const int64_t n3w = 3 * nw;
    double *mem = (double *)calloc((size_t)(51 * nw + 9 * nw * nw), sizeof(double));
    double *msite = mem, *charge = msite + n3w, *dip = charge + n3w, *dip_prev = dip + n3w;
    double *Efd = dip_prev + n3w, *Efq = Efd + n3w, *phi = Efq + n3w, *dM = phi + n3w;
    double *ddt = dM + n3w, *grdq = ddt + 9 * nw * nw;
    double Eint = 0;
    for (int64_t n = 0; n < nw; ++n) { /* :347-393 */
        const int64_t n3 = 3 * n, n9 = 9 * n;
        for (int k = 0; k < 3; ++k)
            msite[n3 + k] = TG_gamma1 * crd[n9 + k] + TG_gamma2 * (crd[n9 + 3 + k] + crd[n9 + 6 + k]);
        Eint += orc_pot_nasa_grad(crd + n9, grad + n9);
        double q3[3], dq3[27];
        orc_dms_nasa_grad(TG_dms1, TG_dms2, TG_dms3, crd + n9, q3, dq3);
        const double tmp = TG_gamma2 / TG_gamma1;
        charge[n3 + 0] = q3[0] / TG_gamma1;
        charge[n3 + 1] = q3[1] + tmp * (q3[1] + q3[2]);
        charge[n3 + 2] = q3[2] + tmp * (q3[1] + q3[2]);
#define DQ3(i, j, k) dq3[(k) + 3 * ((j) + 3 * (i))]
#define GRDQ(i, j, k) grdq[(k) + 3 * ((j) + 3 * ((i) + 3 * n))]
        for (int k = 0; k < 3; ++k) {
            GRDQ(0, 0, k) = DQ3(0, 0, k) + tmp * (DQ3(0, 0, k) + DQ3(0, 1, k));
            GRDQ(1, 0, k) = DQ3(1, 0, k) + tmp * (DQ3(1, 0, k) + DQ3(1, 1, k));
            GRDQ(2, 0, k) = DQ3(2, 0, k) + tmp * (DQ3(2, 0, k) + DQ3(2, 1, k));
            GRDQ(0, 1, k) = DQ3(0, 1, k) + tmp * (DQ3(0, 1, k) + DQ3(0, 0, k));
            GRDQ(1, 1, k) = DQ3(1, 1, k) + tmp * (DQ3(1, 1, k) + DQ3(1, 0, k));
            GRDQ(2, 1, k) = DQ3(2, 1, k) + tmp * (DQ3(2, 1, k) + DQ3(2, 0, k));
            GRDQ(0, 2, k) = DQ3(0, 2, k) - 2 * tmp * (DQ3(0, 0, k) + DQ3(0, 1, k));
            GRDQ(1, 2, k) = DQ3(1, 2, k) - 2 * tmp * (DQ3(1, 0, k) + DQ3(1, 1, k));
            GRDQ(2, 2, k) = DQ3(2, 2, k) - 2 * tmp * (DQ3(2, 0, k) + DQ3(2, 1, k));
        }
#undef DQ3
        for (int k = 0; k < 3; ++k) charge[n3 + k] *= TG_CHARGECON;
        for (int k = 0; k < 27; ++k) grdq[27 * n + k] *= TG_CHARGECON;
    }
    double Evdw = 0;
    for (int64_t i = 0; i < nw; ++i) { /* :397-537 */
        const int64_t i3 = 3 * i, i9 = 9 * i;
        for (int64_t j = 0; j < nw; ++j) {
            if (i == j) continue;
            const int64_t j3 = 3 * j;
            for (int l = 1; l < 3; ++l) {
                const int64_t j3l = j3 + l, jh = 3 * j3l;
                double Rij[3], Rsq = 0;
                for (int k = 0; k < 3; ++k) {
                    Rij[k] = msite[i3 + k] - crd[jh + k];
                    Rsq += Rij[k] * Rij[k];
                }
                double ts0, ts1;
                tg_smear(sqrt(Rsq), AA_HM, &ts0, &ts1, 0, 0);
                phi[i3] += ts0 * charge[j3l];
                phi[j3l] += ts0 * charge[i3];
                const double tmp1 = charge[j3l] * ts1, tmp2 = charge[i3] * tmp1;
                for (int k = 0; k < 3; ++k) {
                    Efq[i3 + k] += tmp1 * Rij[k];
                    grad[jh + k] += tmp2 * Rij[k];
                }
            }
        }
        for (int64_t j = i + 1; j < nw; ++j) {
            const int64_t j3 = 3 * j, j9 = 9 * j;
            for (int a = 1; a < 3; ++a) {
                const int64_t i3a = i3 + a;
                for (int b = 1; b < 3; ++b) {
                    const int64_t j3b = j3 + b;
                    double Rij[3], Rsq = 0;
                    for (int k = 0; k < 3; ++k) {
                        Rij[k] = crd[3 * i3a + k] - crd[3 * j3b + k];
                        Rsq += Rij[k] * Rij[k];
                    }
                    double ts0, ts1;
                    tg_smear(sqrt(Rsq), AA_HH, &ts0, &ts1, 0, 0);
                    phi[i3a] += charge[j3b] * ts0;
                    phi[j3b] += charge[i3a] * ts0;
                    const double qq = charge[i3a] * charge[j3b] * ts1;
                    for (int k = 0; k < 3; ++k) {
                        grad[3 * i3a + k] -= qq * Rij[k];
                        grad[3 * j3b + k] += qq * Rij[k];
                    }
                }
            }
            double Rij[3], Rsq = 0;
            for (int k = 0; k < 3; ++k) {
                Rij[k] = crd[i9 + k] - crd[j9 + k];
                Rsq += Rij[k] * Rij[k];
            }
            const double dRij = sqrt(Rsq);
            const double dR6 = Rsq * Rsq * Rsq;
            const double expon = TG_vdwD * exp(-TG_vdwE * dRij);
            Evdw += TG_vdwC / dR6 + expon;
            const double tmp = -(6.0 * TG_vdwC / dR6) / Rsq - TG_vdwE * expon / dRij;
            for (int k = 0; k < 3; ++k) {
                grad[i9 + k] += tmp * Rij[k];
                grad[j9 + k] -= tmp * Rij[k];
            }
            Rsq = 0.0;
            for (int k = 0; k < 3; ++k) {
                Rij[k] = msite[i3 + k] - msite[j3 + k];
                Rsq += Rij[k] * Rij[k];
            }
            double ts0, ts1, ts2;
            tg_smear(sqrt(Rsq), AA_MM, &ts0, &ts1, &ts2, 0);
            phi[i3] += ts0 * charge[j3];
            phi[j3] += ts0 * charge[i3];
            for (int k = 0; k < 3; ++k) {
                Efq[i3 + k] += charge[j3] * ts1 * Rij[k];
                Efq[j3 + k] -= charge[i3] * ts1 * Rij[k];
            }
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    const double d = 3.0 * ts2 * Rij[k] * Rij[l] - (k == l ? ts1 : 0.0);
                    ddt[n3w * (i3 + k) + j3 + l] = d;
                    ddt[n3w * (j3 + k) + i3 + l] = d;
                }
        }
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) ddt[n3w * (i3 + k) + i3 + l] = 0.0;
    }
    for (int64_t i = 0; i < n3w; ++i) { /* :541-576 */
        dip[i] = TG_polarM * Efq[i];
        dip_prev[i] = dip[i];
    }
    const double dmix = 0.7;
    const double stath = TG_DEBYE / TG_CHARGECON / sqrt((double)nw);
    int it = 0;
    for (; it < 500; ++it) {
        for (int64_t k = 0; k < n3w; ++k)
            for (int64_t l = 0; l < n3w; ++l) Efd[k] += ddt[n3w * k + l] * dip[l];
        double deltadip = 0.0;
        for (int64_t i = 0; i < n3w; ++i) {
            dip[i] = TG_polarM * (Efq[i] + Efd[i]);
            dip[i] = dmix * dip[i] + (1.0 - dmix) * dip_prev[i];
            const double delta = dip[i] - dip_prev[i];
            deltadip += delta * delta;
        }
        deltadip = sqrt(deltadip) * stath;
        if (deltadip < tol) break;
        memcpy(dip_prev, dip, sizeof(double) * n3w);
        memset(Efd, 0, sizeof(double) * n3w);
    }
    if (iters) *iters = it + 1;
    double Eelec = 0, Eind = 0;
    for (int64_t i = 0; i < n3w; ++i) {
        Eelec += phi[i] * charge[i];
        Eind -= dip[i] * Efq[i];
    }
    Eelec /= 2;
    Eind /= 2;
    for (int64_t i = 0; i < nw; ++i) { /* charge-dipole and dipole-dipole, :596-668 */
        const int64_t i3 = 3 * i;
        for (int64_t j = 0; j < nw; ++j) {
            if (i == j) continue;
            const int64_t j3 = 3 * j;
            for (int l = 1; l < 3; ++l) {
                const int64_t j3l = j3 + l, jh = 3 * j3l;
                double Rij[3], Rsq = 0, diR = 0;
                for (int k = 0; k < 3; ++k) {
                    Rij[k] = msite[i3 + k] - crd[jh + k];
                    Rsq += Rij[k] * Rij[k];
                    diR += dip[i3 + k] * Rij[k];
                }
                double ts1, ts2;
                tg_smear(sqrt(Rsq), AA_HM, 0, &ts1, &ts2, 0);
                for (int k = 0; k < 3; ++k) {
                    const double derij = charge[j3l] * (3 * ts2 * diR * Rij[k] - ts1 * dip[i3 + k]);
                    dM[i3 + k] += derij;
                    grad[jh + k] -= derij;
                }
                phi[j3l] -= ts1 * diR;
            }
            double Rij[3], Rsq = 0, diR = 0;
            for (int k = 0; k < 3; ++k) {
                Rij[k] = msite[i3 + k] - msite[j3 + k];
                Rsq += Rij[k] * Rij[k];
                diR += dip[i3 + k] * Rij[k];
            }
            double ts1, ts2, ts3;
            tg_smear(sqrt(Rsq), AA_MM, 0, &ts1, &ts2, &ts3);
            for (int k = 0; k < 3; ++k) {
                const double derij = charge[j3] * (3 * ts2 * diR * Rij[k] - ts1 * dip[i3 + k]);
                dM[i3 + k] += derij;
                dM[j3 + k] -= derij;
            }
            phi[j3] -= ts1 * diR;
            if (j > i) {
                double djR = 0, didj = 0;
                for (int k = 0; k < 3; ++k) {
                    djR += dip[j3 + k] * Rij[k];
                    didj += dip[i3 + k] * dip[j3 + k];
                }
                for (int k = 0; k < 3; ++k) {
                    const double derij = -3 * ts2 * (didj * Rij[k] + djR * dip[i3 + k] + diR * dip[j3 + k]) +
                                         15 * ts3 * diR * djR * Rij[k];
                    dM[i3 + k] += derij;
                    dM[j3 + k] -= derij;
                }
            }
        }
    }
    for (int64_t i = 0; i < nw; ++i) { /* :672-686 */
        const int64_t i3 = 3 * i, i9 = 9 * i;
        for (int k = 0; k < 3; ++k) {
            const double tmp = dM[i3 + k] - charge[i3] * Efq[i3 + k];
            grad[i9 + 0 + k] += TG_gamma1 * tmp;
            grad[i9 + 3 + k] += TG_gamma2 * tmp;
            grad[i9 + 6 + k] += TG_gamma2 * tmp;
        }
    }
    for (int64_t n = 0; n < nw; ++n) { /* :690-713 */
        const int64_t n3 = 3 * n, io = 9 * n, ih1 = 9 * n + 3, ih2 = 9 * n + 6;
        for (int k = 0; k < 3; ++k) {
            grad[ih1 + k] += GRDQ(0, 0, k) * phi[n3 + 1] + GRDQ(0, 1, k) * phi[n3 + 2] + GRDQ(0, 2, k) * phi[n3 + 0];
            grad[ih2 + k] += GRDQ(1, 0, k) * phi[n3 + 1] + GRDQ(1, 1, k) * phi[n3 + 2] + GRDQ(1, 2, k) * phi[n3 + 0];
            grad[io + k] += GRDQ(2, 0, k) * phi[n3 + 1] + GRDQ(2, 1, k) * phi[n3 + 2] + GRDQ(2, 2, k) * phi[n3 + 0];
        }
    }
#undef GRDQ
    free(mem);
    return Eint + Evdw + Eelec + Eind;
}
