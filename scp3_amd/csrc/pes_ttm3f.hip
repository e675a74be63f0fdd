// TTM3-F energies of a batch of configurations, one configuration per wavefront.
//
// Reference: h2o::ttm3f::operator()(nw, crd) (ttm3f.cpp:92-321) with the
// Partridge-Schwenke monomer potential ps::pot_nasa (ps.cpp:486-575), its
// dipole-moment surface ps::dms_nasa with the TTM3 modification
// (ps.cpp:605-759), the smeared Coulomb kernels smear0/01/012 and gammq23
// (ttm3f-bits-smear.cpp:17-171) and the constants of ttm3f.cpp:16-33 and
// ttm3f-bits.h:10-26.
//
// A workgroup is one wave and one configuration.  Its LDS holds the
// configuration: coordinates, M sites and charges, the per-pair partial
// potentials and fields, and the O(nw^2) M-M pair data (R, ts1, ts2) from
// which the dipole tensor T_ij = 3 ts2 R R^T - ts1 I is applied.
//   phase 1  lane w < nw: M site, PS energy, DMS charges of molecule w
//   phase 2  lane p < nw(nw-1)/2 (strided): every site pair of molecule pair
//            p - smeared M-H, H-H, M-M Coulomb, O-O dispersion/repulsion -
//            written to fixed per-pair slots (no atomics)
//   phase 3  lane s < 3nw: potential at site s and field component s summed
//            over the partner molecules in a fixed order (deterministic)
//   phase 4  lane r < 3nw keeps row r of the induced dipoles in registers and
//            iterates mu <- 0.7 polarM (Efq + T mu) + 0.3 mu_prev in FP64,
//            sharing mu through LDS, until the RMS change (Debye, scaled as in
//            ttm3f.cpp:285,296-298) is below tol; the butterfly sum of the
//            change is broadcast from lane 0 so every lane takes the same exit.
// T is the arithmetic of phases 1-3 (double in the FP64 form, float in the
// mixed form); the dipole iteration and all energy sums are FP64.
//
// The reference does not clear its Efd buffer before the first iteration
// (ttm3f.cpp:112 zeroes Efq, phi and part of ddt only); the fixed point does
// not depend on the starting field, and here the iteration starts from the
// zero field.
#include "scp3g_internal.h"

#include <cfloat>
#include <cmath>
#include <vector>

#define SCP3_PS_STORAGE constexpr
namespace scp3g {
namespace psdata {
#include "ps_coeffs.inc"
}  // namespace psdata

namespace ttm {
// ps.cpp:11-18,472-477
constexpr double reoh = 0.958649, b1 = 2.0, roh = 0.9519607159623009;
constexpr double alphaoh = 2.587949757553683, deohA = 42290.92019288289;
constexpr double phh1A = 16.94879431193463, phh2 = 12.66426998162947;
constexpr double f5z = 0.999677885;
constexpr double b1D = 1.0, pa = 0.2999e0, pb = -0.6932e0;
constexpr double pc0 = 1.0099e0, pc1 = -0.1801e0, pc2 = 0.0892e0;
constexpr double costhe = -.24780227221366464506;
constexpr double ath0 = 1.82400520401572996557;
constexpr double cm2kcal = 0.00285914375100642899;
// ttm3f.cpp:20-32, ttm3f-bits.h:10-26
constexpr double polarM = 1.444, polfacH = 0.496, polfacM = 0.837;
constexpr double vdwC = -0.72298855E+03, vdwD = 0.10211829E+06, vdwE = 0.37170376E+01;
constexpr double dms1 = 0.5, dms2 = 0.9578, dms3 = 0.012;
constexpr double gammaM = 0.46, gamma1 = 1.0 - gammaM, gamma2 = gammaM / 2;
constexpr double CHARGECON = 18.22234397655801030455;
constexpr double DEBYE = 4.8033324;
constexpr double dmix = 0.7;
constexpr double aXX = 0.175;
constexpr int maxiter = 500;
}  // namespace ttm

template <typename T>
struct Lim;
template <>
struct Lim<double> {
    static constexpr double eps = DBL_EPSILON;
    static constexpr double fpmin = DBL_MIN / DBL_EPSILON;
};
template <>
struct Lim<float> {
    static constexpr float eps = FLT_EPSILON;
    static constexpr float fpmin = FLT_MIN / FLT_EPSILON;
};

// Gamma(2/3) Q(2/3, x): series below x = 5/3, Lentz continued fraction above
// (ttm3f-bits-smear.cpp:17-80); loops bounded for the device.
template <typename T>
__host__ __device__ T gammq23(T x)
{
    const T a = T(2.0 / 3.0);
    const T eps = Lim<T>::eps, fpmin = Lim<T>::fpmin;
    if (x < T(5.0 / 3.0)) {
        T ap = a, sum = T(1) / a, del = sum;
        for (int it = 0; it < 1000; ++it) {
            ap += T(1);
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * eps) break;
        }
        return T(1.354117939426400e+00) - sum * exp(-x + a * log(x));
    }
    T b = x + T(1) - a, c = T(1) / fpmin, d = T(1) / b, h = d;
    for (int i = 1; i < 1000; ++i) {
        const T an = -T(i) * (T(i) - a);
        b += T(2);
        d = an * d + b;
        if (fabs(d) < fpmin) d = fpmin;
        c = b + an / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = T(1) / d;
        const T del = d * c;
        h *= del;
        if (fabs(del - T(1)) <= eps) break;
    }
    return exp(-x + a * log(x)) * h;
}

// smear0 / smear01 / smear012 (ttm3f-bits-smear.cpp:92-171); want = 0, 1, 2
template <typename T>
__host__ __device__ void smear(T r12, T AA, T a13, int want, T &ts0, T &ts1, T &ts2)
{
    const T dri = T(1) / r12;
    const T drsqi = dri * dri;
    const T rA = r12 / AA;
    const T rA3 = rA * rA * rA;
    const T exp1 = exp(-T(ttm::aXX) * rA3);
    ts0 = (T(1) - exp1 + a13 * rA * gammq23<T>(T(ttm::aXX) * rA3)) * dri;
    if (want >= 1) ts1 = (T(1) - exp1) * dri * drsqi;
    if (want >= 2) ts2 = (ts1 - exp1 * T(ttm::aXX) / (AA * AA * AA)) * drsqi;
}

// PS monomer energy, kcal/mol (ps.cpp:486-575).  The Morse part is written
// as deoh((e-1)^2 - 1) and the constants -2 deoh (both OH), 2 c5z[0] and the
// 0.447 cm^-1 correction are folded in FP64, so that the T-precision part
// is the small, configuration-dependent remainder.
template <typename T>
__host__ __device__ double ps_energy(const T *O, const T *H1, const T *H2)
{
    using namespace ttm;
    T ROH1[3], ROH2[3], RHH[3], d1 = 0, d2 = 0, dh = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = H1[i] - O[i];
        ROH2[i] = H2[i] - O[i];
        RHH[i] = H1[i] - H2[i];
        d1 += ROH1[i] * ROH1[i];
        d2 += ROH2[i] * ROH2[i];
        dh += RHH[i] * RHH[i];
    }
    d1 = sqrt(d1);
    d2 = sqrt(d2);
    dh = sqrt(dh);
    const T costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (d1 * d2);
    const double deoh = f5z * deohA;
    const T e1 = exp(-T(alphaoh) * (d1 - T(roh)));
    const T e2 = exp(-T(alphaoh) * (d2 - T(roh)));
    const T Va_shift = T(deoh) * ((e1 - T(1)) * (e1 - T(1)) + (e2 - T(1)) * (e2 - T(1)));
    // Vb = f5z*phh1A*exp(phh2)*exp(-phh2*dRHH) = f5z*phh1A*exp(phh2*(1-dRHH))
    const T Vb = T(f5z * phh1A) * exp(T(phh2) * (T(1) - dh));
    const T x1 = (d1 - T(reoh)) / T(reoh);
    const T x2 = (d2 - T(reoh)) / T(reoh);
    const T x3 = costh - T(costhe);
    T f0[16], f1[16], f2[16];
    f0[0] = f1[0] = f2[0] = T(0);
    f0[1] = f1[1] = f2[1] = T(1);
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        f0[j] = f0[j - 1] * x1;
        f1[j] = f1[j - 1] * x2;
        f2[j] = f2[j - 1] * x3;
    }
    const T r1 = d1 - T(reoh), r2 = d2 - T(reoh);
    const T efac = exp(-T(b1) * (r1 * r1 + r2 * r2));
    T sum0 = 0;
#pragma unroll
    for (int j = 1; j < 245; ++j) {
        const int inI = psdata::scp3_ps_idx1[j], inJ = psdata::scp3_ps_idx2[j],
                  inK = psdata::scp3_ps_idx3[j];
        sum0 += T(psdata::scp3_ps_c5z[j]) * (f0[inI] * f1[inJ] + f0[inJ] * f1[inI]) * f2[inK];
    }
    const double Kconst = 2.0 * psdata::scp3_ps_c5z[0] - 2.0 * deoh + 0.44739574026257;
    return ((double)(Va_shift + Vb + efac * sum0) + Kconst) * cm2kcal;
}

// DMS charges with the TTM3 modification (ps.cpp:605-759), charge[3] of
// M, H1, H2 in the scaled units of ttm3f.cpp:126-138.
template <typename T>
__host__ __device__ void dms_charges(const T *O, const T *H1, const T *H2, T *chg)
{
    using namespace ttm;
    T ROH1[3], ROH2[3], d1 = 0, d2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = H1[i] - O[i];
        ROH2[i] = H2[i] - O[i];
        d1 += ROH1[i] * ROH1[i];
        d2 += ROH2[i] * ROH2[i];
    }
    d1 = sqrt(d1);
    d2 = sqrt(d2);
    const T costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (d1 * d2);
    const T r1 = d1 - T(reoh), r2 = d2 - T(reoh);
    const T efac = exp(-T(b1D) * (r1 * r1 + r2 * r2));
    const T x1 = r1 / T(reoh), x2 = r2 / T(reoh), x3 = costh - T(costhe);
    T f0[16], f1[16], f2[16];
    f0[0] = f1[0] = f2[0] = T(0);
    f0[1] = f1[1] = f2[1] = T(1);
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        f0[j] = f0[j - 1] * x1;
        f1[j] = f1[j - 1] * x2;
        f2[j] = f2[j - 1] * x3;
    }
    T p1 = 0, p2 = 0;
#pragma unroll
    for (int j = 1; j < 84; ++j) {
        const int inI = psdata::scp3_ps_idxD0[j], inJ = psdata::scp3_ps_idxD1[j],
                  inK = psdata::scp3_ps_idxD2[j];
        p1 += T(psdata::scp3_ps_coefD[j]) * f0[inI] * f1[inJ] * f2[inK];
        p2 += T(psdata::scp3_ps_coefD[j]) * f0[inJ] * f1[inI] * f2[inK];
    }
    const T xx = T(0.529177249);
    const T pl1 = costh, pl2 = T(0.5) * (T(3) * pl1 * pl1 - T(1));
    const T pc = T(pa) * (pow(d1, T(pb)) + pow(d2, T(pb))) * (T(pc0) + pl1 * T(pc1) + pl2 * T(pc2));
    p1 = T(psdata::scp3_ps_coefD[0]) + p1 * efac + pc * xx;
    p2 = T(psdata::scp3_ps_coefD[0]) + p2 * efac + pc * xx;
    T q0 = -(p1 + p2), q1 = p1, q2 = p2;
    // TTM3 modification (ps.cpp:732-759)
    const T ax = ROH1[1] * ROH2[2] - ROH1[2] * ROH2[1];
    const T ay = -ROH1[0] * ROH2[2] + ROH1[2] * ROH2[0];
    const T az = ROH1[0] * ROH2[1] - ROH1[1] * ROH2[0];
    const T sinth = sqrt(ax * ax + ay * ay + az * az) / (d1 * d2);
    const T ang = atan2(sinth, costh);
    const T m1 = T(dms1) * (d1 - T(dms2)) + T(dms3) * (ang - T(ath0));
    const T m2 = T(dms1) * (d2 - T(dms2)) + T(dms3) * (ang - T(ath0));
    q0 = q0 - (m1 + m2);
    q1 = q1 + m1;
    q2 = q2 + m2;
    // ttm3f.cpp:126-138
    const T tmp = T(gamma2 / gamma1);
    chg[0] = T(CHARGECON) * q0 / T(gamma1);
    chg[1] = T(CHARGECON) * (q1 + tmp * (q1 + q2));
    chg[2] = T(CHARGECON) * (q2 + tmp * (q1 + q2));
}

struct SmearConst {
    double AA_HH, AA_HM, AA_MM, a13;
};

static SmearConst smear_const()
{
    SmearConst s;
    s.AA_HH = std::pow(ttm::polfacH * ttm::polfacH, 1.0 / 6.0);
    s.AA_HM = std::pow(ttm::polfacH * ttm::polfacM, 1.0 / 6.0);
    s.AA_MM = std::pow(ttm::polfacM * ttm::polfacM, 1.0 / 6.0);
    s.a13 = std::pow(ttm::aXX, 1.0 / 3.0);
    return s;
}

// All site pairs of molecules i < j; x holds O H H coordinates (9 per
// molecule), ms the M sites, q the charges (M H1 H2 per molecule).  slot[12]:
// potential at M_i H_i1 H_i2, at M_j H_j1 H_j2, field at M_i, field at M_j;
// tens[5]: R = M_i - M_j, ts1, ts2.  Returns the O-O dispersion/repulsion.
template <typename T>
__host__ __device__ double pair_terms(const T *x, const T *ms, const T *q, int i, int j,
                                      const SmearConst &sc, T *slot, T *tens)
{
    using namespace ttm;
    const T AA_HM = T(sc.AA_HM), AA_HH = T(sc.AA_HH), AA_MM = T(sc.AA_MM), a13 = T(sc.a13);
    T ts0, ts1, ts2;
#pragma unroll
    for (int s = 0; s < 12; ++s) slot[s] = T(0);
    // M(a) - H(b) both ways (ttm3f.cpp:148-172)
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const int a = dir ? j : i, b = dir ? i : j;
        T *phiM = slot + (dir ? 3 : 0);   // potential at M_a
        T *phiH = slot + (dir ? 0 : 3);   // potentials of molecule b's sites
        T *E = slot + (dir ? 9 : 6);      // field at M_a
#pragma unroll
        for (int l = 1; l < 3; ++l) {
            T R[3], rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R[k] = ms[3 * a + k] - x[9 * b + 3 * l + k];
                rsq += R[k] * R[k];
            }
            smear<T>(sqrt(rsq), AA_HM, a13, 1, ts0, ts1, ts2);
            phiM[0] += ts0 * q[3 * b + l];
            phiH[l] += ts0 * q[3 * a];
#pragma unroll
            for (int k = 0; k < 3; ++k) E[k] += q[3 * b + l] * ts1 * R[k];
        }
    }
    // H - H (ttm3f.cpp:180-199)
#pragma unroll
    for (int ha = 1; ha < 3; ++ha)
#pragma unroll
        for (int hb = 1; hb < 3; ++hb) {
            T rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T d = x[9 * i + 3 * ha + k] - x[9 * j + 3 * hb + k];
                rsq += d * d;
            }
            smear<T>(sqrt(rsq), AA_HH, a13, 0, ts0, ts1, ts2);
            slot[ha] += q[3 * j + hb] * ts0;
            slot[3 + hb] += q[3 * i + ha] * ts0;
        }
    // O - O (ttm3f.cpp:201-217)
    T rsq = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T d = x[9 * i + k] - x[9 * j + k];
        rsq += d * d;
    }
    const double r = sqrt((double)rsq);
    const double evdw = vdwC / ((double)rsq * rsq * rsq) + vdwD * exp(-vdwE * r);
    // M - M (ttm3f.cpp:219-255)
    T R[3];
    rsq = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R[k] = ms[3 * i + k] - ms[3 * j + k];
        rsq += R[k] * R[k];
    }
    smear<T>(sqrt(rsq), AA_MM, a13, 2, ts0, ts1, ts2);
    slot[0] += ts0 * q[3 * j];
    slot[3] += ts0 * q[3 * i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        slot[6 + k] += q[3 * j] * ts1 * R[k];
        slot[9 + k] -= q[3 * i] * ts1 * R[k];
        tens[k] = R[k];
    }
    tens[3] = ts1;
    tens[4] = ts2;
    return evdw;
}

__host__ __device__ inline int pair_index(int i, int j, int nw)  // i < j
{
    return i * nw - i * (i + 1) / 2 + (j - i - 1);
}

template <typename T>
__host__ __device__ void monomer_terms(const T *x, int w, T *ms, T *q, double &eint)
{
    using namespace ttm;
    const T *O = x + 9 * w, *H1 = O + 3, *H2 = O + 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) ms[3 * w + k] = T(gamma1) * O[k] + T(gamma2) * (H1[k] + H2[k]);
    eint = ps_energy<T>(O, H1, H2);
    dms_charges<T>(O, H1, H2, q + 3 * w);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return __shfl(v, 0, 64);  // one value for every lane
}

template <typename T>
__global__ void __launch_bounds__(64)
    k_ttm3f(const T *__restrict__ crd, int64_t ldc, int nw, int64_t npts, double tol,
            SmearConst sc, double *__restrict__ E)
{
    using namespace ttm;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int64_t cfg = blockIdx.x;
    const int n3 = 3 * nw, P = nw * (nw - 1) / 2;
    double *dipL = smem;                          // [3nw]
    T *x = reinterpret_cast<T *>(smem + 64);      // [9nw]
    T *ms = x + 9 * nw;                           // [3nw]
    T *q = ms + n3;                               // [3nw]
    T *slots = q + n3;                            // [P][12]
    T *tens = slots + 12 * P;                     // [P][5]

    for (int c = lane; c < 9 * nw; c += 64) x[c] = crd[(size_t)c * ldc + cfg];
    __syncthreads();
    double eint = 0.0;
    if (lane < nw) monomer_terms<T>(x, lane, ms, q, eint);
    __syncthreads();
    double evdw = 0.0;
    for (int p = lane; p < P; p += 64) {
        int i = 0, rem = p;
        while (rem >= nw - 1 - i) {
            rem -= nw - 1 - i;
            ++i;
        }
        const int j = i + 1 + rem;
        evdw += pair_terms<T>(x, ms, q, i, j, sc, slots + 12 * p, tens + 5 * p);
    }
    __syncthreads();
    // potentials and charge fields per site / component (fixed partner order)
    double phi = 0.0, efq = 0.0;
    if (lane < n3) {
        const int i = lane / 3, t = lane - 3 * (lane / 3);
        for (int j = 0; j < nw; ++j) {
            if (j == i) continue;
            const bool first = i < j;
            const T *sl = slots + 12 * pair_index(first ? i : j, first ? j : i, nw);
            phi += (double)sl[(first ? 0 : 3) + t];
            efq += (double)sl[(first ? 6 : 9) + t];
        }
    }
    const double eelec = 0.5 * wave_sum(lane < n3 ? phi * (double)q[lane] : 0.0);
    // induced dipoles (ttm3f.cpp:268-307), row r = lane
    const int i = lane / 3, kk = lane - 3 * (lane / 3);
    double dip = lane < n3 ? polarM * efq : 0.0;
    double dip_prev = dip;
    const double stath = DEBYE / CHARGECON / sqrt((double)nw);
    for (int it = 0; it < maxiter; ++it) {
        if (lane < n3) dipL[lane] = dip;
        __syncthreads();
        double efd = 0.0;
        if (lane < n3) {
            for (int j = 0; j < nw; ++j) {
                if (j == i) continue;
                const T *tp = tens + 5 * pair_index(i < j ? i : j, i < j ? j : i, nw);
                const double R0 = tp[0], R1 = tp[1], R2 = tp[2], t1 = tp[3], t2 = tp[4];
                const double m0 = dipL[3 * j], m1 = dipL[3 * j + 1], m2 = dipL[3 * j + 2];
                const double Rk = kk == 0 ? R0 : (kk == 1 ? R1 : R2);
                const double mk = kk == 0 ? m0 : (kk == 1 ? m1 : m2);
                efd += 3.0 * t2 * Rk * (R0 * m0 + R1 * m1 + R2 * m2) - t1 * mk;
            }
        }
        double dn = lane < n3 ? polarM * (efq + efd) : 0.0;
        dn = dmix * dn + (1.0 - dmix) * dip_prev;
        const double delta = lane < n3 ? dn - dip_prev : 0.0;
        const double dd = sqrt(wave_sum(delta * delta)) * stath;
        dip = dn;
        __syncthreads();  // every lane has read dipL before it is rewritten
        if (dd < tol) break;
        dip_prev = dip;
    }
    const double eind = -0.5 * wave_sum(lane < n3 ? dip * efq : 0.0);
    const double ei = wave_sum(eint);
    const double ev = wave_sum(evdw);
    if (lane == 0) E[cfg] = ei + ev + eelec + eind;
}

size_t ttm3f_smem_bytes(int nw, size_t es)
{
    const int P = nw * (nw - 1) / 2;
    return 64 * sizeof(double) + es * (size_t)(9 * nw + 6 * nw + 17 * P);
}

hipError_t launch_ttm3f(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                        int fp32, double *E)
{
    if (npts <= 0) return hipSuccess;
    if (nw < 1 || 3 * nw > 64) return hipErrorInvalidValue;  // one dipole row per lane
    const size_t smem = ttm3f_smem_bytes(nw, fp32 ? sizeof(float) : sizeof(double));
    if (smem > 64 * 1024) return hipErrorInvalidValue;
    const SmearConst sc = smear_const();
    // FP64 form: the reference's tolerance (ttm3f.cpp:32); mixed form 1e-8 D
    const double tol = fp32 ? 1.0e-8 : 1.0e-15;
    if (fp32)
        hipLaunchKernelGGL(k_ttm3f<float>, dim3((unsigned)npts), dim3(64), smem, ctx->stream,
                           (const float *)crd, ldc, nw, npts, tol, sc, E);
    else
        hipLaunchKernelGGL(k_ttm3f<double>, dim3((unsigned)npts), dim3(64), smem, ctx->stream,
                           (const double *)crd, ldc, nw, npts, tol, sc, E);
    return hipGetLastError();
}

// ---- energies and gradients (ttm3f.cpp:325-716) --------------------------------

// smear12 / smear123 without the potential term (ttm3f-bits-smear.cpp:142-190)
__device__ __forceinline__ void smear123(double r12, double AA, double &ts1, double &ts2, double &ts3)
{
    const double dri = 1.0 / r12, drsqi = dri * dri;
    const double rA = r12 / AA, rA3 = rA * rA * rA;
    const double exp1 = exp(-ttm::aXX * rA3);
    const double AA3 = AA * AA * AA;
    ts1 = (1.0 - exp1) * dri * drsqi;
    ts2 = (ts1 - exp1 * ttm::aXX / AA3) * drsqi;
    ts3 = (ts2 - 0.6 * exp1 * r12 * ttm::aXX * ttm::aXX / (AA3 * AA3)) * drsqi;
}

// PS monomer energy (as ps_energy) and gradient dr[9] (ps.cpp:486-603), kcal/mol
__device__ double ps_energy_grad(const double *O, const double *H1, const double *H2, double *dr)
{
    using namespace ttm;
    double ROH1[3], ROH2[3], RHH[3], d1 = 0, d2 = 0, dh = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = H1[i] - O[i];
        ROH2[i] = H2[i] - O[i];
        RHH[i] = H1[i] - H2[i];
        d1 += ROH1[i] * ROH1[i];
        d2 += ROH2[i] * ROH2[i];
        dh += RHH[i] * RHH[i];
    }
    d1 = sqrt(d1);
    d2 = sqrt(d2);
    dh = sqrt(dh);
    const double costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (d1 * d2);
    const double deoh = f5z * deohA;
    const double e1 = exp(-alphaoh * (d1 - roh));
    const double e2 = exp(-alphaoh * (d2 - roh));
    const double Va_shift = deoh * ((e1 - 1.0) * (e1 - 1.0) + (e2 - 1.0) * (e2 - 1.0));
    const double Vb = f5z * phh1A * exp(phh2 * (1.0 - dh));
    const double dVa1 = 2.0 * alphaoh * deoh * e1 * (1.0 - e1) / d1;
    const double dVa2 = 2.0 * alphaoh * deoh * e2 * (1.0 - e2) / d2;
    const double dVb = -phh2 * Vb / dh;
    const double x1 = (d1 - reoh) / reoh, x2 = (d2 - reoh) / reoh, x3 = costh - costhe;
    double f0[16], f1[16], f2[16];
    f0[0] = f1[0] = f2[0] = 0.0;
    f0[1] = f1[1] = f2[1] = 1.0;
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        f0[j] = f0[j - 1] * x1;
        f1[j] = f1[j - 1] * x2;
        f2[j] = f2[j - 1] * x3;
    }
    const double r1 = d1 - reoh, r2 = d2 - reoh;
    const double efac = exp(-b1 * (r1 * r1 + r2 * r2));
    double sum0 = 0, sum1 = 0, sum2 = 0, sum3 = 0;
#pragma unroll
    for (int j = 1; j < 245; ++j) {
        const int inI = psdata::scp3_ps_idx1[j], inJ = psdata::scp3_ps_idx2[j], inK = psdata::scp3_ps_idx3[j];
        const double c = psdata::scp3_ps_c5z[j];
        const double s01 = f0[inI] * f1[inJ] + f0[inJ] * f1[inI];
        sum0 += c * s01 * f2[inK];
        sum1 += c * ((inI - 1) * f0[inI - 1] * f1[inJ] + (inJ - 1) * f0[inJ - 1] * f1[inI]) * f2[inK];
        sum2 += c * ((inJ - 1) * f0[inI] * f1[inJ - 1] + (inI - 1) * f0[inJ] * f1[inI - 1]) * f2[inK];
        sum3 += c * s01 * (inK - 1) * f2[inK - 1];
    }
    const double Kconst = 2.0 * psdata::scp3_ps_c5z[0] - 2.0 * deoh + 0.44739574026257;
    const double dVcdr1 = (-2 * b1 * efac * r1 * sum0 + efac * sum1 / reoh) / d1;
    const double dVcdr2 = (-2 * b1 * efac * r2 * sum0 + efac * sum2 / reoh) / d2;
    const double dVcdcth = efac * sum3;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double gh1 = dVa1 * ROH1[i] + dVb * RHH[i] + dVcdr1 * ROH1[i] +
                           dVcdcth * (ROH2[i] / (d1 * d2) - costh * ROH1[i] / (d1 * d1));
        const double gh2 = dVa2 * ROH2[i] - dVb * RHH[i] + dVcdr2 * ROH2[i] +
                           dVcdcth * (ROH1[i] / (d1 * d2) - costh * ROH2[i] / (d2 * d2));
        dr[3 + i] = gh1 * cm2kcal;
        dr[6 + i] = gh2 * cm2kcal;
        dr[i] = -(gh1 + gh2) * cm2kcal;
    }
    return (Va_shift + Vb + efac * sum0 + Kconst) * cm2kcal;
}

// DMS charges (as dms_charges) and their derivatives, scaled as ttm3f.cpp:357-392:
// grdq[k + 3(j + 3i)] = d charge_j / d atom_i along k, j = (H1, H2, M), i = (H1, H2, O)
__device__ void dms_charges_grad(const double *O, const double *H1, const double *H2, double *chg,
                                 double *grdq)
{
    using namespace ttm;
    double ROH1[3], ROH2[3], d1 = 0, d2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ROH1[i] = H1[i] - O[i];
        ROH2[i] = H2[i] - O[i];
        d1 += ROH1[i] * ROH1[i];
        d2 += ROH2[i] * ROH2[i];
    }
    d1 = sqrt(d1);
    d2 = sqrt(d2);
    const double costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (d1 * d2);
    const double r1 = d1 - reoh, r2 = d2 - reoh;
    const double efac = exp(-b1D * (r1 * r1 + r2 * r2));
    const double x1 = r1 / reoh, x2 = r2 / reoh, x3 = costh - costhe;
    double f0[16], f1[16], f2[16];
    f0[0] = f1[0] = f2[0] = 0.0;
    f0[1] = f1[1] = f2[1] = 1.0;
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        f0[j] = f0[j - 1] * x1;
        f1[j] = f1[j - 1] * x2;
        f2[j] = f2[j - 1] * x3;
    }
    double p1 = 0, p2 = 0, dp1dr1 = 0, dp1dr2 = 0, dp1dc = 0, dp2dr1 = 0, dp2dr2 = 0, dp2dc = 0;
#pragma unroll
    for (int j = 1; j < 84; ++j) {
        const int inI = psdata::scp3_ps_idxD0[j], inJ = psdata::scp3_ps_idxD1[j], inK = psdata::scp3_ps_idxD2[j];
        const double c = psdata::scp3_ps_coefD[j];
        p1 += c * f0[inI] * f1[inJ] * f2[inK];
        p2 += c * f0[inJ] * f1[inI] * f2[inK];
        dp1dr1 += c * (inI - 1) * f0[inI - 1] * f1[inJ] * f2[inK];
        dp1dr2 += c * (inJ - 1) * f0[inI] * f1[inJ - 1] * f2[inK];
        dp1dc += c * (inK - 1) * f0[inI] * f1[inJ] * f2[inK - 1];
        dp2dr1 += c * (inJ - 1) * f0[inJ - 1] * f1[inI] * f2[inK];
        dp2dr2 += c * (inI - 1) * f0[inJ] * f1[inI - 1] * f2[inK];
        dp2dc += c * (inK - 1) * f0[inJ] * f1[inI] * f2[inK - 1];
    }
    const double xx = 0.529177249, xx2 = xx * xx;
    dp1dr1 /= reoh / xx;
    dp1dr2 /= reoh / xx;
    dp2dr1 /= reoh / xx;
    dp2dr2 /= reoh / xx;
    const double pl1 = costh, pl2 = 0.5 * (3.0 * pl1 * pl1 - 1.0);
    const double ang_part = pc0 + pl1 * pc1 + pl2 * pc2;
    const double pc = pa * (pow(d1, pb) + pow(d2, pb)) * ang_part;
    const double dpc0dr1 = pa * pb * pow(d1, pb - 1) * ang_part * xx2;
    const double dpc0dr2 = pa * pb * pow(d2, pb - 1) * ang_part * xx2;
    const double dpc0dc = pa * (pow(d1, pb) + pow(d2, pb)) * (pc1 + 0.5 * (6.0 * pl1) * pc2) * xx;
    const double defacdr1 = -2.0 * b1D * r1 * efac * xx;
    const double defacdr2 = -2.0 * b1D * r2 * efac * xx;
    dp1dr1 = dp1dr1 * efac + p1 * defacdr1 + dpc0dr1;
    dp1dr2 = dp1dr2 * efac + p1 * defacdr2 + dpc0dr2;
    dp1dc = dp1dc * efac + dpc0dc;
    dp2dr1 = dp2dr1 * efac + p2 * defacdr1 + dpc0dr1;
    dp2dr2 = dp2dr2 * efac + p2 * defacdr2 + dpc0dr2;
    dp2dc = dp2dc * efac + dpc0dc;
    p1 = psdata::scp3_ps_coefD[0] + p1 * efac + pc * xx;
    p2 = psdata::scp3_ps_coefD[0] + p2 * efac + pc * xx;
    double q0 = -(p1 + p2), q1 = p1, q2 = p2;
    dp1dr1 /= xx;
    dp1dr2 /= xx;
    dp2dr1 /= xx;
    dp2dr2 /= xx;
    // TTM3 modification (ps.cpp:732-769)
    const double ax = ROH1[1] * ROH2[2] - ROH1[2] * ROH2[1];
    const double ay = -ROH1[0] * ROH2[2] + ROH1[2] * ROH2[0];
    const double az = ROH1[0] * ROH2[1] - ROH1[1] * ROH2[0];
    const double sinth = sqrt(ax * ax + ay * ay + az * az) / (d1 * d2);
    const double ang = atan2(sinth, costh);
    const double m1 = dms1 * (d1 - dms2) + dms3 * (ang - ath0);
    const double m2 = dms1 * (d2 - dms2) + dms3 * (ang - ath0);
    q0 = q0 - (m1 + m2);
    q1 = q1 + m1;
    q2 = q2 + m2;
    dp1dr1 += dms1;
    dp2dr2 += dms1;
    dp1dc -= dms3 / sinth;
    dp2dc -= dms3 / sinth;
    // ps.cpp:773-826, then ttm3f.cpp:357-392
    const double f1q1r13 = (dp1dr1 - (dp1dc * costh / d1)) / d1;
    const double f1q1r23 = dp1dc / (d1 * d2);
    const double f2q1r23 = (dp1dr2 - (dp1dc * costh / d2)) / d2;
    const double f2q1r13 = dp1dc / (d2 * d1);
    const double f1q2r13 = (dp2dr1 - (dp2dc * costh / d1)) / d1;
    const double f1q2r23 = dp2dc / (d1 * d2);
    const double f2q2r23 = (dp2dr2 - (dp2dc * costh / d2)) / d2;
    const double f2q2r13 = dp2dc / (d2 * d1);
    const double tmp = gamma2 / gamma1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double dq[3][3];  // [atom H1 H2 O][charge H1 H2 O]
        dq[0][0] = f1q1r13 * ROH1[k] + f1q1r23 * ROH2[k];
        dq[1][0] = f2q1r13 * ROH1[k] + f2q1r23 * ROH2[k];
        dq[2][0] = -(dq[0][0] + dq[1][0]);
        dq[0][1] = f1q2r13 * ROH1[k] + f1q2r23 * ROH2[k];
        dq[1][1] = f2q2r13 * ROH1[k] + f2q2r23 * ROH2[k];
        dq[2][1] = -(dq[0][1] + dq[1][1]);
        dq[0][2] = -(dq[0][0] + dq[0][1]);
        dq[1][2] = -(dq[1][0] + dq[1][1]);
        dq[2][2] = -(dq[2][0] + dq[2][1]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            grdq[k + 3 * (0 + 3 * i)] = CHARGECON * (dq[i][0] + tmp * (dq[i][0] + dq[i][1]));
            grdq[k + 3 * (1 + 3 * i)] = CHARGECON * (dq[i][1] + tmp * (dq[i][1] + dq[i][0]));
            grdq[k + 3 * (2 + 3 * i)] = CHARGECON * (dq[i][2] - 2 * tmp * (dq[i][0] + dq[i][1]));
        }
    }
    chg[0] = CHARGECON * (q0 / gamma1);
    chg[1] = CHARGECON * (q1 + tmp * (q1 + q2));
    chg[2] = CHARGECON * (q2 + tmp * (q1 + q2));
}

// Charge-charge pair terms of molecules i < j with their gradient: slot and
// tens as pair_terms; gs[18] receives the direct forces on O_i H_i1 H_i2
// O_j H_j1 H_j2 (ttm3f.cpp:397-486); returns the O-O term.
__device__ double pair_terms_grad(const double *x, const double *ms, const double *q, int i, int j,
                                  const SmearConst &sc, double *slot, double *tens, double *gs)
{
    using namespace ttm;
    double ts0, ts1, ts2;
#pragma unroll
    for (int s = 0; s < 12; ++s) slot[s] = 0.0;
#pragma unroll
    for (int s = 0; s < 18; ++s) gs[s] = 0.0;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {  // M(a) - H(b), ttm3f.cpp:403-430
        const int a = dir ? j : i, b = dir ? i : j;
        double *phiM = slot + (dir ? 3 : 0), *phiH = slot + (dir ? 0 : 3), *E = slot + (dir ? 9 : 6);
        double *gH = gs + (dir ? 0 : 9);
#pragma unroll
        for (int l = 1; l < 3; ++l) {
            double R[3], rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R[k] = ms[3 * a + k] - x[9 * b + 3 * l + k];
                rsq += R[k] * R[k];
            }
            smear<double>(sqrt(rsq), sc.AA_HM, sc.a13, 1, ts0, ts1, ts2);
            phiM[0] += ts0 * q[3 * b + l];
            phiH[l] += ts0 * q[3 * a];
            const double t1 = q[3 * b + l] * ts1, t2 = q[3 * a] * t1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                E[k] += t1 * R[k];
                gH[3 * l + k] += t2 * R[k];
            }
        }
    }
#pragma unroll
    for (int ha = 1; ha < 3; ++ha)  // H-H, :439-463
#pragma unroll
        for (int hb = 1; hb < 3; ++hb) {
            double R[3], rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R[k] = x[9 * i + 3 * ha + k] - x[9 * j + 3 * hb + k];
                rsq += R[k] * R[k];
            }
            smear<double>(sqrt(rsq), sc.AA_HH, sc.a13, 1, ts0, ts1, ts2);
            slot[ha] += q[3 * j + hb] * ts0;
            slot[3 + hb] += q[3 * i + ha] * ts0;
            const double qq = q[3 * i + ha] * q[3 * j + hb] * ts1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                gs[3 * ha + k] -= qq * R[k];
                gs[9 + 3 * hb + k] += qq * R[k];
            }
        }
    double R[3], rsq = 0;  // O-O, :466-486
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R[k] = x[9 * i + k] - x[9 * j + k];
        rsq += R[k] * R[k];
    }
    const double r = sqrt(rsq), dR6 = rsq * rsq * rsq;
    const double expon = vdwD * exp(-vdwE * r);
    const double evdw = vdwC / dR6 + expon;
    const double tv = -(6.0 * vdwC / dR6) / rsq - vdwE * expon / r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gs[k] += tv * R[k];
        gs[9 + k] -= tv * R[k];
    }
    rsq = 0;  // M-M, :488-535
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R[k] = ms[3 * i + k] - ms[3 * j + k];
        rsq += R[k] * R[k];
    }
    smear<double>(sqrt(rsq), sc.AA_MM, sc.a13, 2, ts0, ts1, ts2);
    slot[0] += ts0 * q[3 * j];
    slot[3] += ts0 * q[3 * i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        slot[6 + k] += q[3 * j] * ts1 * R[k];
        slot[9 + k] -= q[3 * i] * ts1 * R[k];
        tens[k] = R[k];
    }
    tens[3] = ts1;
    tens[4] = ts2;
    return evdw;
}

// Charge-dipole and dipole-dipole terms of molecules i < j, both directions
// (ttm3f.cpp:596-668).  out[24]: dM_i, dM_j, forces on H_i1 H_i2, on H_j1
// H_j2, potential changes at M_i H_i1 H_i2, at M_j H_j1 H_j2.
__device__ void pair_dipole_grad(const double *x, const double *ms, const double *q, const double *dip,
                                 int i, int j, const SmearConst &sc, double *out)
{
#pragma unroll
    for (int s = 0; s < 24; ++s) out[s] = 0.0;
    double ts1, ts2, ts3;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const int a = dir ? j : i, b = dir ? i : j;
        double *dMa = out + (dir ? 3 : 0), *dMb = out + (dir ? 0 : 3);
        double *gHb = out + (dir ? 6 : 12), *phib = out + (dir ? 18 : 21);
#pragma unroll
        for (int l = 1; l < 3; ++l) {  // dipole of M(a) with the charge of H(b, l)
            double R[3], rsq = 0, diR = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R[k] = ms[3 * a + k] - x[9 * b + 3 * l + k];
                rsq += R[k] * R[k];
                diR += dip[3 * a + k] * R[k];
            }
            smear123(sqrt(rsq), sc.AA_HM, ts1, ts2, ts3);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double derij = q[3 * b + l] * (3 * ts2 * diR * R[k] - ts1 * dip[3 * a + k]);
                dMa[k] += derij;
                gHb[3 * (l - 1) + k] -= derij;
            }
            phib[l] -= ts1 * diR;
        }
        double R[3], rsq = 0, diR = 0;  // dipole of M(a) with the charge of M(b)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R[k] = ms[3 * a + k] - ms[3 * b + k];
            rsq += R[k] * R[k];
            diR += dip[3 * a + k] * R[k];
        }
        smear123(sqrt(rsq), sc.AA_MM, ts1, ts2, ts3);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double derij = q[3 * b] * (3 * ts2 * diR * R[k] - ts1 * dip[3 * a + k]);
            dMa[k] += derij;
            dMb[k] -= derij;
        }
        phib[0] -= ts1 * diR;
        if (dir == 0) {  // dipole-dipole, once per pair
            double djR = 0, didj = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                djR += dip[3 * b + k] * R[k];
                didj += dip[3 * a + k] * dip[3 * b + k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double derij = -3 * ts2 * (didj * R[k] + djR * dip[3 * a + k] + diR * dip[3 * b + k]) +
                                     15 * ts3 * diR * djR * R[k];
                dMa[k] += derij;
                dMb[k] -= derij;
            }
        }
    }
}

// The monomer terms of the gradient, one lane per (configuration,
// molecule) so that every lane works (inside k_ttm3f_grad only nw of 64
// would): PS energy and gradient (ps.cpp:486-603) and DMS charges with
// their derivatives (ps.cpp:605-836, scaled as ttm3f.cpp:357-392).
// mono[f * ldm + cfg * nw + w]: f = 0 energy, 1-9 PS gradient, 10-12
// charges, 13-39 charge derivatives.
constexpr int kMonoFields = 40;

__global__ void __launch_bounds__(256)
    k_ttm3f_monomer_grad(const double *__restrict__ crd, int64_t ldc, int nw, int64_t npts,
                         double *__restrict__ mono, int64_t ldm)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npts * nw) return;
    const int64_t cfg = idx / nw;
    const int w = (int)(idx - cfg * nw);
    double x[9], g[9], q[3], gq[27];
#pragma unroll
    for (int k = 0; k < 9; ++k) x[k] = crd[(size_t)(9 * w + k) * ldc + cfg];
#pragma unroll
    for (int k = 0; k < 27; ++k) gq[k] = 0.0;
    const double e = ps_energy_grad(x, x + 3, x + 6, g);
    dms_charges_grad(x, x + 3, x + 6, q, gq);
    mono[idx] = e;
#pragma unroll
    for (int k = 0; k < 9; ++k) mono[(size_t)(1 + k) * ldm + idx] = g[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) mono[(size_t)(10 + k) * ldm + idx] = q[k];
#pragma unroll
    for (int k = 0; k < 27; ++k) mono[(size_t)(13 + k) * ldm + idx] = gq[k];
}

// The charge and O-O pair terms of the gradient (ttm3f.cpp:403-560), one
// lane per (configuration, molecule pair), for the same reason.
// pairs[f * ldp + cfg * P + p]: f = 0 the pair's O-O energy, 1-12 its
// potential/field slots, 13-17 its smeared dipole tensor, 18-35 its force
// slots (pair_terms_grad's outputs, molecules i < j of pair p).
constexpr int kPairFields = 36;

__global__ void __launch_bounds__(256)
    k_ttm3f_pair_grad(const double *__restrict__ crd, int64_t ldc, int nw, int64_t npts, SmearConst sc,
                      const double *__restrict__ mono, int64_t ldm, double *__restrict__ pairs, int64_t ldp)
{
    using namespace ttm;
    const int P = nw * (nw - 1) / 2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npts * P) return;
    const int64_t cfg = idx / P;
    const int p = (int)(idx - cfg * P);
    int i = 0, rem = p;
    while (rem >= nw - 1 - i) {
        rem -= nw - 1 - i;
        ++i;
    }
    const int j = i + 1 + rem;
    // molecules i and j as molecules 0 and 1 of a two-molecule cluster
    double x[18], ms[6], q[6], slot[12], tens[5], gs[18];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        x[k] = crd[(size_t)(9 * i + k) * ldc + cfg];
        x[9 + k] = crd[(size_t)(9 * j + k) * ldc + cfg];
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            ms[3 * m + k] = gamma1 * x[9 * m + k] + gamma2 * (x[9 * m + 3 + k] + x[9 * m + 6 + k]);
    const size_t m0 = (size_t)cfg * nw;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        q[k] = mono[(size_t)(10 + k) * ldm + m0 + i];
        q[3 + k] = mono[(size_t)(10 + k) * ldm + m0 + j];
    }
    const double ev = pair_terms_grad(x, ms, q, 0, 1, sc, slot, tens, gs);
    pairs[idx] = ev;
#pragma unroll
    for (int f = 0; f < 12; ++f) pairs[(size_t)(1 + f) * ldp + idx] = slot[f];
#pragma unroll
    for (int f = 0; f < 5; ++f) pairs[(size_t)(13 + f) * ldp + idx] = tens[f];
#pragma unroll
    for (int f = 0; f < 18; ++f) pairs[(size_t)(18 + f) * ldp + idx] = gs[f];
}

// One configuration per wave, FP64; phases as k_ttm3f plus the gradient:
// per-pair force slots summed per coordinate in a fixed partner order, the
// dipole terms after the iteration, then the M-site distribution and the
// charge-derivative terms per molecule.  The monomer terms come from
// k_ttm3f_monomer_grad, the first pair terms from k_ttm3f_pair_grad.
__global__ void __launch_bounds__(64)
    k_ttm3f_grad(const double *__restrict__ crd, int64_t ldc, int nw, int64_t npts, double tol,
                 SmearConst sc, const double *__restrict__ mono, int64_t ldm,
                 const double *__restrict__ pairs, int64_t ldp, double *__restrict__ E,
                 double *__restrict__ G, int64_t ldg)
{
    using namespace ttm;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int64_t cfg = blockIdx.x;
    const int n3 = 3 * nw, P = nw * (nw - 1) / 2;
    double *dipL = smem;              // [64]
    double *x = smem + 64;            // [9nw]
    double *ms = x + 9 * nw;          // [3nw]
    double *q = ms + n3;              // [3nw]
    double *g = q + n3;               // [9nw]
    double *grdq = g + 9 * nw;        // [27nw]
    double *phiS = grdq + 27 * nw;    // [3nw]
    double *efqS = phiS + n3;         // [3nw]
    double *dMS = efqS + n3;          // [3nw]
    double *tens = dMS + n3;          // [P][5]
    double *slots = tens + 5 * P;     // [P][12], then [P][18] forces; [P][24] after the iteration
    double *gsl = slots + 12 * P;

    for (int c = lane; c < 9 * nw; c += 64) x[c] = crd[(size_t)c * ldc + cfg];
    __syncthreads();
    double eint = 0.0;
    const size_t m0 = (size_t)cfg * nw;
    if (lane < nw) {
        const int w = lane;
        const double *O = x + 9 * w, *H1 = O + 3, *H2 = O + 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) ms[3 * w + k] = gamma1 * O[k] + gamma2 * (H1[k] + H2[k]);
        eint = mono[m0 + w];
    }
    for (int e = lane; e < (kMonoFields - 1) * nw; e += 64) {
        const int f = 1 + e / nw, w = e - (f - 1) * nw;
        const double v = mono[(size_t)f * ldm + m0 + w];
        if (f < 10)
            g[9 * w + f - 1] = v;
        else if (f < 13)
            q[3 * w + f - 10] = v;
        else
            grdq[27 * w + f - 13] = v;
    }
    __syncthreads();
    double evdw = 0.0;
    for (int p = lane; p < P; p += 64) {
        const size_t pi = (size_t)cfg * P + p;
        evdw += pairs[pi];
#pragma unroll
        for (int f = 0; f < 12; ++f) slots[12 * p + f] = pairs[(size_t)(1 + f) * ldp + pi];
#pragma unroll
        for (int f = 0; f < 5; ++f) tens[5 * p + f] = pairs[(size_t)(13 + f) * ldp + pi];
#pragma unroll
        for (int f = 0; f < 18; ++f) gsl[18 * p + f] = pairs[(size_t)(18 + f) * ldp + pi];
    }
    __syncthreads();
    // forces of the charge-charge and O-O terms, per coordinate over partners
    for (int c = lane; c < 9 * nw; c += 64) {
        const int i = c / 9, e = c - 9 * i;
        double s = 0.0;
        for (int j = 0; j < nw; ++j) {
            if (j == i) continue;
            const bool first = i < j;
            s += gsl[18 * pair_index(first ? i : j, first ? j : i, nw) + (first ? 0 : 9) + e];
        }
        g[c] += s;
    }
    double phi = 0.0, efq = 0.0;
    if (lane < n3) {
        const int i = lane / 3, t = lane - 3 * (lane / 3);
        for (int j = 0; j < nw; ++j) {
            if (j == i) continue;
            const bool first = i < j;
            const double *sl = slots + 12 * pair_index(first ? i : j, first ? j : i, nw);
            phi += sl[(first ? 0 : 3) + t];
            efq += sl[(first ? 6 : 9) + t];
        }
        phiS[lane] = phi;
        efqS[lane] = efq;
    }
    const double eelec = 0.5 * wave_sum(lane < n3 ? phi * q[lane] : 0.0);
    // induced dipoles, as k_ttm3f
    const int ii = lane / 3, kk = lane - 3 * (lane / 3);
    double dip = lane < n3 ? polarM * efq : 0.0;
    double dip_prev = dip;
    const double stath = DEBYE / CHARGECON / sqrt((double)nw);
    for (int it = 0; it < maxiter; ++it) {
        if (lane < n3) dipL[lane] = dip;
        __syncthreads();
        double efd = 0.0;
        if (lane < n3) {
            for (int j = 0; j < nw; ++j) {
                if (j == ii) continue;
                const double *tp = tens + 5 * pair_index(ii < j ? ii : j, ii < j ? j : ii, nw);
                const double R0 = tp[0], R1 = tp[1], R2 = tp[2], t1 = tp[3], t2 = tp[4];
                const double m0 = dipL[3 * j], m1 = dipL[3 * j + 1], m2 = dipL[3 * j + 2];
                const double Rk = kk == 0 ? R0 : (kk == 1 ? R1 : R2);
                const double mk = kk == 0 ? m0 : (kk == 1 ? m1 : m2);
                efd += 3.0 * t2 * Rk * (R0 * m0 + R1 * m1 + R2 * m2) - t1 * mk;
            }
        }
        double dn = lane < n3 ? polarM * (efq + efd) : 0.0;
        dn = dmix * dn + (1.0 - dmix) * dip_prev;
        const double delta = lane < n3 ? dn - dip_prev : 0.0;
        const double dd = sqrt(wave_sum(delta * delta)) * stath;
        dip = dn;
        __syncthreads();
        if (dd < tol) break;
        dip_prev = dip;
    }
    const double eind = -0.5 * wave_sum(lane < n3 ? dip * efq : 0.0);
    if (lane < n3) dipL[lane] = dip;
    __syncthreads();
    // charge-dipole and dipole-dipole terms per pair into [P][24]
    for (int p = lane; p < P; p += 64) {
        int i = 0, rem = p;
        while (rem >= nw - 1 - i) {
            rem -= nw - 1 - i;
            ++i;
        }
        pair_dipole_grad(x, ms, q, dipL, i, i + 1 + rem, sc, slots + 24 * p);
    }
    __syncthreads();
    // sums per molecule: forces on H (6nw), dM (3nw), potential (3nw)
    for (int t = lane; t < 12 * nw; t += 64) {
        int i, e, base_first, base_second;
        if (t < 6 * nw) {
            i = t / 6, e = t - 6 * i;
            base_first = 6, base_second = 12;
        } else if (t < 9 * nw) {
            i = (t - 6 * nw) / 3, e = t - 6 * nw - 3 * i;
            base_first = 0, base_second = 3;
        } else {
            i = (t - 9 * nw) / 3, e = t - 9 * nw - 3 * i;
            base_first = 18, base_second = 21;
        }
        double s = 0.0;
        for (int j = 0; j < nw; ++j) {
            if (j == i) continue;
            const bool first = i < j;
            s += slots[24 * pair_index(first ? i : j, first ? j : i, nw) + (first ? base_first : base_second) + e];
        }
        if (t < 6 * nw) g[9 * i + 3 + e] += s;
        else if (t < 9 * nw) dMS[3 * i + e] = s;
        else phiS[3 * i + e] += s;
    }
    __syncthreads();
    // M-site distribution (:672-686) and the charge derivatives (:690-713)
    if (lane < nw) {
        const int n = lane;
        const double *gq = grdq + 27 * n;
        const double pM = phiS[3 * n], pH1 = phiS[3 * n + 1], pH2 = phiS[3 * n + 2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double tmp = dMS[3 * n + k] - q[3 * n] * efqS[3 * n + k];
            g[9 * n + k] += gamma1 * tmp + gq[k + 3 * (0 + 3 * 2)] * pH1 + gq[k + 3 * (1 + 3 * 2)] * pH2 +
                            gq[k + 3 * (2 + 3 * 2)] * pM;
            g[9 * n + 3 + k] += gamma2 * tmp + gq[k + 3 * (0 + 3 * 0)] * pH1 + gq[k + 3 * (1 + 3 * 0)] * pH2 +
                                gq[k + 3 * (2 + 3 * 0)] * pM;
            g[9 * n + 6 + k] += gamma2 * tmp + gq[k + 3 * (0 + 3 * 1)] * pH1 + gq[k + 3 * (1 + 3 * 1)] * pH2 +
                                gq[k + 3 * (2 + 3 * 1)] * pM;
        }
    }
    __syncthreads();
    for (int c = lane; c < 9 * nw; c += 64) G[(size_t)c * ldg + cfg] = g[c];
    const double ei = wave_sum(eint);
    const double ev = wave_sum(evdw);
    if (lane == 0) E[cfg] = ei + ev + eelec + eind;
}

size_t ttm3f_grad_smem_bytes(int nw)
{
    const int P = nw * (nw - 1) / 2;
    const int pairs = 5 * P + (12 * P + 18 * P > 24 * P ? 12 * P + 18 * P : 24 * P);
    return sizeof(double) * (size_t)(64 + 60 * nw + pairs);
}

hipError_t launch_ttm3f_grad(scp3g_ctx *ctx, int nw, int64_t npts, const double *crd, int64_t ldc,
                             double *E, double *G, int64_t ldg)
{
    if (npts <= 0) return hipSuccess;
    if (nw < 1 || 3 * nw > 64) return hipErrorInvalidValue;
    const size_t smem = ttm3f_grad_smem_bytes(nw);
    if (smem > 64 * 1024) return hipErrorInvalidValue;
    // monomer- and pair-term scratch, kept in the context and grown on
    // demand; configurations in chunks of at most ~1 GiB of scratch
    const int P = nw * (nw - 1) / 2;
    const size_t per_cfg = sizeof(double) * ((size_t)kMonoFields * nw + (size_t)kPairFields * P);
    const int64_t chunk = std::max<int64_t>(256, std::min<int64_t>(npts, (int64_t)((1ull << 30) / per_cfg)));
    const size_t need = per_cfg * (size_t)chunk;
    if (ctx->mono_bytes < need) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
        if (ctx->d_mono) hipFree(ctx->d_mono);
        ctx->d_mono = nullptr;
        ctx->mono_bytes = 0;
        if ((e = hipMalloc(&ctx->d_mono, need)) != hipSuccess) return e;
        ctx->mono_bytes = need;
    }
    const SmearConst sc = smear_const();
    for (int64_t c0 = 0; c0 < npts; c0 += chunk) {
        const int64_t nc = std::min<int64_t>(chunk, npts - c0);
        const int64_t ldm = nc * nw, ldp = nc * P;
        double *mono = ctx->d_mono, *pairs = ctx->d_mono + (size_t)kMonoFields * ldm;
        hipLaunchKernelGGL(k_ttm3f_monomer_grad, dim3((unsigned)((ldm + 255) / 256)), dim3(256), 0, ctx->stream,
                           crd + c0, ldc, nw, nc, mono, ldm);
        if (P > 0)
            hipLaunchKernelGGL(k_ttm3f_pair_grad, dim3((unsigned)((ldp + 255) / 256)), dim3(256), 0, ctx->stream,
                               crd + c0, ldc, nw, nc, sc, (const double *)mono, ldm, pairs, ldp);
        hipLaunchKernelGGL(k_ttm3f_grad, dim3((unsigned)nc), dim3(64), smem, ctx->stream, crd + c0, ldc, nw, nc,
                           1.0e-15, sc, (const double *)mono, ldm, (const double *)pairs, ldp, E + c0, G + c0, ldg);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Host FP64 evaluation of the same terms in the same order as the kernel
// (CPU potential for input preparation).
double ttm3f_energy_host(int nw, const double *crd, int *iters)
{
    using namespace ttm;
    const SmearConst sc = smear_const();
    const int n3 = 3 * nw, P = nw * (nw - 1) / 2;
    std::vector<double> ms(n3), q(n3), slots((size_t)12 * P), tens((size_t)5 * P);
    double eint = 0.0, evdw = 0.0;
    for (int w = 0; w < nw; ++w) {
        double e;
        monomer_terms<double>(crd, w, ms.data(), q.data(), e);
        eint += e;
    }
    for (int i = 0, p = 0; i < nw; ++i)
        for (int j = i + 1; j < nw; ++j, ++p)
            evdw += pair_terms<double>(crd, ms.data(), q.data(), i, j, sc, &slots[12 * p], &tens[5 * p]);
    std::vector<double> phi(n3, 0.0), efq(n3, 0.0), dip(n3), prev(n3), dn(n3);
    for (int s = 0; s < n3; ++s) {
        const int i = s / 3, t = s % 3;
        for (int j = 0; j < nw; ++j) {
            if (j == i) continue;
            const bool first = i < j;
            const double *sl = &slots[12 * pair_index(first ? i : j, first ? j : i, nw)];
            phi[s] += sl[(first ? 0 : 3) + t];
            efq[s] += sl[(first ? 6 : 9) + t];
        }
    }
    double eelec = 0.0;
    for (int s = 0; s < n3; ++s) eelec += phi[s] * q[s];
    eelec *= 0.5;
    for (int s = 0; s < n3; ++s) dip[s] = prev[s] = polarM * efq[s];
    const double stath = DEBYE / CHARGECON / std::sqrt((double)nw);
    int it = 0;
    for (; it < maxiter; ++it) {
        double dd = 0.0;
        for (int r = 0; r < n3; ++r) {
            const int i = r / 3, kk = r % 3;
            double efd = 0.0;
            for (int j = 0; j < nw; ++j) {
                if (j == i) continue;
                const double *tp = &tens[5 * pair_index(i < j ? i : j, i < j ? j : i, nw)];
                const double dot = tp[0] * dip[3 * j] + tp[1] * dip[3 * j + 1] + tp[2] * dip[3 * j + 2];
                efd += 3.0 * tp[4] * tp[kk] * dot - tp[3] * dip[3 * j + kk];
            }
            dn[r] = dmix * (polarM * (efq[r] + efd)) + (1.0 - dmix) * prev[r];
            dd += (dn[r] - prev[r]) * (dn[r] - prev[r]);
        }
        dip = dn;
        if (std::sqrt(dd) * stath < 1.0e-15) break;
        prev = dip;
    }
    if (iters) *iters = it + 1;
    double eind = 0.0;
    for (int s = 0; s < n3; ++s) eind -= dip[s] * efq[s];
    eind *= 0.5;
    return eint + evdw + eelec + eind;
}

}  // namespace scp3g
