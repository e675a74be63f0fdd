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
