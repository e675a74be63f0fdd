// q-TIP4P/F energies of a batch of configurations, one configuration per lane.
//
// Reference: h2o::qtip4pf::operator()(nw, crd), qtip4pf.cpp:152-248, with
// the monomer term of qtip4pf.cpp:57-107 and parameters qtip4pf.cpp:23-43
// (Habershon, Markland, Manolopoulos, JCP 131, 024501 (2009)).
//
// A block of 64 lanes stages its 64 configurations in LDS as [coordinate]
// [lane] (conflict-free: consecutive lanes, consecutive banks) and each lane
// then runs the reference's monomer and pair loops in registers.  T is the
// arithmetic type (double for the FP64 form, float for the mixed form);
// energies are returned in kcal/mol as double.
#include "scp3g_internal.h"

#include <vector>

namespace scp3g {

namespace qtip4pf {
constexpr double CHARGECON = 18.22234397655801030455;
constexpr double qM = -1.1128 * CHARGECON;
constexpr double qH = -qM / 2;
constexpr double rOHeq = 0.9419;
constexpr double aHOHeq = 107.4;
constexpr double epsilon = 0.1852;
constexpr double sigma = 3.1589;
constexpr double alpha_r = 2.287;
constexpr double D_r = 116.09;
constexpr double k_theta = 87.85;
constexpr double gammaM = 0.73612;
constexpr double gamma1 = (1.0 - gammaM) / 2;
}  // namespace qtip4pf

template <typename T>
__host__ __device__ __forceinline__ T monomer(const T *__restrict__ s, int stride)
{
    using namespace qtip4pf;
    T d1[3], d2[3], r1 = 0, r2 = 0, d12 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        d1[i] = s[(i + 3) * stride] - s[i * stride];
        r1 += d1[i] * d1[i];
        d2[i] = s[(i + 6) * stride] - s[i * stride];
        r2 += d2[i] * d2[i];
        d12 += d1[i] * d2[i];
    }
    r1 = sqrt(r1);
    r2 = sqrt(r2);
    const T x1 = T(alpha_r) * (r1 - T(rOHeq));
    const T x2 = T(alpha_r) * (r2 - T(rOHeq));
    const T V1 = x1 * x1 * (T(1) - x1 * (T(1) - T(7.0 / 12.0) * x1)) * T(D_r);
    const T V2 = x2 * x2 * (T(1) - x2 * (T(1) - T(7.0 / 12.0) * x2)) * T(D_r);
    const T r12 = T(1) / r1 / r2;
    const T xi = d12 * r12;
    const T theta = acos(xi);
    const T d_theta = theta - T(aHOHeq * M_PI / 180.0);
    return V1 + V2 + T(0.5 * k_theta) * d_theta * d_theta;
}

// Energy of one configuration whose coordinate c sits at me[c * st]; mm is
// scratch for the 3nw M-site coordinates at the same stride.  Terms are
// evaluated in T and summed in double: the Coulomb terms are O(100) kcal/mol
// and cancel to a few kcal/mol, so a float running sum would lose ~1e-4.
template <typename T>
__host__ __device__ double qtip4pf_energy(const T *__restrict__ me, T *__restrict__ mm, int st, int nw)
{
    using namespace qtip4pf;
    double Eint = 0;
    for (int w = 0; w < nw; ++w) {
        const T *O = me + 9 * w * st;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            mm[(3 * w + k) * st] =
                T(gammaM) * O[k * st] + T(gamma1) * (O[(3 + k) * st] + O[(6 + k) * st]);
        Eint += monomer<T>(O, st);
    }
    double Evdw = 0, Eelec = 0;
    for (int a = 0; a < nw; ++a) {
        // H(b) - M(a), qtip4pf.cpp:179-198
        for (int b = 0; b < nw; ++b) {
            if (a == b) continue;
#pragma unroll
            for (int l = 1; l < 3; ++l) {
                T Rsq = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const T dx = mm[(3 * a + k) * st] - me[(9 * b + 3 * l + k) * st];
                    Rsq += dx * dx;
                }
                Eelec += T(qH * qM) / sqrt(Rsq);
            }
        }
        for (int b = a + 1; b < nw; ++b) {
            // H-H, qtip4pf.cpp:209-224
#pragma unroll
            for (int ha = 1; ha < 3; ++ha)
#pragma unroll
                for (int hb = 1; hb < 3; ++hb) {
                    T Rsq = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const T dx = me[(9 * a + 3 * ha + k) * st] - me[(9 * b + 3 * hb + k) * st];
                        Rsq += dx * dx;
                    }
                    Eelec += T(qH * qH) / sqrt(Rsq);
                }
            // O-O Lennard-Jones, qtip4pf.cpp:227-235
            T Rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T dx = me[(9 * a + k) * st] - me[(9 * b + k) * st];
                Rsq += dx * dx;
            }
            Rsq = T(sigma * sigma) / Rsq;
            const T dR6 = Rsq * Rsq * Rsq;
            Evdw += T(4 * epsilon) * dR6 * (dR6 - T(1));
            // M-M Coulomb, qtip4pf.cpp:237-243
            Rsq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T dx = mm[(3 * a + k) * st] - mm[(3 * b + k) * st];
                Rsq += dx * dx;
            }
            Eelec += T(qM * qM) / sqrt(Rsq);
        }
    }
    return Eint + Evdw + Eelec;
}

template <typename T>
__global__ void __launch_bounds__(64)
    k_qtip4pf(const T *__restrict__ crd, int64_t ldc, int nw, int64_t npts,
              double *__restrict__ E)
{
    extern __shared__ unsigned char smem_raw[];
    T *s = reinterpret_cast<T *>(smem_raw);  // [9nw][64]
    T *ms = s + (size_t)9 * nw * 64;          // M sites [3nw][64]
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const int64_t ic = i < npts ? i : npts - 1;
    for (int c = 0; c < 9 * nw; ++c) s[c * 64 + lane] = crd[(size_t)c * ldc + ic];
    // no barrier needed: every lane only reads its own column
    const double e = qtip4pf_energy<T>(s + lane, ms + lane, 64, nw);
    if (i < npts) E[i] = e;
}

// Energy and gradient of one configuration (qtip4pf.cpp:252-397, monomer
// gradient :84-103): coordinate c at me[c * st], dE/dx written to g[c * st]
// (kcal/mol/A); mm and dm are scratch for the M sites and their gradient
// (3nw each), distributed to O, H, H at the end as the reference does.
__host__ __device__ double qtip4pf_energy_grad(const double *__restrict__ me, double *__restrict__ mm,
                                      double *__restrict__ g, double *__restrict__ dm, int st,
                                      int nw)
{
    using namespace qtip4pf;
    double Eint = 0;
    for (int w = 0; w < nw; ++w) {
        const double *O = me + 9 * w * st;
        double *gw = g + 9 * w * st;
        double d1[3], d2[3], r1 = 0, r2 = 0, d12 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mm[(3 * w + k) * st] = gammaM * O[k * st] + gamma1 * (O[(3 + k) * st] + O[(6 + k) * st]);
            dm[(3 * w + k) * st] = 0.0;
            d1[k] = O[(k + 3) * st] - O[k * st];
            r1 += d1[k] * d1[k];
            d2[k] = O[(k + 6) * st] - O[k * st];
            r2 += d2[k] * d2[k];
            d12 += d1[k] * d2[k];
        }
        r1 = sqrt(r1);
        r2 = sqrt(r2);
        const double x1 = alpha_r * (r1 - rOHeq);
        const double x2 = alpha_r * (r2 - rOHeq);
        const double V1 = x1 * x1 * (1.0 - x1 * (1.0 - (7.0 / 12.0) * x1)) * D_r;
        const double V2 = x2 * x2 * (1.0 - x2 * (1.0 - (7.0 / 12.0) * x2)) * D_r;
        const double r12 = 1.0 / r1 / r2;
        const double xi = d12 * r12;
        const double d_theta = acos(xi) - aHOHeq * M_PI / 180.0;
        // d/dr of the Morse-quartic terms over r; d(theta)/d(xi) = -1/sqrt(1 - xi^2)
        const double dV1 = alpha_r * D_r * x1 * (2.0 - x1 * (3.0 - (7.0 / 3.0) * x1)) / r1;
        const double dV2 = alpha_r * D_r * x2 * (2.0 - x2 * (3.0 - (7.0 / 3.0) * x2)) / r2;
        const double f = -k_theta * d_theta / sqrt(1.0 - xi * xi);
        const double t0 = f * r12;
        const double t1 = dV1 - f * xi / r1 / r1;
        const double t2 = dV2 - f * xi / r2 / r2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a1 = t0 * d2[k] + t1 * d1[k];
            const double a2 = t0 * d1[k] + t2 * d2[k];
            gw[k * st] = -(a1 + a2);
            gw[(k + 3) * st] = a1;
            gw[(k + 6) * st] = a2;
        }
        Eint += V1 + V2 + 0.5 * k_theta * d_theta * d_theta;
    }
    double Evdw = 0, Eelec = 0;
    for (int a = 0; a < nw; ++a) {
        for (int b = 0; b < nw; ++b) {  // H(b) - M(a), qtip4pf.cpp:283-310
            if (a == b) continue;
            for (int l = 1; l < 3; ++l) {
                double dR[3], Rsq = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dR[k] = mm[(3 * a + k) * st] - me[(9 * b + 3 * l + k) * st];
                    Rsq += dR[k] * dR[k];
                }
                const double r1 = 1.0 / sqrt(Rsq);
                const double E1 = qH * qM * r1, g1 = E1 * r1 * r1;
                Eelec += E1;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dm[(3 * a + k) * st] -= g1 * dR[k];
                    g[(9 * b + 3 * l + k) * st] += g1 * dR[k];
                }
            }
        }
        for (int b = a + 1; b < nw; ++b) {
            for (int ha = 1; ha < 3; ++ha)  // H-H, :317-340
                for (int hb = 1; hb < 3; ++hb) {
                    double dR[3], Rsq = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        dR[k] = me[(9 * a + 3 * ha + k) * st] - me[(9 * b + 3 * hb + k) * st];
                        Rsq += dR[k] * dR[k];
                    }
                    const double r1 = 1.0 / sqrt(Rsq);
                    const double E1 = qH * qH * r1, g1 = E1 * r1 * r1;
                    Eelec += E1;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        g[(9 * a + 3 * ha + k) * st] -= g1 * dR[k];
                        g[(9 * b + 3 * hb + k) * st] += g1 * dR[k];
                    }
                }
            double dR[3], Rsq = 0;  // O-O Lennard-Jones, :343-357
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dR[k] = me[(9 * a + k) * st] - me[(9 * b + k) * st];
                Rsq += dR[k] * dR[k];
            }
            const double sR = sigma * sigma / Rsq;
            const double dR6 = sR * sR * sR;
            Evdw += 4 * epsilon * dR6 * (dR6 - 1.0);
            const double gv = 24 * epsilon * dR6 * (2 * dR6 - 1.0) / Rsq;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                g[(9 * a + k) * st] -= gv * dR[k];
                g[(9 * b + k) * st] += gv * dR[k];
            }
            Rsq = 0;  // M-M Coulomb, :360-374
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dR[k] = mm[(3 * a + k) * st] - mm[(3 * b + k) * st];
                Rsq += dR[k] * dR[k];
            }
            const double r1 = 1.0 / sqrt(Rsq);
            const double E1 = qM * qM * r1, g1 = E1 * r1 * r1;
            Eelec += E1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dm[(3 * a + k) * st] -= g1 * dR[k];
                dm[(3 * b + k) * st] += g1 * dR[k];
            }
        }
    }
    for (int a = 0; a < nw; ++a)  // M-site chain rule, :379-390
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double d = dm[(3 * a + k) * st];
            g[(9 * a + k) * st] += gammaM * d;
            g[(9 * a + 3 + k) * st] += gamma1 * d;
            g[(9 * a + 6 + k) * st] += gamma1 * d;
        }
    return Eint + Evdw + Eelec;
}

// One configuration per lane; a block of L lanes keeps its coordinates,
// gradient and M-site scratch in LDS as [coordinate][lane] (24nw x L doubles).
__global__ void __launch_bounds__(64)
    k_qtip4pf_grad(const double *__restrict__ crd, int64_t ldc, int nw, int64_t npts,
                   double *__restrict__ E, double *__restrict__ G, int64_t ldg)
{
    extern __shared__ double sg[];
    const int L = blockDim.x, lane = threadIdx.x;
    double *s = sg, *ms = s + (size_t)9 * nw * L, *gs = ms + (size_t)3 * nw * L,
           *dms = gs + (size_t)9 * nw * L;
    const int64_t i = (int64_t)blockIdx.x * L + lane;
    const int64_t ic = i < npts ? i : npts - 1;
    for (int c = 0; c < 9 * nw; ++c) s[c * L + lane] = crd[(size_t)c * ldc + ic];
    // each lane reads and writes only its own column: no barrier
    const double e = qtip4pf_energy_grad(s + lane, ms + lane, gs + lane, dms + lane, L, nw);
    if (i >= npts) return;
    E[i] = e;
    for (int c = 0; c < 9 * nw; ++c) G[(size_t)c * ldg + i] = gs[c * L + lane];
}

hipError_t launch_qtip4pf_grad(scp3g_ctx *ctx, int nw, int64_t npts, const double *crd, int64_t ldc,
                               double *E, double *G, int64_t ldg)
{
    if (npts <= 0) return hipSuccess;
    // lanes per block: the largest of 64, 32, 16 whose 24nw x L doubles fit 64 KiB
    int L = 64;
    while (L > 16 && (size_t)24 * nw * L * sizeof(double) > 64 * 1024) L /= 2;
    const size_t smem = (size_t)24 * nw * L * sizeof(double);
    if (smem > 64 * 1024) return hipErrorInvalidValue;  // more than 21 waters
    hipLaunchKernelGGL(k_qtip4pf_grad, dim3((unsigned)((npts + L - 1) / L)), dim3(L), smem,
                       ctx->stream, crd, ldc, nw, npts, E, G, ldg);
    return hipGetLastError();
}

// host FP64 energy and gradient (the CPU h2o::PES gradient overload)
double qtip4pf_energy_grad_host(int nw, const double *crd, double *grad)
{
    std::vector<double> mm((size_t)3 * nw), dm((size_t)3 * nw);
    return qtip4pf_energy_grad(crd, mm.data(), grad, dm.data(), 1, nw);
}

double qtip4pf_energy_host(int nw, const double *crd)
{
    std::vector<double> mm((size_t)3 * nw);
    return qtip4pf_energy<double>(crd, mm.data(), 1, nw);
}

hipError_t launch_qtip4pf(scp3g_ctx *ctx, int nw, int64_t npts, const void *crd, int64_t ldc,
                          int fp32, double *E)
{
    if (npts <= 0) return hipSuccess;
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    const size_t smem = (size_t)12 * nw * 64 * es;
    if (smem > 64 * 1024) return hipErrorInvalidValue;  // nw > 10 in FP64
    dim3 grid((unsigned)((npts + 63) / 64));
    if (fp32)
        hipLaunchKernelGGL(k_qtip4pf<float>, grid, dim3(64), smem, ctx->stream,
                           (const float *)crd, ldc, nw, npts, E);
    else
        hipLaunchKernelGGL(k_qtip4pf<double>, grid, dim3(64), smem, ctx->stream,
                           (const double *)crd, ldc, nw, npts, E);
    return hipGetLastError();
}

}  // namespace scp3g
