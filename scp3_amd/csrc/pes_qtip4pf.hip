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
