// Normal modes from the mass-weighted Hessian (SCP3.cpp:292-318).
//
// The reference calls LAPACK through Armadillo's eig_sym; here a cyclic
// Jacobi rotation method (Golub & Van Loan, Matrix Computations, 8.5) is used
// for these small (3N <= 192) dense symmetric matrices.  Eigenvectors are
// determined up to sign; each is normalised so that its largest-magnitude
// component is positive, a deterministic convention (LAPACK's signs are
// implementation-defined).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "scp3g_internal.h"

namespace scp3g {

// A (n x n, column-major, symmetric) -> ascending eigenvalues w and
// column-major eigenvectors V.  Returns the number of sweeps used.
int jacobi_eigh(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int r, int c) -> double & { return A[(size_t)c * n + r]; };
    int sweep = 0;
    for (; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                const double v = a(r, c) * a(r, c);
                tot += v;
                if (r != c) off += v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a(p, q);
                if (apq == 0.0) continue;
                const double app = a(p, p), aqq = a(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double &vkp = V[(size_t)p * n + k], &vkq = V[(size_t)q * n + k];
                    const double x = vkp, y = vkq;
                    vkp = c * x - s * y;
                    vkq = s * x + c * y;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<double> ev(n);
    for (int i = 0; i < n; ++i) ev[i] = a(i, i);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ev[x] < ev[y]; });
    w.resize(n);
    std::vector<double> Vs((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        const int src = idx[j];
        w[j] = ev[src];
        int amax = 0;
        for (int k = 1; k < n; ++k)
            if (std::fabs(V[(size_t)src * n + k]) > std::fabs(V[(size_t)src * n + amax])) amax = k;
        const double sg = V[(size_t)src * n + amax] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < n; ++k) Vs[(size_t)j * n + k] = sg * V[(size_t)src * n + k];
    }
    V.swap(Vs);
    return sweep;
}

}  // namespace scp3g

extern "C" int scp3g_normal_modes(int n3, const double *mass, const double *hess, double *omega,
                                  double *MUa, double *omega_all)
{
    if (n3 < 7 || !mass || !hess || !omega || !MUa) return -4;
    std::vector<double> A(hess, hess + (size_t)n3 * n3), w, V;
    // symmetrise (the input file stores one triangle, DiskIO.cpp:103-113)
    for (int c = 0; c < n3; ++c)
        for (int r = 0; r < c; ++r) {
            const double v = 0.5 * (A[(size_t)c * n3 + r] + A[(size_t)r * n3 + c]);
            A[(size_t)c * n3 + r] = A[(size_t)r * n3 + c] = v;
        }
    scp3g::jacobi_eigh(n3, A, w, V);
    if (omega_all)
        for (int i = 0; i < n3; ++i) omega_all[i] = std::sqrt(std::fabs(w[i]));
    const int M = n3 - 6;
    for (int k = 0; k < M; ++k) {
        if (w[6 + k] <= 0.0) return -7;  // imaginary frequency among kept modes
        omega[k] = std::sqrt(w[6 + k]);
        const double alpha = std::sqrt(omega[k]);
        for (int i = 0; i < n3; ++i)
            MUa[(size_t)k * n3 + i] = V[(size_t)(6 + k) * n3 + i] / (std::sqrt(mass[i]) * alpha);
    }
    return 0;
}
