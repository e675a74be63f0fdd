// Lowest eigenpairs of a small dense symmetric matrix on the host: the
// Rayleigh-Ritz problem of the Davidson solver (m <= a few hundred).
//
//   1. Householder reduction to tridiagonal form T = Q^T A Q (Golub & Van
//      Loan, Matrix Computations, 4th ed., Alg. 5.1.1 and 8.3.1), working
//      on the lower triangle only, the reflectors kept in A's strict lower
//      triangle.
//   2. The lowest nvec eigenvalues of T by bisection on Sturm counts
//      (ibid. Sec. 8.4.1; the count recurrence of LAPACK's dlaebz), all
//      nvec intervals advanced together so that their divisions overlap.
//   3. Their eigenvectors of T by inverse iteration with a pivoted
//      tridiagonal LU (ibid. Sec. 8.2.2), re-orthogonalised against the
//      previous vectors of the same cluster of close eigenvalues.
//   4. Back-transformation z = H_0 H_1 ... H_{m-3} z_T.
//
// Costs 4/3 m^3 + O(nvec m^2) flops, against ~3 sweeps x 6 m^3 for cyclic
// Jacobi; at m = 64..128 that is the difference between ~0.1 ms and 2-18 ms
// per Davidson iteration.  Bisection replaces the implicit QR iteration for
// all m eigenvalues (~3x faster at m = 64: the nvec division chains of one
// bisection round are independent, QR's rotations are one long chain).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace scp3g {

namespace {

// The nvec lowest eigenvalues (ascending) of the symmetric tridiagonal with
// diagonal d and off-diagonal e (e2 = e^2), by bisection.  count(x) =
// #{i : q_i < 0} with q_0 = d_0 - x, q_i = d_i - x - e2_{i-1} / q_{i-1} is
// the number of eigenvalues below x (Sylvester's inertia of the LDL^T of
// T - x I); |q_i| <= pivmin becomes -pivmin (LAPACK dlaebz).  Interval j
// keeps count(lo) <= j < count(hi); all intervals of a group advance
// together, and a group stops when every width is at most tol.
void tridiag_lowest(int m, const double *d, const double *e2, int nvec, double gl, double gu,
                    double pivmin, double tol, double *w)
{
    constexpr int L = 16;
    for (int j0 = 0; j0 < nvec; j0 += L) {
        double lo[L], hi[L], x[L], q[L], cnt[L];
        for (int l = 0; l < L; ++l) {
            lo[l] = gl;
            hi[l] = gu;
        }
        const int nl = std::min(L, nvec - j0);
        for (int round = 0; round < 128; ++round) {
            for (int l = 0; l < L; ++l) {
                x[l] = 0.5 * (lo[l] + hi[l]);
                double t = d[0] - x[l];
                t = std::fabs(t) <= pivmin ? -pivmin : t;
                q[l] = t;
                cnt[l] = t < 0.0 ? 1.0 : 0.0;
            }
            for (int i = 1; i < m; ++i) {
                const double di = d[i], ei = e2[i - 1];
                for (int l = 0; l < L; ++l) {
                    double t = (di - x[l]) - ei / q[l];
                    t = std::fabs(t) <= pivmin ? -pivmin : t;
                    q[l] = t;
                    cnt[l] += t < 0.0 ? 1.0 : 0.0;
                }
            }
            bool done = true;
            for (int l = 0; l < nl; ++l) {
                if (cnt[l] > (double)(j0 + l))
                    hi[l] = x[l];
                else
                    lo[l] = x[l];
                if (hi[l] - lo[l] > tol) done = false;
            }
            if (done) break;
        }
        for (int l = 0; l < nl; ++l) w[j0 + l] = 0.5 * (lo[l] + hi[l]);
    }
}

// Pivoted LU of the tridiagonal T - lam I (Gaussian elimination with
// partial pivoting, top to bottom; a row swap puts fill-in on the second
// superdiagonal).  U is kept as inverse diagonal ui0, first and second
// superdiagonals u1, u2; the eliminations as multipliers and swap flags.
// Pivots below `tiny` in magnitude are replaced by +-tiny.
struct TriLU {
    std::vector<double> ui0, u1, u2, mult;
    std::vector<char> swp;
    explicit TriLU(int m) : ui0(m), u1(m), u2(m), mult(m), swp(m) {}

    void factor(int m, const double *d, const double *e, double lam, double tiny)
    {
        auto inv = [tiny](double p) {
            if (std::fabs(p) < tiny) p = p < 0 ? -tiny : tiny;
            return 1.0 / p;
        };
        double a = d[0] - lam;           // row i, column i
        double c = m > 1 ? e[0] : 0.0;   // row i, column i+1
        double c2 = 0.0;                 // row i, column i+2
        for (int i = 0; i < m - 1; ++i) {
            const double sub = e[i], nd = d[i + 1] - lam, nc = i + 1 < m - 1 ? e[i + 1] : 0.0;
            if (std::fabs(sub) > std::fabs(a)) {
                // pivot on row i+1; the old row i, reduced, moves down
                const double f = a / sub;
                swp[i] = 1;
                mult[i] = f;
                ui0[i] = 1.0 / sub;
                u1[i] = nd;
                u2[i] = nc;
                a = c - f * nd;
                c = c2 - f * nc;
            } else {
                const double ia = inv(a);
                const double f = sub * ia;
                swp[i] = 0;
                mult[i] = f;
                ui0[i] = ia;
                u1[i] = c;
                u2[i] = c2;
                a = nd - f * c;
                c = nc - f * c2;
            }
            c2 = 0.0;
        }
        ui0[m - 1] = inv(a);
        u1[m - 1] = u2[m - 1] = 0.0;
        if (m > 1) u2[m - 2] = 0.0;
    }
    void solve(int m, double *b) const
    {
        for (int i = 0; i < m - 1; ++i) {
            if (swp[i]) std::swap(b[i], b[i + 1]);
            b[i + 1] -= mult[i] * b[i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double s = b[i];
            if (i + 1 < m) s -= u1[i] * b[i + 1];
            if (i + 2 < m) s -= u2[i] * b[i + 2];
            b[i] = s * ui0[i];
        }
    }
};

}  // namespace

// A: m x m column-major symmetric (both triangles), overwritten.  w gets the
// nvec lowest eigenvalues ascending, Z (m x nvec, column-major) their
// orthonormal eigenvectors.  Returns 0.
int small_eigh(int m, std::vector<double> &A, int nvec, std::vector<double> &w, std::vector<double> &Z)
{
    nvec = std::min(nvec, m);
    w.assign(nvec, 0.0);
    Z.assign((size_t)m * nvec, 0.0);
    if (m <= 0 || nvec <= 0) return 0;
    auto a = [&](int r, int c) -> double & { return A[(size_t)c * m + r]; };
    std::vector<double> d(m), e(std::max(m - 1, 1), 0.0), beta(std::max(m - 2, 1), 0.0);
    std::vector<double> v(m), p(m), wv(m);
    // 1. tridiagonalisation: column k's reflector zeroes a(k+2.., k)
    for (int k = 0; k + 2 < m; ++k) {
        const int r = m - k - 1;  // trailing size
        double sigma = 0.0;
        for (int i = k + 2; i < m; ++i) sigma += a(i, k) * a(i, k);
        const double x0 = a(k + 1, k);
        if (sigma == 0.0) {
            beta[k] = 0.0;
            e[k] = x0;
            continue;
        }
        const double mu = std::sqrt(x0 * x0 + sigma);
        const double v0 = x0 <= 0.0 ? x0 - mu : -sigma / (x0 + mu);
        beta[k] = 2.0 * v0 * v0 / (sigma + v0 * v0);
        v[0] = 1.0;
        for (int i = k + 2; i < m; ++i) {
            a(i, k) /= v0;
            v[i - k - 1] = a(i, k);
        }
        e[k] = mu;
        // A22 <- H A22 H with p = beta A22 v, w = p - (beta/2)(p.v) v, on
        // the lower triangle of A22 (columns k+1+j, rows j.. of each)
        for (int i = 0; i < r; ++i) p[i] = 0.0;
        for (int j = 0; j < r; ++j) {
            const double vj = v[j];
            const double *col = &A[(size_t)(k + 1 + j) * m + k + 1];
            double dot = col[j] * vj;
            for (int i = j + 1; i < r; ++i) {
                dot += col[i] * v[i];
                p[i] += col[i] * vj;
            }
            p[j] += dot;
        }
        double pv = 0.0;
        for (int i = 0; i < r; ++i) {
            p[i] *= beta[k];
            pv += p[i] * v[i];
        }
        const double hk = 0.5 * beta[k] * pv;
        for (int i = 0; i < r; ++i) wv[i] = p[i] - hk * v[i];
        for (int j = 0; j < r; ++j) {
            double *col = &A[(size_t)(k + 1 + j) * m + k + 1];
            const double vj = v[j], wj = wv[j];
            for (int i = j; i < r; ++i) col[i] -= v[i] * wj + wv[i] * vj;
        }
    }
    for (int k = 0; k < m; ++k) d[k] = a(k, k);
    if (m >= 2) e[m - 2] = a(m - 1, m - 2);
    // 2. eigenvalues of T
    double tnorm = 0.0;
    for (int i = 0; i < m; ++i) {
        double s = std::fabs(d[i]);
        if (i > 0) s += std::fabs(e[i - 1]);
        if (i + 1 < m) s += std::fabs(e[i]);
        tnorm = std::max(tnorm, s);
    }
    if (tnorm == 0.0) tnorm = 1.0;
    const double eps = 2.220446049250313e-16;
    {
        // Gershgorin interval of T, and the bisection tolerance: 2 eps ||T||,
        // the accuracy of any backward-stable method
        double gl = d[0], gu = d[0], emax2 = 0.0;
        std::vector<double> e2((size_t)std::max(m - 1, 1), 0.0);
        for (int i = 0; i < m; ++i) {
            double r = 0.0;
            if (i > 0) r += std::fabs(e[i - 1]);
            if (i + 1 < m) {
                r += std::fabs(e[i]);
                e2[i] = e[i] * e[i];
                emax2 = std::max(emax2, e2[i]);
            }
            gl = std::min(gl, d[i] - r);
            gu = std::max(gu, d[i] + r);
        }
        const double pivmin = 2.2250738585072014e-308 * std::max(1.0, emax2);
        const double pad = 2.0 * eps * tnorm + pivmin;
        tridiag_lowest(m, d.data(), e2.data(), nvec, gl - pad, gu + pad, pivmin, 2.0 * eps * tnorm, w.data());
    }
    // 3. inverse iteration on T: one factorisation per vector, two solves
    // (three inside a cluster), each followed by orthogonalisation against
    // the cluster's earlier vectors
    TriLU lu(m);
    std::vector<double> b(m);
    const double tiny = eps * tnorm;
    const double ortol = 1e-3 * tnorm;
    int cluster0 = 0;
    double lam_prev = 0.0;
    uint64_t seed = 0x9E3779B97F4A7C15ull;
    for (int j = 0; j < nvec; ++j) {
        if (j > 0 && w[j] - w[j - 1] > ortol) cluster0 = j;
        // shift eigenvalues that coincide with their predecessor's shift a
        // little so that the solves differ
        double lam = w[j];
        if (j > cluster0 && lam - lam_prev < 10 * eps * tnorm) lam = lam_prev + 10 * eps * tnorm;
        lam_prev = lam;
        lu.factor(m, d.data(), e.data(), lam, tiny);
        for (int i = 0; i < m; ++i) {
            seed = seed * 6364136223846793005ull + 1442695040888963407ull;
            b[i] = 0.5 + (double)(seed >> 11) * (1.0 / 9007199254740992.0);
        }
        const int nsolve = j > cluster0 ? 3 : 2;
        for (int it = 0; it < nsolve; ++it) {
            double nb = 0.0;
            for (int i = 0; i < m; ++i) nb = std::max(nb, std::fabs(b[i]));
            const double inb = 1.0 / nb;
            for (int i = 0; i < m; ++i) b[i] *= inb;
            lu.solve(m, b.data());
            for (int q = cluster0; q < j; ++q) {
                const double *zq = &Z[(size_t)q * m];
                double s = 0.0;
                for (int i = 0; i < m; ++i) s += zq[i] * b[i];
                for (int i = 0; i < m; ++i) b[i] -= s * zq[i];
            }
        }
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) nrm += b[i] * b[i];
        nrm = 1.0 / std::sqrt(nrm);
        double *zj = &Z[(size_t)j * m];
        for (int i = 0; i < m; ++i) zj[i] = b[i] * nrm;
    }
    // 4. back-transformation, last reflector first, on a row-major copy so
    // that the nvec dot products of a reflector run side by side
    std::vector<double> Zr((size_t)m * nvec), sj(nvec);
    for (int j = 0; j < nvec; ++j)
        for (int i = 0; i < m; ++i) Zr[(size_t)i * nvec + j] = Z[(size_t)j * m + i];
    for (int k = m - 3; k >= 0; --k) {
        if (beta[k] == 0.0) continue;
        const double *z1 = &Zr[(size_t)(k + 1) * nvec];
        for (int j = 0; j < nvec; ++j) sj[j] = z1[j];
        for (int i = k + 2; i < m; ++i) {
            const double vi = a(i, k);
            const double *zi = &Zr[(size_t)i * nvec];
            for (int j = 0; j < nvec; ++j) sj[j] += vi * zi[j];
        }
        for (int j = 0; j < nvec; ++j) sj[j] *= beta[k];
        double *zk = &Zr[(size_t)(k + 1) * nvec];
        for (int j = 0; j < nvec; ++j) zk[j] -= sj[j];
        for (int i = k + 2; i < m; ++i) {
            const double vi = a(i, k);
            double *zi = &Zr[(size_t)i * nvec];
            for (int j = 0; j < nvec; ++j) zi[j] -= vi * sj[j];
        }
    }
    for (int j = 0; j < nvec; ++j)
        for (int i = 0; i < m; ++i) Z[(size_t)j * m + i] = Zr[(size_t)i * nvec + j];
    return 0;
}

}  // namespace scp3g
