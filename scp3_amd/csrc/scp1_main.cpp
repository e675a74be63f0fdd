// SCP1: self-consistent phonons for a water cluster (SCP1_main.cpp,
// SCP1.cpp), the Gaussian averages on the GPU through libscp3g.
//
//   SCP1 [-N log2(NSobol)] [-p PES] [-L max_iterations] [-c cutoff_cm-1]
//        [-T temperature_K] input
//
// Same options and outputs as the reference: omega_scp1.dat (one line of
// frequencies per iteration), free_energy.dat (F per iteration, Hartree),
// coord.xyz (the current geometry and effective Hessian in the input's
// format), "iter: n" per iteration and the final omega, <U>, <UX>, F and F0
// on stdout.  -S and -r are accepted and, as in the reference (whose SCP1
// object chooses its own skip and never reads the file), have no effect.
// -T is an addition: the reference's main always passes kT = 0
// (SCP1_main.cpp:132) although SCP1::operator() takes any temperature.
// The averages replace SCP1::average (SCP1.cpp:33-86) with
// scp3g_scp1_average: the whole Sobol range on one device instead of an MPI
// split plus Allreduce.  The iteration (SCP1.cpp:89-190) stays on the host.
#include <getopt.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/scp3g.h"
#include "cli_common.h"

namespace {

constexpr double kAutokcalpmol = 627.509474277194;  // Constants.h:21
constexpr double kAutoK = 315775.0428639057;         // Constants.cpp:14

struct Options {
    int64_t nsobol = int64_t(1) << 20;  // SCP1_main.cpp:43
    std::string pes = "qtip4pf";
    int max_iterations = 200;
    double cutoff = -1.0;  // Hartree; <= 0 keeps 50 cm^-1 (SCP1.h:53)
    double kT = 0.0;       // Hartree
    std::string input;
};

void usage_exit(const char *msg)
{
    std::fprintf(stderr, "%s\n", msg);
    std::exit(EXIT_FAILURE);
}

// process_options (SCP1_main.cpp:59-105)
Options parse(int argc, char **argv)
{
    static struct option longopts[] = {{"NSobol", required_argument, nullptr, 'N'},
                                       {"skip", required_argument, nullptr, 'S'},
                                       {"rng-file", required_argument, nullptr, 'r'},
                                       {"PES", required_argument, nullptr, 'p'},
                                       {"max-itertions", required_argument, nullptr, 'L'},
                                       {"cutoff", required_argument, nullptr, 'c'},
                                       {"temperature", required_argument, nullptr, 'T'},
                                       {nullptr, 0, nullptr, 0}};
    Options o;
    int ch;
    while ((ch = getopt_long(argc, argv, "S:N:r:p:L:c:T:", longopts, nullptr)) != -1) {
        switch (ch) {
        case 'N': {
            const int p2 = std::atoi(optarg);
            // the reference stops at 32 and shifts an int; 64-bit here
            if (p2 > 40 || p2 < 0) usage_exit("-N <N>: <N>is too large, NSobol = 2^<N>");
            o.nsobol = int64_t(1) << p2;
            break;
        }
        case 'S':
            if (std::atoi(optarg) > 32) usage_exit("-S <N>: <N>is too large, sobol_skip = 2^<N>");
            break;
        case 'r':
            break;
        case 'p':
            o.pes = optarg;
            break;
        case 'L':
            o.max_iterations = std::atoi(optarg);
            break;
        case 'c':
            o.cutoff = std::atof(optarg) / cli::autocm;
            break;
        case 'T': {
            const double T = std::atof(optarg);
            if (!(T >= 0.0)) usage_exit("-T <K>: the temperature must be >= 0");
            o.kT = T / kAutoK;
            break;
        }
        default:
            std::fprintf(stderr, "Unknown option: %d\n", ch);
            std::exit(EXIT_FAILURE);
        }
    }
    if (optind >= argc) usage_exit("No input file supplied");
    o.input = argv[optind];
    return o;
}

// Symmetric eigenproblem of the upper triangle of A (n x n column-major,
// what Armadillo's eig_sym reads) by cyclic Jacobi rotations: ascending
// eigenvalues w, orthonormal eigenvectors V (columns), each column signed so
// that its largest-magnitude component is positive.
void eig_sym(int n, const std::vector<double> &A0, std::vector<double> &w, std::vector<double> &V)
{
    std::vector<double> A(A0);
    for (int c = 0; c < n; ++c)
        for (int r = c + 1; r < n; ++r) A[(size_t)c * n + r] = A[(size_t)r * n + c];
    std::vector<double> Z((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    auto at = [&](int r, int c) -> double & { return A[(size_t)c * n + r]; };
    double scale = 0.0;
    for (double v : A) scale += v * v;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int c = 1; c < n; ++c)
            for (int r = 0; r < c; ++r) off += at(r, c) * at(r, c);
        if (off <= 1e-32 * scale || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = at(p, q);
                if (apq == 0.0) continue;
                // A' = J^T A J zeroes (p, q) for cot(2 phi) = (a_qq - a_pp) / (2 a_pq)
                const double theta = (at(q, q) - at(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double kp = at(k, p), kq = at(k, q);
                    at(k, p) = c * kp - s * kq;
                    at(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double pk = at(p, k), qk = at(q, k);
                    at(p, k) = c * pk - s * qk;
                    at(q, k) = s * pk + c * qk;
                }
                at(p, q) = at(q, p) = 0.0;
                for (int k = 0; k < n; ++k) {
                    double &zp = Z[(size_t)p * n + k], &zq = Z[(size_t)q * n + k];
                    const double a = zp, b = zq;
                    zp = c * a - s * b;
                    zq = s * a + c * b;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return at(a, a) < at(b, b); });
    w.resize(n);
    V.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        const int o = order[j];
        w[j] = at(o, o);
        int imax = 0;
        for (int k = 1; k < n; ++k)
            if (std::fabs(Z[(size_t)o * n + k]) > std::fabs(Z[(size_t)o * n + imax])) imax = k;
        const double sg = Z[(size_t)o * n + imax] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < n; ++k) V[(size_t)j * n + k] = sg * Z[(size_t)o * n + k];
    }
}

// transrotBasis(q, mass) (Hessian.cpp:89-120): n x 6 column-major
std::vector<double> transrot_basis(const std::vector<double> &q, const std::vector<double> &mass)
{
    const int n = (int)q.size();
    std::vector<double> U((size_t)n * 6, 0.0);
    auto u = [&](int r, int c) -> double & { return U[(size_t)c * n + r]; };
    double total = 0.0;
    for (double m : mass) total += m;
    total /= 3.0;
    for (int i = 0; i < n; i += 3) {
        const double s0 = q[i] * std::sqrt(mass[i]), s1 = q[i + 1] * std::sqrt(mass[i + 1]),
                     s2 = q[i + 2] * std::sqrt(mass[i + 2]);
        u(i, 0) = std::sqrt(mass[i] / total);
        u(i + 1, 1) = std::sqrt(mass[i + 1] / total);
        u(i + 2, 2) = std::sqrt(mass[i + 2] / total);
        u(i + 1, 3) = -s2;
        u(i + 2, 3) = s1;
        u(i, 4) = s2;
        u(i + 2, 4) = -s0;
        u(i, 5) = -s1;
        u(i + 1, 5) = s0;
    }
    for (int i = 3; i < 6; ++i) {
        for (int j = 0; j < i; ++j) {
            double dot = 0.0;
            for (int k = 0; k < n; ++k) dot += u(k, i) * u(k, j);
            for (int k = 0; k < n; ++k) u(k, i) -= u(k, j) * dot;
        }
        double nrm = 0.0;
        for (int k = 0; k < n; ++k) nrm += u(k, i) * u(k, i);
        nrm = std::sqrt(nrm);
        for (int k = 0; k < n; ++k) u(k, i) /= nrm;
    }
    return U;
}

// regtransrot (Hessian.cpp:123-130): H = (1 - U U^T) H (1 - U U^T)
void regtransrot(const std::vector<double> &U, std::vector<double> &H, int n)
{
    std::vector<double> HU((size_t)n * 6, 0.0), HP(H);
    for (int c = 0; c < 6; ++c)
        for (int k = 0; k < n; ++k) {
            const double ukc = U[(size_t)c * n + k];
            for (int r = 0; r < n; ++r) HU[(size_t)c * n + r] += H[(size_t)k * n + r] * ukc;
        }
    for (int j = 0; j < n; ++j)
        for (int c = 0; c < 6; ++c) {
            const double ujc = U[(size_t)c * n + j];
            for (int r = 0; r < n; ++r) HP[(size_t)j * n + r] -= HU[(size_t)c * n + r] * ujc;
        }
    std::vector<double> UtHP((size_t)6 * n, 0.0);  // 6 x n
    for (int j = 0; j < n; ++j)
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += U[(size_t)c * n + k] * HP[(size_t)j * n + k];
            UtHP[(size_t)j * 6 + c] = s;
        }
    for (int j = 0; j < n; ++j)
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int c = 0; c < 6; ++c) s += U[(size_t)c * n + r] * UtHP[(size_t)j * 6 + c];
            H[(size_t)j * n + r] = HP[(size_t)j * n + r] - s;
        }
}

// disp_disp_corr_diag (SCP1.cpp:18-31)
std::vector<double> disp_disp(const std::vector<double> &omega, double kT, double cutoff)
{
    std::vector<double> d(omega.size());
    for (size_t i = 0; i < omega.size(); ++i) {
        d[i] = kT == 0.0 ? 0.5 / omega[i] : 0.5 / (omega[i] * std::tanh(0.5 / kT * omega[i]));
        if (omega[i] < cutoff) d[i] = 0.0;
    }
    return d;
}

// save_for_vladimir (DiskIO.cpp:119-140)
void save_vladimir(const char *name, double F, const std::vector<double> &q, const std::vector<double> &H)
{
    FILE *f = std::fopen(name, "w");
    if (!f) return;
    const int n = (int)q.size();
    std::fprintf(f, "%d\nFree energy = %.17g\n", n / 3, F);
    static const char *sp[3] = {"O", "H", "H"};
    for (int a = 0; a < n / 3; ++a)
        std::fprintf(f, "%s %.17g %.17g %.17g\n", sp[a % 3], q[3 * a] * cli::bohr, q[3 * a + 1] * cli::bohr,
                     q[3 * a + 2] * cli::bohr);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) std::fprintf(f, "%.17g\n", H[(size_t)i * n + j]);
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv)
{
    Options o = parse(argc, argv);
    cli::Input in;
    std::string err;
    if (!cli::load_vladimir(o.input, in, err) || !cli::ohhohh(in, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return EXIT_FAILURE;
    }
    const int n = in.n3;
    if (n % 9) {
        std::fprintf(stderr, "the input is not a water cluster (O H H)\n");
        return EXIT_FAILURE;
    }
    const int nw = n / 9;
    int pes;
    if (o.pes == "qtip4pf")
        pes = SCP3G_PES_QTIP4PF;
    else if (o.pes == "ttm3f")
        pes = SCP3G_PES_TTM3F;
    else {
        std::fprintf(stderr, "PES %s has no GPU implementation (qtip4pf, ttm3f)\n", o.pes.c_str());
        return EXIT_FAILURE;
    }

    scp3g_ctx *g = nullptr;
    if (scp3g_create(&g, 0)) {
        std::fprintf(stderr, "no GPU: %s\n", scp3g_last_error(g));
        return EXIT_FAILURE;
    }
    auto check = [&](int rc) {
        if (rc) {
            std::fprintf(stderr, "scp3g: %s\n", scp3g_last_error(g));
            scp3g_destroy(g);
            std::exit(EXIT_FAILURE);
        }
    };

    // SCP1 constructor (SCP1.h:46-66)
    const bool append = !in.H.empty();
    const double cutoff = o.cutoff > 0 ? o.cutoff : 50.0 / cli::autocm;
    const double qeps = 2.0, kT = o.kT;  // SCP1.h:55; SCP1_main.cpp:132 passes 0
    const int64_t skip = int64_t(1) << ((int)std::floor(std::log2((double)o.nsobol)) + 1);
    FILE *omega_out = std::fopen("omega_scp1.dat", append ? "a" : "w");
    FILE *free_energy_out = std::fopen("free_energy.dat", append ? "a" : "w");
    if (!omega_out || !free_energy_out) {
        std::fprintf(stderr, "cannot open omega_scp1.dat / free_energy.dat\n");
        return EXIT_FAILURE;
    }

    std::vector<double> q = in.x0, mass = in.mass, Ks = in.H;
    if (Ks.empty()) {  // SCP1.cpp:95-102
        Ks.assign((size_t)n * n, 0.0);
        check(scp3g_hessian(g, pes, nw, q.data(), 1e-3, Ks.data()));
        std::vector<double> isq(n);  // massScaleHessian (Hessian.cpp:45-55)
        for (int i = 0; i < n; ++i) isq[i] = 1.0 / std::sqrt(mass[i]);
        for (int i = 0; i < n; ++i)
            for (int r = 0; r < n; ++r) Ks[(size_t)i * n + r] *= isq[r] * isq[i];
        regtransrot(transrot_basis(q, mass), Ks, n);
    }
    std::vector<double> w, U, omega(n), d, iU((size_t)n * n), UX(n), UXq((size_t)n * n), KD((size_t)n * n);
    auto spectrum = [&]() {
        eig_sym(n, Ks, w, U);
        for (int i = 0; i < n; ++i) omega[i] = std::sqrt(std::fabs(w[i]));
        d = disp_disp(omega, kT, cutoff);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) iU[(size_t)c * n + r] = U[(size_t)c * n + r] / std::sqrt(mass[r]);
    };
    spectrum();  // SCP1.cpp:105-110

    double F = 0.0, Upot = 0.0;
    for (int niter = 0; niter < o.max_iterations; ++niter) {
        check(scp3g_scp1_average(g, pes, nw, q.data(), d.data(), iU.data(), skip, o.nsobol, &Upot, UX.data(),
                                 UXq.data()));
        for (int i = 0; i < n; ++i)  // SCP1.cpp:120-124
            for (int j = 0; j < n; ++j) UXq[(size_t)i * n + j] *= std::sqrt(mass[i] / mass[j]);
        for (int c = 0; c < n; ++c) {  // KD = UX_q U, columns / d (:126-135)
            for (int r = 0; r < n; ++r) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += UXq[(size_t)k * n + r] * U[(size_t)c * n + k];
                KD[(size_t)c * n + r] = d[c] != 0.0 ? s / d[c] : 0.0;
            }
        }
        for (int c = 0; c < n; ++c)  // Ks = KD U^T (:136)
            for (int r = 0; r < n; ++r) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += KD[(size_t)k * n + r] * U[(size_t)k * n + c];
                Ks[(size_t)c * n + r] = s;
            }
        for (int c = 0; c < n; ++c)  // (Ks + Ks^T) / 2 (:138-145)
            for (int r = 0; r < c; ++r) {
                const double s = 0.5 * (Ks[(size_t)c * n + r] + Ks[(size_t)r * n + c]);
                Ks[(size_t)c * n + r] = Ks[(size_t)r * n + c] = s;
            }
        regtransrot(transrot_basis(q, mass), Ks, n);  // :147
        spectrum();                                     // :149-158
        std::vector<double> t(n, 0.0);                  // q -= qeps iU (d o (iU^T UX)) (:160)
        for (int c = 0; c < n; ++c) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += iU[(size_t)c * n + k] * UX[k];
            t[c] = d[c] * s;
        }
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int c = 0; c < n; ++c) s += iU[(size_t)c * n + r] * t[c];
            q[r] -= qeps * s;
        }
        F = Upot;  // :165-173
        for (int i = 6; i < n; ++i)
            F += kT > 0.0 ? kT * std::log(2.0 * std::sinh(0.5 / kT * omega[i])) -
                                0.25 * omega[i] / std::tanh(0.5 / kT * omega[i])
                          : 0.25 * omega[i];
        for (int i = 0; i < n; ++i) std::fprintf(omega_out, i ? " %.10e" : "%.10e", omega[i]);
        std::fprintf(omega_out, "\n");
        std::fprintf(free_energy_out, "%.17g\n", F);
        std::fflush(omega_out);
        std::fflush(free_energy_out);
        save_vladimir("coord.xyz", F, q, Ks);
        std::printf("iter: %d\n", niter);
        std::fflush(stdout);
    }
    std::printf("omega =");  // SCP1.cpp:183-188
    for (int i = 0; i < n; ++i) std::printf(" %.6f", omega[i] * cli::autocm);
    std::printf("\n<U> = %.17g\n<UX> =", Upot);
    for (int i = 0; i < n; ++i) std::printf(" %.6e", UX[i]);
    std::printf("\nF = %.17g kcal/mol\n", F * kAutokcalpmol);
    std::printf("F0 = %.17g\n", F);  // SCP1_main.cpp:134
    std::fclose(omega_out);
    std::fclose(free_energy_out);
    scp3g_destroy(g);
    return 0;
}
