// SCP3 command-line driver over libscp3g (both estimators on the GPU).
//
// Options and outputs follow SCP3.cpp:33-151,274-323:
//   SCP3 [-N p] [-S p] [-r file] [-2 n] [-3 n] [-p pes] [-m list] [-G] [-B w] input
//   -N/--NSobol p       2^p Sobol points (default 2^20)
//   -S/--skip p         Sobol skip 2^p (default 2^30)
//   -r/--rng-file f     standard normals, nmodes values per point, instead of Sobol
//   -2/--doubles n      modes with double excitations (default 3N-6)
//   -3/--triples n      modes with triple excitations (default 0)
//   -p/--PES name       qtip4pf (default) or ttm3f
//   -m/--select-modes   "a-b,c" list of modes (default all)
//   -G/--no-gradient    the B^T diag(V) B estimator; without it the gradient
//                       (Stein-identity) estimator, on the GPU in FP64
//   -B/--block-width w  points per block between checkpoints (default 1024)
// Additions: --precision fp64|mixed (default fp64), --levels k (lowest k
// levels printed at the end, default 10, 0 = none), --device d, --cpu-pes
// (the potential is evaluated on the host through h2o::PES, one instance
// per thread, as the reference does; sampling, basis and SYRK stay on the
// GPU: the path for potentials without a device kernel, such as WHBB; with
// the PES's gradient overload without -G), --stein (the gradient estimator with every element from the Stein
// identity instead of the reference's placement of five element families),
// --scramble SEED (affine digit scrambling of the Sobol points, in place of
// the reference's external scrambled -r files), --cv-true p (with -G: the
// learned control variate; an MLP surrogate of the potential is trained on
// the sampler's own (q, V) pairs, the correction term runs over 2^p
// true-potential points first, and the -N points are surrogate points,
// H = S / P + C / 2^p at every checkpoint; --cv-pairs p, 2^p training pairs,
// default 18; --cv-steps s, Adam steps, default 20000; --cv-hidden h, width
// of the three hidden layers, default 256).
// Files: omega0.dat, E0_sd.dat (every 2^14 points), E0_t.dat (empty),
// sM_%07d.h5 (every 2^17 points), levels.dat (this driver's final levels).
#include <getopt.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <thread>
#include <vector>

#include "../../include/PES.h"
#include "../../include/scp3g.h"
#include "cli_common.h"

namespace {

struct Opts {
    int64_t nsobol = (int64_t)1 << 20;
    int64_t skip = (int64_t)1 << 30;
    std::string rng_file, pes = "qtip4pf", input, modes_desc;
    int n2 = 0, n3 = 0, bw = 1024, levels = 10, device = 0;
    bool gradient = true, have_modes = false, cpu_pes = false, stein = false;
    bool scrambled = false;
    unsigned long long scramble = 0;
    int precision = SCP3G_FP64;
    int cv_true = -1, cv_pairs = 18, cv_steps = 20000, cv_hidden = 256;
};

// parse_mode_description (SCP3.cpp:63-89): "a-b,c" -> a..b, c
std::vector<int> parse_modes(const std::string &desc)
{
    std::vector<int> out;
    std::istringstream iss(desc);
    while (iss.good()) {
        if (iss.peek() == ',') iss.get();
        int start, stop;
        if (!(iss >> start)) break;
        stop = start;
        if (iss.peek() == '-') {
            iss.get();
            iss >> stop;
        }
        for (; start <= stop; ++start) out.push_back(start);
    }
    return out;
}

[[noreturn]] void die(const std::string &msg)
{
    std::cerr << msg << "\n";
    std::exit(EXIT_FAILURE);
}

Opts parse(int argc, char **argv)
{
    static option lo[] = {{"NSobol", required_argument, nullptr, 'N'},
                          {"skip", required_argument, nullptr, 'S'},
                          {"rng-file", required_argument, nullptr, 'r'},
                          {"doubles", required_argument, nullptr, '2'},
                          {"triples", required_argument, nullptr, '3'},
                          {"PES", required_argument, nullptr, 'p'},
                          {"select-modes", required_argument, nullptr, 'm'},
                          {"no-gradient", no_argument, nullptr, 'G'},
                          {"block-width", required_argument, nullptr, 'B'},
                          {"precision", required_argument, nullptr, 1000},
                          {"levels", required_argument, nullptr, 1001},
                          {"device", required_argument, nullptr, 1002},
                          {"cpu-pes", no_argument, nullptr, 1003},
                          {"stein", no_argument, nullptr, 1004},
                          {"scramble", required_argument, nullptr, 1005},
                          {"cv-true", required_argument, nullptr, 1006},
                          {"cv-pairs", required_argument, nullptr, 1007},
                          {"cv-steps", required_argument, nullptr, 1008},
                          {"cv-hidden", required_argument, nullptr, 1009},
                          {nullptr, 0, nullptr, 0}};
    Opts o;
    int ch;
    while ((ch = getopt_long(argc, argv, "S:N:2:3:r:p:m:GB:", lo, nullptr)) != -1) {
        int p2;
        switch (ch) {
        case 'N':
            p2 = atoi(optarg);
            // the reference stops at 32 and overflows an int shift at 31, 32
            if (p2 < 0 || p2 > 40) die("-N <N>: <N>is too large, NSobol = 2^<N>");
            o.nsobol = (int64_t)1 << p2;
            break;
        case 'S':
            p2 = atoi(optarg);
            if (p2 < 0 || p2 > 50) die("-S <N>: <N>is too large, sobol_skip = 2^<N>");
            o.skip = (int64_t)1 << p2;
            break;
        case '2': o.n2 = atoi(optarg); break;
        case '3': o.n3 = atoi(optarg); break;
        case 'r': o.rng_file = optarg; break;
        case 'p': o.pes = optarg; break;
        case 'm':
            o.modes_desc = optarg;
            o.have_modes = true;
            break;
        case 'G': o.gradient = false; break;
        case 'B': o.bw = atoi(optarg); break;
        case 1000:
            if (!strcmp(optarg, "fp64")) o.precision = SCP3G_FP64;
            else if (!strcmp(optarg, "mixed")) o.precision = SCP3G_MIXED;
            else die("--precision must be fp64 or mixed");
            break;
        case 1001: o.levels = atoi(optarg); break;
        case 1002: o.device = atoi(optarg); break;
        case 1003: o.cpu_pes = true; break;
        case 1004: o.stein = true; break;
        case 1005:
            o.scrambled = true;
            o.scramble = strtoull(optarg, nullptr, 0);
            break;
        case 1006:
            o.cv_true = atoi(optarg);
            if (o.cv_true < 0 || o.cv_true > 40) die("--cv-true <p>: 2^p true-potential points, p in 0..40");
            break;
        case 1007:
            o.cv_pairs = atoi(optarg);
            if (o.cv_pairs < 0 || o.cv_pairs > 30) die("--cv-pairs <p>: 2^p training pairs, p in 0..30");
            break;
        case 1008: o.cv_steps = atoi(optarg); break;
        case 1009: o.cv_hidden = atoi(optarg); break;
        default: die("Unknown option: " + std::to_string(ch));
        }
    }
    if (optind >= argc) die("No input file supplied");
    o.input = argv[optind];
    if (o.bw <= 0) die("-B <w>: block width must be positive");
    return o;
}

void check(scp3g_ctx *ctx, int rc, const char *what)
{
    if (rc != 0) die(std::string(what) + ": " + scp3g_last_error(ctx));
}

}  // namespace

int main(int argc, char **argv)
{
    Opts o = parse(argc, argv);
    if (o.pes == "whbb" || o.pes == "hbb2-pol") die("Support for Bowman's WHBB was not included.");
    if (o.cpu_pes) {
        // h2o::PESFromString decides which host potentials exist
        h2o::PES *probe = h2o::PESFromString(o.pes);
        if (!probe) die("Unknown H2O potential: " + o.pes);
        delete probe;
        if (!o.rng_file.empty()) die("--cpu-pes takes Sobol points only (no -r)");
    } else if (o.pes != "qtip4pf" && o.pes != "ttm3f") {
        die("Unknown H2O potential: " + o.pes);
    }
    if (o.cv_true >= 0 && (o.gradient || o.cpu_pes || !o.rng_file.empty()))
        die("--cv-true needs -G, the device potential and Sobol points (no --cpu-pes, no -r)");

    cli::Input in;
    std::string err;
    if (!cli::load_vladimir(o.input, in, err)) die(err);
    if (!cli::ohhohh(in, err)) die(err);
    if (in.H.empty()) die("input holds no Hessian");
    const int n3c = in.n3, N = n3c / 3, nm0 = n3c - 6;
    std::vector<double> omega(nm0), MUa((size_t)n3c * nm0), oall(n3c);
    if (scp3g_normal_modes(n3c, in.mass.data(), in.H.data(), omega.data(), MUa.data(), oall.data()))
        die("normal modes failed (imaginary frequency among the kept modes?)");
    {
        std::ofstream sout("omega0.dat");  // SCP3.cpp:300-302
        sout << std::setprecision(10);
        for (double w : oall) sout << w * cli::autocm << "\n";
    }
    if (o.n2 == 0) o.n2 = 3 * N - 6;  // SCP3.cpp:305
    std::vector<int> modes;
    if (o.have_modes) modes = parse_modes(o.modes_desc);
    else
        for (int k = 0; k < nm0; ++k) modes.push_back(k);
    if (o.n2 > (int)modes.size() || o.n3 > (int)modes.size())
        die("Number of double or triple excitations exceed number of modes.");

    scp3g_ctx *ctx = nullptr;
    if (scp3g_create(&ctx, o.device)) die("no usable GPU");
    check(ctx, scp3g_set_system(ctx, N / 3, in.x0.data(), nm0, omega.data(), MUa.data()), "set_system");
    check(ctx, scp3g_set_basis(ctx, (int)modes.size(), modes.data(), o.n2, o.n3), "set_basis");
    if (o.scrambled) check(ctx, scp3g_set_scramble(ctx, 1, o.scramble), "set_scramble");
    if (!o.cpu_pes) check(ctx, scp3g_set_pes(ctx, o.pes.c_str()), "set_pes");
    check(ctx, scp3g_set_precision(ctx, o.precision), "set_precision");
    if (o.gradient)
        check(ctx, scp3g_set_estimator(ctx, o.stein ? SCP3G_EST_GRADIENT_STEIN : SCP3G_EST_GRADIENT),
              "set_estimator");
    check(ctx, scp3g_reset(ctx), "reset");
    if (o.cv_true >= 0) {
        // the surrogate, trained on pairs from far along the sequence; the
        // correction term over the first 2^cv_true points, held as a fixed
        // term; then the surrogate term over the -N points (the one BF16
        // product in the mixed form, matched in the correction)
        check(ctx, scp3g_mlp_init(ctx, o.cv_hidden, 0), "mlp init");
        check(ctx, scp3g_mlp_sample_data(ctx, (int64_t)1 << 40, (int64_t)1 << o.cv_pairs), "mlp training set");
        check(ctx, scp3g_mlp_train(ctx, o.cv_steps, 4096, 1e-3, nullptr), "mlp training");
        if (o.precision == SCP3G_MIXED) check(ctx, scp3g_set_cv_terms(ctx, 1), "cv terms");
        check(ctx, scp3g_accumulate_correction(ctx, o.skip, (int64_t)1 << o.cv_true), "correction term");
        check(ctx, scp3g_hold_as_offset(ctx), "hold correction term");
    }
    int64_t n = 0, n1 = 0, n2s = 0;
    scp3g_basis_size(ctx, &n);
    scp3g_sub_basis_size(ctx, 1, &n1);
    scp3g_sub_basis_size(ctx, 2, &n2s);
    std::cerr << "SCP3 (GPU, " << (o.gradient ? (o.stein ? "gradient, Stein" : "gradient") : "-G") << "): " << N / 3 << " waters, " << modes.size() << " modes, " << n
              << " states, " << o.nsobol << " points, " << (o.precision ? "mixed" : "fp64")
              << " form, " << o.pes;
    if (o.cv_true >= 0) std::cerr << ", control variate (2^" << o.cv_true << " true-potential points)";
    std::cerr << "\n";

    std::ofstream e0sd("E0_sd.dat"), e0t("E0_t.dat");
    e0sd << std::fixed << std::setprecision(10);
    std::ifstream rng;
    if (!o.rng_file.empty()) {
        rng.open(o.rng_file.c_str());
        if (!rng.good()) die("cannot open " + o.rng_file);
    }
    std::vector<double> Mout;
#ifdef SCP3G_HAVE_HDF5
    // sM files are written by a background thread (FP32 conversion and
    // deflate of up to 1.7 GB per file) while the device keeps accumulating;
    // one file at a time, joined before the next and at the end
    std::thread writer;
    std::vector<double> wbuf;
    std::string werr;
    bool wfail = false;
    auto join_writer = [&]() {
        if (writer.joinable()) writer.join();
        if (wfail) die(werr);
    };
#endif
    // blocks of bw points as SCP3.cpp:187; consecutive blocks without a
    // checkpoint are sent to the device together
    const int64_t log_every = (int64_t)1 << 14, save_every = (int64_t)1 << 17;
    int64_t i = 0;
    while (i < o.nsobol) {
        int64_t span = 0;
        do {
            span += o.bw;
        } while (i + span < o.nsobol && (i + span) % log_every != 0);
        if (o.cpu_pes) {
            // SCP3.cpp:205-225 with the potential on the host
            // (SCP3.cpp:210-217 with the gradient overload without -G)
            const int n3 = (int)(3 * N);
            std::vector<double> crd((size_t)span * n3), E((size_t)span), G(o.gradient ? (size_t)span * n3 : 0);
            check(ctx, scp3g_points_coords(ctx, o.skip + i, span, crd.data()), "coords");
#pragma omp parallel
            {
                h2o::PES *pes = h2o::PESFromString(o.pes);
#pragma omp for schedule(static)
                for (int64_t j = 0; j < span; ++j)
                    E[j] = o.gradient ? (*pes)((size_t)(N / 3), crd.data() + (size_t)j * n3, G.data() + (size_t)j * n3)
                                      : (*pes)((size_t)(N / 3), crd.data() + (size_t)j * n3);
                delete pes;
            }
            if (o.gradient)
                check(ctx, scp3g_accumulate_gradients(ctx, o.skip + i, span, E.data(), G.data()), "accumulate");
            else
                check(ctx, scp3g_accumulate_energies(ctx, o.skip + i, span, E.data()), "accumulate");
        } else if (rng.is_open()) {
            std::vector<double> nr((size_t)span * nm0);
            for (double &v : nr) rng >> v;
            if (!rng) die("rng-file ended early");
            check(ctx, scp3g_accumulate_normals(ctx, nr.data(), span), "accumulate");
        } else if (o.cv_true >= 0) {
            check(ctx, scp3g_accumulate_surrogate(ctx, o.skip + i, span), "accumulate");
        } else {
            check(ctx, scp3g_accumulate(ctx, o.skip + i, span), "accumulate");
        }
        i += span;
        if (i % log_every == 0) {
            double e1, e2;
            check(ctx, scp3g_lowest_levels_sub(ctx, n1, 1, 0.0, &e1), "E0 singles");
            check(ctx, scp3g_lowest_levels_sub(ctx, n2s, 1, 0.0, &e2), "E0 doubles");
            e1 *= cli::autocm;
            e2 *= cli::autocm;
            e0sd << i << " " << e1 << " " << e2 << " " << e2 - e1 << std::endl;
            if (i % save_every == 0) {
#ifdef SCP3G_HAVE_HDF5
                Mout.resize((size_t)n * n);
                check(ctx, scp3g_h_download(ctx, Mout.data(), 0), "download");
                char s[128];
                snprintf(s, sizeof(s), "sM_%07lld.h5", (long long)i);
                join_writer();
                wbuf.swap(Mout);
                const std::string fname(s);
                writer = std::thread([&wbuf, &werr, &wfail, fname, n]() {
                    std::string e;
                    if (!cli::save_hdf5(fname, wbuf.data(), n, e)) {
                        werr = e;
                        wfail = true;
                    }
                });
#endif
            }
        }
    }
#ifdef SCP3G_HAVE_HDF5
    join_writer();
#endif
    if (o.levels > 0) {
        const int k = (int)std::min<int64_t>(o.levels, n);
        std::vector<double> ev(k);
        check(ctx, scp3g_lowest_levels(ctx, k, 0.0, ev.data(), nullptr), "levels");
        std::ofstream lv("levels.dat");
        lv << std::setprecision(15);
        lv << "# lowest " << k << " levels: hartree, cm^-1, cm^-1 above the lowest\n";
        for (int j = 0; j < k; ++j)
            lv << ev[j] << " " << ev[j] * cli::autocm << " " << (ev[j] - ev[0]) * cli::autocm << "\n";
        for (int j = 0; j < k; ++j) std::cout << std::setprecision(10) << ev[j] * cli::autocm << "\n";
    }
    scp3g_destroy(ctx);
    return 0;
}
