// eigensolver: lowest K eigenpairs of an sM_*.h5 matrix, on the GPU.
//
// Usage and output follow eigensolver.cpp:83-109:
//   eigensolver sM_XXXXXXX.h5 K
// reads dataset "M" (FP32 in the file, as double), and writes
// eigenpairs_XXXXXXX.dat: "# Lowest K eigenvalues", the K eigenvalues in
// hartree (15 significant digits, one per line), "# Corresponding
// eigenvectors, one per line.", then one eigenvector per line.  The triangle
// read is chosen as eigensolver.cpp:27-30 does.  The dense dsyevr is replaced
// by the block Davidson solver of libscp3g (K <= 28).
#include <cstdlib>
#include <iomanip>
#include <vector>

#include "../../include/scp3g.h"
#include "cli_common.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::cerr << "usage: eigensolver sM_XXXXXXX.h5 K\n";
        return EXIT_FAILURE;
    }
#ifndef SCP3G_HAVE_HDF5
    std::cerr << "eigensolver was built without HDF5\n";
    return EXIT_FAILURE;
#else
    const std::string fin_name(argv[1]);
    const int k = atoi(argv[2]);
    std::vector<double> M;
    int64_t n = 0;
    std::string err;
    if (!cli::load_hdf5(fin_name, M, n, err)) {
        std::cerr << err << "\n";
        return EXIT_FAILURE;
    }
    if (k < 1 || k > n) {
        std::cerr << "K must be in 1..n\n";
        return EXIT_FAILURE;
    }
    scp3g_ctx *ctx = nullptr;
    if (scp3g_create(&ctx, 0)) {
        std::cerr << "no usable GPU\n";
        return EXIT_FAILURE;
    }
    std::vector<double> ev(k), V((size_t)n * k);
    if (scp3g_matrix_lowest_levels(ctx, n, M.data(), k, 0.0, ev.data(), V.data())) {
        std::cerr << "eigensolve failed: " << scp3g_last_error(ctx) << "\n";
        scp3g_destroy(ctx);
        return EXIT_FAILURE;
    }
    scp3g_destroy(ctx);
    // eigensolver.cpp:95-97: "eigenpairs" + name[2 : len-3] + ".dat"
    const std::string base = fin_name.substr(fin_name.find_last_of('/') + 1);
    std::string dir = fin_name.substr(0, fin_name.size() - base.size());
    const std::string fout_name = dir + "eigenpairs" + base.substr(2, base.length() - 5) + ".dat";
    std::ofstream fout(fout_name.c_str());
    fout.precision(15);
    fout << "# Lowest " << k << " eigenvalues\n";
    for (int j = 0; j < k; ++j) fout << ev[j] << "\n";
    fout << "# Corresponding eigenvectors, one per line." << std::endl;
    for (int j = 0; j < k; ++j) {
        for (int64_t r = 0; r < n; ++r) fout << (r ? " " : "") << V[(size_t)j * n + r];
        fout << "\n";
    }
    return 0;
#endif
}
