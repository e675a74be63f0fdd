// Test driver for the sM_*.h5 writer of the SCP3 driver (cli::save_hdf5):
// writes an n x n column-major matrix with the layout of SCP3's Mout (upper
// triangle filled, M(r, c) = r + c / 4 for r <= c, exactly representable in
// FP32; strict lower triangle zero) to argv[2], n = argv[1].
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "../scp3_amd/csrc/cli_common.h"

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    const int64_t n = std::atoll(argv[1]);
    std::vector<double> M((size_t)n * n, 0.0);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = 0; r <= c; ++r) M[(size_t)c * n + r] = (double)r + 0.25 * (double)c;
    std::string err;
    if (!cli::save_hdf5(argv[2], M.data(), n, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    return 0;
}
