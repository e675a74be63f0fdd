// Test driver for the host Rayleigh-Ritz eigensolver: reads m, nvec and an
// m x m column-major matrix (text) from stdin, prints the eigenvalues and
// the eigenvectors (column-major) it returns.
#include <cstdio>
#include <vector>

namespace scp3g {
int small_eigh(int m, std::vector<double> &A, int nvec, std::vector<double> &w, std::vector<double> &Z);
}

int main()
{
    int m, nvec;
    if (std::scanf("%d %d", &m, &nvec) != 2) return 2;
    std::vector<double> A((size_t)m * m), w, Z;
    for (auto &x : A)
        if (std::scanf("%lf", &x) != 1) return 2;
    scp3g::small_eigh(m, A, nvec, w, Z);
    for (double x : w) std::printf("%.17e\n", x);
    for (double x : Z) std::printf("%.17e\n", x);
    return 0;
}
