// C entry points over the reference's own CPU sources, compiled from where
// they lie (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used by tests/
// and tests/golden/make_golden.py to pin the oracle; never by the product.
#include <cstddef>
#include <string>

// sobol.hpp names std::string unqualified
using namespace std;

#include "sobol.hpp"
#include "beasley_springer_moro.h"
#include "HarmonicOscillatorBasis.h"
#include "qtip4pf.h"
#include "ttm3f.h"
#include "ps.h"
#include "ttm3f-bits-smear.h"

extern "C" {

// sobol::i8_sobol(dim, &seed, quasi): one point per call, state kept inside
// the reference (sobol.cpp:14198).  Returns the seed after the call.
long long ref_i8_sobol(int dims, long long seed, double *quasi)
{
    sobol::i8_sobol(dims, &seed, quasi);
    return seed;
}

// npts consecutive points starting at seed (point-major).
void ref_i8_sobol_block(int dims, long long seed, long long npts, double *quasi)
{
    for (long long i = 0; i < npts; i++) sobol::i8_sobol(dims, &seed, quasi + i * dims);
}

// sobol::std_normal for npts consecutive points (sobol.cpp:28706).
void ref_std_normal_block(int dims, long long seed, long long npts, double *out)
{
    for (long long i = 0; i < npts; i++) sobol::std_normal(dims, &seed, out + i * dims);
}

void ref_bsm(int n, double *u, double *x) { beasley_springer_moro(n, u, x); }

double ref_ho(int n, double y) { return ho_basis[n](y); }

double ref_qtip4pf(size_t nw, const double *crd)
{
    static h2o::qtip4pf pot;
    return pot(nw, crd);
}

double ref_qtip4pf_grad(size_t nw, const double *crd, double *grad)
{
    static h2o::qtip4pf pot;
    return pot(nw, crd, grad);
}

double ref_ttm3f(size_t nw, const double *crd)
{
    static h2o::ttm3f pot;
    return pot(nw, crd);
}

double ref_ttm3f_grad(size_t nw, const double *crd, double *grad)
{
    static h2o::ttm3f pot;
    return pot(nw, crd, grad);
}

double ref_pot_nasa(const double *rr) { return ps::pot_nasa(rr, 0); }

void ref_dms_nasa(double p1, double p2, double p3, const double *rr, double *q3, int ttm3)
{
    ps::dms_nasa(p1, p2, p3, rr, q3, 0, ttm3 != 0);
}

void ref_smear012(double r, double AA, double *ts)
{
    ttm3f_bits::smear012(r, AA, ts[0], ts[1], ts[2]);
}

}  // extern "C"
