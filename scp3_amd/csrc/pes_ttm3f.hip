// TTM3-F energies (placeholder until the device kernel lands).
#include "scp3g_internal.h"

#include <cmath>

namespace scp3g {

hipError_t launch_ttm3f(scp3g_ctx *, int, int64_t, const void *, int64_t, int, double *)
{
    return hipErrorNotSupported;
}

double ttm3f_energy_host(int, const double *, int *iters)
{
    if (iters) *iters = 0;
    return NAN;
}

}  // namespace scp3g
