// CPU instances of the potentials behind the h2o::PES interface (PES.h).
//
// The energies are the same host/device functions the GPU kernels run
// (qtip4pf_energy, ttm3f_energy) evaluated in FP64 on the host.  The
// gradient overload is analytic for q-TIP4P/F (the kernel's function,
// qtip4pf.cpp:252-397) and uses central differences (step 1e-5 Angstrom)
// for TTM3-F, whose analytic gradient exists only as a device kernel; the
// device path of the gradient estimator uses the analytic kernels.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/PES.h"
#include "scp3g_internal.h"

namespace scp3g {
double qtip4pf_energy_host(int nw, const double *crd);
double qtip4pf_energy_grad_host(int nw, const double *crd, double *grad);
double ttm3f_energy_host(int nw, const double *crd, int *iters);
}  // namespace scp3g

namespace {

class CpuPES : public h2o::PES {
public:
    CpuPES(int kind) : kind_(kind) {}
    const char *name() const override { return kind_ == 0 ? "q-TIP4P/F" : "TTM3-F"; }
    double operator()(size_t nw, const double *crd) override
    {
        return kind_ == 0 ? scp3g::qtip4pf_energy_host((int)nw, crd)
                          : scp3g::ttm3f_energy_host((int)nw, crd, nullptr);
    }
    double operator()(size_t nw, const double *crd, double *grad) override
    {
        if (kind_ == 0) return scp3g::qtip4pf_energy_grad_host((int)nw, crd, grad);
        const double h = 1e-5;
        std::vector<double> x(crd, crd + 9 * nw);
        for (size_t i = 0; i < 9 * nw; ++i) {
            const double x0 = x[i];
            x[i] = x0 + h;
            const double ep = (*this)(nw, x.data());
            x[i] = x0 - h;
            const double em = (*this)(nw, x.data());
            x[i] = x0;
            grad[i] = (ep - em) / (2 * h);
        }
        return (*this)(nw, crd);
    }

private:
    int kind_;
};

}  // namespace

namespace h2o {
PES *PESFromString(const std::string &name)
{
    if (name == "qtip4pf") return new CpuPES(0);
    if (name == "ttm3f") return new CpuPES(1);
    return nullptr;
}
}  // namespace h2o

// CPU energy (and optional central-difference gradient) of one configuration.
extern "C" int scp3g_pes_cpu(const char *name, int nw, const double *crd, double *E, double *grad)
{
    if (!name || nw <= 0 || !crd || !E) return -4;
    h2o::PES *p = h2o::PESFromString(name);
    if (!p) return -4;
    *E = grad ? (*p)((size_t)nw, crd, grad) : (*p)((size_t)nw, crd);
    delete p;
    return std::isfinite(*E) ? 0 : -10;
}
