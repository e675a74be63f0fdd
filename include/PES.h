/*
 * PES.h - the potential interface of SCP3 (reference PES.h:13-23), kept so
 * that CPU potentials (WHBB, HBB2-pol or a user's own) plug into the host
 * driver unchanged.  Coordinates: 9*nw doubles in O H H order, Angstrom;
 * energy in kcal/mol; gradient (9*nw doubles, kcal/mol/Angstrom) written in
 * full by the gradient overload.  Instances are not re-entrant.
 *
 * The GPU path evaluates qtip4pf and ttm3f in batches behind the C ABI
 * (scp3g.h); PESFromString returns CPU instances of the same potentials,
 * used for input preparation and for potentials without a device kernel.
 */
#ifndef SCP3G_PES_H
#define SCP3G_PES_H

#include <cstddef>
#include <string>

namespace h2o {

class PES {
public:
    PES() {}
    virtual ~PES() {}

    virtual const char *name() const = 0;
    virtual double operator()(size_t nw, const double *crd) = 0;
    virtual double operator()(size_t nw, const double *crd, double *grad) = 0;
};

/* h2o.h:20-48: "qtip4pf" and "ttm3f"; returns nullptr for an unknown name
 * (the reference exits instead).  The caller owns the instance. */
PES *PESFromString(const std::string &name);

}  // namespace h2o

#endif /* SCP3G_PES_H */
