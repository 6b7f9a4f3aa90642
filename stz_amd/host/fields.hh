// fields.hh -- the per-grid-point record of the reference (sb/main/fields.hh:
// 7-52), same members in the same order, used as the host mirror of the
// device planes by the shear_sim facade.
#ifndef STZ_AMD_FIELDS_HH
#define STZ_AMD_FIELDS_HH

#include <cmath>

struct c_field {
    double u, v, p, q, s, tau, chi, X, Y, ddev;
    double cu, cv, cp, cs, cq, ctau, cchi, cX, cY;
    double devsq() const { return 3 * q * q + s * s + tau * tau; }
    double dev() const { return std::sqrt(devsq()); }
    double sxx() const { return -p + s - q; }
    double syy() const { return -p - s - q; }
};

#endif
