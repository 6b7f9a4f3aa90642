// device.h -- shared declarations for the gfx950 kernels of the quasi-static
// STZ step: launch-side structs, error handling, vec2/mat2 device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/stz_amd.h"

namespace stz {

// ---- error handling --------------------------------------------------------
void set_error(const char *fmt, ...);
#define STZ_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::stz::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                             hipGetErrorString(e_));                               \
            return STZ_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// ---- small vector / matrix types ------------------------------------------
template <typename T> struct v2t { T x, y; };
template <typename T> struct alignas(sizeof(T) * 2) v2a { T x, y; };

// ---- field plane set --------------------------------------------------------
// All planes are m*(n+2) values, row r = j+1 for grid row j in [-1,n].
struct FieldPtrs {
    double *u, *v, *p, *q, *s, *tau, *chi, *X, *Y, *ddev;
};

// Physical and derived constants of one simulation (shear_sim.cc:32-36,
// qs_multi.cc:9-15).
struct SimConst {
    int m, n;
    double ax, ay, bx, by, lx, ly, dx, dy, xsp, ysp;
    double mu, mu_inv, K, viscosity, chi_diff, chi_len, t_scale, adapt_fac, u_bdry, bcs;
    double chi_inf;          // shear_sim::chi_inf (shear_sim.hh:92), settable
    stz_model_params stz;
    double thetainv;
};

// launchers (explicit.hip)
// rep: a decision record to replay (one uint32 per cell), or null to evaluate
int launch_advect(const SimConst &c, const FieldPtrs &old_f, const FieldPtrs &new_f, double dt,
                  uint32_t *dec, hipStream_t st, const uint32_t *rep = nullptr);
int launch_direct(const SimConst &c, const FieldPtrs &old_f, const FieldPtrs &new_f, double *un, double *vn,
                  double dt, hipStream_t st);
int launch_chi_diffusion(const SimConst &c, const FieldPtrs &old_f, const FieldPtrs &new_f, double dt,
                         hipStream_t st);
int launch_source(const SimConst &c, const FieldPtrs &f, double dt, double xbc, double bvel, double *src,
                  hipStream_t st);
int launch_correct(const SimConst &c, const FieldPtrs &f, const double *vel, double dt, double bvel,
                   double bpos, hipStream_t st);
int launch_boundaries(const SimConst &c, const FieldPtrs &f, double bvel, double bpos, hipStream_t st);
int launch_init_fields(const SimConst &c, const FieldPtrs &f, int chi_case, double chi0, double chi1,
                       hipStream_t st);

}  // namespace stz
