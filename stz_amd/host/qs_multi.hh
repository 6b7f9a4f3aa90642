// qs_multi.hh -- the projection problem class of the reference's shear_sim
// (sb/main/qs_multi.hh:11-57, qs_multi.cc:9-54) for the gfx950 step.
//
// The public constants, the 2x2-block stencil functions a_dl .. a_ur, a_cc,
// inv_cc and the matrix-free mul_a are the reference's, evaluated on the
// host over the host array z (shear_sim's vel mirror), so host code written
// against the problem class (and tgmg's check_symmetric / check_consistent)
// works unchanged.  init() sets the coefficients and builds the device
// hierarchy; solve() runs the adaptive V-cycles on the device.
#ifndef STZ_AMD_QS_MULTI_HH
#define STZ_AMD_QS_MULTI_HH

#include <cstdlib>

#include "../../include/stz_amd.h"
#include "mat.hh"
#include "tgmg.hh"
#include "vec.hh"

class shear_sim;

struct qs_multi {
    const int m, n, mn;
    const bool x_prd;
    const bool y_prd;
    static const char gs_mode = 0;
    vec *const z;
    const double K, qK, Kmu, mu;
    const double dx, dy, xsp, ysp, xsp2, ysp2, xyfac;
    const double tmufac1, tmufac2;
    const double bcs;
    double viscdt_prev;
    double Kmx, Kmy, mx, my, cxx, cyy;
    const mat regz;
    const double acc;
    mat regh, regv, regc, regci, regcb, regcbi;
    /** the device simulation the hierarchy and solves live on */
    stz_sim *const h;
    /** qs_multi.cc:9-15 (defined in shear_sim.hh, which completes shear_sim) */
    explicit qs_multi(shear_sim &ss);
    ~qs_multi() {}
    /** qs_multi::init (qs_multi.cc:31-54): coefficients for viscosity/dt,
     * then the hierarchy (tgmg::setup) when they changed */
    void init(double viscdt) {
        if (viscdt != viscdt_prev) {
            viscdt_prev = viscdt;
            Kmx = (Kmu + viscdt) * xsp2;
            Kmy = (Kmu + viscdt) * ysp2;
            mx = (mu + viscdt) * xsp2;
            my = (mu + viscdt) * ysp2;
            cxx = -2 * (tmufac1 + viscdt * xyfac);
            cyy = -2 * (tmufac2 + viscdt * xyfac);
            regh = mat(Kmx, 0, 0, my);
            regv = mat(mx, 0, 0, Kmy);
            regc = mat(cxx, 0, 0, cyy);
            regci = mat(1 / cxx, 0, 0, 1 / cyy);
            regcb = mat(bcs * cxx, 0, 0, bcs * cyy);
            regcbi = mat(1 / (bcs * cxx), 0, 0, 1 / (bcs * cyy));
            mg.setup();
        }
    }
    /** the wall rows (qs_multi.hh:30) */
    inline bool bdry(int i, int ij) {
        (void)i;
        return ij >= mn - m || ij < m;
    }
    inline mat a_dl(int i, int ij) { return bdry(i, ij) ? regz : mat(0, qK, qK, 0); }
    inline mat a_dr(int i, int ij) { return bdry(i, ij) ? regz : mat(0, -qK, -qK, 0); }
    inline mat a_ul(int i, int ij) { return bdry(i, ij) ? regz : mat(0, -qK, -qK, 0); }
    inline mat a_ur(int i, int ij) { return bdry(i, ij) ? regz : mat(0, qK, qK, 0); }
    inline mat a_cl(int i, int ij) { return bdry(i, ij) ? regz : regh; }
    inline mat a_cr(int i, int ij) { return bdry(i, ij) ? regz : regh; }
    inline mat a_dc(int i, int ij) { return bdry(i, ij) ? regz : regv; }
    inline mat a_uc(int i, int ij) { return bdry(i, ij) ? regz : regv; }
    inline mat a_cc(int i, int ij) { return bdry(i, ij) ? regcb : regc; }
    inline vec inv_cc(int i, int ij, vec v) { return (bdry(i, ij) ? regcbi : regci) * v; }
    /** the off-diagonal part of the operator at node ij from the host array
     * z (qs_multi.hh:40-50) */
    inline vec mul_a(int i, int ij) {
        if (bdry(i, ij)) return vec(0, 0);
        const vec *zp = z + ij, *zl = zp + (i == 0 ? m - 1 : -1), *zr = zp + (i == m - 1 ? 1 - m : 1);
        return vec(Kmx * (zl->x + zr->x) + my * (zp[-m].x + zp[m].x) +
                       qK * (zl[-m].y + zr[m].y - zr[-m].y - zl[m].y),
                   mx * (zl->y + zr->y) + Kmy * (zp[-m].y + zp[m].y) +
                       qK * (zl[-m].x + zr[m].x - zr[-m].x - zl[m].x));
    }
    /** qs_multi::solve (qs_multi.hh:51-53): exit(1) if the solve fails */
    inline void solve() {
        if (!mg.solve_v_cycle(tp)) std::exit(1);
    }
    tgmg<qs_multi, vec, mat> mg;
    tgmg_predict tp;
};

#endif
