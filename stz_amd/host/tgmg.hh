// tgmg.hh -- the multigrid interface of the reference's tgmg library
// (tgmg/tgmg.hh:286-495, tgmg_predict.hh:6-33, tgmg_config.hh) for the
// gfx950 step.  tgmg<S,V,M> keeps the reference's constructor, public
// members and methods; the hierarchy, smoothing, transfers and solves run on
// the device through the stz_sim handle of the problem class (S::h), which
// the qs_multi problem class of qs_multi.hh provides.  check_symmetric and
// check_consistent run on the host over the problem class's own a_xx /
// mul_a (and its host array z), as the reference's do.
#ifndef STZ_AMD_TGMG_HH
#define STZ_AMD_TGMG_HH

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "../../include/stz_amd.h"
#include "mat.hh"
#include "vec.hh"

// tgmg_config.hh: batch sizes and predictor constants
const int tgmg_max_levels = 16;
const int v_cycle_per_loop = 4;
const int max_v_cycle_loops = 1000;
const int fmg_per_loop = 2;
const int max_fmg_loops = 1000;
const int tgmg_predict_init = 8;
const int tgmg_predict_mult = 16;
const int tgmg_predict_decay = 32;
const int tgmg_predict_max_iters = 2048;
const int tgmg_predict_extra_iters = 1;

inline double mod_sq(double c) { return c * c; }

/** tgmg.hh:41-45: the squared-L2 threshold eps * mult * typical, squared */
template <class F> inline F tgmg_accuracy(F typical, F mult = 1e4) {
    F acc = std::numeric_limits<F>::epsilon() * mult * typical;
    return acc * acc;
}

namespace stz_amd_detail {
inline void check(int code, const char *what) {
    if (code != STZ_OK) {
        std::fprintf(stderr, "stz_amd: %s failed (%d): %s\n", what, code, stz_last_error());
        std::exit(1);
    }
}
}  // namespace stz_amd_detail

/** The adaptive V-cycle count predictor (tgmg_predict.hh:6-33).  The
 * prediction of a device solve runs inside the library, on the simulation
 * the predictor is bound to (h); the counters of a bound predictor are the
 * device's. */
struct tgmg_predict {
    const int mult, decay, max_thresh, extra_iters;
    int lim, solves, vcount;
    stz_sim *h;
    tgmg_predict()
        : mult(tgmg_predict_mult), decay(tgmg_predict_decay), max_thresh(mult * tgmg_predict_max_iters),
          extra_iters(tgmg_predict_extra_iters), lim(mult * tgmg_predict_init), solves(0), vcount(0), h(0) {}
    void add_iters(int n) {
        solves++;
        vcount += n;
    }
    /** mean V-cycles per solve since the last call; resets the counters */
    float avg_iters() {
        if (h) {
            double a = 0;
            stz_amd_detail::check(stz_sim_avg_iters(h, &a), "avg_iters");
            return static_cast<float>(a);
        }
        float ans = static_cast<float>(vcount) / static_cast<float>(solves);
        solves = vcount = 0;
        return ans;
    }
};

/** The top level of a multigrid hierarchy (tgmg.hh:286-495). */
template <class S, class V, class M> class tgmg {
  public:
    S &q;
    V *const b;
    V *const z;
    const int m, n, mn;
    const bool x_prd, y_prd;
    /** number of child grids (tgmg.hh:340), known after setup() */
    int ml;
    int verbose;
    /** convergence rate (digits per V-cycle) of the last solve */
    double conv_rate;
    tgmg(S &q_, V *b_, V *z_)
        : q(q_), b(b_), z(z_), m(q_.m), n(q_.n), mn(q_.mn), x_prd(q_.x_prd), y_prd(q_.y_prd), ml(0), verbose(0),
          conv_rate(0) {}

    /** tgmg::setup (tgmg.hh:331-334): the Galerkin hierarchy for the
     * problem's current coefficients, built on the device */
    void setup() {
        stz_amd_detail::check(stz_mg_setup(q.h, q.viscdt_prev), "tgmg::setup");
        int nl = 0, dims[2 * tgmg_max_levels];
        stz_amd_detail::check(stz_sim_mg_hierarchy(q.h, tgmg_max_levels, &nl, dims), "tgmg::setup");
        ml = nl - 1;
    }
    /** tgmg::solve_v_cycle(tgmg_predict&, ...) (tgmg.hh:439-441,
     * tgmg.in.cc:266-301): adaptive V-cycles to the problem's threshold.
     * The coarsest level is solved exactly, which the reference's
     * 2 hcyc_bottom bottom sweeps approximate. */
    bool solve_v_cycle(tgmg_predict &tp, int cyc_down = 1, int cyc_up = 1, int hcyc_bottom = 10) {
        configure(cyc_down, cyc_up, hcyc_bottom, &tp);
        stz_solve_info info;
        const int code = stz_mg_solve(q.h, &info);
        if (!tp.h) tp.add_iters(info.cycles);
        if (code != STZ_OK) {
            if (verbose >= 1)
                std::printf("Residual=%g not reached %g threshold after %d iterations\n", info.residual,
                            info.tolerance, info.cycles);
            return false;
        }
        return true;
    }
    /** tgmg::solve_v_cycle(...) (tgmg.hh:425-427, tgmg.in.cc:241-255):
     * batches of v_cycle_per_loop V-cycles until the residual is at most
     * the problem's threshold q.acc */
    bool solve_v_cycle(int cyc_down = 1, int cyc_up = 1, int hcyc_bottom = 10) {
        double acc = 0;
        int loops = 0;
        do {
            if (++loops == max_v_cycle_loops) {
                if (verbose == 1)
                    std::printf("Residual=%g not reached %g threshold after %d iterations\n", acc, q.acc,
                                loops * v_cycle_per_loop);
                return false;
            }
            for (int k = 0; k < v_cycle_per_loop; k++) v_cycle(cyc_down, cyc_up, hcyc_bottom);
            acc = l2_error();
        } while (acc > q.acc);
        return !std::isnan(acc);
    }
    /** tgmg::v_cycle (tgmg.in.cc:313-328).  The smoother is the device's
     * four-colour Gauss-Seidel for either value of `symmetric`. */
    void v_cycle(int cyc_down = 1, int cyc_up = 1, int hcyc_bottom = 10, bool symmetric = false) {
        (void)symmetric;
        configure(cyc_down, cyc_up, hcyc_bottom, 0);
        stz_amd_detail::check(stz_mg_v_cycle(q.h), "tgmg::v_cycle");
    }
    /** tgmg::fmg_pre (tgmg.in.cc:369-399) */
    void fmg_pre(int cyc_down = 1, int cyc_up = 1, int hcyc_bottom = 10, bool symmetric = false) {
        (void)symmetric;
        configure(cyc_down, cyc_up, hcyc_bottom, 0);
        stz_amd_detail::check(stz_mg_fmg_pre(q.h), "tgmg::fmg_pre");
    }
    /** tgmg_base::gauss_seidel (tgmg.in.cc:127-165): one sweep on the top
     * level */
    void gauss_seidel() { stz_amd_detail::check(stz_mg_smooth(q.h, 0, 1), "tgmg::gauss_seidel"); }
    /** tgmg_base::clear_z (tgmg.in.cc:792-797), on the device and in the
     * host array z */
    void clear_z() {
        stz_amd_detail::check(stz_mg_clear_z(q.h), "tgmg::clear_z");
        for (int k = 0; k < mn; k++) z[k] = V(0.);
    }
    /** tgmg_base::l2_error (tgmg.hh:146): mean-square residual */
    double l2_error() {
        double e = 0;
        stz_amd_detail::check(stz_mg_l2_error(q.h, &e), "tgmg::l2_error");
        return e;
    }
    /** tgmg::print_hierarchy (tgmg.in.cc:36-41) */
    void print_hierarchy() {
        int nl = 0, dims[2 * tgmg_max_levels];
        stz_amd_detail::check(stz_sim_mg_hierarchy(q.h, tgmg_max_levels, &nl, dims), "tgmg::print_hierarchy");
        for (int k = 0; k < nl; k++) std::printf("# Grid %d: %d by %d\n", k, dims[2 * k], dims[2 * k + 1]);
    }

    /** tgmg::check_symmetric (tgmg_debug.cc:20-43): every pair of
     * off-diagonal entries of the problem class, once, against each other's
     * transpose position */
    void check_symmetric() {
        bool sym = true;
        for (int j = 0, ij = 0; j < n; j++)
            for (int i = 0; i < m; i++, ij++) {
                int li = 0, ri = 0;
                const bool ll = neighbour_left(i, li), rr = neighbour_right(i, ri);
                if (rr) sym &= pair_check(ij, ij + ri, q.a_cr(i, ij), q.a_cl(i + ri, ij + ri), "cr", "cl");
                if (j < n - 1 || y_prd) {
                    const int uij = ij + (j < n - 1 ? m : m - mn);
                    if (ll) sym &= pair_check(ij, uij + li, q.a_ul(i, ij), q.a_dr(i + li, uij + li), "ul", "dr");
                    sym &= pair_check(ij, uij, q.a_uc(i, ij), q.a_dc(i, uij), "uc", "dc");
                    if (rr) sym &= pair_check(ij, uij + ri, q.a_ur(i, ij), q.a_dl(i + ri, uij + ri), "ur", "dl");
                }
            }
        std::printf("The matrix is %ssymmetric\n", sym ? "" : "not ");
    }
    /** tgmg::check_consistent (tgmg_debug.cc:69-103): the problem class's
     * mul_a against its entries, one unit vector at a time in the host
     * array z; clears z (host and device) first, as the reference does */
    void check_consistent() {
        bool cons = true;
        clear_z();
        for (int j = 0, ij = 0; j < n; j++)
            for (int i = 0; i < m; i++, ij++) {
                int li = 0, ri = 0;
                const bool ll = neighbour_left(i, li), rr = neighbour_right(i, ri);
                if (j > 0 || y_prd) {
                    const int dij = ij - (j > 0 ? m : m - mn);
                    if (ll) cons &= unit_check(i, ij, dij + li, q.a_dl(i, ij), "dl");
                    cons &= unit_check(i, ij, dij, q.a_dc(i, ij), "dc");
                    if (rr) cons &= unit_check(i, ij, dij + ri, q.a_dr(i, ij), "dr");
                }
                if (ll) cons &= unit_check(i, ij, ij + li, q.a_cl(i, ij), "cl");
                if (rr) cons &= unit_check(i, ij, ij + ri, q.a_cr(i, ij), "cr");
                if (j < n - 1 || y_prd) {
                    const int uij = ij + (j < n - 1 ? m : m - mn);
                    if (ll) cons &= unit_check(i, ij, uij + li, q.a_ul(i, ij), "ul");
                    cons &= unit_check(i, ij, uij, q.a_uc(i, ij), "uc");
                    if (rr) cons &= unit_check(i, ij, uij + ri, q.a_ur(i, ij), "ur");
                }
            }
        std::printf("The mul_a routine is %sconsistent with the matrix terms\n", cons ? "" : "not ");
    }

  private:
    void configure(int cyc_down, int cyc_up, int hcyc_bottom, const tgmg_predict *tp) {
        stz_mg_config c;
        stz_amd_detail::check(stz_sim_get_mg_config(q.h, &c), "get_mg_config");
        c.nu1 = cyc_down;
        c.nu2 = cyc_up;
        c.bottom_sweeps = hcyc_bottom;
        if (tp) {
            c.predict_mult = tp->mult;
            c.predict_decay = tp->decay;
            c.predict_max_iters = tp->max_thresh / tp->mult;
            c.extra_cycles = tp->extra_iters;
        }
        stz_amd_detail::check(stz_sim_set_mg_config(q.h, &c), "set_mg_config");
    }
    // the column offsets of the left and right neighbours, false where the
    // grid ends (x periodic or not)
    bool neighbour_left(int i, int &li) const {
        if (i > 0) { li = -1; return true; }
        if (x_prd) { li = m - 1; return true; }
        return false;
    }
    bool neighbour_right(int i, int &ri) const {
        if (i < m - 1) { ri = 1; return true; }
        if (x_prd) { ri = 1 - m; return true; }
        return false;
    }
    bool pair_check(int ij, int ij2, M a, M a2, const char *p, const char *p2) {
        if (mod_sq(a - a2) > q.acc) {
            std::printf("a_%s(%d,%d)=(%.10g,%.10g,%.10g,%.10g) a_%s(%d,%d)=(%.10g,%.10g,%.10g,%.10g)\n", p, ij % m,
                        ij / m, a.a, a.b, a.c, a.d, p2, ij2 % m, ij2 / m, a2.a, a2.b, a2.c, a2.d);
            return false;
        }
        return true;
    }
    bool unit_check(int i, int ij, int ij2, M a, const char *p) {
        z[ij2] = V(1.);
        const V v = a * V(1.), v2 = q.mul_a(i, ij), dv = v2 - v;
        z[ij2] = V(0.);
        if (mod_sq(dv) > q.acc) {
            std::printf("a_%s at (%d,%d)=(%.10g,%.10g) but mul_a gives (%.10g,%.10g)\n", p, ij % m, ij / m, v.x, v.y,
                        v2.x, v2.y);
            return false;
        }
        return true;
    }
};

#endif
