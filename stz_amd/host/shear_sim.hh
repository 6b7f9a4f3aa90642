// shear_sim.hh -- C++ drop-in for the reference's quasi-static shear_sim
// (sb/main/shear_sim.hh:20-151) and its multigrid member qs_multi
// (sb/main/qs_multi.hh:11-57, tgmg<qs_multi,vec,mat> + tgmg_predict).
//
// Same constructor, public members and step methods; the state lives in
// HBM on one gfx950 device and every step runs there through the C ABI of
// include/stz_amd.h.  The public arrays `fm`, `vel`, `src` are host mirrors
// with the reference's layout (fm offset by one row so fm[-m..-1] is the
// ghost row j=-1): call download() before reading them and upload() after
// writing them.  Failures print the library's message and exit(1), as the
// reference does on a failed solve (qs_multi.hh:52).
//
// Header-only; link with -lstz_amd (stz_amd/libstz_amd.so).
#ifndef STZ_AMD_SHEAR_SIM_HH
#define STZ_AMD_SHEAR_SIM_HH

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/stz_amd.h"
#include "fields.hh"
#include "mat.hh"
#include "stz_model.hh"
#include "vec.hh"

namespace stz_amd_detail {
inline void check(int code, const char *what) {
    if (code != STZ_OK) {
        std::fprintf(stderr, "stz_amd: %s failed (%d): %s\n", what, code, stz_last_error());
        std::exit(1);
    }
}
}  // namespace stz_amd_detail

class shear_sim;

/** Adaptive V-cycle count predictor (tgmg/tgmg_predict.hh:6-33).  The
 * prediction itself runs inside the library; this keeps the reference's
 * avg_iters() and the tuning constants. */
struct tgmg_predict {
    stz_sim *h;
    explicit tgmg_predict(stz_sim *h_) : h(h_) {}
    double avg_iters() {
        double a = 0;
        stz_amd_detail::check(stz_sim_avg_iters(h, &a), "avg_iters");
        return a;
    }
    void set(int mult, int decay, int init, int max_iters) {
        stz_mg_config c;
        stz_amd_detail::check(stz_sim_get_mg_config(h, &c), "get_mg_config");
        c.predict_mult = mult;
        c.predict_decay = decay;
        c.predict_init = init;
        c.predict_max_iters = max_iters;
        stz_amd_detail::check(stz_sim_set_mg_config(h, &c), "set_mg_config");
    }
};

/** The projection system of shear_sim (qs_multi.hh:11-57): operator setup
 * for viscosity/dt, the adaptive multigrid solve and the residual norm. */
struct qs_multi {
    stz_sim *h;
    tgmg_predict tp;
    explicit qs_multi(stz_sim *h_) : h(h_), tp(h_) {}
    /** qs_multi::init (qs_multi.cc:31-54) */
    void init(double viscdt) { stz_amd_detail::check(stz_mg_setup(h, viscdt), "qs_multi::init"); }
    /** qs_multi::solve (qs_multi.hh:51-53): exit(1) if the solve fails */
    void solve() {
        stz_solve_info info;
        stz_amd_detail::check(stz_mg_solve(h, &info), "tgmg::solve_v_cycle");
    }
    /** tgmg::v_cycle (tgmg.in.cc:313-328) */
    void v_cycle() { stz_amd_detail::check(stz_mg_v_cycle(h), "tgmg::v_cycle"); }
    /** tgmg_base::l2_error (tgmg.hh:146) */
    double l2_error() {
        double e = 0;
        stz_amd_detail::check(stz_mg_l2_error(h, &e), "tgmg::l2_error");
        return e;
    }
    /** tgmg_base::gauss_seidel (tgmg.in.cc:127-165) on hierarchy level lev */
    void gauss_seidel(int lev = 0, int sweeps = 1) {
        stz_amd_detail::check(stz_mg_smooth(h, lev, sweeps), "tgmg::gauss_seidel");
    }
    /** tgmg::print_hierarchy (tgmg.in.cc:36-41) */
    void print_hierarchy() {
        int nl = 0, dims[64];
        stz_amd_detail::check(stz_sim_mg_hierarchy(h, 32, &nl, dims), "mg_hierarchy");
        for (int k = 0; k < nl; k++) std::printf("# Grid %d: %d by %d\n", k, dims[2 * k], dims[2 * k + 1]);
    }
};

class shear_sim {
  public:
    const int m, n, mn;
    const double ax, ay, bx, by, lx, ly, dx, dy, xsp, ysp;
    const double mu, mu_inv, K, viscosity, chi_diff, chi_len, t_scale, tmult, adapt_fac, u_bdry;
    const double bcs;
    stz_dynamics *stz;
    const char *filename;
    /** host mirrors of the device state (see download()/upload()) */
    c_field *const fm;
    vec *const vel;
    vec *const src;
    double time;
    double TZ;
    double chi_inf;
    int f_num;

    shear_sim(const int m_, const int n_, const double ax_, const double bx_, const double ay_, const double by_,
              const double mu_, const double K_, const double visc_, const double chi_diff_,
              const double chi_len_, const double t_scale_, const double tmult_, const double adapt_fac_,
              const double u_bdry_, stz_dynamics *stz_, const char *filename_, int precision = STZ_PREC_FP64)
        : m(m_), n(n_), mn(m_ * n_), ax(ax_), ay(ay_), bx(bx_), by(by_), lx(bx_ - ax_), ly(by_ - ay_),
          dx(lx / m_), dy(ly / n_), xsp(1. / dx), ysp(1. / dy), mu(mu_), mu_inv(1 / mu_), K(K_),
          viscosity(visc_), chi_diff(chi_diff_), chi_len(chi_len_), t_scale(t_scale_), tmult(tmult_),
          adapt_fac(adapt_fac_), u_bdry(u_bdry_), bcs(4), stz(stz_), filename(filename_),
          fm(new c_field[mn + 2 * m] + m), vel(new vec[mn + m]), src(new vec[mn + m]), time(0.0),
          TZ(stz_->TZ), chi_inf(stz_->chi_inf), f_num(0), h(create(precision)), qsm(h) {
        for (int k = -m; k < mn + m; k++) {
            double *d = &fm[k].u;
            for (int f = 0; f < 19; f++) d[f] = 0;
        }
        for (int k = 0; k < mn + m; k++) vel[k] = src[k] = vec(0, 0);
    }
    ~shear_sim() {
        stz_sim_destroy(h);
        delete[] src;
        delete[] vel;
        delete[] (fm - m);
    }

    /** shear_sim::init_fields (shear_sim.cc:60-114), on the device; the
     * host mirrors are refreshed afterwards. */
    void init_fields(int chi_case, double tem_base = 600, double tem_delta = 200) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_init_fields(h, chi_case, tem_base, tem_delta), "init_fields");
        download();
    }
    /** shear_sim::step_forward_quasistatic (shear_sim.cc:280-291) */
    void step_forward_quasistatic(double dt) { step_forward_quasistatic(dt, 1); }
    /** nsteps steps of dt without returning to the host in between */
    void step_forward_quasistatic(double dt, int nsteps) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_step_quasistatic(h, dt, nsteps), "step_forward_quasistatic");
        pull_time();
    }
    /** shear_sim::advection_step (shear_sim.cc:296-361) */
    void advection_step(double dt) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_advection_step(h, dt), "advection_step");
        pull_time();
    }
    /** shear_sim::projection_step (shear_sim.cc:391-438) */
    void projection_step(double dt) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_projection_step(h, dt), "projection_step");
    }
    /** shear_sim::solve_quasistatic (shear_sim.cc:159-181) without file
     * output: frames of `steps` device steps each, with the reference's
     * per-frame diagnostic line; the host mirrors hold the last frame. */
    void solve_quasistatic(double duration, int frames, int steps) {
        double dt = duration / frames * (1.0 / steps);
        for (int k = 1; k <= frames; k++) {
            step_forward_quasistatic(dt, steps);
            download();
            std::printf("# Output frame %d [%d] {%.2f}\n", k + f_num, steps, qsm.tp.avg_iters());
        }
        f_num += frames;
    }
    inline double avg_iters() { return qsm.tp.avg_iters(); }
    inline double bdry_vel() { return time < 1 ? time * u_bdry : u_bdry; }
    inline double bdry_pos() { return time < 1 ? u_bdry * time * time * 0.5 : u_bdry * (time - 0.5); }

    /** copy the device state into fm (rows -1..n), vel and src */
    void download() {
        std::vector<double> plane((size_t)m * (n + 2));
        for (int f = 0; f < STZ_NFIELDS; f++) {
            stz_amd_detail::check(stz_sim_download_field(h, f, plane.data()), "download_field");
            for (int k = 0; k < m * (n + 2); k++) (&fm[k - m].u)[f] = plane[k];
        }
        stz_amd_detail::check(stz_sim_download_vel(h, &vel[0].x), "download_vel");
        stz_amd_detail::check(stz_sim_download_src(h, &src[0].x), "download_src");
        pull_time();
    }
    /** copy fm (rows -1..n), vel, time and chi_inf to the device */
    void upload() {
        std::vector<double> plane((size_t)m * (n + 2));
        for (int f = 0; f < STZ_NFIELDS; f++) {
            for (int k = 0; k < m * (n + 2); k++) plane[k] = (&fm[k - m].u)[f];
            stz_amd_detail::check(stz_sim_upload_field(h, f, plane.data()), "upload_field");
        }
        stz_amd_detail::check(stz_sim_upload_vel(h, &vel[0].x), "upload_vel");
        sync_scalars();
    }
    /** multigrid settings of the projection solve */
    stz_mg_config mg_config() {
        stz_mg_config c;
        stz_amd_detail::check(stz_sim_get_mg_config(h, &c), "get_mg_config");
        return c;
    }
    void set_mg_config(const stz_mg_config &c) {
        stz_amd_detail::check(stz_sim_set_mg_config(h, &c), "set_mg_config");
    }
    stz_sim *handle() { return h; }
    qs_multi &mg() { return qsm; }

  private:
    stz_sim *h;
    qs_multi qsm;
    shear_sim(const shear_sim &);
    shear_sim &operator=(const shear_sim &);

    stz_sim *create(int precision) {
        stz_sim_params p;
        p.m = m;
        p.n = n;
        p.ax = ax;
        p.bx = bx;
        p.ay = ay;
        p.by = by;
        p.mu = mu;
        p.K = K;
        p.viscosity = viscosity;
        p.chi_diff = chi_diff;
        p.chi_len = chi_len;
        p.t_scale = t_scale;
        p.tmult = tmult;
        p.adapt_fac = adapt_fac;
        p.u_bdry = u_bdry;
        p.stz = stz->to_params();
        stz_sim *s = 0;
        stz_amd_detail::check(stz_sim_create(&p, precision, &s), "shear_sim");
        return s;
    }
    void sync_scalars() {
        stz_amd_detail::check(stz_sim_set_time(h, time), "set_time");
        stz_amd_detail::check(stz_sim_set_chi_inf(h, chi_inf), "set_chi_inf");
    }
    void pull_time() { stz_amd_detail::check(stz_sim_get_time(h, &time), "get_time"); }
};

#endif
