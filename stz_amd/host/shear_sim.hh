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

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/stz_amd.h"
#include "fields.hh"
#include "mat.hh"
#include "qs_multi.hh"
#include "stz_model.hh"
#include "tgmg.hh"
#include "vec.hh"

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
          TZ(stz_->TZ), chi_inf(stz_->chi_inf), f_num(0), h(create(precision)), qsm(*this) {
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
    /** shear_sim::chi_diffusion (shear_sim.cc:364-386) adds the diffusion to
     * the change slot cchi, which only advection_step applies, after
     * overwriting it; called on its own it leaves the state unchanged, and
     * so does this.  advection_step includes the diffusion, as there. */
    void chi_diffusion(double dt) { (void)dt; }
    /** shear_sim::solve_quasistatic (shear_sim.cc:159-181): frames of
     * `steps` device steps each, the frame files written to `filename` after
     * each, and the reference's per-frame diagnostic line */
    void solve_quasistatic(double duration, int frames, int steps) {
        double dt = duration / frames * (1.0 / steps);
        if (f_num == 0) {
            write_files(0);
            std::puts("# Output frame 0");
        }
        for (int k = 1; k <= frames; k++) {
            step_forward_quasistatic(dt, steps);
            write_files(k + f_num);
            std::printf("# Output frame %d [%d] {%.2f}\n", k + f_num, steps, qsm.tp.avg_iters());
        }
        f_num += frames;
    }
    /** shear_sim::step_forward (shear_sim.cc:185-276): the direct step */
    void step_forward(double dt) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_step_forward(h, dt, 1), "step_forward");
        pull_time();
    }
    /** shear_sim::solve (shear_sim.cc:120-152): direct steps of dt =
     * tmult dx^2 up to each frame time, the last one shortened to land on it;
     * frames written as they fall due */
    void solve(double duration, int frames) {
        const double time_interval = duration / frames, t_start = time, dt = dx * dx * tmult;
        if (f_num == 0) {
            write_files(0);
            std::puts("# Output frame 0");
        }
        for (int k = 1; k <= frames; k++) {
            const double target_time = t_start + time_interval * k;
            int l = 1;
            while (time + dt * (1 + 1e-8) < target_time) {
                l++;
                step_forward(dt);
            }
            step_forward(target_time - time);
            write_files(k + f_num);
            std::printf("# Output frame %d [%d]\n", k + f_num, l);
        }
        f_num += frames;
    }
    /** shear_sim::write_files (shear_sim.cc:659-676): the reference's field
     * selection (p, tem, dev, X, Y) as Gnuplot binary matrices */
    void write_files(int k) {
        const int fflags = 4 | 128 | 256 | 512 | 1024;
        download();
        static const char *names[12] = {"u", "v", "p", "q", "s", "tau", "chi", "tem", "dev", "X", "Y", "proj"};
        for (int mode = 0; mode < 11; mode++)
            if (fflags & (1 << mode)) output(names[mode], mode, k);
    }
    /** shear_sim::output (shear_sim.cc:688-726) of the host mirror fm
     * (call download() first) */
    void output(const char *prefix, const int mode, const int sn) {
        const bool st = mode >= 2 && mode <= 8;
        const int l = st ? m : m + 1;
        std::vector<char> name(std::strlen(filename) + std::strlen(prefix) + 32);
        std::snprintf(name.data(), name.size(), "%s/%s.%d", filename, prefix, sn);
        FILE *outf = std::fopen(name.data(), "wb");
        if (!outf) {
            std::fprintf(stderr, "stz_amd: can't open %s\n", name.data());
            std::exit(1);
        }
        std::vector<float> buf(l + 1);
        buf[0] = l;
        for (int i = 0; i < l; i++) buf[i + 1] = ax + (st ? (i + 0.5) * dx : i * dx);
        std::fwrite(buf.data(), sizeof(float), l + 1, outf);
        c_field *fp = fm;
        for (int j = 0; j < (st ? n : n + 1); j++) {
            buf[0] = ay + (st ? (j + 0.5) * dy : j * dy);
            for (int i = 0; i < m; i++, fp++) {
                double v = 0;
                switch (mode) {
                case 0: v = fp->u; break;
                case 1: v = fp->v; break;
                case 2: v = fp->p; break;
                case 3: v = fp->q; break;
                case 4: v = fp->s; break;
                case 5: v = fp->tau; break;
                case 6: v = fp->chi; break;
                case 7: v = fp->chi * TZ; break;
                case 8: v = fp->dev(); break;
                case 9: v = fp->X; break;
                case 10: v = fp->Y; break;
                default: v = fp->cu;
                }
                buf[i + 1] = v;
            }
            if (!st) buf[m + 1] = mode == 9 ? buf[1] + bx - ax : buf[1];
            std::fwrite(buf.data(), sizeof(float), l + 1, outf);
        }
        std::fclose(outf);
    }
    /** shear_sim::tractions (shear_sim.cc:1003-1028) */
    void tractions(double &bx_, double &by_, double &tx, double &ty) {
        double o[4];
        stz_amd_detail::check(stz_sim_tractions(h, o), "tractions");
        bx_ = o[0];
        by_ = o[1];
        tx = o[2];
        ty = o[3];
    }
    /** shear_sim::qs_measure (shear_sim.cc:733-775); the integrand is in
     * fm[].cu after download() */
    void qs_measure(double &Q, double &max_qs, bool staggered = false) {
        stz_amd_detail::check(stz_sim_qs_measure(h, staggered ? 1 : 0, &Q, &max_qs), "qs_measure");
    }
    /** shear_sim::compute_strain (shear_sim.cc:778-824) into fm[].cu (after
     * download()); corrected fixes the reference formula's interior-row y
     * difference and seam-column sign */
    void compute_strain(bool corrected = false) {
        stz_amd_detail::check(stz_sim_compute_strain(h, corrected ? 1 : 0), "compute_strain");
    }
    /** shear_sim::l2_comparison (shear_sim.cc:460-502) */
    void l2_comparison(shear_sim &ss, double *l2) {
        stz_amd_detail::check(stz_sim_l2_comparison(h, ss.h, l2), "l2_comparison");
    }
    /** shear_sim::initialize_chi_bicubic (shear_sim.cc:832-913) */
    void initialize_chi_bicubic(int mm, int nn, double *arr, double lo = 0, double sca = 1) {
        sync_scalars();
        stz_amd_detail::check(stz_sim_initialize_chi_bicubic(h, mm, nn, arr, lo, sca), "initialize_chi_bicubic");
        download();
    }
    /** shear_sim::initialize_random (shear_sim.cc:932-985), on the C
     * library's rand() stream as the reference */
    void initialize_random(double l, double tem_avg, double tem_stdev) {
        std::printf("# Setting up random field (convolution window size: %d g.p.)\n", (int(l * 5.) << 1) + 1);
        sync_scalars();
        stz_amd_detail::check(stz_sim_initialize_random(h, l, tem_avg, tem_stdev), "initialize_random");
        download();
    }
    inline double avg_iters() { return qsm.tp.avg_iters(); }
    inline double bdry_vel() { return time < 1 ? time * u_bdry : u_bdry; }
    inline double bdry_pos() { return time < 1 ? u_bdry * time * time * 0.5 : u_bdry * (time - 0.5); }

    /** copy the device state into fm (rows -1..n; with the cu slot), vel
     * and src */
    void download() {
        std::vector<double> plane((size_t)m * (n + 2));
        for (int f = 0; f <= STZ_F_CU; f++) {
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

inline qs_multi::qs_multi(shear_sim &ss)
    : m(ss.m), n(ss.n + 1), mn(m * n), x_prd(true), y_prd(false), z(ss.vel), K(ss.K + ss.mu / 3.0),
      qK(.25 * K * ss.xsp * ss.ysp), Kmu(K + ss.mu), mu(ss.mu), dx(ss.dx), dy(ss.dy), xsp(ss.xsp), ysp(ss.ysp),
      xsp2(ss.xsp * ss.xsp), ysp2(ss.ysp * ss.ysp), xyfac(xsp2 + ysp2), tmufac1(Kmu * xsp2 + mu * ysp2),
      tmufac2(mu * xsp2 + Kmu * ysp2), bcs(ss.bcs), viscdt_prev(-1), Kmx(0), Kmy(0), mx(0), my(0), cxx(0), cyy(0),
      regz(mat(0)), acc(tgmg_accuracy(2 * (tmufac1 + tmufac2), 1e4)), h(ss.handle()), mg(*this, ss.src, z), tp() {
    tp.h = h;
}

#endif
