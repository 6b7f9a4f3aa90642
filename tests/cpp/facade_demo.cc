// facade_demo.cc -- a driver written against the C++ drop-in headers the way
// the reference's sb drivers are written against shear_sim.hh: build the
// model and the simulation, initialise the fields, edit fm on the host,
// upload, step, download, and dump the state for comparison.
//
// usage: facade_demo OUT m n steps dt chi_case ay by u_bdry
//                    mu K visc adapt_fac tmult
//                    TZ c0 tau0 kappa Delta Omegaeps0 chi_inf theta
#include <cstdio>
#include <cmath>
#include <cstdlib>

#include "shear_sim.hh"

int main(int argc, char **argv) {
    if (argc != 23) {
        std::fprintf(stderr, "facade_demo: expected 22 arguments, got %d\n", argc - 1);
        return 2;
    }
    const char *out = argv[1];
    int m = std::atoi(argv[2]), n = std::atoi(argv[3]), steps = std::atoi(argv[4]);
    double dt = std::atof(argv[5]);
    int chi_case = std::atoi(argv[6]);
    double ay = std::atof(argv[7]), by = std::atof(argv[8]);
    double a[14];
    for (int k = 0; k < 14; k++) a[k] = std::atof(argv[9 + k]);
    double u_bdry = a[0], mu = a[1], K = a[2], visc = a[3], adapt_fac = a[4], tmult = a[5];

    stz_dynamics_linear_athermal stz(a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[13], 0);
    shear_sim ss(m, n, -1, 1, ay, by, mu, K, visc, 0, 0, 1.0, tmult, adapt_fac, u_bdry, &stz, "unused");
    stz_mg_config c = ss.mg_config();
    c.tol_mult = 1e-12;
    ss.set_mg_config(c);

    ss.init_fields(chi_case, 600, 200);
    // a host-side edit of fm; square roots of distinct integers, so that no
    // exact ENO ties arise (at a tie, rounding noise picks the branch)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) ss.fm[j * m + i].chi += 1e-4 * std::sqrt(1.0 + (i * 37 + j * 101) % 97);
    ss.time = 0.25;
    ss.upload();
    for (int k = 0; k < steps; k++) ss.step_forward_quasistatic(dt);
    ss.download();

    FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    for (int fld = 0; fld < STZ_NFIELDS; fld++)
        for (int k = -m; k < m * (n + 1); k++) std::fwrite(&ss.fm[k].u + fld, sizeof(double), 1, f);
    std::fwrite(&ss.vel[0].x, sizeof(double), 2 * (size_t)m * (n + 1), f);
    std::fwrite(&ss.time, sizeof(double), 1, f);
    std::fclose(f);
    std::printf("facade_demo: %dx%d, %d steps, time %.6g, avg V-cycles %.2f\n", m, n, steps, ss.time,
                ss.avg_iters());
    return 0;
}
