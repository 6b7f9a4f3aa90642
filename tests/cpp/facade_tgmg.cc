// facade_tgmg.cc -- the tgmg members of the C++ drop-in, used the way host
// code uses the reference's tgmg<qs_multi,vec,mat> (tgmg/tgmg.hh:286-495):
// set up the operator, check it (check_symmetric, check_consistent), run one
// V-cycle from zero, then from zero again fmg_pre and a V-cycle, solve with
// the adaptive predictor and with the fixed-batch loop, and dump the
// residuals and the solution.
//
// usage: facade_tgmg OUT m n viscdt
// The source is a fixed smooth field; OUT receives the velocity of the
// adaptive solve (2*m*(n+1) doubles) followed by the fixed-batch solve's.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "shear_sim.hh"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "facade_tgmg: expected 4 arguments, got %d\n", argc - 1);
        return 2;
    }
    const char *out = argv[1];
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]);
    const double viscdt = std::atof(argv[4]);
    const double TZ = 21000;
    stz_dynamics_linear_athermal stz(TZ, 0.4, 1e-13, 4.2, 8000 / TZ, 0.2638, 900 / TZ, 400 / TZ, 0);
    shear_sim ss(m, n, -1, 1, -1, 1, 20 / 0.85, 143.2 / 0.85, 0.02, 0, 0, 1, 0.5, 0.002, 2e-3, &stz, "unused");
    qs_multi &q = ss.mg();
    q.init(viscdt);
    q.mg.verbose = 1;
    q.mg.check_symmetric();
    q.mg.check_consistent();
    std::printf("levels %d\n", q.mg.ml);

    const int nodes = m * (n + 1);
    std::vector<double> b(2 * nodes);
    for (int j = 0; j <= n; j++)
        for (int i = 0; i < m; i++) {
            const int k = i + m * j;
            b[2 * k] = std::sin(6.283185307179586 * i / m) * std::cos(0.1 * j) + 0.01 * ((k * 37) % 11);
            b[2 * k + 1] = std::cos(6.283185307179586 * 2 * i / m) * std::sin(0.05 * j) - 0.01 * ((k * 13) % 7);
        }
    if (stz_mg_upload_src(ss.handle(), b.data()) != STZ_OK) return 3;

    q.mg.clear_z();
    q.mg.v_cycle();
    const double ev = q.mg.l2_error();
    q.mg.clear_z();
    const double e0 = q.mg.l2_error();
    q.mg.fmg_pre();
    const double e1 = q.mg.l2_error();
    q.mg.v_cycle();
    const double e2 = q.mg.l2_error();
    const bool ok = q.mg.solve_v_cycle(q.tp);
    const double e3 = q.mg.l2_error();
    std::printf("residual zero %.6e fmg_pre %.6e +v_cycle %.6e solved %d %.6e acc %.6e v_cycle_alone %.6e\n", e0,
                e1, e2, ok ? 1 : 0, e3, q.acc, ev);
    ss.download();
    FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    std::fwrite(&ss.vel[0].x, sizeof(double), 2 * (size_t)nodes, f);

    q.mg.clear_z();
    const bool ok2 = q.mg.solve_v_cycle(1, 1, 10);
    const double e4 = q.mg.l2_error();
    std::printf("batch solve %d %.6e\n", ok2 ? 1 : 0, e4);
    ss.download();
    std::fwrite(&ss.vel[0].x, sizeof(double), 2 * (size_t)nodes, f);
    std::fclose(f);
    return ok && ok2 ? 0 : 4;
}
