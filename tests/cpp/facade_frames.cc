// facade_frames.cc -- the diagnostics and frame output of the C++ drop-in:
// load a state into fm/vel, upload, then print tractions and qs_measure,
// dump compute_strain's cu slot, and write the frame files with
// write_files / output as a reference driver would.
//
// usage: facade_frames STATE OUTDIR CUFILE m n ax bx ay by u_bdry
//                      mu K visc adapt_fac tmult TZ c0 tau0 kappa Delta Omegaeps0 chi_inf theta
// STATE: 10 planes of m*(n+2) doubles (rows -1..n) then m*(n+1) vec2.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "shear_sim.hh"

int main(int argc, char **argv) {
    if (argc != 24) {
        std::fprintf(stderr, "facade_frames: expected 23 arguments, got %d\n", argc - 1);
        return 2;
    }
    const char *state = argv[1], *outdir = argv[2], *cufile = argv[3];
    int m = std::atoi(argv[4]), n = std::atoi(argv[5]);
    double a[18];
    for (int k = 0; k < 18; k++) a[k] = std::atof(argv[6 + k]);
    const double ax = a[0], bx = a[1], ay = a[2], by = a[3], u_bdry = a[4], mu = a[5], K = a[6], visc = a[7],
                 adapt_fac = a[8], tmult = a[9];
    stz_dynamics_linear_athermal stz(a[10], a[11], a[12], a[13], a[14], a[15], a[16], a[17], 0);
    shear_sim ss(m, n, ax, bx, ay, by, mu, K, visc, 0, 0, 1.0, tmult, adapt_fac, u_bdry, &stz, outdir);

    FILE *f = std::fopen(state, "rb");
    if (!f) return 3;
    std::vector<double> buf((size_t)10 * m * (n + 2) + 2 * (size_t)m * (n + 1));
    if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return 4;
    std::fclose(f);
    for (int fld = 0; fld < 10; fld++)
        for (int k = 0; k < m * (n + 2); k++) (&ss.fm[k - m].u)[fld] = buf[(size_t)fld * m * (n + 2) + k];
    for (int k = 0; k < m * (n + 1); k++) ss.vel[k] = vec(buf[10 * (size_t)m * (n + 2) + 2 * k],
                                                          buf[10 * (size_t)m * (n + 2) + 2 * k + 1]);
    ss.upload();

    double bxt, byt, txt, tyt, Q, mq, Qs, mqs;
    ss.tractions(bxt, byt, txt, tyt);
    ss.qs_measure(Q, mq);
    ss.qs_measure(Qs, mqs, true);
    std::printf("tractions %.17g %.17g %.17g %.17g\n", bxt, byt, txt, tyt);
    std::printf("qs %.17g %.17g %.17g %.17g\n", Q, mq, Qs, mqs);

    ss.compute_strain();
    ss.download();
    FILE *g = std::fopen(cufile, "wb");
    if (!g) return 5;
    for (int k = -m; k < m * (n + 1); k++) std::fwrite(&ss.fm[k].cu, sizeof(double), 1, g);
    std::fclose(g);

    ss.write_files(3);
    ss.output("u", 0, 3);
    return 0;
}
