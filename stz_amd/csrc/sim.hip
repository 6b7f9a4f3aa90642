// sim.hip -- the C ABI of include/stz_amd.h: simulation state in HBM, the
// per-step kernel sequence, the multigrid hierarchy and its adaptive V-cycle
// driver.
//
// State layout in HBM (one allocation):
//   two plane sets {p,q,s,tau,chi,X,Y} that the advection kernel ping-pongs
//   between, plus u, v, ddev planes, each m*(n+2) doubles (rows -1..n);
//   vel and src (m*(n+1) interleaved vec2) are the top multigrid level's z, b.
// Multigrid: same hierarchy and Galerkin operators as the reference; each
// level stores one 2x2-block stencil per row when the operator is invariant
// along the periodic x direction (always, when every level above has an even
// width), else one per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "device.h"
#include "diag.h"
#include "mg.h"
#include "mg_node.h"
#include "sim_internal.h"

namespace stz {

static thread_local char g_err[512];
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}


// 1D prolongation parents (tgmg.in.cc:552-621): periodic axis with unwrapped
// fine index f; returns count, unwrapped coarse indices and weights
static int parents_periodic(int f, int uf, int sc, int *c, double *w) {
    const int fw = ((f % uf) + uf) % uf, sh = (f - fw) / uf;
    if ((fw & 1) == 0) { c[0] = fw / 2 + sh * sc; w[0] = 1; return 1; }
    c[0] = fw / 2 + sh * sc; w[0] = 0.5;
    c[1] = fw / 2 + 1 + sh * sc; w[1] = 0.5;
    return 2;
}
static int parents_bounded(int f, int uf, int sc, int *c, double *w) {
    if (f < 0 || f >= uf) return 0;
    if ((uf & 1) == 0 && f == uf - 1) { c[0] = sc - 1; w[0] = 1; return 1; }
    if ((f & 1) == 0) { c[0] = f / 2; w[0] = 1; return 1; }
    c[0] = f / 2; w[0] = 0.5; c[1] = f / 2 + 1; w[1] = 0.5;
    return 2;
}
static double weight_of(int np, const int *c, const double *w, int target) {
    for (int k = 0; k < np; k++)
        if (c[k] == target) return w[k];
    return 0;
}

// Galerkin coarse stencil at coarse node (I,J): A_c = P^T A P
// (tgmg.in.cc:667-779), then the inverse of the centre block
static void rap_node(const HostLevel &F, const HostLevel &C, int I, int J, double *t) {
    std::fill(t, t + 40, 0.0);
    const int fx0 = 2 * I, fy0 = ((F.n & 1) == 0 && J == C.n - 1) ? F.n - 1 : 2 * J;
    for (int b = -1; b <= 1; b++) {
        int cy[2], cx[2];
        double wy[2], wx[2];
        const int fy = fy0 + b;
        int np = parents_bounded(fy, F.n, C.n, cy, wy);
        const double wfy = weight_of(np, cy, wy, J);
        if (wfy == 0) continue;
        for (int a = -1; a <= 1; a++) {
            const int fxu = fx0 + a;
            np = parents_periodic(fxu, F.m, C.m, cx, wx);
            const double wfx = weight_of(np, cx, wx, I);
            if (wfx == 0) continue;
            const int fi = ((fxu % F.m) + F.m) % F.m;
            const double *A = F.at(fi, fy);
            for (int k = 0; k < 9; k++) {
                const int gy = fy + k / 3 - 1, gxu = fxu + k % 3 - 1;
                if (gy < 0 || gy >= F.n) continue;
                const int ny = parents_bounded(gy, F.n, C.n, cy, wy);
                const int nx = parents_periodic(gxu, F.m, C.m, cx, wx);
                for (int q = 0; q < ny; q++)
                    for (int p = 0; p < nx; p++) {
                        const int dI = cx[p] - I, dJ = cy[q] - J;
                        if (dI < -1 || dI > 1 || dJ < -1 || dJ > 1) continue;
                        const double w = wfx * wfy * wx[p] * wy[q];
                        double *e = t + 4 * ((dI + 1) + 3 * (dJ + 1));
                        for (int r = 0; r < 4; r++) e[r] += w * A[4 * k + r];
                    }
            }
        }
    }
    const double a = t[16], bb = t[17], c = t[18], d = t[19];
    const double idet = 1. / (a * d - bb * c);
    t[36] = d * idet;
    t[37] = -bb * idet;
    t[38] = -c * idet;
    t[39] = a * idet;
}

// one row stencil of the top operator of qs_multi (qs_multi.hh:30-50,
// qs_multi.cc:31-54): the x row uses Kmx along x and my along y, the y row mx
// along x and Kmy along y (the matrix-free mul_a); wall rows are bcs*diag.
void top_row_stencil(const SimConst &c, double viscdt, bool wall, double *s) {
    const double Kp = c.K + c.mu / 3.0, Kmu = Kp + c.mu;
    const double xsp2 = c.xsp * c.xsp, ysp2 = c.ysp * c.ysp, xyfac = xsp2 + ysp2;
    const double tmufac1 = Kmu * xsp2 + c.mu * ysp2, tmufac2 = c.mu * xsp2 + Kmu * ysp2;
    const double qK = .25 * Kp * c.xsp * c.ysp;
    const double Kmx = (Kmu + viscdt) * xsp2, Kmy = (Kmu + viscdt) * ysp2;
    const double mx = (c.mu + viscdt) * xsp2, my = (c.mu + viscdt) * ysp2;
    const double cxx = -2 * (tmufac1 + viscdt * xyfac), cyy = -2 * (tmufac2 + viscdt * xyfac);
    std::fill(s, s + 40, 0.0);
    if (wall) {
        s[16] = c.bcs * cxx;
        s[19] = c.bcs * cyy;
        s[36] = 1 / (c.bcs * cxx);
        s[39] = 1 / (c.bcs * cyy);
        return;
    }
    const double corner[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
    for (int k : {0, 2, 6, 8}) {
        s[4 * k + 1] = corner[k] * qK;
        s[4 * k + 2] = corner[k] * qK;
    }
    for (int k : {3, 5}) { s[4 * k] = Kmx; s[4 * k + 3] = mx; }
    for (int k : {1, 7}) { s[4 * k] = my; s[4 * k + 3] = Kmy; }
    s[16] = cxx;
    s[19] = cyy;
    s[36] = 1 / cxx;
    s[39] = 1 / cyy;
}

// the top operator as one stencil per row
static void top_stencil(const SimConst &c, double viscdt, HostLevel &L) {
    L.m = c.m;
    L.n = c.n + 1;
    L.xinv = true;
    L.S.assign((size_t)L.n * 40, 0.0);
    for (int j = 0; j < L.n; j++) top_row_stencil(c, viscdt, j == 0 || j == L.n - 1, &L.S[(size_t)j * 40]);
}

static void coarsen(const HostLevel &F, HostLevel &C) {
    C.xinv = F.xinv && (F.m % 2 == 0);
    const size_t pts = C.xinv ? C.n : (size_t)C.m * C.n;
    C.S.assign(pts * 40, 0.0);
    for (int J = 0; J < C.n; J++) {
        if (C.xinv) rap_node(F, C, 0, J, &C.S[(size_t)J * 40]);
        else
            for (int I = 0; I < C.m; I++) rap_node(F, C, I, J, &C.S[(size_t)(I + C.m * J) * 40]);
    }
}

// the constants of a simulation (shear_sim.cc:32-36) and the base of its
// solve threshold, qs_multi.cc:14 with tgmg_accuracy (tgmg.hh:41-45)
void fill_sim_const(const stz_sim_params *p, SimConst &c, double *acc_base) {
    c.m = p->m;
    c.n = p->n;
    c.ax = p->ax; c.bx = p->bx; c.ay = p->ay; c.by = p->by;
    c.lx = p->bx - p->ax;
    c.ly = p->by - p->ay;
    c.dx = c.lx / p->m;
    c.dy = c.ly / p->n;
    c.xsp = 1. / c.dx;
    c.ysp = 1. / c.dy;
    c.mu = p->mu;
    c.mu_inv = 1 / p->mu;
    c.K = p->K;
    c.viscosity = p->viscosity;
    c.chi_diff = p->chi_diff;
    c.chi_len = p->chi_len;
    c.t_scale = p->t_scale;
    c.adapt_fac = p->adapt_fac;
    c.u_bdry = p->u_bdry;
    c.bcs = 4;
    c.chi_inf = p->stz.chi_inf;
    c.stz = p->stz;
    c.thetainv = 1 / p->stz.theta;
    if (acc_base) {
        const double Kp = c.K + c.mu / 3.0, Kmu = Kp + c.mu;
        const double xsp2 = c.xsp * c.xsp, ysp2 = c.ysp * c.ysp;
        const double tmufac1 = Kmu * xsp2 + c.mu * ysp2, tmufac2 = c.mu * xsp2 + Kmu * ysp2;
        const double a = 2.220446049250313080847e-16 * 1e4 * (2 * (tmufac1 + tmufac2));
        *acc_base = a * a;
    }
}

// dense inverse of the coarsest operator (unknown 2*(i+m*j)+comp)
static bool bottom_inverse(const HostLevel &B, std::vector<double> &inv) {
    const int N = 2 * B.m * B.n;
    std::vector<double> A((size_t)N * N, 0.0);
    for (int j = 0; j < B.n; j++)
        for (int i = 0; i < B.m; i++) {
            const double *s = B.at(i, j);
            const int r = 2 * (i + B.m * j);
            for (int k = 0; k < 9; k++) {
                const int jj = j + k / 3 - 1;
                if (jj < 0 || jj >= B.n) continue;
                const int ii = ((i + k % 3 - 1) % B.m + B.m) % B.m;
                const int col = 2 * (ii + B.m * jj);
                A[(size_t)r * N + col] += s[4 * k];
                A[(size_t)r * N + col + 1] += s[4 * k + 1];
                A[(size_t)(r + 1) * N + col] += s[4 * k + 2];
                A[(size_t)(r + 1) * N + col + 1] += s[4 * k + 3];
            }
        }
    inv.assign((size_t)N * N, 0.0);
    for (int r = 0; r < N; r++) inv[(size_t)r * N + r] = 1;
    for (int col = 0; col < N; col++) {  // Gauss-Jordan, partial pivoting
        int piv = col;
        for (int r = col + 1; r < N; r++)
            if (std::fabs(A[(size_t)r * N + col]) > std::fabs(A[(size_t)piv * N + col])) piv = r;
        if (A[(size_t)piv * N + col] == 0) return false;
        if (piv != col)
            for (int k = 0; k < N; k++) {
                std::swap(A[(size_t)piv * N + k], A[(size_t)col * N + k]);
                std::swap(inv[(size_t)piv * N + k], inv[(size_t)col * N + k]);
            }
        const double d = 1 / A[(size_t)col * N + col];
        for (int k = 0; k < N; k++) { A[(size_t)col * N + k] *= d; inv[(size_t)col * N + k] *= d; }
        for (int r = 0; r < N; r++) {
            if (r == col) continue;
            const double f = A[(size_t)r * N + col];
            if (f == 0) continue;
            for (int k = 0; k < N; k++) {
                A[(size_t)r * N + k] -= f * A[(size_t)col * N + k];
                inv[(size_t)r * N + k] -= f * inv[(size_t)col * N + k];
            }
        }
    }
    return true;
}


template <typename T> static int upload(T **dst, const std::vector<double> &src) {
    std::vector<T> tmp(src.begin(), src.end());
    if (!*dst) STZ_HIP(hipMalloc(dst, tmp.size() * sizeof(T)));
    STZ_HIP(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return STZ_OK;
}

}  // namespace stz

using namespace stz;


static void set_defaults(stz_mg_config &g) {
    std::memset(&g, 0, sizeof g);
    g.nu1 = 1;
    g.nu2 = 1;
    g.bottom_sweeps = 10;
    g.extra_cycles = 1;                 // tgmg_config.hh:45
    g.predict_mult = 16;                // tgmg_config.hh:33
    g.predict_decay = 32;               // tgmg_config.hh:37
    g.predict_init = 8;                 // tgmg_config.hh:26
    g.predict_max_iters = 2048;         // tgmg_config.hh:41
    g.tol_mult = 1.0;
    g.max_refine = 4096;
    g.direct_nodes = 4200;
}

static size_t plane_elems(const stz_sim *s) { return (size_t)s->c.m * (s->c.n + 2); }

// hierarchy dimensions (tgmg.in.cc:10-30, tgmg.hh:109-110: x periodic, y not)
static void build_dims(stz_sim *s) {
    s->hl.clear();
    HostLevel top;
    top.m = s->c.m;
    top.n = s->c.n + 1;
    top.xinv = true;
    s->hl.push_back(top);
    int am = (top.m + 1) >> 1, an = (top.n + 2) >> 1;
    while (am * an > 10 && (int)s->hl.size() < kMaxLevels) {
        HostLevel L;
        L.m = am;
        L.n = an;
        s->hl.push_back(L);
        const int um = am, un = an;
        am = (um + 1) >> 1;
        an = (un + 2) >> 1;
    }
}

template <typename T> static int alloc_levels(stz_sim *s, DevMG<T> &mg, bool own_top) {
    mg.L = (int)s->hl.size();
    mg.args.L = mg.L;
    mg.args.ld = -1;
    mg.args.Gd = nullptr;
    for (int l = 0; l < mg.L; l++) {
        LevelArgs<T> &A = mg.args.lv[l];
        A.m = s->hl[l].m;
        A.n = s->hl[l].n;
        A.xodd = A.m & 1;
        const size_t cnt = 2 * (size_t)A.m * A.n;
        if (l > 0 || own_top) {
            STZ_HIP(hipMalloc(&mg.zb[l], 3 * cnt * sizeof(T)));
            STZ_HIP(hipMemset(mg.zb[l], 0, 3 * cnt * sizeof(T)));
            A.z = mg.zb[l];
            A.b = mg.zb[l] + cnt;
            A.z2 = mg.zb[l] + 2 * cnt;
        } else {  // FP64 top level: z is the simulation's vel, z2 its twin
            STZ_HIP(hipMalloc(&mg.zb[l], cnt * sizeof(T)));
            STZ_HIP(hipMemset(mg.zb[l], 0, cnt * sizeof(T)));
            A.z2 = mg.zb[l];
        }
    }
    // first level run by the workgroup solver: the largest suffix of levels
    // whose z, b, stencils and bottom inverse fit the LDS budget, never the
    // top level (stencil sizes assume one per row; mg_build re-checks)
    const size_t budget = 128 * 1024;
    int lc = mg.L - 1;
    while (lc > 1) {
        mg.args.lc = lc - 1;
        if ((size_t)coarse_layout(mg.args) * sizeof(T) > budget) break;
        lc--;
    }
    if (mg.L == 1) lc = 0;
    mg.args.lc = lc;
    mg.lds = (size_t)coarse_layout(mg.args) * sizeof(T);
    return STZ_OK;
}

template <typename T> static int set_lds_attr(const DevMG<T> &mg) {
    if (mg.lds > 160 * 1024) {
        set_error("coarse-level working set (%zu bytes) exceeds LDS", mg.lds);
        return STZ_ERR_STATE;
    }
    return mg_set_lds_limit<T>(mg.lds);
}

// upload the stencils of every level in precision T, recompute the
// workgroup solver's LDS layout and upload its constant part
template <typename T>
static int set_level_stencils(stz_sim *s, DevMG<T> &mg, const std::vector<std::vector<double>> &Sx,
                              const std::vector<double> &binv) {
    int e;
    for (size_t l = 0; l < s->hl.size(); l++) {
        const HostLevel &H = s->hl[l];
        if ((e = upload(&mg.S[l], Sx[l]))) return e;
        LevelArgs<T> &A = mg.args.lv[l];
        A.S = mg.S[l];
        A.rs = H.xinv ? 1 : A.m;
        A.cs = H.xinv ? 0 : 1;
        A.Si = mg.S[l] + H.S.size();
        A.nb = H.nb;
        A.sparse = H.xinv && sparse_pattern(Sx[l].data() + H.S.size()) ? 1 : 0;
    }
    mg.lds = (size_t)coarse_layout(mg.args) * sizeof(T);
    if ((e = set_lds_attr(mg))) return e;
    std::vector<T> pk;
    coarse_pack(mg.args, Sx.data(), binv, pk);
    if (mg.pack && mg.pack_elems != pk.size()) {
        (void)hipFree(mg.pack);
        mg.pack = nullptr;
    }
    if (!mg.pack) STZ_HIP(hipMalloc(&mg.pack, pk.size() * sizeof(T)));
    mg.pack_elems = pk.size();
    STZ_HIP(hipMemcpy(mg.pack, pk.data(), pk.size() * sizeof(T), hipMemcpyHostToDevice));
    mg.args.pack = mg.pack;
    return STZ_OK;
}

// the direct level of hierarchy mg: its mode inverses (FP64 in both
// hierarchies), or none
template <typename T> static int set_direct(DevMG<T> &mg, int ld, const std::vector<double> &G) {
    mg.args.ld = ld;
    mg.args.Gd = nullptr;
    if (ld < 0) return STZ_OK;
    if (mg.G && mg.G_elems != G.size()) {
        (void)hipFree(mg.G);
        mg.G = nullptr;
    }
    if (!mg.G) STZ_HIP(hipMalloc(&mg.G, G.size() * sizeof(double)));
    mg.G_elems = G.size();
    STZ_HIP(hipMemcpy(mg.G, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
    mg.args.Gd = mg.G;
    return mg_direct_set_lds_limit<T>(direct_lds_bytes(mg.args.lv[ld].m, mg.args.lv[ld].n));
}

// the finest level l in 1..lc with at most cfg.direct_nodes nodes, an
// x-invariant operator and a shape the direct launches cover (direct_fits)
// is solved exactly; the V-cycle stops there (both hierarchies)
static int set_direct_level(stz_sim *s) {
    int ld = -1;
    const int lc = s->mg64.args.lc;
    for (int l = 1; l <= lc && l < (int)s->hl.size() && s->cfg.direct_nodes > 0; l++) {
        const HostLevel &H = s->hl[l];
        if (H.xinv && (long long)H.m * H.n <= s->cfg.direct_nodes && direct_fits(H.m, H.n)) {
            ld = l;
            break;
        }
    }
    std::vector<double> G;
    if (ld >= 0 && !direct_inverse(s->hl[ld].S.data(), s->hl[ld].m, s->hl[ld].n, G)) {
        set_error("multigrid level %d (%d x %d) is singular", ld, s->hl[ld].m, s->hl[ld].n);
        return STZ_ERR_STATE;
    }
    int e;
    if ((e = set_direct(s->mg64, ld, G))) return e;
    if (s->precision == STZ_PREC_MIXED && (e = set_direct(s->mg32, ld, G))) return e;
    return STZ_OK;
}

// (re)build operators for viscosity/dt = viscdt (qs_multi::init, qs_multi.cc:31-54)
int mg_build(stz_sim *s, double viscdt) {
    if (viscdt == s->viscdt_prev) return STZ_OK;
    s->viscdt_prev = viscdt;
    top_stencil(s->c, viscdt, s->hl[0]);
    for (size_t l = 1; l < s->hl.size(); l++) coarsen(s->hl[l - 1], s->hl[l]);
    std::vector<double> binv;
    if (!bottom_inverse(s->hl.back(), binv)) {
        set_error("coarsest multigrid operator is singular");
        return STZ_ERR_STATE;
    }
    int e;
    std::vector<std::vector<double>> Sx(s->hl.size());
    for (size_t l = 0; l < s->hl.size(); l++) {
        HostLevel &H = s->hl[l];
        // rows at each wall whose stencil differs from the middle row
        H.nb = H.n;
        if (H.xinv && H.n >= 3) {
            const double *mid = &H.S[(size_t)(H.n / 2) * 40];
            int nb = 0;
            for (int j = 0; j < H.n / 2; j++) {
                const bool lo = std::memcmp(&H.S[(size_t)j * 40], mid, 40 * sizeof(double)) != 0;
                const bool hi = std::memcmp(&H.S[(size_t)(H.n - 1 - j) * 40], mid, 40 * sizeof(double)) != 0;
                if (lo || hi) nb = j + 1;
            }
            H.nb = nb;
        }
        // S holds the row stencils followed by one copy of the interior stencil
        Sx[l] = H.S;
        Sx[l].insert(Sx[l].end(), H.S.begin() + (size_t)(H.xinv ? H.n / 2 : 0) * 40,
                     H.S.begin() + (size_t)(H.xinv ? H.n / 2 : 0) * 40 + 40);
    }
    if ((e = set_level_stencils(s, s->mg64, Sx, binv))) return e;
    if (s->precision == STZ_PREC_MIXED && (e = set_level_stencils(s, s->mg32, Sx, binv))) return e;
    return set_direct_level(s);
}

// mixed form: FP64 residual of the top level -> FP32 source of the next
// correction equation (and zero correction for the per-colour path)
// FP64 residual -> FP32 source of the correction equation; with `norm` also
// its mean square into *dnorm (the host reads it only at convergence tests,
// so the other residuals skip the summing launch)
static int refine_resid(stz_sim *s, bool norm = true) {
    LevelArgs<double> t = s->top64();
    LevelArgs<float> &f = s->mg32.args.lv[0];
    if (s->cfg.kernels == 0 && wall_ok(t)) return mg_refine(t, f.b, s->partial, norm ? s->dnorm : nullptr, s->stream);
    return mg_resid64_to32(t, f.b, f.z, s->partial, s->dnorm, s->stream);
}

// the tiled V-cycle may leave the top FP64 level's result in its second
// buffer; move it back into vel so the public pointer stays valid
static int top_home(stz_sim *s) {
    LevelArgs<double> &t = s->mg64.args.lv[0];
    if (t.z != s->vel) {
        STZ_HIP(hipMemcpyAsync(s->vel, t.z, 2 * sizeof(double) * t.m * t.n, hipMemcpyDeviceToDevice, s->stream));
        t.z2 = t.z;
        t.z = s->vel;
    }
    return STZ_OK;
}

// one cycle: FP64 V-cycle on vel, or (mixed) FP32 V-cycle on the correction
// equation whose source the last residual evaluation wrote, then vel += e
// and a fresh FP64 residual (the next cycle's source)
static int one_cycle(stz_sim *s, bool need_resid, bool need_norm = true) {
    int e;
    const bool tiled = s->cfg.kernels == 0;
    if (s->precision == STZ_PREC_FP64) {
        CoarseArgs<double> &A = s->mg64.args;
        A.nu1 = s->cfg.nu1;
        A.nu2 = s->cfg.nu2;
        if (!tiled) return mg_vcycle(A, s->stream);
        if ((e = mg_vcycle_tiled(A, false, s->stream))) return e;
        return top_home(s);
    }
    CoarseArgs<float> &A = s->mg32.args;
    A.nu1 = s->cfg.nu1;
    A.nu2 = s->cfg.nu2;
    int corrected = 0;  // vel += e done by the top level's last tiled launch
    if (tiled) {
        if ((e = mg_vcycle_tiled(A, true, s->stream, s->vel, &corrected))) return e;
    } else {
        if ((e = mg_vcycle(A, s->stream))) return e;
    }
    if (!corrected && (e = mg_correct64(s->vel, A.lv[0].z, 2 * s->c.m * (s->c.n + 1), s->stream))) return e;
    if (++s->last.refinements > s->cfg.max_refine) {
        set_error("multigrid: more than %d FP64 refinements in one solve", s->cfg.max_refine);
        return STZ_ERR_NOT_CONVERGED;
    }
    if (need_resid) return refine_resid(s, need_norm);
    return STZ_OK;
}

// k cycles; with resid_after the residual after the last one, whose norm
// the convergence test reads
static int run_cycles(stz_sim *s, int k, bool resid_after) {
    int e;
    for (int i = 0; i < k; i++)
        if ((e = one_cycle(s, resid_after || i + 1 < k, resid_after && i + 1 == k))) return e;
    s->last.total_cycles += k;
    return STZ_OK;
}

// enqueue the FP64 form's convergence norm (the mixed form's comes with the
// residual of its last cycle) and mark where it ends
static int enqueue_test(stz_sim *s) {
    int e;
    if (s->precision == STZ_PREC_FP64 && (e = mg_resid_norm(s->top64(), s->partial, s->dnorm, s->stream))) return e;
    STZ_HIP(hipEventRecord(s->norm_ev, s->stream));
    return STZ_OK;
}

// wait for the norm enqueued by enqueue_test (work enqueued after it may
// still run) and read it
static int read_test(stz_sim *s, double *acc) {
    STZ_HIP(hipEventSynchronize(s->norm_ev));
    *acc = *s->hnorm / ((double)s->c.m * (s->c.n + 1));
    return STZ_OK;
}

// tgmg::solve(type=4, tgmg_predict&) (tgmg.in.cc:266-301).  Every test is
// followed by at least one more cycle: the extra cycle when it passes, the
// first of the next k when it fails.  So with extra_cycles >= 1 that cycle is
// enqueued before the host waits for the test's norm, and the device stays
// busy through the wait; the kernels and their order are those of the
// plain schedule, so the result is the same bit for bit.
static int mg_solve_impl(stz_sim *s) {
    int e;
    const stz_mg_config &g = s->cfg;
    const double tol = s->acc_base * g.tol_mult;
    const int max_thresh = g.predict_mult * g.predict_max_iters;
    const bool mixed = s->precision == STZ_PREC_MIXED, ahead = g.extra_cycles > 0;
    std::memset(&s->last, 0, sizeof s->last);
    s->last.tolerance = tol;
    double acc;
    int n = s->lim / g.predict_mult;
    if (mixed) {  // first source of the correction equation
        // (its norm is the first test's when no cycle comes before it)
        if ((e = refine_resid(s, n == 0))) return e;
        s->last.refinements = 0;
    }
    if ((e = run_cycles(s, n, true))) return e;
    if ((e = enqueue_test(s))) return e;
    // a test whose failure would end the solve (the predictor past its
    // limit) gets no cycle ahead, so a failed solve leaves the state the
    // plain schedule leaves
    bool spec = ahead && s->lim + g.predict_mult <= max_thresh;
    const bool unit = g.predict_unit != 0;
    if (spec) {
        if ((e = one_cycle(s, false))) return e;
        s->last.total_cycles++;
    }
    if ((e = read_test(s, &acc))) return e;
    if (acc < tol) {
        if (s->lim > 0) s->lim -= 1 + s->lim / g.predict_decay;
    } else {
        int k = 0;
        do {
            k = unit ? 1 : k + 1;
            s->lim += k * g.predict_mult;
            if (s->lim > max_thresh) {
                s->last.status = STZ_ERR_NOT_CONVERGED;
                s->last.cycles = n;
                s->last.residual = acc;
                set_error("multigrid: residual %g not below %g after %d V-cycles", acc, tol, n);
                return STZ_ERR_NOT_CONVERGED;
            }
            n += k;
            // the cycle enqueued ahead is the first of these k
            const int rest = spec ? k - 1 : k;
            if (spec && mixed && (e = refine_resid(s, rest == 0))) return e;
            if ((e = run_cycles(s, rest, true))) return e;
            if ((e = enqueue_test(s))) return e;
            spec = ahead && s->lim + (unit ? 1 : k + 1) * g.predict_mult <= max_thresh;
            if (spec) {
                if ((e = one_cycle(s, false))) return e;
                s->last.total_cycles++;
            }
            if ((e = read_test(s, &acc))) return e;
        } while (acc >= tol);
    }
    s->solves++;
    s->vcount += n;
    if ((e = run_cycles(s, spec ? g.extra_cycles - 1 : g.extra_cycles, false))) return e;
    s->last.cycles = n;
    s->last.residual = acc;
    s->last.status = std::isnan(acc) ? STZ_ERR_NAN : STZ_OK;
    if (s->last.status != STZ_OK) set_error("multigrid: NaN residual");
    return s->last.status;
}

static inline double bdry_vel(const stz_sim *s) { return s->time < 1 ? s->time * s->p.u_bdry : s->p.u_bdry; }
static inline double bdry_pos(const stz_sim *s) {
    return s->time < 1 ? s->p.u_bdry * s->time * s->time * 0.5 : s->p.u_bdry * (s->time - 0.5);
}

#define CHECK_SIM(s)                                                                                             \
    do {                                                                                                         \
        if (!(s)) { set_error("null simulation"); return STZ_ERR_ARG; }                                         \
        STZ_HIP(hipSetDevice((s)->device));                                                                      \
    } while (0)

static int advection_impl(stz_sim *s, double dt) {
    int e;
    const uint32_t *rep = nullptr;
    if (s->rep) {
        if (s->rep_next >= s->rep_steps) {
            set_error("decision replay: all %d recorded steps are used", s->rep_steps);
            return STZ_ERR_STATE;
        }
        rep = s->rep + (size_t)s->rep_next++ * s->c.m * s->c.n;
    }
    if ((e = launch_advect(s->c, s->cur, s->alt, dt, s->rec_dec ? s->dec : nullptr, s->stream, rep))) return e;
    if (s->c.chi_diff != 0 || s->c.chi_len != 0)
        if ((e = launch_chi_diffusion(s->c, s->cur, s->alt, dt, s->stream))) return e;
    std::swap(s->cur, s->alt);
    // u, v, ddev are single planes shared by both sets
    return STZ_OK;
}

static int projection_impl(stz_sim *s, double dt) {
    int e;
    if ((e = mg_build(s, s->p.viscosity / dt))) return e;
    const double cxx = s->hl[0].S[40 + 16];  // interior centre block, x row
    const double xbc = s->c.bcs * cxx;
    if ((e = launch_source(s->c, s->cur, dt, xbc, bdry_vel(s), s->src, s->stream))) return e;
    if ((e = mg_solve_impl(s))) return e;
    return launch_correct(s->c, s->cur, s->vel, dt, bdry_vel(s), bdry_pos(s), s->stream);
}

extern "C" {

const char *stz_last_error(void) { return g_err; }

int stz_device_count(int *count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return STZ_OK;
}

int stz_sim_create(const stz_sim_params *p, int precision, stz_sim **out) {
    if (!p || !out) { set_error("null argument"); return STZ_ERR_ARG; }
    *out = nullptr;
    if (p->m < 4 || p->n < 4) { set_error("need m >= 4 and n >= 4 (got %d x %d)", p->m, p->n); return STZ_ERR_ARG; }
    if (precision != STZ_PREC_FP64 && precision != STZ_PREC_MIXED) {
        set_error("unknown precision %d", precision);
        return STZ_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return STZ_ERR_NO_DEVICE;
    }
    int dev = 0;
    STZ_HIP(hipGetDevice(&dev));
    stz_sim *s = new stz_sim();
    s->device = dev;
    s->p = *p;
    s->precision = precision;
    fill_sim_const(p, s->c, &s->acc_base);
    const SimConst &c = s->c;
    set_defaults(s->cfg);
    s->lim = s->cfg.predict_mult * s->cfg.predict_init;

    int e = STZ_OK;
    auto fail = [&](int code) {
        stz_sim_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(STZ_ERR_HIP);
    }
    s->own_stream = true;
    const size_t pe = plane_elems(s);
    if (hipMalloc(&s->planes, pe * 20 * sizeof(double)) != hipSuccess) {
        set_error("out of device memory for %zu planes", pe * 20);
        return fail(STZ_ERR_HIP);
    }
    (void)hipMemset(s->planes, 0, pe * 20 * sizeof(double));
    double *P = s->planes;
    s->cur.u = s->alt.u = P + 0 * pe;
    s->cur.v = s->alt.v = P + 1 * pe;
    s->cur.ddev = s->alt.ddev = P + 2 * pe;
    double **cf[7] = {&s->cur.p, &s->cur.q, &s->cur.s, &s->cur.tau, &s->cur.chi, &s->cur.X, &s->cur.Y};
    double **af[7] = {&s->alt.p, &s->alt.q, &s->alt.s, &s->alt.tau, &s->alt.chi, &s->alt.X, &s->alt.Y};
    for (int k = 0; k < 7; k++) {
        *cf[k] = P + (3 + k) * pe;
        *af[k] = P + (10 + k) * pe;
    }
    s->cu = P + 17 * pe;
    s->un = P + 18 * pe;
    s->vn = P + 19 * pe;
    const size_t nodes = (size_t)c.m * (c.n + 1);
    if (hipMalloc(&s->vel, 4 * nodes * sizeof(double)) != hipSuccess) return fail(STZ_ERR_HIP);
    (void)hipMemset(s->vel, 0, 4 * nodes * sizeof(double));
    s->src = s->vel + 2 * nodes;
    if (hipMalloc(&s->dec, (size_t)c.m * c.n * sizeof(uint32_t)) != hipSuccess) return fail(STZ_ERR_HIP);
    (void)hipMemset(s->dec, 0, (size_t)c.m * c.n * sizeof(uint32_t));
    s->nblocks_norm = (int)std::max<size_t>((nodes + 255) / 256,
                                            (size_t)((c.m + 63) / 64) * ((c.n + 1 + 15) / 16));
    {
        const size_t np = std::max<size_t>(s->nblocks_norm + 1, 4 * (size_t)diag_partials_needed(c));
        if (hipMalloc(&s->partial, np * sizeof(double)) != hipSuccess) return fail(STZ_ERR_HIP);
    }
    // the residual norm (and the diagnostics' sums, from element 1 on) are
    // written by the summing kernels straight into coherent pinned host
    // memory: no copy launch before each convergence test
    if (hipHostMalloc(&s->hnorm, 8 * sizeof(double), hipHostMallocCoherent) != hipSuccess) return fail(STZ_ERR_HIP);
    if (hipEventCreateWithFlags(&s->norm_ev, hipEventDisableTiming) != hipSuccess) {
        s->norm_ev = nullptr;
        return fail(STZ_ERR_HIP);
    }
    s->dnorm = s->hnorm;

    build_dims(s);
    if ((e = alloc_levels(s, s->mg64, false))) return fail(e);
    s->mg64.args.lv[0].z = s->vel;
    s->mg64.args.lv[0].b = s->src;
    if (precision == STZ_PREC_MIXED)
        if ((e = alloc_levels(s, s->mg32, true))) return fail(e);
    if ((e = set_lds_attr(s->mg64))) return fail(e);
    if (precision == STZ_PREC_MIXED && (e = set_lds_attr(s->mg32))) return fail(e);
    *out = s;
    return STZ_OK;
}

int stz_sim_destroy(stz_sim *s) {
    if (!s) return STZ_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->mg64.free_all();
    s->mg32.free_all();
    if (s->planes) (void)hipFree(s->planes);
    if (s->vel) (void)hipFree(s->vel);
    if (s->dec) (void)hipFree(s->dec);
    if (s->rep) (void)hipFree(s->rep);
    if (s->partial) (void)hipFree(s->partial);
    if (s->hnorm) (void)hipHostFree(s->hnorm);
    if (s->norm_ev) (void)hipEventDestroy(s->norm_ev);
    if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return STZ_OK;
}

int stz_sim_set_stream(stz_sim *s, void *st) {
    CHECK_SIM(s);
    STZ_HIP(hipStreamSynchronize(s->stream));
    if (st) {
        if (s->own_stream) (void)hipStreamDestroy(s->stream);
        s->stream = (hipStream_t)st;
        s->own_stream = false;
    } else if (!s->own_stream) {
        STZ_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
    }
    return STZ_OK;
}

int stz_sim_synchronize(stz_sim *s) {
    CHECK_SIM(s);
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

static double *field_plane(stz_sim *s, int k) {
    FieldPtrs &f = s->cur;
    double *tab[STZ_NFIELDS + 1] = {f.u, f.v, f.p, f.q, f.s, f.tau, f.chi, f.X, f.Y, f.ddev, s->cu};
    return (k >= 0 && k <= STZ_F_CU) ? tab[k] : nullptr;
}

int stz_sim_upload_field(stz_sim *s, int k, const double *h) {
    CHECK_SIM(s);
    double *d = field_plane(s, k);
    if (!d || !h) { set_error("bad field %d or null pointer", k); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(d, h, plane_elems(s) * sizeof(double), hipMemcpyHostToDevice, s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_download_field(stz_sim *s, int k, double *h) {
    CHECK_SIM(s);
    double *d = field_plane(s, k);
    if (!d || !h) { set_error("bad field %d or null pointer", k); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, d, plane_elems(s) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_field_device_ptr(stz_sim *s, int k, double **dptr) {
    CHECK_SIM(s);
    double *d = field_plane(s, k);
    if (!d || !dptr) { set_error("bad field %d or null pointer", k); return STZ_ERR_ARG; }
    *dptr = d;
    return STZ_OK;
}

int stz_sim_upload_vel(stz_sim *s, const double *h) {
    CHECK_SIM(s);
    if (!h) { set_error("null pointer"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(s->vel, h, 2 * (size_t)s->c.m * (s->c.n + 1) * sizeof(double), hipMemcpyHostToDevice,
                           s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_download_vel(stz_sim *s, double *h) {
    CHECK_SIM(s);
    if (!h) { set_error("null pointer"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, s->vel, 2 * (size_t)s->c.m * (s->c.n + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_download_src(stz_sim *s, double *h) {
    CHECK_SIM(s);
    if (!h) { set_error("null pointer"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, s->src, 2 * (size_t)s->c.m * (s->c.n + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_vel_device_ptr(stz_sim *s, double **d) {
    CHECK_SIM(s);
    if (!d) { set_error("null pointer"); return STZ_ERR_ARG; }
    *d = s->vel;
    return STZ_OK;
}

int stz_sim_set_time(stz_sim *s, double t) {
    if (!s) { set_error("null simulation"); return STZ_ERR_ARG; }
    s->time = t;
    return STZ_OK;
}

int stz_sim_get_time(stz_sim *s, double *t) {
    if (!s || !t) { set_error("null argument"); return STZ_ERR_ARG; }
    *t = s->time;
    return STZ_OK;
}

int stz_sim_set_chi_inf(stz_sim *s, double chi_inf) {
    if (!s) { set_error("null simulation"); return STZ_ERR_ARG; }
    s->c.chi_inf = chi_inf;
    return STZ_OK;
}

int stz_sim_init_fields(stz_sim *s, int chi_case, double tem_base, double tem_delta) {
    CHECK_SIM(s);
    const double TZ = s->p.stz.TZ;
    int e = launch_init_fields(s->c, s->cur, chi_case, tem_base / TZ, tem_delta / TZ, s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_advection_step(stz_sim *s, double dt) {
    CHECK_SIM(s);
    int e = advection_impl(s, dt);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_advection_launch(stz_sim *s, double dt, int count) {
    CHECK_SIM(s);
    if (count < 0) { set_error("negative count"); return STZ_ERR_ARG; }
    int e;
    for (int k = 0; k < count; k++)
        if ((e = advection_impl(s, dt))) return e;
    return STZ_OK;
}

int stz_sim_projection_step(stz_sim *s, double dt) {
    CHECK_SIM(s);
    int e = projection_impl(s, dt);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_step_quasistatic(stz_sim *s, double dt, int nsteps) {
    CHECK_SIM(s);
    if (nsteps < 0) { set_error("nsteps < 0"); return STZ_ERR_ARG; }
    for (int k = 0; k < nsteps; k++) {
        s->time += dt;  // shear_sim.cc:283
        int e;
        if ((e = advection_impl(s, dt))) return e;
        if ((e = projection_impl(s, dt))) return e;
    }
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_set_boundaries(stz_sim *s) {
    CHECK_SIM(s);
    int e = launch_boundaries(s->c, s->cur, bdry_vel(s), bdry_pos(s), s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_get_predictor(stz_sim *s, int32_t *lim) {
    if (!s || !lim) { set_error("null argument"); return STZ_ERR_ARG; }
    *lim = s->lim;
    return STZ_OK;
}

int stz_sim_set_predictor(stz_sim *s, int32_t lim) {
    if (!s) { set_error("null simulation"); return STZ_ERR_ARG; }
    if (lim < 0) { set_error("predictor limit %d < 0", lim); return STZ_ERR_ARG; }
    s->lim = lim;
    return STZ_OK;
}

int stz_sim_get_mg_config(stz_sim *s, stz_mg_config *g) {
    if (!s || !g) { set_error("null argument"); return STZ_ERR_ARG; }
    *g = s->cfg;
    return STZ_OK;
}

int stz_sim_set_mg_config(stz_sim *s, const stz_mg_config *g) {
    if (!s || !g) { set_error("null argument"); return STZ_ERR_ARG; }
    if (g->nu1 < 0 || g->nu2 < 0 || g->nu1 + g->nu2 < 1 || g->extra_cycles < 0 || g->predict_mult < 1 ||
        g->predict_decay < 1 || g->predict_init < 0 || g->predict_max_iters < 1 || !(g->tol_mult > 0)) {
        set_error("invalid multigrid configuration");
        return STZ_ERR_ARG;
    }
    if (g->direct_nodes < 0 || g->max_refine < 1 || (g->predict_unit != 0 && g->predict_unit != 1)) {
        set_error("invalid multigrid configuration (direct_nodes %d, max_refine %d, predict_unit %d)",
                  g->direct_nodes, g->max_refine, g->predict_unit);
        return STZ_ERR_ARG;
    }
    const bool reset = g->predict_mult != s->cfg.predict_mult || g->predict_init != s->cfg.predict_init;
    const bool rebuild = g->direct_nodes != s->cfg.direct_nodes && s->viscdt_prev >= 0;
    s->cfg = *g;
    if (reset) s->lim = g->predict_mult * g->predict_init;
    if (rebuild) {  // a new direct level: the hierarchy is rebuilt now
        const double v = s->viscdt_prev;
        s->viscdt_prev = -1;
        CHECK_SIM(s);
        return mg_build(s, v);
    }
    return STZ_OK;
}

int stz_sim_last_solve(stz_sim *s, stz_solve_info *info) {
    if (!s || !info) { set_error("null argument"); return STZ_ERR_ARG; }
    *info = s->last;
    return STZ_OK;
}

int stz_sim_avg_iters(stz_sim *s, double *avg) {
    if (!s || !avg) { set_error("null argument"); return STZ_ERR_ARG; }
    // tgmg_predict.hh:28-33 (float division, counters reset)
    const float ans = static_cast<float>(s->vcount) / static_cast<float>(s->solves);
    s->solves = s->vcount = 0;
    *avg = ans;
    return STZ_OK;
}

int stz_sim_mg_hierarchy(stz_sim *s, int max_levels, int *nlevels, int *dims) {
    if (!s || !nlevels) { set_error("null argument"); return STZ_ERR_ARG; }
    *nlevels = (int)s->hl.size();
    for (int l = 0; l < (int)s->hl.size() && l < max_levels && dims; l++) {
        dims[2 * l] = s->hl[l].m;
        dims[2 * l + 1] = s->hl[l].n;
    }
    return STZ_OK;
}

int stz_mg_setup(stz_sim *s, double viscdt) {
    CHECK_SIM(s);
    return mg_build(s, viscdt);
}

int stz_mg_solve(stz_sim *s, stz_solve_info *info) {
    CHECK_SIM(s);
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    int e = mg_solve_impl(s);
    if (info) *info = s->last;
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_mg_v_cycle(stz_sim *s) {
    CHECK_SIM(s);
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    int e;
    if (s->precision == STZ_PREC_MIXED && (e = refine_resid(s))) return e;
    if ((e = one_cycle(s, false))) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_mg_l2_error(stz_sim *s, double *err) {
    CHECK_SIM(s);
    if (!err) { set_error("null pointer"); return STZ_ERR_ARG; }
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    int e = mg_resid_norm(s->top64(), s->partial, s->dnorm, s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    *err = *s->hnorm / ((double)s->c.m * (s->c.n + 1));
    return STZ_OK;
}

int stz_mg_smooth(stz_sim *s, int level, int sweeps) {
    CHECK_SIM(s);
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    if (level < 0 || level >= (int)s->hl.size() || sweeps < 0) {
        set_error("bad level %d or sweeps %d", level, sweeps);
        return STZ_ERR_ARG;
    }
    int e;
    if (s->precision == STZ_PREC_MIXED) {
        LevelArgs<float> &L = s->mg32.args.lv[level];
        return s->cfg.kernels == 0 ? mg_smooth_level(L, sweeps, s->stream) : mg_smooth(L, sweeps, s->stream);
    }
    LevelArgs<double> &L = s->mg64.args.lv[level];
    if (s->cfg.kernels != 0) return mg_smooth(L, sweeps, s->stream);
    if ((e = mg_smooth_level(L, sweeps, s->stream))) return e;
    return level == 0 ? top_home(s) : STZ_OK;
}

int stz_mg_vcycle_launch(stz_sim *s, int level, int mode, int count) {
    CHECK_SIM(s);
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    if (s->cfg.kernels != 0) { set_error("the per-colour kernels have no tiled launches"); return STZ_ERR_ARG; }
    if (count < 0) { set_error("negative count"); return STZ_ERR_ARG; }
    int e;
    for (int k = 0; k < count; k++) {
        if (s->precision == STZ_PREC_MIXED) {
            CoarseArgs<float> &A = s->mg32.args;
            A.nu1 = s->cfg.nu1;
            A.nu2 = s->cfg.nu2;
            if ((e = mg_vcycle_launch(A, level, mode, true, s->vel, s->stream))) return e;
        } else {
            CoarseArgs<double> &A = s->mg64.args;
            A.nu1 = s->cfg.nu1;
            A.nu2 = s->cfg.nu2;
            if ((e = mg_vcycle_launch(A, level, mode, false, nullptr, s->stream))) return e;
        }
    }
    return s->precision == STZ_PREC_FP64 && level == 0 ? top_home(s) : STZ_OK;
}

int stz_mg_fmg_pre(stz_sim *s) {
    CHECK_SIM(s);
    if (s->viscdt_prev < 0) { set_error("stz_mg_setup has not been called"); return STZ_ERR_STATE; }
    int e;
    if ((e = top_home(s))) return e;
    CoarseArgs<double> &A = s->mg64.args;
    A.nu1 = s->cfg.nu1;
    A.nu2 = s->cfg.nu2;
    if ((e = mg_fmg_pre(A, s->stream))) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_mg_clear_z(stz_sim *s) {
    CHECK_SIM(s);
    int e;
    if ((e = top_home(s))) return e;
    if ((e = mg_fill_zero(s->vel, 2 * (size_t)s->c.m * (s->c.n + 1), s->stream))) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_mg_upload_src(stz_sim *s, const double *h) {
    CHECK_SIM(s);
    if (!h) { set_error("null pointer"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(s->src, h, 2 * (size_t)s->c.m * (s->c.n + 1) * sizeof(double), hipMemcpyHostToDevice,
                           s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

// ---- direct step, diagnostics, initial conditions --------------------------

int stz_sim_step_forward(stz_sim *s, double dt, int nsteps) {
    CHECK_SIM(s);
    if (nsteps < 0) { set_error("nsteps < 0"); return STZ_ERR_ARG; }
    const SimConst &c = s->c;
    const size_t row = (size_t)c.m;
    int e;
    for (int k = 0; k < nsteps; k++) {
        s->time += dt;  // shear_sim.cc:192
        if ((e = launch_direct(c, s->cur, s->alt, s->un, s->vn, dt, s->stream))) return e;
        if (c.chi_diff != 0 || c.chi_len != 0)
            if ((e = launch_chi_diffusion(c, s->cur, s->alt, dt, s->stream))) return e;
        std::swap(s->cur, s->alt);
        // new u, v of rows 1..n-1 (plane rows 2..n)
        STZ_HIP(hipMemcpyAsync(s->cur.u + 2 * row, s->un + 2 * row, (c.n - 1) * row * sizeof(double),
                               hipMemcpyDeviceToDevice, s->stream));
        STZ_HIP(hipMemcpyAsync(s->cur.v + 2 * row, s->vn + 2 * row, (c.n - 1) * row * sizeof(double),
                               hipMemcpyDeviceToDevice, s->stream));
        if ((e = launch_boundaries(c, s->cur, bdry_vel(s), bdry_pos(s), s->stream))) return e;
    }
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_tractions(stz_sim *s, double *out4) {
    CHECK_SIM(s);
    if (!out4) { set_error("null pointer"); return STZ_ERR_ARG; }
    int e = diag_tractions(s->c, s->cur, s->hnorm + 1, s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    for (int k = 0; k < 4; k++) out4[k] = s->hnorm[1 + k];
    return STZ_OK;
}

int stz_sim_qs_measure(stz_sim *s, int staggered, double *Q, double *max_qs) {
    CHECK_SIM(s);
    if (!Q || !max_qs) { set_error("null pointer"); return STZ_ERR_ARG; }
    int e = diag_qs(s->c, s->cur, s->vel, staggered != 0, s->cu, s->partial, s->hnorm + 1, s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    *Q = std::sqrt(s->hnorm[1] * s->c.dx * s->c.dy);  // shear_sim.cc:775
    *max_qs = s->hnorm[2];
    return STZ_OK;
}

int stz_sim_compute_strain(stz_sim *s, int corrected) {
    CHECK_SIM(s);
    int e = diag_strain(s->c, s->cur, corrected != 0, s->cu, s->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

int stz_sim_l2_comparison(stz_sim *a, stz_sim *b, double *out4) {
    CHECK_SIM(a);
    if (!b || !out4) { set_error("null argument"); return STZ_ERR_ARG; }
    if (a->c.m != b->c.m || a->c.n != b->c.n || a->device != b->device) {
        set_error("l2_comparison needs two simulations of the same size on one device");
        return STZ_ERR_ARG;
    }
    STZ_HIP(hipStreamSynchronize(b->stream));
    int e = diag_l2(a->c, a->cur, b->cur, a->partial, a->hnorm + 1, a->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(a->stream));
    const double fac = a->c.dx * a->c.dy;  // shear_sim.cc:500-501
    for (int k = 0; k < 4; k++) out4[k] = a->hnorm[1 + k] * fac;
    return STZ_OK;
}

// box_muller, shear_sim.cc:987-998, on the C library's rand() stream
static void box_muller(double &r0, double &r1) {
    double x, y, r;
    do {
        x = 2. / RAND_MAX * rand() - 1;
        y = 2. / RAND_MAX * rand() - 1;
        r = x * x + y * y;
    } while (r == 0.0 || r > 1.0);
    const double d = std::sqrt(-2.0 * std::log(r) / r);
    r0 = x * d;
    r1 = y * d;
}

int stz_sim_initialize_random(stz_sim *s, double l, double tem_avg, double tem_stdev) {
    CHECK_SIM(s);
    if (!(l > 0)) { set_error("smoothing length must be positive"); return STZ_ERR_ARG; }
    const SimConst &c = s->c;
    const int m = c.m, n = c.n, cut = int(l * 5.), nx = n + (cut << 1);
    const double llinv = 0.5 / (l * l), TZ = s->p.stz.TZ;
    // gaussian_normal_factor, shear_sim.cc:918-929
    double vfac = 1;
    for (int j = 1; j < cut; j++) {
        const double q = std::exp(-j * j * llinv);
        for (int k = 0; k <= j; k++) {
            const double q2 = q * std::exp(-k * k * llinv);
            vfac += (k == 0 || k == j ? 4 : 8) * q2 * q2;
        }
    }
    const double chi0 = tem_avg / TZ, chi1 = tem_stdev / (std::sqrt(vfac) * TZ);
    // the noise grid, drawn in the reference's order (shear_sim.cc:944-955):
    // the primary rows first, then the padding, outward from them
    std::vector<double> no((size_t)m * nx);
    double *np = no.data() + (size_t)cut * m, *ne = np + (size_t)m * n, *np2 = np;
    for (; np < ne - 1; np += 2) box_muller(*np, np[1]);
    if (np == ne - 1) {
        double r1;
        box_muller(*np, r1);
    }
    while (np < no.data() + (size_t)m * nx) {
        np2--;
        box_muller(*np, *np2);
        np++;
    }
    std::vector<double> w(cut + 1);
    for (int k = 0; k <= cut; k++) w[k] = std::exp(-llinv * (k * k));
    double *d = nullptr;
    const size_t tot = 2 * (size_t)m * nx + w.size();
    STZ_HIP(hipMalloc(&d, tot * sizeof(double)));
    int e = STZ_OK;
    if (hipMemcpyAsync(d, no.data(), no.size() * sizeof(double), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(d + 2 * (size_t)m * nx, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice,
                       s->stream) != hipSuccess) {
        set_error("initialize_random: upload failed");
        e = STZ_ERR_HIP;
    }
    if (!e) e = diag_random_fields(c, s->cur, cut, d + 2 * (size_t)m * nx, d, d + (size_t)m * nx, chi0, chi1,
                                   s->stream);
    const hipError_t se = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e) return e;
    STZ_HIP(se);
    return STZ_OK;
}

int stz_sim_initialize_chi_bicubic(stz_sim *s, int mm, int nn, const double *arr, double lo, double sca) {
    CHECK_SIM(s);
    if (!arr || mm < 1 || nn < 2) { set_error("bad array %d x %d", mm, nn); return STZ_ERR_ARG; }
    const int mmnn = mm * nn;
    // x, y and xy differences, shear_sim.cc:836-865
    std::vector<double> h(4 * (size_t)mmnn);
    double *a = h.data(), *xd = a + mmnn, *yd = xd + mmnn, *xyd = yd + mmnn;
    std::memcpy(a, arr, mmnn * sizeof(double));
    for (int j = 0; j < nn; j++)
        for (int i = 0, ij = j * mm; i < mm; i++, ij++)
            xd[ij] = 0.5 * (arr[i == mm - 1 ? ij - mm + 1 : ij + 1] - arr[i == 0 ? ij + mm - 1 : ij - 1]);
    for (int j = 0; j < nn; j++)
        for (int i = 0, ij = j * mm; i < mm; i++, ij++) {
            if (j == 0) {
                yd[ij] = arr[ij + mm] - arr[ij];
                xyd[ij] = xd[ij + mm] - xd[ij];
            } else if (j == nn - 1) {
                yd[ij] = arr[ij] - arr[ij - mm];
                xyd[ij] = xd[ij] - xd[ij - mm];
            } else {
                yd[ij] = 0.5 * (arr[ij + mm] - arr[ij - mm]);
                xyd[ij] = 0.5 * (xd[ij + mm] - xd[ij - mm]);
            }
        }
    double *d = nullptr;
    STZ_HIP(hipMalloc(&d, h.size() * sizeof(double)));
    int e = STZ_OK;
    if (hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s->stream) != hipSuccess) {
        set_error("initialize_chi_bicubic: upload failed");
        e = STZ_ERR_HIP;
    }
    if (!e) e = diag_bicubic(s->c, s->cur, mm, nn, d, lo, sca, s->stream);
    const hipError_t se = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e) return e;
    STZ_HIP(se);
    return STZ_OK;
}

int stz_sim_record_decisions(stz_sim *s, int enable) {
    if (!s) { set_error("null simulation"); return STZ_ERR_ARG; }
    s->rec_dec = enable != 0;
    return STZ_OK;
}

int stz_sim_replay_decisions(stz_sim *s, int nsteps, const uint32_t *h) {
    CHECK_SIM(s);
    if (nsteps < 0 || (nsteps > 0 && !h)) { set_error("need nsteps >= 0 and a record"); return STZ_ERR_ARG; }
    STZ_HIP(hipStreamSynchronize(s->stream));  // a step in flight may still read the old record
    if (s->rep) (void)hipFree(s->rep);
    s->rep = nullptr;
    s->rep_steps = s->rep_next = 0;
    if (nsteps == 0) return STZ_OK;
    const size_t bytes = (size_t)nsteps * s->c.m * s->c.n * sizeof(uint32_t);
    if (hipMalloc(&s->rep, bytes) != hipSuccess) {
        s->rep = nullptr;
        set_error("out of device memory for a %d-step decision record", nsteps);
        return STZ_ERR_HIP;
    }
    STZ_HIP(hipMemcpy(s->rep, h, bytes, hipMemcpyHostToDevice));
    s->rep_steps = nsteps;
    return STZ_OK;
}

int stz_sim_download_decisions(stz_sim *s, uint32_t *h) {
    CHECK_SIM(s);
    if (!h) { set_error("null pointer"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, s->dec, (size_t)s->c.m * s->c.n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s->stream));
    STZ_HIP(hipStreamSynchronize(s->stream));
    return STZ_OK;
}

}  // extern "C"
