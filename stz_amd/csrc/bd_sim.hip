// bd_sim.hip -- C ABI of the free-boundary simulation (include/stz_amd.h,
// stz_bd_*): the state of bd_sim (sb/main/bd_sim.hh:27-190) in HBM and the
// order of its quasi-static step (bd_sim.cc:240-259, 374-545).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bd.h"
#include "bd_mg.h"

using namespace stz;

struct stz_bd {
    stz_bd_params p{};
    BdConst c{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *mem = nullptr;          // all FP64 planes
    double *cell[2][6] = {};        // p q s tau chi ddev, two sets
    double *node[2][4] = {};        // u v X Y, two sets
    int cc = 0, nc = 0;             // current set of each
    double *phi = nullptr, *vp = nullptr, *tmp = nullptr;
    double *vel = nullptr, *src = nullptr, *force = nullptr;
    double *nkey = nullptr;         // node extrapolation keys (3 doubles per node)
    int8_t *st = nullptr, *nst = nullptr;
    int *rowt = nullptr, *done_c = nullptr, *done_n = nullptr, *counters = nullptr, *xlist = nullptr;
    long xspan = 0;                 // max(cells, nodes): the extrapolation lists hold 6 xspan ints
    double *fmats = nullptr;
    int fcap = 0;
    unsigned long long *dmin = nullptr;
    int *hint = nullptr;            // coherent pinned: flags and counts
    BdSolver mg;
    double time = 0;
    double march_delta = 0;
    double acc = 1e-20;             // qs_multi_bd.cc:16
    int max_iter = 500;
    bool have_band = false;
    stz_solve_info last{};
    int march_rounds = 0, waves_c = 0, waves_n = 0;

    BdCells cells(int set) const {
        return BdCells{cell[set][0], cell[set][1], cell[set][2], cell[set][3], cell[set][4], cell[set][5]};
    }
    BdNodes nodes(int set) const { return BdNodes{node[set][0], node[set][1], node[set][2], node[set][3]}; }
};

namespace {

#define CHECK_BD(b)                                                                                              \
    do {                                                                                                         \
        if (!(b)) { set_error("null simulation"); return STZ_ERR_ARG; }                                         \
        STZ_HIP(hipSetDevice((b)->device));                                                                      \
    } while (0)

int read_int(stz_bd *b, const int *dev, int *out) {
    STZ_HIP(hipMemcpyAsync(b->hint + 4, dev, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    *out = b->hint[4];
    return STZ_OK;
}

// levelset::build_band (levelset.cc:209-345) + set_staggered_status
int build_band(stz_bd *b) {
    const BdConst &c = b->c;
    int e;
    STZ_HIP(hipMemsetAsync(b->counters, 0, 8 * sizeof(int), b->stream));
    if ((e = bd_ls_scan(c, b->phi, b->st, b->stream))) return e;
    if ((e = bd_ls_newton_distance(c, b->phi, b->st, b->tmp, b->counters + 7, b->stream))) return e;
    int fails = 0;
    if ((e = read_int(b, b->counters + 7, &fails))) return e;
    if (fails) {
        set_error("Newton iteration failed at %d interface cells (levelset.cc:285)", fails);
        return STZ_ERR_STATE;
    }
    if ((e = bd_ls_march(c, b->phi, b->st, b->march_delta, b->dmin, b->counters, (volatile int *)b->hint,
                         &b->march_rounds, b->stream)))
        return e;
    if ((e = bd_ls_finish(c, b->phi, b->vp, b->st, 1, b->stream))) return e;
    if ((e = bd_ls_node_status(c, b->phi, b->st, b->nst, b->stream))) return e;
    b->have_band = true;
    return STZ_OK;
}

// levelset::move_and_smooth + bd_sim::set_staggered_status
// (bd_sim.cc:379-381)
int move_level_set(stz_bd *b, double dt) {
    const BdConst &c = b->c;
    int e;
    if ((e = bd_ls_interface_velocity(c, b->phi, b->st, b->nodes(b->nc), b->vp, b->stream))) return e;
    if ((e = bd_ls_advance(c, b->phi, b->vp, b->st, dt, b->stream))) return e;
    if ((e = bd_ls_march(c, b->phi, b->st, b->march_delta, b->dmin, b->counters, (volatile int *)b->hint,
                         &b->march_rounds, b->stream)))
        return e;
    if ((e = bd_ls_finish(c, b->phi, b->vp, b->st, 0, b->stream))) return e;
    return bd_ls_node_status(c, b->phi, b->st, b->nst, b->stream);
}

int advection(stz_bd *b, double dt) {
    int e;
    if ((e = move_level_set(b, dt))) return e;
    const int oc = b->cc, on = b->nc, wc = 1 - oc, wn = 1 - on;
    if ((e = bd_advect(b->c, b->st, b->nst, b->cells(oc), b->nodes(on), b->cells(wc), b->nodes(wn), dt,
                       b->stream)))
        return e;
    if (b->c.chi_len > 0 && (e = bd_chi_diffusion(b->c, b->st, b->cells(oc), b->cells(wc), dt, b->stream)))
        return e;
    b->cc = wc;
    b->nc = wn;
    return STZ_OK;
}

// bd_sim::post_process (bd_sim.cc:256-259)
// the cell and (all) node extrapolations; the two walks are independent
// and share their wave launches
static int extrapolate(stz_bd *b, bool all) {
    int e, fails = 0;
    if (all)
        e = bd_extrap_both(b->c, b->phi, b->st, b->nst, b->cells(b->cc), b->nodes(b->nc), b->done_c, b->done_n,
                           b->nkey, b->counters, b->counters + 16, b->xlist, b->xlist + 3 * b->xspan,
                           (volatile int *)b->hint, &b->waves_c, &b->waves_n, b->stream);
    else
        e = bd_extrap_cells(b->c, b->phi, b->st, b->cells(b->cc), b->done_c, b->counters, b->xlist,
                            (volatile int *)b->hint, &b->waves_c, b->stream);
    if (e) return e;
    if ((e = read_int(b, b->counters + 2, &fails))) return e;
    if (fails) {
        set_error("cell extrapolation: singular fit at %d cells (traverse.hh:214-230)", fails);
        return STZ_ERR_STATE;
    }
    if (!all) return STZ_OK;
    if ((e = read_int(b, b->counters + 16 + 2, &fails))) return e;
    if (fails) {
        set_error("node extrapolation: singular fit at %d nodes (traverse.hh:153-164)", fails);
        return STZ_ERR_STATE;
    }
    return STZ_OK;
}

int post_process(stz_bd *b, bool all) {
    const int e = extrapolate(b, all);
    return e ? e : bd_boundaries(b->c, b->nodes(b->nc), 0, b->stream);
}

int grow_free_rows(stz_bd *b, int need) {
    if (b->fmats) (void)hipFree(b->fmats);
    b->fcap = std::max(need + need / 2, 2 * b->fcap);
    STZ_HIP(hipMalloc(&b->fmats, 32 * (size_t)b->fcap * sizeof(double)));
    return STZ_OK;
}

// bd_sim::projection_step (bd_sim.cc:510-545)
int projection(stz_bd *b, double dt) {
    const BdConst &c = b->c;
    int e, nfree = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
        STZ_HIP(hipMemsetAsync(b->counters + 6, 0, sizeof(int), b->stream));
        if ((e = bd_operator(c, b->phi, b->st, b->nst, b->cells(b->cc), b->nodes(b->nc), b->force, dt, b->rowt,
                             b->src, b->fmats, b->fcap, b->counters + 6, b->stream)))
            return e;
        if ((e = read_int(b, b->counters + 6, &nfree))) return e;
        if (nfree <= b->fcap) break;
        if ((e = grow_free_rows(b, nfree))) return e;
    }
    const BdOp op = bd_op_constants(c, dt);
    BdTop t;
    t.m = c.me;
    t.n = c.ne;
    t.rowt = b->rowt;
    t.fmats = b->fmats;
    t.Kmx = op.Kmx;
    t.Kmy = op.Kmy;
    t.mx = op.mx;
    t.my = op.my;
    t.qK = op.qK;
    t.cxx = op.cxx;
    t.cyy = op.cyy;
    t.bcs = op.bcs;
    t.ocs = op.ocs;
    t.ncs = op.ncs;
    if ((e = b->mg.setup(t, b->stream))) return e;
    if ((e = b->mg.solve(b->vel, b->src, b->acc, b->max_iter, b->last, b->stream))) return e;
    if (b->last.status == STZ_ERR_NAN) {
        set_error("free-boundary solve: residual is NaN");
        return STZ_ERR_NAN;
    }
    if (b->last.status != STZ_OK) {
        set_error("free-boundary solve did not reach %g in %d iterations (residual %g)", b->acc, b->max_iter,
                  b->last.residual);
        return STZ_ERR_NOT_CONVERGED;
    }
    return bd_correct(c, b->st, b->cells(b->cc), b->nodes(b->nc), b->vel, dt, b->stream);
}

}  // namespace

extern "C" {

int stz_bd_create(const stz_bd_params *p, stz_bd **out) {
    if (!p || !out) { set_error("null argument"); return STZ_ERR_ARG; }
    *out = nullptr;
    if (p->m < 8 || p->n < 8) { set_error("need m >= 8 and n >= 8 (got %d x %d)", p->m, p->n); return STZ_ERR_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return STZ_ERR_NO_DEVICE;
    }
    int dev = 0;
    STZ_HIP(hipGetDevice(&dev));
    stz_bd *b = new stz_bd();
    b->device = dev;
    b->p = *p;
    BdConst &c = b->c;
    // bd_sim.cc:25-53
    c.m = p->m;
    c.n = p->n;
    c.me = p->m + 1;
    c.ne = p->n + 1;
    c.ax = p->ax; c.bx = p->bx; c.ay = p->ay; c.by = p->by;
    c.dx = (p->bx - p->ax) / p->m;
    c.dy = (p->by - p->ay) / p->n;
    c.xsp = 1 / c.dx;
    c.ysp = 1 / c.dy;
    c.mu = p->mu;
    c.mu_inv = 1 / p->mu;
    c.K = p->K;
    c.viscosity = p->viscosity;
    c.chi_len = p->chi_len;
    c.filter_stress = p->filter_stress;
    c.t_scale = p->t_scale;
    c.tmult = p->tmult;
    c.adapt_fac = p->adapt_fac;
    c.stz = p->stz;
    c.TZ = p->stz.TZ;
    c.chi_inf = p->stz.chi_inf;
    c.thetainv = 1 / p->stz.theta;
    c.bc_smooth = 0;
    c.lax = c.ax + 0.5 * c.dx;
    c.lay = c.ay + 0.5 * c.dy;
    c.phi_lim = 5 * c.dx;
    b->march_delta = 0.3 * std::min(c.dx, c.dy);

    auto fail = [&](int code) {
        stz_bd_destroy(b);
        return code;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(STZ_ERR_HIP);
    }
    b->own_stream = true;
    const size_t mn = (size_t)c.m * c.n, mne = (size_t)c.me * c.ne;
    const size_t doubles = 12 * mn + 3 * mn + 8 * mne + 6 * mne + 3 * mne;
    if (hipMalloc(&b->mem, doubles * sizeof(double)) != hipSuccess) {
        set_error("out of device memory (%zu doubles)", doubles);
        return fail(STZ_ERR_HIP);
    }
    (void)hipMemset(b->mem, 0, doubles * sizeof(double));
    double *P = b->mem;
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 6; k++) b->cell[s][k] = P + (6 * s + k) * mn;
    P += 12 * mn;
    b->phi = P;
    b->vp = P + mn;
    b->tmp = P + 2 * mn;
    P += 3 * mn;
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 4; k++) b->node[s][k] = P + (4 * s + k) * mne;
    P += 8 * mne;
    b->vel = P;
    b->src = P + 2 * mne;
    b->force = nullptr;
    P += 6 * mne;  // vel, src, spare
    b->nkey = P;
    b->xspan = std::max(mn, mne);
    if (hipMalloc(&b->st, mn) != hipSuccess || hipMalloc(&b->nst, mne) != hipSuccess ||
        hipMalloc(&b->rowt, mne * sizeof(int)) != hipSuccess || hipMalloc(&b->done_c, mn * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->done_n, mne * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->xlist, 6 * std::max(mn, mne) * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->counters, 32 * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->dmin, sizeof(unsigned long long)) != hipSuccess) {
        set_error("out of device memory");
        return fail(STZ_ERR_HIP);
    }
    (void)hipMemset(b->st, 0, mn);
    (void)hipMemset(b->nst, 0, mne);
    (void)hipMemset(b->counters, 0, 32 * sizeof(int));
    if (hipHostMalloc((void **)&b->hint, 16 * sizeof(int), hipHostMallocCoherent) != hipSuccess) {
        set_error("hipHostMalloc failed");
        return fail(STZ_ERR_HIP);
    }
    b->fcap = 0;
    int e;
    if ((e = grow_free_rows(b, 8 * (c.me + c.ne)))) return fail(e);
    if ((e = b->mg.init(c.me, c.ne))) return fail(e);
    *out = b;
    return STZ_OK;
}

int stz_bd_destroy(stz_bd *b) {
    if (!b) return STZ_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->mg.free_all();
    if (b->mem) (void)hipFree(b->mem);
    if (b->force) (void)hipFree(b->force);
    if (b->st) (void)hipFree(b->st);
    if (b->nst) (void)hipFree(b->nst);
    if (b->rowt) (void)hipFree(b->rowt);
    if (b->done_c) (void)hipFree(b->done_c);
    if (b->done_n) (void)hipFree(b->done_n);
    if (b->xlist) (void)hipFree(b->xlist);
    if (b->counters) (void)hipFree(b->counters);
    if (b->dmin) (void)hipFree(b->dmin);
    if (b->fmats) (void)hipFree(b->fmats);
    if (b->hint) (void)hipHostFree(b->hint);
    if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return STZ_OK;
}

int stz_bd_set_stream(stz_bd *b, void *st) {
    CHECK_BD(b);
    STZ_HIP(hipStreamSynchronize(b->stream));
    if (st) {
        if (b->own_stream) (void)hipStreamDestroy(b->stream);
        b->stream = (hipStream_t)st;
        b->own_stream = false;
    } else if (!b->own_stream) {
        STZ_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->own_stream = true;
    }
    return STZ_OK;
}

int stz_bd_synchronize(stz_bd *b) {
    CHECK_BD(b);
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_set_level_set(stz_bd *b, const double *phi) {
    CHECK_BD(b);
    if (!phi) { set_error("null phi"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(b->phi, phi, (size_t)b->c.m * b->c.n * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
    int e = build_band(b);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_upload_level_set(stz_bd *b, const double *phi, const int32_t *status) {
    CHECK_BD(b);
    if (!phi || !status) { set_error("null argument"); return STZ_ERR_ARG; }
    const size_t mn = (size_t)b->c.m * b->c.n;
    std::vector<int8_t> t(mn);
    for (size_t q = 0; q < mn; q++) {
        if (status[q] < 0 || status[q] > 7) { set_error("cell status %d at %zu out of 0..7", status[q], q); return STZ_ERR_ARG; }
        t[q] = (int8_t)status[q];
    }
    STZ_HIP(hipMemcpyAsync(b->phi, phi, mn * sizeof(double), hipMemcpyHostToDevice, b->stream));
    STZ_HIP(hipMemcpyAsync(b->st, t.data(), mn, hipMemcpyHostToDevice, b->stream));
    int e = bd_ls_node_status(b->c, b->phi, b->st, b->nst, b->stream);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    b->have_band = true;
    return STZ_OK;
}

int stz_bd_interface_velocity(stz_bd *b, double *vp) {
    CHECK_BD(b);
    if (!vp) { set_error("null argument"); return STZ_ERR_ARG; }
    const size_t mn = (size_t)b->c.m * b->c.n;
    STZ_HIP(hipMemsetAsync(b->tmp, 0, mn * sizeof(double), b->stream));
    int e = bd_ls_interface_velocity(b->c, b->phi, b->st, b->nodes(b->nc), b->tmp, b->stream);
    if (e) return e;
    STZ_HIP(hipMemcpyAsync(vp, b->tmp, mn * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_remarch(stz_bd *b) {
    CHECK_BD(b);
    if (!b->have_band) { set_error("no level set: call stz_bd_set_level_set first"); return STZ_ERR_STATE; }
    const BdConst &c = b->c;
    int e;
    STZ_HIP(hipMemsetAsync(b->vp, 0, (size_t)c.m * c.n * sizeof(double), b->stream));
    if ((e = bd_ls_advance(c, b->phi, b->vp, b->st, 0.0, b->stream))) return e;
    if ((e = bd_ls_march(c, b->phi, b->st, b->march_delta, b->dmin, b->counters, (volatile int *)b->hint,
                         &b->march_rounds, b->stream)))
        return e;
    if ((e = bd_ls_finish(c, b->phi, b->vp, b->st, 0, b->stream))) return e;
    if ((e = bd_ls_node_status(c, b->phi, b->st, b->nst, b->stream))) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_set_solver(stz_bd *b, int nu1, int nu2, int bottom_sweeps, int max_iter, double tol) {
    CHECK_BD(b);
    if (nu1 < 0 || nu2 < 0 || nu1 + nu2 < 1 || bottom_sweeps < 1 || max_iter < 1 || !(tol > 0)) {
        set_error("bad solver settings");
        return STZ_ERR_ARG;
    }
    b->mg.nu1 = nu1;
    b->mg.nu2 = nu2;
    b->mg.bottom_half = (bottom_sweeps + 1) / 2;
    b->max_iter = max_iter;
    b->acc = tol;
    return STZ_OK;
}

int stz_bd_set_precision(stz_bd *b, int precision) {
    CHECK_BD(b);
    if (precision != STZ_PREC_FP64 && precision != STZ_PREC_MIXED) {
        set_error("unknown precision %d", precision);
        return STZ_ERR_ARG;
    }
    STZ_HIP(hipStreamSynchronize(b->stream));
    return b->mg.set_mixed(precision == STZ_PREC_MIXED);
}

int stz_bd_step_stats(stz_bd *b, int32_t *out4) {
    CHECK_BD(b);
    if (!out4) { set_error("null argument"); return STZ_ERR_ARG; }
    out4[0] = b->march_rounds;
    out4[1] = b->waves_c;
    out4[2] = b->waves_n;
    out4[3] = b->last.total_cycles / 2;
    return STZ_OK;
}

int stz_bd_operator(stz_bd *b, double dt, int32_t *rowt, double *blocks36, double *src) {
    CHECK_BD(b);
    const BdConst &c = b->c;
    int e, nfree = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
        STZ_HIP(hipMemsetAsync(b->counters + 6, 0, sizeof(int), b->stream));
        if ((e = bd_operator(c, b->phi, b->st, b->nst, b->cells(b->cc), b->nodes(b->nc), b->force, dt, b->rowt,
                             b->src, b->fmats, b->fcap, b->counters + 6, b->stream)))
            return e;
        if ((e = read_int(b, b->counters + 6, &nfree))) return e;
        if (nfree <= b->fcap) break;
        if ((e = grow_free_rows(b, nfree))) return e;
    }
    const size_t mne = (size_t)c.me * c.ne;
    std::vector<int> rt(mne);
    std::vector<double> fm(32 * (size_t)nfree);
    STZ_HIP(hipMemcpy(rt.data(), b->rowt, mne * sizeof(int), hipMemcpyDeviceToHost));
    if (nfree) STZ_HIP(hipMemcpy(fm.data(), b->fmats, fm.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (src) STZ_HIP(hipMemcpy(src, b->src, 2 * mne * sizeof(double), hipMemcpyDeviceToHost));
    const BdOp op = bd_op_constants(c, dt);
    for (size_t ij = 0; ij < mne; ij++) {
        const int d = rt[ij];
        if (rowt) rowt[ij] = d >= 0 ? 0 : d;
        if (!blocks36) continue;
        double *o = blocks36 + 36 * ij;
        std::fill(o, o + 36, 0.0);
        if (d == -1) {
            const double rc[4] = {0, op.qK, op.qK, 0}, mrc[4] = {0, -op.qK, -op.qK, 0};
            const double rv[4] = {op.mx, 0, 0, op.Kmy}, rh[4] = {op.Kmx, 0, 0, op.my}, ru[4] = {op.cxx, 0, 0, op.cyy};
            const double *blk[9] = {rc, rv, mrc, rh, ru, rh, mrc, rv, rc};
            for (int k = 0; k < 9; k++) std::copy(blk[k], blk[k] + 4, o + 4 * k);
        } else {
            const double sc = d == -2 ? op.bcs : (d == -3 ? op.ocs : op.ncs);
            o[16] = sc * op.cxx;
            o[19] = sc * op.cyy;
            if (d >= 0)
                for (int q = 0; q < 8; q++) std::copy(&fm[32 * (size_t)d + 4 * q], &fm[32 * (size_t)d + 4 * q] + 4,
                                                      o + 4 * (q < 4 ? q : q + 1));
        }
    }
    return STZ_OK;
}

int stz_bd_init_bar(stz_bd *b, int ca) {
    CHECK_BD(b);
    int e;
    if ((e = bd_init_bar_phi(b->c, ca, b->phi, b->stream))) return e;
    if ((e = build_band(b))) return e;
    if ((e = bd_init_bar_fields(b->c, ca, b->st, b->nst, b->cells(b->cc), b->nodes(b->nc), b->stream))) return e;
    // bd_sim::extrapolate (bd_sim.hh:162-165), no boundary conditions
    if ((e = extrapolate(b, true))) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_upload_cells(stz_bd *b, int k, const double *h) {
    CHECK_BD(b);
    if (!h || k < 0 || k >= STZ_BD_PHI) { set_error("bad cell plane %d", k); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(b->cell[b->cc][k], h, (size_t)b->c.m * b->c.n * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_download_cells(stz_bd *b, int k, double *h) {
    CHECK_BD(b);
    if (!h || k < 0 || k > STZ_BD_PHI) { set_error("bad cell plane %d", k); return STZ_ERR_ARG; }
    const double *src = k == STZ_BD_PHI ? b->phi : b->cell[b->cc][k];
    STZ_HIP(hipMemcpyAsync(h, src, (size_t)b->c.m * b->c.n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_upload_nodes(stz_bd *b, int k, const double *h) {
    CHECK_BD(b);
    if (!h || k < 0 || k >= STZ_BD_NNODE) { set_error("bad node plane %d", k); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(b->node[b->nc][k], h, (size_t)b->c.me * b->c.ne * sizeof(double),
                           hipMemcpyHostToDevice, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_download_nodes(stz_bd *b, int k, double *h) {
    CHECK_BD(b);
    if (!h || k < 0 || k >= STZ_BD_NNODE) { set_error("bad node plane %d", k); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, b->node[b->nc][k], (size_t)b->c.me * b->c.ne * sizeof(double),
                           hipMemcpyDeviceToHost, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_download_vel(stz_bd *b, double *h) {
    CHECK_BD(b);
    if (!h) { set_error("null argument"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, b->vel, 2 * (size_t)b->c.me * b->c.ne * sizeof(double), hipMemcpyDeviceToHost,
                           b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_download_src(stz_bd *b, double *h) {
    CHECK_BD(b);
    if (!h) { set_error("null argument"); return STZ_ERR_ARG; }
    STZ_HIP(hipMemcpyAsync(h, b->src, 2 * (size_t)b->c.me * b->c.ne * sizeof(double), hipMemcpyDeviceToHost,
                           b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_download_status(stz_bd *b, int32_t *cs, int32_t *ns) {
    CHECK_BD(b);
    const size_t mn = (size_t)b->c.m * b->c.n, mne = (size_t)b->c.me * b->c.ne;
    std::vector<int8_t> t(std::max(mn, mne));
    if (cs) {
        STZ_HIP(hipMemcpy(t.data(), b->st, mn, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < mn; q++) cs[q] = t[q] & 15;
    }
    if (ns) {
        STZ_HIP(hipMemcpy(t.data(), b->nst, mne, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < mne; q++) ns[q] = t[q];
    }
    return STZ_OK;
}

int stz_bd_set_time(stz_bd *b, double t) {
    CHECK_BD(b);
    b->time = t;
    return STZ_OK;
}

int stz_bd_get_time(stz_bd *b, double *t) {
    CHECK_BD(b);
    if (!t) { set_error("null argument"); return STZ_ERR_ARG; }
    *t = b->time;
    return STZ_OK;
}

int stz_bd_advection_step(stz_bd *b, double dt) {
    CHECK_BD(b);
    if (!b->have_band) { set_error("no level set: call stz_bd_set_level_set first"); return STZ_ERR_STATE; }
    int e = advection(b, dt);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_projection_step(stz_bd *b, double dt) {
    CHECK_BD(b);
    if (!b->have_band) { set_error("no level set: call stz_bd_set_level_set first"); return STZ_ERR_STATE; }
    int e = projection(b, dt);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_post_process(stz_bd *b, int all_fields) {
    CHECK_BD(b);
    if (!b->have_band) { set_error("no level set: call stz_bd_set_level_set first"); return STZ_ERR_STATE; }
    int e = post_process(b, all_fields != 0);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_step_quasistatic(stz_bd *b, double dt, int nsteps) {
    CHECK_BD(b);
    if (!b->have_band) { set_error("no level set: call stz_bd_set_level_set first"); return STZ_ERR_STATE; }
    if (nsteps < 0) { set_error("nsteps < 0"); return STZ_ERR_ARG; }
    int e;
    for (int k = 0; k < nsteps; k++) {
        // bd_sim.cc:240-254
        b->time += dt;
        b->c.x_left -= b->c.u_left * dt;
        b->c.x_right -= b->c.u_right * dt;
        if ((e = advection(b, dt))) return e;
        if ((e = post_process(b, false))) return e;
        if ((e = projection(b, dt))) return e;
        if ((e = post_process(b, true))) return e;
    }
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_last_solve(stz_bd *b, stz_solve_info *info) {
    CHECK_BD(b);
    if (!info) { set_error("null argument"); return STZ_ERR_ARG; }
    *info = b->last;
    return STZ_OK;
}

int stz_bd_set_body_force(stz_bd *b, const double *f) {
    CHECK_BD(b);
    const size_t n2 = 2 * (size_t)b->c.me * b->c.ne;
    if (!f) {
        if (b->force) (void)hipFree(b->force);
        b->force = nullptr;
        return STZ_OK;
    }
    if (!b->force) STZ_HIP(hipMalloc(&b->force, n2 * sizeof(double)));
    STZ_HIP(hipMemcpyAsync(b->force, f, n2 * sizeof(double), hipMemcpyHostToDevice, b->stream));
    STZ_HIP(hipStreamSynchronize(b->stream));
    return STZ_OK;
}

int stz_bd_set_edge_velocity(stz_bd *b, double u_left, double u_right) {
    CHECK_BD(b);
    b->c.u_left = u_left;
    b->c.u_right = u_right;
    return STZ_OK;
}

}  // extern "C"
