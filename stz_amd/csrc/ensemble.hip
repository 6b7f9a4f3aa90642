// ensemble.hip -- the C ABI of stz_ensemble (include/stz_amd.h): many
// independent quasi-static simulations stepped by batched launches.
//
// One step of the ensemble is the single simulation's step
// (stz_sim_step_quasistatic, sim.hip) with every launch widened to all
// members:
//   head   time += dt, advection (+ chi diffusion), source, first residual
//          evaluation (mixed: FP64 residual -> FP32 source; FP64 form: norm),
//          control: start each member's tgmg_predict count
//   slot   one V-cycle of every member whose count asks for one (the others'
//          workgroups return at once), its residual, control: count and test
//   tail   stress correction of the members whose solve converged, control:
//          mark them corrected
// The host launches head, the number of slots the members' predictors
// announce, and tail, then synchronises once and reads the members' states;
// a member that needs more cycles than planned gets further slots and
// another tail.  Launch arguments never change between steps (the state they
// depend on is in EnsState on the device), so a step is captured once per
// step parity and slot count as one HIP graph and replayed; further slots
// and tails replay graphs of their own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "batch.h"
#include "sim_internal.h"

using namespace stz;

namespace {

// a device array built from a host vector, freed by the owner
struct DevBuf {
    void *p = nullptr;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
};

template <typename T> struct DevTable {
    T *E = nullptr;
    int *map = nullptr;  // (entry, workgroup within it) per workgroup
    int ne = 0, total = 0;
};

template <typename T> struct TileLaunch {
    TileKey key;
    DevTable<TileEntry<T>> t;
};

template <typename T> struct VCyclePlan {
    std::vector<TileLaunch<T>> down, up;
    CoarseEntry<T> *coarse = nullptr;  // members solving their coarse levels in one workgroup
    int ncoarse = 0;
    size_t lds = 0;
    DevTable<DirectEntry<T>> direct;   // members whose V-cycle stops at a direct level
    int *dinv_map = nullptr;           // and the workgroups of their inverse launch
    int dinv_total = 0;
    size_t dlds = 0;
};

}  // namespace

struct stz_ensemble {
    int device = 0, precision = 0;
    std::vector<stz_sim *> mem;
    hipStream_t stream = nullptr;
    bool use_graphs = true;
    // tables (valid for dt_tab and the members' plane-set parity below)
    bool built = false;
    double dt_tab = 0;
    stz_mg_config cfg{};
    std::vector<double *> cur0;   // members' current p plane when the tables were built
    std::vector<SimConst> c0;     // and their constants and decision-record flags
    std::vector<int> rec0;
    int parity = 0;               // steps since then, mod 2
    std::vector<DevBuf> bufs;
    DevTable<CellEntry> cells, srcs, chis;
    DevTable<RefineEntry> refs;
    NormEntry *norms = nullptr;
    int top_sparse = 0;
    VCyclePlan<double> plan64;
    VCyclePlan<float> plan32;
    EnsPredict pred{};
    // state: device array and a pinned host copy
    EnsState *dst = nullptr, *hst = nullptr;
    // graphs
    hipGraphExec_t g_tail[2] = {nullptr, nullptr}, g_slot = nullptr;
    std::map<std::pair<int, int>, hipGraphExec_t> g_step;  // whole step by (parity, planned slots)
    int64_t stats[4] = {0, 0, 0, 0};
    int tuned[2] = {0, 0};        // launch positions that kept 32x16 / 64x34 tiles

    int nmem() const { return (int)mem.size(); }
    void free_graphs() {
        for (int p = 0; p < 2; p++) {
            if (g_tail[p]) (void)hipGraphExecDestroy(g_tail[p]);
            g_tail[p] = nullptr;
        }
        if (g_slot) (void)hipGraphExecDestroy(g_slot);
        g_slot = nullptr;
        for (auto &g : g_step) (void)hipGraphExecDestroy(g.second);
        g_step.clear();
    }
    void free_tables() {
        free_graphs();
        for (auto &b : bufs) b.release();
        bufs.clear();
        plan64 = VCyclePlan<double>();
        plan32 = VCyclePlan<float>();
        tuned[0] = tuned[1] = 0;
        built = false;
    }
};

namespace {

template <typename T> int upload_vec(stz_ensemble *e, const std::vector<T> &h, T **out) {
    *out = nullptr;
    if (h.empty()) return STZ_OK;
    DevBuf b;
    STZ_HIP(hipMalloc(&b.p, h.size() * sizeof(T)));
    e->bufs.push_back(b);
    STZ_HIP(hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = reinterpret_cast<T *>(b.p);
    return STZ_OK;
}

// workgroup -> (entry, index within the entry) for entries of wgs[k] workgroups
int make_map(stz_ensemble *e, const std::vector<int> &wgs, int **map, int *total) {
    std::vector<int> m;
    for (size_t k = 0; k < wgs.size(); k++)
        for (int w = 0; w < wgs[k]; w++) {
            m.push_back((int)k);
            m.push_back(w);
        }
    *total = (int)(m.size() / 2);
    return upload_vec(e, m, map);
}

// entries with their workgroup counts -> device table
template <typename T>
int make_table(stz_ensemble *e, const std::vector<T> &E, const std::vector<int> &wgs, DevTable<T> &t) {
    int rc;
    if ((rc = upload_vec(e, E, &t.E))) return rc;
    if ((rc = make_map(e, wgs, &t.map, &t.total))) return rc;
    t.ne = (int)E.size();
    return STZ_OK;
}

// workgroups of the tiled smoother on a level (launch_tile's grid)
template <typename T> void tile_grid(const LevelArgs<T> &L, bool big, int &gx, int &gy) {
    gx = big ? L.m / 64 : L.m / 32;
    gy = big ? (L.n + 33) / 34 : (L.n + 15) / 16;
}

// average time of one pass over `ls` (ungated launches, after one untimed
// pass).  Between steps the buffers a tiled launch writes (a level's z
// buffers, the coarser level's source) are rewritten by the next cycle
// before they are read, except vel (the FP64 top level's z, or the target
// of the mixed correction): the caller keeps that aside.
template <typename T> int time_launches(stz_ensemble *e, const std::vector<TileLaunch<T>> &ls, double *us) {
    hipEvent_t a, b;
    STZ_HIP(hipEventCreate(&a));
    STZ_HIP(hipEventCreate(&b));
    int rc = STZ_OK;
    for (int rep = 0; rep < 4 && !rc; rep++) {
        if (rep == 1) (void)hipEventRecord(a, e->stream);
        for (const TileLaunch<T> &l : ls)
            if ((rc = batch_tile<T>(l.key, l.t.E, l.t.map, l.t.ne, l.t.total, nullptr, e->stream))) break;
    }
    float ms = 0;
    if (!rc) {
        (void)hipEventRecord(b, e->stream);
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
            set_error("timing the tile launches failed");
            rc = STZ_ERR_HIP;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *us = 1e3 * ms / 3;
    return rc;
}

// every member's V-cycle plan, launches grouped by position: pre-smoothing
// by level downwards, the workgroup solves, post-smoothing upwards
template <typename T>
int build_plan(stz_ensemble *e, VCyclePlan<T> &P, DevMG<T> stz_sim::*mgp, bool top_zero) {
    using Pos = std::tuple<int, int, int>;  // (phase, level key, index)
    std::map<Pos, std::vector<std::pair<TileRec<T>, int>>> groups;
    std::vector<CoarseEntry<T>> coarse;
    std::vector<DirectEntry<T>> direct;
    std::vector<int> dwgs, iwgs;
    size_t lds = 0, dlds = 0;
    int rc;
    for (int k = 0; k < e->nmem(); k++) {
        stz_sim *s = e->mem[k];
        DevMG<T> &mg = s->*mgp;
        CoarseArgs<T> A = mg.args;
        A.nu1 = e->cfg.nu1;
        A.nu2 = e->cfg.nu2;
        std::vector<TileRec<T>> recs;
        // mixed: the top level's last launch applies vel += e (one_cycle)
        if ((rc = mg_vcycle_plan(A, top_zero, recs, top_zero ? s->vel : nullptr))) {
            char msg[400];
            std::snprintf(msg, sizeof msg, "ensemble member %d (%d x %d): %s", k, s->c.m, s->c.n, stz_last_error());
            set_error("%s", msg);
            return rc;
        }
        for (const TileRec<T> &r : recs) {
            const int lk = r.phase == 0 ? r.level : -r.level;  // up: deeper levels first
            groups[Pos(r.phase, lk, r.index)].push_back({r, k});
        }
        if (A.ld >= 0) {
            DirectEntry<T> de;
            de.member = k;
            de.M = A.lv[A.ld].m;
            de.Nn = A.lv[A.ld].n;
            de.D = A.Gd;
            de.b = A.lv[A.ld].b;
            de.z = A.lv[A.ld].z;
            direct.push_back(de);
            dwgs.push_back(direct_mode_workgroups(de.M, de.Nn));
            iwgs.push_back(direct_inverse_workgroups(de.M, de.Nn));
            dlds = std::max(dlds, direct_lds_bytes(de.M, de.Nn));
            continue;
        }
        CoarseEntry<T> ce;
        ce.member = k;
        ce.A = A;
        coarse.push_back(ce);
        lds = std::max(lds, coarse_lds_bytes(A));
    }
    for (auto &g : groups) {
        // one launch per variant present at this position
        std::vector<TileKey> keys;
        for (auto &rk : g.second)
            if (std::find(keys.begin(), keys.end(), rk.first.key) == keys.end()) keys.push_back(rk.first.key);
        for (const TileKey &key : keys) {
            std::vector<const TileRec<T> *> R;
            std::vector<int> who;
            for (auto &rk : g.second)
                if (rk.first.key == key) {
                    R.push_back(&rk.first);
                    who.push_back(rk.second);
                }
            // candidates: 32x16 tiles for all; 64x34 tiles where the level
            // width allows them (the rest 32x16).  Tile shape does not change
            // what a launch computes, only how much of the halo is recomputed
            // and how many workgroups share the chip, so the faster is kept.
            std::vector<std::vector<TileLaunch<T>>> cands;
            for (int big = 0; big < 2; big++) {
                std::vector<TileLaunch<T>> cand;
                for (int part = 0; part < 2; part++) {
                    const bool want_big = big && part == 0;
                    std::vector<TileEntry<T>> E;
                    std::vector<int> wgs;
                    for (size_t q = 0; q < R.size(); q++) {
                        const TileRec<T> &r = *R[q];
                        const bool can_big = r.L.m % 64 == 0;
                        if ((big ? (can_big ? 0 : 1) : 0) != part) continue;
                        TileEntry<T> te;
                        te.member = who[q];
                        int gx, gy;
                        tile_grid(r.L, want_big, gx, gy);
                        te.gx = gx;
                        te.L = r.L;
                        te.zout = r.zout;
                        te.Cz = r.Cz;
                        te.Cr = r.Cr;
                        te.zero_z = r.zero_z;
                        te.corr = r.corr;
                        E.push_back(te);
                        wgs.push_back(gx * gy);
                    }
                    if (E.empty()) continue;
                    TileLaunch<T> tl;
                    tl.key = key;
                    tl.key.big = want_big ? 1 : 0;
                    if ((rc = make_table(e, E, wgs, tl.t))) return rc;
                    cand.push_back(tl);
                }
                bool any_big = false;
                for (auto &tl : cand) any_big |= tl.key.big != 0;
                if (big == 0 || any_big) cands.push_back(cand);
            }
            int best = 0;
            if (cands.size() > 1) {
                double t[2];
                for (int c = 0; c < 2; c++)
                    if ((rc = time_launches<T>(e, cands[c], &t[c]))) return rc;
                best = t[1] < t[0] ? 1 : 0;
            }
            e->tuned[best]++;
            auto &dst = std::get<0>(g.first) == 0 ? P.down : P.up;
            dst.insert(dst.end(), cands[best].begin(), cands[best].end());
        }
    }
    if ((rc = upload_vec(e, coarse, &P.coarse))) return rc;
    P.ncoarse = (int)coarse.size();
    P.lds = lds;
    if (lds > 160 * 1024) {
        set_error("workgroup solver needs %zu bytes of LDS", lds);
        return STZ_ERR_STATE;
    }
    if (!direct.empty()) {
        if ((rc = make_table(e, direct, dwgs, P.direct))) return rc;
        if ((rc = make_map(e, iwgs, &P.dinv_map, &P.dinv_total))) return rc;
        P.dlds = dlds;
        if ((rc = mg_direct_set_lds_limit<T>(dlds))) return rc;
    }
    return batch_set_lds_limit<T>(lds);
}

int build_tables(stz_ensemble *e, double dt) {
    e->free_tables();
    int rc;
    const int N = e->nmem();
    // settings: one for all members
    e->cfg = e->mem[0]->cfg;
    for (int k = 1; k < N; k++)
        if (std::memcmp(&e->mem[k]->cfg, &e->cfg, sizeof e->cfg) != 0) {
            set_error("ensemble member %d has multigrid settings different from member 0", k);
            return STZ_ERR_ARG;
        }
    if (e->cfg.kernels != 0) {
        set_error("ensembles run the tiled kernels only (kernels = 0)");
        return STZ_ERR_ARG;
    }
    e->pred.mult = e->cfg.predict_mult;
    e->pred.decay = e->cfg.predict_decay;
    e->pred.max_thresh = e->cfg.predict_mult * e->cfg.predict_max_iters;
    e->pred.unit = e->cfg.predict_unit;
    e->pred.extra = e->cfg.extra_cycles;
    e->pred.mixed = e->precision == STZ_PREC_MIXED ? 1 : 0;
    // operators of every member for this dt (projection_impl)
    for (int k = 0; k < N; k++)
        if ((rc = mg_build(e->mem[k], e->mem[k]->p.viscosity / dt))) return rc;

    std::vector<CellEntry> cells, chis;
    std::vector<int> wcell, wsrc, wchi, wref;
    std::vector<RefineEntry> refs;
    std::vector<NormEntry> norms;
    e->cur0.assign(N, nullptr);
    e->c0.resize(N);
    e->rec0.assign(N, 0);
    int sparse = -1;
    for (int k = 0; k < N; k++) {
        stz_sim *s = e->mem[k];
        const SimConst &c = s->c;
        CellEntry ce;
        std::memset(&ce, 0, sizeof ce);
        ce.member = k;
        ce.gx = (c.m + 63) / 64;
        ce.c = c;
        ce.f[0] = s->cur;
        ce.f[1] = s->alt;
        ce.vel = s->vel;
        ce.src = s->src;
        ce.dec = s->rec_dec ? s->dec : nullptr;
        ce.xbc = c.bcs * s->hl[0].S[40 + 16];  // projection_impl: bcs * cxx
        cells.push_back(ce);
        wcell.push_back(ce.gx * ((c.n + 3) / 4));
        wsrc.push_back(ce.gx * ((c.n + 1 + 3) / 4));
        if (c.chi_diff != 0 || c.chi_len != 0) {
            chis.push_back(ce);
            wchi.push_back(wcell.back());
        }
        e->cur0[k] = s->cur.p;
        e->c0[k] = s->c;
        e->rec0[k] = s->rec_dec ? 1 : 0;
        const LevelArgs<double> t = s->top64();
        if (!wall_ok(t)) {
            set_error("ensemble member %d: top level not covered by the tiled kernels", k);
            return STZ_ERR_ARG;
        }
        RefineEntry re;
        std::memset(&re, 0, sizeof re);
        re.member = k;
        re.L = t;
        re.partial = s->partial;
        NormEntry ne;
        ne.member = k;
        ne.partial = s->partial;
        if (e->precision == STZ_PREC_MIXED) {
            // mg_refine's grid: 64 x 16 tiles
            re.gx = (t.m + 63) / 64;
            re.b32 = s->mg32.args.lv[0].b;
            wref.push_back(re.gx * ((t.n + 15) / 16));
            if (sparse < 0) sparse = t.sparse;
            if (sparse != t.sparse) {
                set_error("ensemble members differ in the top-level stencil pattern");
                return STZ_ERR_ARG;
            }
        } else {
            // mg_resid_norm's grid: 256 nodes per block
            re.gx = (t.m * t.n + 255) / 256;
            wref.push_back(re.gx);
        }
        ne.np = wref.back();
        refs.push_back(re);
        norms.push_back(ne);
    }
    e->top_sparse = sparse > 0 ? 1 : 0;
    if ((rc = make_table(e, cells, wcell, e->cells))) return rc;
    e->srcs = e->cells;
    {
        if ((rc = make_map(e, wsrc, &e->srcs.map, &e->srcs.total))) return rc;
    }
    if ((rc = make_table(e, chis, wchi, e->chis))) return rc;
    if ((rc = make_table(e, refs, wref, e->refs))) return rc;
    if ((rc = upload_vec(e, norms, &e->norms))) return rc;
    {
        // the tile-shape timing runs the top level's last smoothing launch,
        // which writes vel (FP64: it is the level's z; mixed: vel += e):
        // keep the velocities aside meanwhile
        std::vector<size_t> off(N + 1, 0);
        for (int k = 0; k < N; k++) off[k + 1] = off[k] + 2 * (size_t)e->mem[k]->c.m * (e->mem[k]->c.n + 1);
        double *keep = nullptr;
        STZ_HIP(hipMalloc(&keep, off[N] * sizeof(double)));
        for (int k = 0; k < N && !rc; k++)
            if (hipMemcpyAsync(keep + off[k], e->mem[k]->vel, (off[k + 1] - off[k]) * sizeof(double),
                               hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
                rc = STZ_ERR_HIP;
        if (!rc) {
            if (e->precision == STZ_PREC_MIXED) rc = build_plan<float>(e, e->plan32, &stz_sim::mg32, true);
            else rc = build_plan<double>(e, e->plan64, &stz_sim::mg64, false);
        }
        for (int k = 0; k < N; k++)
            if (hipMemcpyAsync(e->mem[k]->vel, keep + off[k], (off[k + 1] - off[k]) * sizeof(double),
                               hipMemcpyDeviceToDevice, e->stream) != hipSuccess && !rc)
                rc = STZ_ERR_HIP;
        if (hipStreamSynchronize(e->stream) != hipSuccess && !rc) rc = STZ_ERR_HIP;
        (void)hipFree(keep);
        if (rc == STZ_ERR_HIP) set_error("saving the velocities around the tile timing failed");
    }
    if (rc) return rc;
    e->dt_tab = dt;
    e->parity = 0;
    e->built = true;
    return STZ_OK;
}

// ---- the three parts of a step --------------------------------------------
int launch_head(stz_ensemble *e, int parity, double dt) {
    int rc;
    hipStream_t st = e->stream;
    const int N = e->nmem();
    if ((rc = batch_time(e->dst, N, dt, st))) return rc;
    const DevTable<CellEntry> &C = e->cells, &X = e->chis, &S = e->srcs;
    if ((rc = batch_advect(C.E, C.map, C.ne, C.total, parity, dt, st))) return rc;
    if ((rc = batch_chi_diffusion(X.E, X.map, X.ne, X.total, parity, dt, st))) return rc;
    // the advected plane set is current from here on
    if ((rc = batch_source(S.E, S.map, S.ne, S.total, parity ^ 1, e->dst, dt, st))) return rc;
    const DevTable<RefineEntry> &R = e->refs;
    if (e->precision == STZ_PREC_MIXED) rc = batch_refine(R.E, R.map, R.ne, R.total, e->top_sparse, nullptr, st);
    else rc = batch_resid_norm(R.E, R.map, R.ne, R.total, nullptr, st);
    if (rc) return rc;
    return batch_control(e->dst, e->norms, N, e->pred, 0, st);
}

template <typename T> int launch_vcycle(stz_ensemble *e, const VCyclePlan<T> &P) {
    int rc;
    for (const TileLaunch<T> &l : P.down)
        if ((rc = batch_tile<T>(l.key, l.t.E, l.t.map, l.t.ne, l.t.total, e->dst, e->stream))) return rc;
    if ((rc = batch_coarse<T>(P.coarse, P.ncoarse, P.lds, e->dst, e->stream))) return rc;
    const DevTable<DirectEntry<T>> &D = P.direct;
    if ((rc = batch_direct<T>(D.E, D.map, D.total, P.dinv_map, P.dinv_total, D.ne, P.dlds, e->dst, e->stream)))
        return rc;
    for (const TileLaunch<T> &l : P.up)
        if ((rc = batch_tile<T>(l.key, l.t.E, l.t.map, l.t.ne, l.t.total, e->dst, e->stream))) return rc;
    return STZ_OK;
}

int launch_slot(stz_ensemble *e) {
    int rc;
    hipStream_t st = e->stream;
    const DevTable<RefineEntry> &R = e->refs;
    if (e->precision == STZ_PREC_MIXED) {
        // vel += e is applied by the top level's last smoothing launch
        if ((rc = launch_vcycle<float>(e, e->plan32))) return rc;
        if ((rc = batch_refine(R.E, R.map, R.ne, R.total, e->top_sparse, e->dst, st))) return rc;
    } else {
        if ((rc = launch_vcycle<double>(e, e->plan64))) return rc;
        if ((rc = batch_resid_norm(R.E, R.map, R.ne, R.total, e->dst, st))) return rc;
    }
    return batch_control(e->dst, e->norms, e->nmem(), e->pred, 1, st);
}

int launch_tail(stz_ensemble *e, int parity, double dt) {
    int rc;
    const DevTable<CellEntry> &C = e->cells;
    if ((rc = batch_correct(C.E, C.map, C.ne, C.total, parity ^ 1, e->dst, dt, e->stream))) return rc;
    return batch_control(e->dst, e->norms, e->nmem(), e->pred, 2, e->stream);
}

// capture fn() on the ensemble's stream into an executable graph
template <typename F> int capture(stz_ensemble *e, hipGraphExec_t *out, F fn) {
    hipGraph_t g = nullptr;
    STZ_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = fn();
    const hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(ec));
        return STZ_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
        return STZ_ERR_HIP;
    }
    e->stats[3]++;
    return STZ_OK;
}

int run_slot(stz_ensemble *e) {
    e->stats[1]++;
    if (!e->use_graphs) return launch_slot(e);
    int rc;
    if (!e->g_slot && (rc = capture(e, &e->g_slot, [&] { return launch_slot(e); }))) return rc;
    STZ_HIP(hipGraphLaunch(e->g_slot, e->stream));
    return STZ_OK;
}

// head, `slots` cycle slots and tail as one graph
int run_step(stz_ensemble *e, int parity, int slots, double dt) {
    int rc;
    e->stats[1] += slots;
    if (!e->use_graphs) {
        if ((rc = launch_head(e, parity, dt))) return rc;
        for (int q = 0; q < slots; q++)
            if ((rc = launch_slot(e))) return rc;
        return launch_tail(e, parity, dt);
    }
    const std::pair<int, int> key(parity, slots);
    auto it = e->g_step.find(key);
    if (it == e->g_step.end()) {
        hipGraphExec_t g = nullptr;
        rc = capture(e, &g, [&] {
            int r;
            if ((r = launch_head(e, parity, dt))) return r;
            for (int q = 0; q < slots; q++)
                if ((r = launch_slot(e))) return r;
            return launch_tail(e, parity, dt);
        });
        if (rc) return rc;
        it = e->g_step.emplace(key, g).first;
    }
    STZ_HIP(hipGraphLaunch(it->second, e->stream));
    return STZ_OK;
}

int run_tail(stz_ensemble *e, int parity, double dt) {
    if (!e->use_graphs) return launch_tail(e, parity, dt);
    int rc;
    if (!e->g_tail[parity] && (rc = capture(e, &e->g_tail[parity], [&] { return launch_tail(e, parity, dt); })))
        return rc;
    STZ_HIP(hipGraphLaunch(e->g_tail[parity], e->stream));
    return STZ_OK;
}

int sync_state(stz_ensemble *e) {
    STZ_HIP(hipMemcpyAsync(e->hst, e->dst, e->nmem() * sizeof(EnsState), hipMemcpyDeviceToHost, e->stream));
    STZ_HIP(hipStreamSynchronize(e->stream));
    return STZ_OK;
}

bool in_progress(const EnsState &s) { return s.phase == ENS_FIRST || s.phase == ENS_LOOP || s.phase == ENS_EXTRA; }

// cycles the predictor of a member will run if its first test passes
int planned_slots(const stz_ensemble *e) {
    int most = 0;
    for (int k = 0; k < e->nmem(); k++) most = std::max(most, e->mem[k]->lim / e->pred.mult);
    return most + e->pred.extra;
}

}  // namespace

extern "C" {

int stz_ensemble_create(const stz_sim_params *params, int count, int precision, stz_ensemble **out) {
    if (!params || !out || count < 1) {
        set_error("need parameters for at least one member");
        return STZ_ERR_ARG;
    }
    *out = nullptr;
    stz_ensemble *e = new stz_ensemble();
    e->precision = precision;
    auto fail = [&](int code) {
        stz_ensemble_destroy(e);
        return code;
    };
    for (int k = 0; k < count; k++) {
        stz_sim *s = nullptr;
        const int rc = stz_sim_create(&params[k], precision, &s);
        if (rc) return fail(rc);
        e->mem.push_back(s);
    }
    e->device = e->mem[0]->device;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(STZ_ERR_HIP);
    }
    if (hipMalloc(&e->dst, count * sizeof(EnsState)) != hipSuccess ||
        hipHostMalloc(&e->hst, count * sizeof(EnsState), hipHostMallocDefault) != hipSuccess) {
        set_error("out of memory for the ensemble state");
        return fail(STZ_ERR_HIP);
    }
    *out = e;
    return STZ_OK;
}

int stz_ensemble_destroy(stz_ensemble *e) {
    if (!e) return STZ_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->free_tables();
    for (stz_sim *s : e->mem) stz_sim_destroy(s);
    if (e->dst) (void)hipFree(e->dst);
    if (e->hst) (void)hipHostFree(e->hst);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return STZ_OK;
}

int stz_ensemble_size(stz_ensemble *e, int *count) {
    if (!e || !count) { set_error("null argument"); return STZ_ERR_ARG; }
    *count = e->nmem();
    return STZ_OK;
}

int stz_ensemble_member(stz_ensemble *e, int k, stz_sim **member) {
    if (!e || !member) { set_error("null argument"); return STZ_ERR_ARG; }
    if (k < 0 || k >= e->nmem()) { set_error("member %d out of range [0, %d)", k, e->nmem()); return STZ_ERR_ARG; }
    *member = e->mem[k];
    return STZ_OK;
}

int stz_ensemble_set_mg_config(stz_ensemble *e, const stz_mg_config *cfg) {
    if (!e || !cfg) { set_error("null argument"); return STZ_ERR_ARG; }
    for (stz_sim *s : e->mem) {
        const int rc = stz_sim_set_mg_config(s, cfg);
        if (rc) return rc;
    }
    e->built = false;
    return STZ_OK;
}

int stz_ensemble_use_graphs(stz_ensemble *e, int on) {
    if (!e) { set_error("null argument"); return STZ_ERR_ARG; }
    e->use_graphs = on != 0;
    return STZ_OK;
}

int stz_ensemble_tile_choice(stz_ensemble *e, int32_t *out2) {
    if (!e || !out2) { set_error("null argument"); return STZ_ERR_ARG; }
    out2[0] = e->tuned[0];
    out2[1] = e->tuned[1];
    return STZ_OK;
}

int stz_ensemble_stats(stz_ensemble *e, int64_t *out4) {
    if (!e || !out4) { set_error("null argument"); return STZ_ERR_ARG; }
    for (int k = 0; k < 4; k++) out4[k] = e->stats[k];
    return STZ_OK;
}

int stz_ensemble_step_quasistatic(stz_ensemble *e, double dt, int nsteps) {
    if (!e) { set_error("null ensemble"); return STZ_ERR_ARG; }
    if (nsteps < 0) { set_error("nsteps < 0"); return STZ_ERR_ARG; }
    if (nsteps == 0) return STZ_OK;
    STZ_HIP(hipSetDevice(e->device));
    const int N = e->nmem();
    int rc;
    // the tables hold for this dt, these settings and the members' current
    // plane sets; anything done to a member on its own may have changed them
    bool fresh = e->built && e->dt_tab == dt;
    for (int k = 0; fresh && k < N; k++) {
        stz_sim *s = e->mem[k];
        fresh = std::memcmp(&s->cfg, &e->cfg, sizeof e->cfg) == 0 && s->viscdt_prev == s->p.viscosity / dt &&
                (e->parity ? s->alt.p : s->cur.p) == e->cur0[k] &&
                std::memcmp(&s->c, &e->c0[k], sizeof s->c) == 0 && s->rec_dec == (bool)e->rec0[k];
    }
    if (!fresh) {
        for (stz_sim *s : e->mem) STZ_HIP(hipStreamSynchronize(s->stream));
        if ((rc = build_tables(e, dt))) {
            e->free_tables();
            return rc;
        }
    }
    // member state -> device
    for (int k = 0; k < N; k++) {
        stz_sim *s = e->mem[k];
        EnsState &h = e->hst[k];
        std::memset(&h, 0, sizeof h);
        h.time = s->time;
        h.tol = s->acc_base * e->cfg.tol_mult;
        h.denom = (double)s->c.m * (s->c.n + 1);
        h.lim = s->lim;
        h.solves = s->solves;
        h.vcount = s->vcount;
        h.phase = ENS_CORRECTED;
    }
    STZ_HIP(hipMemcpyAsync(e->dst, e->hst, N * sizeof(EnsState), hipMemcpyHostToDevice, e->stream));
    int status = STZ_OK;
    // a launch or copy failing inside a step leaves the members' device state
    // partly advanced and their host mirrors (time, predictor, plane sets)
    // behind it: the tables are marked stale and the members' state is
    // undefined until the caller sets it again (include/stz_amd.h)
    auto abort_step = [&](int code) {
        e->built = false;
        return code;
    };
    for (int step = 0; step < nsteps && status == STZ_OK; step++) {
        const int p = e->parity;
        const int plan = planned_slots(e);
        if ((rc = run_step(e, p, plan, dt))) return abort_step(rc);
        if ((rc = sync_state(e))) return abort_step(rc);
        // members whose predictor asked for more cycles than planned
        for (;;) {
            bool more = false;
            for (int k = 0; k < N; k++) more |= in_progress(e->hst[k]);
            if (!more) break;
            for (int q = 0; q < 2; q++) {
                e->stats[2]++;
                if ((rc = run_slot(e))) return abort_step(rc);
            }
            if ((rc = run_tail(e, p, dt))) return abort_step(rc);
            if ((rc = sync_state(e))) return abort_step(rc);
        }
        e->stats[0]++;
        // host mirrors of every member, as stz_sim_step_quasistatic leaves them
        for (int k = 0; k < N; k++) {
            stz_sim *s = e->mem[k];
            const EnsState &h = e->hst[k];
            s->time = h.time;
            s->lim = h.lim;
            s->solves = h.solves;
            s->vcount = h.vcount;
            std::swap(s->cur, s->alt);
            std::memset(&s->last, 0, sizeof s->last);
            s->last.status = h.status;
            s->last.cycles = h.n;
            s->last.total_cycles = h.cycles_run;
            s->last.refinements = e->precision == STZ_PREC_MIXED ? h.cycles_run : 0;
            s->last.residual = h.acc;
            s->last.tolerance = h.tol;
            if (h.status != STZ_OK && status == STZ_OK) {
                status = h.status;
                set_error("ensemble member %d: %s after %d V-cycles (residual %g, tolerance %g)", k,
                          h.status == STZ_ERR_NAN ? "NaN residual" : "not converged", h.n, h.acc, h.tol);
            }
        }
        e->parity ^= 1;
    }
    return status;
}

}  // extern "C"
