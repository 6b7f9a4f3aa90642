// sim_internal.h -- the simulation object behind the C ABI's stz_sim handle
// (sim.hip), shared with the ensemble driver (ensemble.hip).
#pragma once

#include <vector>

#include "device.h"
#include "mg.h"

namespace stz {

// ---------------------------------------------------------------------------
// host-side operator hierarchy
// ---------------------------------------------------------------------------
struct HostLevel {
    int m, n;
    int nb = 0;              // rows at each wall whose stencil is not the interior one
    bool xinv;               // stencil independent of i
    std::vector<double> S;   // 40 per row (xinv) or per node
    const double *at(int i, int j) const { return &S[(size_t)(xinv ? j : (i + m * j)) * 40]; }
};

// ---------------------------------------------------------------------------
// device hierarchy for one precision
// ---------------------------------------------------------------------------
template <typename T> struct DevMG {
    int L = 0;
    CoarseArgs<T> args{};
    T *S[kMaxLevels] = {};
    T *zb[kMaxLevels] = {};  // z and b of levels (own allocations)
    T *pack = nullptr;       // constant part of the workgroup solver's LDS image
    size_t pack_elems = 0;
    size_t lds = 0;
    double *G = nullptr;     // the direct level's mode inverses (mg_direct.hip, FP64)
    size_t G_elems = 0;

    void free_all() {
        for (int l = 0; l < kMaxLevels; l++) {
            if (S[l]) (void)hipFree(S[l]);
            if (zb[l]) (void)hipFree(zb[l]);
            S[l] = zb[l] = nullptr;
        }
        if (pack) (void)hipFree(pack);
        pack = nullptr;
        pack_elems = 0;
        if (G) (void)hipFree(G);
        G = nullptr;
        G_elems = 0;
    }
};

}  // namespace stz

// ---------------------------------------------------------------------------
// the simulation object
// ---------------------------------------------------------------------------
struct stz_sim {
    stz_sim_params p;
    stz::SimConst c;
    int precision;
    int device;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double time = 0;
    // fields
    double *planes = nullptr;
    stz::FieldPtrs cur{}, alt{};
    double *cu = nullptr;              // c_field::cu: diagnostic output plane
    double *un = nullptr, *vn = nullptr;  // new u, v of the direct step
    double *vel = nullptr, *src = nullptr;
    uint32_t *dec = nullptr;
    bool rec_dec = false;
    uint32_t *rep = nullptr;  // decision record being replayed (rep_steps x m*n)
    int rep_steps = 0, rep_next = 0;
    // multigrid
    std::vector<stz::HostLevel> hl;
    double viscdt_prev = -1;
    stz::DevMG<double> mg64;
    stz::DevMG<float> mg32;
    double *partial = nullptr, *dnorm = nullptr, *hnorm = nullptr;
    hipEvent_t norm_ev = nullptr;  // marks the end of a convergence test's norm
    int nblocks_norm = 0;
    double acc_base = 0;
    stz_mg_config cfg{};
    int lim = 0, solves = 0, vcount = 0;
    stz_solve_info last{};

    stz::LevelArgs<double> top64() const {
        stz::LevelArgs<double> L = mg64.args.lv[0];
        L.m = c.m;
        L.n = c.n + 1;
        L.z = vel;
        L.b = src;
        L.S = mg64.S[0];
        L.rs = 1;
        L.cs = 0;
        L.xodd = c.m & 1;
        return L;
    }
};

namespace stz {
// the constants of a simulation and, with acc_base, the base of its solve
// threshold (sim.hip)
void fill_sim_const(const stz_sim_params *p, SimConst &c, double *acc_base);
// one row stencil of the top operator for viscosity/dt = viscdt (sim.hip)
void top_row_stencil(const SimConst &c, double viscdt, bool wall, double *s40);
}  // namespace stz

// (re)build the operators for viscosity/dt = viscdt (qs_multi::init,
// qs_multi.cc:31-54): host Galerkin hierarchy, uploads; no-op when unchanged
int mg_build(stz_sim *s, double viscdt);
