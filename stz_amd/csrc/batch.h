// batch.h -- batched launches for an ensemble of independent quasi-static
// simulations.  One launch runs one kernel of the step on every member it
// lists: its workgroups are split among the members by a table that gives
// each workgroup its entry and its index within that entry, and each workgroup
// runs the single-simulation kernel body on its member's arrays, so every
// member computes exactly what its own simulation would.  The per-member
// tables live in device memory and do not change from step to step, which
// lets a whole step be captured in a HIP graph; what does change (time, the
// adaptive V-cycle bookkeeping, which members still cycle) is in EnsState,
// written only by the device.
#pragma once

#include <vector>

#include "device.h"
#include "mg.h"

namespace stz {

// the adaptive V-cycle count of tgmg::solve(type, tgmg_predict&)
// (tgmg.in.cc:266-301), one state machine per member
enum { ENS_FIRST = 0, ENS_LOOP = 1, ENS_EXTRA = 2, ENS_DONE = 3, ENS_CORRECTED = 4, ENS_FAILED = 5 };

struct EnsState {
    double time;       // simulation time (incremented at the start of a step)
    double tol;        // threshold of the mean-square residual
    double denom;      // number of top-level nodes: residual = |r|^2 / denom
    double acc;        // residual at the last test
    int lim, n, k, remaining;  // tgmg_predict limit, cycles counted, loop batch, cycles left
    int phase, active;         // ENS_*; 1 while the member runs the current cycle slot
    int resid;                 // 1 if the slot's cycle is followed by a residual evaluation
    int cycles_run, solves, vcount;
    int status;        // STZ_OK, STZ_ERR_NOT_CONVERGED, STZ_ERR_NAN
};

// predictor settings shared by all members (tgmg_config.hh:26-45)
struct EnsPredict {
    int mult, decay, max_thresh, extra;
    int unit;   // one more cycle per failed test (stz_mg_config.predict_unit)
    int mixed;  // residual after every cycle but the last (mixed) or only before a test (FP64)
};

// entry and workgroup index within it of workgroup w: map holds one
// (entry, local index) pair per workgroup, a single scalar load
__device__ __forceinline__ static int2 batch_slot(const int *map, int w) {
    return reinterpret_cast<const int2 *>(map)[w];
}

// one member's cell-grid kernels (advection, chi diffusion, source, correction)
struct CellEntry {
    int member, gx;     // gx: workgroups along x (64 x 4 threads each)
    SimConst c;
    FieldPtrs f[2];     // plane sets: f[parity] is the current one
    double *vel, *src;
    uint32_t *dec;      // decision record or null
    double xbc;         // wall-row source factor bcs * cxx
};

// one member's launch of the LDS-tiled smoother (mg_tiled.hip)
template <typename T> struct TileEntry {
    int member, gx;
    LevelArgs<T> L;
    T *zout;
    LevelArgs<T> Cz, Cr;
    int zero_z;
    double *corr;  // mixed top level's last launch: vel += result
};

// FP64 residual of a member's top level (refinement or norm)
struct RefineEntry {
    int member, gx;
    LevelArgs<double> L;
    float *b32;         // mixed: FP32 source of the correction equation
    double *partial;    // |r|^2 of each workgroup
};

template <typename T> struct CoarseEntry {
    int member;
    CoarseArgs<T> A;
};

// one member's exact solve of its direct level (mg_direct.hip)
template <typename T> struct DirectEntry {
    int member;
    int M, Nn;
    double *D;     // the level's mode inverses, twiddles and mode scratch (FP64)
    const T *b;
    T *z;
};

// workgroup-partial sums of one member's residual
struct NormEntry {
    int member, np;
    const double *partial;
};

// launchers: E is a device array of ne entries, map one (entry, index)
// pair per workgroup, total workgroups.  With a non-null state array,
// members whose `active` flag is clear return at once (cycle slots; the
// residual launches test `resid`); correct runs only members in ENS_DONE.
int batch_time(EnsState *S, int nmem, double dt, hipStream_t st);
int batch_advect(const CellEntry *E, const int *map, int ne, int total, int parity, double dt, hipStream_t st);
int batch_chi_diffusion(const CellEntry *E, const int *map, int ne, int total, int parity, double dt,
                        hipStream_t st);
int batch_source(const CellEntry *E, const int *map, int ne, int total, int parity, const EnsState *S, double dt,
                 hipStream_t st);
int batch_correct(const CellEntry *E, const int *map, int ne, int total, int parity, const EnsState *S, double dt,
                  hipStream_t st);

// tiled smoother launch of variant (S sweeps, mode, fuse, sparse, 64x32 tiles)
struct TileKey {
    int sweeps, mode, fuse, sparse, big;
    bool operator==(const TileKey &o) const {
        return sweeps == o.sweeps && mode == o.mode && fuse == o.fuse && sparse == o.sparse && big == o.big;
    }
};
template <typename T>
int batch_tile(const TileKey &key, const TileEntry<T> *E, const int *map, int ne, int total, const EnsState *S,
               hipStream_t st);
// one launch of a member's tiled V-cycle: phase 0 pre-smoothing (with the
// fused restriction), 2 post-smoothing; level and launch index within it
template <typename T> struct TileRec {
    int phase, level, index;
    TileKey key;
    LevelArgs<T> L;
    T *zout;
    LevelArgs<T> Cz, Cr;
    int zero_z;
    double *corr;
};
// the launches mg_vcycle_tiled(A, top_zero, st, corr) would issue
// (mg_tiled.hip)
template <typename T>
int mg_vcycle_plan(CoarseArgs<T> A, bool top_zero, std::vector<TileRec<T>> &out, double *corr);

// mixed: FP64 residual -> FP32 source of the correction equation + norm partials
int batch_refine(const RefineEntry *E, const int *map, int ne, int total, int sparse, const EnsState *S,
                 hipStream_t st);
template <typename T> int batch_coarse(const CoarseEntry<T> *E, int ne, size_t lds, const EnsState *S, hipStream_t st);
// direct-level solves: a mode launch (direct_mode_workgroups(M, Nn)
// workgroups per entry) and an inverse launch (direct_inverse_workgroups)
template <typename T>
int batch_direct(const DirectEntry<T> *E, const int *mode_map, int mode_total, const int *inv_map, int inv_total,
                 int ne, size_t lds, const EnsState *S, hipStream_t st);
int direct_mode_workgroups(int M, int Nn);
int direct_inverse_workgroups(int M, int Nn);
template <typename T> int batch_set_lds_limit(size_t bytes);
// fp64: |r|^2 partials of the top level, blocks as mg_resid_norm
int batch_resid_norm(const RefineEntry *E, const int *map, int ne, int total, const EnsState *S, hipStream_t st);

// the per-member V-cycle count (mg.hip): mode 0 starts a solve, mode 1
// follows a cycle slot, mode 2 marks corrected members
int batch_control(EnsState *S, const NormEntry *N, int nmem, EnsPredict P, int mode, hipStream_t st);

}  // namespace stz
