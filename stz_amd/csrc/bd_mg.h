// bd_mg.h -- the free-boundary projection solver (bd_mg.hip).
#pragma once

#include <utility>
#include <vector>

#include "bd.h"
#include "mg.h"

namespace stz {

// top operator: row types and the constants of qs_multi_bd::init
struct BdTop {
    int m, n;              // nodes (m+1 and n+1 of the cell grid)
    const int *rowt;       // -1 interior, -2 box edge, -3 outside, >= 0 free-row slot
    const double *fmats;   // 32 doubles per free row: dl dc dr cl cr ul uc ur
    double Kmx, Kmy, mx, my, qK, cxx, cyy, bcs, ocs, ncs;
};

// Coarse level.  Rows whose fine rows were all interior rows share one
// stencil (type -1, Sc[0..39]); rows whose fine rows were all outside rows
// and whose product equals the shared outside stencil bit for bit share
// that one (type -2, Sc[40..79]); the others keep their own 40 entries in S
// (type 0).  The Galerkin hierarchy is built in FP64; the mixed form keeps
// an FP32 copy of every level (stencils rounded once per solve) for its
// preconditioner.
template <typename T> struct BdLevelT {
    int m = 0, n = 0;
    T *z = nullptr, *b = nullptr, *r = nullptr;
    T *S = nullptr;        // 40 per node, read only where type >= 0
    int *type = nullptr;   // shared by the FP64 and FP32 copies
    T *Sc = nullptr;       // 80: the shared interior and outside stencils
};
constexpr int kSharedStencils = 80;
using BdLevel = BdLevelT<double>;

enum { kR = 0, kRhat, kP, kV, kS, kT, kPhat, kShat, kTmp, kPre, kBdVecs };
constexpr int kMaxPartial = 4096;    // workgroups of the grid-stride product kernels
constexpr int kPartialHalf = 32768;  // partial sums per product (the second at this offset)

// The operator application fused into the top level's last post-smoothing
// launch: y = A z on every node in FP64, with the sums of (u, y) and, when
// yy, of (y, y) per tile into partial[tile] and partial[kPartialHalf + tile]
struct ApplyArgs {
    double *y;
    const double *u;
    double *partial;
    int yy;
};

struct BdSolver {
    int top_m = 0, top_n = 0, nl = 0;
    int nu1 = 1, nu2 = 1, bottom_half = 10;
    BdTop top{};
    BdLevel lv[kMaxLevels];
    double *vec[kBdVecs] = {};
    double *partial = nullptr, *sc = nullptr;
    double *hscal = nullptr;   // coherent pinned
    // mixed form: the V-cycle (the preconditioner) in FP32 on FP32 copies of
    // the levels; BiCGStab, its residuals and the convergence test stay FP64
    bool mixed = false;
    BdLevelT<float> lv32[kMaxLevels];
    float *top32[4] = {};      // top level: preconditioned p^, s^, the residual, the pre-smoothed iterate

    static int dims_for(int tm, int tn, std::vector<std::pair<int, int>> &out);
    int init(int tm, int tn);
    void free_all();
    int set_mixed(bool on);
    int setup(const BdTop &t, hipStream_t s);          // Galerkin hierarchy
    int solve(double *x, const double *b, double acc, int max_iter, stz_solve_info &info, hipStream_t s);
    int l2_error(const double *x, const double *b, double *out, hipStream_t s);
    int vcycle(const double *rhs, double *z, hipStream_t s);

  private:
    int alloc_mixed();
    void free_mixed();
    // with ap, the last top-level launch also applies the operator where it
    // can (*tiles: the partial sums it wrote, 0 where it did not)
    template <typename T, typename TB>
    int vcycle_t(const TB *rhs, T *z, BdLevelT<T> *L, T *r0, T *zpre, hipStream_t s, const ApplyArgs *ap = nullptr,
                 int *tiles = nullptr);
    template <typename T, typename TB>
    int smooth_top(T *zout, const TB *b, int sweeps, int mode, const T *zin, const T *zc, int cm, int cn,
                   hipStream_t s);
    template <typename T>
    int smooth_level(const BdLevelT<T> &L, bool bottom, int sweeps, int mode, const T *zin, T *zout, const T *zc,
                     int cm, int cn, hipStream_t s);
    int dot(const double *a, const double *b, int slot, int op, int host_slot, hipStream_t s);
};

}  // namespace stz
