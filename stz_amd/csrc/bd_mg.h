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
// stencil Sc (type -1); the others keep their own 40 doubles in S (type 0).
struct BdLevel {
    int m = 0, n = 0;
    double *z = nullptr, *b = nullptr, *r = nullptr;
    double *S = nullptr;   // 40 per node, read only where type >= 0
    int *type = nullptr;
    double *Sc = nullptr;  // 40
};

enum { kR = 0, kRhat, kP, kV, kS, kT, kPhat, kShat, kTmp, kBdVecs };
constexpr int kMaxPartial = 4096;

struct BdSolver {
    int top_m = 0, top_n = 0, nl = 0;
    int nu1 = 1, nu2 = 1, bottom_half = 10;
    BdTop top{};
    BdLevel lv[kMaxLevels];
    double *vec[kBdVecs] = {};
    double *partial = nullptr, *sc = nullptr;
    double *hscal = nullptr;   // coherent pinned

    static int dims_for(int tm, int tn, std::vector<std::pair<int, int>> &out);
    int init(int tm, int tn);
    void free_all();
    int setup(const BdTop &t, hipStream_t s);          // Galerkin hierarchy
    int solve(double *x, const double *b, double acc, int max_iter, stz_solve_info &info, hipStream_t s);
    int l2_error(const double *x, const double *b, double *out, hipStream_t s);
    int vcycle(const double *rhs, double *z, hipStream_t s);

  private:
    int smooth_top(double *z, const double *b, int sweeps, hipStream_t s);
    int smooth_level(int l, int sweeps, hipStream_t s);
    int dot(const double *a, const double *b, int slot, int op, int host_slot, hipStream_t s);
};

}  // namespace stz
