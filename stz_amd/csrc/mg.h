// mg.h -- device-side multigrid structures and launchers (mg.hip).
#pragma once

#include <vector>

#include "device.h"

namespace stz {

constexpr int kMaxLevels = 16;

// One level of the hierarchy: m*n nodes, periodic in x, bounded in y.  Its
// operator is a 9-point 2x2-block stencil: entry k = 0..8 (dl,dc,dr,cl,cc,cr,
// ul,uc,ur) plus k = 9 the inverse of cc, 4 values (a,b,c,d) each, stored for
// node (i,j) at S[(j*rs + i*cs)*40 + 4k + c].  x-invariant levels use cs = 0,
// rs = 1 (one stencil per row).  z and b are interleaved (x,y) pairs.
template <typename T> struct LevelArgs {
    int m, n;
    T *z;   // current solution
    T *z2;  // second buffer: tiled sweeps read z and write z2, then swap
    T *b;
    const T *S;
    int rs, cs;
    int xodd;  // m odd: the seam column i = m-1 gets a colour of its own
    // x-invariant levels: every row j with nb <= j < n-nb has the stencil Si
    // (the rows near the walls differ), so kernels can keep it in registers
    const T *Si;
    int nb;
    // Si has the sparsity of the top-level operator (offdiag9_top in
    // mg_node.h): the tiled kernels then skip its zero entries
    int sparse;
};

// A whole V-cycle: levels 0..L-1 (0 = top); levels >= lc are run by a single
// workgroup with z and b in LDS, the coarsest exactly through binv.
template <typename T> struct CoarseArgs {
    int lc, L;
    int nu1, nu2;
    LevelArgs<T> lv[kMaxLevels];
    // LDS layout of the workgroup solver, in elements of T: first the part
    // that is the same for every solve (row stencils of levels lc..L-1, the
    // coarsest level's dense inverse, transposed), copied from `pack`; then
    // z and b of every level and a residual scratch
    const T *pack;
    int pack_elems;
    int zoff[kMaxLevels], boff[kMaxLevels], soff[kMaxLevels], roff, binv_off;
    int lds_elems;
    // ld >= 0: level ld (an x-invariant level, lc or finer) is solved
    // exactly through its Fourier modes (FP64 buffer Gd, mg_direct.hip, in
    // both hierarchies) and the levels below it are not visited; -1: the
    // workgroup solver runs levels lc..L-1
    int ld;
    double *Gd;
};

// the level a V-cycle bottoms out at, and its solve
template <typename T> inline int bottom_level(const CoarseArgs<T> &A) { return A.ld >= 0 ? A.ld : A.lc; }
template <typename T> int mg_bottom(const CoarseArgs<T> &A, hipStream_t st);

// mg_direct.hip: exact solve of an x-invariant level through its Fourier
// modes, in FP64 whatever the level's T; D is the level's device buffer of
// direct_elems(M, Nn) doubles (host: direct_inverse fills it), with the LDS
// the mode launch needs
template <typename T> int mg_direct_solve(const LevelArgs<T> &L, double *D, hipStream_t st);
size_t direct_lds_bytes(int M, int Nn);
template <typename T> int mg_direct_set_lds_limit(size_t bytes);
size_t direct_elems(int M, int Nn);
bool direct_inverse(const double *S, int M, int Nn, std::vector<double> &D);
// whether the two-launch solve covers an M x Nn level (M even, 2 Nn and M
// within the launches' fixed register and LDS tiles)
bool direct_fits(int M, int Nn);

// LDS layout for levels lc..L-1; returns the number of elements of T
template <typename T> int coarse_layout(CoarseArgs<T> &A);
// host: the constant part of the layout, to be uploaded as A.pack.  S[l] are
// the stencils of level l as stored on the device, binv the coarsest
// level's dense inverse (row-major)
template <typename T>
void coarse_pack(const CoarseArgs<T> &A, const std::vector<double> *S, const std::vector<double> &binv,
                 std::vector<T> &out);

template <typename T> int mg_smooth(const LevelArgs<T> &L, int sweeps, hipStream_t st);
template <typename T> int mg_restrict(const LevelArgs<T> &F, const LevelArgs<T> &C, hipStream_t st);
template <typename T> int mg_prolong(const LevelArgs<T> &C, const LevelArgs<T> &F, hipStream_t st);
template <typename T> int mg_coarse(const CoarseArgs<T> &A, hipStream_t st);
template <typename T> int mg_vcycle(const CoarseArgs<T> &A, hipStream_t st);
// the V-cycle from level l0 down (level l0 from its current z)
template <typename T> int mg_vcycle_from(const CoarseArgs<T> &A, int l0, hipStream_t st);
// tgmg::fmg_pre: the full-multigrid pass before a solve's V-cycles
template <typename T> int mg_fmg_pre(const CoarseArgs<T> &A, hipStream_t st);
template <typename T> int mg_resid_norm(const LevelArgs<T> &L, double *partial, double *out, hipStream_t st);
template <typename T> size_t coarse_lds_bytes(const CoarseArgs<T> &A);
// allow the workgroup solver this much dynamic LDS
template <typename T> int mg_set_lds_limit(size_t bytes);
int mg_resid64_to32(const LevelArgs<double> &L, float *b32, float *z32, double *partial, double *out,
                    hipStream_t st);
int mg_correct64(double *z, const float *e, int count, hipStream_t st);
int mg_sum_partials(const double *partial, int np, double *out, hipStream_t st);
template <typename T> int mg_fill_zero(T *p, size_t count, hipStream_t st);

// tiled path (mg_tiled.hip)
// row stencils with at most two distinct wall rows at each wall: the LDS
// tiles of the tiled kernels and the refinement kernel apply
template <typename T> bool wall_ok(const LevelArgs<T> &L);
template <typename T> bool tiled_ok(const LevelArgs<T> &L);
// one V-cycle with the tiled kernels where a level allows them; swaps the
// z/z2 pointers of the levels it smooths; top_zero starts the top level
// from z = 0 (correction equation)
// corr (mixed form): the top level's last smoothing launch adds its result
// to this FP64 array instead of storing it (the refinement's vel += e);
// *corrected says whether it did (the top level was tiled)
template <typename T>
int mg_vcycle_tiled(CoarseArgs<T> &A, bool top_zero, hipStream_t st, double *corr = nullptr,
                    int *corrected = nullptr);
// one smoothing launch of a tiled V-cycle on `level` alone: mode 1 the
// pre-smoothing with its restriction, mode 2 the post-smoothing with its
// prolongation (and, with corr on level 0, the mixed form's vel += e)
template <typename T>
int mg_vcycle_launch(CoarseArgs<T> &A, int level, int mode, bool top_zero, double *corr, hipStream_t st);
// `sweeps` smoothing sweeps on one level (tiled where possible; swaps z/z2)
template <typename T> int mg_smooth_level(LevelArgs<T> &L, int sweeps, hipStream_t st);
// residual of the FP64 top level -> FP32 source, |r|^2 partials summed to *out
// (out == nullptr: partials only)
int mg_refine(const LevelArgs<double> &L, float *b32, double *partial, double *out, hipStream_t st);
int mg_refine_blocks(const LevelArgs<double> &L);

}  // namespace stz
