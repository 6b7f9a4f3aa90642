// bd_mg.hip -- gfx950 solver for the free-boundary projection: the
// operator of qs_multi_bd (sb/main/qs_multi_bd.hh:31-57) on the nodes, a
// Galerkin hierarchy of tgmg (tgmg.in.cc:10-30, 667-779; coarse sizes
// (m+2)>>1 on both bounded axes, tgmg.hh:109-110, and two levels fewer than
// tgmg would build, bd_sim.cc:51) rebuilt on the device at every solve, and
// V-cycles of four-colour Gauss-Seidel sweeps.
//
// The reference iterates bare V-cycles to a mean-square residual of 1e-20
// (qs_multi_bd.cc:16, tgmg.in.cc:242-255).  Its free-surface rows are not
// diagonally dominant, and on some geometries (a plate with a hole at 64^2
// or 256^2 cells, for one) those V-cycles diverge although the linear system
// is well posed.  Here one V-cycle is the preconditioner of BiCGStab
// (van der Vorst 1992, right preconditioning), run to the same residual
// threshold: where the reference converges both reach the same solution,
// and where it diverges this one still converges.
//
// Top level: matrix-free from the row types (-1 interior with the constant
// stencil of qs_multi_bd::mul_a, -2 box edge, -3 outside, >= 0 a
// free-surface row whose eight blocks sit in a compact table).  Coarse
// levels: 40 doubles per node, nine 2x2 blocks (dl,dc,dr,cl,cc,cr,ul,uc,ur)
// and the inverse of cc.  Vectors are interleaved (x,y) pairs.
#include "bd_mg.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace stz {

namespace {

constexpr int kThreads = 256;
inline dim3 blocks(long n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

template <typename T> struct V2T {
    T x, y;
};
using V2 = V2T<double>;

// one entry (2x2 block k of row (i,j)) of the top operator, as the
// reference's a_xx accessors give it (qs_multi_bd.hh:31-39)
__device__ __forceinline__ void top_entry(const BdTop &t, int ij, int k, double *o) {
    const int d = t.rowt[ij];
    o[0] = o[1] = o[2] = o[3] = 0;
    if (d == -1) {
        switch (k) {
        case 0: case 8: o[1] = o[2] = t.qK; break;            // rc
        case 2: case 6: o[1] = o[2] = -t.qK; break;           // mrc
        case 1: case 7: o[0] = t.mx; o[3] = t.Kmy; break;     // rv
        case 3: case 5: o[0] = t.Kmx; o[3] = t.my; break;     // rh
        default: o[0] = t.cxx; o[3] = t.cyy;                  // ru
        }
        return;
    }
    if (k == 4) {
        const double s = d == -2 ? t.bcs : (d == -3 ? t.ocs : t.ncs);
        o[0] = s * t.cxx;
        o[3] = s * t.cyy;
        return;
    }
    if (d >= 0) {
        const double *a = t.fmats + 32 * (size_t)d + 4 * (k < 4 ? k : k - 1);
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
        o[3] = a[3];
    }
}

// off-diagonal part of row ij applied to z (qs_multi_bd::mul_a,
// qs_multi_bd.hh:41-56) and the diagonal block; in T, with the operator's
// FP64 constants rounded to T
// zp: the row's own pair of z, in rows of `m` pairs (the vector itself, or
// a tile of it in LDS)
template <typename T, typename TZ = T>
__device__ __forceinline__ V2T<T> top_offdiag_at(const BdTop &t, const TZ *zp, int m, int d) {
    if (d == -1) {
        const TZ *zl = zp - 2, *zr = zp + 2, *zd = zp - 2 * m, *zu = zp + 2 * m;
        const TZ *zdl = zd - 2, *zdr = zd + 2, *zul = zu - 2, *zur = zu + 2;
        const T Kmx = (T)t.Kmx, Kmy = (T)t.Kmy, mx = (T)t.mx, my = (T)t.my, qK = (T)t.qK;
        return V2T<T>{Kmx * ((T)zl[0] + (T)zr[0]) + my * ((T)zd[0] + (T)zu[0]) +
                          qK * ((T)zdl[1] + (T)zur[1] - (T)zdr[1] - (T)zul[1]),
                      mx * ((T)zl[1] + (T)zr[1]) + Kmy * ((T)zd[1] + (T)zu[1]) +
                          qK * ((T)zdl[0] + (T)zur[0] - (T)zdr[0] - (T)zul[0])};
    }
    if (d < 0) return V2T<T>{0, 0};
    const double *a = t.fmats + 32 * (size_t)d;
    const int off[8] = {-m - 1, -m, -m + 1, -1, 1, m - 1, m, m + 1};
    T rx = 0, ry = 0;
    for (int q = 0; q < 8; q++) {
        const TZ *zz = zp + 2 * (ptrdiff_t)off[q];
        rx += (T)a[4 * q] * (T)zz[0] + (T)a[4 * q + 1] * (T)zz[1];
        ry += (T)a[4 * q + 2] * (T)zz[0] + (T)a[4 * q + 3] * (T)zz[1];
    }
    return V2T<T>{rx, ry};
}
template <typename T, typename TZ = T>
__device__ __forceinline__ V2T<T> top_offdiag(const BdTop &t, const TZ *z, int ij, int d) {
    return top_offdiag_at<T, TZ>(t, z + 2 * (size_t)ij, t.m, d);
}
__device__ __forceinline__ double top_diag_scale(const BdTop &t, int d) {
    return d == -1 ? 1.0 : (d == -2 ? t.bcs : (d == -3 ? t.ocs : t.ncs));
}

// the stencil of row ij: shared (type -1 interior, -2 outside) or its own
template <typename T> __device__ __forceinline__ const T *lvl_stencil(const BdLevelT<T> &L, int ij) {
    const int t = L.type[ij];
    return t < 0 ? L.Sc + 40 * (-1 - t) : L.S + 40 * (size_t)ij;
}

// off-diagonal part of a coarse row: stencil s, zp the row's own pair of z
// in rows of `stride` pairs, (i,j) the row in the m x n level
// (the loads are unconditional, from the row itself where a neighbour is
// missing, so that they are all in flight at once)
template <typename T>
__device__ __forceinline__ V2T<T> lvl_offdiag_at(const T *s, const T *zp, int stride, int i, int j, int m, int n) {
    T rx = 0, ry = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if (k == 4) continue;
        const int di = k % 3 - 1, dj = k / 3 - 1;
        const bool in = i + di >= 0 && i + di < m && j + dj >= 0 && j + dj < n;
        const T *zz = zp + 2 * (in ? di + stride * dj : 0);
        const T z0 = zz[0], z1 = zz[1];
        if (!in) continue;
        rx += s[4 * k] * z0 + s[4 * k + 1] * z1;
        ry += s[4 * k + 2] * z0 + s[4 * k + 3] * z1;
    }
    return V2T<T>{rx, ry};
}
// the 40 entries of a coarse row's stencil, loaded together
template <typename T> __device__ __forceinline__ void lvl_row(const BdLevelT<T> &L, int ij, T (&s)[40]) {
    const T *p = lvl_stencil(L, ij);
#pragma unroll
    for (int q = 0; q < 40; q++) s[q] = p[q];
}

// ---------------------------------------------------------------------------
// Gauss-Seidel sweeps, one colour (i&1, j&1) per launch
// ---------------------------------------------------------------------------
template <typename T, typename TB>
__global__ void __launch_bounds__(kThreads) gs_top_kernel(BdTop t, T *__restrict__ z, const TB *__restrict__ b,
                                                          int colour) {
    const int hm = (t.m + 1 - (colour & 1)) >> 1, hn = (t.n + 1 - (colour >> 1)) >> 1;
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= hm * hn) return;
    const int i = 2 * (q % hm) + (colour & 1), j = 2 * (q / hm) + (colour >> 1), ij = i + t.m * j;
    const int d = t.rowt[ij];
    const V2T<T> o = top_offdiag<T>(t, z, ij, d);
    const double sc = top_diag_scale(t, d);
    z[2 * (size_t)ij] = ((T)b[2 * (size_t)ij] - o.x) / (T)(sc * t.cxx);
    z[2 * (size_t)ij + 1] = ((T)b[2 * (size_t)ij + 1] - o.y) / (T)(sc * t.cyy);
}

template <typename T> __global__ void __launch_bounds__(kThreads) gs_level_kernel(BdLevelT<T> L, int colour) {
    const int hm = (L.m + 1 - (colour & 1)) >> 1, hn = (L.n + 1 - (colour >> 1)) >> 1;
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= hm * hn) return;
    const int i = 2 * (q % hm) + (colour & 1), j = 2 * (q / hm) + (colour >> 1), ij = i + L.m * j;
    T s[40];
    lvl_row(L, ij, s);
    const V2T<T> o = lvl_offdiag_at(s, L.z + 2 * (size_t)ij, L.m, i, j, L.m, L.n);
    const T rx = L.b[2 * (size_t)ij] - o.x, ry = L.b[2 * (size_t)ij + 1] - o.y;
    L.z[2 * (size_t)ij] = s[36] * rx + s[37] * ry;
    L.z[2 * (size_t)ij + 1] = s[38] * rx + s[39] * ry;
}

// The bottom level in one workgroup: z in LDS (from zero when `zero`),
// `sweeps` four-colour sweeps.  Up to 1024 nodes, each thread keeps its
// node's stencil and right-hand side in registers for all the sweeps; the
// host launches as many threads as nodes (rounded up to whole waves), so
// each of the 4 x sweeps barriers waits on as few waves as it can.
template <typename T>
__global__ void __launch_bounds__(1024) gs_bottom_kernel(BdLevelT<T> L, int sweeps, int zero) {
    extern __shared__ __align__(16) unsigned char bd_smem[];
    T *zs = reinterpret_cast<T *>(bd_smem);
    const int N = L.m * L.n;
    for (int q = threadIdx.x; q < 2 * N; q += blockDim.x) zs[q] = zero ? (T)0 : L.z[q];
    if (N <= (int)blockDim.x) {
        const int q = threadIdx.x, own = q < N;
        const int i = own ? q % L.m : 0, j = own ? q / L.m : 0, col = (i & 1) + 2 * (j & 1);
        T s[40], bx = 0, by = 0;
        lvl_row(L, own ? q : 0, s);
        if (own) {
            bx = L.b[2 * q];
            by = L.b[2 * q + 1];
        }
        __syncthreads();
        for (int sw = 0; sw < sweeps; sw++)
            for (int colour = 0; colour < 4; colour++) {
                if (own && col == colour) {
                    const V2T<T> o = lvl_offdiag_at(s, zs + 2 * q, L.m, i, j, L.m, L.n);
                    const T rx = bx - o.x, ry = by - o.y;
                    zs[2 * q] = s[36] * rx + s[37] * ry;
                    zs[2 * q + 1] = s[38] * rx + s[39] * ry;
                }
                __syncthreads();
            }
    } else {
        __syncthreads();
        for (int sw = 0; sw < sweeps; sw++)
            for (int colour = 0; colour < 4; colour++) {
                const int hm = (L.m + 1 - (colour & 1)) >> 1, hn = (L.n + 1 - (colour >> 1)) >> 1;
                for (int q = threadIdx.x; q < hm * hn; q += blockDim.x) {
                    const int i = 2 * (q % hm) + (colour & 1), j = 2 * (q / hm) + (colour >> 1), ij = i + L.m * j;
                    T s[40];
                    lvl_row(L, ij, s);
                    const V2T<T> o = lvl_offdiag_at(s, zs + 2 * ij, L.m, i, j, L.m, L.n);
                    const T rx = L.b[2 * ij] - o.x, ry = L.b[2 * ij + 1] - o.y;
                    zs[2 * ij] = s[36] * rx + s[37] * ry;
                    zs[2 * ij + 1] = s[38] * rx + s[39] * ry;
                }
                __syncthreads();
            }
    }
    for (int q = threadIdx.x; q < 2 * N; q += blockDim.x) L.z[q] = zs[q];
}

// r = b - A z on the top level; optional per-block |r|^2 partial sums
template <typename T, typename TB>
__global__ void __launch_bounds__(kThreads) resid_top_kernel(BdTop t, const T *__restrict__ z, const TB *__restrict__ b,
                                                             T *__restrict__ r, double *__restrict__ partial) {
    __shared__ double red[kThreads / 64];
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    double acc = 0;
    if (ij < t.m * t.n) {
        const int d = t.rowt[ij];
        const V2T<T> o = top_offdiag<T>(t, z, ij, d);
        const double sc = top_diag_scale(t, d);
        const T rx = (T)b[2 * (size_t)ij] - o.x - (T)(sc * t.cxx) * z[2 * (size_t)ij];
        const T ry = (T)b[2 * (size_t)ij + 1] - o.y - (T)(sc * t.cyy) * z[2 * (size_t)ij + 1];
        if (r) {
            r[2 * (size_t)ij] = rx;
            r[2 * (size_t)ij + 1] = ry;
        }
        acc = (double)rx * rx + (double)ry * ry;
    }
    if (!partial) return;
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kThreads / 64; w++) s += red[w];
        partial[blockIdx.x] = s;
    }
}

template <typename T> __global__ void __launch_bounds__(kThreads) resid_level_kernel(BdLevelT<T> L) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= L.m * L.n) return;
    const int i = ij % L.m, j = ij / L.m;
    T s[40];
    lvl_row(L, ij, s);
    const V2T<T> o = lvl_offdiag_at(s, L.z + 2 * (size_t)ij, L.m, i, j, L.m, L.n);
    const T zx = L.z[2 * (size_t)ij], zy = L.z[2 * (size_t)ij + 1];
    L.r[2 * (size_t)ij] = L.b[2 * (size_t)ij] - o.x - (s[16] * zx + s[17] * zy);
    L.r[2 * (size_t)ij + 1] = L.b[2 * (size_t)ij + 1] - o.y - (s[18] * zx + s[19] * zy);
}

// 1D prolongation parents on a bounded axis (tgmg.in.cc:552-621)
__device__ __forceinline__ int parents(int f, int uf, int sc, int *c, double *w) {
    if (f < 0 || f >= uf) return 0;
    if ((uf & 1) == 0 && f == uf - 1) { c[0] = sc - 1; w[0] = 1; return 1; }
    if ((f & 1) == 0) { c[0] = f / 2; w[0] = 1; return 1; }
    c[0] = f / 2; w[0] = 0.5; c[1] = f / 2 + 1; w[1] = 0.5;
    return 2;
}
// the weight of parent `target` among the np parents (c, w); written out
// so that c and w stay in registers
__device__ __forceinline__ double weight_of(int np, const int *c, const double *w, int target) {
    return np > 0 && c[0] == target ? w[0] : (np > 1 && c[1] == target ? w[1] : 0);
}

// b_c = P^T r_f (restriction with the unnormalised weights 1, 1/2, 1/4)
template <typename T>
__global__ void __launch_bounds__(kThreads) restrict_kernel(const T *__restrict__ rf, int fm, int fn,
                                                            T *__restrict__ bc, int cm, int cn) {
    const int IJ = blockIdx.x * kThreads + threadIdx.x;
    if (IJ >= cm * cn) return;
    const int I = IJ % cm, J = IJ / cm;
    T sx = 0, sy = 0;
    int c[2];
    double w[2];
    for (int fy = 2 * J - 1; fy <= 2 * J + 1; fy++) {
        const double wy = weight_of(parents(fy, fn, cn, c, w), c, w, J);
        if (wy == 0) continue;
        for (int fx = 2 * I - 1; fx <= 2 * I + 1; fx++) {
            const double wx = weight_of(parents(fx, fm, cm, c, w), c, w, I);
            if (wx == 0) continue;
            const T *r = rf + 2 * (size_t)(fx + fm * fy);
            sx += (T)(wx * wy) * r[0];
            sy += (T)(wx * wy) * r[1];
        }
    }
    bc[2 * (size_t)IJ] = sx;
    bc[2 * (size_t)IJ + 1] = sy;
}

// z_f += P z_c
template <typename T>
__global__ void __launch_bounds__(kThreads) prolong_kernel(const T *__restrict__ zc, int cm, int cn,
                                                           T *__restrict__ zf, int fm, int fn) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= fm * fn) return;
    const int i = ij % fm, j = ij / fm;
    int cx[2], cy[2];
    double wx[2], wy[2];
    const int nx = parents(i, fm, cm, cx, wx), ny = parents(j, fn, cn, cy, wy);
    T sx = 0, sy = 0;
    for (int b = 0; b < ny; b++)
        for (int a = 0; a < nx; a++) {
            const T w = (T)(wx[a] * wy[b]);
            const T *z = zc + 2 * (size_t)(cx[a] + cm * cy[b]);
            sx += w * z[0];
            sy += w * z[1];
        }
    zf[2 * (size_t)ij] += sx;
    zf[2 * (size_t)ij + 1] += sy;
}

// ---------------------------------------------------------------------------
// Tiled Gauss-Seidel: one four-colour sweep of a tile in one launch, with
// the result of the colour-by-colour launches.  The tile (64 x 32 nodes in
// FP32, 32 x 32 in FP64) is staged in LDS with a halo of 4 nodes: z, the
// right-hand side (in T) and the row type of every node; colour pass p
// (p = 0..3) updates the nodes of its colour within 3 - p of the tile, which
// read only nodes within 4 - p, all brought up to date by the passes
// before.  z starts from memory (MODE 0), from zero (MODE 1, pre-smoothing a
// zero guess), or from memory plus the coarse correction (MODE 2,
// z_f += P z_c before post-smoothing, tgmg.in.cc:552-621, from the coarse
// nodes under the tile, staged first).  The result goes to another vector
// (zout), since the tiles around read the input's halo while this one
// writes.
// ---------------------------------------------------------------------------
constexpr int kTileThreads = 1024;

// Sums of a1 and a2 over a workgroup of NT threads in a fixed order (the
// lanes by shuffles, then the waves in turn) into partial[idx] and
// partial[kPartialHalf + idx]
template <int NT> __device__ __forceinline__ void group_sums2(double a1, double a2, double *partial, int idx) {
    __shared__ double red[2][NT / 64];
    for (int o = 32; o > 0; o >>= 1) {
        a1 += __shfl_down(a1, o, 64);
        a2 += __shfl_down(a2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a1;
        red[1][threadIdx.x >> 6] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s1 = 0, s2 = 0;
        for (int w = 0; w < NT / 64; w++) {
            s1 += red[0][w];
            s2 += red[1][w];
        }
        partial[idx] = s1;
        partial[kPartialHalf + idx] = s2;
    }
}
// MODE 3: MODE 1 and then the residual on the tile and one node around it
// (which needs z final two nodes out: a halo of 6) and its restriction to
// the coarse nodes under the tile, written to the coarse right-hand side
// cb (tgmg.in.cc:515-551)
// MODE 4: MODE 2 and then y = A z on the tile (the top level's operator in
// FP64, ApplyArgs), which needs z final one node out: a halo of 5
constexpr int halo_of(int mode) { return mode == 3 ? 6 : (mode == 4 ? 5 : 4); }
// (a halo of 5 takes 30 rows, so that its region stays within three staged
// nodes per thread: four held the launch above 64 VGPRs, one tile per CU)
template <typename T, int kHalo> struct Tile {
    static constexpr int X = sizeof(T) == 4 ? 64 : 32, Y = kHalo == 5 ? 30 : 32;
    static constexpr int LW = X + 2 * kHalo, LH = Y + 2 * kHalo, N = LW * LH;
    static constexpr int P = (N + kTileThreads - 1) / kTileThreads;  // staged nodes per thread
    // the coarse nodes under the tile and its halo start at x0/2 - (H+1)/2
    // (the parent of fine node x0 - H) and likewise in y
    static constexpr int CW = LW / 2 + 3, CH = LH / 2 + 3, NC = CW * CH;
    static size_t lds(int mode) {
        return (4 * (size_t)N + (mode == 2 || mode == 4 ? 2 * NC : 0) + kSharedStencils) * sizeof(T) +
               (size_t)N * sizeof(int);
    }
};

// the top level: z in T, b in TB staged as T (qs_multi_bd::mul_a rows).
// In FP32 (the mixed form's preconditioner) an update multiplies by the
// diagonal's reciprocal for its row type, formed once in FP64 and rounded
// (ix*, iy*: interior, box edge, outside, free row); FP64 divides, as
// qs_multi_bd's smoother.
template <typename T, typename TB> struct TopRows {
    BdTop t;
    const TB *b;
    T ix0 = 0, ix1 = 0, ix2 = 0, ix3 = 0, iy0 = 0, iy1 = 0, iy2 = 0, iy3 = 0;
    using Src = TB;
    static constexpr int kOutside = -2;  // no row: nothing is read
    static TopRows make(const BdTop &t, const TB *b) {
        TopRows r{t, b};
        r.ix0 = (T)(1 / t.cxx);
        r.iy0 = (T)(1 / t.cyy);
        r.ix1 = (T)(1 / (t.bcs * t.cxx));
        r.iy1 = (T)(1 / (t.bcs * t.cyy));
        r.ix2 = (T)(1 / (t.ocs * t.cxx));
        r.iy2 = (T)(1 / (t.ocs * t.cyy));
        r.ix3 = (T)(1 / (t.ncs * t.cxx));
        r.iy3 = (T)(1 / (t.ncs * t.cyy));
        return r;
    }
    __device__ int m() const { return t.m; }
    __device__ int n() const { return t.n; }
    __device__ const TB *rhs() const { return b; }
    __device__ const int *types() const { return t.rowt; }
    __device__ void begin(T *) const {}
    __device__ void update(int, int, int l, T *zs, const T *bs, const int *ts, const T *, int LW) const {
        const int d = ts[l];
        const V2T<T> o = top_offdiag_at<T>(t, zs + 2 * l, LW, d);
        if constexpr (sizeof(T) == 4) {
            const T rx = d == -1 ? ix0 : (d == -2 ? ix1 : (d == -3 ? ix2 : ix3));
            const T ry = d == -1 ? iy0 : (d == -2 ? iy1 : (d == -3 ? iy2 : iy3));
            zs[2 * l] = (bs[2 * l] - o.x) * rx;
            zs[2 * l + 1] = (bs[2 * l + 1] - o.y) * ry;
        } else {
            const double sc = top_diag_scale(t, d);
            zs[2 * l] = (bs[2 * l] - o.x) / (T)(sc * t.cxx);
            zs[2 * l + 1] = (bs[2 * l + 1] - o.y) / (T)(sc * t.cyy);
        }
    }
    // y = A z in FP64 at staged node l, as apply_top_dot_kernel (the same
    // sums in the same order; a free row's eight blocks one at a time, which
    // keeps the launch within 64 VGPRs, two tiles per CU)
    __device__ void apply(int l, const T *zs, const int *ts, int LW, double &y0, double &y1) const {
        const int d = ts[l];
        const T *zp = zs + 2 * l;
        V2 o;
        if (d >= 0) {
            const double *a = t.fmats + 32 * (size_t)d;
            const int off[8] = {-LW - 1, -LW, -LW + 1, -1, 1, LW - 1, LW, LW + 1};
            double rx = 0, ry = 0;
#pragma unroll 1
            for (int q = 0; q < 8; q++) {
                const T *zz = zp + 2 * (ptrdiff_t)off[q];
                rx += a[4 * q] * (double)zz[0] + a[4 * q + 1] * (double)zz[1];
                ry += a[4 * q + 2] * (double)zz[0] + a[4 * q + 3] * (double)zz[1];
            }
            o = V2{rx, ry};
        } else {
            o = top_offdiag_at<double, T>(t, zp, LW, d);
        }
        const double sc = top_diag_scale(t, d);
        y0 = o.x + sc * t.cxx * (double)zp[0];
        y1 = o.y + sc * t.cyy * (double)zp[1];
    }
    // b - A z, as resid_top_kernel
    __device__ V2T<T> resid(int, int, int l, const T *zs, const T *bs, const int *ts, const T *, int LW) const {
        const int d = ts[l];
        const V2T<T> o = top_offdiag_at<T>(t, zs + 2 * l, LW, d);
        const double sc = top_diag_scale(t, d);
        return V2T<T>{bs[2 * l] - o.x - (T)(sc * t.cxx) * zs[2 * l],
                      bs[2 * l + 1] - o.y - (T)(sc * t.cyy) * zs[2 * l + 1]};
    }
};

// a coarse level; its two shared stencils are staged too
template <typename T> struct LevelRows {
    BdLevelT<T> L;
    using Src = T;
    static constexpr int kOutside = -1;
    __device__ int m() const { return L.m; }
    __device__ int n() const { return L.n; }
    __device__ const T *rhs() const { return L.b; }
    __device__ const int *types() const { return L.type; }
    __device__ void begin(T *sc) const {
        for (int q = threadIdx.x; q < kSharedStencils; q += blockDim.x) sc[q] = L.Sc[q];
    }
    // the 40 stencil entries are loaded together before they are used
    __device__ void relax(const T *s0, int i, int j, int l, T *zs, const T *bs, int LW) const {
        T s[40];
#pragma unroll
        for (int q = 0; q < 40; q++) s[q] = s0[q];
        const V2T<T> o = lvl_offdiag_at(s, zs + 2 * l, LW, i, j, L.m, L.n);
        const T rx = bs[2 * l] - o.x, ry = bs[2 * l + 1] - o.y;
        zs[2 * l] = s[36] * rx + s[37] * ry;
        zs[2 * l + 1] = s[38] * rx + s[39] * ry;
    }
    __device__ void update(int i, int j, int l, T *zs, const T *bs, const int *ts, const T *sc, int LW) const {
        const int t = ts[l];
        if (t < 0) relax(sc + 40 * (-1 - t), i, j, l, zs, bs, LW);
        else relax(L.S + 40 * (i + (size_t)L.m * j), i, j, l, zs, bs, LW);
    }
    // b - A z, as resid_level_kernel
    __device__ V2T<T> resid_at(const T *s0, int i, int j, int l, const T *zs, const T *bs, int LW) const {
        T s[40];
#pragma unroll
        for (int q = 0; q < 40; q++) s[q] = s0[q];
        const V2T<T> o = lvl_offdiag_at(s, zs + 2 * l, LW, i, j, L.m, L.n);
        const T zx = zs[2 * l], zy = zs[2 * l + 1];
        return V2T<T>{bs[2 * l] - o.x - (s[16] * zx + s[17] * zy), bs[2 * l + 1] - o.y - (s[18] * zx + s[19] * zy)};
    }
    __device__ V2T<T> resid(int i, int j, int l, const T *zs, const T *bs, const int *ts, const T *sc, int LW) const {
        const int t = ts[l];
        if (t < 0) return resid_at(sc + 40 * (-1 - t), i, j, l, zs, bs, LW);
        return resid_at(L.S + 40 * (i + (size_t)L.m * j), i, j, l, zs, bs, LW);
    }
};

template <typename T, int MODE, class Rows>
__device__ __forceinline__ static void gs_tile_body(Rows rows, const T *__restrict__ z,
                                                               T *__restrict__ zout, const T *__restrict__ zc,
                                                               T *__restrict__ cb, int cm, int cn, ApplyArgs ap) {
    constexpr int kHalo = halo_of(MODE);
    using D = Tile<T, kHalo>;
    using TB = typename Rows::Src;
    constexpr int P = D::P;
    extern __shared__ __align__(16) unsigned char bd_tile_smem[];
    T *zs = reinterpret_cast<T *>(bd_tile_smem), *bs = zs + 2 * D::N, *cs = bs + 2 * D::N,
      *sc = cs + (MODE == 2 || MODE == 4 ? 2 * D::NC : 0);
    int *ts = reinterpret_cast<int *>(sc + kSharedStencils);
    const int m = rows.m(), n = rows.n(), tid = threadIdx.x;
    const int x0 = blockIdx.x * D::X, y0 = blockIdx.y * D::Y, cx0 = x0 / 2 - (kHalo + 1) / 2,
              cy0 = y0 / 2 - (kHalo + 1) / 2;
    const TB *b = rows.rhs();
    const int *tp = rows.types();
    // staged nodes q = tid + k kTileThreads; every load of a thread is issued
    // before the first is used
    T vx[P], vy[P];
    TB bx[P], by[P];
    int tt[P];
#pragma unroll
    for (int k = 0; k < P; k++) {
        const int q = tid + k * kTileThreads, i = x0 - kHalo + q % D::LW, j = y0 - kHalo + q / D::LW;
        const bool in = q < D::N && i >= 0 && i < m && j >= 0 && j < n;
        const size_t ij = in ? i + (size_t)m * j : 0;
        if (MODE == 0 || MODE == 2 || MODE == 4) {
            const T a0 = z[2 * ij], a1 = z[2 * ij + 1];
            vx[k] = in ? a0 : (T)0;
            vy[k] = in ? a1 : (T)0;
        } else {
            vx[k] = vy[k] = 0;
        }
        const TB b0 = b[2 * ij], b1 = b[2 * ij + 1];
        const int t0 = tp[ij];
        bx[k] = in ? b0 : (TB)0;
        by[k] = in ? b1 : (TB)0;
        tt[k] = in ? t0 : Rows::kOutside;
    }
    if (MODE == 2 || MODE == 4) {
        constexpr int PC = (D::NC + kTileThreads - 1) / kTileThreads;
        T cx[PC], cy[PC];
#pragma unroll
        for (int k = 0; k < PC; k++) {
            const int q = tid + k * kTileThreads, I = cx0 + q % D::CW, J = cy0 + q / D::CW;
            const bool in = q < D::NC && I >= 0 && I < cm && J >= 0 && J < cn;
            const size_t IJ = in ? I + (size_t)cm * J : 0;
            const T a0 = zc[2 * IJ], a1 = zc[2 * IJ + 1];
            cx[k] = in ? a0 : (T)0;
            cy[k] = in ? a1 : (T)0;
        }
#pragma unroll
        for (int k = 0; k < PC; k++) {
            const int q = tid + k * kTileThreads;
            if (q < D::NC) {
                cs[2 * q] = cx[k];
                cs[2 * q + 1] = cy[k];
            }
        }
        __syncthreads();
    }
    rows.begin(sc);
#pragma unroll
    for (int k = 0; k < P; k++) {
        const int q = tid + k * kTileThreads, i = x0 - kHalo + q % D::LW, j = y0 - kHalo + q / D::LW;
        if (q >= D::N) continue;
        T wxv = vx[k], wyv = vy[k];
        if ((MODE == 2 || MODE == 4) && i >= 0 && i < m && j >= 0 && j < n) {  // as prolong_kernel
            int px[2], py[2];
            double wx[2], wy[2];
            const int nx = parents(i, m, cm, px, wx), ny = parents(j, n, cn, py, wy);
            T sx = 0, sy = 0;
            for (int bb = 0; bb < ny; bb++)
                for (int aa = 0; aa < nx; aa++) {
                    const T w = (T)(wx[aa] * wy[bb]);
                    const T *zz = cs + 2 * ((px[aa] - cx0) + D::CW * (py[bb] - cy0));
                    sx += w * zz[0];
                    sy += w * zz[1];
                }
            wxv += sx;
            wyv += sy;
        }
        zs[2 * q] = wxv;
        zs[2 * q + 1] = wyv;
        bs[2 * q] = (T)bx[k];
        bs[2 * q + 1] = (T)by[k];
        ts[q] = tt[k];
    }
    __syncthreads();
    for (int c = 0; c < 4; c++) {
        const int h = kHalo - 1 - c;
        const int ax = max(x0 - h, 0), ex = min(x0 + D::X + h, m), ay = max(y0 - h, 0), ey = min(y0 + D::Y + h, n);
        const int ix = ax + ((ax ^ c) & 1), jy = ay + ((ay ^ (c >> 1)) & 1);
        const int nx = ex > ix ? (ex - ix + 1) >> 1 : 0, ny = ey > jy ? (ey - jy + 1) >> 1 : 0;
        for (int q = tid; q < nx * ny; q += kTileThreads) {
            const int i = ix + 2 * (q % nx), j = jy + 2 * (q / nx);
            rows.update(i, j, (i - x0 + kHalo) + D::LW * (j - y0 + kHalo), zs, bs, ts, sc, D::LW);
        }
        __syncthreads();
    }
    double a1 = 0, a2 = 0;  // MODE 4: the tile's sums of (u, y) and (y, y)
    for (int q = tid; q < D::X * D::Y; q += kTileThreads) {
        const int i = x0 + q % D::X, j = y0 + q / D::X;
        if (i >= m || j >= n) continue;
        const int l = (i - x0 + kHalo) + D::LW * (j - y0 + kHalo);
        const size_t k = 2 * (i + (size_t)m * j);
        zout[k] = zs[2 * l];
        zout[k + 1] = zs[2 * l + 1];
    }
    if constexpr (MODE == 4) {
#pragma unroll 1
        for (int q = tid; q < D::X * D::Y; q += kTileThreads) {
            const int i = x0 + q % D::X, j = y0 + q / D::X;
            if (i >= m || j >= n) continue;
            const int l = (i - x0 + kHalo) + D::LW * (j - y0 + kHalo);
            const size_t k = 2 * (i + (size_t)m * j);
            double y0v, y1v;
            rows.apply(l, zs, ts, D::LW, y0v, y1v);
            ap.y[k] = y0v;
            ap.y[k + 1] = y1v;
            a1 += ap.u[k] * y0v + ap.u[k + 1] * y1v;
            if (ap.yy) a2 += y0v * y0v + y1v * y1v;
        }
        // the tile's sums in a fixed order (lanes by shuffles, then the
        // waves in turn) through the right-hand side's LDS, which no pass
        // reads any more
        for (int o = 32; o > 0; o >>= 1) {
            a1 += __shfl_down(a1, o, 64);
            a2 += __shfl_down(a2, o, 64);
        }
        double *red = reinterpret_cast<double *>(bs);
        if ((tid & 63) == 0) {
            red[tid >> 6] = a1;
            red[kTileThreads / 64 + (tid >> 6)] = a2;
        }
        __syncthreads();
        if (tid == 0) {
            double s1 = 0, s2 = 0;
            for (int w = 0; w < kTileThreads / 64; w++) {
                s1 += red[w];
                s2 += red[kTileThreads / 64 + w];
            }
            const int tile = blockIdx.x + gridDim.x * blockIdx.y;
            ap.partial[tile] = s1;
            ap.partial[kPartialHalf + tile] = s2;
        }
    }
    if (MODE != 3) return;
    // the residual on the tile and one node around it, over b in LDS (a
    // node's residual reads its own b and the z around it)
    {
        constexpr int RW = D::X + 2, RH = D::Y + 2;
        for (int q = tid; q < RW * RH; q += kTileThreads) {
            const int i = x0 - 1 + q % RW, j = y0 - 1 + q / RW;
            if (i < 0 || i >= m || j < 0 || j >= n) continue;
            const int l = (i - x0 + kHalo) + D::LW * (j - y0 + kHalo);
            const V2T<T> r = rows.resid(i, j, l, zs, bs, ts, sc, D::LW);
            bs[2 * l] = r.x;
            bs[2 * l + 1] = r.y;
        }
    }
    __syncthreads();
    // the coarse nodes under the tile (the last tile of a row or column also
    // takes the coarse node past the fine grid's end, when there is one), as
    // restrict_kernel: b_c = P^T r
    const int I0 = x0 / 2, I1 = blockIdx.x == gridDim.x - 1 ? cm : min(cm, I0 + D::X / 2);
    const int J0 = y0 / 2, J1 = blockIdx.y == gridDim.y - 1 ? cn : min(cn, J0 + D::Y / 2);
    const int nI = I1 - I0;
    for (int q = tid; q < nI * (J1 - J0); q += kTileThreads) {
        const int I = I0 + q % nI, J = J0 + q / nI;
        T sx = 0, sy = 0;
        int c[2];
        double w[2];
        for (int fy = 2 * J - 1; fy <= 2 * J + 1; fy++) {
            const double wy = weight_of(parents(fy, n, cn, c, w), c, w, J);
            if (wy == 0) continue;
            for (int fx = 2 * I - 1; fx <= 2 * I + 1; fx++) {
                const double wx = weight_of(parents(fx, m, cm, c, w), c, w, I);
                if (wx == 0) continue;
                const T *r = bs + 2 * ((fx - x0 + kHalo) + D::LW * (fy - y0 + kHalo));
                sx += (T)(wx * wy) * r[0];
                sy += (T)(wx * wy) * r[1];
            }
        }
        cb[2 * (I + (size_t)cm * J)] = sx;
        cb[2 * (I + (size_t)cm * J) + 1] = sy;
    }
}

template <typename T, int MODE, class Rows>
__global__ void __launch_bounds__(kTileThreads) gs_tile_kernel(Rows rows, const T *__restrict__ z,
                                                               T *__restrict__ zout, const T *__restrict__ zc,
                                                               T *__restrict__ cb, int cm, int cn, ApplyArgs ap) {
    gs_tile_body<T, MODE, Rows>(rows, z, zout, zc, cb, cm, cn, ap);
}
// MODE 4 (the top level's last post-smoothing sweep with the operator
// application).  Two tiles per CU need 64 VGPRs or fewer (eight waves per
// SIMD); the hint keeps the compiler's schedule within that (104 without
// it, one tile per CU)
template <typename T, class Rows>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
gs_tile_apply_kernel(Rows rows, const T *__restrict__ z, T *__restrict__ zout, const T *__restrict__ zc,
                     T *__restrict__ cb, int cm, int cn, ApplyArgs ap) {
    gs_tile_body<T, 4, Rows>(rows, z, zout, zc, cb, cm, cn, ap);
}

// Sources of fine-level entries for the Galerkin product: the top operator,
// a stored coarse level, or (for a shared stencil) a level whose every row
// is its interior row (KIND -1) or its outside row (KIND -2)
struct TopSrc {
    BdTop t;
    __device__ int m() const { return t.m; }
    __device__ int n() const { return t.n; }
    __device__ bool interior(int f) const { return t.rowt[f] == -1; }
    __device__ bool outside(int f) const { return t.rowt[f] == -3; }
    __device__ void entry(int f, int k, double *A) const { top_entry(t, f, k, A); }
};
struct LvlSrc {
    BdLevel F;
    __device__ int m() const { return F.m; }
    __device__ int n() const { return F.n; }
    __device__ bool interior(int f) const { return F.type[f] == -1; }
    __device__ bool outside(int f) const { return F.type[f] == -2; }
    __device__ void entry(int f, int k, double *A) const {
        const double *s = lvl_stencil(F, f) + 4 * k;
        A[0] = s[0]; A[1] = s[1]; A[2] = s[2]; A[3] = s[3];
    }
};
template <int KIND> struct TopConstSrc {
    BdTop t;
    __device__ int m() const { return t.m; }
    __device__ int n() const { return t.n; }
    __device__ bool interior(int) const { return KIND == -1; }
    __device__ bool outside(int) const { return KIND == -2; }
    __device__ void entry(int, int k, double *A) const {
        // rows of type -1 (interior) and -3 (outside) do not depend on the
        // row index
        BdTop u = t;
        const int d = KIND == -1 ? -1 : -3;
        u.rowt = &d;
        top_entry(u, 0, k, A);
    }
};
template <int KIND> struct LvlConstSrc {
    BdLevel F;
    __device__ int m() const { return F.m; }
    __device__ int n() const { return F.n; }
    __device__ bool interior(int) const { return KIND == -1; }
    __device__ bool outside(int) const { return KIND == -2; }
    __device__ void entry(int, int k, double *A) const {
        const double *s = F.Sc + 40 * (-1 - KIND) + 4 * k;
        A[0] = s[0]; A[1] = s[1]; A[2] = s[2]; A[3] = s[3];
    }
};

// Galerkin coarse block row at coarse node (I,J): sum over fine nodes f
// with weight P(f,(I,J)) of P(f,.)^T A(f,g) P(g,.), then the inverse of cc.
// Returns -1 when every fine row used was an interior row, -2 when every
// one was an outside row, 0 otherwise.
template <class Src>
__device__ int rap_node(const Src &src, int cm, int cn, int I, int J, double *o) {
    const int fm = src.m(), fn = src.n();
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0;
    bool all_interior = true, all_outside = true;
    int c[2], cx[2], cy[2];
    double w[2], wx[2], wy[2];
    for (int fy = 2 * J - 1; fy <= 2 * J + 1; fy++) {
        const double wfy = weight_of(parents(fy, fn, cn, c, w), c, w, J);
        if (wfy == 0) continue;
        for (int fx = 2 * I - 1; fx <= 2 * I + 1; fx++) {
            const double wfx = weight_of(parents(fx, fm, cm, c, w), c, w, I);
            if (wfx == 0) continue;
            const int fij = fx + fm * fy;
            all_interior = all_interior && src.interior(fij);
            all_outside = all_outside && src.outside(fij);
            for (int k = 0; k < 9; k++) {
                const int gx = fx + k % 3 - 1, gy = fy + k / 3 - 1;
                if (gx < 0 || gx >= fm || gy < 0 || gy >= fn) continue;
                double A[4];
                src.entry(fij, k, A);
                if (A[0] == 0 && A[1] == 0 && A[2] == 0 && A[3] == 0) continue;
                const int ny = parents(gy, fn, cn, cy, wy), nx = parents(gx, fm, cm, cx, wx);
                // every index below is a compile-time one after unrolling,
                // so acc stays in registers; each entry's sum runs in the
                // order of the parents (q, then p)
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int p = 0; p < 2; p++) {
                        if (q >= ny || p >= nx) continue;
                        const int dI = cx[p] - I, dJ = cy[q] - J;
                        if (dI < -1 || dI > 1 || dJ < -1 || dJ > 1) continue;
                        const double ww = wfx * wfy * wx[p] * wy[q];
                        const int blk = (dI + 1) + 3 * (dJ + 1);
#pragma unroll
                        for (int bb = 0; bb < 9; bb++)
                            if (bb == blk)
#pragma unroll
                                for (int r = 0; r < 4; r++) acc[4 * bb + r] += ww * A[r];
                    }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 36; q++) o[q] = acc[q];
    const double a = acc[16], b = acc[17], cc = acc[18], d = acc[19];
    const double idet = 1. / (a * d - b * cc);
    o[36] = d * idet;
    o[37] = -b * idet;
    o[38] = -cc * idet;
    o[39] = a * idet;
    return all_interior ? -1 : (all_outside ? -2 : 0);
}

// The kind rap_node returns, from the same fine rows, without the product
template <class Src> __device__ int rap_kind(const Src &src, int cm, int cn, int I, int J) {
    const int fm = src.m(), fn = src.n();
    bool all_interior = true, all_outside = true;
    int c[2];
    double w[2];
    for (int fy = 2 * J - 1; fy <= 2 * J + 1; fy++) {
        if (weight_of(parents(fy, fn, cn, c, w), c, w, J) == 0) continue;
        for (int fx = 2 * I - 1; fx <= 2 * I + 1; fx++) {
            if (weight_of(parents(fx, fm, cm, c, w), c, w, I) == 0) continue;
            const int fij = fx + fm * fy;
            all_interior = all_interior && src.interior(fij);
            all_outside = all_outside && src.outside(fij);
        }
    }
    return all_interior ? -1 : (all_outside ? -2 : 0);
}

// Every row's stencil: a row whose fine rows were all interior rows takes
// the shared interior stencil (its product is not formed), one whose fine
// rows were all outside rows takes the shared outside stencil if its
// product equals it bit for bit (near the level's edges the prolongation
// weights differ); the others store their own product.  The shared
// stencils are formed first (rap_const_kernel).
template <class Src>
__global__ void __launch_bounds__(kThreads) rap_kernel(Src src, BdLevel C, int shared) {
    const int IJ = blockIdx.x * kThreads + threadIdx.x;
    if (IJ >= C.m * C.n) return;
    const int I = IJ % C.m, J = IJ / C.m, kind = shared ? rap_kind(src, C.m, C.n, I, J) : 0;
    if (kind == -1) {  // the shared interior stencil, no product needed
        C.type[IJ] = -1;
        return;
    }
    double o[40];
    (void)rap_node(src, C.m, C.n, I, J, o);
    int type = 0;
    if (kind == -2) {
        bool same = true;
        for (int q = 0; q < 40; q++) same = same && o[q] == C.Sc[40 + q];
        type = same ? -2 : 0;
    }
    C.type[IJ] = type;
    if (type < 0) return;
    double *S = C.S + 40 * (size_t)IJ;
    for (int q = 0; q < 40; q++) S[q] = o[q];
}

// a shared stencil of a level (KIND -1 interior, -2 outside): the product
// at a node whose footprint lies away from the level's edges, every fine row
// of the kind (same arithmetic as rap_kernel)
template <int KIND, class Src>
__global__ void rap_const_kernel(Src src, BdLevel C) {
    if (threadIdx.x != 0) return;
    double o[40];
    (void)rap_node(src, C.m, C.n, 2, 2, o);
    for (int q = 0; q < 40; q++) C.Sc[40 * (-1 - KIND) + q] = o[q];
}

template <typename T> __global__ void __launch_bounds__(kThreads) fill_kernel(T *p, long n, T v) {
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q < n) p[q] = v;
}

// precision changes at the mixed preconditioner's boundary
template <typename S, typename D>
__global__ void __launch_bounds__(kThreads) convert_kernel(const S *__restrict__ a, D *__restrict__ b, long n) {
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q < n) b[q] = (D)a[q];
}

// ---------------------------------------------------------------------------
// BiCGStab vector kernels; scalars live on the device (sc[]):
// 0 rho, 1 rho_prev, 2 alpha, 3 omega, 4 (rhat,v), 5 (t,s), 6 (t,t), 7 (r,r)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) dot_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                                       long n, double *__restrict__ partial) {
    __shared__ double red[kThreads / 64];
    double acc = 0;
    for (long q = (long)blockIdx.x * kThreads + threadIdx.x; q < n; q += (long)gridDim.x * kThreads)
        acc += a[q] * b[q];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kThreads / 64; w++) s += red[w];
        partial[blockIdx.x] = s;
    }
}

// sum partials into sc[slot]; then the step's scalar update:
// op 2: alpha = rho/(rhat,v); op 3: omega = (t,s)/(t,t); op 0: none
__global__ void __launch_bounds__(kThreads) finish_dot_kernel(const double *__restrict__ partial, int np,
                                                              double *sc, int slot, int op,
                                                              volatile double *host) {
    __shared__ double red[kThreads / 64];
    double acc = 0;
    for (int q = threadIdx.x; q < np; q += kThreads) acc += partial[q];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s = 0;
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
    sc[slot] = s;
    if (op == 2) sc[2] = sc[0] / sc[4];
    else if (op == 3) sc[3] = sc[5] / sc[6];
    if (host) *host = s;
}

// Products fused into the kernels that form one of their factors.  Block
// sums go to partial[block] and partial[kPartialHalf + block].
__device__ __forceinline__ void block_sums2(double a1, double a2, double *partial) {
    group_sums2<kThreads>(a1, a2, partial, blockIdx.x);
}

// y = A x on the top level in FP64 (x FP64, or the mixed form's FP32
// preconditioned vector), with the sums of (u, y) and, when yy, of (y, y)
template <typename TX>
__global__ void __launch_bounds__(kThreads) apply_top_dot_kernel(BdTop t, const TX *__restrict__ x,
                                                                 double *__restrict__ y,
                                                                 const double *__restrict__ u, int yy,
                                                                 double *__restrict__ partial) {
    double a1 = 0, a2 = 0;
    const int mn = t.m * t.n;
    for (int ij = blockIdx.x * kThreads + threadIdx.x; ij < mn; ij += gridDim.x * kThreads) {
        const int d = t.rowt[ij];
        const V2 o = top_offdiag<double, TX>(t, x, ij, d);
        const double sc = top_diag_scale(t, d);
        const double y0 = o.x + sc * t.cxx * (double)x[2 * (size_t)ij];
        const double y1 = o.y + sc * t.cyy * (double)x[2 * (size_t)ij + 1];
        y[2 * (size_t)ij] = y0;
        y[2 * (size_t)ij + 1] = y1;
        a1 += u[2 * (size_t)ij] * y0 + u[2 * (size_t)ij + 1] * y1;
        if (yy) a2 += y0 * y0 + y1 * y1;
    }
    block_sums2(a1, a2, partial);
}

// the two sums of a fused kernel into sc[slot1] and (slot2 >= 0) sc[slot2],
// host copies where given, then the step's scalar update (op as
// finish_dot_kernel)
__global__ void __launch_bounds__(kThreads) finish_dots_kernel(const double *__restrict__ partial, int np,
                                                               double *sc, int slot1, int slot2, int op,
                                                               volatile double *host1, volatile double *host2) {
    __shared__ double red[2][kThreads / 64];
    double a1 = 0, a2 = 0;
    for (int q = threadIdx.x; q < np; q += kThreads) {
        a1 += partial[q];
        a2 += partial[kPartialHalf + q];
    }
    for (int o = 32; o > 0; o >>= 1) {
        a1 += __shfl_down(a1, o, 64);
        a2 += __shfl_down(a2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a1;
        red[1][threadIdx.x >> 6] = a2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s1 = 0, s2 = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        s1 += red[0][w];
        s2 += red[1][w];
    }
    sc[slot1] = s1;
    if (slot2 >= 0) sc[slot2] = s2;
    if (op == 2) sc[2] = sc[0] / sc[4];
    else if (op == 3) sc[3] = sc[5] / sc[6];
    if (host1) *host1 = s1;
    if (host2) *host2 = s2;
}

// p = r + beta (p - omega v), beta = (rho/rho_prev)(alpha/omega)
__global__ void __launch_bounds__(kThreads) bicg_p_kernel(double *p, const double *__restrict__ r,
                                                          const double *__restrict__ v, const double *sc, long n,
                                                          int first) {
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= n) return;
    const double beta = (sc[0] / sc[1]) * (sc[2] / sc[3]);
    p[q] = first ? r[q] : r[q] + beta * (p[q] - sc[3] * v[q]);
}
// s = r - alpha v
__global__ void __launch_bounds__(kThreads) bicg_s_kernel(double *s, const double *__restrict__ r,
                                                          const double *__restrict__ v, const double *sc, long n) {
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q < n) s[q] = r[q] - sc[2] * v[q];
}
// x += alpha phat + omega shat; r = s - omega t; rho_prev = rho; with the
// sums of (r, r) and (r_hat, r)
template <typename TP>
__global__ void __launch_bounds__(kThreads) bicg_x_dot_kernel(double *x, double *r, const TP *__restrict__ ph,
                                                              const TP *__restrict__ sh,
                                                              const double *__restrict__ s,
                                                              const double *__restrict__ t,
                                                              const double *__restrict__ rh, double *sc, long n,
                                                              double *__restrict__ partial) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[1] = sc[0];
    const double a = sc[2], w = sc[3];
    double a1 = 0, a2 = 0;
    const long stride = (long)gridDim.x * kThreads;
    long q = (long)blockIdx.x * kThreads + threadIdx.x;
    // four elements at a time, their loads issued together
    for (; q + 3 * stride < n; q += 4 * stride) {
        double xv[4], pv[4], sv[4], s_[4], tv[4], hv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long e = q + u * stride;
            xv[u] = x[e];
            pv[u] = (double)ph[e];
            sv[u] = (double)sh[e];
            s_[u] = s[e];
            tv[u] = t[e];
            hv[u] = rh[e];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long e = q + u * stride;
            x[e] = xv[u] + (a * pv[u] + w * sv[u]);
            const double rq = s_[u] - w * tv[u];
            r[e] = rq;
            a1 += rq * rq;
            a2 += hv[u] * rq;
        }
    }
    for (; q < n; q += stride) {
        x[q] += a * (double)ph[q] + w * (double)sh[q];
        const double rq = s[q] - w * t[q];
        r[q] = rq;
        a1 += rq * rq;
        a2 += rh[q] * rq;
    }
    block_sums2(a1, a2, partial);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int BdSolver::dims_for(int tm, int tn, std::vector<std::pair<int, int>> &out) {
    // tgmg.in.cc:13-29 with tgmg_grid_min = 10, then two levels fewer
    out.clear();
    int am = (tm + 2) >> 1, an = (tn + 2) >> 1;
    std::vector<std::pair<int, int>> all;
    while ((long)am * an > 10) {
        all.push_back({am, an});
        const int nm = (am + 2) >> 1, nn = (an + 2) >> 1;
        am = nm;
        an = nn;
    }
    const int keep = (int)all.size() - 2;
    for (int l = 0; l < keep; l++) out.push_back(all[l]);
    return (int)out.size();
}

int BdSolver::init(int tm, int tn) {
    const bool was_mixed = mixed;
    free_all();
    top_m = tm;
    top_n = tn;
    std::vector<std::pair<int, int>> d;
    nl = dims_for(tm, tn, d);
    const size_t nv = 2 * (size_t)tm * tn;
    for (int k = 0; k < kBdVecs; k++) STZ_HIP(hipMalloc(&vec[k], nv * sizeof(double)));
    STZ_HIP(hipMalloc(&partial, 2 * (size_t)kPartialHalf * sizeof(double)));
    STZ_HIP(hipMalloc(&sc, 16 * sizeof(double)));
    STZ_HIP(hipHostMalloc((void **)&hscal, 4 * sizeof(double), hipHostMallocCoherent));
    for (int l = 0; l < nl; l++) {
        BdLevel &L = lv[l];
        L.m = d[l].first;
        L.n = d[l].second;
        const size_t n2 = 2 * (size_t)L.m * L.n;
        STZ_HIP(hipMalloc(&L.z, n2 * sizeof(double)));
        STZ_HIP(hipMalloc(&L.b, n2 * sizeof(double)));
        STZ_HIP(hipMalloc(&L.r, n2 * sizeof(double)));
        STZ_HIP(hipMalloc(&L.S, 20 * n2 * sizeof(double)));
        STZ_HIP(hipMalloc(&L.type, (size_t)L.m * L.n * sizeof(int)));
        STZ_HIP(hipMalloc(&L.Sc, kSharedStencils * sizeof(double)));
    }
    return was_mixed ? set_mixed(true) : STZ_OK;
}

int BdSolver::alloc_mixed() {
    const size_t nv = 2 * (size_t)top_m * top_n;
    for (int k = 0; k < 4; k++) STZ_HIP(hipMalloc(&top32[k], nv * sizeof(float)));
    for (int l = 0; l < nl; l++) {
        BdLevelT<float> &L = lv32[l];
        L.m = lv[l].m;
        L.n = lv[l].n;
        const size_t n2 = 2 * (size_t)L.m * L.n;
        STZ_HIP(hipMalloc(&L.z, n2 * sizeof(float)));
        STZ_HIP(hipMalloc(&L.b, n2 * sizeof(float)));
        STZ_HIP(hipMalloc(&L.r, n2 * sizeof(float)));
        STZ_HIP(hipMalloc(&L.S, 20 * n2 * sizeof(float)));
        STZ_HIP(hipMalloc(&L.Sc, kSharedStencils * sizeof(float)));
        L.type = lv[l].type;
    }
    return STZ_OK;
}

void BdSolver::free_mixed() {
    for (int k = 0; k < 4; k++) {
        if (top32[k]) (void)hipFree(top32[k]);
        top32[k] = nullptr;
    }
    for (int l = 0; l < kMaxLevels; l++) {
        BdLevelT<float> &L = lv32[l];
        for (float *q : {L.z, L.b, L.r, L.S, L.Sc})
            if (q) (void)hipFree(q);
        L = BdLevelT<float>{};
    }
}

// FP32 copies for the mixed form's preconditioner (allocated once; the
// stencils are rounded from the FP64 hierarchy at every setup)
int BdSolver::set_mixed(bool on) {
    if (on && !top32[0]) {
        const int e = alloc_mixed();
        if (e) {
            free_mixed();
            return e;
        }
    }
    if (!on) free_mixed();
    mixed = on;
    return STZ_OK;
}

void BdSolver::free_all() {
    free_mixed();
    for (int k = 0; k < kBdVecs; k++) {
        if (vec[k]) (void)hipFree(vec[k]);
        vec[k] = nullptr;
    }
    for (int l = 0; l < kMaxLevels; l++) {
        BdLevel &L = lv[l];
        if (L.z) (void)hipFree(L.z);
        if (L.b) (void)hipFree(L.b);
        if (L.r) (void)hipFree(L.r);
        if (L.S) (void)hipFree(L.S);
        if (L.type) (void)hipFree(L.type);
        if (L.Sc) (void)hipFree(L.Sc);
        L = BdLevel{};
    }
    if (partial) (void)hipFree(partial);
    if (sc) (void)hipFree(sc);
    if (hscal) (void)hipHostFree(hscal);
    partial = sc = nullptr;
    hscal = nullptr;
    nl = 0;
}

int BdSolver::setup(const BdTop &t, hipStream_t s) {
    top = t;
    for (int l = 0; l < nl; l++) {
        const dim3 g = blocks((long)lv[l].m * lv[l].n);
        // the shared stencil needs a fine footprint of 7x7 nodes
        const bool cst = (l == 0 ? top_m : lv[l - 1].m) >= 7 && (l == 0 ? top_n : lv[l - 1].n) >= 7;
        // the shared stencils first: rap_kernel compares outside rows with one
        if (l == 0) {
            if (cst) {
                hipLaunchKernelGGL((rap_const_kernel<-1, TopConstSrc<-1>>), dim3(1), dim3(64), 0, s,
                                   TopConstSrc<-1>{top}, lv[0]);
                hipLaunchKernelGGL((rap_const_kernel<-2, TopConstSrc<-2>>), dim3(1), dim3(64), 0, s,
                                   TopConstSrc<-2>{top}, lv[0]);
            }
            hipLaunchKernelGGL(rap_kernel<TopSrc>, g, dim3(kThreads), 0, s, TopSrc{top}, lv[0], cst ? 1 : 0);
        } else {
            if (cst) {
                hipLaunchKernelGGL((rap_const_kernel<-1, LvlConstSrc<-1>>), dim3(1), dim3(64), 0, s,
                                   LvlConstSrc<-1>{lv[l - 1]}, lv[l]);
                hipLaunchKernelGGL((rap_const_kernel<-2, LvlConstSrc<-2>>), dim3(1), dim3(64), 0, s,
                                   LvlConstSrc<-2>{lv[l - 1]}, lv[l]);
            }
            hipLaunchKernelGGL(rap_kernel<LvlSrc>, g, dim3(kThreads), 0, s, LvlSrc{lv[l - 1]}, lv[l], cst ? 1 : 0);
        }
        STZ_HIP(hipGetLastError());
    }
    if (mixed)
        for (int l = 0; l < nl; l++) {
            const long ns = 40L * lv[l].m * lv[l].n;
            hipLaunchKernelGGL((convert_kernel<double, float>), blocks(ns), dim3(kThreads), 0, s, lv[l].S, lv32[l].S, ns);
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3(1), dim3(kThreads), 0, s, lv[l].Sc, lv32[l].Sc,
                               (long)kSharedStencils);
            STZ_HIP(hipGetLastError());
        }
    return STZ_OK;
}

// one sweep of the tiled kernel over an m x n level from zin into zout;
// mode 1 starts from zero, mode 2 adds the correction of the cm x cn level
// zc first
// one launch of gs_tile_kernel<T, MODE, Rows> with `lds` bytes of LDS (the
// kernel's limit raised once where that is above 64 KiB)
template <typename T, int MODE, class Rows>
static int tile_launch(dim3 g, size_t lds, hipStream_t s, const Rows &rows, const T *zin, T *zout, const T *zc,
                       T *cb, int cm, int cn, const ApplyArgs &ap = ApplyArgs{}) {
    void (*kern)(Rows, const T *, T *, const T *, T *, int, int, ApplyArgs);
    if constexpr (MODE == 4) kern = &gs_tile_apply_kernel<T, Rows>;
    else kern = &gs_tile_kernel<T, MODE, Rows>;
    static const hipError_t lim =
        lds > 65536 ? hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                    : hipSuccess;
    STZ_HIP(lim);
    hipLaunchKernelGGL(kern, g, dim3(kTileThreads), lds, s, rows, zin, zout, zc, cb, cm, cn, ap);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T, class Rows>
static int tile_sweep(const Rows &rows, int m, int n, const T *zin, T *zout, int mode, const T *zc, int cm, int cn,
                      hipStream_t s, T *cb = nullptr) {
    using D = Tile<T, 4>;
    const dim3 g((m + D::X - 1) / D::X, (n + D::Y - 1) / D::Y);
    if (mode == 3) return tile_launch<T, 3>(g, Tile<T, 6>::lds(1), s, rows, zin, zout, zc, cb, cm, cn);
    if (mode == 1) return tile_launch<T, 1>(g, D::lds(1), s, rows, zin, zout, zc, cb, cm, cn);
    if (mode == 2) return tile_launch<T, 2>(g, D::lds(2), s, rows, zin, zout, zc, cb, cm, cn);
    return tile_launch<T, 0>(g, D::lds(0), s, rows, zin, zout, zc, cb, cm, cn);
}

// the top level's last post-smoothing sweep (mode 2) with the operator
// application fused (mode 4); returns the partial sums written through
// *tiles
template <typename T, typename TB>
static int tile_sweep_apply(const TopRows<T, TB> &rows, int m, int n, const T *zin, T *zout, const T *zc, int cm,
                            int cn, const ApplyArgs &ap, int *tiles, hipStream_t s) {
    using D = Tile<T, 5>;
    const dim3 g((m + D::X - 1) / D::X, (n + D::Y - 1) / D::Y);
    *tiles = (int)(g.x * g.y);
    return tile_launch<T, 4>(g, Tile<T, 5>::lds(4), s, rows, zin, zout, zc, (T *)nullptr, cm, cn, ap);
}

// the start of a smoothing that the tiled kernel does not take, in zout:
// zero (mode 1), zin (mode 0), zin + P zc (mode 2)
template <typename T>
static int start_in_place(T *zout, const T *zin, int m, int n, int mode, const T *zc, int cm, int cn, hipStream_t s) {
    const long n2 = 2L * m * n;
    if (mode == 1) {
        hipLaunchKernelGGL(fill_kernel<T>, blocks(n2), dim3(kThreads), 0, s, zout, n2, (T)0);
    } else if (zin != zout) {
        STZ_HIP(hipMemcpyAsync(zout, zin, n2 * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    if (mode == 2)
        hipLaunchKernelGGL(prolong_kernel<T>, blocks((long)m * n), dim3(kThreads), 0, s, zc, cm, cn, zout, m, n);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// Smoothing of the top level from zin (mode as tile_sweep) into zout
// (zout != zin): one sweep by the tiled kernel, more by one launch per
// colour in place
template <typename T, typename TB>
int BdSolver::smooth_top(T *zout, const TB *b, int sweeps, int mode, const T *zin, const T *zc, int cm, int cn,
                         hipStream_t s) {
    if (sweeps == 1) return tile_sweep(TopRows<T, TB>::make(top, b), top_m, top_n, zin, zout, mode, zc, cm, cn, s);
    const int e = start_in_place(zout, zin, top_m, top_n, mode, zc, cm, cn, s);
    if (e) return e;
    for (int k = 0; k < sweeps; k++)
        for (int c = 0; c < 4; c++) {
            const long cnt = (long)((top_m + 1 - (c & 1)) >> 1) * ((top_n + 1 - (c >> 1)) >> 1);
            hipLaunchKernelGGL((gs_top_kernel<T, TB>), blocks(cnt), dim3(kThreads), 0, s, top, zout, b, c);
        }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// Smoothing of a coarse level from zin into zout, as smooth_top; the bottom
// level, when it fits in LDS, in one workgroup in place
template <typename T>
int BdSolver::smooth_level(const BdLevelT<T> &L, bool bottom, int sweeps, int mode, const T *zin, T *zout,
                           const T *zc, int cm, int cn, hipStream_t s) {
    const size_t lds = 2 * (size_t)L.m * L.n * sizeof(T);
    if (bottom && lds <= 64 * 1024) {
        // a zero start is the kernel's own; other starts are staged in zout
        int e = mode == 1 ? STZ_OK : start_in_place(zout, zin, L.m, L.n, mode, zc, cm, cn, s);
        if (e) return e;
        BdLevelT<T> B = L;
        B.z = zout;
        const long nodes = (long)L.m * L.n;
        const int nt = nodes <= 1024 ? (int)((nodes + 63) / 64 * 64) : 1024;
        hipLaunchKernelGGL(gs_bottom_kernel<T>, dim3(1), dim3(nt), lds, s, B, sweeps, mode == 1 ? 1 : 0);
        STZ_HIP(hipGetLastError());
        return STZ_OK;
    }
    if (sweeps == 1) return tile_sweep(LevelRows<T>{L}, L.m, L.n, zin, zout, mode, zc, cm, cn, s);
    const int e = start_in_place(zout, zin, L.m, L.n, mode, zc, cm, cn, s);
    if (e) return e;
    BdLevelT<T> B = L;
    B.z = zout;
    for (int k = 0; k < sweeps; k++)
        for (int c = 0; c < 4; c++) {
            const long cnt = (long)((L.m + 1 - (c & 1)) >> 1) * ((L.n + 1 - (c >> 1)) >> 1);
            hipLaunchKernelGGL(gs_level_kernel<T>, blocks(cnt), dim3(kThreads), 0, s, B, c);
        }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// One V-cycle with zero start on the top level, in precision T (the source
// read in its own precision TB): z ~= A^-1 rhs.  zpre holds the top level's
// pre-smoothed iterate.  A coarse level keeps its pre-smoothed iterate in z
// and its post-smoothed one in r (the residual it held is restricted by
// then); the bottom level solves in z.
template <typename T, typename TB>
int BdSolver::vcycle_t(const TB *rhs, T *z, BdLevelT<T> *Lv, T *r0, T *zpre, hipStream_t s, const ApplyArgs *ap,
                       int *tiles) {
    const T *none = nullptr;
    int e;
    if (tiles) *tiles = 0;
    if (nl == 0) {
        if ((e = smooth_top(zpre, rhs, nu1, 1, none, none, 0, 0, s))) return e;
        return smooth_top(z, rhs, nu2, 0, (const T *)zpre, none, 0, 0, s);
    }
    // one pre-smoothing sweep per level: the tiled launch also forms the
    // residual and restricts it to the next level (mode 3)
    const bool fused = nu1 == 1;
    if (fused) {
        e = tile_sweep(TopRows<T, TB>::make(top, rhs), top_m, top_n, none, zpre, 3, none, Lv[0].m, Lv[0].n, s, Lv[0].b);
        if (e) return e;
    } else {
        if ((e = smooth_top(zpre, rhs, nu1, 1, none, none, 0, 0, s))) return e;
        hipLaunchKernelGGL((resid_top_kernel<T, TB>), blocks((long)top_m * top_n), dim3(kThreads), 0, s, top, zpre,
                           rhs, r0, (double *)nullptr);
        hipLaunchKernelGGL(restrict_kernel<T>, blocks((long)Lv[0].m * Lv[0].n), dim3(kThreads), 0, s, r0, top_m,
                           top_n, Lv[0].b, Lv[0].m, Lv[0].n);
    }
    for (int l = 0; l + 1 < nl; l++) {
        BdLevelT<T> &L = Lv[l], &C = Lv[l + 1];
        if (fused) {
            if ((e = tile_sweep(LevelRows<T>{L}, L.m, L.n, none, L.z, 3, none, C.m, C.n, s, C.b))) return e;
            continue;
        }
        if ((e = smooth_level(L, false, nu1, 1, none, L.z, none, 0, 0, s))) return e;
        hipLaunchKernelGGL(resid_level_kernel<T>, blocks((long)L.m * L.n), dim3(kThreads), 0, s, L);
        hipLaunchKernelGGL(restrict_kernel<T>, blocks((long)C.m * C.n), dim3(kThreads), 0, s, L.r, L.m, L.n, C.b,
                           C.m, C.n);
    }
    BdLevelT<T> &B = Lv[nl - 1];
    if ((e = smooth_level(B, true, 2 * bottom_half, 1, none, B.z, none, 0, 0, s))) return e;
    const T *res = B.z;  // the corrected iterate of the level below
    for (int l = nl - 1; l > 0; l--) {
        const BdLevelT<T> &C = Lv[l], &F = Lv[l - 1];
        if ((e = smooth_level(F, false, nu2, 2, (const T *)F.z, F.r, res, C.m, C.n, s))) return e;
        res = F.r;
    }
    // the operator application fused into the last launch: one sweep, a
    // partial sum per tile within the buffer
    {
        using D = Tile<T, 5>;
        const long nt = (long)((top_m + D::X - 1) / D::X) * ((top_n + D::Y - 1) / D::Y);
        if (ap && tiles && nu2 == 1 && nt <= kPartialHalf)
            return tile_sweep_apply(TopRows<T, TB>::make(top, rhs), top_m, top_n, (const T *)zpre, z, res, Lv[0].m,
                                    Lv[0].n, *ap, tiles, s);
    }
    return smooth_top(z, rhs, nu2, 2, (const T *)zpre, res, Lv[0].m, Lv[0].n, s);
}

// the FP64 preconditioner (the mixed form applies vcycle_t<float, double>
// inside solve and keeps its result in FP32)
int BdSolver::vcycle(const double *rhs, double *z, hipStream_t s) {
    return vcycle_t<double, double>(rhs, z, lv, vec[kTmp], vec[kPre], s);
}

int BdSolver::dot(const double *a, const double *b, int slot, int op, int host_slot, hipStream_t s) {
    const long n2 = 2L * top_m * top_n;
    const int nb = (int)std::min<long>((n2 + kThreads - 1) / kThreads, kMaxPartial);
    hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(kThreads), 0, s, a, b, n2, partial);
    hipLaunchKernelGGL(finish_dot_kernel, dim3(1), dim3(kThreads), 0, s, partial, nb, sc, slot, op,
                       host_slot >= 0 ? hscal + host_slot : (volatile double *)nullptr);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// Grid of a grid-stride kernel: the workgroups resident on the device at
// once (occupancy x CUs), at most kMaxPartial and at most one per kThreads
// elements.  Every workgroup does the same work, so a grid of several
// resident sets leaves its last set partly empty (4096 workgroups of
// bicg_x_dot_kernel at 7 per CU: 2.3 sets run as 3).
template <typename K> static int resident_grid(K kernel, long elems) {
    int dev = 0, occ = 0, cus = 0;
    long g = kMaxPartial;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kThreads, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && occ > 0 && cus > 0)
        g = std::min<long>((long)occ * cus, kMaxPartial);
    return (int)std::max<long>(1, std::min<long>(g, (elems + kThreads - 1) / kThreads));
}

// BiCGStab (van der Vorst 1992) with one V-cycle as right preconditioner.
// When the shadow residual becomes orthogonal to the residual (the box-edge
// rows are solved exactly by the first sweep, so a residual that starts
// there loses all overlap with it) the iteration restarts from the current
// residual.
int BdSolver::solve(double *x, const double *b, double acc, int max_iter, stz_solve_info &info, hipStream_t s) {
    const long mn = (long)top_m * top_n, n2 = 2 * mn;
    const dim3 g = blocks(n2), b1(kThreads);
    // grids of the kernels with fused products: at most kMaxPartial blocks;
    // the x/r update in one resident set (417 -> 397 us at 4096^2; the
    // operator applications, a node per thread for their stencils, lose
    // 4 % with the same grid)
    const int nd = (int)std::min<long>((mn + kThreads - 1) / kThreads, kMaxPartial),
              nx = mixed ? resident_grid(bicg_x_dot_kernel<float>, n2) : resident_grid(bicg_x_dot_kernel<double>, n2);
    const dim3 gd(nd), gx(nx);
    double *r = vec[kR], *rh = vec[kRhat], *p = vec[kP], *v = vec[kV], *sv = vec[kS], *t = vec[kT],
           *ph = vec[kPhat], *sh = vec[kShat];
    info = stz_solve_info{};
    info.tolerance = acc;
    int e;
    // restart: r = b - A x, r_hat = r, rho = (r,r)
    auto restart = [&](bool recompute, double &rr) -> int {
        if (recompute)
            hipLaunchKernelGGL((resid_top_kernel<double, double>), blocks(mn), b1, 0, s, top, x, b, r,
                               (double *)nullptr);
        STZ_HIP(hipMemcpyAsync(rh, r, n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
        int e2 = dot(r, r, 0, 0, 0, s);
        if (e2) return e2;
        STZ_HIP(hipStreamSynchronize(s));
        rr = *hscal;
        return STZ_OK;
    };
    double rr = 0, rhr = 0, rh2 = 0;
    if ((e = restart(true, rr))) return e;
    rh2 = rr;
    double res = rr / mn;
    info.residual = res;
    if (std::isnan(res)) {
        info.status = STZ_ERR_NAN;
        return STZ_OK;
    }
    if (res <= acc) {
        info.status = STZ_OK;
        return STZ_OK;
    }
    bool first = true;
    for (int total = 1; total <= max_iter; total++) {
        hipLaunchKernelGGL(bicg_p_kernel, g, b1, 0, s, p, r, v, sc, n2, first ? 1 : 0);
        first = false;
        // p^ = M p: FP64, or (mixed) an FP32 V-cycle reading p in FP64 and
        // leaving p^ in FP32 for the FP64 kernels that use it
        // v = A p^ with (r_hat, v), alpha
        int np = nd;  // partial sums of the product
        if (mixed) {
            const ApplyArgs ap{v, rh, partial, 0};
            int tiles = 0;
            if ((e = vcycle_t<float, double>(p, top32[0], lv32, top32[2], top32[3], s, &ap, &tiles))) return e;
            if (tiles) np = tiles;
            else hipLaunchKernelGGL(apply_top_dot_kernel<float>, gd, b1, 0, s, top, top32[0], v, rh, 0, partial);
        } else {
            const ApplyArgs ap{v, rh, partial, 0};
            int tiles = 0;
            if ((e = vcycle_t<double, double>(p, ph, lv, vec[kTmp], vec[kPre], s, &ap, &tiles))) return e;
            if (tiles) np = tiles;
            else hipLaunchKernelGGL(apply_top_dot_kernel<double>, gd, b1, 0, s, top, ph, v, rh, 0, partial);
        }
        hipLaunchKernelGGL(finish_dots_kernel, dim3(1), b1, 0, s, partial, np, sc, 4, -1, 2, nullptr, nullptr);
        hipLaunchKernelGGL(bicg_s_kernel, g, b1, 0, s, sv, r, v, sc, n2);
        // t = A s^ with (t, s) and (t, t), omega
        np = nd;
        if (mixed) {
            const ApplyArgs ap{t, sv, partial, 1};
            int tiles = 0;
            if ((e = vcycle_t<float, double>(sv, top32[1], lv32, top32[2], top32[3], s, &ap, &tiles))) return e;
            if (tiles) np = tiles;
            else hipLaunchKernelGGL(apply_top_dot_kernel<float>, gd, b1, 0, s, top, top32[1], t, sv, 1, partial);
        } else {
            const ApplyArgs ap{t, sv, partial, 1};
            int tiles = 0;
            if ((e = vcycle_t<double, double>(sv, sh, lv, vec[kTmp], vec[kPre], s, &ap, &tiles))) return e;
            if (tiles) np = tiles;
            else hipLaunchKernelGGL(apply_top_dot_kernel<double>, gd, b1, 0, s, top, sh, t, sv, 1, partial);
        }
        hipLaunchKernelGGL(finish_dots_kernel, dim3(1), b1, 0, s, partial, np, sc, 5, 6, 3, nullptr, nullptr);
        // x, r with (r, r) -> host[0] and the next rho = (r_hat, r) -> sc[0], host[1]
        if (mixed)
            hipLaunchKernelGGL(bicg_x_dot_kernel<float>, gx, b1, 0, s, x, r, top32[0], top32[1], sv, t, rh, sc, n2,
                               partial);
        else
            hipLaunchKernelGGL(bicg_x_dot_kernel<double>, gx, b1, 0, s, x, r, ph, sh, sv, t, rh, sc, n2, partial);
        hipLaunchKernelGGL(finish_dots_kernel, dim3(1), b1, 0, s, partial, nx, sc, 7, 0, 0,
                           (volatile double *)hscal, (volatile double *)hscal + 1);
        STZ_HIP(hipGetLastError());
        info.cycles = info.total_cycles = 2 * total;
        STZ_HIP(hipStreamSynchronize(s));
        rr = hscal[0];
        rhr = hscal[1];
        res = rr / mn;
        info.residual = res;
        if (std::isnan(res)) {
            info.status = STZ_ERR_NAN;
            return STZ_OK;
        }
        bool again = false;
        if (res <= acc) {
            // confirm with the true residual
            if ((e = restart(true, rr))) return e;
            res = rr / mn;
            info.residual = res;
            info.refinements++;
            if (res <= acc) {
                info.status = STZ_OK;
                return STZ_OK;
            }
            again = true;
        } else if (!(std::fabs(rhr) > 1e-12 * std::sqrt(rh2 * rr))) {
            if ((e = restart(false, rr))) return e;
            again = true;
        }
        if (again) {
            rh2 = rr;
            first = true;
        }
    }
    info.status = STZ_ERR_NOT_CONVERGED;
    return STZ_OK;
}

int BdSolver::l2_error(const double *x, const double *b, double *out, hipStream_t s) {
    const long mn = (long)top_m * top_n;
    double *r = vec[kTmp];
    hipLaunchKernelGGL((resid_top_kernel<double, double>), blocks(mn), dim3(kThreads), 0, s, top, x, b, r,
                       (double *)nullptr);
    int e = dot(r, r, 7, 0, 0, s);
    if (e) return e;
    STZ_HIP(hipStreamSynchronize(s));
    *out = *hscal / mn;
    return STZ_OK;
}

}  // namespace stz
