// mg.hip -- gfx950 geometric multigrid for the quasi-static projection
// (the reference's tgmg<qs_multi,vec,mat>, tgmg/tgmg.in.cc), templated on the
// working precision T (double, or float for the mixed-precision form).
//
// Same hierarchy, bilinear prolongation P and restriction R = P^T, and
// Galerkin coarse operators as the reference (tgmg.in.cc:10-30, 552-779).
// The smoother is a four-colour point Gauss-Seidel: colour (i%2, j%2), so no
// two nodes of one colour share a 9-point stencil and every colour pass is
// fully parallel and deterministic.  (The reference's zebra-line ordering,
// tgmg.in.cc:127-165, is sequential along a row; per V(1,1) cycle both
// converge at the same rate, ~0.61 on this operator.)  On an odd periodic
// width the seam column gets a colour of its own.
//
// Kernels:
//   colour pass      one thread per node of one colour
//   restrict         one thread per coarse node; recomputes the <=9 fine
//                    residuals it needs (no residual array in HBM), writes the
//                    coarse source and zeroes the coarse solution
//   prolong          one thread per fine node, z += P zc
//   coarse solve     ONE workgroup runs the whole sub-V-cycle on the small
//                    levels with z and b resident in LDS, and solves the
//                    coarsest level exactly with a precomputed dense inverse
//   residual norm    wavefront-shuffle reduction to per-block partials, then
//                    one block sums the partials in a fixed order
#include <cstring>

#include "batch.h"
#include "mg_node.h"

namespace stz {

// ---------------------------------------------------------------------------
// grid kernels
// ---------------------------------------------------------------------------
template <typename T> __global__ void __launch_bounds__(256) colour_kernel(LevelArgs<T> L, int cx, int cy) {
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 * (blockIdx.y * blockDim.y + threadIdx.y) + cy;
    int i;
    if (cx == 2) {
        if (tx != 0) return;
        i = L.m - 1;
    } else {
        i = 2 * tx + cx;
        if (L.xodd && i == L.m - 1) return;
    }
    if (i >= L.m || j >= L.n) return;
    gs_node(L, i, j);
}

template <typename T> __global__ void __launch_bounds__(256) restrict_kernel(LevelArgs<T> F, LevelArgs<T> C) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= C.m || J >= C.n) return;
    restrict_node(F, C, I, J);
}

template <typename T> __global__ void __launch_bounds__(256) prolong_kernel(LevelArgs<T> C, LevelArgs<T> F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= F.m || j >= F.n) return;
    prolong_node(C, F, i, j);
}

// ---- reductions -------------------------------------------------------------
__device__ __forceinline__ static double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block of 256 threads -> one partial, stored at partial[slot]
__device__ __forceinline__ static void block_partial(double v, double *partial, int slot) {
    __shared__ double ws[4];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[slot] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}
__device__ __forceinline__ static void block_partial(double v, double *partial) {
    block_partial(v, partial, blockIdx.x);
}

template <typename T>
__device__ __forceinline__ static void resid_norm_body(const LevelArgs<T> &L, double *partial, int blk) {
    const int idx = blk * 256 + threadIdx.x;
    double v = 0;
    if (idx < L.m * L.n) {
        const int j = idx / L.m, i = idx - j * L.m;
        T rx, ry;
        resid_node(L, i, j, rx, ry);
        v = (double)rx * (double)rx + (double)ry * (double)ry;
    }
    block_partial(v, partial, blk);
}

template <typename T> __global__ void __launch_bounds__(256) resid_norm_kernel(LevelArgs<T> L, double *partial) {
    resid_norm_body(L, partial, blockIdx.x);
}

// fixed-order sum of the partials by one block of 256 threads
__device__ __forceinline__ static double sum_partials_body(const double *partial, int np) {
    double v = 0;
    for (int k = threadIdx.x; k < np; k += 256) v += partial[k];
    __shared__ double ws[4];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    return (ws[0] + ws[1]) + (ws[2] + ws[3]);  // valid in thread 0
}

__global__ void __launch_bounds__(256) sum_partials_kernel(const double *partial, int np, double *out) {
    const double v = sum_partials_body(partial, np);
    if (threadIdx.x == 0) *out = v;
}

// mixed precision: FP64 residual of the top level -> FP32 source of the
// correction equation, zero FP32 correction, norm partials
__global__ void __launch_bounds__(256) resid64_to32_kernel(LevelArgs<double> L, float *b32, float *z32,
                                                           double *partial) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    double v = 0;
    if (idx < L.m * L.n) {
        const int j = idx / L.m, i = idx - j * L.m;
        double rx, ry;
        resid_node(L, i, j, rx, ry);
        v = rx * rx + ry * ry;
        b32[2 * idx] = (float)rx;
        b32[2 * idx + 1] = (float)ry;
        z32[2 * idx] = 0.f;
        z32[2 * idx + 1] = 0.f;
    }
    block_partial(v, partial);
}

template <typename T> __global__ void __launch_bounds__(256) fill_zero_kernel(T *p, size_t count) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx < count) p[idx] = 0;
}

__global__ void __launch_bounds__(256) correct64_kernel(double *z, const float *e, int count) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < count) z[idx] += (double)e[idx];
}

// ---------------------------------------------------------------------------
// workgroup coarse solver: levels lc..L-1 entirely in LDS (z, b, stencils, a
// residual scratch and the coarsest level's dense inverse), kCoarseThreads
// threads, one barrier per pass.  z and b of a level are stored colour-split:
// node (i,j) in parity array (i&1) + 2(j&1) at ((j>>1), (i>>1)), so the lanes
// of a colour pass, and the neighbours they read, are at consecutive
// addresses.  Each node update issues all of its LDS loads before any
// arithmetic (no branch between them: rows beyond a wall are read from a
// clamped row and zeroed), which keeps one pass to one LDS round trip.
// ---------------------------------------------------------------------------
constexpr int kCoarseThreads = 512;

// stencil values stored for a level: one 40-value stencil per row, or per node
template <typename T> __host__ __device__ __forceinline__ static int stencil_count(const LevelArgs<T> &L) {
    return (L.cs ? L.m * L.n : L.n) * 40;
}

template <typename T> struct WGLevel {
    int m, n, rs, cs, xodd, hm, hn, nb;
    int zo, bo, so;
};

template <typename T> __device__ __forceinline__ static WGLevel<T> wg_level(const CoarseArgs<T> &A, int l) {
    WGLevel<T> W;
    W.m = A.lv[l].m;
    W.n = A.lv[l].n;
    W.rs = A.lv[l].rs;
    W.cs = A.lv[l].cs;
    W.xodd = A.lv[l].xodd;
    W.nb = A.lv[l].cs ? W.n : A.lv[l].nb;
    W.hm = (W.m + 1) >> 1;
    W.hn = (W.n + 1) >> 1;
    W.zo = A.zoff[l];
    W.bo = A.boff[l];
    W.so = A.soff[l];
    return W;
}

// element offset of node (i,j)'s x component in a colour-split level array
template <typename T> __device__ __forceinline__ static int cidx(const WGLevel<T> &W, int i, int j) {
    return 2 * ((((i & 1) + 2 * (j & 1)) * W.hn + (j >> 1)) * W.hm + (i >> 1));
}

template <typename T> __device__ __forceinline__ static vec2s<T> lds2(const T *sm, int k) {
    return *reinterpret_cast<const vec2s<T> *>(sm + k);
}

// off-diagonal sum at node (i,j) of an LDS-resident level
template <typename T, typename SP>
__device__ __forceinline__ static void wg_offdiag(const T *sm, const WGLevel<T> &W, int i, int j, const SP &s, T &ax,
                                                  T &ay) {
    const int m = W.m;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const int jd = j > 0 ? j - 1 : j, ju = j < W.n - 1 ? j + 1 : j;
    const T *z = sm + W.zo;
    vec2s<T> dl = lds2(z, cidx(W, il, jd)), dc = lds2(z, cidx(W, i, jd)), dr = lds2(z, cidx(W, ir, jd));
    const vec2s<T> cl = lds2(z, cidx(W, il, j)), cr = lds2(z, cidx(W, ir, j));
    vec2s<T> ul = lds2(z, cidx(W, il, ju)), uc = lds2(z, cidx(W, i, ju)), ur = lds2(z, cidx(W, ir, ju));
    const vec2s<T> zero = {0, 0};
    if (j == 0) dl = dc = dr = zero;
    if (j == W.n - 1) ul = uc = ur = zero;
    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
}

// residuals of every node of level W into r (natural order): rows nb..n-nb-1
// with the interior stencil si held in registers, the wall rows with their
// own stencils from LDS (two loops, so that the register path never goes
// through a pointer)
template <typename T, typename SP>
__device__ __forceinline__ static void wg_resid_node(T *sm, const WGLevel<T> &W, const SP &s, int ro, int i, int j) {
    T ax, ay, rx, ry;
    const int k = cidx(W, i, j);
    wg_offdiag(sm, W, i, j, s, ax, ay);
    resid_value<T>(s, sm[W.bo + k], sm[W.bo + k + 1], ax, ay, lds2(sm + W.zo, k), rx, ry);
    sm[ro + 2 * (i + W.m * j)] = rx;
    sm[ro + 2 * (i + W.m * j) + 1] = ry;
}

template <typename T>
__device__ __forceinline__ static void wg_resid_all(T *sm, const WGLevel<T> &W, const T (&si)[40], int ro) {
    const int lo = W.nb < W.n ? W.nb : W.n, hi = W.n - W.nb > lo ? W.n - W.nb : lo, m = W.m;
    for (int t = threadIdx.x; t < (hi - lo) * m; t += kCoarseThreads) {
        const int r = t / m, i = t - r * m;
        wg_resid_node(sm, W, si, ro, i, lo + r);
    }
    for (int t = threadIdx.x; t < (lo + W.n - hi) * m; t += kCoarseThreads) {
        const int r = t / m, i = t - r * m, j = r < lo ? r : hi + (r - lo);
        const T *s = sm + W.so + (j * W.rs + i * W.cs) * 40;
        wg_resid_node(sm, W, s, ro, i, j);
    }
}

template <typename T>
__device__ __forceinline__ static void wg_smooth(T *sm, const WGLevel<T> &W, const T (&si)[40], int sweeps) {
    const int hm = W.hm, hn = W.hn;
    for (int sw = 0; sw < sweeps; sw++)
        for (int cy = 0; cy < 2; cy++)
            for (int cx = 0; cx < 2 + W.xodd; cx++) {
                const int cnt = cx == 2 ? hn : hm * hn;
                for (int t = threadIdx.x; t < cnt; t += kCoarseThreads) {
                    int i, j;
                    if (cx == 2) { i = W.m - 1; j = 2 * t + cy; }
                    else { const int ty = t / hm; i = 2 * (t - ty * hm) + cx; j = 2 * ty + cy; }
                    if (i >= W.m || j >= W.n || (cx < 2 && W.xodd && i == W.m - 1)) continue;
                    T ax, ay;
                    const int k = cidx(W, i, j);
                    vec2s<T> o;
                    if (j < W.nb || j >= W.n - W.nb) {
                        const T *s = sm + W.so + (j * W.rs + i * W.cs) * 40;
                        wg_offdiag(sm, W, i, j, s, ax, ay);
                        o = gs_value<T>(s, sm[W.bo + k], sm[W.bo + k + 1], ax, ay);
                    } else {
                        wg_offdiag(sm, W, i, j, si, ax, ay);
                        o = gs_value<T>(si, sm[W.bo + k], sm[W.bo + k + 1], ax, ay);
                    }
                    *reinterpret_cast<vec2s<T> *>(sm + W.zo + k) = o;
                }
                __syncthreads();
            }
}

// Fast path of wg_smooth for even m when every colour has at most one node
// per thread: the LDS offsets of a thread's four nodes (one per colour), of
// their neighbours and stencils, and their sources b, do not change between
// sweeps, so they are set up once per visit of the level and each colour
// pass is loads, the FMA tree and one store.
template <typename T>
__device__ __forceinline__ static void wg_smooth_fast(T *sm, const WGLevel<T> &W, const T (&si)[40], int sweeps) {
    const int t = threadIdx.x, hm = W.hm;
    const int ty = t / hm, tx = t - ty * hm;
    int kz[4], nbo[4][8], so[4];
    unsigned fl[4];  // 1 active, 2 wall-row stencil, 4 no row below, 8 no row above
    T bx[4], by[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int cx = c & 1, cy = c >> 1;
        const int i = 2 * tx + cx, j = 2 * ty + cy;
        const bool act = t < hm * W.hn && j < W.n;
        const int ii = act ? i : 0, jj = act ? j : 0;
        const int il = ii == 0 ? W.m - 1 : ii - 1, ir = ii == W.m - 1 ? 0 : ii + 1;
        const int jd = jj > 0 ? jj - 1 : jj, ju = jj < W.n - 1 ? jj + 1 : jj;
        kz[c] = cidx(W, ii, jj);
        nbo[c][0] = cidx(W, il, jd);
        nbo[c][1] = cidx(W, ii, jd);
        nbo[c][2] = cidx(W, ir, jd);
        nbo[c][3] = cidx(W, il, jj);
        nbo[c][4] = cidx(W, ir, jj);
        nbo[c][5] = cidx(W, il, ju);
        nbo[c][6] = cidx(W, ii, ju);
        nbo[c][7] = cidx(W, ir, ju);
        so[c] = W.so + (jj * W.rs + ii * W.cs) * 40;
        fl[c] = (act ? 1u : 0u) | ((jj < W.nb || jj >= W.n - W.nb) ? 2u : 0u) | (jj == 0 ? 4u : 0u) |
                (jj == W.n - 1 ? 8u : 0u);
        bx[c] = sm[W.bo + kz[c]];
        by[c] = sm[W.bo + kz[c] + 1];
    }
    const T *z = sm + W.zo;
    const vec2s<T> zero = {0, 0};
    for (int sw = 0; sw < sweeps; sw++) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (fl[c] & 1u) {
                vec2s<T> dl = lds2(z, nbo[c][0]), dc = lds2(z, nbo[c][1]), dr = lds2(z, nbo[c][2]);
                const vec2s<T> cl = lds2(z, nbo[c][3]), cr = lds2(z, nbo[c][4]);
                vec2s<T> ul = lds2(z, nbo[c][5]), uc = lds2(z, nbo[c][6]), ur = lds2(z, nbo[c][7]);
                if (fl[c] & 4u) dl = dc = dr = zero;
                if (fl[c] & 8u) ul = uc = ur = zero;
                T ax, ay;
                vec2s<T> o;
                if (fl[c] & 2u) {
                    const T *s = sm + so[c];
                    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    o = gs_value<T>(s, bx[c], by[c], ax, ay);
                } else {
                    offdiag9<T>(si, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    o = gs_value<T>(si, bx[c], by[c], ax, ay);
                }
                *reinterpret_cast<vec2s<T> *>(sm + W.zo + kz[c]) = o;
            }
            __syncthreads();
        }
    }
}

template <typename T>
__device__ __forceinline__ static void wg_smooth_level(T *sm, const WGLevel<T> &W, const T (&si)[40], int sweeps) {
    if (!W.xodd && W.hm * W.hn <= kCoarseThreads) wg_smooth_fast(sm, W, si, sweeps);
    else wg_smooth(sm, W, si, sweeps);
}

// ---------------------------------------------------------------------------
// Block path of the workgroup solver: m even, one stencil per row (cs == 0)
// and at most kCoarseThreads 2x2 blocks.  Thread t owns block (tx, ty), the
// nodes (2tx..2tx+1, 2ty..2ty+1), one of each colour.  For the whole visit of
// a level it keeps the stencils of its two rows and its four sources in
// registers, so a colour pass is eight LDS loads, the FMA tree and a store.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ static bool block_ok(const WGLevel<T> &W) {
    return !W.xodd && W.cs == 0 && W.hm * W.hn <= kCoarseThreads;
}

// element offset of block (row, col) in colour array c
template <typename T> __device__ __forceinline__ static int boff(const WGLevel<T> &W, int c, int row, int col) {
    return 2 * ((c * W.hn + row) * W.hm + col);
}

struct BlockPos {
    bool act;
    int tx, ty, txl, txr, tyd, tyu;
};

template <typename T> __device__ __forceinline__ static BlockPos block_pos(const WGLevel<T> &W) {
    BlockPos b;
    const int t = threadIdx.x, hm = W.hm, hn = W.hn;
    b.act = t < hm * hn;
    b.ty = b.act ? t / hm : 0;
    b.tx = b.act ? t - b.ty * hm : 0;
    b.txl = b.tx == 0 ? hm - 1 : b.tx - 1;
    b.txr = b.tx == hm - 1 ? 0 : b.tx + 1;
    b.tyd = b.ty > 0 ? b.ty - 1 : 0;
    b.tyu = b.ty < hn - 1 ? b.ty + 1 : hn - 1;
    return b;
}

// the eight neighbours of this thread's node of colour (CX, CY)
template <typename T, int CX, int CY>
__device__ __forceinline__ static void block_neighbours(const T *z, const WGLevel<T> &W, const BlockPos &b, vec2s<T> &dl,
                                                        vec2s<T> &dc, vec2s<T> &dr, vec2s<T> &cl, vec2s<T> &cr,
                                                        vec2s<T> &ul, vec2s<T> &uc, vec2s<T> &ur) {
    constexpr int cX = (1 - CX) + 2 * CY, cY = CX + 2 * (1 - CY), cD = (1 - CX) + 2 * (1 - CY);
    const int colL = CX ? b.tx : b.txl, colR = CX ? b.txr : b.tx;
    const int rowD = CY ? b.ty : b.tyd, rowU = CY ? b.tyu : b.ty;
    dl = lds2(z, boff(W, cD, rowD, colL));
    dc = lds2(z, boff(W, cY, rowD, b.tx));
    dr = lds2(z, boff(W, cD, rowD, colR));
    cl = lds2(z, boff(W, cX, b.ty, colL));
    cr = lds2(z, boff(W, cX, b.ty, colR));
    ul = lds2(z, boff(W, cD, rowU, colL));
    uc = lds2(z, boff(W, cY, rowU, b.tx));
    ur = lds2(z, boff(W, cD, rowU, colR));
    const int j = 2 * b.ty + CY;
    const vec2s<T> zero = {0, 0};
    if (j == 0) dl = dc = dr = zero;
    if (j == W.n - 1) ul = uc = ur = zero;
}

template <typename T, int CX, int CY>
__device__ __forceinline__ static void block_gs(T *sm, const WGLevel<T> &W, const BlockPos &b, const T (&s)[40], T bx,
                                                T by) {
    if (b.act && 2 * b.ty + CY < W.n) {
        vec2s<T> dl, dc, dr, cl, cr, ul, uc, ur;
        block_neighbours<T, CX, CY>(sm + W.zo, W, b, dl, dc, dr, cl, cr, ul, uc, ur);
        T ax, ay;
        offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
        *reinterpret_cast<vec2s<T> *>(sm + W.zo + boff(W, CX + 2 * CY, b.ty, b.tx)) = gs_value<T>(s, bx, by, ax, ay);
    }
    __syncthreads();
}

// residuals of this thread's four nodes from one load of their 4x4
// neighbourhood (rows 2ty-1..2ty+2, columns 2tx-1..2tx+2; zero beyond a wall)
template <typename T>
__device__ __forceinline__ static void block_resid4(T *sm, const WGLevel<T> &W, const BlockPos &b, const T (&s0)[40],
                                                    const T (&s1)[40], const T (&bx)[4], const T (&by)[4], int ro) {
    if (!b.act) return;
    const T *z = sm + W.zo;
    const int rows[4] = {b.tyd, b.ty, b.ty, b.tyu}, cols[4] = {b.txl, b.tx, b.tx, b.txr};
    vec2s<T> Z[4][4];
    const vec2s<T> zero = {0, 0};
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int j = 2 * b.ty - 1 + a, cy = (a + 1) & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int cx = (c + 1) & 1;
            Z[a][c] = lds2(z, boff(W, cx + 2 * cy, rows[a], cols[c]));
            if (j < 0 || j >= W.n) Z[a][c] = zero;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int cx = c & 1, cy = c >> 1, i = 2 * b.tx + cx, j = 2 * b.ty + cy;
        if (j >= W.n) continue;
        T ax, ay, rx, ry;
        if (cy == 0) {
            offdiag9<T>(s0, Z[0][cx], Z[0][cx + 1], Z[0][cx + 2], Z[1][cx], Z[1][cx + 2], Z[2][cx], Z[2][cx + 1],
                        Z[2][cx + 2], ax, ay);
            resid_value<T>(s0, bx[c], by[c], ax, ay, Z[1][cx + 1], rx, ry);
        } else {
            offdiag9<T>(s1, Z[1][cx], Z[1][cx + 1], Z[1][cx + 2], Z[2][cx], Z[2][cx + 2], Z[3][cx], Z[3][cx + 1],
                        Z[3][cx + 2], ax, ay);
            resid_value<T>(s1, bx[c], by[c], ax, ay, Z[2][cx + 1], rx, ry);
        }
        sm[ro + 2 * (i + W.m * j)] = rx;
        sm[ro + 2 * (i + W.m * j) + 1] = ry;
    }
}

// z(i,j) += (P zc)(i,j) for one fine node
template <typename T>
__device__ __forceinline__ static void wg_prolong_node(T *sm, const WGLevel<T> &F, const WGLevel<T> &C, int i, int j) {
    const int sc = C.m;
    int cx0, cx1, cy0, cy1;
    T wx0, wx1, wy0, wy1;
    if ((i & 1) == 0) { cx0 = i >> 1; wx0 = 1; cx1 = cx0; wx1 = 0; }
    else { cx0 = i >> 1; cx1 = cx0 + 1 == sc ? 0 : cx0 + 1; wx0 = wx1 = (T)0.5; }
    if ((F.n & 1) == 0 && j == F.n - 1) { cy0 = C.n - 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
    else if ((j & 1) == 0) { cy0 = j >> 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
    else { cy0 = j >> 1; cy1 = cy0 + 1; wy0 = wy1 = (T)0.5; }
    const T *zc = sm + C.zo;
    const vec2s<T> a00 = lds2(zc, cidx(C, cx0, cy0)), a10 = lds2(zc, cidx(C, cx1, cy0)),
                   a01 = lds2(zc, cidx(C, cx0, cy1)), a11 = lds2(zc, cidx(C, cx1, cy1));
    const int k = cidx(F, i, j);
    sm[F.zo + k] += wy0 * (wx0 * a00.x + wx1 * a10.x) + wy1 * (wx0 * a01.x + wx1 * a11.x);
    sm[F.zo + k + 1] += wy0 * (wx0 * a00.y + wx1 * a10.y) + wy1 * (wx0 * a01.y + wx1 * a11.y);
}

// One visit of a block-path level: [z += P zc], `sweeps` four-colour sweeps,
// [residuals into r].  Ends with a barrier unless residuals were written
// (the caller's next barrier covers them).
template <typename T>
__device__ __forceinline__ static void wg_blocks(T *sm, const WGLevel<T> &W, int sweeps, const WGLevel<T> *Cp, int ro) {
    const BlockPos b = block_pos(W);
    const int j0 = 2 * b.ty, j1 = j0 + 1 < W.n ? j0 + 1 : j0;
    // a wall row reads its own stencil; every other row reads the interior
    // one (row n/2), so most lanes of a wave load the same LDS address
    const int r0 = (j0 < W.nb || j0 >= W.n - W.nb) ? j0 : W.n / 2;
    const int r1 = (j1 < W.nb || j1 >= W.n - W.nb) ? j1 : W.n / 2;
    T s0[40], s1[40];
#pragma unroll
    for (int k = 0; k < 40; k++) {
        s0[k] = sm[W.so + r0 * 40 + k];
        s1[k] = sm[W.so + r1 * 40 + k];
    }
    if (Cp) {
        if (b.act)
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int i = 2 * b.tx + (c & 1), j = 2 * b.ty + (c >> 1);
                if (j < W.n) wg_prolong_node(sm, W, *Cp, i, j);
            }
        __syncthreads();
    }
    T bx[4], by[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int k = boff(W, c, b.ty, b.tx);
        bx[c] = sm[W.bo + k];
        by[c] = sm[W.bo + k + 1];
    }
    for (int sw = 0; sw < sweeps; sw++) {
        block_gs<T, 0, 0>(sm, W, b, s0, bx[0], by[0]);
        block_gs<T, 1, 0>(sm, W, b, s0, bx[1], by[1]);
        block_gs<T, 0, 1>(sm, W, b, s1, bx[2], by[2]);
        block_gs<T, 1, 1>(sm, W, b, s1, bx[3], by[3]);
    }
    if (ro >= 0) block_resid4(sm, W, b, s0, s1, bx, by, ro);
}

// copy cnt values from global memory into LDS through dst(t, value), eight
// loads in flight per thread before the first store
template <typename T, typename D> __device__ __forceinline__ static void g2lds(const T *src, int cnt, D dst) {
    constexpr int K = 8;
    for (int base = threadIdx.x; base < cnt; base += K * kCoarseThreads) {
        T v[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int t = base + k * kCoarseThreads;
            v[k] = t < cnt ? src[t] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int t = base + k * kCoarseThreads;
            if (t < cnt) dst(t, v[k]);
        }
    }
}

// the level's interior stencil into registers (uniform across the workgroup)
template <typename T> __device__ __forceinline__ static void load_si(const LevelArgs<T> &L, T *si) {
#pragma unroll
    for (int k = 0; k < 40; k++) si[k] = L.Si[k];
}

template <typename T>
__device__ __forceinline__ static void coarse_solve_body(const CoarseArgs<T> &A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    const int tid = threadIdx.x, NT = kCoarseThreads;
    const int ro = A.roff;
    // prologue: the constant part (stencils, transposed bottom inverse) is
    // one contiguous copy of A.pack; the source of level lc goes colour-split
    // with z = 0.  Every load of a batch is issued before its stores.
    {
        const WGLevel<T> W = wg_level(A, A.lc);
        using V2 = vec2s<T>;
        const int np = A.pack_elems / 2, nb = W.m * W.n;  // pack_elems is a multiple of 4
        const V2 *pk = reinterpret_cast<const V2 *>(A.pack);
        const V2 *gb = reinterpret_cast<const V2 *>(A.lv[A.lc].b);
        constexpr int K = 8;
        for (int base = tid; base < np + nb; base += K * NT) {
            V2 v[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int e = base + k * NT;
                v[k] = e < np ? pk[e] : (e < np + nb ? gb[e - np] : V2{0, 0});
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int e = base + k * NT;
                if (e < np) {
                    reinterpret_cast<V2 *>(sm)[e] = v[k];
                } else if (e < np + nb) {
                    const int node = e - np, j = node / W.m, i = node - j * W.m, c = cidx(W, i, j);
                    *reinterpret_cast<V2 *>(sm + W.bo + c) = v[k];
                    *reinterpret_cast<V2 *>(sm + W.zo + c) = V2{0, 0};
                }
            }
        }
    }
    __syncthreads();
    // down: smooth, fine residuals into r (one pass), P^T-weighted sums
    for (int l = A.lc; l < A.L - 1; l++) {
        const WGLevel<T> F = wg_level(A, l), C = wg_level(A, l + 1);
        {
            if (block_ok(F)) {
                wg_blocks(sm, F, A.nu1, (const WGLevel<T> *)nullptr, ro);
            } else {
                T si[40];
                load_si(A.lv[l], si);
                wg_smooth_level(sm, F, si, A.nu1);
                wg_resid_all(sm, F, si, ro);
            }
        }
        __syncthreads();
        for (int t = tid; t < C.m * C.n; t += NT) {
            const int J = t / C.m, I = t - J * C.m;
            const int fx0 = 2 * I, fy0 = ((F.n & 1) == 0 && J == C.n - 1) ? F.n - 1 : 2 * J;
            T bx = 0, by = 0;
            if (!F.xodd) {
                // even fine width: x weights 1/2, 1, 1/2 at 2I-1 (periodic), 2I, 2I+1
                const int fl = fx0 == 0 ? F.m - 1 : fx0 - 1;
#pragma unroll
                for (int b = -1; b <= 1; b++) {
                    const int fy = fy0 + b;
                    const T wy = (T)pw_bounded(fy, F.n, C.n, J);
                    const int row = ro + 2 * F.m * (fy < 0 ? 0 : (fy >= F.n ? F.n - 1 : fy));
                    const vec2s<T> rl = lds2(sm, row + 2 * fl), rc = lds2(sm, row + 2 * fx0),
                                   rr = lds2(sm, row + 2 * fx0 + 2);
                    bx = fmav(wy, fmav((T)0.5, rl.x + rr.x, rc.x), bx);
                    by = fmav(wy, fmav((T)0.5, rl.y + rr.y, rc.y), by);
                }
            } else {
#pragma unroll
                for (int b = -1; b <= 1; b++) {
                    const int fy = fy0 + b;
                    const float wy = pw_bounded(fy, F.n, C.n, J);
                    const int fyc = fy < 0 ? 0 : (fy >= F.n ? F.n - 1 : fy);
#pragma unroll
                    for (int a = -1; a <= 1; a++) {
                        const int fx = fx0 + a;
                        const float wx = pw_periodic(fx, F.m, C.m, I);
                        const int fi = fx < 0 ? fx + F.m : (fx >= F.m ? fx - F.m : fx);
                        const vec2s<T> r = lds2(sm + ro, 2 * (fi + F.m * fyc));
                        const T w = (T)(wx * wy);
                        bx = fmav(w, r.x, bx);
                        by = fmav(w, r.y, by);
                    }
                }
            }
            const int k = cidx(C, I, J);
            sm[C.bo + k] = bx;
            sm[C.bo + k + 1] = by;
            sm[C.zo + k] = 0;
            sm[C.zo + k + 1] = 0;
        }
        __syncthreads();
    }
    // bottom: exact solve with the dense inverse (natural unknown order)
    {
        const WGLevel<T> B = wg_level(A, A.L - 1);
        const int N = 2 * B.m * B.n;
        for (int t = tid; t < B.m * B.n; t += NT) {
            const int j = t / B.m, i = t - j * B.m, k = cidx(B, i, j);
            sm[ro + 2 * t] = sm[B.bo + k];
            sm[ro + 2 * t + 1] = sm[B.bo + k + 1];
        }
        __syncthreads();
        // in FP64 in both forms: the inverse is held in doubles
        const double *bi = reinterpret_cast<const double *>(sm + A.binv_off);
        for (int r = tid; r < N; r += NT) {
            double acc = 0;
            for (int c = 0; c < N; c++) acc = fma(bi[c * N + r], (double)sm[ro + c], acc);
            const int node = r >> 1, j = node / B.m, i = node - j * B.m;
            sm[B.zo + cidx(B, i, j) + (r & 1)] = (T)acc;
        }
        __syncthreads();
    }
    // up: z += P zc, then smooth
    for (int l = A.L - 2; l >= A.lc; l--) {
        const WGLevel<T> F = wg_level(A, l), C = wg_level(A, l + 1);
        if (block_ok(F)) {
            wg_blocks(sm, F, A.nu2, &C, -1);
        } else {
            T si[40];
            load_si(A.lv[l], si);
            for (int t = tid; t < F.m * F.n; t += NT) {
                const int j = t / F.m, i = t - j * F.m;
                wg_prolong_node(sm, F, C, i, j);
            }
            __syncthreads();
            wg_smooth_level(sm, F, si, A.nu2);
        }
    }
    // store the level-lc solution in natural order
    {
        const WGLevel<T> W = wg_level(A, A.lc);
        T *gz = A.lv[A.lc].z;
        for (int t = tid; t < W.m * W.n; t += NT) {
            const int j = t / W.m, i = t - j * W.m, k = cidx(W, i, j);
            gz[2 * t] = sm[W.zo + k];
            gz[2 * t + 1] = sm[W.zo + k + 1];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) coarse_solve_kernel(CoarseArgs<T> A) {
    coarse_solve_body(A);
}

// one workgroup per member of an ensemble (batch.h)
template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) coarse_solve_batch_kernel(const CoarseEntry<T> *E,
                                                                            const EnsState *St) {
    const CoarseEntry<T> &x = E[blockIdx.x];
    if (St && !St[x.member].active) return;
    coarse_solve_body(x.A);
}

template <typename T> int coarse_layout(CoarseArgs<T> &A) {
    int off = 0;
    for (int l = A.lc; l < A.L; l++) {
        A.soff[l] = off;
        off += stencil_count(A.lv[l]);
        off = (off + 3) & ~3;
    }
    // the coarsest level's dense inverse in doubles (two elements per value
    // when T is float), 8-byte aligned
    const int N = 2 * A.lv[A.L - 1].m * A.lv[A.L - 1].n;
    constexpr int per = (int)(sizeof(double) / sizeof(T));
    A.binv_off = off;
    off += per * N * N;
    off = (off + 3) & ~3;
    A.pack_elems = off;
    for (int l = A.lc; l < A.L; l++) {
        // colour-split storage: four parity arrays of ceil(m/2) x ceil(n/2)
        const int cnt = 8 * ((A.lv[l].m + 1) >> 1) * ((A.lv[l].n + 1) >> 1);
        A.zoff[l] = off;
        off += cnt;
        A.boff[l] = off;
        off += cnt;
    }
    // residual scratch: natural order, the largest level (also holds the
    // coarsest source for the dense solve)
    A.roff = off;
    off += 2 * A.lv[A.lc].m * A.lv[A.lc].n;
    off = (off + 3) & ~3;
    A.lds_elems = off;
    return off;
}

template <typename T>
void coarse_pack(const CoarseArgs<T> &A, const std::vector<double> *S, const std::vector<double> &binv,
                 std::vector<T> &out) {
    out.assign((size_t)A.pack_elems, (T)0);
    for (int l = A.lc; l < A.L; l++) {
        const int cnt = stencil_count(A.lv[l]);
        for (int k = 0; k < cnt; k++) out[(size_t)A.soff[l] + k] = (T)S[l][(size_t)k];
    }
    const int N = 2 * A.lv[A.L - 1].m * A.lv[A.L - 1].n;
    double *bi = reinterpret_cast<double *>(out.data() + A.binv_off);
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) {
            const double v = binv[(size_t)r * N + c];
            std::memcpy(bi + (size_t)c * N + r, &v, sizeof v);
        }
}
template void coarse_pack<double>(const CoarseArgs<double> &, const std::vector<double> *, const std::vector<double> &,
                                  std::vector<double> &);
template void coarse_pack<float>(const CoarseArgs<float> &, const std::vector<double> *, const std::vector<double> &,
                                 std::vector<float> &);
template int coarse_layout<double>(CoarseArgs<double> &);
template int coarse_layout<float>(CoarseArgs<float> &);

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T> int mg_smooth(const LevelArgs<T> &L, int sweeps, hipStream_t st) {
    const int hm = (L.m + 1) >> 1, hn = (L.n + 1) >> 1;
    const dim3 b(64, 4), g((hm + 63) / 64, (hn + 3) / 4);
    for (int sw = 0; sw < sweeps; sw++)
        for (int cy = 0; cy < 2; cy++) {
            hipLaunchKernelGGL(colour_kernel<T>, g, b, 0, st, L, 0, cy);
            hipLaunchKernelGGL(colour_kernel<T>, g, b, 0, st, L, 1, cy);
            if (L.xodd) hipLaunchKernelGGL(colour_kernel<T>, dim3(1, g.y), b, 0, st, L, 2, cy);
        }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_restrict(const LevelArgs<T> &F, const LevelArgs<T> &C, hipStream_t st) {
    const dim3 b(64, 4), g((C.m + 63) / 64, (C.n + 3) / 4);
    hipLaunchKernelGGL(restrict_kernel<T>, g, b, 0, st, F, C);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_prolong(const LevelArgs<T> &C, const LevelArgs<T> &F, hipStream_t st) {
    const dim3 b(64, 4), g((F.m + 63) / 64, (F.n + 3) / 4);
    hipLaunchKernelGGL(prolong_kernel<T>, g, b, 0, st, C, F);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> size_t coarse_lds_bytes(const CoarseArgs<T> &A) { return (size_t)A.lds_elems * sizeof(T); }

template <typename T> int mg_coarse(const CoarseArgs<T> &A, hipStream_t st) {
    hipLaunchKernelGGL(coarse_solve_kernel<T>, dim3(1), dim3(kCoarseThreads), coarse_lds_bytes(A), st, A);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_resid_norm(const LevelArgs<T> &L, double *partial, double *out, hipStream_t st) {
    const int nb = (L.m * L.n + 255) / 256;
    hipLaunchKernelGGL(resid_norm_kernel<T>, dim3(nb), dim3(256), 0, st, L, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, nb, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int mg_resid64_to32(const LevelArgs<double> &L, float *b32, float *z32, double *partial, double *out,
                    hipStream_t st) {
    const int nb = (L.m * L.n + 255) / 256;
    hipLaunchKernelGGL(resid64_to32_kernel, dim3(nb), dim3(256), 0, st, L, b32, z32, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, nb, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_fill_zero(T *p, size_t count, hipStream_t st) {
    hipLaunchKernelGGL(fill_zero_kernel<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
template int mg_fill_zero<double>(double *, size_t, hipStream_t);
template int mg_fill_zero<float>(float *, size_t, hipStream_t);

int mg_sum_partials(const double *partial, int np, double *out, hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, np, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int mg_correct64(double *z, const float *e, int count, hipStream_t st) {
    hipLaunchKernelGGL(correct64_kernel, dim3((count + 255) / 256), dim3(256), 0, st, z, e, count);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// the V-cycle from level l0 down (v_cycle_internal, tgmg.in.cc:341-359, with
// the smoothing of level l0 itself): grid kernels above lc, workgroup below.
// Level l0 starts from its current z; coarser levels from zero.
template <typename T> int mg_bottom(const CoarseArgs<T> &A, hipStream_t st) {
    return A.ld >= 0 ? mg_direct_solve(A.lv[A.ld], A.Gd, st) : mg_coarse(A, st);
}

template <typename T> int mg_vcycle_from(const CoarseArgs<T> &A, int l0, hipStream_t st) {
    int e;
    const int lb = bottom_level(A);
    if (l0 >= lb) return mg_bottom(A, st);
    if ((e = mg_smooth(A.lv[l0], A.nu1, st))) return e;
    for (int l = l0; l < lb; l++) {
        if ((e = mg_restrict(A.lv[l], A.lv[l + 1], st))) return e;
        if (l + 1 < lb && (e = mg_smooth(A.lv[l + 1], A.nu1, st))) return e;
    }
    if ((e = mg_bottom(A, st))) return e;
    for (int l = lb; l >= l0 + 1; l--) {
        if ((e = mg_prolong(A.lv[l], A.lv[l - 1], st))) return e;
        if ((e = mg_smooth(A.lv[l - 1], A.nu2, st))) return e;
    }
    return STZ_OK;
}

// one V-cycle on levels 0..L-1
template <typename T> int mg_vcycle(const CoarseArgs<T> &A, hipStream_t st) { return mg_vcycle_from(A, 0, st); }

// tgmg::fmg_pre (tgmg.in.cc:369-399): smooth the top level, restrict the
// residual down the hierarchy smoothing each level from zero, solve the
// levels of the workgroup solver, then on the way up prolong, smooth (nu2
// then nu1 sweeps) and run a V-cycle from each level, and finally add the
// prolongation of level 1 to the top level (no top-level smoothing after
// it: the solve's next V-cycle does that)
template <typename T> int mg_fmg_pre(const CoarseArgs<T> &A, hipStream_t st) {
    int e;
    if ((e = mg_smooth(A.lv[0], A.nu1, st))) return e;
    if (A.L == 1) return STZ_OK;
    const int lb = bottom_level(A);
    for (int l = 0; l < lb; l++) {
        if ((e = mg_restrict(A.lv[l], A.lv[l + 1], st))) return e;
        if (l + 1 < lb && (e = mg_smooth(A.lv[l + 1], A.nu1, st))) return e;
    }
    if ((e = mg_bottom(A, st))) return e;
    for (int l = lb - 1; l >= 1; l--) {
        if ((e = mg_prolong(A.lv[l + 1], A.lv[l], st))) return e;
        if ((e = mg_smooth(A.lv[l], A.nu2, st))) return e;
        if ((e = mg_vcycle_from(A, l, st))) return e;
    }
    return mg_prolong(A.lv[1], A.lv[0], st);
}

// The attribute belongs to the kernel, not to a simulation: it only ever
// grows, so that simulations of different sizes in one process all fit.
template <typename T> int mg_set_lds_limit(size_t bytes) {
    static size_t set = 0;
    if (bytes <= set) return STZ_OK;
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&coarse_solve_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    set = bytes;
    return STZ_OK;
}
template int mg_set_lds_limit<double>(size_t);
template int mg_set_lds_limit<float>(size_t);

template int mg_restrict<double>(const LevelArgs<double> &, const LevelArgs<double> &, hipStream_t);
template int mg_restrict<float>(const LevelArgs<float> &, const LevelArgs<float> &, hipStream_t);
template int mg_prolong<double>(const LevelArgs<double> &, const LevelArgs<double> &, hipStream_t);
template int mg_prolong<float>(const LevelArgs<float> &, const LevelArgs<float> &, hipStream_t);
template int mg_coarse<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_coarse<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_smooth<double>(const LevelArgs<double> &, int, hipStream_t);
template int mg_smooth<float>(const LevelArgs<float> &, int, hipStream_t);
template int mg_vcycle<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_vcycle_from<double>(const CoarseArgs<double> &, int, hipStream_t);
template int mg_bottom<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_bottom<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_vcycle_from<float>(const CoarseArgs<float> &, int, hipStream_t);
template int mg_fmg_pre<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_fmg_pre<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_vcycle<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_resid_norm<double>(const LevelArgs<double> &, double *, double *, hipStream_t);
template int mg_resid_norm<float>(const LevelArgs<float> &, double *, double *, hipStream_t);
template size_t coarse_lds_bytes<double>(const CoarseArgs<double> &);
template size_t coarse_lds_bytes<float>(const CoarseArgs<float> &);

// ---------------------------------------------------------------------------
// batched launches over the members of an ensemble (batch.h)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) resid_norm_batch_kernel(const RefineEntry *E, const int *map, int ne,
                                                               const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const RefineEntry &x = E[e];
    if (St && !St[x.member].resid) return;
    const LevelArgs<double> L = x.L;
    resid_norm_body(L, x.partial, l);
}

// tgmg::solve(type, tgmg_predict&) (tgmg.in.cc:266-301) as a state machine:
// the same tests, in the same order, on the same residuals as the host loop
// of the single simulation, advanced once per cycle slot.  acc is the
// residual after the cycles counted so far.
__device__ static void ens_resolve(EnsState &st, double acc, const EnsPredict &P) {
    for (;;) {
        if (st.phase == ENS_FIRST || st.phase == ENS_LOOP) {
            if (st.remaining > 0) {
                st.active = 1;
                return;
            }
            st.acc = acc;
            if (acc < st.tol) {
                if (st.phase == ENS_FIRST && st.lim > 0) st.lim -= 1 + st.lim / P.decay;
                st.phase = ENS_EXTRA;
                st.remaining = P.extra;
                continue;
            }
            st.k = P.unit ? 1 : st.k + 1;
            st.lim += st.k * P.mult;
            if (st.lim > P.max_thresh) {
                st.phase = ENS_FAILED;
                st.status = STZ_ERR_NOT_CONVERGED;
                st.active = 0;
                return;
            }
            st.n += st.k;
            st.remaining = st.k;
            st.phase = ENS_LOOP;
            st.active = 1;
            return;
        }
        if (st.phase == ENS_EXTRA) {
            if (st.remaining > 0) {
                st.active = 1;
                return;
            }
            st.phase = ENS_DONE;
            st.solves++;
            st.vcount += st.n;
            st.status = isnan(st.acc) ? STZ_ERR_NAN : STZ_OK;
        }
        st.active = 0;
        return;
    }
}

// one workgroup per member: mode 0 starts the solve from the residual of the
// step's first evaluation, mode 1 counts the slot's cycle of every member
// that ran it and tests its residual, mode 2 marks members corrected
__global__ void __launch_bounds__(256) control_batch_kernel(EnsState *S, const NormEntry *N, EnsPredict P,
                                                            int mode) {
    const NormEntry &x = N[blockIdx.x];
    EnsState &st = S[x.member];
    if (mode == 2) {
        if (threadIdx.x == 0 && st.phase == ENS_DONE) st.phase = ENS_CORRECTED;
        return;
    }
    if (mode == 1 && !st.active) return;
    const double sum = sum_partials_body(x.partial, x.np);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        st.n = st.lim / P.mult;
        st.k = 0;
        st.remaining = st.n;
        st.phase = ENS_FIRST;
        st.cycles_run = 0;
        st.status = STZ_OK;
    } else {
        st.remaining--;
        st.cycles_run++;
    }
    ens_resolve(st, sum / st.denom, P);
    // the single simulation evaluates the residual after every mixed cycle
    // but the solve's last (one_cycle's need_resid), and in FP64 only where
    // it tests it (top_l2)
    const bool last_extra = st.phase == ENS_EXTRA && st.remaining == 1;
    const bool before_test = st.phase != ENS_EXTRA && st.remaining == 1;
    st.resid = st.active && (P.mixed ? !last_extra : before_test) ? 1 : 0;
}

template <typename T> int batch_coarse(const CoarseEntry<T> *E, int ne, size_t lds, const EnsState *St,
                                       hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(coarse_solve_batch_kernel<T>, dim3(ne), dim3(kCoarseThreads), lds, st, E, St);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
template int batch_coarse<double>(const CoarseEntry<double> *, int, size_t, const EnsState *, hipStream_t);
template int batch_coarse<float>(const CoarseEntry<float> *, int, size_t, const EnsState *, hipStream_t);

template <typename T> int batch_set_lds_limit(size_t bytes) {
    static size_t set = 0;
    if (bytes <= set) return STZ_OK;
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&coarse_solve_batch_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    set = bytes;
    return STZ_OK;
}
template int batch_set_lds_limit<double>(size_t);
template int batch_set_lds_limit<float>(size_t);

int batch_resid_norm(const RefineEntry *E, const int *map, int ne, int total, const EnsState *St, hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(resid_norm_batch_kernel<double>, dim3(total), dim3(256), 0, st, E, map, ne, St);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int batch_control(EnsState *S, const NormEntry *N, int nmem, EnsPredict P, int mode, hipStream_t st) {
    if (nmem == 0) return STZ_OK;
    hipLaunchKernelGGL(control_batch_kernel, dim3(nmem), dim3(256), 0, st, S, N, P, mode);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

}  // namespace stz
