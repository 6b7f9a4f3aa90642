// mg_node.h -- node-level multigrid arithmetic shared by the grid kernels,
// the tiled kernels and the workgroup coarse solver.
#pragma once

#include "mg.h"

namespace stz {

// ---------------------------------------------------------------------------
// node-level device helpers (shared by the grid kernels and the workgroup
// coarse solver)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ static void acc2(T &ax, T &ay, const T *s, const T *zz) {
    const T zx = zz[0], zy = zz[1];
    ax += s[0] * zx + s[1] * zy;
    ay += s[2] * zx + s[3] * zy;
}

// sum of the off-diagonal blocks times z (the reference's mul_a)
template <typename T>
__device__ __forceinline__ static void offdiag(const LevelArgs<T> &L, int i, int j, const T *s, T &ax, T &ay) {
    const int m = L.m;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const T *z = L.z;
    ax = 0;
    ay = 0;
    if (j > 0) {
        const T *zr = z + 2 * m * (j - 1);
        acc2(ax, ay, s + 0, zr + 2 * il);
        acc2(ax, ay, s + 4, zr + 2 * i);
        acc2(ax, ay, s + 8, zr + 2 * ir);
    }
    {
        const T *zr = z + 2 * m * j;
        acc2(ax, ay, s + 12, zr + 2 * il);
        acc2(ax, ay, s + 20, zr + 2 * ir);
    }
    if (j < L.n - 1) {
        const T *zr = z + 2 * m * (j + 1);
        acc2(ax, ay, s + 24, zr + 2 * il);
        acc2(ax, ay, s + 28, zr + 2 * i);
        acc2(ax, ay, s + 32, zr + 2 * ir);
    }
}

template <typename T> __device__ __forceinline__ static const T *stencil_at(const LevelArgs<T> &L, int i, int j) {
    return L.S + (size_t)(j * L.rs + i * L.cs) * 40;
}

// Gauss-Seidel update of node (i,j): z = inv(cc) (b - A'z)
template <typename T> __device__ __forceinline__ static void gs_node(const LevelArgs<T> &L, int i, int j) {
    const T *s = stencil_at(L, i, j);
    T ax, ay;
    offdiag(L, i, j, s, ax, ay);
    const int k = 2 * (i + L.m * j);
    const T rx = L.b[k] - ax, ry = L.b[k + 1] - ay;
    L.z[k] = s[36] * rx + s[37] * ry;
    L.z[k + 1] = s[38] * rx + s[39] * ry;
}

// residual b - A z at node (i,j)
template <typename T> __device__ __forceinline__ static void resid_node(const LevelArgs<T> &L, int i, int j, T &rx,
                                                                        T &ry) {
    const T *s = stencil_at(L, i, j);
    T ax, ay;
    offdiag(L, i, j, s, ax, ay);
    const int k = 2 * (i + L.m * j);
    const T zx = L.z[k], zy = L.z[k + 1];
    rx = L.b[k] - ax - (s[16] * zx + s[17] * zy);
    ry = L.b[k + 1] - ay - (s[18] * zx + s[19] * zy);
}

// 1D prolongation weight of coarse c in fine f.  Periodic axis: f unwrapped,
// coarse indices compared unwrapped (tgmg.in.cc:552-621).
__device__ __forceinline__ static float pw_periodic(int f, int uf, int sc, int c) {
    const int fw = ((f % uf) + uf) % uf, sh = (f - fw) / uf;
    const int c0 = (fw >> 1) + sh * sc;
    if ((fw & 1) == 0) return c0 == c ? 1.f : 0.f;
    return (c0 == c || c0 + 1 == c) ? 0.5f : 0.f;
}
__device__ __forceinline__ static float pw_bounded(int f, int uf, int sc, int c) {
    if (f < 0 || f >= uf) return 0.f;
    if ((uf & 1) == 0 && f == uf - 1) return c == sc - 1 ? 1.f : 0.f;
    const int c0 = f >> 1;
    if ((f & 1) == 0) return c0 == c ? 1.f : 0.f;
    return (c0 == c || c0 + 1 == c) ? 0.5f : 0.f;
}

// coarse source at (I,J) = sum of P-weighted fine residuals; zero coarse z
template <typename T>
__device__ __forceinline__ static void restrict_node(const LevelArgs<T> &F, const LevelArgs<T> &C, int I, int J) {
    const int fx0 = 2 * I;
    const int fy0 = ((F.n & 1) == 0 && J == C.n - 1) ? F.n - 1 : 2 * J;
    T bx = 0, by = 0;
#pragma unroll
    for (int b = -1; b <= 1; b++) {
        const int fy = fy0 + b;
        const float wy = pw_bounded(fy, F.n, C.n, J);
        if (wy == 0.f) continue;
#pragma unroll
        for (int a = -1; a <= 1; a++) {
            const int fx = fx0 + a;
            const float wx = pw_periodic(fx, F.m, C.m, I);
            if (wx == 0.f) continue;
            const int fi = fx < 0 ? fx + F.m : (fx >= F.m ? fx - F.m : fx);
            T rx, ry;
            resid_node(F, fi, fy, rx, ry);
            const T w = (T)(wx * wy);
            bx += w * rx;
            by += w * ry;
        }
    }
    const int k = 2 * (I + C.m * J);
    C.b[k] = bx;
    C.b[k + 1] = by;
    C.z[k] = 0;
    C.z[k + 1] = 0;
}

// fine z at (i,j) += P zc
template <typename T>
__device__ __forceinline__ static void prolong_node(const LevelArgs<T> &C, const LevelArgs<T> &F, int i, int j) {
    const int sc = C.m;
    int cx0, cx1, cy0, cy1;
    T wx0, wx1, wy0, wy1;
    if ((i & 1) == 0) { cx0 = i >> 1; wx0 = 1; cx1 = cx0; wx1 = 0; }
    else { cx0 = i >> 1; cx1 = cx0 + 1 == sc ? 0 : cx0 + 1; wx0 = wx1 = (T)0.5; }
    if ((F.n & 1) == 0 && j == F.n - 1) { cy0 = C.n - 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
    else if ((j & 1) == 0) { cy0 = j >> 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
    else { cy0 = j >> 1; cy1 = cy0 + 1; wy0 = wy1 = (T)0.5; }
    const T *z0 = C.z + 2 * sc * cy0, *z1 = C.z + 2 * sc * cy1;
    const T ex = wy0 * (wx0 * z0[2 * cx0] + wx1 * z0[2 * cx1]) + wy1 * (wx0 * z1[2 * cx0] + wx1 * z1[2 * cx1]);
    const T ey = wy0 * (wx0 * z0[2 * cx0 + 1] + wx1 * z0[2 * cx1 + 1]) +
                 wy1 * (wx0 * z1[2 * cx0 + 1] + wx1 * z1[2 * cx1 + 1]);
    const int k = 2 * (i + F.m * j);
    F.z[k] += ex;
    F.z[k + 1] += ey;
}

}  // namespace stz
