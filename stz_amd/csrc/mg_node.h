// mg_node.h -- node-level multigrid arithmetic shared by the grid kernels,
// the tiled kernels and the workgroup coarse solver.
#pragma once

#include "mg.h"

namespace stz {

template <typename T> struct alignas(2 * sizeof(T)) vec2s { T x, y; };

__device__ __forceinline__ static float fmav(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ static double fmav(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------------------
// (ax, ay) = A'z, the off-diagonal part of the 9-point 2x2-block stencil, from
// the eight neighbours in stencil order dl,dc,dr,cl,cr,ul,uc,ur (block k has
// entries s[4k..4k+3] = (a, b; c, d)).  The sum is formed as three row sums
// (below, centre, above) of fused multiply-adds, so that the six chains of
// each node are independent and the dependent depth is 8 instead of 16.
// Every kernel (per-colour, tiled, workgroup coarse solver, residuals) uses
// this one routine, so all paths round identically.
// ---------------------------------------------------------------------------
template <typename T, typename SP>
__device__ __forceinline__ static void offdiag9(const SP &s, const vec2s<T> &dl, const vec2s<T> &dc,
                                                const vec2s<T> &dr, const vec2s<T> &cl, const vec2s<T> &cr,
                                                const vec2s<T> &ul, const vec2s<T> &uc, const vec2s<T> &ur, T &ax,
                                                T &ay) {
    T xd = s[0] * dl.x, yd = s[2] * dl.x;
    xd = fmav(s[1], dl.y, xd);
    yd = fmav(s[3], dl.y, yd);
    xd = fmav(s[4], dc.x, xd);
    yd = fmav(s[6], dc.x, yd);
    xd = fmav(s[5], dc.y, xd);
    yd = fmav(s[7], dc.y, yd);
    xd = fmav(s[8], dr.x, xd);
    yd = fmav(s[10], dr.x, yd);
    xd = fmav(s[9], dr.y, xd);
    yd = fmav(s[11], dr.y, yd);
    T xc = s[12] * cl.x, yc = s[14] * cl.x;
    xc = fmav(s[13], cl.y, xc);
    yc = fmav(s[15], cl.y, yc);
    xc = fmav(s[20], cr.x, xc);
    yc = fmav(s[22], cr.x, yc);
    xc = fmav(s[21], cr.y, xc);
    yc = fmav(s[23], cr.y, yc);
    T xu = s[24] * ul.x, yu = s[26] * ul.x;
    xu = fmav(s[25], ul.y, xu);
    yu = fmav(s[27], ul.y, yu);
    xu = fmav(s[28], uc.x, xu);
    yu = fmav(s[30], uc.x, yu);
    xu = fmav(s[29], uc.y, xu);
    yu = fmav(s[31], uc.y, yu);
    xu = fmav(s[32], ur.x, xu);
    yu = fmav(s[34], ur.x, yu);
    xu = fmav(s[33], ur.y, xu);
    yu = fmav(s[35], ur.y, yu);
    ax = (xd + xu) + xc;
    ay = (yd + yu) + yc;
}

// Gauss-Seidel value from the off-diagonal sum: inv(cc) (b - A'z)
template <typename T, typename SP>
__device__ __forceinline__ static vec2s<T> gs_value(const SP &s, T bx, T by, T ax, T ay) {
    const T rx = bx - ax, ry = by - ay;
    vec2s<T> o;
    o.x = fmav(s[37], ry, s[36] * rx);
    o.y = fmav(s[39], ry, s[38] * rx);
    return o;
}

// residual b - A'z - cc z
template <typename T, typename SP>
__device__ __forceinline__ static void resid_value(const SP &s, T bx, T by, T ax, T ay, const vec2s<T> &z, T &rx,
                                                   T &ry) {
    rx = bx - ax - fmav(s[17], z.y, s[16] * z.x);
    ry = by - ay - fmav(s[19], z.y, s[18] * z.x);
}

// ---------------------------------------------------------------------------
// The top level's interior stencil (qs_multi.cc:31-54, qs_multi.hh:40-50)
// couples x to x and y to y only through the four edge neighbours (dc, cl,
// cr, uc) and x to y only through the four corners (dl, dr, ul, ur); its
// centre block and that block's inverse are diagonal.  The *_top variants
// drop exactly the entries that are zero in this pattern and keep the order
// of the others: a dropped fma(0, z, a) returns a unchanged, and a chain
// that starts from 0 * z continues with fma(s, z', +-0) = s * z', so on such
// a stencil they round as offdiag9 / gs_value / resid_value do (up to the
// sign of a zero) with half the multiply-adds.  sparse_pattern() is the
// host's test for the pattern.
// ---------------------------------------------------------------------------
template <typename T, typename SP>
__device__ __forceinline__ static void offdiag9_top(const SP &s, const vec2s<T> &dl, const vec2s<T> &dc,
                                                    const vec2s<T> &dr, const vec2s<T> &cl, const vec2s<T> &cr,
                                                    const vec2s<T> &ul, const vec2s<T> &uc, const vec2s<T> &ur,
                                                    T &ax, T &ay) {
    T xd = s[1] * dl.y, yd = s[2] * dl.x;
    xd = fmav(s[4], dc.x, xd);
    yd = fmav(s[7], dc.y, yd);
    xd = fmav(s[9], dr.y, xd);
    yd = fmav(s[10], dr.x, yd);
    T xc = s[12] * cl.x, yc = s[15] * cl.y;
    xc = fmav(s[20], cr.x, xc);
    yc = fmav(s[23], cr.y, yc);
    T xu = s[25] * ul.y, yu = s[26] * ul.x;
    xu = fmav(s[28], uc.x, xu);
    yu = fmav(s[31], uc.y, yu);
    xu = fmav(s[33], ur.y, xu);
    yu = fmav(s[34], ur.x, yu);
    ax = (xd + xu) + xc;
    ay = (yd + yu) + yc;
}

template <typename T, typename SP>
__device__ __forceinline__ static vec2s<T> gs_value_top(const SP &s, T bx, T by, T ax, T ay) {
    vec2s<T> o;
    o.x = s[36] * (bx - ax);
    o.y = s[39] * (by - ay);
    return o;
}

template <typename T, typename SP>
__device__ __forceinline__ static void resid_value_top(const SP &s, T bx, T by, T ax, T ay, const vec2s<T> &z,
                                                       T &rx, T &ry) {
    rx = bx - ax - s[16] * z.x;
    ry = by - ay - s[19] * z.y;
}

// dense or top-pattern arithmetic, chosen at compile time
template <bool SPARSE, typename T, typename SP>
__device__ __forceinline__ static void offdiag9_any(const SP &s, const vec2s<T> &dl, const vec2s<T> &dc,
                                                    const vec2s<T> &dr, const vec2s<T> &cl, const vec2s<T> &cr,
                                                    const vec2s<T> &ul, const vec2s<T> &uc, const vec2s<T> &ur,
                                                    T &ax, T &ay) {
    if constexpr (SPARSE) offdiag9_top<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
    else offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
}
template <bool SPARSE, typename T, typename SP>
__device__ __forceinline__ static vec2s<T> gs_value_any(const SP &s, T bx, T by, T ax, T ay) {
    if constexpr (SPARSE) return gs_value_top<T>(s, bx, by, ax, ay);
    else return gs_value<T>(s, bx, by, ax, ay);
}
template <bool SPARSE, typename T, typename SP>
__device__ __forceinline__ static void resid_value_any(const SP &s, T bx, T by, T ax, T ay, const vec2s<T> &z,
                                                       T &rx, T &ry) {
    if constexpr (SPARSE) resid_value_top<T>(s, bx, by, ax, ay, z, rx, ry);
    else resid_value<T>(s, bx, by, ax, ay, z, rx, ry);
}

// host: does the 40-entry stencil s have the top-level pattern?
inline bool sparse_pattern(const double *s) {
    static const int zero[] = {0, 3, 5, 6, 8, 11, 13, 14, 17, 18, 21, 22, 24, 27, 29, 30, 32, 35, 37, 38};
    for (int k : zero)
        if (s[k] != 0.0) return false;
    return true;
}

// the eight neighbours of node (i,j) of a natural-order level (zero outside
// the walls; x periodic)
template <typename T>
__device__ __forceinline__ static void neighbours(const LevelArgs<T> &L, int i, int j, vec2s<T> &dl, vec2s<T> &dc,
                                                  vec2s<T> &dr, vec2s<T> &cl, vec2s<T> &cr, vec2s<T> &ul,
                                                  vec2s<T> &uc, vec2s<T> &ur) {
    const int m = L.m;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const int jd = j > 0 ? j - 1 : j, ju = j < L.n - 1 ? j + 1 : j;
    const vec2s<T> *z = reinterpret_cast<const vec2s<T> *>(L.z);
    const vec2s<T> *zd = z + (size_t)m * jd, *zc = z + (size_t)m * j, *zu = z + (size_t)m * ju;
    dl = zd[il];
    dc = zd[i];
    dr = zd[ir];
    cl = zc[il];
    cr = zc[ir];
    ul = zu[il];
    uc = zu[i];
    ur = zu[ir];
    const vec2s<T> zero = {0, 0};
    if (j == 0) dl = dc = dr = zero;
    if (j == L.n - 1) ul = uc = ur = zero;
}

template <typename T> __device__ __forceinline__ static const T *stencil_at(const LevelArgs<T> &L, int i, int j) {
    return L.S + (size_t)(j * L.rs + i * L.cs) * 40;
}

// Gauss-Seidel update of node (i,j): z = inv(cc) (b - A'z)
template <typename T> __device__ __forceinline__ static void gs_node(const LevelArgs<T> &L, int i, int j) {
    const T *s = stencil_at(L, i, j);
    vec2s<T> dl, dc, dr, cl, cr, ul, uc, ur;
    neighbours(L, i, j, dl, dc, dr, cl, cr, ul, uc, ur);
    T ax, ay;
    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
    const int k = 2 * (i + L.m * j);
    const vec2s<T> o = gs_value<T>(s, L.b[k], L.b[k + 1], ax, ay);
    L.z[k] = o.x;
    L.z[k + 1] = o.y;
}

// residual b - A z at node (i,j)
template <typename T> __device__ __forceinline__ static void resid_node(const LevelArgs<T> &L, int i, int j, T &rx,
                                                                        T &ry) {
    const T *s = stencil_at(L, i, j);
    vec2s<T> dl, dc, dr, cl, cr, ul, uc, ur;
    neighbours(L, i, j, dl, dc, dr, cl, cr, ul, uc, ur);
    T ax, ay;
    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
    const int k = 2 * (i + L.m * j);
    const vec2s<T> z0 = {L.z[k], L.z[k + 1]};
    resid_value<T>(s, L.b[k], L.b[k + 1], ax, ay, z0, rx, ry);
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
