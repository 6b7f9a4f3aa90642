// bd_levelset.hip -- gfx950 kernels of the narrow-band level set that carries
// the free boundary of the free-boundary step: interface velocity by the
// Newton closest point (levelset.cc:117-207, bd_sim velocity/bicubic_velocity
// from sb/main/bicubic_vel.pl), the band rebuilt outward from the moved
// interface (levelset.cc:209-345, 701-772, heap.cc), cell and node statuses
// (bd_sim.cc:125-172), and the extrapolation of the fields across the
// interface (traverse.hh:62-240).
//
// What is serial in the reference and how it runs here:
//  * The fast march (heap.cc:444-473) accepts one trial cell at a time from a
//    binary heap.  Here a round accepts every trial cell whose tentative value
//    lies within `delta` of the smallest, then recomputes the trial cells next
//    to the newly accepted ones.  Each tentative value comes from the
//    reference's own local update (heap.cc:183-268 and its corner/threeway
//    cases); cells accepted in one round do not see each other, so values can
//    differ from the heap order's by a fraction of the cell size where two
//    neighbours fall in the same round.
//  * The extrapolation walks the band in increasing phi, each cell fitting a
//    weighted plane to the cells already visited (traverse.hh:176-240).  The
//    set a cell fits to is fixed by phi alone (band cells with smaller phi),
//    so waves that run every cell whose smaller-phi neighbours are finished
//    give the reference's values exactly.  The node extrapolation
//    (traverse.hh:62-174) visits a node at the first cell of the band order
//    among its four cells; that order is rebuilt per node from phi (ties in
//    phi broken by cell index) and the same waves reproduce it.
//
// Floating-point expressions keep the reference's evaluation order and FMA
// contraction is off, so results agree bit for bit where the algorithm is
// the same.
#include "bd.h"

#include <algorithm>
#include <climits>

#pragma clang fp contract(off)

namespace stz {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int code(int8_t s) { return s & 15; }
constexpr int8_t kNew = 16;  // "accepted in this round" flag on top of the code

// ---------------------------------------------------------------------------
// Piecewise-cubic interpolation on a grid (levelset/built/bicubic.pl and
// sb/main/bicubic_vel.pl): Hermite cubics with centred-difference slopes in
// the interior, one-sided quadratics in the first and last interval.  Case a
// (x) / b (y): 0 first interval, 1 last, 2 interior.  Entry k of a row:
// 0 slope at the left node, 1 left value, 2 right value, 3 slope at the
// right node.  Each product is formed in the reference's order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool wpresent(int a, int k) { return !((a == 0 && k == 0) || (a == 1 && k == 3)); }
// value weight applied to data d: d*factors, sign separately
__device__ __forceinline__ double wval(int a, int k, double d, double e, double me) {
    switch (a * 4 + k) {
    case 1: return d * me * me;
    case 2: return d * e * (2 - e);
    case 3: return d * e * me;  // negative
    case 4: return d * e * me;
    case 5: return d * (1 - e * e);
    case 6: return d * e * e;
    case 8: return d * e * me * me;
    case 9: return d * me * me * (1 + 2 * e);
    case 10: return d * e * e * (3 - 2 * e);
    default: return d * e * e * me;  // 11, negative
    }
}
__device__ __forceinline__ bool wneg(int a, int k) { return k == 3 && a != 1; }
__device__ __forceinline__ double wder(int a, int k, double d, double e, double me) {
    switch (a * 4 + k) {
    case 1: return d * 2 * me;  // negative
    case 2: return d * 2 * me;
    case 3: return d * (2 * e - 1);
    case 4: return d * (1 - 2 * e);
    case 5: return d * 2 * e;  // negative
    case 6: return d * 2 * e;
    case 8: return d * me * (1 - 3 * e);
    case 9: return d * 6 * e * me;  // negative
    case 10: return d * 6 * e * me;
    default: return d * e * (3 * e - 2);
    }
}
__device__ __forceinline__ bool dneg(int k) { return k == 1; }
// weights alone (the y direction multiplies a whole row after forming them)
__device__ __forceinline__ double yval(int b, int l, double f, double mf) {
    switch (b * 4 + l) {
    case 1: return mf * mf;
    case 2: return f * (2 - f);
    case 3: return f * mf;
    case 4: return f * mf;
    case 5: return (1 - f * f);
    case 6: return f * f;
    case 8: return f * mf * mf;
    case 9: return mf * mf * (1 + 2 * f);
    case 10: return f * f * (3 - 2 * f);
    default: return f * f * mf;
    }
}
__device__ __forceinline__ double yder(int b, int l, double f, double mf) {
    switch (b * 4 + l) {
    case 1: return -2 * mf;
    case 2: return 2 * mf;
    case 3: return (2 * f - 1);
    case 4: return (1 - 2 * f);
    case 5: return -2 * f;
    case 6: return 2 * f;
    case 8: return mf * (1 - 3 * f);
    case 9: return -6 * f * mf;
    case 10: return 6 * f * mf;
    default: return f * (3 * f - 2);
    }
}
__device__ __forceinline__ double acc_term(double acc, double t, bool neg, bool first) {
    return first ? (neg ? -t : t) : (neg ? acc - t : acc + t);
}

// The 4x4 table of a case: row 0 slope-y row, rows 1,2 values, row 3 top
// slope-y row; columns as the entries above.  F at (col,row) = F[col+w*row].
struct Table {
    double t[4][4];
};
__device__ static void fill_table(const double *F, int w, int a, int b, int ij, double sx, double sy, double s2,
                                  Table &T) {
    const double *p = F + ij;
    if (a != 0) {
        T.t[1][0] = (p[1] - p[-1]) * sx;
        T.t[2][0] = (p[w + 1] - p[w - 1]) * sx;
    }
    if (a != 1) {
        T.t[1][3] = (p[2] - p[0]) * sx;
        T.t[2][3] = (p[w + 2] - p[w]) * sx;
    }
    if (b != 0) {
        T.t[0][1] = (p[w] - p[-w]) * sy;
        T.t[0][2] = (p[w + 1] - p[-w + 1]) * sy;
    }
    if (b != 1) {
        T.t[3][1] = (p[2 * w] - p[0]) * sy;
        T.t[3][2] = (p[2 * w + 1] - p[1]) * sy;
    }
    if (b != 0 && a != 0) T.t[0][0] = (p[w + 1] - p[-w + 1] - p[w - 1] + p[-w - 1]) * s2;
    if (b != 1 && a != 0) T.t[3][0] = (p[2 * w + 1] - p[1] - p[2 * w - 1] + p[-1]) * s2;
    if (b != 0 && a != 1) T.t[0][3] = (p[w + 2] - p[-w + 2] - p[w] + p[-w]) * s2;
    if (b != 1 && a != 1) T.t[3][3] = (p[2 * w + 2] - p[2] - p[2 * w] + p[0]) * s2;
    T.t[1][1] = p[0];
    T.t[1][2] = p[1];
    T.t[2][1] = p[w];
    T.t[2][2] = p[w + 1];
}

// locate (x,y) on a grid of nx x ny points: case, base index, local coords
__device__ __forceinline__ void locate(double x, double y, double ox, double oy, double xsp, double ysp, int nx,
                                       int ny, int &a, int &b, int &ij, double &e, double &f) {
    e = (x - ox) * xsp;
    f = (y - oy) * ysp;
    const int i = int(e), j = int(f);
    int bj;
    if (j < 1) { b = 0; bj = 0; }
    else if (j >= ny - 2) { b = 1; f -= double(ny - 2); bj = ny - 2; }
    else { b = 2; f -= double(j); bj = j; }
    int bi;
    if (i < 1) { a = 0; bi = 0; }
    else if (i >= nx - 2) { a = 1; e -= double(nx - 2); bi = nx - 2; }
    else { a = 2; e -= double(i); bi = i; }
    ij = bi + nx * bj;
}

// value of a row l with x-case a
__device__ __forceinline__ double row_val(const Table &T, int l, int a, double e, double me) {
    double acc = 0;
    bool first = true;
    for (int k = 0; k < 4; k++) {
        if (!wpresent(a, k)) continue;
        acc = acc_term(acc, wval(a, k, T.t[l][k], e, me), wneg(a, k), first);
        first = false;
    }
    return acc;
}
__device__ __forceinline__ double row_der(const Table &T, int l, int a, double e, double me) {
    double acc = 0;
    bool first = true;
    for (int k = 0; k < 4; k++) {
        if (!wpresent(a, k)) continue;
        acc = acc_term(acc, wder(a, k, T.t[l][k], e, me), dneg(k), first);
        first = false;
    }
    return acc;
}

// levelset::bicubic(x,y,phix,phiy) (levelset.cc:77-93 with built/bicubic.cc)
__device__ static double ls_bicubic(const BdConst &c, const double *phi, double x, double y, double &phix,
                                    double &phiy) {
    int a, b, ij;
    double e, f;
    locate(x, y, c.lax, c.lay, c.xsp, c.ysp, c.m, c.n, a, b, ij, e, f);
    const double me = 1 - e, mf = 1 - f;
    Table T;
    fill_table(phi, c.m, a, b, ij, 0.5, 0.5, 0.25, T);
    // phix = xsp*(sum_l yweight_l * (sum_k d * xslope_k)); rows multiply the
    // weight first, then the bracket
    double sx = 0, sv = 0, sy = 0;
    bool first = true;
    for (int l = 0; l < 4; l++) {
        if (!wpresent(b, l)) continue;
        const double w = yval(b, l, f, mf), rv = row_val(T, l, a, e, me);
        sx = acc_term(sx, w * row_der(T, l, a, e, me), wneg(b, l), first);
        sv = acc_term(sv, w * rv, wneg(b, l), first);
        // y slopes: literal signs are inside the weights
        sy = first ? yder(b, l, f, mf) * rv : sy + yder(b, l, f, mf) * rv;
        first = false;
    }
    phix = c.xsp * sx;
    phiy = c.ysp * sy;
    return sv;
}

// bd_sim::bicubic_velocity (generated by bicubic_vel.pl) on the node grid
__device__ static void vel_bicubic(const BdConst &c, const double *U, const double *V, double x, double y,
                                   double &qu, double &qv) {
    int a, b, ij;
    double e, f;
    locate(x, y, c.ax, c.ay, c.xsp, c.ysp, c.me, c.ne, a, b, ij, e, f);
    const double me = 1 - e, mf = 1 - f;
    Table TU, TV;
    fill_table(U, c.me, a, b, ij, 0.5, 0.5, 0.25, TU);
    fill_table(V, c.me, a, b, ij, 0.5, 0.5, 0.25, TV);
    double su = 0, sv = 0;
    bool first = true;
    for (int l = 0; l < 4; l++) {
        if (!wpresent(b, l)) continue;
        const double w = yval(b, l, f, mf);
        su = acc_term(su, w * row_val(TU, l, a, e, me), wneg(b, l), first);
        sv = acc_term(sv, w * row_val(TV, l, a, e, me), wneg(b, l), first);
        first = false;
    }
    qu = su;
    qv = sv;
}

// bd_sim::velocity (bicubic_vel.pl:6-11): normal speed of the boundary
__device__ static double bd_velocity(const BdConst &c, const double *phi, const BdNodes &nd, double x, double y) {
    double px, py, qu, qv;
    (void)ls_bicubic(c, phi, x, y, px, py);
    vel_bicubic(c, nd.u, nd.v, x, y, qu, qv);
    return -qu * px - qv * py;
}

// levelset::newton_failsafe (levelset.cc:148-207)
__device__ static bool newton_failsafe(const BdConst &c, const double *phi, const int8_t *st, int i, int ij,
                                       double x0, double y0, double &x, double &y) {
    const int m = c.m, mn = c.m * c.n;
    const int ts = 7 - code(st[ij]);
    bool hf, vf;
    double tdis, cphi = phi[ij];
    if (i > 0 && code(st[ij - 1]) == ts) {
        x = cphi / (phi[ij - 1] - cphi);
        if (i < m - 1 && code(st[ij + 1]) == ts) {
            tdis = cphi / (cphi - phi[ij + 1]);
            if (tdis < -x) x = tdis;
        }
        hf = true;
    } else if (i < m - 1 && code(st[ij + 1]) == ts) {
        x = cphi / (cphi - phi[ij + 1]);
        hf = true;
    } else hf = false;
    if (ij >= m && code(st[ij - m]) == ts) {
        y = cphi / (phi[ij - m] - cphi);
        if (ij < mn - m && code(st[ij + m]) == ts) {
            tdis = cphi / (cphi - phi[ij + m]);
            if (tdis < -y) y = tdis;
        }
        vf = true;
    } else if (ij < mn - m && code(st[ij + m]) == ts) {
        y = cphi / (cphi - phi[ij + m]);
        vf = true;
    } else vf = false;
    if (hf) {
        if (vf) {
            y *= c.dy;
            tdis = x * c.dx;
            cphi = tdis * y / (tdis * tdis + y * y);
            x = y * cphi + x0;
            y = tdis * cphi + y0;
            return true;
        }
        x = x * c.dx + x0;
        y = y0;
        return true;
    }
    if (vf) {
        x = x0;
        y = y * c.dy + y0;
        return true;
    }
    return false;
}

// levelset::newton (levelset.cc:117-146).  The reference compares with the
// C library's integer abs (its headers resolve abs(double) to abs(int)), so
// the convergence count advances whenever |phi| < 1 and the closing
// distance check passes whenever |x0-x| < 1: the loop runs eight Newton
// steps.  Reproduced here as written.
__device__ static bool ls_newton(const BdConst &c, const double *phi, const int8_t *st, int i, int ij,
                                 double &x0, double &y0, double &x, double &y) {
    const double dyy = c.dy / (double(c.m)), step_limit = sqrt(c.dx * c.dx + c.dy * c.dy);
    int iter = 0, cutoff = 0;
    double ph, phix, phiy, delx, dely, m0, m1;
    x0 = x = c.lax + i * c.dx;
    y0 = y = c.lay + (ij - i) * dyy;
    ph = ls_bicubic(c, phi, x, y, phix, phiy);
    do {
        delx = x0 - x;
        dely = y0 - y;
        m0 = delx * phix + dely * phiy + ph;
        m1 = phix * phix + phiy * phiy;
        iter++;
        m0 /= m1;
        if (iter > 100 || !isfinite(m0) || m0 > step_limit)
            return newton_failsafe(c, phi, st, i, ij, x0, y0, x, y);
        x += delx - m0 * phix;
        y += dely - m0 * phiy;
        ph = ls_bicubic(c, phi, x, y, phix, phiy);
        if (abs((int)ph) < 1e-10) cutoff++;
        else cutoff = 0;
    } while (cutoff < 8);
    return abs((int)(x0 - x)) < 3 * c.dx && abs((int)(y0 - y)) < 3 * c.dy
               ? true
               : newton_failsafe(c, phi, st, i, ij, x0, y0, x, y);
}

// ---------------------------------------------------------------------------
// statuses
// ---------------------------------------------------------------------------

// levelset::build_band scan (levelset.cc:217-273): 4/7 where phi > 0, 3/0
// elsewhere, 3/4 where a neighbour has the other sign
__global__ void __launch_bounds__(kThreads) scan_kernel(BdConst c, const double *__restrict__ phi, int8_t *st) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, mn = c.m * c.n;
    if (ij >= mn) return;
    const int i = ij % c.m;
    const bool pos = phi[ij] > 0;
    bool iface = false;
    if (i > 0) iface |= (phi[ij - 1] > 0) != pos;
    if (i < c.m - 1) iface |= (phi[ij + 1] > 0) != pos;
    if (ij >= c.m) iface |= (phi[ij - c.m] > 0) != pos;
    if (ij < mn - c.m) iface |= (phi[ij + c.m] > 0) != pos;
    st[ij] = pos ? (iface ? 4 : 7) : (iface ? 3 : 0);
}

// build_band (levelset.cc:279-296): interface cells get +-(distance to the
// Newton closest point); fail counts cells where Newton and its fallback fail
__global__ void __launch_bounds__(kThreads) newton_distance_kernel(BdConst c, const double *__restrict__ phi,
                                                                   const int8_t *__restrict__ st,
                                                                   double *__restrict__ out, int *fail) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, mn = c.m * c.n;
    if (ij >= mn) return;
    const int s = code(st[ij]);
    if (s != 3 && s != 4) return;
    double x0, y0, x, y;
    if (!ls_newton(c, phi, st, ij % c.m, ij, x0, y0, x, y)) {
        atomicAdd(fail, 1);
        out[ij] = phi[ij];
        return;
    }
    x -= x0;
    y -= y0;
    const double d = sqrt(x * x + y * y);
    out[ij] = s == 4 ? d : -d;
}

__global__ void __launch_bounds__(kThreads) copy_interface_kernel(BdConst c, const double *__restrict__ src,
                                                                  const int8_t *__restrict__ st, double *phi) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    const int s = code(st[ij]);
    if (s == 3 || s == 4) phi[ij] = src[ij];
}

// levelset::laplacian (levelset.cc:696-699)
__device__ __forceinline__ double ls_laplacian(const BdConst &c, const double *phi, int i, int ij) {
    const int m = c.m, mn = c.m * c.n;
    return (i != 0 && i != m - 1 ? c.xsp * c.xsp * (phi[ij + 1] + phi[ij - 1] - 2 * phi[ij]) : 0) +
           (ij >= m && ij < mn - m ? c.ysp * c.ysp * (phi[ij + m] + phi[ij - m] - 2 * phi[ij]) : 0);
}

// move_and_smooth velocity extension (levelset.hh:159-205)
__global__ void __launch_bounds__(kThreads) interface_velocity_kernel(BdConst c, const double *__restrict__ phi,
                                                                      const int8_t *__restrict__ st, BdNodes nd,
                                                                      double *__restrict__ vp) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    const int s = code(st[ij]);
    if (s != 3 && s != 4) return;
    const int i = ij % c.m;
    double x0, y0, x, y;
    double v = ls_newton(c, phi, st, i, ij, x0, y0, x, y) ? bd_velocity(c, phi, nd, x, y)
                                                           : bd_velocity(c, phi, nd, x0, y0);
    v += c.bc_smooth * ls_laplacian(c, phi, i, ij);
    vp[ij] = v;
}

// post_move_update (levelset.cc:701-715) moves the interface cells; the
// rest of the band is forgotten (move loops, levelset.hh:169-200)
__global__ void __launch_bounds__(kThreads) advance_kernel(BdConst c, double *__restrict__ phi,
                                                           const double *__restrict__ vp, int8_t *st, double dt) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    const int s = code(st[ij]);
    if (s == 3 || s == 4) phi[ij] += vp[ij] * dt;
    else st[ij] = s < 4 ? 0 : 7;
}

// ---------------------------------------------------------------------------
// the march (heap.cc), one direction at a time
// ---------------------------------------------------------------------------
template <bool POS> struct Dir {
    static constexpr int iface = POS ? 4 : 3, acc = POS ? 5 : 2, trial = POS ? 6 : 1, far = POS ? 7 : 0;
    __device__ static bool cm(double a, double b) { return POS ? a < b : a > b; }
    __device__ static double ct(double a, double b) { return POS ? a + b : a - b; }
    __device__ static bool in(int8_t s) { return POS ? code(s) < 6 : code(s) > 1; }
};

template <bool POS> struct Eik {
    using Z = Dir<POS>;
    const BdConst &c;
    const double *phi;
    const int8_t *st;
    double dx, dy, dx2, dy2, dxbydy, dx2bydy2;
    __device__ Eik(const BdConst &c_, const double *p, const int8_t *s) : c(c_), phi(p), st(s) {
        dx = c.dx;
        dy = c.dy;
        dx2 = dx * dx;
        dy2 = dy * dy;
        dxbydy = dx / dy;
        dx2bydy2 = dxbydy * dxbydy;
    }
    __device__ bool in(int ij) const { return Z::in(st[ij]); }
    // heap.cc:396-442
    __device__ bool look_l(int i, int ij, double &e) const {
        if (i > 1 && in(ij - 2) && Z::cm(phi[ij - 2], phi[ij - 1])) {
            e = 2 * phi[ij - 1] - 0.5 * phi[ij - 2];
            return true;
        }
        e = phi[ij - 1];
        return false;
    }
    __device__ bool look_r(int i, int ij, double &e) const {
        if (i < c.m - 2 && in(ij + 2) && Z::cm(phi[ij + 2], phi[ij + 1])) {
            e = 2 * phi[ij + 1] - 0.5 * phi[ij + 2];
            return true;
        }
        e = phi[ij + 1];
        return false;
    }
    __device__ bool look_d(int ij, double &e) const {
        const int m = c.m;
        if (ij >= 2 * m && in(ij - 2 * m) && Z::cm(phi[ij - 2 * m], phi[ij - m])) {
            e = 2 * phi[ij - m] - 0.5 * phi[ij - 2 * m];
            return true;
        }
        e = phi[ij - m];
        return false;
    }
    __device__ bool look_u(int ij, double &e) const {
        const int m = c.m, mn = c.m * c.n;
        if (ij < mn - 2 * m && in(ij + 2 * m) && Z::cm(phi[ij + 2 * m], phi[ij + m])) {
            e = 2 * phi[ij + m] - 0.5 * phi[ij + 2 * m];
            return true;
        }
        e = phi[ij + m];
        return false;
    }
    // heap.cc:385-394
    __device__ double corner_bail(bool h, bool v, double a, double b) const {
        const double ti = 2.0 / 3.0;
        double ht = h ? Z::ct(a, dx) * ti : Z::ct(a, dx);
        double vt = v ? Z::ct(b, dy) * ti : Z::ct(b, dy);
        return Z::cm(ht, vt) ? ht : vt;
    }
    // heap.cc:358-383
    __device__ double corner(bool h, bool v, double a, double b) const {
        const double f2a = 2.25 * (dy2 + dx2), f2b = 2.25 * dy2 + dx2, f2c = dy2 + 2.25 * dx2, f2d = dy2 + dx2;
        const double b1 = 2.25 * dx2 > dy2 ? 1 + 2.25 * dx2bydy2 : 1 + dy2 / (2.25 * dx2);
        const double b2 = dx2 > dy2 ? 1 + dx2bydy2 : 1 + dy2 / dx2;
        const double b3 = dx2 > 2.25 * dy2 ? 1 + dx2bydy2 / 2.25 : 1 + 2.25 * dy2 / dx2;
        const double i2a = dy2 / f2a, i2b = dy2 / f2b, i2c = dy2 / f2c, i2d = dy2 / f2d;
        double f;
        if (h) {
            if (v) {
                f = 1.5 * (a - b);
                f *= f;
                return f * b2 < f2a ? Z::ct(a * 1.5 + b * 1.5 * dx2bydy2, dxbydy * sqrt(f2a - f)) * i2a
                                    : corner_bail(true, true, a, b);
            }
            f = a - 1.5 * b;
            f *= f;
            return f * b3 < f2b ? Z::ct(a * 1.5 + b * dx2bydy2, dxbydy * sqrt(f2b - f)) * i2b
                                : corner_bail(true, false, a, b);
        }
        if (v) {
            f = 1.5 * a - b;
            f *= f;
            return f * b1 < f2c ? Z::ct(a + b * 1.5 * dx2bydy2, dxbydy * sqrt(f2c - f)) * i2c
                                : corner_bail(false, true, a, b);
        }
        f = a - b;
        f *= f;
        return f * b2 < f2d ? Z::ct(a + b * dx2bydy2, dxbydy * sqrt(f2d - f)) * i2d
                            : corner_bail(false, false, a, b);
    }
    __device__ static void clamp0(double &g) {
        if (Z::cm(g, 0)) g = 0;
    }
    // heap.cc:270-296; a,c horizontal (first, second order), b,d vertical
    __device__ double fourway(double a, double b, double c_, double d) const {
        double e = 2 * (c_ - a), f = 2 * (d - b), g, h;
        if (Z::cm(e, f)) {
            g = e - a;
            h = e - b;
            clamp0(g);
            clamp0(h);
            if (g * g + h * h * dx2bydy2 < dx2) {
                g = 1.5 * f - c_;
                h = f - b;
                clamp0(g);
                clamp0(h);
                if (g * g + h * h * dx2bydy2 < dx2) return corner(true, true, c_, d);
                return corner(true, false, c_, b);
            }
            return corner(false, false, a, b);
        }
        g = f - a;
        h = f - b;
        clamp0(g);
        clamp0(h);
        if (g * g + h * h * dx2bydy2 < dx2) {
            g = e - a;
            h = 1.5 * e - d;
            clamp0(g);
            clamp0(h);
            if (g * g + h * h * dx2bydy2 < dx2) return corner(true, true, c_, d);
            return corner(false, true, a, d);
        }
        return corner(false, false, a, b);
    }
    // heap.cc:298-326
    __device__ double threeway1(bool h1, bool h2, bool v, double a, double c_, double b) const {
        double d;
        if (h1) {
            if (h2) {
                d = Z::cm(a, c_) ? a : c_;
                return corner(true, v, d, b);
            }
            d = 2 * (a - c_);
            double e = d - c_, f = (v ? 1.5 * d : d) - b;
            clamp0(e);
            clamp0(f);
            return e * e + f * f * dx2bydy2 < dx2 ? corner(true, v, a, b) : corner(false, v, c_, b);
        }
        if (h2) {
            d = 2 * (c_ - a);
            double e = d - a, f = (v ? 1.5 * d : d) - b;
            clamp0(e);
            clamp0(f);
            return e * e + f * f * dx2bydy2 < dx2 ? corner(true, v, c_, b) : corner(false, v, a, b);
        }
        d = Z::cm(a, c_) ? a : c_;
        return corner(false, v, d, b);
    }
    // heap.cc:328-356
    __device__ double threeway2(bool h, bool v1, bool v2, double b, double a, double c_) const {
        double d;
        if (v1) {
            if (v2) {
                d = Z::cm(a, c_) ? a : c_;
                return corner(h, true, b, d);
            }
            d = 2 * (a - c_);
            double e = d - c_, f = (h ? 1.5 * d : d) - b;
            clamp0(e);
            clamp0(f);
            return f * f + e * e * dx2bydy2 < dx2 ? corner(h, true, b, a) : corner(h, false, b, c_);
        }
        if (v2) {
            d = 2 * (c_ - a);
            double e = d - a, f = (h ? 1.5 * d : d) - b;
            clamp0(e);
            clamp0(f);
            return f * f + e * e * dx2bydy2 < dx2 ? corner(h, true, b, c_) : corner(h, false, b, a);
        }
        d = Z::cm(a, c_) ? a : c_;
        return corner(h, false, b, d);
    }
    // heap.cc:183-268: tentative value of cell (i,ij) from its accepted
    // neighbours.  The look_* helpers overwrite their output before the
    // called case reads it, as in the reference's argument lists.
    __device__ double calc(int i, int ij) const {
        const int m = c.m, mn = c.m * c.n;
        const double ti = 2.0 / 3.0;
        double a, b, cc, d;
        const bool D = ij >= m && in(ij - m), U = ij < mn - m && in(ij + m);
        const bool R = i < m - 1 && in(ij + 1), L = i > 0 && in(ij - 1);
        if (D) {
            if (U) {
                if (R) {
                    if (L) {
                        if (look_d(ij, a)) {
                            if (look_u(ij, cc)) {
                                d = Z::cm(a, cc) ? a : cc;
                                const bool hl = look_l(i, ij, a), hr = look_r(i, ij, cc);
                                return threeway1(hl, hr, true, a, cc, d);
                            }
                            if (look_l(i, ij, b)) {
                                if (look_r(i, ij, d)) {
                                    b = Z::cm(b, d) ? b : d;
                                    return threeway2(true, true, false, b, a, cc);
                                }
                                return fourway(d, cc, b, a);
                            }
                            if (look_r(i, ij, d)) return fourway(b, cc, d, a);
                            b = Z::cm(b, d) ? b : d;
                            return threeway2(false, true, false, b, a, cc);
                        }
                        if (look_u(ij, cc)) {
                            if (look_l(i, ij, b)) {
                                if (look_r(i, ij, d)) {
                                    b = Z::cm(b, d) ? b : d;
                                    return threeway2(true, true, false, b, cc, a);
                                }
                                return fourway(d, a, b, cc);
                            }
                            if (look_r(i, ij, d)) return fourway(b, a, d, cc);
                            b = Z::cm(b, d) ? b : d;
                            return threeway2(false, true, false, b, cc, a);
                        }
                        d = Z::cm(a, cc) ? a : cc;
                        const bool hl = look_l(i, ij, a), hr = look_r(i, ij, cc);
                        return threeway1(hl, hr, false, a, cc, d);
                    }
                    const bool h = look_r(i, ij, b), v1 = look_u(ij, a), v2 = look_d(ij, cc);
                    return threeway2(h, v1, v2, b, a, cc);
                }
                if (L) {
                    const bool h = look_l(i, ij, b), v1 = look_u(ij, a), v2 = look_d(ij, cc);
                    return threeway2(h, v1, v2, b, a, cc);
                }
                cc = look_d(ij, a) ? Z::ct(a, dy) * ti : Z::ct(a, dy);
                d = look_u(ij, b) ? Z::ct(b, dy) * ti : Z::ct(b, dy);
                return Z::cm(cc, d) ? cc : d;
            }
            if (R) {
                if (L) {
                    const bool hl = look_l(i, ij, a), hr = look_r(i, ij, cc), v = look_d(ij, b);
                    return threeway1(hl, hr, v, a, cc, b);
                }
                const bool h = look_r(i, ij, a), v = look_d(ij, b);
                return corner(h, v, a, b);
            }
            if (L) {
                const bool h = look_l(i, ij, a), v = look_d(ij, b);
                return corner(h, v, a, b);
            }
            return look_d(ij, a) ? Z::ct(a, dy) * ti : Z::ct(a, dy);
        }
        if (U) {
            if (R) {
                if (L) {
                    const bool hl = look_l(i, ij, a), hr = look_r(i, ij, cc), v = look_u(ij, b);
                    return threeway1(hl, hr, v, a, cc, b);
                }
                const bool h = look_r(i, ij, a), v = look_u(ij, b);
                return corner(h, v, a, b);
            }
            if (L) {
                const bool h = look_l(i, ij, a), v = look_u(ij, b);
                return corner(h, v, a, b);
            }
            return look_u(ij, a) ? Z::ct(a, dy) * ti : Z::ct(a, dy);
        }
        if (R) {
            if (L) {
                cc = look_l(i, ij, a) ? Z::ct(a, dx) * ti : Z::ct(a, dx);
                d = look_r(i, ij, b) ? Z::ct(b, dx) * ti : Z::ct(b, dx);
                return Z::cm(cc, d) ? cc : d;
            }
            return look_r(i, ij, a) ? Z::ct(a, dx) * ti : Z::ct(a, dx);
        }
        if (L) return look_l(i, ij, a) ? Z::ct(a, dx) * ti : Z::ct(a, dx);
        return 0;
    }
};

// order-preserving map of a double to an unsigned key
__device__ __forceinline__ unsigned long long okey(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_inv(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// check_gridpoint from the interface cells (heap.cc:26-46)
template <bool POS>
__global__ void __launch_bounds__(kThreads) march_seed_kernel(BdConst c, double *phi, int8_t *st,
                                                              unsigned long long *dmin, int *counter) {
    using Z = Dir<POS>;
    const int ij = blockIdx.x * kThreads + threadIdx.x, mn = c.m * c.n;
    if (ij == 0) {
        *dmin = ~0ull;
        *counter = 0;
    }
    if (ij >= mn) return;
    if (code(st[ij]) != Z::far) return;
    const int i = ij % c.m, m = c.m;
    const bool seed = (i > 0 && code(st[ij - 1]) == Z::iface) || (i < m - 1 && code(st[ij + 1]) == Z::iface) ||
                      (ij >= m && code(st[ij - m]) == Z::iface) || (ij < mn - m && code(st[ij + m]) == Z::iface);
    if (!seed) return;
    Eik<POS> E(c, phi, st);
    const double v = E.calc(i, ij);
    phi[ij] = v;
    st[ij] = Z::trial;
}

// The march's per-round kernels run over the whole grid, of which only the
// band's cells take part: the minimum and the acceptance take cells a grid
// stride apart (kMarchBlocks workgroups), so they are bound by the one byte
// of status per cell they read rather than by dispatching a workgroup per
// 256 cells.  The update, which reads a cell's neighbours, keeps a thread
// per cell.
constexpr int kMarchBlocks = 2048;

template <bool POS>
__global__ void __launch_bounds__(kThreads) march_min_kernel(BdConst c, const double *__restrict__ phi,
                                                             const int8_t *__restrict__ st,
                                                             unsigned long long *dmin) {
    using Z = Dir<POS>;
    __shared__ unsigned long long red[kThreads / 64];
    const int mn = c.m * c.n;
    unsigned long long k = ~0ull;
    for (int ij = blockIdx.x * kThreads + threadIdx.x; ij < mn; ij += gridDim.x * kThreads)
        if (code(st[ij]) == Z::trial) {
            const unsigned long long t = okey(POS ? phi[ij] : -phi[ij]);
            k = t < k ? t : k;
        }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_down(k, o, 64);
        k = t < k ? t : k;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; w++) k = red[w] < k ? red[w] : k;
        if (k != ~0ull) atomicMin(dmin, k);
    }
}

// accept every trial cell within delta of the smallest (and below the band
// limit, heap.cc:448); clear last round's flags; count acceptances
template <bool POS>
__global__ void __launch_bounds__(kThreads) march_accept_kernel(BdConst c, const double *__restrict__ phi,
                                                                int8_t *st, const unsigned long long *dmin,
                                                                double delta, int *counter) {
    using Z = Dir<POS>;
    const int mn = c.m * c.n;
    const unsigned long long dm = *dmin;
    int acc = 0;
    for (int ij = blockIdx.x * kThreads + threadIdx.x; ij < mn; ij += gridDim.x * kThreads) {
        const int8_t s = st[ij];
        if (s & kNew) {
            st[ij] = s & 15;
            continue;
        }
        if (code(s) != Z::trial || dm == ~0ull) continue;
        const double kmin = okey_inv(dm), key = POS ? phi[ij] : -phi[ij];
        if (key < kmin + delta && key < c.phi_lim) {
            st[ij] = Z::acc | kNew;
            acc++;
        }
    }
    if (acc) atomicAdd(counter, acc);
}

// heap::update / scan after acceptances (heap.cc:166-181, 466-470):
// neighbours along the axes recompute, far cells next to an accepted cell
// become trial
template <bool POS>
__global__ void __launch_bounds__(kThreads) march_update_kernel(BdConst c, double *phi, int8_t *st,
                                                                unsigned long long *dmin, int *counter,
                                                                int *last) {
    using Z = Dir<POS>;
    const int ij = blockIdx.x * kThreads + threadIdx.x, mn = c.m * c.n, m = c.m;
    if (ij == 0) {
        *last = *counter;
        *counter = 0;
        *dmin = ~0ull;
    }
    if (ij >= mn) return;
    const int s = code(st[ij]);
    if (s != Z::trial && s != Z::far) return;
    const int i = ij % m;
    auto isnew = [&](int k) { return (st[k] & kNew) != 0; };
    const bool n1 = (i > 0 && isnew(ij - 1)) || (i < m - 1 && isnew(ij + 1)) || (ij >= m && isnew(ij - m)) ||
                    (ij < mn - m && isnew(ij + m));
    bool go = n1;
    if (s == Z::trial && !go)
        go = (i > 1 && isnew(ij - 2)) || (i < m - 2 && isnew(ij + 2)) || (ij >= 2 * m && isnew(ij - 2 * m)) ||
             (ij < mn - 2 * m && isnew(ij + 2 * m));
    if (!go) return;
    Eik<POS> E(c, phi, st);
    phi[ij] = E.calc(i, ij);
    if (s == Z::far) st[ij] = Z::trial;
}

__global__ void flag_kernel(const int *last, volatile int *hflag) {
    if (threadIdx.x == 0) *hflag = *last;
}

// final statuses and clean (levelset.cc:416-421, 740-771): band cells are
// those accepted or on the interface; move mode takes the side from the
// sign of phi (phi >= 0 outside, the bisection of levelset.cc:743-746) and
// marks interface cells by an opposite-side neighbour
__global__ void __launch_bounds__(kThreads) finish_kernel(BdConst c, double *phi, double *vp, int8_t *st,
                                                          int build) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, mn = c.m * c.n, m = c.m;
    if (ij >= mn) return;
    const int s = code(st[ij]);
    const bool band = s >= 2 && s <= 5;
    if (!band) {
        phi[ij] = s <= 1 ? -c.phi_lim : c.phi_lim;
        vp[ij] = 0;
        st[ij] = (int8_t)s;
        return;
    }
    if (build) {
        st[ij] = (int8_t)s;
        return;
    }
    auto side = [&](int k) {
        const int t = code(st[k]);
        return (t >= 2 && t <= 5) ? phi[k] >= 0 : t >= 4;
    };
    const bool pos = phi[ij] >= 0;
    const int i = ij % m;
    const bool iface = (i > 0 && side(ij - 1) != pos) || (i < m - 1 && side(ij + 1) != pos) ||
                       (ij >= m && side(ij - m) != pos) || (ij < mn - m && side(ij + m) != pos);
    st[ij] = pos ? (iface ? 4 : 5) : (iface ? 3 : 2);
}

// bd_sim::set_staggered_status (bd_sim.cc:125-172)
__global__ void __launch_bounds__(kThreads) node_status_kernel(BdConst c, const double *__restrict__ phi,
                                                               const int8_t *__restrict__ st, int8_t *nst) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= c.me * c.ne) return;
    const int ie = k % c.me, je = k / c.me, m = c.m, n = c.n;
    const int cb = ie + m * je;  // cell (ie,je)
    int kk = 0, l = 0;
    if (ie > 0) {
        if (je > 0) { kk++; if (code(st[cb - m - 1]) > 3) l++; }
        if (je < n) { kk++; if (code(st[cb - 1]) > 3) l++; }
    }
    if (ie < m) {
        if (je > 0) { kk++; if (code(st[cb - m]) > 3) l++; }
        if (je < n) { kk++; if (code(st[cb]) > 3) l++; }
    }
    int8_t out;
    if (l == 0) out = 0;
    else if (l == kk) out = 7;
    else {
        double phit = 0;
        if (ie > 0) {
            if (je > 0) phit += phi[cb - m - 1];
            if (je < n) phit += phi[cb - 1];
        }
        if (ie < m) {
            if (je > 0) phit += phi[cb - m];
            if (je < n) phit += phi[cb];
        }
        out = phit > 0 ? 7 : 0;
    }
    nst[k] = out;
}

// ---------------------------------------------------------------------------
// extrapolation of the cell and node fields (traverse.hh:62-240,
// extrap.hh:26-72)
//
// The reference walks the entries outside the interface in one order
// (cells: increasing phi; nodes: the key below), and each takes a weighted
// linear least-squares fit over the entries of its window that precede it.
// Here the walk is a dependency graph: a pending entry depends on the
// pending entries of its window that precede it.  An entry is set once
// everything it depends on is set (Kahn's topological order), so each fit
// reads exactly the values the reference's reads and sums them in the same
// order.  The walk goes in waves, one launch each: one wavefront (a team)
// per ready entry sets it and releases the entries that depend on it into
// the next wave's list.  The number of waves is the longest dependency
// chain.
//
// done[]: 0 source, -1 pending, INT_MAX not in the band; a set entry holds
// the wave that set it.  deg[] (pending entries): the count of unset
// dependencies in the low 16 bits and, for cells, the window half-width
// above them.  counter: [2] failed fits, [3] pending entries, [5] order
// violations (an entry fitted before one it depends on, which would be a
// fault of this code: reported as an error), [8] waves, [9] entries set,
// [10] the largest cell window half-width beyond 3, [12..14] the wave list
// counts.
// ---------------------------------------------------------------------------
__global__ void zero_counters_kernel(int *counter) {
    const int t = threadIdx.x;
    if ((t >= 2 && t <= 5) || (t >= 8 && t <= 14)) counter[t] = 0;
}

// *cnt += the number of lanes with pred, one atomic per wavefront
__device__ __forceinline__ void wave_count(bool pred, int *cnt) {
    const unsigned long long act = __ballot(1), msk = __ballot(pred);
    if (msk && (int)__lane_id() == __builtin_ctzll(act)) atomicAdd(cnt, (int)__popcll(msk));
}

// list[(*cnt)++] = v for the lanes with pred, one atomic per wavefront
__device__ __forceinline__ void wave_push(bool pred, int v, int *list, int *cnt) {
    const unsigned long long act = __ballot(1), msk = __ballot(pred);
    if (!msk) return;
    const int leader = __builtin_ctzll(act);
    int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(cnt, (int)__popcll(msk));
    base = __shfl(base, leader);
    if (pred)
        list[base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u))] = v;
}

// LDS written by some lanes of a wavefront and read by others
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void window(int i, int j, int m, int n, int r, int &ilo, int &ihi, int &jlo, int &jhi) {
    ilo = i <= r ? -i : -r;
    ihi = i + r >= m ? m - 1 - i : r;
    jlo = j <= r ? -j : -r;
    jhi = j + r >= n ? n - 1 - j : r;
}

// the determinant of the fit's 3x3 system from its weighted sums
__device__ __forceinline__ double fit_det(double s, double sx, double sy, double sxx, double sxy, double syy) {
    return s * (sxy * sxy - sxx * syy) + sx * sx * syy - 2 * sx * sxy * sy + sy * sy * sxx;
}

// The window half-width of pending cell ij (traverse.hh:189-214): the
// smallest r in 3..7 whose band cells with a smaller phi give a nonsingular
// system, 8 if none does (the fit fails).  The system depends on phi and
// band membership only, so this is known before the walk.
__device__ static int xc_radius(const BdConst &c, const double *__restrict__ phi, const int8_t *__restrict__ st,
                                int ij) {
    const int m = c.m, n = c.n, i = ij % m, j = ij / m;
    const double ph = phi[ij];
    for (int r = 3; r <= 7; r++) {
        int ilo, ihi, jlo, jhi;
        window(i, j, m, n, r, ilo, ihi, jlo, jhi);
        double s = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0;
        for (int dj = jlo; dj <= jhi; dj++)
            for (int di = ilo; di <= ihi; di++) {
                const int k = ij + dj * m + di, sk = code(st[k]);
                if (sk < 2 || sk > 5) continue;
                double w = phi[k];
                if (!(w < ph)) continue;
                w = ph - w;
                s += w;
                sx += di * w;
                sy += dj * w;
                sxx += di * di * w;
                sxy += di * dj * w;
                syy += dj * dj * w;
            }
        if (!(fabs(fit_det(s, sx, sy, sxx, sxy, syy)) < 1e-10)) return r;
    }
    return 8;
}

// statuses 2-3: source, 4-5: pending, others outside the band; a pending
// cell gets its window, its dependency count, and the ready list if that is
// zero
__global__ void __launch_bounds__(kThreads) xc_prepare_kernel(BdConst c, const double *__restrict__ phi,
                                                              const int8_t *__restrict__ st, int *done, int *deg,
                                                              int *counter, int *ready) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, m = c.m, n = c.n;
    int d = INT_MAX;
    if (ij < m * n) {
        const int s = code(st[ij]);
        d = (s >= 2 && s <= 3) ? 0 : ((s >= 4 && s <= 5) ? -1 : INT_MAX);
        done[ij] = d;
    }
    const bool pend = d == -1;
    wave_count(pend, &counter[3]);
    int cnt = 0;
    if (pend) {
        const int r = xc_radius(c, phi, st, ij), rw = r > 7 ? 7 : r;
        if (rw > 3) atomicMax(&counter[10], rw);
        const int i = ij % m, j = ij / m;
        const double ph = phi[ij];
        int ilo, ihi, jlo, jhi;
        window(i, j, m, n, rw, ilo, ihi, jlo, jhi);
        for (int dj = jlo; dj <= jhi; dj++)
            for (int di = ilo; di <= ihi; di++) {
                const int k = ij + dj * m + di, sk = code(st[k]);
                cnt += (sk >= 4 && sk <= 5 && phi[k] < ph) ? 1 : 0;
            }
        deg[ij] = cnt | r << 16;
    }
    wave_push(pend && cnt == 0, ij, ready, &counter[13]);
}

// one team's staging for a fit: the addend of each of its (at most 21)
// sums for each of 64 window entries, rows padded against LDS bank
// conflicts
constexpr int kSums = 21, kPad = 65;
struct TeamBuf {
    double a[kSums][kPad];
};

// The addends of one window entry, formed as the sequential loop forms them
// (traverse.hh:196-212 for cells, 145-152 for nodes): 0..5 the weight sums
// s, sx, sy, sxx, sxy, syy (cells di * w, nodes w * di) and 6 + 3q + e field
// q times w, w x, w y.  Zero for an entry not used: a sum that starts at +0
// never holds -0, so adding +0 leaves it as it was.
template <bool CELL, int NV>
__device__ __forceinline__ void stage_addends(TeamBuf &B, int lane, bool used, int di, int dj, double w,
                                              const double *v) {
    double t[6];
    if (CELL) {
        t[0] = w;
        t[1] = di * w;
        t[2] = dj * w;
        t[3] = di * di * w;
        t[4] = di * dj * w;
        t[5] = dj * dj * w;
    } else {
        t[0] = w;
        t[1] = w * di;
        t[2] = w * dj;
        t[3] = w * di * di;
        t[4] = w * di * dj;
        t[5] = w * dj * dj;
    }
#pragma unroll
    for (int a = 0; a < 6; a++) B.a[a][lane] = used ? t[a] : 0.0;
    const double wx = w * di, wy = w * dj;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        B.a[6 + 3 * q][lane] = used ? w * v[q] : 0.0;
        B.a[7 + 3 * q][lane] = used ? wx * v[q] : 0.0;
        B.a[8 + 3 * q][lane] = used ? wy * v[q] : 0.0;
    }
}

// lane a < nsums adds up row a of the staging over cnt entries, in order
__device__ __forceinline__ double team_sum(const TeamBuf &B, int lane, int nsums, int cnt, double acc) {
    if (lane < nsums) {
        const double *row = B.a[lane];
        int u = 0;
        for (; u + 8 <= cnt; u += 8) {
            double x[8];
#pragma unroll
            for (int e = 0; e < 8; e++) x[e] = row[u + e];
#pragma unroll
            for (int e = 0; e < 8; e++) acc += x[e];
        }
        for (; u < cnt; u++) acc += row[u];
    }
    return acc;
}

// the fit from the team's sums: lane 6 + 3q returns field q, and false if
// the system is singular
__device__ __forceinline__ bool team_solve(double acc, int lane, double &out) {
    const double s = __shfl(acc, 0), sx = __shfl(acc, 1), sy = __shfl(acc, 2), sxx = __shfl(acc, 3),
                 sxy = __shfl(acc, 4), syy = __shfl(acc, 5);
    const double p1 = __shfl(acc, (lane + 1) & 63), p2 = __shfl(acc, (lane + 2) & 63);
    double det = fit_det(s, sx, sy, sxx, sxy, syy);
    if (fabs(det) < 1e-10) return false;
    det = 1 / det;
    const double al = (sxy * sxy - sxx * syy) * det, be = (sx * syy - sxy * sy) * det,
                 ga = (sxx * sy - sxy * sx) * det;
    out = acc * al + p1 * be + p2 * ga;
    return true;
}

// what a lane of a team reads about one neighbour of the entry it releases,
// loaded before the entry's fit so that the two loads overlap (none of it
// changes in between: a dependant stays pending until its dependencies are
// released)
struct CellProbe {
    int k, dx, dy, d, g;
    double p;
    bool in;
};

// the cell walk (traverse.hh:176-240): a cell depends on the pending cells
// of its window with a smaller phi
struct CellWalk {
    BdConst c;
    const double *phi;
    BdCells f;
    int *done, *deg, *counter;
    int R;  // the largest window half-width

    __device__ void fit(int wave, int ij, int lane, TeamBuf &B) const {
        const int m = c.m, n = c.n, i = ij % m, j = ij / m, r = deg[ij] >> 16;
        if (r > 7) {  // no window gives a fit (traverse.hh:214-218)
            if (lane == 0) {
                atomicAdd(&counter[2], 1);
                done[ij] = wave;
            }
            return;
        }
        const double ph = phi[ij];
        int ilo, ihi, jlo, jhi;
        window(i, j, m, n, r, ilo, ihi, jlo, jhi);
        const int wd = ihi - ilo + 1, nwin = wd * (jhi - jlo + 1);
        double acc = 0;
        bool bad = false;
        for (int base = 0; base < nwin; base += 64) {
            const int t = base + lane, tc = t < nwin ? t : base;
            const int di = ilo + tc % wd, dj = jlo + tc / wd, k = ij + dj * m + di, d = done[k];
            const double pk = phi[k], v[5] = {f.p[k], f.q[k], f.s[k], f.tau[k], f.chi[k]};
            const bool used = t < nwin && d != INT_MAX && pk < ph;
            bad |= used && (d < 0 || d >= wave);
            stage_addends<true, 5>(B, lane, used, di, dj, ph - pk, v);
            wave_sync();
            acc = team_sum(B, lane, 21, nwin - base < 64 ? nwin - base : 64, acc);
            wave_sync();
        }
        double out = 0;
        bad |= !team_solve(acc, lane, out);  // the window was chosen nonsingular
        if (lane >= 6 && lane < 21 && (lane - 6) % 3 == 0) {
            switch ((lane - 6) / 3) {
            case 0: f.p[ij] = out; break;
            case 1: f.q[ij] = out; break;
            case 2: f.s[ij] = out; break;
            case 3: f.tau[ij] = out; break;
            default: f.chi[ij] = out;
            }
        }
        if (__any(bad) && lane == 0) atomicAdd(&counter[5], 1);
        if (lane == 0) done[ij] = wave;
    }

    // neighbour base + lane of the (2R+1)^2 square around ij
    __device__ CellProbe probe(int ij, int lane, int base) const {
        const int m = c.m, n = c.n, i = ij % m, j = ij / m, span = 2 * R + 1;
        const int t = base + lane, dx = t % span - R, dy = t / span - R;
        const bool in = t < span * span && (dx | dy) != 0 && i + dx >= 0 && i + dx < m && j + dy >= 0 && j + dy < n;
        const int k = in ? ij + dy * m + dx : ij;
        return CellProbe{k, dx, dy, done[k], deg[k], phi[k], in};
    }

    // a pending cell k depends on ij if ij lies in its window and has the
    // smaller phi: one count less, and k is ready at zero
    __device__ void release(int ij, int lane, const CellProbe &first, int *next, int *nnext) const {
        const int span = 2 * R + 1, area = span * span;
        const double ph = phi[ij];
        for (int base = 0; base < area; base += 64) {
            const CellProbe P = base == 0 ? first : probe(ij, lane, base);
            const int rk = (P.g >> 16) > 7 ? 7 : P.g >> 16;
            const bool dep = P.in && P.d == -1 && ph < P.p && abs(P.dx) <= rk && abs(P.dy) <= rk;
            const bool ready = dep && (atomicSub(&deg[P.k], 1) & 0xffff) == 1;
            wave_push(ready, P.k, next, nnext);
        }
    }
};

// traverse::staggered_phi (traverse.hh:109-117)
__device__ __forceinline__ double staggered_phi(const BdConst &c, const double *phi, int ie, int je) {
    const int m = c.m, n = c.n;
    int ij = ie + m * je;
    const int di = ie == 0 ? 1 : (ie == m ? -1 : 0), dj = je == 0 ? 1 : (je == n ? -1 : 0);
    if (di == 0 && dj == 0) return 0.25 * (phi[ij - m - 1] + phi[ij - m] + phi[ij - 1] + phi[ij]);
    ij += di + m * dj;
    return (0.5 + dj) * ((0.5 + di) * phi[ij - m - 1] + (0.5 - di) * phi[ij - m]) +
           (0.5 - dj) * ((0.5 + di) * phi[ij - 1] + (0.5 - di) * phi[ij]);
}

// Node keys.  The staggered pass walks the band cells with phi >= s_point
// in increasing phi; a node (status 7) is set when the first of its in-box
// cells is walked, provided all its in-box cells belong to that walk
// (traverse.hh:62-107, the mask taken from the regular pass just before it);
// nodes completed by one cell go in increasing staggered phi, then in corner
// order.  key = (phi of that cell, its index, staggered phi, corner).
// done: 0 source (status 0), -1 pending, INT_MAX never set.
__global__ void __launch_bounds__(kThreads) xn_prepare_kernel(BdConst c, const double *__restrict__ phi,
                                                              const int8_t *__restrict__ st,
                                                              const int8_t *__restrict__ nst, int *done,
                                                              double *keyd, int *keyi, int *counter) {
    const int k = blockIdx.x * kThreads + threadIdx.x, mne = c.me * c.ne;
    bool pend = false;
    if (k < mne) {
        const int ie = k % c.me, je = k / c.me, m = c.m, n = c.n;
        const double spt = -1.2 * sqrt(c.dx * c.dx + c.dy * c.dy);
        keyd[2 * (size_t)k + 1] = staggered_phi(c, phi, ie, je);
        int d = 0;
        if (nst[k] != 0) {
            bool ok = true;
            double best = 0;
            int bidx = -1, corner = 0;
            for (int q = 0; q < 4; q++) {  // cells (ie-1,je-1), (ie,je-1), (ie-1,je), (ie,je)
                const int ci = ie - 1 + (q & 1), cj = je - 1 + (q >> 1);
                if (ci < 0 || ci >= m || cj < 0 || cj >= n) continue;
                const int cidx = ci + m * cj, s = code(st[cidx]);
                const double cp = phi[cidx];
                if (!(s >= 2 && s <= 5) || cp < spt) {
                    ok = false;
                    break;
                }
                if (bidx < 0 || cp < best || (cp == best && cidx < bidx)) {
                    best = cp;
                    bidx = cidx;
                    corner = 3 - q;  // the node is this corner of the cell
                }
            }
            if (!ok || bidx < 0) {
                d = INT_MAX;
            } else {
                d = -1;
                keyd[2 * (size_t)k] = best;
                keyi[2 * (size_t)k] = bidx;
                keyi[2 * (size_t)k + 1] = corner;
            }
        }
        done[k] = d;
        pend = d == -1;
    }
    wave_count(pend, &counter[3]);
}

// key order (cell phi, cell index, staggered phi, corner) of two nodes
__device__ __forceinline__ bool key_lt(double ca, int ia, double sa, int oa, double cb, int ib, double sb, int ob) {
    if (ca != cb) return ca < cb;
    if (ia != ib) return ia < ib;
    if (sa != sb) return sa < sb;
    return oa < ob;
}
__device__ __forceinline__ bool key_less(const double *keyd, const int *keyi, int a, int b) {
    return key_lt(keyd[2 * (size_t)a], keyi[2 * (size_t)a], keyd[2 * (size_t)a + 1], keyi[2 * (size_t)a + 1],
                  keyd[2 * (size_t)b], keyi[2 * (size_t)b], keyd[2 * (size_t)b + 1], keyi[2 * (size_t)b + 1]);
}

// A node uses the neighbours of its 5x5 window with a smaller staggered phi
// that are sources or precede it in key order (traverse.hh:133-145); the
// pending ones among the latter are its dependencies.
__global__ void __launch_bounds__(kThreads) xn_deps_kernel(BdConst c, const int *__restrict__ done,
                                                           const double *__restrict__ keyd,
                                                           const int *__restrict__ keyi, int *deg, int *counter,
                                                           int *ready) {
    const int k = blockIdx.x * kThreads + threadIdx.x, me = c.me, ne = c.ne;
    const bool pend = k < me * ne && done[k] == -1;
    int cnt = 0;
    if (pend) {
        const int ie = k % me, je = k / me;
        const double sp = keyd[2 * (size_t)k + 1];
        int ilo, ihi, jlo, jhi;
        window(ie, je, me, ne, 2, ilo, ihi, jlo, jhi);
        for (int dj = jlo; dj <= jhi; dj++)
            for (int di = ilo; di <= ihi; di++) {
                const int q = k + dj * me + di;
                cnt += (done[q] == -1 && keyd[2 * (size_t)q + 1] < sp && key_less(keyd, keyi, q, k)) ? 1 : 0;
            }
        deg[k] = cnt;
    }
    wave_push(pend && cnt == 0, k, ready, &counter[13]);
}

struct NodeProbe {
    int q, d, qi, qo;
    double qc, qs;
    bool in;
};

// the node walk (traverse.hh:119-174)
struct NodeWalk {
    BdConst c;
    BdNodes f;
    int *done, *deg, *counter;
    const double *keyd;
    const int *keyi;

    __device__ void fit(int wave, int k, int lane, TeamBuf &B) const {
        const int me = c.me, ne = c.ne, ie = k % me, je = k / me;
        const double sp = keyd[2 * (size_t)k + 1], kc = keyd[2 * (size_t)k];
        const int ki = keyi[2 * (size_t)k], ko = keyi[2 * (size_t)k + 1];
        int ilo, ihi, jlo, jhi;
        window(ie, je, me, ne, 2, ilo, ihi, jlo, jhi);
        const int wd = ihi - ilo + 1, nwin = wd * (jhi - jlo + 1);  // at most 25
        const int lc = lane < nwin ? lane : 0, di = ilo + lc % wd, dj = jlo + lc / wd, q = k + dj * me + di;
        const int d = done[q], qi = keyi[2 * (size_t)q], qo = keyi[2 * (size_t)q + 1];
        const double sq = keyd[2 * (size_t)q + 1], qc = keyd[2 * (size_t)q],
                     v[4] = {f.u[q], f.v[q], f.X[q], f.Y[q]};
        // used: a source, or set before k in key order (traverse.hh:133-145)
        const bool cand = lane < nwin && d != INT_MAX && sq < sp, prec = key_lt(qc, qi, sq, qo, kc, ki, sp, ko);
        const bool used = cand && (d == 0 || prec), bad = cand && d != 0 && prec && !(d > 0 && d < wave);
        stage_addends<false, 4>(B, lane, used, di, dj, sp - sq, v);
        wave_sync();
        const double acc = team_sum(B, lane, 18, nwin, 0.0);
        wave_sync();
        double out = 0;
        const bool ok = team_solve(acc, lane, out);
        if (__any(bad) && lane == 0) atomicAdd(&counter[5], 1);
        if (!ok) {  // singular fit (traverse.hh:153-164)
            if (lane == 0) atomicAdd(&counter[2], 1);
        } else if (lane >= 6 && lane < 18 && (lane - 6) % 3 == 0) {
            switch ((lane - 6) / 3) {
            case 0: f.u[k] = out; break;
            case 1: f.v[k] = out; break;
            case 2: f.X[k] = out; break;
            default: f.Y[k] = out;
            }
        }
        if (lane == 0) done[k] = wave;
    }

    // neighbour lane of the 5x5 square around k
    __device__ NodeProbe probe(int k, int lane, int) const {
        const int me = c.me, ne = c.ne, ie = k % me, je = k / me, dx = lane % 5 - 2, dy = lane / 5 - 2;
        const bool in = lane < 25 && (dx | dy) != 0 && ie + dx >= 0 && ie + dx < me && je + dy >= 0 && je + dy < ne;
        const int q = in ? k + dy * me + dx : k;
        return NodeProbe{q, done[q], keyi[2 * (size_t)q], keyi[2 * (size_t)q + 1], keyd[2 * (size_t)q],
                         keyd[2 * (size_t)q + 1], in};
    }

    // a pending node q depends on k if k lies in its window, has the smaller
    // staggered phi and precedes it in key order
    __device__ void release(int k, int lane, const NodeProbe &P, int *next, int *nnext) const {
        const double sp = keyd[2 * (size_t)k + 1], kc = keyd[2 * (size_t)k];
        const int ki = keyi[2 * (size_t)k], ko = keyi[2 * (size_t)k + 1];
        const bool dep = P.in && P.d == -1 && sp < P.qs && key_lt(kc, ki, sp, ko, P.qc, P.qi, P.qs, P.qo);
        const bool ready = dep && (atomicSub(&deg[P.q], 1) & 0xffff) == 1;
        wave_push(ready, P.q, next, nnext);
    }
};

constexpr int kWaveThreads = 256, kWaveTeams = kWaveThreads / 64, kWaveBlocks = 256, kMaxWaves = 1 << 20;

// One wave of a walk: the teams of the grid set the entries of this wave's
// list and release their dependants into the next one.  Lists alternate by
// wave parity; the counts sit in three slots, wave w reading slot w % 3,
// filling slot (w + 1) % 3 and clearing slot (w + 2) % 3 for wave w + 1 (the
// wave before read it, and the wave after fills it, each in a launch of its
// own).  A launch after the last wave finds a zero count and returns.
// blk of nblk: this workgroup's place among the wave's workgroups
template <class Walk>
__device__ __forceinline__ static void walk_wave(const Walk &W, int wave, int *lists, long span, int *counter,
                                                 TeamBuf *buf, int blk, int nblk) {
    int *slot = counter + 12;
    const int count = slot[wave % 3], lane = threadIdx.x & 63, team = threadIdx.x >> 6;
    if (blk == 0 && threadIdx.x == 0) {
        slot[(wave + 2) % 3] = 0;
        if (count > 0) {
            counter[8] = wave;
            counter[9] += count;
        }
    }
    if (count == 0) return;
    const int *cur = lists + (wave & 1) * span;
    int *nxt = lists + ((wave + 1) & 1) * span;
    for (int t = blk * kWaveTeams + team; t < count; t += nblk * kWaveTeams) {
        const int e = cur[t];
        const auto first = W.probe(e, lane, 0);
        W.fit(wave, e, lane, buf[team]);
        W.release(e, lane, first, nxt, slot + (wave + 1) % 3);
    }
}

__global__ void __launch_bounds__(kWaveThreads) xc_wave_kernel(BdConst c, const double *__restrict__ phi, BdCells f,
                                                               int *done, int *deg, int *counter, int *lists,
                                                               long span, int wave) {
    __shared__ TeamBuf buf[kWaveTeams];
    const int R = counter[10] > 3 ? counter[10] : 3;
    walk_wave(CellWalk{c, phi, f, done, deg, counter, R}, wave, lists, span, counter, buf, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kWaveThreads) xn_wave_kernel(BdConst c, BdNodes f, int *done,
                                                               const double *__restrict__ keyd,
                                                               const int *__restrict__ keyi, int *deg, int *counter,
                                                               int *lists, long span, int wave) {
    __shared__ TeamBuf buf[kWaveTeams];
    walk_wave(NodeWalk{c, f, done, deg, counter, keyd, keyi}, wave, lists, span, counter, buf, blockIdx.x, gridDim.x);
}

// a wave of the cell walk and one of the node walk in one launch (the two
// walks do not depend on each other): the first kWaveBlocks workgroups take
// the cells, the others the nodes
__global__ void __launch_bounds__(kWaveThreads) both_wave_kernel(BdConst c, const double *__restrict__ phi, BdCells fc,
                                                                 int *done_c, int *deg_c, int *counter_c,
                                                                 int *lists_c, BdNodes fn, int *done_n,
                                                                 const double *__restrict__ keyd,
                                                                 const int *__restrict__ keyi, int *deg_n,
                                                                 int *counter_n, int *lists_n, long span, int wave) {
    __shared__ TeamBuf buf[kWaveTeams];
    if (blockIdx.x < kWaveBlocks) {
        const int R = counter_c[10] > 3 ? counter_c[10] : 3;
        walk_wave(CellWalk{c, phi, fc, done_c, deg_c, counter_c, R}, wave, lists_c, span, counter_c, buf, blockIdx.x,
                  kWaveBlocks);
    } else {
        walk_wave(NodeWalk{c, fn, done_n, deg_n, counter_n, keyd, keyi}, wave, lists_n, span, counter_n, buf,
                  blockIdx.x - kWaveBlocks, gridDim.x - kWaveBlocks);
    }
}

// the count of wave w (zero once the walk has ended) for the host; with a
// second walk's counter, the sum of both
__global__ void wave_count_kernel(const int *counter, const int *counter2, int wave, volatile int *hflag) {
    if (threadIdx.x == 0) hflag[12] = counter[12 + wave % 3] + (counter2 ? counter2[12 + wave % 3] : 0);
}

// a walk's counts for the host: pending, order violations, waves, set
__global__ void walk_result_kernel(const int *counter, volatile int *hflag) {
    if (threadIdx.x == 0) {
        hflag[8] = counter[3];
        hflag[9] = counter[5];
        hflag[10] = counter[8];
        hflag[11] = counter[9];
    }
}

inline dim3 blocks(long n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

}  // namespace

int bd_ls_scan(const BdConst &c, const double *phi, int8_t *st, hipStream_t s) {
    hipLaunchKernelGGL(scan_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, phi, st);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_ls_newton_distance(const BdConst &c, const double *phi, const int8_t *st, double *out, int *fail,
                          hipStream_t s) {
    hipLaunchKernelGGL(newton_distance_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, phi, st, out,
                       fail);
    STZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(copy_interface_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, out, st,
                       const_cast<double *>(phi));
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_ls_interface_velocity(const BdConst &c, const double *phi, const int8_t *st, const BdNodes &nd,
                             double *vp, hipStream_t s) {
    hipLaunchKernelGGL(interface_velocity_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, phi, st, nd,
                       vp);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_ls_advance(const BdConst &c, double *phi, const double *vp, int8_t *st, double dt, hipStream_t s) {
    hipLaunchKernelGGL(advance_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, phi, vp, st, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <bool POS>
static int march_dir(const BdConst &c, double *phi, int8_t *st, double delta, unsigned long long *dmin,
                     int *counter, volatile int *hflag, int *rounds, hipStream_t s) {
    const dim3 g = blocks((long)c.m * c.n), b(kThreads);
    int *last = counter + 1;
    hipLaunchKernelGGL(march_seed_kernel<POS>, g, b, 0, s, c, phi, st, dmin, counter);
    STZ_HIP(hipGetLastError());
    const int batch = 4, max_rounds = 4096;
    const dim3 gm((unsigned)std::min<long>((long)g.x, kMarchBlocks));
    int r = 0;
    for (;;) {
        for (int q = 0; q < batch; q++) {
            hipLaunchKernelGGL(march_min_kernel<POS>, gm, b, 0, s, c, phi, st, dmin);
            hipLaunchKernelGGL(march_accept_kernel<POS>, gm, b, 0, s, c, phi, st, dmin, delta, counter);
            hipLaunchKernelGGL(march_update_kernel<POS>, g, b, 0, s, c, phi, st, dmin, counter, last);
        }
        STZ_HIP(hipGetLastError());
        r += batch;
        hipLaunchKernelGGL(flag_kernel, dim3(1), dim3(64), 0, s, last, hflag);
        STZ_HIP(hipStreamSynchronize(s));
        if (*hflag == 0) break;
        if (r >= max_rounds) {
            set_error("level-set march did not finish in %d rounds", max_rounds);
            return STZ_ERR_NOT_CONVERGED;
        }
    }
    // clear the flags of the last acceptances
    hipLaunchKernelGGL(march_accept_kernel<POS>, gm, b, 0, s, c, phi, st, dmin, 0.0, counter);
    STZ_HIP(hipGetLastError());
    *rounds += r;
    return STZ_OK;
}

int bd_ls_march(const BdConst &c, double *phi, int8_t *st, double delta, unsigned long long *dmin, int *counter,
                volatile int *hflag, int *rounds, hipStream_t s) {
    *rounds = 0;
    int e = march_dir<true>(c, phi, st, delta, dmin, counter, hflag, rounds, s);
    if (e) return e;
    return march_dir<false>(c, phi, st, delta, dmin, counter, hflag, rounds, s);
}

int bd_ls_finish(const BdConst &c, double *phi, double *vp, int8_t *st, int build, hipStream_t s) {
    hipLaunchKernelGGL(finish_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, phi, vp, st, build);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_ls_node_status(const BdConst &c, const double *phi, const int8_t *st, int8_t *nst, hipStream_t s) {
    hipLaunchKernelGGL(node_status_kernel, blocks((long)c.me * c.ne), dim3(kThreads), 0, s, c, phi, st, nst);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// Waves in batches of launches that double from 16 to 256; after each
// batch the host reads whether the walk has ended.
template <class Launch>
static int run_walk(Launch launch, const int *counter, const int *counter2, volatile int *hflag, hipStream_t s) {
    int w = 1, batch = 16;
    for (;;) {
        for (int q = 0; q < batch; q++, w++) launch(w);
        hipLaunchKernelGGL(wave_count_kernel, dim3(1), dim3(64), 0, s, counter, counter2, w, hflag);
        STZ_HIP(hipGetLastError());
        STZ_HIP(hipStreamSynchronize(s));
        if (hflag[12] == 0) break;
        if (w >= kMaxWaves) break;
        batch = batch < 256 ? 2 * batch : 256;
    }
    return STZ_OK;
}

// a finished walk's counts: entries set = entries pending, none fitted out
// of order
static int walk_result(const int *counter, volatile int *hflag, int *waves, const char *what, hipStream_t s) {
    hipLaunchKernelGGL(walk_result_kernel, dim3(1), dim3(64), 0, s, counter, hflag);
    STZ_HIP(hipGetLastError());
    STZ_HIP(hipStreamSynchronize(s));
    if (hflag[9] != 0 || hflag[11] != hflag[8]) {
        set_error("%s extrapolation: %d of %d entries set in %d waves, %d fitted out of order", what, hflag[11],
                  hflag[8], hflag[10], hflag[9]);
        return STZ_ERR_NOT_CONVERGED;
    }
    *waves = hflag[10];
    return STZ_OK;
}

// list: 3 max(cells, nodes) ints (dependency counts, two wave lists)
int bd_extrap_cells(const BdConst &c, const double *phi, const int8_t *st, const BdCells &f, int *done,
                    int *counter, int *list, volatile int *hflag, int *waves, hipStream_t s) {
    const long mn = (long)c.m * c.n, span = std::max(mn, (long)c.me * c.ne);
    int *lists = list + span;
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counter);
    // the first wave's list: lists + span (wave 1 is odd), its count slot 13
    hipLaunchKernelGGL(xc_prepare_kernel, blocks(mn), dim3(kThreads), 0, s, c, phi, st, done, list, counter,
                       lists + span);
    STZ_HIP(hipGetLastError());
    int e = run_walk(
        [&](int w) {
            hipLaunchKernelGGL(xc_wave_kernel, dim3(kWaveBlocks), dim3(kWaveThreads), 0, s, c, phi, f, done, list,
                               counter, lists, span, w);
        },
        counter, (const int *)nullptr, hflag, s);
    return e ? e : walk_result(counter, hflag, waves, "cell", s);
}

int bd_extrap_nodes(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdNodes &f,
                    int *done, double *key, int *counter, int *list, volatile int *hflag, int *waves,
                    hipStream_t s) {
    const long mne = (long)c.me * c.ne, span = std::max((long)c.m * c.n, mne);
    int *keyi = reinterpret_cast<int *>(key + 2 * mne), *lists = list + span;
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL(xn_prepare_kernel, blocks(mne), dim3(kThreads), 0, s, c, phi, st, nst, done, key, keyi,
                       counter);
    hipLaunchKernelGGL(xn_deps_kernel, blocks(mne), dim3(kThreads), 0, s, c, done, key, keyi, list, counter,
                       lists + span);
    STZ_HIP(hipGetLastError());
    int e = run_walk(
        [&](int w) {
            hipLaunchKernelGGL(xn_wave_kernel, dim3(kWaveBlocks), dim3(kWaveThreads), 0, s, c, f, done, key, keyi,
                               list, counter, lists, span, w);
        },
        counter, (const int *)nullptr, hflag, s);
    return e ? e : walk_result(counter, hflag, waves, "node", s);
}

// both walks, their waves sharing launches; cnt_n and list_n are the node
// walk's own counters (16 ints) and lists (3 max(cells, nodes) ints)
int bd_extrap_both(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdCells &fc,
                   const BdNodes &fn, int *done_c, int *done_n, double *key, int *cnt_c, int *cnt_n, int *list_c,
                   int *list_n, volatile int *hflag, int *waves_c, int *waves_n, hipStream_t s) {
    const long mn = (long)c.m * c.n, mne = (long)c.me * c.ne, span = std::max(mn, mne);
    int *keyi = reinterpret_cast<int *>(key + 2 * mne), *lc = list_c + span, *ln = list_n + span;
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, cnt_c);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, cnt_n);
    hipLaunchKernelGGL(xc_prepare_kernel, blocks(mn), dim3(kThreads), 0, s, c, phi, st, done_c, list_c, cnt_c,
                       lc + span);
    hipLaunchKernelGGL(xn_prepare_kernel, blocks(mne), dim3(kThreads), 0, s, c, phi, st, nst, done_n, key, keyi,
                       cnt_n);
    hipLaunchKernelGGL(xn_deps_kernel, blocks(mne), dim3(kThreads), 0, s, c, done_n, key, keyi, list_n, cnt_n,
                       ln + span);
    STZ_HIP(hipGetLastError());
    int e = run_walk(
        [&](int w) {
            hipLaunchKernelGGL(both_wave_kernel, dim3(2 * kWaveBlocks), dim3(kWaveThreads), 0, s, c, phi, fc, done_c,
                               list_c, cnt_c, lc, fn, done_n, key, keyi, list_n, cnt_n, ln, span, w);
        },
        cnt_c, cnt_n, hflag, s);
    if (e || (e = walk_result(cnt_c, hflag, waves_c, "cell", s))) return e;
    return walk_result(cnt_n, hflag, waves_n, "node", s);
}

}  // namespace stz
