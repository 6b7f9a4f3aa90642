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

template <bool POS>
__global__ void __launch_bounds__(kThreads) march_min_kernel(BdConst c, const double *__restrict__ phi,
                                                             const int8_t *__restrict__ st,
                                                             unsigned long long *dmin) {
    using Z = Dir<POS>;
    __shared__ unsigned long long red[kThreads / 64];
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    unsigned long long k = ~0ull;
    if (ij < c.m * c.n && code(st[ij]) == Z::trial) k = okey(POS ? phi[ij] : -phi[ij]);
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
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    int8_t s = st[ij];
    if (s & kNew) {
        st[ij] = s & 15;
        return;
    }
    if (code(s) != Z::trial || *dmin == ~0ull) return;
    const double kmin = okey_inv(*dmin), key = POS ? phi[ij] : -phi[ij];
    if (key < kmin + delta && key < c.phi_lim) {
        st[ij] = Z::acc | kNew;
        atomicAdd(counter, 1);
    }
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
// extrapolation of the cell fields (traverse.hh:176-240, extrap.hh:49-72)
// done[]: 0 source (inside band), -1 pending (outside band), INT_MAX not in
// the band; a finished cell holds the wave that computed it
// ---------------------------------------------------------------------------
__global__ void zero_counters_kernel(int *counter) {
    if (threadIdx.x < 4) counter[threadIdx.x] = 0;
}

// keep the still pending entries of a list (counter[3] = new length)
__global__ void __launch_bounds__(kThreads) compact_kernel(const int *__restrict__ done,
                                                           const int *__restrict__ list, int count,
                                                           int *__restrict__ out, int *counter) {
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= count) return;
    const int k = list[q];
    if (done[k] == -1) out[atomicAdd(&counter[3], 1)] = k;
}

__global__ void zero_one_kernel(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// the pending cells are also listed (counter[3] entries of list) so that the
// waves run over them only
__global__ void __launch_bounds__(kThreads) xc_prepare_kernel(BdConst c, const int8_t *__restrict__ st, int *done,
                                                              int *counter, int *list) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    const int s = code(st[ij]);
    const int d = (s >= 2 && s <= 3) ? 0 : ((s >= 4 && s <= 5) ? -1 : INT_MAX);
    done[ij] = d;
    if (d == -1) list[atomicAdd(&counter[3], 1)] = ij;
}

// one wave over the listed pending cells; counter[2] counts failed fits
__global__ void __launch_bounds__(kThreads) xc_wave_kernel(BdConst c, const double *__restrict__ phi, BdCells f,
                                                           int *done, int wave, int *counter,
                                                           const int *__restrict__ list, int count) {
    const int q = blockIdx.x * kThreads + threadIdx.x, m = c.m, n = c.n;
    if (q >= count) return;
    const int ij = list[q];
    if (done[ij] != -1) return;
    const int i = ij % m, j = ij / m;
    const double ph = phi[ij];
    double s, sx, sy, sxx, sxy, syy, det = 0, p[15];
    int r;
    for (r = 3; r <= 7; r++) {
        const int ilo = i <= r ? -i : -r, ihi = i + r >= m ? m - 1 - i : r, jlo = j <= r ? -j : -r,
                  jhi = j + r >= n ? n - 1 - j : r;
        for (int dj = jlo; dj <= jhi; dj++)
            for (int di = ilo; di <= ihi; di++) {
                const int k = ij + dj * m + di, d = done[k];
                if (d == INT_MAX || !(phi[k] < ph)) continue;
                if (d < 0 || d >= wave) return;
            }
        s = sx = sy = sxx = sxy = syy = 0;
        for (int q = 0; q < 15; q++) p[q] = 0;
        for (int dj = jlo; dj <= jhi; dj++)
            for (int di = ilo; di <= ihi; di++) {
                const int k = ij + dj * m + di;
                if (done[k] == INT_MAX) continue;
                double w = phi[k];
                if (!(w < ph)) continue;
                w = ph - w;
                s += w;
                sx += di * w;
                sy += dj * w;
                sxx += di * di * w;
                sxy += di * dj * w;
                syy += dj * dj * w;
                const double x = di, y = dj;
                const double v[5] = {f.p[k], f.q[k], f.s[k], f.tau[k], f.chi[k]};
                for (int q = 0; q < 5; q++) {
                    p[3 * q] += w * v[q];
                    p[3 * q + 1] += w * x * v[q];
                    p[3 * q + 2] += w * y * v[q];
                }
            }
        det = s * (sxy * sxy - sxx * syy) + sx * sx * syy - 2 * sx * sxy * sy + sy * sy * sxx;
        if (!(fabs(det) < 1e-10)) break;
    }
    if (r > 7) {
        atomicAdd(&counter[2], 1);
        done[ij] = wave;
        return;
    }
    det = 1 / det;
    const double al = (sxy * sxy - sxx * syy) * det, be = (sx * syy - sxy * sy) * det,
                 ga = (sxx * sy - sxy * sx) * det;
    f.p[ij] = p[0] * al + p[1] * be + p[2] * ga;
    f.q[ij] = p[3] * al + p[4] * be + p[5] * ga;
    f.s[ij] = p[6] * al + p[7] * be + p[8] * ga;
    f.tau[ij] = p[9] * al + p[10] * be + p[11] * ga;
    f.chi[ij] = p[12] * al + p[13] * be + p[14] * ga;
    done[ij] = wave;
}

// ---------------------------------------------------------------------------
// extrapolation of the node fields (traverse.hh:62-174, extrap.hh:26-47)
// ---------------------------------------------------------------------------

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
struct NodeKey {
    double cphi, sphi;
    int cidx, corner;
};

__global__ void __launch_bounds__(kThreads) xn_prepare_kernel(BdConst c, const double *__restrict__ phi,
                                                              const int8_t *__restrict__ st,
                                                              const int8_t *__restrict__ nst, int *done,
                                                              double *keyd, int *keyi, int *counter, int *list) {
    const int k = blockIdx.x * kThreads + threadIdx.x, mne = c.me * c.ne;
    if (k >= mne) return;
    const int ie = k % c.me, je = k / c.me, m = c.m, n = c.n;
    const double spt = -1.2 * sqrt(c.dx * c.dx + c.dy * c.dy);
    keyd[2 * (size_t)k + 1] = staggered_phi(c, phi, ie, je);
    if (nst[k] == 0) {
        done[k] = 0;
        return;
    }
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
        done[k] = INT_MAX;
        return;
    }
    done[k] = -1;
    list[atomicAdd(&counter[3], 1)] = k;
    keyd[2 * (size_t)k] = best;
    keyi[2 * (size_t)k] = bidx;
    keyi[2 * (size_t)k + 1] = corner;
}

__device__ __forceinline__ bool key_less(const double *keyd, const int *keyi, int a, int b) {
    const double ca = keyd[2 * (size_t)a], cb = keyd[2 * (size_t)b];
    if (ca != cb) return ca < cb;
    const int ia = keyi[2 * (size_t)a], ib = keyi[2 * (size_t)b];
    if (ia != ib) return ia < ib;
    const double sa = keyd[2 * (size_t)a + 1], sb = keyd[2 * (size_t)b + 1];
    if (sa != sb) return sa < sb;
    return keyi[2 * (size_t)a + 1] < keyi[2 * (size_t)b + 1];
}

__global__ void __launch_bounds__(kThreads) xn_wave_kernel(BdConst c, BdNodes f, int *done,
                                                           const double *__restrict__ keyd,
                                                           const int *__restrict__ keyi, int wave, int *counter,
                                                           const int *__restrict__ list, int count) {
    const int q0 = blockIdx.x * kThreads + threadIdx.x, me = c.me, ne = c.ne;
    if (q0 >= count) return;
    const int k = list[q0];
    if (done[k] != -1) return;
    const int ie = k % me, je = k / me;
    const double sp = keyd[2 * (size_t)k + 1];
    const int r = 2;
    const int ilo = ie <= r ? -ie : -r, ihi = ie + r >= me ? me - 1 - ie : r, jlo = je <= r ? -je : -r,
              jhi = je + r >= ne ? ne - 1 - je : r;
    // a neighbour is used if it is a source or set before this node, and has
    // a smaller staggered phi (traverse.hh:133-145)
    auto used = [&](int q) {
        const int d = done[q];
        if (d == INT_MAX) return 0;
        if (!(keyd[2 * (size_t)q + 1] < sp)) return 0;
        if (d == 0) return 1;
        if (!key_less(keyd, keyi, q, k)) return 0;
        return (d > 0 && d < wave) ? 1 : -1;
    };
    for (int dj = jlo; dj <= jhi; dj++)
        for (int di = ilo; di <= ihi; di++)
            if (used(k + dj * me + di) < 0) return;
    double s = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, p[12];
    for (int q = 0; q < 12; q++) p[q] = 0;
    for (int dj = jlo; dj <= jhi; dj++)
        for (int di = ilo; di <= ihi; di++) {
            const int q = k + dj * me + di;
            if (used(q) != 1) continue;
            double w = keyd[2 * (size_t)q + 1];
            w = sp - w;
            s += w;
            sx += w * di;
            sy += w * dj;
            sxx += w * di * di;
            sxy += w * di * dj;
            syy += w * dj * dj;
            const double x = di, y = dj;
            const double v[4] = {f.u[q], f.v[q], f.X[q], f.Y[q]};
            for (int t = 0; t < 4; t++) {
                p[3 * t] += w * v[t];
                p[3 * t + 1] += w * x * v[t];
                p[3 * t + 2] += w * y * v[t];
            }
        }
    double det = s * (sxy * sxy - sxx * syy) + sx * sx * syy - 2 * sx * sxy * sy + sy * sy * sxx;
    if (fabs(det) < 1e-10) {
        atomicAdd(&counter[2], 1);
        done[k] = wave;
        return;
    }
    det = 1 / det;
    const double al = (sxy * sxy - sxx * syy) * det, be = (sx * syy - sxy * sy) * det,
                 ga = (sxx * sy - sxy * sx) * det;
    f.u[k] = p[0] * al + p[1] * be + p[2] * ga;
    f.v[k] = p[3] * al + p[4] * be + p[5] * ga;
    f.X[k] = p[6] * al + p[7] * be + p[8] * ga;
    f.Y[k] = p[9] * al + p[10] * be + p[11] * ga;
    done[k] = wave;
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
    int r = 0;
    for (;;) {
        for (int q = 0; q < batch; q++) {
            hipLaunchKernelGGL(march_min_kernel<POS>, g, b, 0, s, c, phi, st, dmin);
            hipLaunchKernelGGL(march_accept_kernel<POS>, g, b, 0, s, c, phi, st, dmin, delta, counter);
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
    hipLaunchKernelGGL(march_accept_kernel<POS>, g, b, 0, s, c, phi, st, dmin, 0.0, counter);
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

// Waves in batches that double from 4 to 64 launches; after each batch the
// list keeps only the cells still pending (so late waves, when a few long
// dependency chains remain, touch a few cells).  list has room for twice
// the longest list.
template <class Launch>
static int run_waves(Launch launch, const int *done, int *counter, int *list, int count, volatile int *hflag,
                     int *waves, const char *what, hipStream_t s) {
    const int max_waves = 1 << 16;
    int w = 1, batch = 4;
    int *cur = list, *nxt = list + count;
    for (;;) {
        for (int q = 0; q < batch; q++, w++) launch(w, cur, count);
        hipLaunchKernelGGL(zero_one_kernel, dim3(1), dim3(64), 0, s, counter + 3);
        hipLaunchKernelGGL(compact_kernel, blocks(count), dim3(kThreads), 0, s, done, cur, count, nxt, counter);
        hipLaunchKernelGGL(flag_kernel, dim3(1), dim3(64), 0, s, counter + 3, hflag);
        STZ_HIP(hipGetLastError());
        STZ_HIP(hipStreamSynchronize(s));
        count = *hflag;
        if (count == 0) break;
        if (w >= max_waves) {
            set_error("%s extrapolation did not finish in %d waves", what, max_waves);
            return STZ_ERR_NOT_CONVERGED;
        }
        std::swap(cur, nxt);
        batch = batch < 64 ? 2 * batch : 64;
    }
    *waves = w - 1;
    return STZ_OK;
}

static int pending_count(const int *counter, volatile int *hflag, hipStream_t s, int *count) {
    hipLaunchKernelGGL(flag_kernel, dim3(1), dim3(64), 0, s, counter + 3, hflag);
    STZ_HIP(hipStreamSynchronize(s));
    *count = *hflag;
    return STZ_OK;
}

// counter: [0],[1] pending counts by wave parity, [2] failed fits, [3] list length
int bd_extrap_cells(const BdConst &c, const double *phi, const int8_t *st, const BdCells &f, int *done,
                    int *counter, int *list, volatile int *hflag, int *waves, hipStream_t s) {
    const dim3 g = blocks((long)c.m * c.n), b(kThreads);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL(xc_prepare_kernel, g, b, 0, s, c, st, done, counter, list);
    STZ_HIP(hipGetLastError());
    int count = 0, e = pending_count(counter, hflag, s, &count);
    if (e) return e;
    *waves = 0;
    if (count == 0) return STZ_OK;
    return run_waves(
        [&](int w, const int *l, int n) {
            hipLaunchKernelGGL(xc_wave_kernel, blocks(n), b, 0, s, c, phi, f, done, w, counter, l, n);
        },
        done, counter, list, count, hflag, waves, "cell", s);
}

int bd_extrap_nodes(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdNodes &f,
                    int *done, double *key, int *counter, int *list, volatile int *hflag, int *waves,
                    hipStream_t s) {
    const long mne = (long)c.me * c.ne;
    const dim3 g = blocks(mne), b(kThreads);
    int *keyi = reinterpret_cast<int *>(key + 2 * mne);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL(xn_prepare_kernel, g, b, 0, s, c, phi, st, nst, done, key, keyi, counter, list);
    STZ_HIP(hipGetLastError());
    int count = 0, e = pending_count(counter, hflag, s, &count);
    if (e) return e;
    *waves = 0;
    if (count == 0) return STZ_OK;
    return run_waves(
        [&](int w, const int *l, int n) {
            hipLaunchKernelGGL(xn_wave_kernel, blocks(n), b, 0, s, c, f, done, key, keyi, w, counter, l, n);
        },
        done, counter, list, count, hflag, waves, "node", s);
}

}  // namespace stz
