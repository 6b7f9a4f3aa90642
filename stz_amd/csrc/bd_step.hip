// bd_step.hip -- gfx950 kernels of the free-boundary step outside the level
// set: the explicit reference-map / stress / chi update
// (bd_sim::advection_step, bd_sim.cc:374-469, chi_diffusion :472-505), the
// operator rows and sources of the projection (qs_multi_bd::init,
// qs_multi_bd.cc:26-196), the velocity / stress correction
// (bd_sim.cc:510-545), the box boundary conditions (bd_sim.cc:732-761) and
// the initial geometry and fields of init_bar (bd_sim.cc:63-123).
//
// One thread per cell or node, 64 lanes of a wavefront along a row.  The
// reference updates c_field slots in place after computing change slots; here
// the update reads one plane set and writes the other, which leaves the same
// state.  Expressions keep the reference's evaluation order and FMA
// contraction is off, so results agree bit for bit away from the
// transcendental functions of the STZ law.
#include "bd.h"
#include "stz_law.h"

#pragma clang fp contract(off)

namespace stz {

namespace {

constexpr int kThreads = 256;
inline dim3 blocks(long n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

__device__ __forceinline__ double eno2(double p0, double p1, double p2, double p3) {
    return fabs(p0 - 2 * p1 + p2) > fabs(p1 - 2 * p2 + p3) ? 3 * p1 - 4 * p2 + p3 : p0 - p2;
}

// five staggered values of one cell
struct Stag {
    double p, q, s, tau, chi;
};
__device__ __forceinline__ Stag load_cell(const BdCells &f, int k) {
    return Stag{f.p[k], f.q[k], f.s[k], f.tau[k], f.chi[k]};
}
// st_one_sided / st_eno2 (bd_sim.cc:814-832)
__device__ __forceinline__ Stag st_one_sided(double hs, const Stag &a, const Stag &b) {
    hs *= 2;
    return Stag{hs * (a.p - b.p), hs * (a.q - b.q), hs * (a.s - b.s), hs * (a.tau - b.tau), hs * (a.chi - b.chi)};
}
__device__ __forceinline__ Stag st_eno2(double hs, const Stag &a, const Stag &b, const Stag &c, const Stag &d) {
    return Stag{hs * eno2(a.p, b.p, c.p, d.p), hs * eno2(a.q, b.q, c.q, d.q), hs * eno2(a.s, b.s, c.s, d.s),
                hs * eno2(a.tau, b.tau, c.tau, d.tau), hs * eno2(a.chi, b.chi, c.chi, d.chi)};
}
// bc_deriv_proj ghost (bd_sim.cc:587-600): 2 f - f(mirror)
__device__ __forceinline__ Stag ghost(const Stag &f, const Stag &b) {
    return Stag{2 * f.p - b.p, 2 * f.q - b.q, 2 * f.s - b.s, 2 * f.tau - b.tau, 2 * f.chi - b.chi};
}

// reference map at the interior nodes (bd_sim.cc:387-409, update :460-468).
// The y-direction status tests read cc[k -+ 2m] (the node two rows away and
// two columns across), as the reference's ccp[-2*m] on the node array does.
__global__ void __launch_bounds__(kThreads) advect_nodes_kernel(BdConst c, const int8_t *__restrict__ nst,
                                                                BdNodes o, BdNodes w, double dt) {
    const int k = blockIdx.x * kThreads + threadIdx.x, me = c.me, m = c.m, n = c.n;
    if (k >= me * c.ne) return;
    const int i = k % me, j = k / me;
    double X = o.X[k], Y = o.Y[k];
    if (j > 0 && i > 0 && i < m && j < n && nst[k] < 4) {
        const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
        const double uc = o.u[k], vc = o.v[k];
        double Xx, Yx, Xy, Yy;
        if (uc > 0) {
            if (i > 1 && nst[k - 2] < 4) {
                Xx = hx * eno2(o.X[k + 1], X, o.X[k - 1], o.X[k - 2]);
                Yx = hx * eno2(o.Y[k + 1], Y, o.Y[k - 1], o.Y[k - 2]);
            } else {
                const double hs = hx * 2;
                Xx = hs * (X - o.X[k - 1]);
                Yx = hs * (Y - o.Y[k - 1]);
            }
        } else {
            if (i < m - 1 && nst[k + 2] < 4) {
                Xx = -hx * eno2(o.X[k - 1], X, o.X[k + 1], o.X[k + 2]);
                Yx = -hx * eno2(o.Y[k - 1], Y, o.Y[k + 1], o.Y[k + 2]);
            } else {
                const double hs = -hx * 2;
                Xx = hs * (X - o.X[k + 1]);
                Yx = hs * (Y - o.Y[k + 1]);
            }
        }
        if (vc > 0) {
            if (j > 1 && nst[k - 2 * m] < 4) {
                Xy = hy * eno2(o.X[k + me], X, o.X[k - me], o.X[k - 2 * me]);
                Yy = hy * eno2(o.Y[k + me], Y, o.Y[k - me], o.Y[k - 2 * me]);
            } else {
                const double hs = hy * 2;
                Xy = hs * (X - o.X[k - me]);
                Yy = hs * (Y - o.Y[k - me]);
            }
        } else {
            if (j < n - 1 && nst[k + 2 * m] < 4) {
                Xy = -hy * eno2(o.X[k - me], X, o.X[k + me], o.X[k + 2 * me]);
                Yy = -hy * eno2(o.Y[k - me], Y, o.Y[k + me], o.Y[k + 2 * me]);
            } else {
                const double hs = -hy * 2;
                Xy = hs * (X - o.X[k + me]);
                Yy = hs * (Y - o.Y[k + me]);
            }
        }
        X += -uc * Xx - vc * Xy;
        Y += -uc * Yx - vc * Yy;
    }
    w.X[k] = X;
    w.Y[k] = Y;
    w.u[k] = o.u[k];
    w.v[k] = o.v[k];
}

// stress and chi at the cells inside the body (bd_sim.cc:411-458, update
// :460-468).  Where a ghost replaces a missing left (right) neighbour and the
// same cell also needs a bottom (top) ghost, the reference stores both in one
// slot, so its filter term reads the y ghost in place of the x one; kept.
__global__ void __launch_bounds__(kThreads) advect_cells_kernel(BdConst c, const int8_t *__restrict__ st,
                                                                BdCells o, BdNodes on, BdCells w, double dt) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, m = c.m, n = c.n, me = c.me;
    if (ij >= m * n) return;
    const int i = ij % m, j = ij / m;
    const Stag f = load_cell(o, ij);
    if (st[ij] >= 4) {
        w.p[ij] = f.p;
        w.q[ij] = f.q;
        w.s[ij] = f.s;
        w.tau[ij] = f.tau;
        w.chi[ij] = f.chi;
        w.ddev[ij] = o.ddev[ij];
        return;
    }
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    const double hxxv = c.filter_stress, hyyv = c.filter_stress, hcv = 2 * (hxxv + hyyv);
    const int k = i + me * j;  // node (i,j), lower-left corner of the cell
    const double u0 = on.u[k], u1 = on.u[k + 1], u2 = on.u[k + me], u3 = on.u[k + me + 1];
    const double v0 = on.v[k], v1 = on.v[k + 1], v2 = on.v[k + me], v3 = on.v[k + me + 1];
    const double uc = 0.25 * (u0 + u1 + u2 + u3), vc = 0.25 * (v0 + v1 + v2 + v3),
                 two_omega = hx * (-v0 + v1 - v2 + v3) + hy * (-u0 - u1 + u2 + u3);
    // x neighbours, ghosts at the box edges
    const bool dd = i > 0, uu = i < m - 1;
    const Stag gl = dd ? Stag{} : ghost(f, load_cell(o, ij + 1));
    const Stag gr = uu ? Stag{} : ghost(f, load_cell(o, ij - 1));
    const Stag fl = dd ? load_cell(o, ij - 1) : gl, fr = uu ? load_cell(o, ij + 1) : gr;
    Stag f1;
    if (uc > 0) f1 = (dd && i > 1 && st[ij - 2] < 4) ? st_eno2(hx, fr, f, fl, load_cell(o, ij - 2))
                                                      : st_one_sided(hx, f, fl);
    else f1 = (uu && i < m - 2 && st[ij + 2] < 4) ? st_eno2(-hx, fl, f, fr, load_cell(o, ij + 2))
                                                   : st_one_sided(-hx, f, fr);
    const bool yd = j > 0, yu = j < n - 1;
    const Stag gd = yd ? Stag{} : ghost(f, load_cell(o, ij + m));
    const Stag gu = yu ? Stag{} : ghost(f, load_cell(o, ij - m));
    const Stag fd = yd ? load_cell(o, ij - m) : gd, fu = yu ? load_cell(o, ij + m) : gu;
    Stag f2;
    if (vc > 0) f2 = (yd && j > 1 && st[ij - 2 * m] < 4) ? st_eno2(hy, fu, f, fd, load_cell(o, ij - 2 * m))
                                                          : st_one_sided(hy, f, fd);
    else f2 = (yu && j < n - 2 && st[ij + 2 * m] < 4) ? st_eno2(-hy, fd, f, fu, load_cell(o, ij + 2 * m))
                                                       : st_one_sided(-hy, f, fu);
    // the shared ghost slots (bd_sim.cc:424-429)
    const Stag flf = (!dd && !yd) ? gd : fl, frf = (!uu && !yu) ? gu : fr;
    const double pfil = hxxv * (frf.p + flf.p) + hyyv * (fu.p + fd.p) - hcv * f.p,
                 qfil = hxxv * (frf.q + flf.q) + hyyv * (fu.q + fd.q) - hcv * f.q,
                 sfil = hxxv * (frf.s + flf.s) + hyyv * (fu.s + fd.s) - hcv * f.s,
                 taufil = hxxv * (frf.tau + flf.tau) + hyyv * (fu.tau + fd.tau) - hcv * f.tau;
    double dtmp = sqrt(3 * f.q * f.q + f.s * f.s + f.tau * f.tau);
    double cchi = f.chi, ddev;
    int trips;
    bool last_yield;
    dtmp = plastic_term(c, dtmp, cchi, ddev, dt, trips, last_yield);
    cchi -= uc * f1.chi + vc * f2.chi;
    const double cp = -uc * f1.p - vc * f2.p + pfil, cq = -uc * f1.q - vc * f2.q - dtmp * f.q + qfil,
                 cs = -uc * f1.s - vc * f2.s - two_omega * f.tau - dtmp * f.s + sfil,
                 ctau = -uc * f1.tau - vc * f2.tau + two_omega * f.s - dtmp * f.tau + taufil;
    w.p[ij] = f.p + cp;
    w.q[ij] = f.q + cq;
    w.s[ij] = f.s + cs;
    w.tau[ij] = f.tau + ctau;
    w.chi[ij] = cchi;
    w.ddev[ij] = ddev;
}

// bd_sim::chi_diffusion (bd_sim.cc:472-505): old chi, new ddev, added to the
// new chi before it replaces the old one
__global__ void __launch_bounds__(kThreads) chi_diffusion_kernel(BdConst c, const int8_t *__restrict__ st,
                                                                 const double *__restrict__ chi_old,
                                                                 const double *__restrict__ ddev,
                                                                 double *__restrict__ chi_new) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, m = c.m, n = c.n;
    if (ij >= m * n || st[ij] >= 4) return;
    const int i = ij % m, j = ij / m;
    const double dfac = 0.25 * c.chi_len * c.chi_len * c.mu_inv, dfacx = dfac * c.xsp * c.xsp,
                 dfacy = dfac * c.ysp * c.ysp;
    double cchi = chi_new[ij];
    const double fc = chi_old[ij], fd0 = ddev[ij];
    if (i > 0 && i < m - 1 && st[ij - 1] < 4 && st[ij + 1] < 4) {
        const double Dl = fd0 + ddev[ij - 1], Dr = fd0 + ddev[ij + 1];
        cchi += dfacx * (Dr * chi_old[ij + 1] - (Dr + Dl) * fc + Dl * chi_old[ij - 1]);
    }
    if (j > 0 && j < n - 1 && st[ij - m] < 4 && st[ij + m] < 4) {
        const double Dd = fd0 + ddev[ij - m], Du = fd0 + ddev[ij + m];
        cchi += dfacy * (Du * chi_old[ij + m] - (Du + Dd) * fc + Dd * chi_old[ij - m]);
    }
    chi_new[ij] = cchi;
}

// bd_sim::set_boundaries (bd_sim.cc:732-761), with the edge grips: columns 0
// and m move at u_left / u_right and carry the reference map of the
// material held there (shift = -integral of the grip velocity)
__global__ void __launch_bounds__(kThreads) boundaries_kernel(BdConst c, BdNodes f) {
    const int t = blockIdx.x * kThreads + threadIdx.x, me = c.me, n = c.n, m = c.m;
    if (t < me) {  // bottom and top rows
        const int i = t;
        const double ug = i == 0 ? c.u_left : (i == m ? c.u_right : 0.0);
        const double xs = i == 0 ? c.x_left : (i == m ? c.x_right : 0.0);
        f.u[i] = 0 + ug;
        f.v[i] = -0.000;
        f.X[i] = c.ax + i * c.dx + xs;
        f.Y[i] = c.ay;
        const int k = i + me * n;
        f.u[k] = 0 + ug;
        f.v[k] = 0;
        f.X[k] = c.ax + i * c.dx + xs;
        f.Y[k] = c.by;
    } else if (t < me + n - 1) {  // left and right columns, rows 1..n-1
        const int j = t - me + 1, kl = me * j, kr = kl + m;
        f.u[kl] = 0 + c.u_left;
        f.v[kl] = -0.000;
        f.X[kl] = c.ax + c.x_left;
        f.Y[kl] = c.ay + j * c.dy;
        f.u[kr] = 0 + c.u_right;
        f.v[kr] = -0.000;
        f.X[kr] = c.bx + c.x_right;
        f.Y[kr] = c.ay + j * c.dy;
    }
}

// ---------------------------------------------------------------------------
// qs_multi_bd::init (qs_multi_bd.cc:26-196)
// ---------------------------------------------------------------------------
struct M2 {
    double a, b, c, d;
};
__device__ __forceinline__ M2 mz() { return M2{0, 0, 0, 0}; }
__device__ __forceinline__ M2 madd(const M2 &x, const M2 &y) { return M2{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
__device__ __forceinline__ M2 msub(const M2 &x, const M2 &y) { return M2{x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
__device__ __forceinline__ M2 smul(double e, const M2 &f) { return M2{e * f.a, e * f.b, e * f.c, e * f.d}; }
__device__ __forceinline__ M2 mneg(const M2 &f) { return M2{-f.a, -f.b, -f.c, -f.d}; }
__device__ __forceinline__ M2 mmul(const M2 &x, const M2 &e) {
    return M2{x.a * e.a + x.b * e.c, x.a * e.b + x.b * e.d, x.c * e.a + x.d * e.c, x.c * e.b + x.d * e.d};
}
// operator/(double, mat) of shared/mat.hh:89-92
__device__ __forceinline__ M2 sdiv(double e, const M2 &f) {
    const double idet = e / (f.a * f.d - f.b * f.c);
    return M2{f.d * idet, -f.b * idet, -f.c * idet, f.a * idet};
}

// qs_multi_bd.cc:158-178
__device__ static void square1(double z, const M2 &h, const M2 &v, M2 &a0, M2 &a1, M2 &a2, M2 &a3, bool pos) {
    const double fz = 1 - z;
    a0 = smul(-fz, madd(h, v));
    a1 = msub(smul(fz, h), smul(z, v));
    a2 = msub(smul(fz, v), smul(z, h));
    a3 = smul(z, madd(h, v));
    const M2 co = smul(8 * z * fz, pos ? a1 : a2);
    a1 = msub(a1, co);
    a2 = msub(a2, co);
    a0 = madd(a0, co);
    a3 = madd(a3, co);
}
__device__ static void square2(double z, const M2 &h, const M2 &v, M2 &a0, M2 &a1, M2 &a2, M2 &a3, bool pos) {
    const double fz = 1 - z;
    a0 = msub(smul(-fz, h), smul(z, v));
    a1 = smul(fz, msub(h, v));
    a2 = smul(z, msub(v, h));
    a3 = madd(smul(fz, v), smul(z, h));
    const M2 co = smul(8 * z * fz, pos ? a3 : a0);
    a0 = msub(a0, co);
    a3 = msub(a3, co);
    a1 = madd(a1, co);
    a2 = madd(a2, co);
}

// qs_multi_bd::contrib (qs_multi_bd.cc:180-184): traction of cell (i,j)
__device__ __forceinline__ void contrib(const BdCells &f, int m, int i, int j, double nx, double ny, double &x,
                                        double &y) {
    const int k = i + m * j;
    const double t0 = f.p[k] + f.q[k], s = f.s[k], tau = f.tau[k];
    x = (-t0 + s) * nx + tau * ny;
    y = tau * nx - (t0 + s) * ny;
}

__global__ void __launch_bounds__(kThreads) operator_kernel(BdConst c, BdOp op, const double *__restrict__ phi,
                                                            const int8_t *__restrict__ st,
                                                            const int8_t *__restrict__ nst, BdCells f, BdNodes nd,
                                                            const double *__restrict__ force, double dt,
                                                            int *__restrict__ rowt, double *__restrict__ src,
                                                            double *__restrict__ fmats, int fcap, int *nfree) {
    const int M = c.me, N = c.ne, ml = c.m;
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= M * N) return;
    const int i = ij % M, j = ij / M;
    const double idt = 1 / dt;
    const double xfac = 0.5 / dt * c.xsp, yfac = 0.5 / dt * c.ysp, xbc = op.bcs * op.cxx, ybc = op.bcs * op.cyy;
    if (i == 0 || i == M - 1 || j == 0 || j == N - 1) {
        rowt[ij] = -2;
        src[2 * ij] = xbc * nd.u[ij];
        src[2 * ij + 1] = ybc * nd.v[ij];
        return;
    }
    const int ijl = i + ml * j;  // cell (i,j)
    if (nst[ij] < 4) {
        rowt[ij] = -1;
        const int c0 = ijl - ml - 1, c1 = ijl - ml, c2 = ijl - 1, c3 = ijl;
        const double p0 = f.p[c0], p1 = f.p[c1], p2 = f.p[c2], p3 = f.p[c3];
        const double q0 = f.q[c0], q1 = f.q[c1], q2 = f.q[c2], q3 = f.q[c3];
        const double s0 = f.s[c0], s1 = f.s[c1], s2 = f.s[c2], s3 = f.s[c3];
        const double t0 = f.tau[c0], t1 = f.tau[c1], t2 = f.tau[c2], t3 = f.tau[c3];
        double bx = (-p0 + p1 - p2 + p3 - q0 + q1 - q2 + q3 + s0 - s1 + s2 - s3) * xfac + (t0 + t1 - t2 - t3) * yfac;
        double by = (t0 - t1 + t2 - t3) * xfac + (-p0 - p1 + p2 + p3 - q0 - q1 + q2 + q3 - s0 - s1 + s2 + s3) * yfac;
        if (force) {
            bx += idt * force[2 * ij];
            by += idt * force[2 * ij + 1];
        }
        src[2 * ij] = bx;
        src[2 * ij + 1] = by;
        return;
    }
    // Neumann-type free boundary (qs_multi_bd.cc:74-149)
    const bool near = st[ijl] < 4 || st[ijl - 1] < 4 || st[ijl - ml] < 4 || st[ijl - M] < 4 ||
                      nst[ij - M - 1] < 4 || nst[ij - M] < 4 || nst[ij - ml] < 4 || nst[ij - 1] < 4 ||
                      nst[ij + 1] < 4 || nst[ij + ml] < 4 || nst[ij + M] < 4 || nst[ij + M + 1] < 4;
    if (!near) {
        rowt[ij] = -3;
        src[2 * ij] = 0;
        src[2 * ij + 1] = 0;
        return;
    }
    const double *phip = phi + ijl;
    double nx = c.xsp * (*phip - phip[-1] + phip[-ml] - phip[-M]),
           ny = c.ysp * (*phip + phip[-1] - phip[-ml] - phip[-M]), nsq = nx * nx + ny * ny;
    if (nsq > 1e-50) {
        nsq = 1 / sqrt(nsq);
        nx *= nsq;
        ny *= nsq;
    }
    const double Kp = c.K + c.mu / 3.0, Kmu = Kp + c.mu, mu = c.mu;
    const double fx = fabs(nx), fy = fabs(ny);
    double z, zz;
    int k;
    const M2 h = smul(c.xsp, M2{-Kmu * nx, -mu * ny, (mu - Kp) * ny, -mu * nx}),
             v = smul(c.ysp, M2{-mu * ny, (mu - Kp) * nx, -mu * nx, -Kmu * ny});
    M2 dl = mz(), dc = mz(), dr = mz(), cl = mz(), cc = mz(), cr = mz(), ul = mz(), uc = mz(), ur = mz();
    double vx, vy, ax0, ay0, ax1, ay1;
    auto fz = [](double t) { return t < 0.5 ? 0 : (t > 1.5 ? 2 : 2 * (t - 0.5)); };
    if (fx > fy) {
        zz = z = 1 - ny / fx;
        if (nx < 0) {
            k = i;
            if (z < 0.5) square1(z + 0.5, h, v, dc, dr, cc, cr, true);
            else if (z < 1.5) {
                z -= 0.5;
                cc = mneg(h);
                cr = madd(h, smul(1 - 2 * z, v));
                dr = smul(-(1 - z), v);
                ur = smul(z, v);
            } else square2(z - 1.5, h, v, cc, cr, uc, ur, true);
        } else {
            k = i - 1;
            if (z < 0.5) square2(z + 0.5, h, v, dl, dc, cl, cc, false);
            else if (z < 1.5) {
                z -= 0.5;
                cc = h;
                cl = madd(mneg(h), smul(1 - 2 * z, v));
                dl = smul(-(1 - z), v);
                ul = smul(z, v);
            } else square1(z - 1.5, h, v, cl, cc, ul, uc, false);
        }
        zz = fz(zz);
        contrib(f, ml, k, j - 1, nx, ny, ax0, ay0);
        contrib(f, ml, k, j, nx, ny, ax1, ay1);
    } else {
        zz = z = 1 - nx / fy;
        if (ny < 0) {
            k = j;
            if (z < 0.5) square1(z + 0.5, h, v, cl, cc, ul, uc, false);
            else if (z < 1.5) {
                z -= 0.5;
                cc = mneg(v);
                uc = madd(v, smul(1 - 2 * z, h));
                ul = smul(-(1 - z), h);
                ur = smul(z, h);
            } else square2(2.5 - z, h, v, cc, cr, uc, ur, true);
        } else {
            k = j - 1;
            if (z < 0.5) square2(0.5 - z, h, v, dl, dc, cl, cc, false);
            else if (z < 1.5) {
                z -= 0.5;
                cc = v;
                dc = madd(mneg(v), smul(1 - 2 * z, h));
                dl = smul(-(1 - z), h);
                dr = smul(z, h);
            } else square1(z - 1.5, h, v, dc, dr, cc, cr, true);
        }
        zz = fz(zz);
        contrib(f, ml, i - 1, k, nx, ny, ax0, ay0);
        contrib(f, ml, i, k, nx, ny, ax1, ay1);
    }
    // vsr = 0.5*((2-zz)*contrib0 + zz*contrib1)
    vx = 0.5 * ((2 - zz) * ax0 + zz * ax1);
    vy = 0.5 * ((2 - zz) * ay0 + zz * ay1);
    // precondition: cc = ru*(ncs/cc), source idt*cc*vsr, blocks cc*a
    const M2 ru{op.cxx, 0, 0, op.cyy};
    const M2 P = mmul(ru, sdiv(op.ncs, cc));
    const M2 Pi = smul(idt, P);
    src[2 * ij] = Pi.a * vx + Pi.b * vy;
    src[2 * ij + 1] = Pi.c * vx + Pi.d * vy;
    const M2 out[8] = {mmul(P, dl), mmul(P, dc), mmul(P, dr), mmul(P, cl),
                       mmul(P, cr), mmul(P, ul), mmul(P, uc), mmul(P, ur)};
    const int slot = atomicAdd(nfree, 1);
    if (slot >= fcap) {
        rowt[ij] = -4;  // no room: the host grows the table and runs again
        return;
    }
    rowt[ij] = slot;
    double *o = fmats + 32 * (size_t)slot;
    for (int q = 0; q < 8; q++) {
        o[4 * q] = out[q].a;
        o[4 * q + 1] = out[q].b;
        o[4 * q + 2] = out[q].c;
        o[4 * q + 3] = out[q].d;
    }
}

// velocity from the solve at the interior nodes (bd_sim.cc:521-525)
__global__ void __launch_bounds__(kThreads) correct_nodes_kernel(BdConst c, BdNodes nd,
                                                                 const double *__restrict__ vel) {
    const int k = blockIdx.x * kThreads + threadIdx.x, me = c.me;
    if (k >= me * c.ne) return;
    const int i = k % me, j = k / me;
    if (i > 0 && j > 0 && i < c.m && j < c.n) {
        nd.u[k] = vel[2 * k];
        nd.v[k] = vel[2 * k + 1];
    }
}

// stress correction at the cells inside the body (bd_sim.cc:526-542)
__global__ void __launch_bounds__(kThreads) correct_cells_kernel(BdConst c, const int8_t *__restrict__ st,
                                                                 BdCells f, const double *__restrict__ vel,
                                                                 double dt) {
    const int ij = blockIdx.x * kThreads + threadIdx.x, m = c.m, me = c.me;
    if (ij >= m * c.n || st[ij] >= 4) return;
    const int i = ij % m, j = ij / m, k = i + me * j;
    const double third = 1 / 3.0, hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    const double *vp = vel + 2 * (size_t)k;
    const double x0 = vp[0], y0 = vp[1], x1 = vp[2], y1 = vp[3];
    const double x2 = vp[2 * me], y2 = vp[2 * me + 1], x3 = vp[2 * me + 2], y3 = vp[2 * me + 3];
    const double ux = hx * (-x0 + x1 - x2 + x3), uy = hy * (-x0 - x1 + x2 + x3), vx = hx * (-y0 + y1 - y2 + y3),
                 vy = hy * (-y0 - y1 + y2 + y3);
    f.p[ij] += -c.K * (ux + vy);
    f.q[ij] += -c.mu * third * (ux + vy);
    f.s[ij] += c.mu * (ux - vy);
    f.tau[ij] += c.mu * (uy + vx);
}

// init_bar geometry (bd_sim.cc:63-98)
__global__ void __launch_bounds__(kThreads) bar_phi_kernel(BdConst c, int ca, double *phi) {
    const int ij = blockIdx.x * kThreads + threadIdx.x;
    if (ij >= c.m * c.n) return;
    const int i = ij % c.m, j = ij / c.m;
    const double Bx = -1, Bw = 0.65, Bc = 0.3, Bspc = 0.4;
    const double y = c.ay + (j + 0.5) * c.dy;
    double x, out;
    if (ca == 1) {
        x = c.ax + (i + 0.5) * c.dx - Bx;
        out = x > -Bc ? (-fabs(y) > x ? -x - Bw : fabs(y) - Bw)
                      : (fabs(y) < Bc ? -x - Bw
                                      : Bc - Bw + sqrt((x + Bc) * (x + Bc) + (fabs(y) - Bc) * (fabs(y) - Bc)));
    } else if (ca == 2) {
        x = c.ax + (i + 0.5) * c.dx;
        double sp, tphi = fabs(x) - 0.25;
        sp = 0.15 - sqrt(x * x + (y - Bspc) * (y - Bspc));
        if (sp > tphi) tphi = sp;
        sp = 0.15 - sqrt(x * x + (y - 2 * Bspc) * (y - 2 * Bspc));
        if (sp > tphi) tphi = sp;
        sp = 0.15 - sqrt(x * x + (y + Bspc) * (y + Bspc));
        if (sp > tphi) tphi = sp;
        sp = 0.15 - sqrt(x * x + (y + 2 * Bspc) * (y + 2 * Bspc));
        if (sp > tphi) tphi = sp;
        out = tphi;
    } else {
        x = c.ax + (i + 0.5) * c.dx;
        out = sqrt(4 * y * y + x * x) - 1;
    }
    phi[ij] = out;
}

// init_bar fields (bd_sim.cc:107-120)
__global__ void __launch_bounds__(kThreads) bar_fields_kernel(BdConst c, int ca, const int8_t *__restrict__ st,
                                                              const int8_t *__restrict__ nst, BdCells f,
                                                              BdNodes nd) {
    const int k = blockIdx.x * kThreads + threadIdx.x, me = c.me;
    if (k >= me * c.ne) return;
    const int i = k % me, j = k / me;
    const double chi0 = 620 / c.TZ;
    if (i < c.m && j < c.n) {
        const int ij = i + c.m * j;
        if (st[ij] < 4) {
            f.p[ij] = f.q[ij] = f.s[ij] = f.tau[ij] = 0;
            f.chi[ij] = chi0;
        }
    }
    if (nst[k] <= 4) {
        const double x = c.ax + i * c.dx;
        nd.X[k] = x;
        nd.Y[k] = c.ay + j * c.dy;
        if (ca < 3) nd.u[k] = nd.v[k] = 0;
        else {
            nd.u[k] = -0.2 * nd.Y[k];
            nd.v[k] = 0.2 * nd.X[k];
        }
    }
}

}  // namespace

BdOp bd_op_constants(const BdConst &c, double dt) {
    // qs_multi_bd.cc:9-20, 29-51
    BdOp o;
    const double Kp = c.K + c.mu / 3.0, Kmu = Kp + c.mu;
    const double xsp2 = c.xsp * c.xsp, ysp2 = c.ysp * c.ysp, xyfac = xsp2 + ysp2;
    const double tmufac1 = Kmu * xsp2 + c.mu * ysp2, tmufac2 = c.mu * xsp2 + Kmu * ysp2;
    const double idt = 1 / dt, viscdt = c.viscosity * idt;
    o.qK = .25 * Kp * c.xsp * c.ysp;
    o.Kmx = (Kmu + viscdt) * xsp2;
    o.Kmy = (Kmu + viscdt) * ysp2;
    o.mx = (c.mu + viscdt) * xsp2;
    o.my = (c.mu + viscdt) * ysp2;
    o.cxx = -2 * (tmufac1 + viscdt * xyfac);
    o.cyy = -2 * (tmufac2 + viscdt * xyfac);
    o.bcs = 1;
    o.ocs = 1e-3;
    o.ncs = 0.14;
    return o;
}

int bd_advect(const BdConst &c, const int8_t *st, const int8_t *nst, const BdCells &o, const BdNodes &on,
              const BdCells &w, const BdNodes &wn, double dt, hipStream_t s) {
    hipLaunchKernelGGL(advect_nodes_kernel, blocks((long)c.me * c.ne), dim3(kThreads), 0, s, c, nst, on, wn, dt);
    STZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(advect_cells_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, st, o, on, w, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_chi_diffusion(const BdConst &c, const int8_t *st, const BdCells &o, const BdCells &w, double dt,
                     hipStream_t s) {
    (void)dt;
    hipLaunchKernelGGL(chi_diffusion_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, st, o.chi, w.ddev,
                       w.chi);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_boundaries(const BdConst &c, const BdNodes &f, double edge_shift, hipStream_t s) {
    (void)edge_shift;
    hipLaunchKernelGGL(boundaries_kernel, blocks((long)c.me + c.n), dim3(kThreads), 0, s, c, f);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_operator(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdCells &f,
                const BdNodes &nd, const double *force, double dt, int *rowt, double *src, double *fmats, int fcap,
                int *nfree, hipStream_t s) {
    const BdOp op = bd_op_constants(c, dt);
    hipLaunchKernelGGL(operator_kernel, blocks((long)c.me * c.ne), dim3(kThreads), 0, s, c, op, phi, st, nst, f, nd,
                       force, dt, rowt, src, fmats, fcap, nfree);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_correct(const BdConst &c, const int8_t *st, const BdCells &f, const BdNodes &nd, const double *vel,
               double dt, hipStream_t s) {
    hipLaunchKernelGGL(correct_nodes_kernel, blocks((long)c.me * c.ne), dim3(kThreads), 0, s, c, nd, vel);
    STZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(correct_cells_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, st, f, vel, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_init_bar_phi(const BdConst &c, int ca, double *phi, hipStream_t s) {
    hipLaunchKernelGGL(bar_phi_kernel, blocks((long)c.m * c.n), dim3(kThreads), 0, s, c, ca, phi);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int bd_init_bar_fields(const BdConst &c, int ca, const int8_t *st, const int8_t *nst, const BdCells &f,
                       const BdNodes &nd, hipStream_t s) {
    hipLaunchKernelGGL(bar_fields_kernel, blocks((long)c.me * c.ne), dim3(kThreads), 0, s, c, ca, st, nst, f, nd);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

}  // namespace stz
