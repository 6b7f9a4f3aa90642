// adjoint.hip -- hand-written adjoint (reverse mode) of the FP64
// quasi-static step, for gradients of a function of the final state with
// respect to the initial state and to the STZ parameters (c0, Omega*eps0,
// chi_inf).
//
// The forward step (shear_sim.cc:280-438) is
//   F' = advect(F)          ENO2 advection + adaptive STZ plasticity
//   b  = source(F')         div(sigma)/dt, wall rows constant in F
//   z  = A^{-1} b           the multigrid solve (converged: exact inverse)
//   F''= correct(F', z)     stress increments from grad z, u,v <- z, ghosts
// and its adjoint runs the transposes in reverse order:
//   adj_correct  lambda(F'') -> lambda(F') and g = dL/dz
//   the adjoint solve  w = A^{-T} g: the interior block of A is symmetric
//                (the interior stencil's blocks are symmetric and mirrored)
//                and the wall rows hold only their diagonal, so w is the
//                forward solve of A w = g with g zeroed on the wall rows
//   adj_source   lambda(F') += source^T w
//   adj_advect   lambda(F) = advect'(F)^T lambda(F'), parameter gradient
// with every discrete choice of the forward (upwind side, ENO branch,
// one-sided wall rule, yield branch, plastic substep count) taken again from
// the recomputed forward values: the adjoint kernels re-evaluate the
// forward's expressions in the same order, so they take the same branches.
// The plastic substep loop is differentiated in forward mode (tangents of
// sbar, chi and the remaining time with respect to five inputs) along the
// recomputed loop.
//
// Layout: adjoint planes as the fields (m columns, n+2 rows, FieldPtrs),
// node vectors as vel (m*(n+1) interleaved vec2).  The advection adjoint
// scatters into the adjoint of the pre-step fields with FP64 atomics.
#include "adjoint.h"
#include "dual.h"
#include "stz_law.h"

#pragma clang fp contract(off)

namespace stz {

#define PL(F, i, j) (F)[(i) + m * ((j) + 1)]

// eno2 (shear_sim.cc:653-655) with its branch frozen: value and the weights
// of its four arguments
__device__ __forceinline__ static double eno2w(double p0, double p1, double p2, double p3, double w[4]) {
    const bool b = fabs(p0 - 2 * p1 + p2) > fabs(p1 - 2 * p2 + p3);
    if (b) {
        w[0] = 0; w[1] = 3; w[2] = -4; w[3] = 1;
        return 3 * p1 - 4 * p2 + p3;
    }
    w[0] = 1; w[1] = 0; w[2] = -1; w[3] = 0;
    return p0 - p2;
}

__device__ __forceinline__ static void aadd(double *p, double v) {
    if (v != 0) atomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// adaptive plastic term (stz_law.h plastic_term) with its Jacobian.
// Inputs: 0 sbar, 1 chi, 2 c0, 3 Omegaeps0, 4 chi_inf.  Outputs dtmp =
// ddev/sbar, chi after the substeps, ddev.  The loop runs once in dual
// numbers (dual.h) through the same law code as the step (dplastic_t), so
// every STZ law and truncation model is covered and the values, hence the
// yield branch and the substep count, are the forward's own.
// ---------------------------------------------------------------------------
struct PlasticJac {
    double dtmp, chi, ddev;
    double J[3][5];
};

__device__ static void plastic_jac(const SimConst &c, double sbar0, double chi0, double dt, PlasticJac &P) {
    typedef Dual<5> D5;
    const double small_dev_cutoff = 1e-12;
    const double theta = 0;
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 5; k++) P.J[r][k] = 0;
    if (sbar0 < small_dev_cutoff) {
        P.dtmp = 0;
        P.ddev = 0;
        P.chi = chi0;
        P.J[1][1] = 1;
        return;
    }
    const stz_model_params &mp = c.stz;
    const D5 s0 = D5::var(sbar0, 0), c0 = D5::var(mp.c0, 2), Oe = D5::var(mp.Omegaeps0, 3),
             chi_inf = D5::var(c.chi_inf, 4);
    D5 s = s0, chi = D5::var(chi0, 1), dtl(dt), dplas, dchi1, dchi2, adt;
    bool adapt = true;
    do {
        // the forward's own expressions and order (stz_law.h plastic_term)
        dplas = 2 * c.t_scale * c.mu * dplastic_t(mp, c.thetainv, s, chi, Oe, c0, dchi1, dchi2);
        if (fabs(dtl * dplas) > c.adapt_fac) {
            adt = c.adapt_fac / fabs(dplas);
            dtl -= adt;
        } else {
            adapt = false;
            adt = dtl;
        }
        s -= adt * dplas;
        chi += (dchi1 * (chi_inf - chi) + dchi2 * (theta - chi)) * c.t_scale * adt;
    } while (adapt);
    const D5 ddev = s0 - s, dtmp = ddev / s0;
    P.ddev = ddev.v;
    P.dtmp = dtmp.v;
    P.chi = chi.v;
    for (int k = 0; k < 5; k++) {
        P.J[0][k] = dtmp.d[k];
        P.J[1][k] = chi.d[k];
        P.J[2][k] = ddev.d[k];
    }
}

// ---------------------------------------------------------------------------
// adjoint of correct_kernel (explicit.hip, shear_sim.cc:418-434 and the
// set_boundaries rows): one thread per cell column entry (i, j), j < n.
// L: adjoint of the state after the step; A: adjoint of the state after the
// advection (written here for p,q,s,tau,chi,X,Y,ddev,u,v); coef: per cell
// the adjoints of (ux, uy, vx, vy) for the node gather below
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) adj_correct_cells(SimConst c, FieldPtrs L, FieldPtrs A, double *coef) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const double third = 1 / 3.0;
    // the corrected value of row j also feeds the ghost rows (2 P0 - P1,
    // 2 P(n-1) - P(n-2))
    auto eff = [&](const double *F) {
        double v = PL(F, i, j);
        if (j == 0) v += 2 * PL(F, i, -1);
        if (j == 1) v -= PL(F, i, -1);
        if (j == n - 1) v += 2 * PL(F, i, n);
        if (j == n - 2) v -= PL(F, i, n);
        return v;
    };
    const double lp = eff(L.p), lq = eff(L.q), ls = eff(L.s), lt = eff(L.tau), lc = eff(L.chi);
    PL(A.p, i, j) = lp;
    PL(A.q, i, j) = lq;
    PL(A.s, i, j) = ls;
    PL(A.tau, i, j) = lt;
    PL(A.chi, i, j) = lc;
    PL(A.ddev, i, j) = PL(L.ddev, i, j);
    // p += -K (ux+vy), q += -(mu/3)(ux+vy), s += mu (ux-vy), tau += mu (uy+vx)
    const double a = -c.K * lp - c.mu * third * lq, bs = c.mu * ls, bt = c.mu * lt;
    double *cf = coef + 4 * (size_t)(i + m * j);
    cf[0] = a + bs;  // ux
    cf[1] = bt;      // uy
    cf[2] = bt;      // vx
    cf[3] = a - bs;  // vy
    // ghost rows, overwritten by the correction: no adjoint flows back
    if (j == 0) {
        PL(A.p, i, -1) = 0; PL(A.q, i, -1) = 0; PL(A.s, i, -1) = 0; PL(A.tau, i, -1) = 0; PL(A.chi, i, -1) = 0;
        PL(A.ddev, i, -1) = PL(L.ddev, i, -1);
    }
    if (j == n - 1) {
        PL(A.p, i, n) = 0; PL(A.q, i, n) = 0; PL(A.s, i, n) = 0; PL(A.tau, i, n) = 0; PL(A.chi, i, n) = 0;
        PL(A.ddev, i, n) = PL(L.ddev, i, n);
    }
}

// nodes: u, v (overwritten by z on rows 1..n-1, by the wall values on rows
// 0 and n) and X, Y (wall rows overwritten); g = dL/dz on interior nodes
__global__ void __launch_bounds__(256) adj_correct_nodes(SimConst c, FieldPtrs L, FieldPtrs A, const double *coef,
                                                         double *g, double dt) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;  // node row 0..n
    if (i >= m || j > n) return;
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    PL(A.u, i, j) = 0;
    PL(A.v, i, j) = 0;
    const bool wall = j == 0 || j == n;
    PL(A.X, i, j) = wall ? 0 : PL(L.X, i, j);
    PL(A.Y, i, j) = wall ? 0 : PL(L.Y, i, j);
    if (j == 0) {  // plane row -1 of the node fields: not touched by the correction
        PL(A.X, i, -1) = PL(L.X, i, -1);
        PL(A.Y, i, -1) = PL(L.Y, i, -1);
        PL(A.u, i, -1) = PL(L.u, i, -1);
        PL(A.v, i, -1) = PL(L.v, i, -1);
    }
    double gx = 0, gy = 0;
    if (!wall) {
        gx = PL(L.u, i, j);
        gy = PL(L.v, i, j);
        const int il = i == 0 ? m - 1 : i - 1;
        // the four cells reading node (i,j): as vp (i,j), vo (i-1,j), vpu
        // (i,j-1), vou (i-1,j-1)
        struct Role { int ci, cj; double sx_ux, sx_uy; };
        const Role roles[4] = {{i, j, -1, -1}, {il, j, 1, -1}, {i, j - 1, -1, 1}, {il, j - 1, 1, 1}};
        for (int r = 0; r < 4; r++) {
            const int cj = roles[r].cj;
            if (cj < 0 || cj >= n) continue;
            const double *cf = coef + 4 * (size_t)(roles[r].ci + m * cj);
            // ux = hx(-vpx + vox - vpux + voux), uy = hy(-vpx - vox + vpux + voux), same for v
            gx += roles[r].sx_ux * hx * cf[0] + roles[r].sx_uy * hy * cf[1];
            gy += roles[r].sx_ux * hx * cf[2] + roles[r].sx_uy * hy * cf[3];
        }
    }
    g[2 * (i + m * j)] = gx;
    g[2 * (i + m * j) + 1] = gy;
}

// adjoint of source_kernel on the interior node rows (the wall rows of the
// source do not depend on the fields): A.{p,q,s,tau} += source^T w
__global__ void __launch_bounds__(256) adj_source(SimConst c, FieldPtrs A, const double *w, double dt) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const double xfac = 0.5 / dt * c.xsp, yfac = 0.5 / dt * c.ysp;
    const int ir = i == m - 1 ? 0 : i + 1;
    // cell (i,j) is f3 of node (i,j), f2 of node (i+1,j), f1 of node
    // (i,j+1), f0 of node (i+1,j+1).  Weights of the cell's values in that
    // role: px, sx of p (and q), s in fx (times xfac); ty of tau in fx
    // (times yfac); tx of tau in fy (times xfac); py of p, q, s in fy (times
    // yfac)
    struct Role { int ni, nj; double px, sx, tx, ty, py; };
    const Role roles[4] = {{i, j, 1, -1, -1, -1, 1}, {ir, j, -1, 1, 1, -1, 1},
                           {i, j + 1, 1, -1, -1, 1, -1}, {ir, j + 1, -1, 1, 1, 1, -1}};
    double lp = 0, ls = 0, lt = 0;
    for (int r = 0; r < 4; r++) {
        const int nj = roles[r].nj;
        if (nj < 1 || nj > n - 1) continue;
        const double wx = w[2 * (roles[r].ni + m * nj)], wy = w[2 * (roles[r].ni + m * nj) + 1];
        lp += roles[r].px * xfac * wx + roles[r].py * yfac * wy;
        ls += roles[r].sx * xfac * wx + roles[r].py * yfac * wy;
        lt += roles[r].ty * yfac * wx + roles[r].tx * xfac * wy;
    }
    PL(A.p, i, j) += lp;
    PL(A.q, i, j) += lp;
    PL(A.s, i, j) += ls;
    PL(A.tau, i, j) += lt;
}

// ---------------------------------------------------------------------------
// adjoint of advect_kernel (explicit.hip, shear_sim.cc:296-361): the forward
// values are recomputed from the pre-step fields o; A is the adjoint of the
// post-advection state; G (zeroed by the caller) receives the adjoint of o
// by atomic scatter; grad[0..2] += d/d(c0, Omegaeps0, chi_inf)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) adj_advect(SimConst c, FieldPtrs o, FieldPtrs A, FieldPtrs G, double dt,
                                                  double *grad) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    double g0 = 0, g1 = 0, g2 = 0;
    if (i < m && j < n) {
        const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt, lx = c.lx;
        const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
        const int ill = i <= 1 ? i + m - 2 : i - 2, irr = i >= m - 2 ? i + 2 - m : i + 2;

        if (j > 0) {
            // node (i,j): X' = X0 - uc Xx - vc Xy, likewise Y
            const double uc = PL(o.u, i, j), vc = PL(o.v, i, j);
            const double lX = PL(A.X, i, j), lY = PL(A.Y, i, j);
            const double X0 = PL(o.X, i, j), Y0 = PL(o.Y, i, j);
            int xi[4];
            double hxs, Xx, Yx, wX[4], wY[4];
            if (uc > 0) {
                xi[0] = ir; xi[1] = i; xi[2] = il; xi[3] = ill;
                hxs = hx;
                Xx = hx * eno2w(PL(o.X, ir, j) + (i == m - 1 ? lx : 0), X0, PL(o.X, il, j) + (i == 0 ? -lx : 0),
                                PL(o.X, ill, j) + (i <= 1 ? -lx : 0), wX);
                Yx = hx * eno2w(PL(o.Y, ir, j), Y0, PL(o.Y, il, j), PL(o.Y, ill, j), wY);
            } else {
                xi[0] = il; xi[1] = i; xi[2] = ir; xi[3] = irr;
                hxs = -hx;
                Xx = -hx * eno2w(PL(o.X, il, j) + (i == 0 ? -lx : 0), X0, PL(o.X, ir, j) + (i == m - 1 ? lx : 0),
                                 PL(o.X, irr, j) + (i >= m - 2 ? lx : 0), wX);
                Yx = -hx * eno2w(PL(o.Y, il, j), Y0, PL(o.Y, ir, j), PL(o.Y, irr, j), wY);
            }
            int yj[4];
            int ny;
            double hys, Xy, Yy, vX[4], vY[4];
            if (vc > 0) {
                if (j > 1) {
                    yj[0] = j + 1; yj[1] = j; yj[2] = j - 1; yj[3] = j - 2;
                    ny = 4;
                    hys = hy;
                    Xy = hy * eno2w(PL(o.X, i, j + 1), X0, PL(o.X, i, j - 1), PL(o.X, i, j - 2), vX);
                    Yy = hy * eno2w(PL(o.Y, i, j + 1), Y0, PL(o.Y, i, j - 1), PL(o.Y, i, j - 2), vY);
                } else {
                    const double hs = hy * 2;
                    yj[0] = j; yj[1] = j - 1;
                    ny = 2;
                    hys = hs;
                    vX[0] = vY[0] = 1; vX[1] = vY[1] = -1;
                    Xy = hs * (X0 - PL(o.X, i, j - 1));
                    Yy = hs * (Y0 - PL(o.Y, i, j - 1));
                }
            } else {
                if (j < n - 1) {
                    yj[0] = j - 1; yj[1] = j; yj[2] = j + 1; yj[3] = j + 2;
                    ny = 4;
                    hys = -hy;
                    Xy = -hy * eno2w(PL(o.X, i, j - 1), X0, PL(o.X, i, j + 1), PL(o.X, i, j + 2), vX);
                    Yy = -hy * eno2w(PL(o.Y, i, j - 1), Y0, PL(o.Y, i, j + 1), PL(o.Y, i, j + 2), vY);
                } else {
                    const double hs = -hy * 2;
                    yj[0] = j; yj[1] = j + 1;
                    ny = 2;
                    hys = hs;
                    vX[0] = vY[0] = 1; vX[1] = vY[1] = -1;
                    Xy = hs * (X0 - PL(o.X, i, j + 1));
                    Yy = hs * (Y0 - PL(o.Y, i, j + 1));
                }
            }
            aadd(&PL(G.X, i, j), lX);
            aadd(&PL(G.Y, i, j), lY);
            for (int k = 0; k < 4; k++) {
                aadd(&PL(G.X, xi[k], j), -uc * hxs * wX[k] * lX);
                aadd(&PL(G.Y, xi[k], j), -uc * hxs * wY[k] * lY);
            }
            for (int k = 0; k < ny; k++) {
                aadd(&PL(G.X, i, yj[k]), -vc * hys * vX[k] * lX);
                aadd(&PL(G.Y, i, yj[k]), -vc * hys * vY[k] * lY);
            }
            aadd(&PL(G.u, i, j), -Xx * lX - Yx * lY);
            aadd(&PL(G.v, i, j), -Xy * lX - Yy * lY);
        } else {
            // copies into the new plane set (row 0 and the ghost row -1)
            aadd(&PL(G.X, i, 0), PL(A.X, i, 0));
            aadd(&PL(G.Y, i, 0), PL(A.Y, i, 0));
            aadd(&PL(G.X, i, -1), PL(A.X, i, -1));
            aadd(&PL(G.Y, i, -1), PL(A.Y, i, -1));
            aadd(&PL(G.p, i, -1), PL(A.p, i, -1));
            aadd(&PL(G.q, i, -1), PL(A.q, i, -1));
            aadd(&PL(G.s, i, -1), PL(A.s, i, -1));
            aadd(&PL(G.tau, i, -1), PL(A.tau, i, -1));
            aadd(&PL(G.chi, i, -1), PL(A.chi, i, -1));
        }
        if (j == n - 1) {
            aadd(&PL(G.X, i, n), PL(A.X, i, n));
            aadd(&PL(G.Y, i, n), PL(A.Y, i, n));
            aadd(&PL(G.p, i, n), PL(A.p, i, n));
            aadd(&PL(G.q, i, n), PL(A.q, i, n));
            aadd(&PL(G.s, i, n), PL(A.s, i, n));
            aadd(&PL(G.tau, i, n), PL(A.tau, i, n));
            aadd(&PL(G.chi, i, n), PL(A.chi, i, n));
        }
        // u, v (all plane rows) and the ghost rows of ddev are not changed by
        // the advection (u, v, ddev are single planes shared by both plane
        // sets): their adjoint passes
        aadd(&PL(G.u, i, j), PL(A.u, i, j));
        aadd(&PL(G.v, i, j), PL(A.v, i, j));
        if (j == 0) {
            aadd(&PL(G.u, i, -1), PL(A.u, i, -1));
            aadd(&PL(G.v, i, -1), PL(A.v, i, -1));
            aadd(&PL(G.ddev, i, -1), PL(A.ddev, i, -1));
        }
        if (j == n - 1) {
            aadd(&PL(G.u, i, n), PL(A.u, i, n));
            aadd(&PL(G.v, i, n), PL(A.v, i, n));
            aadd(&PL(G.ddev, i, n), PL(A.ddev, i, n));
        }

        // ---- cell (i,j)
        const double u00 = PL(o.u, i, j), u10 = PL(o.u, ir, j), u01 = PL(o.u, i, j + 1), u11 = PL(o.u, ir, j + 1);
        const double v00 = PL(o.v, i, j), v10 = PL(o.v, ir, j), v01 = PL(o.v, i, j + 1), v11 = PL(o.v, ir, j + 1);
        const double uc = 0.25 * (u00 + u10 + u01 + u11);
        const double vc = 0.25 * (v00 + v10 + v01 + v11);
        const double two_omega = hx * (-v00 + v10 - v01 + v11) - hy * (-u00 - u10 + u01 + u11);
        const double *F[5] = {o.p, o.q, o.s, o.tau, o.chi};
        double *GF[5] = {G.p, G.q, G.s, G.tau, G.chi};
        double f0[5], f1[5], f2[5], w1[5][4], w2[5][4];
        int xi[4], yj[4], ny;
        double hxs, hys;
        for (int k = 0; k < 5; k++) f0[k] = PL(F[k], i, j);
        if (uc > 0) {
            xi[0] = ir; xi[1] = i; xi[2] = il; xi[3] = ill;
            hxs = hx;
        } else {
            xi[0] = il; xi[1] = i; xi[2] = ir; xi[3] = irr;
            hxs = -hx;
        }
        for (int k = 0; k < 5; k++)
            f1[k] = hxs * eno2w(PL(F[k], xi[0], j), f0[k], PL(F[k], xi[2], j), PL(F[k], xi[3], j), w1[k]);
        if (vc > 0) {
            if (j > 0) {
                yj[0] = j + 1; yj[1] = j; yj[2] = j - 1; yj[3] = j - 2;
                ny = 4;
                hys = hy;
            } else {
                yj[0] = j; yj[1] = j - 1;
                ny = 2;
                hys = hy * 2;
            }
        } else {
            if (j < n - 1) {
                yj[0] = j - 1; yj[1] = j; yj[2] = j + 1; yj[3] = j + 2;
                ny = 4;
                hys = -hy;
            } else {
                yj[0] = j; yj[1] = j + 1;
                ny = 2;
                hys = -hy * 2;
            }
        }
        for (int k = 0; k < 5; k++) {
            if (ny == 4) {
                f2[k] = hys * eno2w(PL(F[k], i, yj[0]), f0[k], PL(F[k], i, yj[2]), PL(F[k], i, yj[3]), w2[k]);
            } else {
                f2[k] = hys * (f0[k] - PL(F[k], i, yj[1]));
                w2[k][0] = 1;
                w2[k][1] = -1;
            }
        }
        const double cchi = f0[4] - uc * f1[4] - vc * f2[4];
        const double sbar = sqrt(3 * f0[1] * f0[1] + f0[2] * f0[2] + f0[3] * f0[3]);
        PlasticJac P;
        plastic_jac(c, sbar, cchi, dt, P);
        const double dtmp = P.dtmp;

        const double lp = PL(A.p, i, j), lq = PL(A.q, i, j), ls = PL(A.s, i, j), lt = PL(A.tau, i, j);
        const double lchi = PL(A.chi, i, j), lddev = PL(A.ddev, i, j);
        const double ldtmp = -(lq * f0[1] + ls * f0[2] + lt * f0[3]);
        double lin[5];
        for (int k = 0; k < 5; k++) lin[k] = ldtmp * P.J[0][k] + lchi * P.J[1][k] + lddev * P.J[2][k];
        const double lsbar = lin[0], lcchi = lin[1];
        g0 = lin[2];
        g1 = lin[3];
        g2 = lin[4];
        const double lom = ls * f0[3] - lt * f0[2];
        const double lk[5] = {lp, lq, ls, lt, lcchi};
        double luc = 0, lvc = 0;
        for (int k = 0; k < 5; k++) {
            luc -= lk[k] * f1[k];
            lvc -= lk[k] * f2[k];
        }
        double lf0[5];
        lf0[0] = lp;
        lf0[1] = lq * (1 - dtmp);
        lf0[2] = ls * (1 - dtmp) - lt * two_omega;
        lf0[3] = lt * (1 - dtmp) + ls * two_omega;
        lf0[4] = lcchi;
        if (sbar > 0) {
            lf0[1] += lsbar * 3 * f0[1] / sbar;
            lf0[2] += lsbar * f0[2] / sbar;
            lf0[3] += lsbar * f0[3] / sbar;
        }
        for (int k = 0; k < 5; k++) {
            aadd(&PL(GF[k], i, j), lf0[k]);
            const double l1 = -uc * lk[k] * hxs, l2 = -vc * lk[k] * hys;
            for (int q = 0; q < 4; q++) aadd(&PL(GF[k], xi[q], j), w1[k][q] * l1);
            for (int q = 0; q < ny; q++) aadd(&PL(GF[k], i, yj[q]), w2[k][q] * l2);
        }
        // uc, vc = corner averages; two_omega = hx(-v00+v10-v01+v11) - hy(-u00-u10+u01+u11)
        aadd(&PL(G.u, i, j), 0.25 * luc + hy * lom);
        aadd(&PL(G.u, ir, j), 0.25 * luc + hy * lom);
        aadd(&PL(G.u, i, j + 1), 0.25 * luc - hy * lom);
        aadd(&PL(G.u, ir, j + 1), 0.25 * luc - hy * lom);
        aadd(&PL(G.v, i, j), 0.25 * lvc - hx * lom);
        aadd(&PL(G.v, ir, j), 0.25 * lvc + hx * lom);
        aadd(&PL(G.v, i, j + 1), 0.25 * lvc - hx * lom);
        aadd(&PL(G.v, ir, j + 1), 0.25 * lvc + hx * lom);
    }
    // parameter gradient: block sums, one atomic per block and parameter
    __shared__ double red[3][4];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y, lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        g0 += __shfl_xor(g0, off, 64);
        g1 += __shfl_xor(g1, off, 64);
        g2 += __shfl_xor(g2, off, 64);
    }
    if (lane == 0) {
        red[0][wid] = g0;
        red[1][wid] = g1;
        red[2][wid] = g2;
    }
    __syncthreads();
    if (tid < 3) atomicAdd(grad + tid, (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]));
}

// ---------------------------------------------------------------------------
// adjoint of chi_diffusion (explicit.hip, shear_sim.cc:364-386): chi after
// the step = advected chi + the constant and ddev-weighted diffusion of the
// pre-step chi, with the advection's new ddev as weights.  o: the pre-step
// state; dnew: the advection's ddev plane.  The pre-step chi's adjoint goes
// into G (before adj_advect adds its own part); the new ddev's into A.ddev,
// which adj_advect then reads.  One thread per cell, scattered with atomics.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) adj_chi_diffusion(SimConst c, FieldPtrs o, const double *dnew, FieldPtrs A,
                                                         FieldPtrs G, double dt) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const double r = PL(A.chi, i, j);
    if (r == 0) return;
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    const double hxxcd = 2 * c.xsp * hx * c.chi_diff, hyycd = 2 * c.ysp * hy * c.chi_diff,
                 hccd = 2 * (hxxcd + hyycd);
    const double dfac = 0.25 * c.chi_len * c.chi_len * c.mu_inv, dfacx = dfac * c.xsp * c.xsp,
                 dfacy = dfac * c.ysp * c.ysp;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const double D0 = PL(dnew, i, j);
    const double Dd = D0 + PL(dnew, i, j - 1), Dl = D0 + PL(dnew, il, j), Dr = D0 + PL(dnew, ir, j),
                 Du = D0 + PL(dnew, i, j + 1);
    const double cf = PL(o.chi, i, j), cl = PL(o.chi, il, j), cr = PL(o.chi, ir, j), cd = PL(o.chi, i, j - 1),
                 cu = PL(o.chi, i, j + 1);
    aadd(&PL(G.chi, i, j), r * (-hccd - dfacx * (Dr + Dl) - dfacy * (Du + Dd)));
    aadd(&PL(G.chi, ir, j), r * (hxxcd + dfacx * Dr));
    aadd(&PL(G.chi, il, j), r * (hxxcd + dfacx * Dl));
    aadd(&PL(G.chi, i, j + 1), r * (hyycd + dfacy * Du));
    aadd(&PL(G.chi, i, j - 1), r * (hyycd + dfacy * Dd));
    const double gr = r * dfacx * (cr - cf), gl = r * dfacx * (cl - cf), gu = r * dfacy * (cu - cf),
                 gd = r * dfacy * (cd - cf);
    aadd(&PL(A.ddev, i, j), gr + gl + gu + gd);
    aadd(&PL(A.ddev, ir, j), gr);
    aadd(&PL(A.ddev, il, j), gl);
    aadd(&PL(A.ddev, i, j + 1), gu);
    aadd(&PL(A.ddev, i, j - 1), gd);
}

__global__ void __launch_bounds__(256) adj_fill(double *p, size_t count) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < count) p[k] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

int adj_zero(double *p, size_t count, hipStream_t st) {
    hipLaunchKernelGGL(adj_fill, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int adj_launch_correct(const SimConst &c, const FieldPtrs &L, const FieldPtrs &A, double *coef, double *g,
                       double dt, hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(adj_correct_cells, grid2(c.m, c.n, b), b, 0, st, c, L, A, coef);
    hipLaunchKernelGGL(adj_correct_nodes, grid2(c.m, c.n + 1, b), b, 0, st, c, L, A, coef, g, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int adj_launch_source(const SimConst &c, const FieldPtrs &A, const double *w, double dt, hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(adj_source, grid2(c.m, c.n, b), b, 0, st, c, A, w, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int adj_launch_chi_diffusion(const SimConst &c, const FieldPtrs &o, const double *dnew, const FieldPtrs &A,
                             const FieldPtrs &G, double dt, hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(adj_chi_diffusion, grid2(c.m, c.n, b), b, 0, st, c, o, dnew, A, G, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int adj_launch_advect(const SimConst &c, const FieldPtrs &o, const FieldPtrs &A, const FieldPtrs &G, double dt,
                      double *grad, hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(adj_advect, grid2(c.m, c.n, b), b, 0, st, c, o, A, G, dt, grad);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

}  // namespace stz
