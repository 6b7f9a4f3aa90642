// explicit.hip -- gfx950 kernels for the explicit part of the quasi-static STZ
// step (shear_sim::advection_step, chi_diffusion, the projection source and
// correction, set_boundaries, init_fields).
//
// Layout: structure-of-arrays FP64 planes, m columns by n+2 rows (rows -1..n),
// one thread per cell, 64 threads of a wavefront along a row so every plane
// access is a coalesced 512-byte row segment.  The advection kernel reads the
// old planes and writes a second plane set, which replaces the reference's
// change slots cp..cY (fields.hh:28-51) and its separate update loop
// (shear_sim.cc:356-360): the state after the kernel is the same, with seven
// fewer plane reads and writes.
//
// Every arithmetic expression keeps the reference's evaluation order, and FMA
// contraction is off in this file, so results agree bit for bit with the
// reference wherever no transcendental function is involved (the ENO, upwind
// and yield decisions in particular); exp/tanh/cosh differ from glibc by at
// most a few ulp.
#include "batch.h"
#include "device.h"
#include "stz_law.h"

#pragma clang fp contract(off)

namespace stz {

#define PL(F, i, j) (F)[(i) + m * ((j) + 1)]

// shear_sim.cc:653-655; br = 1 when the one-sided second-order branch is taken.
// force >= 0 takes that branch (1) or the centred one (0) without the test.
__device__ __forceinline__ static double eno2(double p0, double p1, double p2, double p3, uint32_t &br,
                                              int force = -1) {
    bool b = force < 0 ? fabs(p0 - 2 * p1 + p2) > fabs(p1 - 2 * p2 + p3) : force != 0;
    br = b ? 1u : 0u;
    return b ? 3 * p1 - 4 * p2 + p3 : p0 - p2;
}

// ---------------------------------------------------------------------------
// advection + plasticity: shear_sim.cc:296-349 (+ update, :356-360)
//
// REPLAY: every decision (upwind sides, ENO branches, the wall rules that
// follow from the upwind side, the plastic cut-off, substep count and last
// yield branch) is read from rep, a record of the same cell written by an
// earlier step, instead of being evaluated.  A substep count of 255 means the
// record saturated; such a cell evaluates its plastic loop as usual.
// ---------------------------------------------------------------------------
// advection tile: 64 x 4 cells per workgroup (one wave per row), staged
// with a halo of two columns and two rows on each side (the ENO stencils
// reach +-2, the staggered velocity +1)
constexpr int kAdvTileX = 64, kAdvTileY = 4;
constexpr int kAdvRows = kAdvTileY + 4, kAdvCols = kAdvTileX + 4, kAdvPlanes = 9;
constexpr int kAdvPlaneElems = kAdvRows * kAdvCols;
constexpr int kAdvThreads = kAdvTileX * kAdvTileY;
constexpr int kAdvPasses = (kAdvPlaneElems + kAdvThreads - 1) / kAdvThreads;
enum { kU = 0, kV, kX, kY, kP, kQ, kS, kTau, kChi };

// stage rows j0-2 .. j0+5 and columns i0-2 .. i0+65 (x periodic) of u, v,
// X, Y, p, q, s, tau, chi into tile[plane][row][column]; rows outside -1..n
// hold a copy of the nearest one and are never read
__device__ __forceinline__ static void stage_advect_tile(const FieldPtrs &o, int m, int n, int i0, int j0,
                                                         double *tile) {
    const int t = threadIdx.x + kAdvTileX * threadIdx.y;
    int off[kAdvPasses];
#pragma unroll
    for (int w = 0; w < kAdvPasses; w++) {
        const int e = min(t + kAdvThreads * w, kAdvPlaneElems - 1);
        const int row = e / kAdvCols, col = e - row * kAdvCols;
        const int r = min(max(j0 - 2 + row, -1), n);
        int ca = i0 - 2 + col;
        if (ca < 0) ca += m;
        while (ca >= m) ca -= m;
        off[w] = ca + m * (r + 1);
    }
    const double *const P[kAdvPlanes] = {o.u, o.v, o.X, o.Y, o.p, o.q, o.s, o.tau, o.chi};
    double v[kAdvPlanes][kAdvPasses];
#pragma unroll
    for (int k = 0; k < kAdvPlanes; k++)
#pragma unroll
        for (int w = 0; w < kAdvPasses; w++) v[k][w] = P[k][off[w]];
#pragma unroll
    for (int k = 0; k < kAdvPlanes; k++)
#pragma unroll
        for (int w = 0; w < kAdvPasses; w++)
            if (t + kAdvThreads * w < kAdvPlaneElems) tile[k * kAdvPlaneElems + t + kAdvThreads * w] = v[k][w];
    __syncthreads();
}

template <bool REPLAY>
__device__ __forceinline__ static void advect_body(const SimConst &c, const FieldPtrs &o, const FieldPtrs &w,
                                                   double dt, uint32_t *dec, const uint32_t *rep, int bx,
                                                   int by) {
    const int m = c.m, n = c.n;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = bx * kAdvTileX + tx;
    const int j = by * kAdvTileY + ty;
    // the block's tile of the nine input planes with its +-2 halo, staged
    // with every load of a thread in flight at once; each read below is an
    // LDS access at a fixed offset from the thread's own cell
    __shared__ double tile[kAdvPlanes * kAdvRows * kAdvCols];
    stage_advect_tile(o, m, n, bx * kAdvTileX, by * kAdvTileY, tile);
    if (i >= m || j >= n) return;
#define RD(k, dx, dy) tile[((k) * kAdvRows + ty + 2 + (dy)) * kAdvCols + tx + 2 + (dx)]
    const uint32_t rb = REPLAY ? rep[i + m * j] : 0u;
    // the recorded branch of decision bit b, or -1 (evaluate) without replay
#define RB(b) (REPLAY ? (int)((rb >> (b)) & 1u) : -1)
#define UP(natural, mask) (REPLAY ? (rb & (mask)) != 0 : (natural))
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt, lx = c.lx;
    uint32_t bits = 0, br;
    double uc, vc, two_omega, dtmp;

    if (j > 0) {
        double Xx, Xy, Yx, Yy;
        uc = RD(kU, 0, 0);
        vc = RD(kV, 0, 0);
        const double X0 = RD(kX, 0, 0), Y0 = RD(kY, 0, 0);
        if (UP(uc > 0, STZ_D_NODE_UPX)) {
            bits |= STZ_D_NODE_UPX;
            Xx = hx * eno2(RD(kX, 1, 0) + (i == m - 1 ? lx : 0), X0, RD(kX, -1, 0) + (i == 0 ? -lx : 0),
                           RD(kX, -2, 0) + (i <= 1 ? -lx : 0), br, RB(2));
            bits |= br ? STZ_D_NODE_ENO_XX : 0;
            Yx = hx * eno2(RD(kY, 1, 0), Y0, RD(kY, -1, 0), RD(kY, -2, 0), br, RB(3));
            bits |= br ? STZ_D_NODE_ENO_YX : 0;
        } else {
            Xx = -hx * eno2(RD(kX, -1, 0) + (i == 0 ? -lx : 0), X0, RD(kX, 1, 0) + (i == m - 1 ? lx : 0),
                            RD(kX, 2, 0) + (i >= m - 2 ? lx : 0), br, RB(2));
            bits |= br ? STZ_D_NODE_ENO_XX : 0;
            Yx = -hx * eno2(RD(kY, -1, 0), Y0, RD(kY, 1, 0), RD(kY, 2, 0), br, RB(3));
            bits |= br ? STZ_D_NODE_ENO_YX : 0;
        }
        if (UP(vc > 0, STZ_D_NODE_UPY)) {
            bits |= STZ_D_NODE_UPY;
            if (j > 1) {
                Xy = hy * eno2(RD(kX, 0, 1), X0, RD(kX, 0, -1), RD(kX, 0, -2), br, RB(4));
                bits |= br ? STZ_D_NODE_ENO_XY : 0;
                Yy = hy * eno2(RD(kY, 0, 1), Y0, RD(kY, 0, -1), RD(kY, 0, -2), br, RB(5));
                bits |= br ? STZ_D_NODE_ENO_YY : 0;
            } else {
                const double hs = hy * 2;
                Xy = hs * (X0 - RD(kX, 0, -1));
                Yy = hs * (Y0 - RD(kY, 0, -1));
                bits |= STZ_D_NODE_ONESIDED;
            }
        } else {
            if (j < n - 1) {
                Xy = -hy * eno2(RD(kX, 0, -1), X0, RD(kX, 0, 1), RD(kX, 0, 2), br, RB(4));
                bits |= br ? STZ_D_NODE_ENO_XY : 0;
                Yy = -hy * eno2(RD(kY, 0, -1), Y0, RD(kY, 0, 1), RD(kY, 0, 2), br, RB(5));
                bits |= br ? STZ_D_NODE_ENO_YY : 0;
            } else {
                const double hs = -hy * 2;
                Xy = hs * (X0 - RD(kX, 0, 1));
                Yy = hs * (Y0 - RD(kY, 0, 1));
                bits |= STZ_D_NODE_ONESIDED;
            }
        }
        // X += cX, Y += cY (fields.hh:68-71)
        PL(w.X, i, j) = X0 + (-uc * Xx - vc * Xy);
        PL(w.Y, i, j) = Y0 + (-uc * Yx - vc * Yy);
    } else {
        // row 0 reference map is left to set_boundaries; keep it, and the
        // ghost row -1, in the new plane set
        PL(w.X, i, 0) = RD(kX, 0, 0);
        PL(w.Y, i, 0) = RD(kY, 0, 0);
        PL(w.X, i, -1) = RD(kX, 0, -1);
        PL(w.Y, i, -1) = RD(kY, 0, -1);
        PL(w.p, i, -1) = RD(kP, 0, -1);
        PL(w.q, i, -1) = RD(kQ, 0, -1);
        PL(w.s, i, -1) = RD(kS, 0, -1);
        PL(w.tau, i, -1) = RD(kTau, 0, -1);
        PL(w.chi, i, -1) = RD(kChi, 0, -1);
    }
    if (j == n - 1) {
        PL(w.X, i, n) = RD(kX, 0, 1);
        PL(w.Y, i, n) = RD(kY, 0, 1);
        PL(w.p, i, n) = RD(kP, 0, 1);
        PL(w.q, i, n) = RD(kQ, 0, 1);
        PL(w.s, i, n) = RD(kS, 0, 1);
        PL(w.tau, i, n) = RD(kTau, 0, 1);
        PL(w.chi, i, n) = RD(kChi, 0, 1);
    }

    // staggered velocity and spin, shear_sim.cc:330-333
    {
        const double u00 = RD(kU, 0, 0), u10 = RD(kU, 1, 0), u01 = RD(kU, 0, 1), u11 = RD(kU, 1, 1);
        const double v00 = RD(kV, 0, 0), v10 = RD(kV, 1, 0), v01 = RD(kV, 0, 1), v11 = RD(kV, 1, 1);
        uc = 0.25 * (u00 + u10 + u01 + u11);
        vc = 0.25 * (v00 + v10 + v01 + v11);
        two_omega = hx * (-v00 + v10 - v01 + v11) - hy * (-u00 - u10 + u01 + u11);
    }

    // ENO derivatives of p,q,s,tau,chi, shear_sim.cc:336-338
    double f0[5], f1[5], f2[5];
#pragma unroll
    for (int k = 0; k < 5; k++) f0[k] = RD(kP + k, 0, 0);
    if (UP(uc > 0, STZ_D_CELL_UPX)) {
        bits |= STZ_D_CELL_UPX;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            f1[k] = hx * eno2(RD(kP + k, 1, 0), f0[k], RD(kP + k, -1, 0), RD(kP + k, -2, 0), br,
                              RB(STZ_D_CELL_ENO_X0 + k));
            bits |= br << (STZ_D_CELL_ENO_X0 + k);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) {
            f1[k] = -hx * eno2(RD(kP + k, -1, 0), f0[k], RD(kP + k, 1, 0), RD(kP + k, 2, 0), br,
                               RB(STZ_D_CELL_ENO_X0 + k));
            bits |= br << (STZ_D_CELL_ENO_X0 + k);
        }
    }
    if (UP(vc > 0, STZ_D_CELL_UPY)) {
        bits |= STZ_D_CELL_UPY;
        if (j > 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) {
                f2[k] = hy * eno2(RD(kP + k, 0, 1), f0[k], RD(kP + k, 0, -1), RD(kP + k, 0, -2), br,
                                  RB(STZ_D_CELL_ENO_Y0 + k));
                bits |= br << (STZ_D_CELL_ENO_Y0 + k);
            }
        } else {
            const double hs = hy * 2;
#pragma unroll
            for (int k = 0; k < 5; k++) f2[k] = hs * (f0[k] - RD(kP + k, 0, -1));
            bits |= STZ_D_CELL_ONESIDED;
        }
    } else {
        if (j < n - 1) {
#pragma unroll
            for (int k = 0; k < 5; k++) {
                f2[k] = -hy * eno2(RD(kP + k, 0, -1), f0[k], RD(kP + k, 0, 1), RD(kP + k, 0, 2), br,
                                   RB(STZ_D_CELL_ENO_Y0 + k));
                bits |= br << (STZ_D_CELL_ENO_Y0 + k);
            }
        } else {
            const double hs = -hy * 2;
#pragma unroll
            for (int k = 0; k < 5; k++) f2[k] = hs * (f0[k] - RD(kP + k, 0, 1));
            bits |= STZ_D_CELL_ONESIDED;
        }
    }

    // chi advection and plasticity, shear_sim.cc:341-342; fields.hh:85-93
    double cchi = f0[4] - uc * f1[4] - vc * f2[4];
    const double sbar = sqrt(3 * f0[1] * f0[1] + f0[2] * f0[2] + f0[3] * f0[3]);
    double ddev;
    int trips;
    bool last_yield;
    if (REPLAY && (rb >> STZ_D_SUBSTEP_SHIFT) != 255u) {
        trips = (int)(rb >> STZ_D_SUBSTEP_SHIFT);
        last_yield = (rb & STZ_D_YIELD_LAST) != 0;
        dtmp = plastic_term_replay(c, sbar, cchi, ddev, dt, (rb & STZ_D_SMALLDEV) != 0, trips, last_yield);
    } else {
        dtmp = plastic_term(c, sbar, cchi, ddev, dt, trips, last_yield);
    }
    PL(w.ddev, i, j) = ddev;
    PL(w.chi, i, j) = cchi;

    // stress updates, shear_sim.cc:345-348, applied as in fields.hh:61-67
    PL(w.p, i, j) = f0[0] + (-uc * f1[0] - vc * f2[0]);
    PL(w.q, i, j) = f0[1] + (-uc * f1[1] - vc * f2[1] - dtmp * f0[1]);
    PL(w.s, i, j) = f0[2] + (-uc * f1[2] - vc * f2[2] + two_omega * f0[3] - dtmp * f0[2]);
    PL(w.tau, i, j) = f0[3] + (-uc * f1[3] - vc * f2[3] - two_omega * f0[2] - dtmp * f0[3]);

    if (dec) {
        const bool athermal = c.stz.kind == STZ_MODEL_LINEAR_ATHERMAL || c.stz.kind == STZ_MODEL_NONLINEAR_ATHERMAL;
        if (REPLAY ? (rb & STZ_D_SMALLDEV) != 0 : sbar < 1e-12) bits |= STZ_D_SMALLDEV;
        else if (athermal && (REPLAY ? (rb & STZ_D_YIELD) != 0 : !(sbar < 1))) bits |= STZ_D_YIELD;
        if (athermal && last_yield) bits |= STZ_D_YIELD_LAST;
        bits |= (uint32_t)(trips > 255 ? 255 : trips) << STZ_D_SUBSTEP_SHIFT;
        dec[i + m * j] = bits;
    }
#undef RB
#undef UP
#undef RD
}
// The dispatcher hands workgroups to the eight XCDs round robin (linear id
// mod 8), and each XCD has its own L2.  The row halo a 64x4 block reads (+-2
// rows for the advection's ENO stencils, one row for the source and the
// correction) is the interior of its vertical neighbours, so the block id is
// remapped to give each XCD one contiguous band of tile rows: a halo row is
// then read from HBM once per band rather than once per block.
__device__ __forceinline__ static void xcd_band_tile(int &bx, int &by) {
    const int gx = gridDim.x, total = gx * gridDim.y, b = blockIdx.x + gx * blockIdx.y;
    const int q = total >> 3, r = total & 7, x = b & 7, k = b >> 3;
    const int t = x * q + min(x, r) + k;
    bx = t % gx;
    by = t / gx;
}

__global__ void __launch_bounds__(256) advect_kernel(SimConst c, FieldPtrs o, FieldPtrs w, double dt,
                                                     uint32_t *dec) {
    int bx, by;
    xcd_band_tile(bx, by);
    advect_body<false>(c, o, w, dt, dec, nullptr, bx, by);
}
__global__ void __launch_bounds__(256) advect_replay_kernel(SimConst c, FieldPtrs o, FieldPtrs w, double dt,
                                                            uint32_t *dec, const uint32_t *rep) {
    int bx, by;
    xcd_band_tile(bx, by);
    advect_body<true>(c, o, w, dt, dec, rep, bx, by);
}

// ---------------------------------------------------------------------------
// direct (dynamic) step, shear_sim.cc:185-276: the advection kernel's work
// plus the velocity update (ENO2 advection of u,v, the stress-divergence
// force net_force(..,-hx,-hy,..) and viscosity), the compressive terms of
// the stress update, and constant chi diffusion inside the cell update.
// New u,v go to un,vn (rows 1..n-1; the caller copies them back), all other
// new fields to the plane set w.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) direct_kernel(SimConst c, FieldPtrs o, FieldPtrs w, double *un, double *vn,
                                                     double dt) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const double third = 1. / 3;
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt, lx = c.lx;
    const double hxxv = 2 * c.xsp * hx * c.viscosity, hyyv = 2 * c.ysp * hy * c.viscosity, hcv = 2 * (hxxv + hyyv);
    const double hxxcd = 2 * c.xsp * hx * c.chi_diff, hyycd = 2 * c.ysp * hy * c.chi_diff,
                 hccd = 2 * (hxxcd + hyycd);
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const int ill = i <= 1 ? i + m - 2 : i - 2, irr = i >= m - 2 ? i + 2 - m : i + 2;
    uint32_t br;
    double uc, vc, ux, vx, uy, vy, two_omega, dtmp;

    if (j > 0) {
        double Xx, Xy, Yx, Yy, fx, fy;
        uc = PL(o.u, i, j);
        vc = PL(o.v, i, j);
        const double X0 = PL(o.X, i, j), Y0 = PL(o.Y, i, j);
        // rmv_eno2 / rmv_one_sided, shear_sim.cc:203-219, 588-625
        if (uc > 0) {
            ux = hx * eno2(PL(o.u, ir, j), uc, PL(o.u, il, j), PL(o.u, ill, j), br);
            vx = hx * eno2(PL(o.v, ir, j), vc, PL(o.v, il, j), PL(o.v, ill, j), br);
            Xx = hx * eno2(PL(o.X, ir, j) + (i == m - 1 ? lx : 0), X0, PL(o.X, il, j) + (i == 0 ? -lx : 0),
                           PL(o.X, ill, j) + (i <= 1 ? -lx : 0), br);
            Yx = hx * eno2(PL(o.Y, ir, j), Y0, PL(o.Y, il, j), PL(o.Y, ill, j), br);
        } else {
            ux = -hx * eno2(PL(o.u, il, j), uc, PL(o.u, ir, j), PL(o.u, irr, j), br);
            vx = -hx * eno2(PL(o.v, il, j), vc, PL(o.v, ir, j), PL(o.v, irr, j), br);
            Xx = -hx * eno2(PL(o.X, il, j) + (i == 0 ? -lx : 0), X0, PL(o.X, ir, j) + (i == m - 1 ? lx : 0),
                            PL(o.X, irr, j) + (i >= m - 2 ? lx : 0), br);
            Yx = -hx * eno2(PL(o.Y, il, j), Y0, PL(o.Y, ir, j), PL(o.Y, irr, j), br);
        }
        if (vc > 0) {
            if (j > 1) {
                uy = hy * eno2(PL(o.u, i, j + 1), uc, PL(o.u, i, j - 1), PL(o.u, i, j - 2), br);
                vy = hy * eno2(PL(o.v, i, j + 1), vc, PL(o.v, i, j - 1), PL(o.v, i, j - 2), br);
                Xy = hy * eno2(PL(o.X, i, j + 1), X0, PL(o.X, i, j - 1), PL(o.X, i, j - 2), br);
                Yy = hy * eno2(PL(o.Y, i, j + 1), Y0, PL(o.Y, i, j - 1), PL(o.Y, i, j - 2), br);
            } else {
                uy = 2 * hy * (uc - PL(o.u, i, j - 1));
                vy = 2 * hy * (vc - PL(o.v, i, j - 1));
                const double hs = hy * 2;
                Xy = hs * (X0 - PL(o.X, i, j - 1));
                Yy = hs * (Y0 - PL(o.Y, i, j - 1));
            }
        } else {
            if (j < n - 1) {
                uy = -hy * eno2(PL(o.u, i, j - 1), uc, PL(o.u, i, j + 1), PL(o.u, i, j + 2), br);
                vy = -hy * eno2(PL(o.v, i, j - 1), vc, PL(o.v, i, j + 1), PL(o.v, i, j + 2), br);
                Xy = -hy * eno2(PL(o.X, i, j - 1), X0, PL(o.X, i, j + 1), PL(o.X, i, j + 2), br);
                Yy = -hy * eno2(PL(o.Y, i, j - 1), Y0, PL(o.Y, i, j + 1), PL(o.Y, i, j + 2), br);
            } else {
                uy = 2 * -hy * (uc - PL(o.u, i, j + 1));
                vy = 2 * -hy * (vc - PL(o.v, i, j + 1));
                const double hs = -hy * 2;
                Xy = hs * (X0 - PL(o.X, i, j + 1));
                Yy = hs * (Y0 - PL(o.Y, i, j + 1));
            }
        }
        // net_force(fx,fy,-hx,-hy,flp,fp), shear_sim.cc:222, 447-453
        {
            const double xf = -hx, yf = -hy;
            const double p0 = PL(o.p, il, j - 1), p1 = PL(o.p, i, j - 1), p2 = PL(o.p, il, j), p3 = PL(o.p, i, j);
            const double q0 = PL(o.q, il, j - 1), q1 = PL(o.q, i, j - 1), q2 = PL(o.q, il, j), q3 = PL(o.q, i, j);
            const double s0 = PL(o.s, il, j - 1), s1 = PL(o.s, i, j - 1), s2 = PL(o.s, il, j), s3 = PL(o.s, i, j);
            const double t0 = PL(o.tau, il, j - 1), t1 = PL(o.tau, i, j - 1), t2 = PL(o.tau, il, j),
                         t3 = PL(o.tau, i, j);
            fx = (-p0 + p1 - p2 + p3 - q0 + q1 - q2 + q3 + s0 - s1 + s2 - s3) * xf + (t0 + t1 - t2 - t3) * yf;
            fy = (t0 - t1 + t2 - t3) * xf + (-p0 - p1 + p2 + p3 - q0 - q1 + q2 + q3 - s0 - s1 + s2 + s3) * yf;
        }
        const double uvisc = hxxv * (PL(o.u, ir, j) + PL(o.u, il, j)) + hyyv * (PL(o.u, i, j + 1) + PL(o.u, i, j - 1)) -
                             hcv * uc;
        const double vvisc = hxxv * (PL(o.v, ir, j) + PL(o.v, il, j)) + hyyv * (PL(o.v, i, j + 1) + PL(o.v, i, j - 1)) -
                             hcv * vc;
        // update() (fields.hh:56-60): u += cu, v += cv, X += cX, Y += cY
        PL(un, i, j) = uc + (-uc * ux - vc * uy + c.mu_inv * (fx + uvisc));
        PL(vn, i, j) = vc + (-uc * vx - vc * vy + c.mu_inv * (fy + vvisc));
        PL(w.X, i, j) = X0 + (-uc * Xx - vc * Xy);
        PL(w.Y, i, j) = Y0 + (-uc * Yx - vc * Yy);
    } else {
        // row 0 keeps u, v, X, Y until set_boundaries; ghost row -1 carried over
        PL(w.X, i, 0) = PL(o.X, i, 0);
        PL(w.Y, i, 0) = PL(o.Y, i, 0);
        PL(w.X, i, -1) = PL(o.X, i, -1);
        PL(w.Y, i, -1) = PL(o.Y, i, -1);
        PL(w.p, i, -1) = PL(o.p, i, -1);
        PL(w.q, i, -1) = PL(o.q, i, -1);
        PL(w.s, i, -1) = PL(o.s, i, -1);
        PL(w.tau, i, -1) = PL(o.tau, i, -1);
        PL(w.chi, i, -1) = PL(o.chi, i, -1);
    }
    if (j == n - 1) {
        PL(w.X, i, n) = PL(o.X, i, n);
        PL(w.Y, i, n) = PL(o.Y, i, n);
        PL(w.p, i, n) = PL(o.p, i, n);
        PL(w.q, i, n) = PL(o.q, i, n);
        PL(w.s, i, n) = PL(o.s, i, n);
        PL(w.tau, i, n) = PL(o.tau, i, n);
        PL(w.chi, i, n) = PL(o.chi, i, n);
    }

    // deformation rate, staggered velocity and spin, shear_sim.cc:240-247
    {
        const double u00 = PL(o.u, i, j), u10 = PL(o.u, ir, j), u01 = PL(o.u, i, j + 1), u11 = PL(o.u, ir, j + 1);
        const double v00 = PL(o.v, i, j), v10 = PL(o.v, ir, j), v01 = PL(o.v, i, j + 1), v11 = PL(o.v, ir, j + 1);
        uc = 0.25 * (u00 + u10 + u01 + u11);
        vc = 0.25 * (v00 + v10 + v01 + v11);
        ux = hx * (-u00 + u10 - u01 + u11);
        uy = hy * (-u00 - u10 + u01 + u11);
        vx = hx * (-v00 + v10 - v01 + v11);
        vy = hy * (-v00 - v10 + v01 + v11);
        two_omega = vx - uy;
    }

    // ENO derivatives of p,q,s,tau,chi, shear_sim.cc:250-252
    const double *F[5] = {o.p, o.q, o.s, o.tau, o.chi};
    double f0[5], f1[5], f2[5];
#pragma unroll
    for (int k = 0; k < 5; k++) f0[k] = PL(F[k], i, j);
    if (uc > 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) f1[k] = hx * eno2(PL(F[k], ir, j), f0[k], PL(F[k], il, j), PL(F[k], ill, j), br);
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) f1[k] = -hx * eno2(PL(F[k], il, j), f0[k], PL(F[k], ir, j), PL(F[k], irr, j), br);
    }
    if (vc > 0) {
        if (j > 0) {
#pragma unroll
            for (int k = 0; k < 5; k++)
                f2[k] = hy * eno2(PL(F[k], i, j + 1), f0[k], PL(F[k], i, j - 1), PL(F[k], i, j - 2), br);
        } else {
            const double hs = hy * 2;
#pragma unroll
            for (int k = 0; k < 5; k++) f2[k] = hs * (f0[k] - PL(F[k], i, j - 1));
        }
    } else {
        if (j < n - 1) {
#pragma unroll
            for (int k = 0; k < 5; k++)
                f2[k] = -hy * eno2(PL(F[k], i, j - 1), f0[k], PL(F[k], i, j + 1), PL(F[k], i, j + 2), br);
        } else {
            const double hs = -hy * 2;
#pragma unroll
            for (int k = 0; k < 5; k++) f2[k] = hs * (f0[k] - PL(F[k], i, j + 1));
        }
    }

    // chi with its constant diffusion, then plasticity, shear_sim.cc:256-257
    double cchi = f0[4] - uc * f1[4] - vc * f2[4] + hxxcd * (PL(o.chi, ir, j) + PL(o.chi, il, j)) +
                  hyycd * (PL(o.chi, i, j + 1) + PL(o.chi, i, j - 1)) - hccd * f0[4];
    const double sbar = sqrt(3 * f0[1] * f0[1] + f0[2] * f0[2] + f0[3] * f0[3]);
    double ddev;
    int trips;
    bool last_yield;
    dtmp = plastic_term(c, sbar, cchi, ddev, dt, trips, last_yield);
    PL(w.chi, i, j) = cchi;
    // f.ddev *= f.dev() (:265), with the stresses not yet updated
    PL(w.ddev, i, j) = ddev * sbar;

    // stress updates, shear_sim.cc:260-263, applied as in fields.hh:61-67
    PL(w.p, i, j) = f0[0] + (-uc * f1[0] - vc * f2[0] - c.K * (ux + vy));
    PL(w.q, i, j) = f0[1] + (-uc * f1[1] - vc * f2[1] - c.mu * third * (ux + vy) - dtmp * f0[1]);
    PL(w.s, i, j) = f0[2] + (-uc * f1[2] - vc * f2[2] + two_omega * f0[3] + c.mu * (ux - vy) - dtmp * f0[2]);
    PL(w.tau, i, j) = f0[3] + (-uc * f1[3] - vc * f2[3] - two_omega * f0[2] + c.mu * (vx + uy) - dtmp * f0[3]);
}

// chi_diffusion, shear_sim.cc:364-386: w.chi holds the post-plastic chi
// (cchi) and receives the diffused value; o.chi is the previous chi.
__device__ __forceinline__ static void chi_diffusion_body(const SimConst &c, const double *__restrict__ chi_old,
                                                          const double *__restrict__ ddev, double *chi_new,
                                                          double dt, int bx, int by) {
    const int m = c.m, n = c.n;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    const double hxxcd = 2 * c.xsp * hx * c.chi_diff, hyycd = 2 * c.ysp * hy * c.chi_diff,
                 hccd = 2 * (hxxcd + hyycd);
    const double dfac = 0.25 * c.chi_len * c.chi_len * c.mu_inv, dfacx = dfac * c.xsp * c.xsp,
                 dfacy = dfac * c.ysp * c.ysp;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const double D0 = PL(ddev, i, j);
    const double Dd = D0 + PL(ddev, i, j - 1), Dl = D0 + PL(ddev, il, j), Dr = D0 + PL(ddev, ir, j),
                 Du = D0 + PL(ddev, i, j + 1);
    const double cf = PL(chi_old, i, j), cl = PL(chi_old, il, j), cr = PL(chi_old, ir, j),
                 cd = PL(chi_old, i, j - 1), cu = PL(chi_old, i, j + 1);
    PL(chi_new, i, j) += hxxcd * (cr + cl) + hyycd * (cu + cd) - hccd * cf
                         + dfacx * (Dr * cr - (Dr + Dl) * cf + Dl * cl)
                         + dfacy * (Du * cu - (Du + Dd) * cf + Dd * cd);
}
__global__ void __launch_bounds__(256) chi_diffusion_kernel(SimConst c, const double *__restrict__ chi_old,
                                                            const double *__restrict__ ddev, double *chi_new,
                                                            double dt) {
    chi_diffusion_body(c, chi_old, ddev, chi_new, dt, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// projection source b = div(sigma)/dt and wall rows, shear_sim.cc:399-409,
// net_force :447-453.  One thread per node (i,j), j in [0,n].
// ---------------------------------------------------------------------------
__device__ __forceinline__ static void source_body(const SimConst &c, const FieldPtrs &f, double xfac,
                                                   double yfac, double xbc, double bvel, double *src, int bx,
                                                   int by) {
    const int m = c.m, n = c.n;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    if (i >= m || j > n) return;
    double fx, fy;
    if (j == 0) {
        fx = -xbc * bvel;
        fy = 0;
    } else if (j == n) {
        fx = xbc * bvel;
        fy = 0;
    } else {
        const int il = i == 0 ? m - 1 : i - 1;
        // f0=(i-1,j-1) f1=(i,j-1) f2=(i-1,j) f3=(i,j)
        const double p0 = PL(f.p, il, j - 1), p1 = PL(f.p, i, j - 1), p2 = PL(f.p, il, j), p3 = PL(f.p, i, j);
        const double q0 = PL(f.q, il, j - 1), q1 = PL(f.q, i, j - 1), q2 = PL(f.q, il, j), q3 = PL(f.q, i, j);
        const double s0 = PL(f.s, il, j - 1), s1 = PL(f.s, i, j - 1), s2 = PL(f.s, il, j), s3 = PL(f.s, i, j);
        const double t0 = PL(f.tau, il, j - 1), t1 = PL(f.tau, i, j - 1), t2 = PL(f.tau, il, j),
                     t3 = PL(f.tau, i, j);
        fx = (-p0 + p1 - p2 + p3 - q0 + q1 - q2 + q3 + s0 - s1 + s2 - s3) * xfac + (t0 + t1 - t2 - t3) * yfac;
        fy = (t0 - t1 + t2 - t3) * xfac + (-p0 - p1 + p2 + p3 - q0 - q1 + q2 + q3 - s0 - s1 + s2 + s3) * yfac;
    }
    const int ij = i + m * j;
    src[2 * ij] = fx;
    src[2 * ij + 1] = fy;
}
__global__ void __launch_bounds__(256) source_kernel(SimConst c, FieldPtrs f, double xfac, double yfac,
                                                     double xbc, double bvel, double *src) {
    int bx, by;
    xcd_band_tile(bx, by);
    source_body(c, f, xfac, yfac, xbc, bvel, src, bx, by);
}

// ---------------------------------------------------------------------------
// stress correction from the solved velocity, shear_sim.cc:418-434, fused
// with set_boundaries (shear_sim.cc:552-581).  Rows 0/1 are done by the row-0
// threads and rows n-1/n-2 by the row-(n-1) threads, so that the ghost-row
// extrapolation reads corrected values without a second kernel (n >= 4).
// ---------------------------------------------------------------------------
__device__ __forceinline__ static void correct_cell(const SimConst &c, const FieldPtrs &f, const double *vel,
                                                    int i, int j, double hx, double hy, double &p, double &q,
                                                    double &s, double &t) {
    const int m = c.m;
    const double third = 1 / 3.0;
    const int io = i == m - 1 ? 0 : i + 1;
    const double *vp = vel + 2 * (i + m * j), *vo = vel + 2 * (io + m * j);
    const double vpx = vp[0], vpy = vp[1], vox = vo[0], voy = vo[1];
    const double vpux = vp[2 * m], vpuy = vp[2 * m + 1], voux = vo[2 * m], vouy = vo[2 * m + 1];
    const double ux = hx * (-vpx + vox - vpux + voux), uy = hy * (-vpx - vox + vpux + voux),
                 vx = hx * (-vpy + voy - vpuy + vouy), vy = hy * (-vpy - voy + vpuy + vouy);
    p = PL(f.p, i, j) + -c.K * (ux + vy);
    q = PL(f.q, i, j) + -c.mu * third * (ux + vy);
    s = PL(f.s, i, j) + c.mu * (ux - vy);
    t = PL(f.tau, i, j) + c.mu * (uy + vx);
    PL(f.u, i, j) = vpx;
    PL(f.v, i, j) = vpy;
}

__device__ __forceinline__ static void correct_body(const SimConst &c, const FieldPtrs &f, const double *vel,
                                                    double dt, double bvel, double bpos, int bx, int by) {
    const int m = c.m, n = c.n;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    if (j == 1 || j == n - 2) return;  // done by the wall-row threads
    const double hx = 0.5 * c.xsp * dt, hy = 0.5 * c.ysp * dt;
    double p, q, s, t;
    correct_cell(c, f, vel, i, j, hx, hy, p, q, s, t);
    if (j == 0 || j == n - 1) {
        const int j1 = j == 0 ? 1 : n - 2, jg = j == 0 ? -1 : n;
        double p1, q1, s1, t1;
        correct_cell(c, f, vel, i, j1, hx, hy, p1, q1, s1, t1);
        PL(f.p, i, j1) = p1;
        PL(f.q, i, j1) = q1;
        PL(f.s, i, j1) = s1;
        PL(f.tau, i, j1) = t1;
        PL(f.p, i, jg) = 2 * p - p1;
        PL(f.q, i, jg) = 2 * q - q1;
        PL(f.s, i, jg) = 2 * s - s1;
        PL(f.tau, i, jg) = 2 * t - t1;
        PL(f.chi, i, jg) = 2 * PL(f.chi, i, j) - PL(f.chi, i, j1);
        const int jw = j == 0 ? 0 : n;
        PL(f.u, i, jw) = j == 0 ? -bvel : bvel;
        PL(f.v, i, jw) = 0;
        PL(f.X, i, jw) = j == 0 ? c.ax + i * c.dx + bpos : c.ax + i * c.dx - bpos;
        // both walls' y in registers before the pick: a select of two
        // loads from the argument block is otherwise folded into one load
        // through a selected address, which costs a private (scratch) copy
        // of the two values in every thread of the launch
        double ya = c.ay, yb = c.by;
        asm volatile("" : "+v"(ya), "+v"(yb));
        PL(f.Y, i, jw) = j == 0 ? ya : yb;
    }
    PL(f.p, i, j) = p;
    PL(f.q, i, j) = q;
    PL(f.s, i, j) = s;
    PL(f.tau, i, j) = t;
}
__global__ void __launch_bounds__(256) correct_kernel(SimConst c, FieldPtrs f, const double *vel, double dt,
                                                      double bvel, double bpos) {
    int bx, by;
    xcd_band_tile(bx, by);
    correct_body(c, f, vel, dt, bvel, bpos, bx, by);
}

// set_boundaries on its own, shear_sim.cc:552-581
__global__ void boundaries_kernel(SimConst c, FieldPtrs f, double bvel, double bpos) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    PL(f.u, i, 0) = -bvel;
    PL(f.v, i, 0) = 0;
    PL(f.X, i, 0) = c.ax + i * c.dx + bpos;
    PL(f.Y, i, 0) = c.ay;
    double *F[5] = {f.p, f.q, f.s, f.tau, f.chi};
#pragma unroll
    for (int k = 0; k < 5; k++) PL(F[k], i, -1) = 2 * PL(F[k], i, 0) - PL(F[k], i, 1);
    PL(f.u, i, n) = bvel;
    PL(f.v, i, n) = 0;
    PL(f.X, i, n) = c.ax + i * c.dx - bpos;
    PL(f.Y, i, n) = c.by;
#pragma unroll
    for (int k = 0; k < 5; k++) PL(F[k], i, n) = 2 * PL(F[k], i, n - 1) - PL(F[k], i, n - 2);
}

// shear_sim.cc:60-114
__global__ void init_fields_kernel(SimConst c, FieldPtrs f, int chi_case, double chi0, double chi1) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y - 1;
    if (i >= m || j > n) return;
    const double pi = 3.1415926535897932384626433832795;
    double Y = c.ay + j * c.dy, Ys = Y + 0.5 * c.dy;
    double X = c.ax + i * c.dx, Xs = X + 0.5 * c.dx, chi;
    PL(f.u, i, j) = 0;
    PL(f.v, i, j) = 0;
    PL(f.ddev, i, j) = 0;
    PL(f.p, i, j) = 0;
    PL(f.q, i, j) = 0;
    PL(f.s, i, j) = 0;
    PL(f.tau, i, j) = 0;
    if (chi_case == 0) chi = chi0 + chi1 * exp(-200 * (Xs * Xs + Ys * Ys));
    else if (chi_case == 1) {
        const double csa = cos(30 * pi / 180.), sia = sin(30 * pi / 180.);
        double Xr = csa * Xs + sia * Ys, Yr = -sia * Xs + csa * Ys;
        Xr = Xr < -1 ? Xr + 1 : (Xr > 1 ? Xr - 1 : 0);
        chi = chi0 + chi1 * exp(-200 * (Xr * Xr + Yr * Yr));
    } else {
        double R = 1e100, Rt, Xr, Yr;
        for (int k = 0; k < 16; k++) {
            double sxk = c.ax + c.lx * (k + 0.5) * 0.0625, syk = -0.2 * sin(pi * 0.0625 * (2 * k + 1));
            Xr = sxk - Xs;
            if (Xr < -c.lx * 0.5) Xr += c.lx;
            if (Xr > c.lx * 0.5) Xr -= c.lx;
            Yr = syk - Ys;
            Rt = Xr * Xr + Yr * Yr;
            if (Rt < R) R = Rt;
        }
        chi = chi0 + chi1 * exp(-200 * R);
    }
    PL(f.chi, i, j) = chi;
    PL(f.X, i, j) = X;
    PL(f.Y, i, j) = Y;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

int launch_advect(const SimConst &c, const FieldPtrs &o, const FieldPtrs &w, double dt, uint32_t *dec,
                  hipStream_t st, const uint32_t *rep) {
    dim3 b(64, 4);
    if (rep) hipLaunchKernelGGL(advect_replay_kernel, grid2(c.m, c.n, b), b, 0, st, c, o, w, dt, dec, rep);
    else hipLaunchKernelGGL(advect_kernel, grid2(c.m, c.n, b), b, 0, st, c, o, w, dt, dec);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_direct(const SimConst &c, const FieldPtrs &o, const FieldPtrs &w, double *un, double *vn, double dt,
                  hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(direct_kernel, grid2(c.m, c.n, b), b, 0, st, c, o, w, un, vn, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_chi_diffusion(const SimConst &c, const FieldPtrs &o, const FieldPtrs &w, double dt, hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(chi_diffusion_kernel, grid2(c.m, c.n, b), b, 0, st, c, o.chi, w.ddev, w.chi, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_source(const SimConst &c, const FieldPtrs &f, double dt, double xbc, double bvel, double *src,
                  hipStream_t st) {
    dim3 b(64, 4);
    const double xfac = 0.5 / dt * c.xsp, yfac = 0.5 / dt * c.ysp;
    hipLaunchKernelGGL(source_kernel, grid2(c.m, c.n + 1, b), b, 0, st, c, f, xfac, yfac, xbc, bvel, src);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_correct(const SimConst &c, const FieldPtrs &f, const double *vel, double dt, double bvel, double bpos,
                   hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(correct_kernel, grid2(c.m, c.n, b), b, 0, st, c, f, vel, dt, bvel, bpos);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_boundaries(const SimConst &c, const FieldPtrs &f, double bvel, double bpos, hipStream_t st) {
    hipLaunchKernelGGL(boundaries_kernel, dim3((c.m + 255) / 256), dim3(256), 0, st, c, f, bvel, bpos);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int launch_init_fields(const SimConst &c, const FieldPtrs &f, int chi_case, double chi0, double chi1,
                       hipStream_t st) {
    dim3 b(64, 4);
    hipLaunchKernelGGL(init_fields_kernel, grid2(c.m, c.n + 2, b), b, 0, st, c, f, chi_case, chi0, chi1);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// ---------------------------------------------------------------------------
// batched launches over the members of an ensemble (batch.h): the same
// bodies, one member per workgroup range; the wall velocity and position
// come from the member's device-side time, evaluated as the host evaluates
// bdry_vel / bdry_pos (shear_sim.hh:120-124)
// ---------------------------------------------------------------------------
__device__ __forceinline__ static double ens_bvel(const SimConst &c, double t) {
    return t < 1 ? t * c.u_bdry : c.u_bdry;
}
__device__ __forceinline__ static double ens_bpos(const SimConst &c, double t) {
    return t < 1 ? c.u_bdry * t * t * 0.5 : c.u_bdry * (t - 0.5);
}

__global__ void __launch_bounds__(256) time_batch_kernel(EnsState *S, int nmem, double dt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nmem) S[k].time += dt;  // shear_sim.cc:283
}

__global__ void __launch_bounds__(256) advect_batch_kernel(const CellEntry *E, const int *map, int ne, int parity,
                                                           double dt) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    // local copies: the body's stores cannot alias them, so each field is
    // loaded once
    const CellEntry &x = E[e];
    const SimConst c = x.c;
    const FieldPtrs o = x.f[parity], w = x.f[parity ^ 1];
    const int gx = x.gx;
    advect_body<false>(c, o, w, dt, x.dec, nullptr, l % gx, l / gx);
}

__global__ void __launch_bounds__(256) chi_diffusion_batch_kernel(const CellEntry *E, const int *map, int ne,
                                                                  int parity, double dt) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const CellEntry &x = E[e];
    const SimConst c = x.c;
    const FieldPtrs o = x.f[parity], w = x.f[parity ^ 1];
    const int gx = x.gx;
    chi_diffusion_body(c, o.chi, w.ddev, w.chi, dt, l % gx, l / gx);
}

__global__ void __launch_bounds__(256) source_batch_kernel(const CellEntry *E, const int *map, int ne, int parity,
                                                           const EnsState *S, double dt) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const CellEntry &x = E[e];
    const SimConst c = x.c;
    const FieldPtrs f = x.f[parity];
    const int gx = x.gx;
    const double xfac = 0.5 / dt * c.xsp, yfac = 0.5 / dt * c.ysp;
    source_body(c, f, xfac, yfac, x.xbc, ens_bvel(c, S[x.member].time), x.src, l % gx, l / gx);
}

__global__ void __launch_bounds__(256) correct_batch_kernel(const CellEntry *E, const int *map, int ne, int parity,
                                                            const EnsState *S, double dt) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const CellEntry &x = E[e];
    const EnsState &st = S[x.member];
    if (st.phase != ENS_DONE) return;
    const SimConst c = x.c;
    const FieldPtrs f = x.f[parity];
    const int gx = x.gx;
    const double t = st.time;
    correct_body(c, f, x.vel, dt, ens_bvel(c, t), ens_bpos(c, t), l % gx, l / gx);
}

int batch_time(EnsState *S, int nmem, double dt, hipStream_t st) {
    hipLaunchKernelGGL(time_batch_kernel, dim3((nmem + 255) / 256), dim3(256), 0, st, S, nmem, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int batch_advect(const CellEntry *E, const int *map, int ne, int total, int parity, double dt, hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(advect_batch_kernel, dim3(total), dim3(64, 4), 0, st, E, map, ne, parity, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int batch_chi_diffusion(const CellEntry *E, const int *map, int ne, int total, int parity, double dt,
                        hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(chi_diffusion_batch_kernel, dim3(total), dim3(64, 4), 0, st, E, map, ne, parity, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int batch_source(const CellEntry *E, const int *map, int ne, int total, int parity, const EnsState *S, double dt,
                 hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(source_batch_kernel, dim3(total), dim3(64, 4), 0, st, E, map, ne, parity, S, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int batch_correct(const CellEntry *E, const int *map, int ne, int total, int parity, const EnsState *S, double dt,
                  hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(correct_batch_kernel, dim3(total), dim3(64, 4), 0, st, E, map, ne, parity, S, dt);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

}  // namespace stz
