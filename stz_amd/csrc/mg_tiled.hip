// mg_tiled.hip -- LDS-tiled multigrid kernels for the large levels.
//
// smooth_tile_kernel: one workgroup per TX x TY tile.  The tile plus a halo
//   of H = 4*S nodes is loaded once into LDS (z) and registers (b); S
//   four-colour Gauss-Seidel sweeps then run in LDS.  A node of the colour
//   being updated is correct when its eight neighbours are, so each colour
//   pass spoils one more ring at the halo edge; after 4*S passes exactly the
//   interior is still exact, and only it is written back.  S sweeps therefore
//   cost one read and one write of z and one read of b, instead of 4*S
//   kernel launches each reading all of z.  Each thread owns 2x2 blocks of
//   nodes (one node of every colour) and keeps their b values in registers.
//   Rows away from the walls all share one stencil, which every thread keeps
//   in registers; only the few wall-adjacent rows read their stencil from
//   LDS.  Optionally the kernel starts from z = 0 (coarse levels, correction
//   equation) or adds the bilinear prolongation of the coarser level's z
//   while loading (fused coarse-grid correction, tgmg.in.cc:552-621).
// restrict_tile_kernel: one workgroup per CX x CY coarse tile; computes every
//   fine residual it needs exactly once into LDS, then the P^T-weighted sums
//   (tgmg.in.cc:626-647).
// refine_kernel: the FP64 half of one iterative-refinement step of the mixed
//   form: r = src - A vel in FP64 (one-node halo in LDS), r -> FP32 source of
//   the next correction equation, |r|^2 partials.
// Tiled kernels never write the array they read halos from: smoothing reads
// z and writes a second buffer (the caller swaps them).
//
// The colour order inside a sweep, (0,0),(1,0),(0,1),(1,1) in (i%2, j%2), is
// the one the per-colour kernels of mg.hip use, so both paths apply the same
// smoother.
#include <utility>

#include "batch.h"
#include "mg_node.h"

namespace stz {

enum { TILE_LOAD = 0, TILE_ZERO = 1, TILE_PROLONG = 2 };

// Tile geometry.  The halo is 4S nodes for S fused sweeps, two more when the
// kernel also restricts (the residual ring around the tile needs z valid one
// node further out).  Each thread owns one 2x2 block of the tile-plus-halo
// region (two when that would exceed 1024 threads); in the 64-wide tiles
// each thread owns two blocks, which shortens the colour passes (two
// independent updates per thread between barriers): 11 % less time per
// launch than one block per thread at 1024^2, while three or four blocks
// per thread are slower again.
__host__ __device__ constexpr int tile_halo(int S, bool fuse) { return 4 * S + (fuse ? 2 : 0); }
__host__ __device__ constexpr int tile_blocks(int TX, int TY, int H) { return ((TX + 2 * H) / 2) * ((TY + 2 * H) / 2); }
__host__ __device__ constexpr int tile_bpt(int TX, int TY, int H) {
    return (tile_blocks(TX, TY, H) + 1023) / 1024 < (TX >= 64 ? 2 : 1) ? (TX >= 64 ? 2 : 1)
                                                                       : (tile_blocks(TX, TY, H) + 1023) / 1024;
}
__host__ __device__ constexpr int tile_threads(int TX, int TY, int H) {
    return ((tile_blocks(TX, TY, H) + tile_bpt(TX, TY, H) - 1) / tile_bpt(TX, TY, H) + 63) / 64 * 64;
}

// the eight neighbours of block blk's node of colour C in colour-split LDS:
// x-neighbours have colour cX, y-neighbours cY, diagonal ones cD, at block
// offsets that depend only on the colour
template <typename T, int BX, int NB, int C>
__device__ __forceinline__ static void tile_neighbours(const vec2s<T> *zs, int blk, vec2s<T> &dl, vec2s<T> &dc,
                                                       vec2s<T> &dr, vec2s<T> &cl, vec2s<T> &cr, vec2s<T> &ul,
                                                       vec2s<T> &uc, vec2s<T> &ur) {
    constexpr int cx = C & 1, cy = C >> 1;
    constexpr int cX = (1 - cx) + 2 * cy, cY = cx + 2 * (1 - cy), cD = (1 - cx) + 2 * (1 - cy);
    constexpr int xl = cx ? 0 : -1, xr = cx ? 1 : 0, yd = cy ? 0 : -BX, yu = cy ? BX : 0;
    const vec2s<T> *zX = zs + cX * NB + blk, *zY = zs + cY * NB + blk, *zD = zs + cD * NB + blk;
    dl = zD[yd + xl];
    dc = zY[yd];
    dr = zD[yd + xr];
    cl = zX[xl];
    cr = zX[xr];
    ul = zD[yu + xl];
    uc = zY[yu];
    ur = zD[yu + xr];
}

// the same, with the three neighbours that lie in the thread's own 2x2
// block (offset 0) taken from `own`, the block's current values, which the
// thread keeps in registers: five LDS reads per node instead of eight
template <int OFF, typename T>
__device__ __forceinline__ static vec2s<T> near_or_own(const vec2s<T> *p, const vec2s<T> &own) {
    if constexpr (OFF == 0) return own;
    else return p[OFF];
}

template <typename T, int BX, int NB, int C>
__device__ __forceinline__ static void tile_neighbours_own(const vec2s<T> *zs, int blk, const vec2s<T> (&own)[4],
                                                           vec2s<T> &dl, vec2s<T> &dc, vec2s<T> &dr, vec2s<T> &cl,
                                                           vec2s<T> &cr, vec2s<T> &ul, vec2s<T> &uc, vec2s<T> &ur) {
    constexpr int cx = C & 1, cy = C >> 1;
    constexpr int cX = (1 - cx) + 2 * cy, cY = cx + 2 * (1 - cy), cD = (1 - cx) + 2 * (1 - cy);
    constexpr int xl = cx ? 0 : -1, xr = cx ? 1 : 0, yd = cy ? 0 : -BX, yu = cy ? BX : 0;
    const vec2s<T> *zX = zs + cX * NB + blk, *zY = zs + cY * NB + blk, *zD = zs + cD * NB + blk;
    dl = near_or_own<yd + xl>(zD, own[cD]);
    dc = near_or_own<yd>(zY, own[cY]);
    dr = near_or_own<yd + xr>(zD, own[cD]);
    cl = near_or_own<xl>(zX, own[cX]);
    cr = near_or_own<xr>(zX, own[cX]);
    ul = near_or_own<yu + xl>(zD, own[cD]);
    uc = near_or_own<yu>(zY, own[cY]);
    ur = near_or_own<yu + xr>(zD, own[cD]);
}

// Wall rows (j < nb or j >= n - nb, nb <= kMaxWall) have stencils of their
// own; every other row has the interior one.  A tile keeps the 2 nb wall
// stencils in LDS, wherever it lies.
constexpr int kMaxWall = 2;

template <typename T>
__device__ __forceinline__ static void load_wall_stencils(T *ss, const LevelArgs<T> &L, int tid, int nt) {
    // the host launches tiled kernels only when wall_ok(L); the bound keeps
    // the stores inside ss regardless
    const int cnt = 2 * (L.nb < kMaxWall ? L.nb : kMaxWall) * 40;
    for (int t = tid; t < cnt; t += nt) {
        const int r = t / 40, k = t - r * 40, j = r < L.nb ? r : L.n - 2 * L.nb + r;
        ss[t] = L.S[(size_t)j * 40 + k];
    }
}

// the wall stencil of row gj (gj < nb or gj >= n - nb)
template <typename T> __device__ __forceinline__ static const T *wall_stencil(const T *ss, int n, int nb, int gj) {
    return ss + (gj < nb ? gj : nb + gj - (n - nb)) * 40;
}

// S four-colour Gauss-Seidel sweeps of one tile.  With FUSE the kernel then
// also forms the residual on the tile and its first ring and restricts it to
// the coarse nodes whose centre lies in the tile (R = P^T, full weighting in
// x, tgmg.in.cc:626-647), writing Cr's source (and zeroing its z when
// zero_z), so the V-cycle needs no separate restriction launch.
template <typename T, int TX, int TY, int S, int MODE, bool FUSE, bool SP>
__device__ __forceinline__ static void smooth_tile_body(const LevelArgs<T> &L, T *__restrict__ zout,
                                                        const LevelArgs<T> &Cz, const LevelArgs<T> &Cr, int zero_z,
                                                        double *corr, int tbx, int tby) {
    constexpr int H = tile_halo(S, FUSE), RX = TX + 2 * H, RY = TY + 2 * H;
    constexpr int BX = RX / 2, BY = RY / 2, NB = BX * BY, NT = tile_threads(TX, TY, H);
    constexpr int BPT = (NB + NT - 1) / NT;
    static_assert(RX % 2 == 0 && RY % 2 == 0, "even region");
    // colour-split LDS: node (li,lj) lives in array c = (li&1) + 2(lj&1) at
    // block index (lj>>1)*BX + (li>>1); a colour pass then reads each
    // neighbour from another colour's array at consecutive addresses across
    // the lanes (no bank conflicts), with compile-time offsets per colour
    __shared__ vec2s<T> zs[4 * NB];
    __shared__ T ss[2 * kMaxWall * 40];
    const int tid = threadIdx.x;
    const int m = L.m, n = L.n, nb = L.nb;
    const int i0 = tbx * TX - H, j0 = tby * TY - H;

    // interior stencil: wave-uniform, held in registers
    T si[40];
#pragma unroll
    for (int k = 0; k < 40; k++) si[k] = L.Si[k];
    // wall rows' stencils, only if this tile reaches them
    const bool walls = j0 < nb || j0 + RY > n - nb;
    if (walls) load_wall_stencils(ss, L, tid, NT);
    // z into LDS (zero outside the level) and into the owner's registers,
    // b into registers
    T bx[BPT][4], by[BPT][4];
    vec2s<T> zo[BPT][4];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int blk = tid + q * NT;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            bx[q][c] = 0;
            by[q][c] = 0;
        }
        if (blk >= NB) continue;
        const int bxi = blk % BX, byi = blk / BX;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int li = 2 * bxi + (c & 1), lj = 2 * byi + (c >> 1);
            const int gj = j0 + lj;
            int gi = i0 + li;
            gi = gi < 0 ? gi + m : (gi >= m ? gi - m : gi);
            T zx = 0, zy = 0;
            if (gj >= 0 && gj < n) {
                const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
                bx[q][c] = L.b[k];
                by[q][c] = L.b[k + 1];
                if (MODE == TILE_LOAD || MODE == TILE_PROLONG) {
                    zx = L.z[k];
                    zy = L.z[k + 1];
                }
                if (MODE == TILE_PROLONG) {
                    // z += P zc at fine node (gi, gj); Cz is the coarser level
                    const int sc = Cz.m;
                    int cx0, cx1, cy0, cy1;
                    T wx0, wx1, wy0, wy1;
                    if ((gi & 1) == 0) { cx0 = gi >> 1; cx1 = cx0; wx0 = 1; wx1 = 0; }
                    else { cx0 = gi >> 1; cx1 = cx0 + 1 == sc ? 0 : cx0 + 1; wx0 = wx1 = (T)0.5; }
                    if ((n & 1) == 0 && gj == n - 1) { cy0 = cy1 = Cz.n - 1; wy0 = 1; wy1 = 0; }
                    else if ((gj & 1) == 0) { cy0 = cy1 = gj >> 1; wy0 = 1; wy1 = 0; }
                    else { cy0 = gj >> 1; cy1 = cy0 + 1; wy0 = wy1 = (T)0.5; }
                    const T *z0 = Cz.z + 2 * (size_t)sc * cy0, *z1 = Cz.z + 2 * (size_t)sc * cy1;
                    zx += wy0 * (wx0 * z0[2 * cx0] + wx1 * z0[2 * cx1]) +
                          wy1 * (wx0 * z1[2 * cx0] + wx1 * z1[2 * cx1]);
                    zy += wy0 * (wx0 * z0[2 * cx0 + 1] + wx1 * z0[2 * cx1 + 1]) +
                          wy1 * (wx0 * z1[2 * cx0 + 1] + wx1 * z1[2 * cx1 + 1]);
                }
            }
            zs[c * NB + blk].x = zx;
            zs[c * NB + blk].y = zy;
            zo[q][c] = vec2s<T>{zx, zy};
        }
    }
    __syncthreads();

    // S sweeps of four colour passes
#pragma unroll 1
    for (int sw = 0; sw < S; sw++) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
#pragma unroll
            for (int q = 0; q < BPT; q++) {
                const int blk = tid + q * NT;
                if (blk >= NB) continue;
                const int bxi = blk % BX, byi = blk / BX;
                const int li = 2 * bxi + (c & 1), lj = 2 * byi + (c >> 1);
                const int gj = j0 + lj;
                if (li == 0 || li == RX - 1 || lj == 0 || lj == RY - 1 || gj < 0 || gj >= n) continue;
                vec2s<T> dl, dc, dr, cl, cr, ul, uc, ur;
                if (c == 0) tile_neighbours_own<T, BX, NB, 0>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else if (c == 1) tile_neighbours_own<T, BX, NB, 1>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else if (c == 2) tile_neighbours_own<T, BX, NB, 2>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else tile_neighbours_own<T, BX, NB, 3>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                T ax, ay;
                vec2s<T> out;
                if (walls && (gj < nb || gj >= n - nb)) {
                    const T *s = wall_stencil(ss, n, nb, gj);
                    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    out = gs_value<T>(s, bx[q][c], by[q][c], ax, ay);
                } else {
                    offdiag9_any<SP, T>(si, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    out = gs_value_any<SP, T>(si, bx[q][c], by[q][c], ax, ay);
                }
                zs[c * NB + blk] = out;
                zo[q][c] = out;
            }
            __syncthreads();
        }
    }

    // residuals of the tile and its first ring (z is valid two rings out)
    T rxs[FUSE ? BPT : 1][4], rys[FUSE ? BPT : 1][4];
    if (FUSE) {
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int blk = tid + q * NT;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                rxs[q][c] = 0;
                rys[q][c] = 0;
                if (blk >= NB) continue;
                const int bxi = blk % BX, byi = blk / BX;
                const int li = 2 * bxi + (c & 1), lj = 2 * byi + (c >> 1);
                const int gj = j0 + lj;
                if (li < H - 1 || li > H + TX || lj < H - 1 || lj > H + TY || gj < 0 || gj >= n) continue;
                vec2s<T> dl, dc, dr, cl, cr, ul, uc, ur;
                if (c == 0) tile_neighbours_own<T, BX, NB, 0>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else if (c == 1) tile_neighbours_own<T, BX, NB, 1>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else if (c == 2) tile_neighbours_own<T, BX, NB, 2>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                else tile_neighbours_own<T, BX, NB, 3>(zs, blk, zo[q], dl, dc, dr, cl, cr, ul, uc, ur);
                const vec2s<T> z0 = zo[q][c];
                T ax, ay;
                if (walls && (gj < nb || gj >= n - nb)) {
                    const T *s = wall_stencil(ss, n, nb, gj);
                    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    resid_value<T>(s, bx[q][c], by[q][c], ax, ay, z0, rxs[q][c], rys[q][c]);
                } else {
                    offdiag9_any<SP, T>(si, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    resid_value_any<SP, T>(si, bx[q][c], by[q][c], ax, ay, z0, rxs[q][c], rys[q][c]);
                }
            }
        }
    }

    // write back the interior
    for (int t = tid; t < TX * TY; t += NT) {
        const int li = H + t % TX, lj = H + t / TX;
        const int gj = j0 + lj, gi = i0 + li;
        if (gj >= n) continue;
        const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
        const vec2s<T> v = zs[((li & 1) + 2 * (lj & 1)) * NB + (lj >> 1) * BX + (li >> 1)];
        if (corr) {  // the mixed form's correction vel += e, in place of storing e
            corr[k] += (double)v.x;
            corr[k + 1] += (double)v.y;
        } else {
            zout[k] = v.x;
            zout[k + 1] = v.y;
        }
    }
    if (!FUSE) return;

    // the residual ring region, natural order, in the z storage
    constexpr int RW = TX + 2;
    static_assert((TX + 2) * (TY + 2) <= 4 * NB, "residual region fits the tile storage");
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int blk = tid + q * NT;
        if (blk >= NB) continue;
        const int bxi = blk % BX, byi = blk / BX;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int li = 2 * bxi + (c & 1), lj = 2 * byi + (c >> 1);
            if (li < H - 1 || li > H + TX || lj < H - 1 || lj > H + TY) continue;
            zs[(lj - (H - 1)) * RW + (li - (H - 1))] = vec2s<T>{rxs[q][c], rys[q][c]};
        }
    }
    __syncthreads();
    // coarse nodes of this tile: I in [fi0/2, fi0/2 + TX/2), J with its fine
    // centre row fy0(J) in [fj0, fj0 + TY) (for an even fine height the last
    // coarse row is centred on the last fine row)
    const int fi0 = tbx * TX, fj0 = tby * TY;
    for (int t = tid; t < (TX / 2) * (TY / 2 + 1); t += NT) {
        const int tI = t % (TX / 2), lr = t / (TX / 2);
        int J;
        if (lr < TY / 2) {
            J = fj0 / 2 + lr;
        } else {
            J = Cr.n - 1;
            if ((n & 1) || (J >= fj0 / 2 && J < fj0 / 2 + TY / 2)) continue;
        }
        if (J >= Cr.n) continue;
        const int fy0 = ((n & 1) == 0 && J == Cr.n - 1) ? n - 1 : 2 * J;
        if (fy0 < fj0 || fy0 >= fj0 + TY) continue;
        const int I = fi0 / 2 + tI;
        T bxs = 0, bys = 0;
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            const int fy = fy0 + b;
            const T wy = (T)pw_bounded(fy, n, Cr.n, J);
            const vec2s<T> *rp = zs + (fy - (fj0 - 1)) * RW + 2 * tI;  // fine columns 2I-1 .. 2I+1
            if (wy != (T)0) {
                bxs = fmav(wy, fmav((T)0.5, rp[0].x + rp[2].x, rp[1].x), bxs);
                bys = fmav(wy, fmav((T)0.5, rp[0].y + rp[2].y, rp[1].y), bys);
            }
        }
        const size_t k = 2 * ((size_t)I + (size_t)Cr.m * J);
        Cr.b[k] = bxs;
        Cr.b[k + 1] = bys;
        if (zero_z) {
            Cr.z[k] = 0;
            Cr.z[k + 1] = 0;
        }
    }
}

template <typename T, int TX, int TY, int S, int MODE, bool FUSE, bool SP>
__global__ void __launch_bounds__(tile_threads(TX, TY, tile_halo(S, FUSE)))
    smooth_tile_kernel(LevelArgs<T> L, T *__restrict__ zout, LevelArgs<T> Cz, LevelArgs<T> Cr, int zero_z,
                       double *corr) {
    smooth_tile_body<T, TX, TY, S, MODE, FUSE, SP>(L, zout, Cz, Cr, zero_z, corr, blockIdx.x, blockIdx.y);
}

// the same over the members of an ensemble (batch.h); members whose slot
// flag is clear return at once
template <typename T, int TX, int TY, int S, int MODE, bool FUSE, bool SP>
__global__ void __launch_bounds__(tile_threads(TX, TY, tile_halo(S, FUSE)))
    smooth_tile_batch_kernel(const TileEntry<T> *E, const int *map, int ne, const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const TileEntry<T> &x = E[e];
    if (St && !St[x.member].active) return;
    const LevelArgs<T> L = x.L, Cz = x.Cz, Cr = x.Cr;
    const int gx = x.gx;
    smooth_tile_body<T, TX, TY, S, MODE, FUSE, SP>(L, x.zout, Cz, Cr, x.zero_z, x.corr, l % gx, l / gx);
}

// restriction: coarse tile CX x CY of C from fine level F
template <typename T, int CX, int CY, bool SP>
__global__ void __launch_bounds__(256) restrict_tile_kernel(LevelArgs<T> F, LevelArgs<T> C, int zero_z) {
    // fine residual region: cols [2I0-1, 2I0+2CX-1], rows [2J0-2, 2J0+2CY]
    constexpr int RXr = 2 * CX + 1, RYr = 2 * CY + 3;
    // z region one node wider on every side
    constexpr int ZX = RXr + 2, ZY = RYr + 2, NT = 256;
    __shared__ vec2s<T> zs[ZY * ZX];
    __shared__ vec2s<T> rs[RYr * RXr];
    __shared__ T ss[2 * kMaxWall * 40];
    const int tid = threadIdx.x, m = F.m, n = F.n, nb = F.nb;
    const int I0 = blockIdx.x * CX, J0 = blockIdx.y * CY;
    const int zi0 = 2 * I0 - 2, zj0 = 2 * J0 - 3;
    T si[40];
#pragma unroll
    for (int k = 0; k < 40; k++) si[k] = F.Si[k];
    for (int t = tid; t < ZX * ZY; t += NT) {
        const int lj = t / ZX, li = t - lj * ZX;
        const int gj = zj0 + lj;
        int gi = zi0 + li;
        gi = ((gi % m) + m) % m;
        vec2s<T> v;
        v.x = 0;
        v.y = 0;
        if (gj >= 0 && gj < n) {
            const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
            v.x = F.z[k];
            v.y = F.z[k + 1];
        }
        zs[t] = v;
    }
    if (zj0 + 1 < nb || zj0 + 1 + RYr > n - nb) load_wall_stencils(ss, F, tid, NT);
    __syncthreads();
    for (int t = tid; t < RXr * RYr; t += NT) {
        const int lj = t / RXr, li = t - lj * RXr;
        const int gj = zj0 + 1 + lj;
        vec2s<T> r;
        r.x = 0;
        r.y = 0;
        if (gj >= 0 && gj < n) {
            int gi = zi0 + 1 + li;
            gi = ((gi % m) + m) % m;
            const vec2s<T> *zd = zs + lj * ZX + li + 1, *zc = zd + ZX, *zu = zc + ZX;
            T ax, ay;
            const vec2s<T> z0 = zc[0];
            const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
            if (gj < nb || gj >= n - nb) {
                const T *s = wall_stencil(ss, n, nb, gj);
                offdiag9<T>(s, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
                resid_value<T>(s, F.b[k], F.b[k + 1], ax, ay, z0, r.x, r.y);
            } else {
                offdiag9_any<SP, T>(si, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
                resid_value_any<SP, T>(si, F.b[k], F.b[k + 1], ax, ay, z0, r.x, r.y);
            }
        }
        rs[t] = r;
    }
    __syncthreads();
    for (int t = tid; t < CX * CY; t += NT) {
        const int I = I0 + t % CX, J = J0 + t / CX;
        if (I >= C.m || J >= C.n) continue;
        const int fx0 = 2 * I;
        const int fy0 = ((n & 1) == 0 && J == C.n - 1) ? n - 1 : 2 * J;
        T bxs = 0, bys = 0;
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            const int fy = fy0 + b;
            const float wy = pw_bounded(fy, n, C.n, J);
            if (wy == 0.f) continue;
#pragma unroll
            for (int a = -1; a <= 1; a++) {
                const float wx = pw_periodic(fx0 + a, m, C.m, I);
                if (wx == 0.f) continue;
                const vec2s<T> r = rs[(fy - (zj0 + 1)) * RXr + (fx0 + a - (zi0 + 1))];
                const T w = (T)(wx * wy);
                bxs += w * r.x;
                bys += w * r.y;
            }
        }
        const size_t k = 2 * ((size_t)I + (size_t)C.m * J);
        C.b[k] = bxs;
        C.b[k + 1] = bys;
        if (zero_z) {
            C.z[k] = 0;
            C.z[k + 1] = 0;
        }
    }
}

// FP64 residual of the top level (64x16 tiles) -> FP32 source, |r|^2 partials
template <bool SP>
__device__ __forceinline__ static void refine_body(const LevelArgs<double> &L, float *b32, double *partial,
                                                   int bx, int by, int gx) {
    constexpr int TX = 64, TY = 16, ZX = TX + 2, ZY = TY + 2, NT = 256;
    __shared__ vec2s<double> zs[ZY * ZX];
    __shared__ double ss[2 * kMaxWall * 40];
    __shared__ double ws[4];
    const int tid = threadIdx.x, m = L.m, n = L.n, nb = L.nb;
    const int i0 = bx * TX, j0 = by * TY;
    double si[40];
#pragma unroll
    for (int k = 0; k < 40; k++) si[k] = L.Si[k];
    for (int t = tid; t < ZX * ZY; t += NT) {
        const int lj = t / ZX, li = t - lj * ZX;
        const int gj = j0 - 1 + lj;
        int gi = i0 - 1 + li;
        gi = gi < 0 ? gi + m : (gi >= m ? gi - m : gi);
        vec2s<double> v;
        v.x = 0;
        v.y = 0;
        if (gj >= 0 && gj < n) {
            const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
            v.x = L.z[k];
            v.y = L.z[k + 1];
        }
        zs[t] = v;
    }
    if (j0 < nb || j0 + TY > n - nb) load_wall_stencils(ss, L, tid, NT);
    __syncthreads();
    double acc = 0;
    for (int t = tid; t < TX * TY; t += NT) {
        const int li = t % TX, lj = t / TX;
        const int gi = i0 + li, gj = j0 + lj;
        if (gi >= m || gj >= n) continue;
        const vec2s<double> *zd = zs + lj * ZX + li + 1, *zc = zd + ZX, *zu = zc + ZX;
        const vec2s<double> z0 = zc[0];
        const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
        double ax, ay, rx, ry;
        if (gj < nb || gj >= n - nb) {
            const double *s = wall_stencil(ss, n, nb, gj);
            offdiag9<double>(s, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
            resid_value<double>(s, L.b[k], L.b[k + 1], ax, ay, z0, rx, ry);
        } else {
            offdiag9_any<SP, double>(si, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
            resid_value_any<SP, double>(si, L.b[k], L.b[k + 1], ax, ay, z0, rx, ry);
        }
        b32[k] = (float)rx;
        b32[k + 1] = (float)ry;
        acc += rx * rx + ry * ry;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((tid & 63) == 0) ws[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) partial[bx + gx * by] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

template <bool SP>
__global__ void __launch_bounds__(256) refine_kernel(LevelArgs<double> L, float *b32, double *partial) {
    refine_body<SP>(L, b32, partial, blockIdx.x, blockIdx.y, gridDim.x);
}

template <bool SP>
__global__ void __launch_bounds__(256) refine_batch_kernel(const RefineEntry *E, const int *map, int ne,
                                                           const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const int e = el.x, l = el.y;
    const RefineEntry &x = E[e];
    if (St && !St[x.member].resid) return;
    const LevelArgs<double> L = x.L;
    const int gx = x.gx;
    refine_body<SP>(L, x.b32, x.partial, l % gx, l / gx, gx);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T> bool wall_ok(const LevelArgs<T> &L) {
    return L.cs == 0 && L.nb <= kMaxWall && L.n >= 2 * L.nb;
}
template <typename T> bool tiled_ok(const LevelArgs<T> &L) { return wall_ok(L) && L.m % 32 == 0 && L.m >= 32; }

// large tiles when they give at least one workgroup per CU, else 32x16
// tiles: on the small levels the kernel is latency-bound and more, lighter
// workgroups finish sooner.  The large tiles are 64x34: the 1025 node rows
// of a 1024^2 level make 31 rows of them (496 tiles, at most two per CU),
// where 64x32 tiles leave a 33rd row holding one node row and put a third
// tile on some CUs (528 tiles), which set the launch's time: 15.8 -> 13.1 us
// for the FP32 pre-smoothing launch, 35.2 -> 30.2 us in FP64
template <typename T> static bool big_tiles(const LevelArgs<T> &L) {
    return L.m % 64 == 0 && (L.m / 64) * ((L.n + 31) / 32) >= 256;
}
// 16x16 tiles on the smallest tiled levels (fewer than 100 tiles of 32x16):
// twice the workgroups of lighter work, 0.6 % (mixed) and 1.7 % (FP64) off
// the 1024^2 step
template <typename T> static bool small_tiles(const LevelArgs<T> &L) { return (L.m / 32) * ((L.n + 15) / 16) < 100; }


template <typename T, int S, int MODE, bool FUSE>
static void launch_tile(const LevelArgs<T> &L, T *zout, const LevelArgs<T> &Cz, const LevelArgs<T> &Cr, int zero_z,
                        double *corr, hipStream_t st) {
    if (big_tiles(L)) {
        const dim3 g(L.m / 64, (L.n + 33) / 34), b(tile_threads(64, 34, tile_halo(S, FUSE)));
        if (L.sparse)
            hipLaunchKernelGGL((smooth_tile_kernel<T, 64, 34, S, MODE, FUSE, true>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
        else
            hipLaunchKernelGGL((smooth_tile_kernel<T, 64, 34, S, MODE, FUSE, false>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
    } else if (small_tiles(L)) {
        const dim3 g(L.m / 16, (L.n + 15) / 16), b(tile_threads(16, 16, tile_halo(S, FUSE)));
        if (L.sparse)
            hipLaunchKernelGGL((smooth_tile_kernel<T, 16, 16, S, MODE, FUSE, true>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
        else
            hipLaunchKernelGGL((smooth_tile_kernel<T, 16, 16, S, MODE, FUSE, false>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
    } else {
        const dim3 g(L.m / 32, (L.n + 15) / 16), b(tile_threads(32, 16, tile_halo(S, FUSE)));
        if (L.sparse)
            hipLaunchKernelGGL((smooth_tile_kernel<T, 32, 16, S, MODE, FUSE, true>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
        else
            hipLaunchKernelGGL((smooth_tile_kernel<T, 32, 16, S, MODE, FUSE, false>), g, b, 0, st, L, zout, Cz, Cr,
                               zero_z, corr);
    }
}

template <typename T, int S, bool FUSE>
static void launch_tile_mode(int mode, const LevelArgs<T> &L, T *zout, const LevelArgs<T> &Cz, const LevelArgs<T> &Cr,
                             int zero_z, double *corr, hipStream_t st) {
    if (mode == TILE_ZERO) launch_tile<T, S, TILE_ZERO, FUSE>(L, zout, Cz, Cr, zero_z, corr, st);
    else if (mode == TILE_PROLONG) launch_tile<T, S, TILE_PROLONG, FUSE>(L, zout, Cz, Cr, zero_z, corr, st);
    else launch_tile<T, S, TILE_LOAD, FUSE>(L, zout, Cz, Cr, zero_z, corr, st);
}

// the coarse z must start at zero unless its first smoothing is tiled
// (which starts from zero by itself)
template <typename T> static int coarse_needs_zero(const LevelArgs<T> &C) { return tiled_ok(C) ? 0 : 1; }

// `sweeps` sweeps on level L; the first kernel starts from zero (TILE_ZERO),
// loads z (TILE_LOAD) or adds the prolongation of C (TILE_PROLONG).  Tiled
// kernels read L.z and write L.z2, then the two are swapped.  With R, the
// last tiled launch also restricts the residual into R (returns 1 in
// *restricted).  Levels the tiled kernel does not cover use the per-colour
// kernels in place.
template <typename T>
static int smooth_level(LevelArgs<T> &L, const LevelArgs<T> *C, int sweeps, int mode, hipStream_t st,
                        const LevelArgs<T> *R = nullptr, int *restricted = nullptr, double *corr = nullptr,
                        int *corrected = nullptr) {
    int e;
    if (restricted) *restricted = 0;
    if (corrected) *corrected = 0;
    if (!tiled_ok(L)) {
        if (mode == TILE_PROLONG && (e = mg_prolong(*C, L, st))) return e;
        if (mode == TILE_ZERO && (e = mg_fill_zero(L.z, 2 * (size_t)L.m * L.n, st))) return e;
        return mg_smooth(L, sweeps, st);
    }
    const LevelArgs<T> Cv = C ? *C : L;
    const LevelArgs<T> Rv = R ? *R : L;
    const int zz = R ? coarse_needs_zero(*R) : 0;
    while (sweeps > 0) {
        const int s = sweeps >= 2 ? 2 : 1;
        const bool fuse = R && s == sweeps;
        double *cr = s == sweeps ? corr : nullptr;
        if (s == 2) {
            if (fuse) launch_tile_mode<T, 2, true>(mode, L, L.z2, Cv, Rv, zz, cr, st);
            else launch_tile_mode<T, 2, false>(mode, L, L.z2, Cv, Rv, zz, cr, st);
        } else {
            if (fuse) launch_tile_mode<T, 1, true>(mode, L, L.z2, Cv, Rv, zz, cr, st);
            else launch_tile_mode<T, 1, false>(mode, L, L.z2, Cv, Rv, zz, cr, st);
        }
        if (fuse && restricted) *restricted = 1;
        if (cr && corrected) *corrected = 1;
        std::swap(L.z, L.z2);
        sweeps -= s;
        mode = TILE_LOAD;
    }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> static int restrict_level(const LevelArgs<T> &F, const LevelArgs<T> &C, hipStream_t st) {
    if (!wall_ok(F)) return mg_restrict(F, C, st);  // per-node stencils: simple kernel (zeroes C.z)
    const int zero_z = coarse_needs_zero(C);
    constexpr int CY = sizeof(T) == 8 ? 8 : 16;
    if (((C.m + 31) / 32) * ((C.n + CY - 1) / CY) >= 256) {
        const dim3 g((C.m + 31) / 32, (C.n + CY - 1) / CY);
        if (F.sparse) hipLaunchKernelGGL((restrict_tile_kernel<T, 32, CY, true>), g, dim3(256), 0, st, F, C, zero_z);
        else hipLaunchKernelGGL((restrict_tile_kernel<T, 32, CY, false>), g, dim3(256), 0, st, F, C, zero_z);
    } else {
        const dim3 g((C.m + 15) / 16, (C.n + 7) / 8);
        if (F.sparse) hipLaunchKernelGGL((restrict_tile_kernel<T, 16, 8, true>), g, dim3(256), 0, st, F, C, zero_z);
        else hipLaunchKernelGGL((restrict_tile_kernel<T, 16, 8, false>), g, dim3(256), 0, st, F, C, zero_z);
    }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_vcycle_tiled(CoarseArgs<T> &A, bool top_zero, hipStream_t st, double *corr,
                                          int *corrected) {
    int e;
    if (corrected) *corrected = 0;
    const int lb = bottom_level(A);
    for (int l = 0; l < lb; l++) {
        // pre-smoothing, fused with the restriction into level l+1 when tiled
        int done = 0;
        if ((e = smooth_level(A.lv[l], (const LevelArgs<T> *)nullptr, A.nu1,
                              l == 0 ? (top_zero ? TILE_ZERO : TILE_LOAD) : TILE_ZERO, st, &A.lv[l + 1], &done)))
            return e;
        if (!done && (e = restrict_level(A.lv[l], A.lv[l + 1], st))) return e;
    }
    if (lb == 0 && (e = smooth_level(A.lv[0], (const LevelArgs<T> *)nullptr, A.nu1,
                                     top_zero ? TILE_ZERO : TILE_LOAD, st)))
        return e;
    if ((e = mg_bottom(A, st))) return e;
    for (int l = lb; l >= 1; l--)
        if ((e = smooth_level(A.lv[l - 1], &A.lv[l], A.nu2, TILE_PROLONG, st, (const LevelArgs<T> *)nullptr,
                              nullptr, l == 1 ? corr : nullptr, l == 1 ? corrected : nullptr)))
            return e;
    return STZ_OK;
}

// One of the smoothing launches a tiled V-cycle issues on `level`, alone:
// mode 1 the pre-smoothing (nu1 sweeps from zero on the correction equation
// or a coarse level, else from z; the residual restricted into level+1),
// mode 2 the post-smoothing (nu2 sweeps after adding the prolongation of
// level+1; on the mixed form's top level also vel += e into corr).  For
// timing the kernels that mg_vcycle_tiled runs; the level's z/z2 are left
// swapped as the cycle leaves them.
template <typename T>
int mg_vcycle_launch(CoarseArgs<T> &A, int level, int mode, bool top_zero, double *corr, hipStream_t st) {
    if (level < 0 || level >= bottom_level(A) || !tiled_ok(A.lv[level])) {
        set_error("level %d is not a tiled V-cycle level", level);
        return STZ_ERR_ARG;
    }
    int done = 0;
    if (mode == 1) {
        const int m0 = level == 0 ? (top_zero ? TILE_ZERO : TILE_LOAD) : TILE_ZERO;
        int e = smooth_level(A.lv[level], (const LevelArgs<T> *)nullptr, A.nu1, m0, st, &A.lv[level + 1], &done);
        if (e) return e;
        return done ? STZ_OK : restrict_level(A.lv[level], A.lv[level + 1], st);
    }
    if (mode == 2)
        return smooth_level(A.lv[level], &A.lv[level + 1], A.nu2, TILE_PROLONG, st, (const LevelArgs<T> *)nullptr,
                            nullptr, level == 0 ? corr : nullptr, &done);
    set_error("unknown launch mode %d", mode);
    return STZ_ERR_ARG;
}
template int mg_vcycle_launch<double>(CoarseArgs<double> &, int, int, bool, double *, hipStream_t);
template int mg_vcycle_launch<float>(CoarseArgs<float> &, int, int, bool, double *, hipStream_t);

template <typename T> int mg_smooth_level(LevelArgs<T> &L, int sweeps, hipStream_t st) {
    return smooth_level(L, (const LevelArgs<T> *)nullptr, sweeps, TILE_LOAD, st);
}

int mg_refine_blocks(const LevelArgs<double> &L) { return ((L.m + 63) / 64) * ((L.n + 15) / 16); }

int mg_refine(const LevelArgs<double> &L, float *b32, double *partial, double *out, hipStream_t st) {
    const dim3 g((L.m + 63) / 64, (L.n + 15) / 16);
    if (L.sparse) hipLaunchKernelGGL(refine_kernel<true>, g, dim3(256), 0, st, L, b32, partial);
    else hipLaunchKernelGGL(refine_kernel<false>, g, dim3(256), 0, st, L, b32, partial);
    STZ_HIP(hipGetLastError());
    return out ? mg_sum_partials(partial, (int)(g.x * g.y), out, st) : STZ_OK;
}

template bool wall_ok<double>(const LevelArgs<double> &);
template bool wall_ok<float>(const LevelArgs<float> &);
template bool tiled_ok<double>(const LevelArgs<double> &);
template bool tiled_ok<float>(const LevelArgs<float> &);
template int mg_vcycle_tiled<double>(CoarseArgs<double> &, bool, hipStream_t, double *, int *);
template int mg_vcycle_tiled<float>(CoarseArgs<float> &, bool, hipStream_t, double *, int *);
template int mg_smooth_level<double>(LevelArgs<double> &, int, hipStream_t);
template int mg_smooth_level<float>(LevelArgs<float> &, int, hipStream_t);

// ---------------------------------------------------------------------------
// batched launchers (batch.h).  Tile shape does not change what a tiled
// launch computes (every node sees the same colour order and operands), so
// batches use 32x16 tiles, or the large 64x34 ones when key.big
// ---------------------------------------------------------------------------
template <typename T, int S, int MODE, bool FUSE, bool SP>
static void launch_tile_batch(bool big, const TileEntry<T> *E, const int *map, int ne, int total,
                              const EnsState *St, hipStream_t st) {
    if (big)
        hipLaunchKernelGGL((smooth_tile_batch_kernel<T, 64, 34, S, MODE, FUSE, SP>), dim3(total),
                           dim3(tile_threads(64, 34, tile_halo(S, FUSE))), 0, st, E, map, ne, St);
    else
        hipLaunchKernelGGL((smooth_tile_batch_kernel<T, 32, 16, S, MODE, FUSE, SP>), dim3(total),
                           dim3(tile_threads(32, 16, tile_halo(S, FUSE))), 0, st, E, map, ne, St);
}

template <typename T, int S, int MODE>
static void launch_tile_batch_fs(const TileKey &k, const TileEntry<T> *E, const int *map, int ne, int total,
                                 const EnsState *St, hipStream_t st) {
    if (k.fuse) {
        if (k.sparse) launch_tile_batch<T, S, MODE, true, true>(k.big, E, map, ne, total, St, st);
        else launch_tile_batch<T, S, MODE, true, false>(k.big, E, map, ne, total, St, st);
    } else {
        if (k.sparse) launch_tile_batch<T, S, MODE, false, true>(k.big, E, map, ne, total, St, st);
        else launch_tile_batch<T, S, MODE, false, false>(k.big, E, map, ne, total, St, st);
    }
}

template <typename T, int S>
static void launch_tile_batch_m(const TileKey &k, const TileEntry<T> *E, const int *map, int ne, int total,
                                const EnsState *St, hipStream_t st) {
    if (k.mode == TILE_ZERO) launch_tile_batch_fs<T, S, TILE_ZERO>(k, E, map, ne, total, St, st);
    else if (k.mode == TILE_PROLONG) launch_tile_batch_fs<T, S, TILE_PROLONG>(k, E, map, ne, total, St, st);
    else launch_tile_batch_fs<T, S, TILE_LOAD>(k, E, map, ne, total, St, st);
}

template <typename T>
int batch_tile(const TileKey &key, const TileEntry<T> *E, const int *map, int ne, int total, const EnsState *St,
               hipStream_t st) {
    if (ne == 0) return STZ_OK;
    if (key.sweeps == 2) launch_tile_batch_m<T, 2>(key, E, map, ne, total, St, st);
    else if (key.sweeps == 1) launch_tile_batch_m<T, 1>(key, E, map, ne, total, St, st);
    else {
        set_error("batched smoother launch with %d sweeps", key.sweeps);
        return STZ_ERR_ARG;
    }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
template int batch_tile<double>(const TileKey &, const TileEntry<double> *, const int *, int, int, const EnsState *,
                                hipStream_t);
template int batch_tile<float>(const TileKey &, const TileEntry<float> *, const int *, int, int, const EnsState *,
                               hipStream_t);

int batch_refine(const RefineEntry *E, const int *map, int ne, int total, int sparse, const EnsState *St,
                 hipStream_t st) {
    if (ne == 0) return STZ_OK;
    if (sparse) hipLaunchKernelGGL(refine_batch_kernel<true>, dim3(total), dim3(256), 0, st, E, map, ne, St);
    else hipLaunchKernelGGL(refine_batch_kernel<false>, dim3(total), dim3(256), 0, st, E, map, ne, St);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// the tiled V-cycle of one member as a list of launches, for the ensemble:
// the same sequence mg_vcycle_tiled issues (and the same z/z2 exchanges),
// recorded instead of launched.  Fails for a level the batched path does
// not cover (per-colour kernels, separate restriction).
template <typename T>
static int plan_level(LevelArgs<T> &L, const LevelArgs<T> *C, int sweeps, int mode, const LevelArgs<T> *R,
                      int level, int phase, std::vector<TileRec<T>> &out, double *corr = nullptr) {
    if (!tiled_ok(L)) {
        set_error("level %d (%d x %d) is not covered by the tiled kernels", level, L.m, L.n);
        return STZ_ERR_ARG;
    }
    const LevelArgs<T> Cv = C ? *C : L;
    const LevelArgs<T> Rv = R ? *R : L;
    const int zz = R ? coarse_needs_zero(*R) : 0;
    int idx = 0;
    while (sweeps > 0) {
        const int s = sweeps >= 2 ? 2 : 1;
        const bool fuse = R && s == sweeps;
        TileRec<T> r;
        r.phase = phase;
        r.level = level;
        r.index = idx++;
        r.key = TileKey{s, mode, fuse ? 1 : 0, L.sparse ? 1 : 0, 0};
        r.L = L;
        r.zout = L.z2;
        r.Cz = Cv;
        r.Cr = Rv;
        r.zero_z = zz;
        r.corr = s == sweeps ? corr : nullptr;
        out.push_back(r);
        std::swap(L.z, L.z2);
        sweeps -= s;
        mode = TILE_LOAD;
    }
    return STZ_OK;
}

template <typename T>
int mg_vcycle_plan(CoarseArgs<T> A, bool top_zero, std::vector<TileRec<T>> &out, double *corr) {
    int e;
    out.clear();
    // the z/z2 exchanges of one cycle must cancel on every level, so that
    // the recorded pointers hold for every cycle of every member; the
    // restriction rides on the last pre-smoothing launch
    if ((((A.nu1 + 1) / 2 + (A.nu2 + 1) / 2) & 1) || A.nu1 < 1) {
        set_error("batched V-cycle needs nu1 >= 1 and an even number of smoother launches per level (nu1 %d, "
                  "nu2 %d)", A.nu1, A.nu2);
        return STZ_ERR_ARG;
    }
    if (A.lc < 1) {
        set_error("batched V-cycle needs at least one grid level above the workgroup solver");
        return STZ_ERR_ARG;
    }
    const int lb = bottom_level(A);
    for (int l = 0; l < lb; l++) {
        if ((e = plan_level(A.lv[l], (const LevelArgs<T> *)nullptr, A.nu1,
                            l == 0 ? (top_zero ? TILE_ZERO : TILE_LOAD) : TILE_ZERO, &A.lv[l + 1], l, 0, out)))
            return e;
    }
    for (int l = lb; l >= 1; l--)
        if ((e = plan_level(A.lv[l - 1], &A.lv[l], A.nu2, TILE_PROLONG, (const LevelArgs<T> *)nullptr, l - 1, 2,
                            out, l == 1 ? corr : nullptr)))
            return e;
    return STZ_OK;
}
template int mg_vcycle_plan<double>(CoarseArgs<double>, bool, std::vector<TileRec<double>> &, double *);
template int mg_vcycle_plan<float>(CoarseArgs<float>, bool, std::vector<TileRec<float>> &, double *);

}  // namespace stz
