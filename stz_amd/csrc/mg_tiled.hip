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

#include "mg_node.h"

namespace stz {

enum { TILE_LOAD = 0, TILE_ZERO = 1, TILE_PROLONG = 2 };

// one thread per 2x2 block of the tile-plus-halo region
__host__ __device__ constexpr int tile_threads(int TX, int TY, int S) {
    return (((TX + 8 * S) / 2) * ((TY + 8 * S) / 2) + 63) / 64 * 64;
}

template <typename T, int TX, int TY, int S, int MODE>
__global__ void __launch_bounds__(tile_threads(TX, TY, S))
    smooth_tile_kernel(LevelArgs<T> L, T *__restrict__ zout, LevelArgs<T> Cz) {
    constexpr int H = 4 * S, RX = TX + 2 * H, RY = TY + 2 * H;
    constexpr int BX = RX / 2, BY = RY / 2, NB = BX * BY, NT = tile_threads(TX, TY, S);
    constexpr int BPT = (NB + NT - 1) / NT;
    static_assert(RX % 2 == 0 && RY % 2 == 0, "even region");
    // colour-split LDS: node (li,lj) lives in array c = (li&1) + 2(lj&1) at
    // block index (lj>>1)*BX + (li>>1); a colour pass then reads each
    // neighbour from another colour's array at consecutive addresses across
    // the lanes (no bank conflicts), with compile-time offsets per colour
    __shared__ vec2s<T> zs[4 * NB];
    __shared__ T ss[RY * 40];
    const int tid = threadIdx.x;
    const int m = L.m, n = L.n, nb = L.nb;
    const int i0 = blockIdx.x * TX - H, j0 = blockIdx.y * TY - H;

    // interior stencil: wave-uniform, held in registers
    T si[40];
#pragma unroll
    for (int k = 0; k < 40; k++) si[k] = L.Si[k];
    // wall-adjacent rows' stencils, only if this tile reaches them
    const bool edge_tile = j0 < nb || j0 + RY > n - nb;
    if (edge_tile)
        for (int t = tid; t < RY * 40; t += NT) {
            const int r = t / 40, k = t - r * 40, j = j0 + r;
            ss[t] = (j >= 0 && j < n) ? L.S[(size_t)j * 40 + k] : (T)0;
        }
    // z into LDS (zero outside the level), b into registers
    T bx[BPT][4], by[BPT][4];
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
                // neighbours: x-neighbours have colour cX, y-neighbours cY,
                // diagonal ones cD; block offsets depend only on (cx, cy)
                const int cx = c & 1, cy = c >> 1;
                const int cX = (1 - cx) + 2 * cy, cY = cx + 2 * (1 - cy), cD = (1 - cx) + 2 * (1 - cy);
                const int xl = cx ? 0 : -1, xr = cx ? 1 : 0, yd = cy ? 0 : -BX, yu = cy ? BX : 0;
                const vec2s<T> *zX = zs + cX * NB + blk, *zY = zs + cY * NB + blk, *zD = zs + cD * NB + blk;
                const vec2s<T> dl = zD[yd + xl], dc = zY[yd], dr = zD[yd + xr];
                const vec2s<T> cl = zX[xl], cr = zX[xr];
                const vec2s<T> ul = zD[yu + xl], uc = zY[yu], ur = zD[yu + xr];
                T ax, ay;
                vec2s<T> out;
                if (gj < nb || gj >= n - nb) {
                    const T *s = ss + lj * 40;
                    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    out = gs_value<T>(s, bx[q][c], by[q][c], ax, ay);
                } else {
                    offdiag9<T>(si, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
                    out = gs_value<T>(si, bx[q][c], by[q][c], ax, ay);
                }
                zs[c * NB + blk] = out;
            }
            __syncthreads();
        }
    }

    // write back the interior
    for (int t = tid; t < TX * TY; t += NT) {
        const int li = H + t % TX, lj = H + t / TX;
        const int gj = j0 + lj, gi = i0 + li;
        if (gj >= n) continue;
        const size_t k = 2 * ((size_t)gi + (size_t)m * gj);
        const vec2s<T> v = zs[((li & 1) + 2 * (lj & 1)) * NB + (lj >> 1) * BX + (li >> 1)];
        zout[k] = v.x;
        zout[k + 1] = v.y;
    }
}

// restriction: coarse tile CX x CY of C from fine level F
template <typename T, int CX, int CY>
__global__ void __launch_bounds__(256) restrict_tile_kernel(LevelArgs<T> F, LevelArgs<T> C, int zero_z) {
    // fine residual region: cols [2I0-1, 2I0+2CX-1], rows [2J0-2, 2J0+2CY]
    constexpr int RXr = 2 * CX + 1, RYr = 2 * CY + 3;
    // z region one node wider on every side
    constexpr int ZX = RXr + 2, ZY = RYr + 2, NT = 256;
    __shared__ vec2s<T> zs[ZY * ZX];
    __shared__ vec2s<T> rs[RYr * RXr];
    __shared__ T ss[RYr * 40];
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
    if (zj0 + 1 < nb || zj0 + 1 + RYr > n - nb)
        for (int t = tid; t < RYr * 40; t += NT) {
            const int r = t / 40, k = t - r * 40, j = zj0 + 1 + r;
            ss[t] = (j >= 0 && j < n) ? F.S[(size_t)j * 40 + k] : (T)0;
        }
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
                const T *s = ss + lj * 40;
                offdiag9<T>(s, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
                resid_value<T>(s, F.b[k], F.b[k + 1], ax, ay, z0, r.x, r.y);
            } else {
                offdiag9<T>(si, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
                resid_value<T>(si, F.b[k], F.b[k + 1], ax, ay, z0, r.x, r.y);
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
__global__ void __launch_bounds__(256) refine_kernel(LevelArgs<double> L, float *b32, double *partial) {
    constexpr int TX = 64, TY = 16, ZX = TX + 2, ZY = TY + 2, NT = 256;
    __shared__ vec2s<double> zs[ZY * ZX];
    __shared__ double ss[TY * 40];
    __shared__ double ws[4];
    const int tid = threadIdx.x, m = L.m, n = L.n, nb = L.nb;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
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
    if (j0 < nb || j0 + TY > n - nb)
        for (int t = tid; t < TY * 40; t += NT) {
            const int r = t / 40, k = t - r * 40, j = j0 + r;
            ss[t] = j < n ? L.S[(size_t)j * 40 + k] : 0.0;
        }
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
            const double *s = ss + lj * 40;
            offdiag9<double>(s, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
            resid_value<double>(s, L.b[k], L.b[k + 1], ax, ay, z0, rx, ry);
        } else {
            offdiag9<double>(si, zd[-1], zd[0], zd[1], zc[-1], zc[1], zu[-1], zu[0], zu[1], ax, ay);
            resid_value<double>(si, L.b[k], L.b[k + 1], ax, ay, z0, rx, ry);
        }
        b32[k] = (float)rx;
        b32[k + 1] = (float)ry;
        acc += rx * rx + ry * ry;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((tid & 63) == 0) ws[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x + gridDim.x * blockIdx.y] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T> bool tiled_ok(const LevelArgs<T> &L) { return L.cs == 0 && L.m % 32 == 0 && L.m >= 32; }

// large tiles (64x32) when they give at least one workgroup per CU, else
// 32x16 tiles: on the small levels the kernel is latency-bound and more,
// lighter workgroups finish sooner
template <typename T, int S, int MODE>
static void launch_tile(const LevelArgs<T> &L, T *zout, const LevelArgs<T> &C, hipStream_t st) {
    if (L.m % 64 == 0 && (L.m / 64) * ((L.n + 31) / 32) >= 256) {
        const dim3 g(L.m / 64, (L.n + 31) / 32);
        hipLaunchKernelGGL((smooth_tile_kernel<T, 64, 32, S, MODE>), g, dim3(tile_threads(64, 32, S)), 0, st, L, zout,
                           C);
    } else {
        const dim3 g(L.m / 32, (L.n + 15) / 16);
        hipLaunchKernelGGL((smooth_tile_kernel<T, 32, 16, S, MODE>), g, dim3(tile_threads(32, 16, S)), 0, st, L, zout,
                           C);
    }
}

// `sweeps` sweeps on level L; the first kernel starts from zero (TILE_ZERO),
// loads z (TILE_LOAD) or adds the prolongation of C (TILE_PROLONG).  Tiled
// kernels read L.z and write L.z2, then the two are swapped.  Levels the
// tiled kernel does not cover use the per-colour kernels in place.
template <typename T> static int smooth_level(LevelArgs<T> &L, const LevelArgs<T> *C, int sweeps, int mode,
                                              hipStream_t st) {
    int e;
    if (!tiled_ok(L)) {
        if (mode == TILE_PROLONG && (e = mg_prolong(*C, L, st))) return e;
        if (mode == TILE_ZERO && (e = mg_fill_zero(L.z, 2 * (size_t)L.m * L.n, st))) return e;
        return mg_smooth(L, sweeps, st);
    }
    const LevelArgs<T> Cv = C ? *C : L;
    while (sweeps > 0) {
        const int s = sweeps >= 2 ? 2 : 1;
        if (s == 2) {
            if (mode == TILE_ZERO) launch_tile<T, 2, TILE_ZERO>(L, L.z2, Cv, st);
            else if (mode == TILE_PROLONG) launch_tile<T, 2, TILE_PROLONG>(L, L.z2, Cv, st);
            else launch_tile<T, 2, TILE_LOAD>(L, L.z2, Cv, st);
        } else {
            if (mode == TILE_ZERO) launch_tile<T, 1, TILE_ZERO>(L, L.z2, Cv, st);
            else if (mode == TILE_PROLONG) launch_tile<T, 1, TILE_PROLONG>(L, L.z2, Cv, st);
            else launch_tile<T, 1, TILE_LOAD>(L, L.z2, Cv, st);
        }
        std::swap(L.z, L.z2);
        sweeps -= s;
        mode = TILE_LOAD;
    }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> static int restrict_level(const LevelArgs<T> &F, const LevelArgs<T> &C, hipStream_t st) {
    if (F.cs != 0) return mg_restrict(F, C, st);  // per-node stencils: simple kernel (zeroes C.z)
    // the coarse z must start at zero unless its first smoothing is tiled
    const int zero_z = tiled_ok(C) ? 0 : 1;
    constexpr int CY = sizeof(T) == 8 ? 8 : 16;
    if (((C.m + 31) / 32) * ((C.n + CY - 1) / CY) >= 256) {
        const dim3 g((C.m + 31) / 32, (C.n + CY - 1) / CY);
        hipLaunchKernelGGL((restrict_tile_kernel<T, 32, CY>), g, dim3(256), 0, st, F, C, zero_z);
    } else {
        const dim3 g((C.m + 15) / 16, (C.n + 7) / 8);
        hipLaunchKernelGGL((restrict_tile_kernel<T, 16, 8>), g, dim3(256), 0, st, F, C, zero_z);
    }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_vcycle_tiled(CoarseArgs<T> &A, bool top_zero, hipStream_t st) {
    int e;
    if ((e = smooth_level(A.lv[0], (const LevelArgs<T> *)nullptr, A.nu1, top_zero ? TILE_ZERO : TILE_LOAD, st)))
        return e;
    for (int l = 0; l < A.lc; l++) {
        if ((e = restrict_level(A.lv[l], A.lv[l + 1], st))) return e;
        if (l + 1 < A.lc && (e = smooth_level(A.lv[l + 1], (const LevelArgs<T> *)nullptr, A.nu1, TILE_ZERO, st)))
            return e;
    }
    if ((e = mg_coarse(A, st))) return e;
    for (int l = A.lc; l >= 1; l--)
        if ((e = smooth_level(A.lv[l - 1], &A.lv[l], A.nu2, TILE_PROLONG, st))) return e;
    return STZ_OK;
}

template <typename T> int mg_smooth_level(LevelArgs<T> &L, int sweeps, hipStream_t st) {
    return smooth_level(L, (const LevelArgs<T> *)nullptr, sweeps, TILE_LOAD, st);
}

int mg_refine_blocks(const LevelArgs<double> &L) { return ((L.m + 63) / 64) * ((L.n + 15) / 16); }

int mg_refine(const LevelArgs<double> &L, float *b32, double *partial, double *out, hipStream_t st) {
    const dim3 g((L.m + 63) / 64, (L.n + 15) / 16);
    hipLaunchKernelGGL(refine_kernel, g, dim3(256), 0, st, L, b32, partial);
    STZ_HIP(hipGetLastError());
    return mg_sum_partials(partial, (int)(g.x * g.y), out, st);
}

template bool tiled_ok<double>(const LevelArgs<double> &);
template bool tiled_ok<float>(const LevelArgs<float> &);
template int mg_vcycle_tiled<double>(CoarseArgs<double> &, bool, hipStream_t);
template int mg_vcycle_tiled<float>(CoarseArgs<float> &, bool, hipStream_t);
template int mg_smooth_level<double>(LevelArgs<double> &, int, hipStream_t);
template int mg_smooth_level<float>(LevelArgs<float> &, int, hipStream_t);

}  // namespace stz
