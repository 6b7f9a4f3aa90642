// mg_direct.hip -- exact solve of one coarse multigrid level whose operator
// is invariant along the periodic x direction (every level of the periodic
// simple-shear projection: qs_multi's operator does not depend on i, and
// bilinear Galerkin coarsening of an even-width x-invariant level stays
// x-invariant).
//
// A discrete Fourier transform along x turns such a level of M x Nn nodes
// into M/2 + 1 independent complex systems of N2 = 2 Nn unknowns (mode k:
// the row stencils' blocks summed as sum_dx S_j[dx,dy] e^{i theta_k dx},
// theta_k = 2 pi k / M; the other modes are their conjugates).  The host
// inverts each once per operator (Gauss-Jordan with partial pivoting) and
// folds the inverse transform's weights into the inverses; the device then
// solves the level in two launches:
//   mode launch     one workgroup per (mode, block of rows): the DFT of the
//                   source along x for its mode, then its rows of the mode's
//                   inverse applied to it
//   inverse launch  one workgroup per row (j, component): the real inverse
//                   DFT of the solved modes
// Every sum runs in a fixed order, so the result is deterministic.  Per
// solve: M N2 (M/2 + 1) complex multiply-adds for the transforms and
// (M/2 + 1) N2^2 for the modes, instead of a V-cycle's chain of
// barrier-separated smoothing passes on ever smaller levels.
//
// Both launches are latency bound (a 64 x 65 level: 66 KB of source, 9 MB of
// mode inverses in FP64 over 132 workgroups), so each is laid out for as few
// dependent memory round trips as possible: the mode launch issues its rows
// of the inverse before it stages the source, stages the source with all of
// a thread's loads in flight at once, and transforms it from LDS (transposed,
// so a wave reads consecutive addresses); the twiddle factors come from a
// table of the M roots of unity indexed by k i mod M.
//
// Device buffer of a direct level (direct_elems values of T):
//   Gh  [k][r][q]  complex inverse of mode k, weighted for the inverse DFT
//   tw  [m]        (cos, sin) of 2 pi m / M, m < M
//   Xh  [r][k]     the solved modes (written by the mode launch)
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "batch.h"
#include "device.h"
#include "mg.h"
#include "mg_node.h"

namespace stz {

constexpr int kModeThreads = 256;
constexpr int kModeWaves = kModeThreads / 64;
constexpr int kRowBlocks = 4;      // row blocks per mode in the mode launch
constexpr int kMaxRows = 9;        // rows of a mode's inverse per wave
constexpr int kMaxQ = 3;           // entries of a row per lane (q = lane + 64 s)
constexpr int kStage = 17;         // source nodes staged per thread per pass
constexpr int kInvThreads = 64;
template <typename T> struct Pair;
template <> struct Pair<double> { typedef double2 type; };
template <> struct Pair<float> { typedef float2 type; };
// the largest N2 = 2 Nn the mode launch covers
constexpr int kMaxN2 = kRowBlocks * kModeWaves * kMaxRows < 64 * kMaxQ ? kRowBlocks * kModeWaves * kMaxRows
                                                                        : 64 * kMaxQ;

__host__ __device__ static size_t gh_elems(int M, int Nn) { return (size_t)2 * (M / 2 + 1) * (2 * Nn) * (2 * Nn); }
__host__ __device__ static size_t tw_elems(int M) { return (size_t)2 * M; }
__host__ __device__ static size_t xh_elems(int M, int Nn) { return (size_t)2 * (M / 2 + 1) * (2 * Nn); }
size_t direct_elems(int M, int Nn) { return gh_elems(M, Nn) + tw_elems(M) + xh_elems(M, Nn); }

// LDS of the mode launch: the source transposed (M rows of N2 + 1), the
// twiddle table, the transformed source, all FP64
size_t direct_lds_bytes(int M, int Nn) { return (size_t)(M * (2 * Nn + 1) + 2 * M + 2 * 2 * Nn) * sizeof(double); }

bool direct_fits(int M, int Nn) {
    return M >= 2 && M % 2 == 0 && M <= 2 * kInvThreads && Nn >= 1 && 2 * Nn <= kMaxN2 &&
           direct_lds_bytes(M, Nn) <= 150 * 1024;
}

// mode launch body: workgroup blk = k * kRowBlocks + row block; the
// level's source is in T, everything else in FP64
template <typename T>
__device__ __forceinline__ static void direct_mode_body(int M, int Nn, double *__restrict__ D,
                                                        const T *__restrict__ b, int blk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N2 = 2 * Nn, K = M / 2 + 1, P = N2 + 1, nodes = M * Nn;
    const int k = blk / kRowBlocks, rb = blk % kRowBlocks;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const double *Gh = D + (size_t)2 * k * N2 * N2;
    const double *tw = D + gh_elems(M, Nn);
    double *Xh = D + gh_elems(M, Nn) + tw_elems(M);
    double *bt = reinterpret_cast<double *>(smem_raw);  // bt[i P + q] = b(i, q), q = 2 j' + c'
    double *ts = bt + M * P;                            // (cos, sin) of 2 pi m / M
    double *bh = ts + 2 * M;                            // transformed source, per q

    typedef typename Pair<T>::type V2;
    // this wave's rows of the mode's inverse: r = r0 + wave + 4 s
    const int rows = (N2 + kRowBlocks - 1) / kRowBlocks, r0 = rb * rows, r1 = min(r0 + rows, N2);
    double g[kMaxRows][kMaxQ][2];
#pragma unroll
    for (int s = 0; s < kMaxRows; s++) {
        const int r = r0 + wave + kModeWaves * s;
#pragma unroll
        for (int u = 0; u < kMaxQ; u++) {
            const int q = lane + 64 * u;
            const bool ok = r < r1 && q < N2;
            const double2 gv = reinterpret_cast<const double2 *>(Gh)[(size_t)(ok ? r : 0) * N2 + (ok ? q : 0)];
            g[s][u][0] = ok ? gv.x : 0.0;
            g[s][u][1] = ok ? gv.y : 0.0;
        }
    }
    // stage the source (vec2 nodes) transposed, and the twiddle table
    const V2 *bv = reinterpret_cast<const V2 *>(b);
    for (int base = 0; base < nodes; base += kStage * kModeThreads) {
        V2 v[kStage];
        // unconditional loads (of node 0 past the end), so that all of a
        // thread's loads are in flight together
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int nd = base + threadIdx.x + kModeThreads * u;
            v[u] = bv[nd < nodes ? nd : 0];
        }
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int nd = base + threadIdx.x + kModeThreads * u;
            if (nd < nodes) {
                const int i = nd % M, jj = nd / M;
                bt[i * P + 2 * jj] = (double)v[u].x;
                bt[i * P + 2 * jj + 1] = (double)v[u].y;
            }
        }
    }
    for (int t = threadIdx.x; t < 2 * M; t += kModeThreads) ts[t] = tw[t];
    __syncthreads();
    // bhat(q) = sum_i b(i, q) e^{-i theta_k i}, i ascending
    for (int q = threadIdx.x; q < N2; q += kModeThreads) {
        double re = 0, im = 0;
        int m = 0;
#pragma unroll 8
        for (int i = 0; i < M; i++) {
            const double v = bt[i * P + q];
            re = fmav(v, ts[2 * m], re);
            im = fmav(-v, ts[2 * m + 1], im);
            m += k;
            if (m >= M) m -= M;
        }
        bh[2 * q] = re;
        bh[2 * q + 1] = im;
    }
    __syncthreads();
    // rows of the mode's (weighted) inverse applied to bhat: each lane sums
    // its q in ascending order, then a fixed xor tree over the wave
    double br[kMaxQ], bi[kMaxQ];
#pragma unroll
    for (int u = 0; u < kMaxQ; u++) {
        const int q = lane + 64 * u;
        br[u] = q < N2 ? bh[2 * q] : 0.0;
        bi[u] = q < N2 ? bh[2 * q + 1] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kMaxRows; s++) {
        const int r = r0 + wave + kModeWaves * s;
        if (r0 + wave + kModeWaves * s >= r1) break;  // uniform over the wave
        double re = 0, im = 0;
#pragma unroll
        for (int u = 0; u < kMaxQ; u++) {
            re = fmav(g[s][u][0], br[u], re);
            re = fmav(-g[s][u][1], bi[u], re);
            im = fmav(g[s][u][0], bi[u], im);
            im = fmav(g[s][u][1], br[u], im);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off);
            im += __shfl_xor(im, off);
        }
        if (lane == 0) reinterpret_cast<double2 *>(Xh)[(size_t)r * K + k] = make_double2(re, im);
    }
}

// inverse launch body: workgroup r = 2 j + c; z(i, r) = Re sum_k Xh[r][k]
// e^{i theta_k i}, k ascending, in FP64, rounded to the level's T (the
// weights 1/M and 2/M are folded into Gh)
template <typename T>
__device__ __forceinline__ static void direct_inverse_body(int M, int Nn, const double *__restrict__ D,
                                                           T *__restrict__ z, int r) {
    __shared__ double xs[2 * (kInvThreads + 1)];
    __shared__ double ts[2 * 2 * kInvThreads];
    const int K = M / 2 + 1;
    const double *tw = D + gh_elems(M, Nn);
    const double *Xh = tw + tw_elems(M) + (size_t)2 * r * K;
    const int j = r >> 1, c = r & 1;
    for (int t = threadIdx.x; t < 2 * K; t += kInvThreads) xs[t] = Xh[t];
    for (int t = threadIdx.x; t < 2 * M; t += kInvThreads) ts[t] = tw[t];
    __syncthreads();
    for (int i = threadIdx.x; i < M; i += kInvThreads) {
        double s = 0;
        int m = 0;
#pragma unroll 8
        for (int kk = 0; kk < K; kk++) {
            s = fmav(xs[2 * kk], ts[2 * m], s);
            s = fmav(-xs[2 * kk + 1], ts[2 * m + 1], s);
            m += i;
            if (m >= M) m -= M;
        }
        z[2 * (i + M * j) + c] = (T)s;
    }
}
template <typename T>
__global__ void __launch_bounds__(kModeThreads) direct_mode_kernel(int M, int Nn, double *D, const T *b) {
    direct_mode_body(M, Nn, D, b, blockIdx.x);
}
template <typename T>
__global__ void __launch_bounds__(kInvThreads) direct_inverse_kernel(int M, int Nn, const double *D, T *z) {
    direct_inverse_body(M, Nn, D, z, blockIdx.x);
}

// the same over the members of an ensemble (batch.h); members whose slot
// flag is clear return at once
template <typename T>
__global__ void __launch_bounds__(kModeThreads) direct_mode_batch_kernel(const DirectEntry<T> *E, const int *map,
                                                                         int ne, const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const DirectEntry<T> &x = E[el.x];
    if (St && !St[x.member].active) return;
    direct_mode_body(x.M, x.Nn, x.D, x.b, el.y);
}
template <typename T>
__global__ void __launch_bounds__(kInvThreads) direct_inverse_batch_kernel(const DirectEntry<T> *E, const int *map,
                                                                           int ne, const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const DirectEntry<T> &x = E[el.x];
    if (St && !St[x.member].active) return;
    direct_inverse_body(x.M, x.Nn, x.D, x.z, el.y);
}

int direct_mode_workgroups(int M, int Nn) {
    (void)Nn;
    return (M / 2 + 1) * kRowBlocks;
}
int direct_inverse_workgroups(int M, int Nn) {
    (void)M;
    return 2 * Nn;
}

template <typename T> int mg_direct_solve(const LevelArgs<T> &L, double *D, hipStream_t st) {
    const int M = L.m, Nn = L.n;
    hipLaunchKernelGGL(direct_mode_kernel<T>, dim3(direct_mode_workgroups(M, Nn)), dim3(kModeThreads),
                       direct_lds_bytes(M, Nn), st, M, Nn, D, L.b);
    hipLaunchKernelGGL(direct_inverse_kernel<T>, dim3(direct_inverse_workgroups(M, Nn)), dim3(kInvThreads), 0, st,
                       M, Nn, D, L.z);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_direct_set_lds_limit(size_t bytes) {
    static size_t set = 0;
    if (bytes <= set) return STZ_OK;
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&direct_mode_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&direct_mode_batch_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    set = bytes;
    return STZ_OK;
}

template <typename T>
int batch_direct(const DirectEntry<T> *E, const int *mode_map, int mode_total, const int *inv_map, int inv_total,
                 int ne, size_t lds, const EnsState *St, hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(direct_mode_batch_kernel<T>, dim3(mode_total), dim3(kModeThreads), lds, st, E, mode_map, ne,
                       St);
    hipLaunchKernelGGL(direct_inverse_batch_kernel<T>, dim3(inv_total), dim3(kInvThreads), 0, st, E, inv_map, ne,
                       St);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// host: the device buffer of an x-invariant level with row stencils S (40
// per row: blocks dl,dc,dr,cl,cc,cr,ul,uc,ur, then the centre inverse), M
// nodes per row (even), Nn rows, in the layout above (Xh zero).  Returns
// false if a mode's system is singular.
bool direct_inverse(const double *S, int M, int Nn, std::vector<double> &D) {
    typedef std::complex<double> cd;
    const int N2 = 2 * Nn, K = M / 2 + 1;
    D.assign(direct_elems(M, Nn), 0.0);
    std::vector<cd> A((size_t)N2 * N2), I((size_t)N2 * N2);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < K; k++) {
        const double th = 2 * pi * k / M;
        const cd e[3] = {std::polar(1.0, -th), cd(1, 0), std::polar(1.0, th)};
        std::fill(A.begin(), A.end(), cd(0, 0));
        std::fill(I.begin(), I.end(), cd(0, 0));
        for (int r = 0; r < N2; r++) I[(size_t)r * N2 + r] = 1;
        for (int j = 0; j < Nn; j++) {
            const double *s = S + (size_t)j * 40;
            for (int kb = 0; kb < 9; kb++) {
                const int dx = kb % 3 - 1, jj = j + kb / 3 - 1;
                if (jj < 0 || jj >= Nn) continue;
                const cd w = e[dx + 1];
                A[(size_t)(2 * j) * N2 + 2 * jj] += s[4 * kb] * w;
                A[(size_t)(2 * j) * N2 + 2 * jj + 1] += s[4 * kb + 1] * w;
                A[(size_t)(2 * j + 1) * N2 + 2 * jj] += s[4 * kb + 2] * w;
                A[(size_t)(2 * j + 1) * N2 + 2 * jj + 1] += s[4 * kb + 3] * w;
            }
        }
        // Gauss-Jordan with partial pivoting
        for (int col = 0; col < N2; col++) {
            int piv = col;
            for (int r = col + 1; r < N2; r++)
                if (std::abs(A[(size_t)r * N2 + col]) > std::abs(A[(size_t)piv * N2 + col])) piv = r;
            if (std::abs(A[(size_t)piv * N2 + col]) == 0) return false;
            if (piv != col)
                for (int c = 0; c < N2; c++) {
                    std::swap(A[(size_t)piv * N2 + c], A[(size_t)col * N2 + c]);
                    std::swap(I[(size_t)piv * N2 + c], I[(size_t)col * N2 + c]);
                }
            const cd d = 1.0 / A[(size_t)col * N2 + col];
            for (int c = 0; c < N2; c++) {
                A[(size_t)col * N2 + c] *= d;
                I[(size_t)col * N2 + c] *= d;
            }
            for (int r = 0; r < N2; r++) {
                if (r == col) continue;
                const cd f = A[(size_t)r * N2 + col];
                if (f == cd(0, 0)) continue;
                for (int c = 0; c < N2; c++) {
                    A[(size_t)r * N2 + c] -= f * A[(size_t)col * N2 + c];
                    I[(size_t)r * N2 + c] -= f * I[(size_t)col * N2 + c];
                }
            }
        }
        // z = (1/M) [X_0 + X_{M/2} (-1)^i] + (2/M) sum_{0<k<M/2} Re(X_k e^{i theta_k i})
        const double w = (k == 0 || 2 * k == M) ? 1.0 / M : 2.0 / M;
        double *g = D.data() + (size_t)2 * k * N2 * N2;
        for (size_t t = 0; t < (size_t)N2 * N2; t++) {
            g[2 * t] = w * I[t].real();
            g[2 * t + 1] = w * I[t].imag();
        }
    }
    double *tw = D.data() + gh_elems(M, Nn);
    for (int i = 0; i < M; i++) {
        const double a = 2 * pi * (double)i / M;
        tw[2 * i] = std::cos(a);
        tw[2 * i + 1] = std::sin(a);
    }
    return true;
}

template int mg_direct_solve<double>(const LevelArgs<double> &, double *, hipStream_t);
template int mg_direct_solve<float>(const LevelArgs<float> &, double *, hipStream_t);
template int mg_direct_set_lds_limit<double>(size_t);
template int mg_direct_set_lds_limit<float>(size_t);
template int batch_direct<double>(const DirectEntry<double> *, const int *, int, const int *, int, int, size_t,
                                  const EnsState *, hipStream_t);
template int batch_direct<float>(const DirectEntry<float> *, const int *, int, const int *, int, int, size_t,
                                 const EnsState *, hipStream_t);

}  // namespace stz
