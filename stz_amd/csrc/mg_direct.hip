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
// Device buffer of a direct level (direct_elems values of T):
//   Gh  [k][r][q]  complex inverse of mode k, weighted for the inverse DFT
//   tw  [k][i]     (cos, sin) of theta_k i
//   Xh  [k][r]     the solved modes (written by the mode launch)
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
constexpr int kRowBlocks = 4;      // row blocks per mode in the mode launch
constexpr int kRowLanes = 8;       // lanes that share one row of a mode's matvec
constexpr int kInvThreads = 64;

__host__ __device__ static size_t gh_elems(int M, int Nn) { return (size_t)2 * (M / 2 + 1) * (2 * Nn) * (2 * Nn); }
__host__ __device__ static size_t tw_elems(int M) { return (size_t)2 * (M / 2 + 1) * M; }
__host__ __device__ static size_t xh_elems(int M, int Nn) { return (size_t)2 * (M / 2 + 1) * (2 * Nn); }
size_t direct_elems(int M, int Nn) { return gh_elems(M, Nn) + tw_elems(M) + xh_elems(M, Nn); }

template <typename T> size_t direct_lds_bytes(int M, int Nn) {
    return (size_t)(M * 2 * Nn + 2 * M + 2 * 2 * Nn) * sizeof(T);
}

// mode launch body: workgroup blk = k * kRowBlocks + row block
template <typename T>
__device__ __forceinline__ static void direct_mode_body(int M, int Nn, T *__restrict__ D, const T *__restrict__ b,
                                                        int blk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N2 = 2 * Nn, nb = M * N2;
    const int k = blk / kRowBlocks, rb = blk % kRowBlocks;
    const T *Gh = D + (size_t)2 * k * N2 * N2;
    const T *tw = D + gh_elems(M, Nn) + (size_t)2 * k * M;
    T *Xh = D + gh_elems(M, Nn) + tw_elems(M) + (size_t)2 * k * N2;
    T *bs = reinterpret_cast<T *>(smem_raw);  // b, natural order: 2 (i + M j') + c'
    T *ts = bs + nb;                          // (cos, sin) of theta_k i
    T *bh = ts + 2 * M;                       // DFT of the source, per q = 2 j' + c'
    for (int t = threadIdx.x; t < nb; t += kModeThreads) bs[t] = b[t];
    for (int t = threadIdx.x; t < 2 * M; t += kModeThreads) ts[t] = tw[t];
    __syncthreads();
    // bhat(q) = sum_i b(i, q) e^{-i theta_k i}
    // (two threads per q, each summing half of the i; kModeThreads >= 2 N2
    // is not required: the loop strides over q)
    for (int t = threadIdx.x; t < 2 * ((N2 + kModeThreads / 2 - 1) / (kModeThreads / 2)) * (kModeThreads / 2);
         t += kModeThreads) {
        const int q = t >> 1, h = t & 1;
        T re = 0, im = 0;
        if (q < N2) {
            const T *bq = bs + 2 * M * (q >> 1) + (q & 1);
            const int i0 = h * (M / 2), i1 = h ? M : M / 2;
#pragma unroll 8
            for (int i = i0; i < i1; i++) {
                const T v = bq[2 * i];
                re = fmav(v, ts[2 * i], re);
                im = fmav(-v, ts[2 * i + 1], im);
            }
        }
        re += __shfl_xor(re, 1, 2);
        im += __shfl_xor(im, 1, 2);
        if (q < N2 && h == 0) {
            bh[2 * q] = re;
            bh[2 * q + 1] = im;
        }
    }
    __syncthreads();
    // rows r0 .. r0 + rows of the mode's (weighted) inverse applied to bhat
    const int rows = (N2 + kRowBlocks - 1) / kRowBlocks, r0 = rb * rows;
    const int lane = threadIdx.x % kRowLanes;
    for (int rr = threadIdx.x / kRowLanes; rr < rows; rr += kModeThreads / kRowLanes) {
        const int r = r0 + rr;
        const bool act = r < N2;
        T re = 0, im = 0;
        if (act) {
            const T *g = Gh + (size_t)2 * r * N2;
            for (int q = lane; q < N2; q += kRowLanes) {
                const T gr = g[2 * q], gi = g[2 * q + 1], br = bh[2 * q], bi = bh[2 * q + 1];
                re = fmav(gr, br, re);
                re = fmav(-gi, bi, re);
                im = fmav(gr, bi, im);
                im = fmav(gi, br, im);
            }
        }
        // fixed-order sum over the row's lanes (consecutive lanes of one wave)
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) {
            re += __shfl_xor(re, off, kRowLanes);
            im += __shfl_xor(im, off, kRowLanes);
        }
        if (act && lane == 0) {
            Xh[2 * r] = re;
            Xh[2 * r + 1] = im;
        }
    }
}

// inverse launch body: workgroup r = 2 j + c; z(i, r) = Re sum_k Xh[k][r]
// e^{i theta_k i} (the weights 1/M and 2/M are folded into Gh)
template <typename T>
__device__ __forceinline__ static void direct_inverse_body(int M, int Nn, const T *__restrict__ D, T *__restrict__ z,
                                                           int r) {
    const int N2 = 2 * Nn, K = M / 2 + 1;
    const T *tw = D + gh_elems(M, Nn);
    const T *Xh = tw + tw_elems(M);
    const int j = r >> 1, c = r & 1;
    for (int i = threadIdx.x; i < M; i += kInvThreads) {
        T s = 0;
        for (int k = 0; k < K; k++) {
            const T xr = Xh[2 * (k * N2 + r)], xi = Xh[2 * (k * N2 + r) + 1];
            const T co = tw[2 * (k * M + i)], si = tw[2 * (k * M + i) + 1];
            s = fmav(xr, co, s);
            s = fmav(-xi, si, s);
        }
        z[2 * (i + M * j) + c] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(kModeThreads) direct_mode_kernel(int M, int Nn, T *D, const T *b) {
    direct_mode_body(M, Nn, D, b, blockIdx.x);
}
template <typename T>
__global__ void __launch_bounds__(kInvThreads) direct_inverse_kernel(int M, int Nn, const T *D, T *z) {
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

template <typename T> int mg_direct_solve(const LevelArgs<T> &L, T *D, hipStream_t st) {
    const int M = L.m, Nn = L.n;
    hipLaunchKernelGGL(direct_mode_kernel<T>, dim3(direct_mode_workgroups(M, Nn)), dim3(kModeThreads),
                       direct_lds_bytes<T>(M, Nn), st, M, Nn, D, L.b);
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
        double *tw = D.data() + gh_elems(M, Nn) + (size_t)2 * k * M;
        for (int i = 0; i < M; i++) {
            const double a = 2 * pi * (double)((k * i) % M) / M;
            tw[2 * i] = std::cos(a);
            tw[2 * i + 1] = std::sin(a);
        }
    }
    return true;
}

template int mg_direct_solve<double>(const LevelArgs<double> &, double *, hipStream_t);
template int mg_direct_solve<float>(const LevelArgs<float> &, float *, hipStream_t);
template size_t direct_lds_bytes<double>(int, int);
template size_t direct_lds_bytes<float>(int, int);
template int mg_direct_set_lds_limit<double>(size_t);
template int mg_direct_set_lds_limit<float>(size_t);
template int batch_direct<double>(const DirectEntry<double> *, const int *, int, const int *, int, int, size_t,
                                  const EnsState *, hipStream_t);
template int batch_direct<float>(const DirectEntry<float> *, const int *, int, const int *, int, int, size_t,
                                 const EnsState *, hipStream_t);

}  // namespace stz
