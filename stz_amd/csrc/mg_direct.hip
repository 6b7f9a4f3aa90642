// mg_direct.hip -- exact solve of one coarse multigrid level whose operator
// is invariant along the periodic x direction (every level of the periodic
// simple-shear projection: qs_multi's operator does not depend on i, and
// bilinear Galerkin coarsening of an even-width x-invariant level stays
// x-invariant).
//
// On such a level the inverse is block-circulant in x: z(i,j) = sum over
// (delta, j') of G[delta](j, j') b(i - delta mod M, j'), with 2x2 blocks.
// The host forms G once per operator from the level's row stencils by a
// discrete Fourier transform in x: mode k turns the level into a
// (2 Nn) x (2 Nn) complex block-tridiagonal system A_k (blocks
// sum_dx S_j[dx,dy] e^{i theta_k dx}), whose inverse is taken by Gauss-Jordan
// elimination; G[delta] = (1/M) sum_k inv(A_k) e^{i theta_k delta}, using
// inv(A_{M-k}) = conj(inv(A_k)).
//
// The device applies G in one launch: the V-cycle's coarse solve becomes one
// parallel matrix-vector product instead of a chain of barrier-separated
// smoothing passes on ever smaller levels.  Each workgroup owns 32 outputs
// (i0..i0+31 of one row j, one component c); it stages all of b and its
// (j,c) slice of G in LDS, each of its 8 waves sums over one eighth of the
// (j',c') range, and the partial sums meet in LDS in a fixed order, so the
// result is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "batch.h"
#include "device.h"
#include "mg.h"
#include "mg_node.h"

namespace stz {

constexpr int kDirThreads = 256;
constexpr int kDirGroups = kDirThreads / 32;  // partial sums per output

// z of the level from its b; workgroup blk = (j * 2 + c) * ceil(M / 32) +
// block of 32 outputs i (those beyond M idle but take part in the barriers)
template <typename T>
__device__ __forceinline__ static void direct_solve_body(int M, int Nn, const T *__restrict__ G,
                                                         const T *__restrict__ b, T *__restrict__ z, int blk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *bs = reinterpret_cast<T *>(smem_raw);  // b, natural order: 2 (i + M j') + c'
    const int N2 = 2 * Nn, nb = M * N2;
    T *gs = bs + nb;                           // G slice: delta * N2 + q
    T *ps = gs + nb;                           // partial sums [group][output]
    const int nib = (M + 31) / 32, ib = blk % nib, jc = blk / nib;
    const T *Gj = G + (size_t)jc * nb;
    for (int t = threadIdx.x; t < nb; t += kDirThreads) {
        bs[t] = b[t];
        gs[t] = Gj[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const int i = ib * 32 + lane;
    // this group's share of q = 2 j' + c'
    const int qper = (N2 + kDirGroups - 1) / kDirGroups, q0 = grp * qper,
              q1 = i < M ? (q0 + qper < N2 ? q0 + qper : N2) : q0;
    T acc = 0;
    for (int d = 0; d < M && q0 < q1; d++) {
        int is = i - d;
        is = is < 0 ? is + M : is;
        const T *g = gs + d * N2;
        const T *bb = bs + 2 * is;
        for (int q = q0; q < q1; q++) {
            const int jp = q >> 1, cp = q & 1;
            acc = fmav(g[q], bb[2 * M * jp + cp], acc);
        }
    }
    ps[grp * 32 + lane] = acc;
    __syncthreads();
    if (grp == 0 && i < M) {
        T s = ps[lane];
        for (int g2 = 1; g2 < kDirGroups; g2++) s += ps[g2 * 32 + lane];
        const int j = jc >> 1, c = jc & 1;
        z[2 * (i + M * j) + c] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(kDirThreads) direct_solve_kernel(int M, int Nn, const T *__restrict__ G,
                                                                   const T *__restrict__ b, T *__restrict__ z) {
    direct_solve_body(M, Nn, G, b, z, blockIdx.x);
}

// the same over the members of an ensemble (batch.h); members whose slot
// flag is clear return at once
template <typename T>
__global__ void __launch_bounds__(kDirThreads) direct_solve_batch_kernel(const DirectEntry<T> *E, const int *map,
                                                                         int ne, const EnsState *St) {
    const int2 el = batch_slot(map, blockIdx.x);
    const DirectEntry<T> &x = E[el.x];
    if (St && !St[x.member].active) return;
    direct_solve_body(x.M, x.Nn, x.G, x.b, x.z, el.y);
}

template <typename T> size_t direct_lds_bytes(int M, int Nn) {
    return (size_t)(2 * M * 2 * Nn + kDirGroups * 32) * sizeof(T);
}

template <typename T> int mg_direct_solve(const LevelArgs<T> &L, const T *G, hipStream_t st) {
    const int M = L.m, Nn = L.n;
    hipLaunchKernelGGL(direct_solve_kernel<T>, dim3(2 * Nn * ((M + 31) / 32)), dim3(kDirThreads),
                       direct_lds_bytes<T>(M, Nn), st, M, Nn, G, L.b, L.z);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_direct_set_lds_limit(size_t bytes) {
    static size_t set = 0;
    if (bytes <= set) return STZ_OK;
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&direct_solve_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&direct_solve_batch_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    set = bytes;
    return STZ_OK;
}

template <typename T>
int batch_direct(const DirectEntry<T> *E, const int *map, int ne, int total, size_t lds, const EnsState *St,
                 hipStream_t st) {
    if (ne == 0) return STZ_OK;
    hipLaunchKernelGGL(direct_solve_batch_kernel<T>, dim3(total), dim3(kDirThreads), lds, st, E, map, ne, St);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
template int batch_direct<double>(const DirectEntry<double> *, const int *, int, int, size_t, const EnsState *,
                                  hipStream_t);
template int batch_direct<float>(const DirectEntry<float> *, const int *, int, int, size_t, const EnsState *,
                                 hipStream_t);
int direct_workgroups(int M, int Nn) { return 2 * Nn * ((M + 31) / 32); }

// host: the block-circulant inverse of an x-invariant level with row
// stencils S (40 per row: blocks dl,dc,dr,cl,cc,cr,ul,uc,ur, then the
// centre inverse), M nodes per row (even), Nn rows.  Layout of G:
// [(2 j + c) * M + delta] * (2 Nn) + (2 j' + c').  Returns false if a mode's
// system is singular.
bool direct_inverse(const double *S, int M, int Nn, std::vector<double> &G) {
    typedef std::complex<double> cd;
    const int N2 = 2 * Nn, K = M / 2 + 1;
    std::vector<cd> Gk((size_t)K * N2 * N2);
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
        std::copy(I.begin(), I.end(), Gk.begin() + (size_t)k * N2 * N2);
    }
    // G[delta] = (1/M) sum_k Gk e^{i theta_k delta}, conjugate pairs folded
    G.assign((size_t)N2 * M * N2, 0.0);
    std::vector<cd> ph((size_t)K * M);
    for (int k = 0; k < K; k++)
        for (int d = 0; d < M; d++) ph[(size_t)k * M + d] = std::polar(1.0, 2 * pi * (double)((k * d) % M) / M);
    for (int r = 0; r < N2; r++)
        for (int d = 0; d < M; d++)
            for (int q = 0; q < N2; q++) {
                double sum = 0;
                for (int k = 0; k < K; k++) {
                    const cd v = Gk[((size_t)k * N2 + r) * N2 + q] * ph[(size_t)k * M + d];
                    const bool self = k == 0 || 2 * k == M;
                    sum += self ? v.real() : 2 * v.real();
                }
                G[((size_t)r * M + d) * N2 + q] = sum / M;
            }
    return true;
}

template int mg_direct_solve<double>(const LevelArgs<double> &, const double *, hipStream_t);
template int mg_direct_solve<float>(const LevelArgs<float> &, const float *, hipStream_t);
template size_t direct_lds_bytes<double>(int, int);
template size_t direct_lds_bytes<float>(int, int);
template int mg_direct_set_lds_limit<double>(size_t);
template int mg_direct_set_lds_limit<float>(size_t);

}  // namespace stz
