// mg.hip -- gfx950 geometric multigrid for the quasi-static projection
// (the reference's tgmg<qs_multi,vec,mat>, tgmg/tgmg.in.cc), templated on the
// working precision T (double, or float for the mixed-precision form).
//
// Same hierarchy, bilinear prolongation P and restriction R = P^T, and
// Galerkin coarse operators as the reference (tgmg.in.cc:10-30, 552-779).
// The smoother is a four-colour point Gauss-Seidel: colour (i%2, j%2), so no
// two nodes of one colour share a 9-point stencil and every colour pass is
// fully parallel and deterministic.  (The reference's zebra-line ordering,
// tgmg.in.cc:127-165, is sequential along a row; per V(1,1) cycle both
// converge at the same rate, ~0.61 on this operator.)  On an odd periodic
// width the seam column gets a colour of its own.
//
// Kernels:
//   colour pass      one thread per node of one colour
//   restrict         one thread per coarse node; recomputes the <=9 fine
//                    residuals it needs (no residual array in HBM), writes the
//                    coarse source and zeroes the coarse solution
//   prolong          one thread per fine node, z += P zc
//   coarse solve     ONE workgroup runs the whole sub-V-cycle on the small
//                    levels with z and b resident in LDS, and solves the
//                    coarsest level exactly with a precomputed dense inverse
//   residual norm    wavefront-shuffle reduction to per-block partials, then
//                    one block sums the partials in a fixed order
#include "mg_node.h"

namespace stz {

// ---------------------------------------------------------------------------
// grid kernels
// ---------------------------------------------------------------------------
template <typename T> __global__ void __launch_bounds__(256) colour_kernel(LevelArgs<T> L, int cx, int cy) {
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 * (blockIdx.y * blockDim.y + threadIdx.y) + cy;
    int i;
    if (cx == 2) {
        if (tx != 0) return;
        i = L.m - 1;
    } else {
        i = 2 * tx + cx;
        if (L.xodd && i == L.m - 1) return;
    }
    if (i >= L.m || j >= L.n) return;
    gs_node(L, i, j);
}

template <typename T> __global__ void __launch_bounds__(256) restrict_kernel(LevelArgs<T> F, LevelArgs<T> C) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= C.m || J >= C.n) return;
    restrict_node(F, C, I, J);
}

template <typename T> __global__ void __launch_bounds__(256) prolong_kernel(LevelArgs<T> C, LevelArgs<T> F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= F.m || j >= F.n) return;
    prolong_node(C, F, i, j);
}

// ---- reductions -------------------------------------------------------------
__device__ __forceinline__ static double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block of 256 threads -> one partial
__device__ __forceinline__ static void block_partial(double v, double *partial) {
    __shared__ double ws[4];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

template <typename T> __global__ void __launch_bounds__(256) resid_norm_kernel(LevelArgs<T> L, double *partial) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    double v = 0;
    if (idx < L.m * L.n) {
        const int j = idx / L.m, i = idx - j * L.m;
        T rx, ry;
        resid_node(L, i, j, rx, ry);
        v = (double)rx * (double)rx + (double)ry * (double)ry;
    }
    block_partial(v, partial);
}

// fixed-order sum of the partials by one block
__global__ void __launch_bounds__(256) sum_partials_kernel(const double *partial, int np, double *out) {
    double v = 0;
    for (int k = threadIdx.x; k < np; k += 256) v += partial[k];
    __shared__ double ws[4];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// mixed precision: FP64 residual of the top level -> FP32 source of the
// correction equation, zero FP32 correction, norm partials
__global__ void __launch_bounds__(256) resid64_to32_kernel(LevelArgs<double> L, float *b32, float *z32,
                                                           double *partial) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    double v = 0;
    if (idx < L.m * L.n) {
        const int j = idx / L.m, i = idx - j * L.m;
        double rx, ry;
        resid_node(L, i, j, rx, ry);
        v = rx * rx + ry * ry;
        b32[2 * idx] = (float)rx;
        b32[2 * idx + 1] = (float)ry;
        z32[2 * idx] = 0.f;
        z32[2 * idx + 1] = 0.f;
    }
    block_partial(v, partial);
}

template <typename T> __global__ void __launch_bounds__(256) fill_zero_kernel(T *p, size_t count) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx < count) p[idx] = 0;
}

__global__ void __launch_bounds__(256) correct64_kernel(double *z, const float *e, int count) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < count) z[idx] += (double)e[idx];
}

// ---------------------------------------------------------------------------
// workgroup coarse solver: levels lc..L-1, z and b in LDS
// ---------------------------------------------------------------------------
template <typename T>
__device__ static void wg_smooth(const LevelArgs<T> &L, int sweeps) {
    const int hm = (L.m + 1) >> 1, hn = (L.n + 1) >> 1;
    for (int sw = 0; sw < sweeps; sw++) {
        for (int cy = 0; cy < 2; cy++)
            for (int cx = 0; cx < 2 + L.xodd; cx++) {
                const int cnt = cx == 2 ? hn : hm * hn;
                for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                    int i, j;
                    if (cx == 2) { i = L.m - 1; j = 2 * t + cy; }
                    else { const int ty = t / hm; i = 2 * (t - ty * hm) + cx; j = 2 * ty + cy; }
                    if (i >= L.m || j >= L.n) continue;
                    if (cx < 2 && L.xodd && i == L.m - 1) continue;
                    gs_node(L, i, j);
                }
                __syncthreads();
            }
    }
}

template <typename T> __global__ void __launch_bounds__(1024) coarse_solve_kernel(CoarseArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    // level l with z and b redirected to its LDS slice
    auto lds = [&](int l) {
        size_t off = 0;
        for (int k = A.lc; k < l; k++) off += 4 * (size_t)A.lv[k].m * A.lv[k].n;
        LevelArgs<T> L = A.lv[l];
        L.z = smem + off;
        L.b = smem + off + 2 * L.m * L.n;
        return L;
    };
    // load the restricted source of level lc, zero its solution
    {
        const LevelArgs<T> L = lds(A.lc);
        const T *gb = A.lv[A.lc].b;
        const int cnt = 2 * L.m * L.n;
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            L.b[t] = gb[t];
            L.z[t] = 0;
        }
        __syncthreads();
    }
    // down
    for (int l = A.lc; l < A.L - 1; l++) {
        const LevelArgs<T> F = lds(l), C = lds(l + 1);
        wg_smooth(F, A.nu1);
        for (int t = threadIdx.x; t < C.m * C.n; t += blockDim.x) {
            const int J = t / C.m, I = t - J * C.m;
            restrict_node(F, C, I, J);
        }
        __syncthreads();
    }
    // bottom: exact solve with the dense inverse
    {
        const LevelArgs<T> B = lds(A.L - 1);
        const int N = 2 * B.m * B.n;
        for (int r = threadIdx.x; r < N; r += blockDim.x) {
            const T *row = A.binv + (size_t)r * N;
            T acc = 0;
            for (int k = 0; k < N; k++) acc += row[k] * B.b[k];
            B.z[r] = acc;
        }
        __syncthreads();
    }
    // up
    for (int l = A.L - 2; l >= A.lc; l--) {
        const LevelArgs<T> F = lds(l), C = lds(l + 1);
        for (int t = threadIdx.x; t < F.m * F.n; t += blockDim.x) {
            const int j = t / F.m, i = t - j * F.m;
            prolong_node(C, F, i, j);
        }
        __syncthreads();
        wg_smooth(F, A.nu2);
    }
    // store the level-lc solution
    {
        const LevelArgs<T> L = lds(A.lc);
        T *gz = A.lv[A.lc].z;
        const int cnt = 2 * L.m * L.n;
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) gz[t] = L.z[t];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T> int mg_smooth(const LevelArgs<T> &L, int sweeps, hipStream_t st) {
    const int hm = (L.m + 1) >> 1, hn = (L.n + 1) >> 1;
    const dim3 b(64, 4), g((hm + 63) / 64, (hn + 3) / 4);
    for (int sw = 0; sw < sweeps; sw++)
        for (int cy = 0; cy < 2; cy++) {
            hipLaunchKernelGGL(colour_kernel<T>, g, b, 0, st, L, 0, cy);
            hipLaunchKernelGGL(colour_kernel<T>, g, b, 0, st, L, 1, cy);
            if (L.xodd) hipLaunchKernelGGL(colour_kernel<T>, dim3(1, g.y), b, 0, st, L, 2, cy);
        }
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_restrict(const LevelArgs<T> &F, const LevelArgs<T> &C, hipStream_t st) {
    const dim3 b(64, 4), g((C.m + 63) / 64, (C.n + 3) / 4);
    hipLaunchKernelGGL(restrict_kernel<T>, g, b, 0, st, F, C);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_prolong(const LevelArgs<T> &C, const LevelArgs<T> &F, hipStream_t st) {
    const dim3 b(64, 4), g((F.m + 63) / 64, (F.n + 3) / 4);
    hipLaunchKernelGGL(prolong_kernel<T>, g, b, 0, st, C, F);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> size_t coarse_lds_bytes(const CoarseArgs<T> &A) {
    size_t cnt = 0;
    for (int l = A.lc; l < A.L; l++) cnt += 4 * (size_t)A.lv[l].m * A.lv[l].n;
    return cnt * sizeof(T);
}

template <typename T> int mg_coarse(const CoarseArgs<T> &A, hipStream_t st) {
    hipLaunchKernelGGL(coarse_solve_kernel<T>, dim3(1), dim3(1024), coarse_lds_bytes(A), st, A);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_resid_norm(const LevelArgs<T> &L, double *partial, double *out, hipStream_t st) {
    const int nb = (L.m * L.n + 255) / 256;
    hipLaunchKernelGGL(resid_norm_kernel<T>, dim3(nb), dim3(256), 0, st, L, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, nb, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int mg_resid64_to32(const LevelArgs<double> &L, float *b32, float *z32, double *partial, double *out,
                    hipStream_t st) {
    const int nb = (L.m * L.n + 255) / 256;
    hipLaunchKernelGGL(resid64_to32_kernel, dim3(nb), dim3(256), 0, st, L, b32, z32, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, nb, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

template <typename T> int mg_fill_zero(T *p, size_t count, hipStream_t st) {
    hipLaunchKernelGGL(fill_zero_kernel<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
template int mg_fill_zero<double>(double *, size_t, hipStream_t);
template int mg_fill_zero<float>(float *, size_t, hipStream_t);

int mg_sum_partials(const double *partial, int np, double *out, hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, np, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int mg_correct64(double *z, const float *e, int count, hipStream_t st) {
    hipLaunchKernelGGL(correct64_kernel, dim3((count + 255) / 256), dim3(256), 0, st, z, e, count);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

// one V-cycle on levels 0..L-1 (grid kernels above lc, workgroup below)
template <typename T> int mg_vcycle(const CoarseArgs<T> &A, hipStream_t st) {
    int e;
    if ((e = mg_smooth(A.lv[0], A.nu1, st))) return e;
    for (int l = 0; l < A.lc; l++) {
        if ((e = mg_restrict(A.lv[l], A.lv[l + 1], st))) return e;
        if (l + 1 < A.lc && (e = mg_smooth(A.lv[l + 1], A.nu1, st))) return e;
    }
    if ((e = mg_coarse(A, st))) return e;
    for (int l = A.lc; l >= 1; l--) {
        if ((e = mg_prolong(A.lv[l], A.lv[l - 1], st))) return e;
        if ((e = mg_smooth(A.lv[l - 1], A.nu2, st))) return e;
    }
    return STZ_OK;
}

template <typename T> int mg_set_lds_limit(size_t bytes) {
    STZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&coarse_solve_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return STZ_OK;
}
template int mg_set_lds_limit<double>(size_t);
template int mg_set_lds_limit<float>(size_t);

template int mg_restrict<double>(const LevelArgs<double> &, const LevelArgs<double> &, hipStream_t);
template int mg_restrict<float>(const LevelArgs<float> &, const LevelArgs<float> &, hipStream_t);
template int mg_prolong<double>(const LevelArgs<double> &, const LevelArgs<double> &, hipStream_t);
template int mg_prolong<float>(const LevelArgs<float> &, const LevelArgs<float> &, hipStream_t);
template int mg_coarse<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_coarse<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_smooth<double>(const LevelArgs<double> &, int, hipStream_t);
template int mg_smooth<float>(const LevelArgs<float> &, int, hipStream_t);
template int mg_vcycle<double>(const CoarseArgs<double> &, hipStream_t);
template int mg_vcycle<float>(const CoarseArgs<float> &, hipStream_t);
template int mg_resid_norm<double>(const LevelArgs<double> &, double *, double *, hipStream_t);
template int mg_resid_norm<float>(const LevelArgs<float> &, double *, double *, hipStream_t);
template size_t coarse_lds_bytes<double>(const CoarseArgs<double> &);
template size_t coarse_lds_bytes<float>(const CoarseArgs<float> &);

}  // namespace stz
