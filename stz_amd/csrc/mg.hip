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
// workgroup coarse solver: levels lc..L-1 entirely in LDS (z, b, stencils, a
// residual scratch and the coarsest level's dense inverse), kCoarseThreads
// threads, one barrier per pass.  z and b of a level are stored colour-split:
// node (i,j) in parity array (i&1) + 2(j&1) at ((j>>1), (i>>1)), so the lanes
// of a colour pass, and the neighbours they read, are at consecutive
// addresses.  Each node update issues all of its LDS loads before any
// arithmetic (no branch between them: rows beyond a wall are read from a
// clamped row and zeroed), which keeps one pass to one LDS round trip.
// ---------------------------------------------------------------------------
constexpr int kCoarseThreads = 512;

template <typename T> struct WGLevel {
    int m, n, rs, cs, xodd, hm, hn;
    int zo, bo, so;
};

template <typename T> __device__ __forceinline__ static WGLevel<T> wg_level(const CoarseArgs<T> &A, int l) {
    WGLevel<T> W;
    W.m = A.lv[l].m;
    W.n = A.lv[l].n;
    W.rs = A.lv[l].rs;
    W.cs = A.lv[l].cs;
    W.xodd = A.lv[l].xodd;
    W.hm = (W.m + 1) >> 1;
    W.hn = (W.n + 1) >> 1;
    W.zo = A.zoff[l];
    W.bo = A.boff[l];
    W.so = A.soff[l];
    return W;
}

// element offset of node (i,j)'s x component in a colour-split level array
template <typename T> __device__ __forceinline__ static int cidx(const WGLevel<T> &W, int i, int j) {
    return 2 * ((((i & 1) + 2 * (j & 1)) * W.hn + (j >> 1)) * W.hm + (i >> 1));
}

template <typename T> __device__ __forceinline__ static vec2s<T> lds2(const T *sm, int k) {
    return *reinterpret_cast<const vec2s<T> *>(sm + k);
}

// off-diagonal sum at node (i,j) of an LDS-resident level
template <typename T>
__device__ __forceinline__ static void wg_offdiag(const T *sm, const WGLevel<T> &W, int i, int j, const T *s, T &ax,
                                                  T &ay) {
    const int m = W.m;
    const int il = i == 0 ? m - 1 : i - 1, ir = i == m - 1 ? 0 : i + 1;
    const int jd = j > 0 ? j - 1 : j, ju = j < W.n - 1 ? j + 1 : j;
    const T *z = sm + W.zo;
    vec2s<T> dl = lds2(z, cidx(W, il, jd)), dc = lds2(z, cidx(W, i, jd)), dr = lds2(z, cidx(W, ir, jd));
    const vec2s<T> cl = lds2(z, cidx(W, il, j)), cr = lds2(z, cidx(W, ir, j));
    vec2s<T> ul = lds2(z, cidx(W, il, ju)), uc = lds2(z, cidx(W, i, ju)), ur = lds2(z, cidx(W, ir, ju));
    const vec2s<T> zero = {0, 0};
    if (j == 0) dl = dc = dr = zero;
    if (j == W.n - 1) ul = uc = ur = zero;
    offdiag9<T>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
}

// residual at (i,j) written in natural order to r
template <typename T> __device__ __forceinline__ static void wg_resid(T *sm, const WGLevel<T> &W, int ro, int i,
                                                                      int j) {
    const T *s = sm + W.so + (j * W.rs + i * W.cs) * 40;
    T ax, ay, rx, ry;
    wg_offdiag(sm, W, i, j, s, ax, ay);
    const int k = cidx(W, i, j);
    resid_value<T>(s, sm[W.bo + k], sm[W.bo + k + 1], ax, ay, lds2(sm + W.zo, k), rx, ry);
    sm[ro + 2 * (i + W.m * j)] = rx;
    sm[ro + 2 * (i + W.m * j) + 1] = ry;
}

template <typename T> __device__ static void wg_smooth(T *sm, const WGLevel<T> &W, int sweeps) {
    const int hm = W.hm, hn = W.hn;
    for (int sw = 0; sw < sweeps; sw++)
        for (int cy = 0; cy < 2; cy++)
            for (int cx = 0; cx < 2 + W.xodd; cx++) {
                const int cnt = cx == 2 ? hn : hm * hn;
                for (int t = threadIdx.x; t < cnt; t += kCoarseThreads) {
                    int i, j;
                    if (cx == 2) { i = W.m - 1; j = 2 * t + cy; }
                    else { const int ty = t / hm; i = 2 * (t - ty * hm) + cx; j = 2 * ty + cy; }
                    if (i >= W.m || j >= W.n || (cx < 2 && W.xodd && i == W.m - 1)) continue;
                    const T *s = sm + W.so + (j * W.rs + i * W.cs) * 40;
                    T ax, ay;
                    wg_offdiag(sm, W, i, j, s, ax, ay);
                    const int k = cidx(W, i, j);
                    const vec2s<T> o = gs_value<T>(s, sm[W.bo + k], sm[W.bo + k + 1], ax, ay);
                    *reinterpret_cast<vec2s<T> *>(sm + W.zo + k) = o;
                }
                __syncthreads();
            }
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) coarse_solve_kernel(CoarseArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    const int tid = threadIdx.x, NT = kCoarseThreads;
    const int ro = A.roff;
    // prologue: stencils of every level, source of level lc, bottom inverse
    // (stored transposed, so that the threads of the bottom solve, one per
    // row, read consecutive addresses)
    for (int l = A.lc; l < A.L; l++) {
        const int cnt = (A.lv[l].cs ? A.lv[l].m * A.lv[l].n : A.lv[l].n) * 40;
        const T *src = A.lv[l].S;
        for (int t = tid; t < cnt; t += NT) sm[A.soff[l] + t] = src[t];
    }
    {
        const WGLevel<T> W = wg_level(A, A.lc);
        const T *gb = A.lv[A.lc].b;
        for (int t = tid; t < W.m * W.n; t += NT) {
            const int j = t / W.m, i = t - j * W.m, k = cidx(W, i, j);
            sm[W.bo + k] = gb[2 * t];
            sm[W.bo + k + 1] = gb[2 * t + 1];
            sm[W.zo + k] = 0;
            sm[W.zo + k + 1] = 0;
        }
        const int N = 2 * A.lv[A.L - 1].m * A.lv[A.L - 1].n;
        for (int t = tid; t < N * N; t += NT) {
            const int r = t / N, c = t - r * N;
            sm[A.binv_off + c * N + r] = A.binv[t];
        }
    }
    __syncthreads();
    // down: smooth, fine residuals into r (one pass), P^T-weighted sums
    for (int l = A.lc; l < A.L - 1; l++) {
        const WGLevel<T> F = wg_level(A, l), C = wg_level(A, l + 1);
        wg_smooth(sm, F, A.nu1);
        for (int t = tid; t < F.m * F.n; t += NT) {
            const int j = t / F.m, i = t - j * F.m;
            wg_resid(sm, F, ro, i, j);
        }
        __syncthreads();
        for (int t = tid; t < C.m * C.n; t += NT) {
            const int J = t / C.m, I = t - J * C.m;
            const int fx0 = 2 * I, fy0 = ((F.n & 1) == 0 && J == C.n - 1) ? F.n - 1 : 2 * J;
            T bx = 0, by = 0;
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                const int fy = fy0 + b;
                const float wy = pw_bounded(fy, F.n, C.n, J);
                const int fyc = fy < 0 ? 0 : (fy >= F.n ? F.n - 1 : fy);
#pragma unroll
                for (int a = -1; a <= 1; a++) {
                    const int fx = fx0 + a;
                    const float wx = pw_periodic(fx, F.m, C.m, I);
                    const int fi = fx < 0 ? fx + F.m : (fx >= F.m ? fx - F.m : fx);
                    const vec2s<T> r = lds2(sm + ro, 2 * (fi + F.m * fyc));
                    const T w = (T)(wx * wy);
                    bx = fmav(w, r.x, bx);
                    by = fmav(w, r.y, by);
                }
            }
            const int k = cidx(C, I, J);
            sm[C.bo + k] = bx;
            sm[C.bo + k + 1] = by;
            sm[C.zo + k] = 0;
            sm[C.zo + k + 1] = 0;
        }
        __syncthreads();
    }
    // bottom: exact solve with the dense inverse (natural unknown order)
    {
        const WGLevel<T> B = wg_level(A, A.L - 1);
        const int N = 2 * B.m * B.n;
        for (int t = tid; t < B.m * B.n; t += NT) {
            const int j = t / B.m, i = t - j * B.m, k = cidx(B, i, j);
            sm[ro + 2 * t] = sm[B.bo + k];
            sm[ro + 2 * t + 1] = sm[B.bo + k + 1];
        }
        __syncthreads();
        for (int r = tid; r < N; r += NT) {
            T acc = 0;
            for (int c = 0; c < N; c++) acc = fmav(sm[A.binv_off + c * N + r], sm[ro + c], acc);
            const int node = r >> 1, j = node / B.m, i = node - j * B.m;
            sm[B.zo + cidx(B, i, j) + (r & 1)] = acc;
        }
        __syncthreads();
    }
    // up: z += P zc, then smooth
    for (int l = A.L - 2; l >= A.lc; l--) {
        const WGLevel<T> F = wg_level(A, l), C = wg_level(A, l + 1);
        for (int t = tid; t < F.m * F.n; t += NT) {
            const int j = t / F.m, i = t - j * F.m;
            const int sc = C.m;
            int cx0, cx1, cy0, cy1;
            T wx0, wx1, wy0, wy1;
            if ((i & 1) == 0) { cx0 = i >> 1; wx0 = 1; cx1 = cx0; wx1 = 0; }
            else { cx0 = i >> 1; cx1 = cx0 + 1 == sc ? 0 : cx0 + 1; wx0 = wx1 = (T)0.5; }
            if ((F.n & 1) == 0 && j == F.n - 1) { cy0 = C.n - 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
            else if ((j & 1) == 0) { cy0 = j >> 1; wy0 = 1; cy1 = cy0; wy1 = 0; }
            else { cy0 = j >> 1; cy1 = cy0 + 1; wy0 = wy1 = (T)0.5; }
            const T *zc = sm + C.zo;
            const vec2s<T> a00 = lds2(zc, cidx(C, cx0, cy0)), a10 = lds2(zc, cidx(C, cx1, cy0)),
                           a01 = lds2(zc, cidx(C, cx0, cy1)), a11 = lds2(zc, cidx(C, cx1, cy1));
            const int k = cidx(F, i, j);
            sm[F.zo + k] += wy0 * (wx0 * a00.x + wx1 * a10.x) + wy1 * (wx0 * a01.x + wx1 * a11.x);
            sm[F.zo + k + 1] += wy0 * (wx0 * a00.y + wx1 * a10.y) + wy1 * (wx0 * a01.y + wx1 * a11.y);
        }
        __syncthreads();
        wg_smooth(sm, F, A.nu2);
    }
    // store the level-lc solution in natural order
    {
        const WGLevel<T> W = wg_level(A, A.lc);
        T *gz = A.lv[A.lc].z;
        for (int t = tid; t < W.m * W.n; t += NT) {
            const int j = t / W.m, i = t - j * W.m, k = cidx(W, i, j);
            gz[2 * t] = sm[W.zo + k];
            gz[2 * t + 1] = sm[W.zo + k + 1];
        }
    }
}

template <typename T> int coarse_layout(CoarseArgs<T> &A) {
    int off = 0;
    for (int l = A.lc; l < A.L; l++) {
        // colour-split storage: four parity arrays of ceil(m/2) x ceil(n/2)
        const int cnt = 8 * ((A.lv[l].m + 1) >> 1) * ((A.lv[l].n + 1) >> 1);
        A.zoff[l] = off;
        off += cnt;
        A.boff[l] = off;
        off += cnt;
        A.soff[l] = off;
        off += (A.lv[l].cs ? A.lv[l].m * A.lv[l].n : A.lv[l].n) * 40;
        off = (off + 3) & ~3;
    }
    // residual scratch: natural order, the largest level (also holds the
    // coarsest source for the dense solve)
    A.roff = off;
    off += 2 * A.lv[A.lc].m * A.lv[A.lc].n;
    off = (off + 3) & ~3;
    const int N = 2 * A.lv[A.L - 1].m * A.lv[A.L - 1].n;
    A.binv_off = off;
    off += N * N;
    A.lds_elems = off;
    return off;
}
template int coarse_layout<double>(CoarseArgs<double> &);
template int coarse_layout<float>(CoarseArgs<float> &);

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

template <typename T> size_t coarse_lds_bytes(const CoarseArgs<T> &A) { return (size_t)A.lds_elems * sizeof(T); }

template <typename T> int mg_coarse(const CoarseArgs<T> &A, hipStream_t st) {
    hipLaunchKernelGGL(coarse_solve_kernel<T>, dim3(1), dim3(kCoarseThreads), coarse_lds_bytes(A), st, A);
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
