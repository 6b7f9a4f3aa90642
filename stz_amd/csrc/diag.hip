// diag.hip -- gfx950 kernels for the frame diagnostics and initial
// conditions of shear_sim: tractions, qs_measure, compute_strain,
// l2_comparison, initialize_random (Gaussian smoothing of host-generated
// noise) and initialize_chi_bicubic.
//
// Reductions are deterministic: one partial per workgroup in a fixed lane
// and wave order, then one workgroup sums the partials in index order.  The
// reference sums rows in OpenMP order, so sums agree to rounding only.
// FMA contraction is off, so per-node values (the cu slot, the bicubic chi)
// keep the reference's rounding wherever no transcendental is involved.
#include "diag.h"

#pragma clang fp contract(off)

namespace stz {

#define PL(P, i, j) (P)[(size_t)(i) + (size_t)m * ((j) + 1)]

__device__ __forceinline__ static double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ static double wmax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// K values per thread summed over a 256-thread block into partial[K*block + k]
template <int K> __device__ __forceinline__ static void block_sums(const double (&v)[K], double *partial, int block) {
    __shared__ double ws[4][K];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double s = wsum(v[k]);
        if (lane == 0) ws[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        partial[K * block + k] = (ws[0][k] + ws[1][k]) + (ws[2][k] + ws[3][k]);
    }
}

// fixed-order sum of np K-component partials (component `maxk` is a max)
template <int K> __global__ void __launch_bounds__(256) reduce_kernel(const double *partial, int np, int maxk,
                                                                      double *out) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = 0;
    for (int p = threadIdx.x; p < np; p += 256)
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = k == maxk ? fmax(v[k], partial[K * p + k]) : v[k] + partial[K * p + k];
    __shared__ double ws[4][K];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double s = k == maxk ? wmax(v[k]) : wsum(v[k]);
        if (lane == 0) ws[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        out[k] = k == maxk ? fmax(fmax(ws[0][k], ws[1][k]), fmax(ws[2][k], ws[3][k]))
                           : (ws[0][k] + ws[1][k]) + (ws[2][k] + ws[3][k]);
    }
}

// ---------------------------------------------------------------------------
// tractions, shear_sim.cc:1003-1028: (bx, by, tx, ty)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) tractions_kernel(SimConst c, FieldPtrs f, double *out) {
    const int m = c.m, n = c.n;
    double v[4] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < m; i += 256) {
        const double syy0 = -PL(f.p, i, 0) - PL(f.s, i, 0) - PL(f.q, i, 0);
        const double syy1 = -PL(f.p, i, 1) - PL(f.s, i, 1) - PL(f.q, i, 1);
        v[0] += 3 * PL(f.tau, i, 0) - PL(f.tau, i, 1);
        v[1] += 3 * syy0 - syy1;
        const double syyb = -PL(f.p, i, n - 2) - PL(f.s, i, n - 2) - PL(f.q, i, n - 2);
        const double syyt = -PL(f.p, i, n - 1) - PL(f.s, i, n - 1) - PL(f.q, i, n - 1);
        v[2] -= 3 * PL(f.tau, i, n - 1) - PL(f.tau, i, n - 2);
        v[3] -= 3 * syyt - syyb;
    }
    __shared__ double ws[4][4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = wsum(v[k]);
        if (lane == 0) ws[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        out[k] = ((ws[0][k] + ws[1][k]) + (ws[2][k] + ws[3][k])) * (0.5 * c.dx);
    }
}

// ---------------------------------------------------------------------------
// qs_measure, shear_sim.cc:733-775: integrand into cu (rows 1..n-2), block
// partials (sum, max)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) qs_kernel(SimConst c, FieldPtrs f, const double *vel, int staggered,
                                                 double *cu, double *partial) {
    const int m = c.m, n = c.n;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int rows = n - 2;
    double v[2] = {0, 0};
    if (idx < m * rows) {
        const int j = 1 + idx / m, i = idx - (j - 1) * m;
        const double qsf1 = 6 * c.K * c.mu * (c.K + (1 / 3.0) * c.mu), qsf2 = 6 * c.K * c.mu * c.mu;
        double ux, uy, vx, vy;
        if (staggered) {
            const int d = i == m - 1 ? 1 - m : 1;
            const double *vp = vel + 2 * ((size_t)i + (size_t)m * j);
            const double p0x = vp[0], p0y = vp[1], pdx = vp[2 * d], pdy = vp[2 * d + 1];
            const double pmx = vp[2 * m], pmy = vp[2 * m + 1], pmdx = vp[2 * (m + d)], pmdy = vp[2 * (m + d) + 1];
            ux = 0.5 * c.xsp * (-p0x + pdx - pmx + pmdx);
            uy = 0.5 * c.ysp * (-p0x - pdx + pmx + pmdx);
            vx = 0.5 * c.xsp * (-p0y + pdy - pmy + pmdy);
            vy = 0.5 * c.ysp * (-p0y - pdy + pmy + pmdy);
        } else {
            const int ir = i == m - 1 ? 0 : i + 1, il = i == 0 ? m - 1 : i - 1;
            ux = 0.5 * c.xsp * (PL(f.u, ir, j) - PL(f.u, il, j));
            vx = 0.5 * c.xsp * (PL(f.v, ir, j) - PL(f.v, il, j));
            uy = 0.5 * c.ysp * (PL(f.u, i, j + 1) - PL(f.u, i, j - 1));
            vy = 0.5 * c.ysp * (PL(f.v, i, j + 1) - PL(f.v, i, j - 1));
        }
        const double l1 = ux + vy, l2 = ux - vy, l3 = uy + vx;
        const double q = qsf1 * l1 * l1 + qsf2 * (l2 * l2 + l3 * l3);
        PL(cu, i, j) = q;
        v[0] = q;
        v[1] = q;
    }
    // sum and max over the block
    __shared__ double ws[4][2];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const double s0 = wsum(v[0]), s1 = wmax(v[1]);
    if (lane == 0) {
        ws[wid][0] = s0;
        ws[wid][1] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = (ws[0][0] + ws[1][0]) + (ws[2][0] + ws[3][0]);
        partial[2 * blockIdx.x + 1] = fmax(fmax(ws[0][1], ws[1][1]), fmax(ws[2][1], ws[3][1]));
    }
}

// ---------------------------------------------------------------------------
// compute_strain, shear_sim.cc:778-824 (mat::sym_eigenvectors,
// shared/mat.cc:8-26; mat::devmod, mat.hh:51-53): the deviatoric modulus of
// the Hencky strain into cu (rows 0..n-1).  corrected = 0 reproduces the
// reference, including its interior-row y difference that takes both ends
// from row j+1 (shear_sim.cc:804, so 0 there) and its seam column i = m-1
// that subtracts the periodic shift (:793); corrected = 1 uses rows j+1 and
// j-1 and adds the shift.
// ---------------------------------------------------------------------------
__device__ __forceinline__ static void mmul(const double *A, const double *B, double *o) {
    o[0] = A[0] * B[0] + A[1] * B[2];
    o[1] = A[0] * B[1] + A[1] * B[3];
    o[2] = A[2] * B[0] + A[3] * B[2];
    o[3] = A[2] * B[1] + A[3] * B[3];
}

__global__ void __launch_bounds__(256) strain_kernel(SimConst c, FieldPtrs f, int corrected, double *cu) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    double Xx, Yx, Xy, Yy;
    const double lxs = c.bx - c.ax;
    if (i == 0) {
        Xx = PL(f.X, 1, j) - PL(f.X, m - 1, j) + lxs;
        Yx = PL(f.Y, 1, j) - PL(f.Y, m - 1, j);
    } else if (i == m - 1) {
        Xx = PL(f.X, 0, j) - PL(f.X, m - 2, j) + (corrected ? lxs : -lxs);
        Yx = PL(f.Y, 0, j) - PL(f.Y, m - 2, j);
    } else {
        Xx = PL(f.X, i + 1, j) - PL(f.X, i - 1, j);
        Yx = PL(f.Y, i + 1, j) - PL(f.Y, i - 1, j);
    }
    Xx *= 0.5 * c.xsp;
    Yx *= 0.5 * c.xsp;
    int ju, jd;
    double yfac;
    if (j == 0 || j == n - 1) {
        yfac = c.ysp;
        ju = j + 1;
        jd = j;
    } else {
        yfac = 0.5 * c.ysp;
        ju = j + 1;
        jd = corrected ? j - 1 : j + 1;
    }
    Xy = yfac * (PL(f.X, i, ju) - PL(f.X, i, jd));
    Yy = yfac * (PL(f.Y, i, ju) - PL(f.Y, i, jd));
    const double J = 1 / (Xx * Yy - Xy * Yx);
    const double F[4] = {Yy * J, -Xy * J, -Yx * J, Xx * J}, FT[4] = {F[0], F[2], F[1], F[3]};
    double FTF[4];
    mmul(FT, F, FTF);
    // sym_eigenvectors of FTF (l1 >= l2)
    double l1, l2, L[4];
    const double a = FTF[0], b = FTF[1], d = FTF[3];
    if (fabs(b) < 1e-12 * (fabs(a) + fabs(d))) {
        if (a > d) {
            l1 = a;
            l2 = d;
            L[0] = 1; L[1] = 0; L[2] = 0; L[3] = 1;
        } else {
            l1 = d;
            l2 = a;
            L[0] = 0; L[1] = 1; L[2] = 1; L[3] = 0;
        }
    } else {
        const double det = sqrt((a - d) * (a - d) + 4 * b * b);
        l1 = 0.5 * (a + d + det);
        l2 = 0.5 * (a + d - det);
        const double r = b, s = a - l1;
        const double nor1 = 1 / sqrt(r * r + s * s);
        L[0] = r * nor1; L[1] = -s * nor1; L[2] = s * nor1; L[3] = r * nor1;
    }
    double out = 0;
    if (l1 > 0 || l2 > 0) {
        const double LT[4] = {L[0], L[2], L[1], L[3]}, D[4] = {0.5 * log(l1), 0, 0, 0.5 * log(l2)};
        double t[4], sg[4];
        mmul(LT, D, t);
        mmul(t, L, sg);
        out = sqrt(0.5 * (sg[0] - sg[3]) * (sg[0] - sg[3]) + sg[1] * sg[1] + sg[2] * sg[2]);
    }
    PL(cu, i, j) = out;
}

// ---------------------------------------------------------------------------
// l2_comparison, shear_sim.cc:460-502: block partials of the four sums over
// rows 0..n (row 0 and the extra node row n weighted 1/2 for node fields)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) l2_kernel(SimConst c, FieldPtrs a, FieldPtrs b, double *partial) {
    const int m = c.m, n = c.n;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (idx < m * (n + 1)) {
        const int j = idx / m, i = idx - j * m;
        const double du = PL(a.u, i, j) - PL(b.u, i, j), dv = PL(a.v, i, j) - PL(b.v, i, j);
        const double dX = PL(a.X, i, j) - PL(b.X, i, j), dY = PL(a.Y, i, j) - PL(b.Y, i, j);
        const double w = (j == 0 || j == n) ? 0.5 : 1.0;
        v[0] = w * (du * du + dv * dv);
        v[3] = w * (dX * dX + dY * dY);
        if (j < n) {
            const double dp = PL(a.p, i, j) - PL(b.p, i, j), dq = PL(a.q, i, j) - PL(b.q, i, j);
            const double ds = PL(a.s, i, j) - PL(b.s, i, j), dt = PL(a.tau, i, j) - PL(b.tau, i, j);
            const double dc = PL(a.chi, i, j) - PL(b.chi, i, j);
            v[1] = 2 * (dt * dt + ds * ds) + 3 * dp * dp + 6 * dq * dq;
            v[2] = dc * dc;
        }
    }
    block_sums<4>(v, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// initialize_random, shear_sim.cc:915-998: chi = chi0 + chi1 * (Gaussian
// smoothing of the noise).  The noise (m x (n + 2 cut) values, rows -cut..
// n+cut-1) is drawn on the host exactly as the reference draws it; the
// convolution with exp(-llinv (di^2 + dj^2)) is done separably, first along
// x (periodic) into tmp, then along y.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) smooth_x_kernel(int m, int rows, int cut, const double *w, const double *noise,
                                                       double *tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || r >= rows) return;
    const double *row = noise + (size_t)r * m;
    double s = 0;
    for (int d = -cut; d <= cut; d++) {
        int ci = i + d;
        ci = ci >= 0 ? ci % m : m - 1 - (m - 1 - ci) % m;
        s += row[ci] * w[d < 0 ? -d : d];
    }
    tmp[(size_t)r * m + i] = s;
}

__global__ void __launch_bounds__(256) random_fields_kernel(SimConst c, FieldPtrs f, int cut, const double *w,
                                                            const double *tmp, double chi0, double chi1) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j > n) return;
    PL(f.u, i, j) = 0;
    PL(f.v, i, j) = 0;
    PL(f.X, i, j) = c.ax + i * c.dx;
    PL(f.Y, i, j) = c.ay + j * c.dy;
    if (j < n) {
        double g = 0;
        for (int d = -cut; d <= cut; d++) g += tmp[(size_t)(j + cut + d) * m + i] * w[d < 0 ? -d : d];
        PL(f.chi, i, j) = chi0 + g * chi1;
        PL(f.p, i, j) = 0;
        PL(f.q, i, j) = 0;
        PL(f.s, i, j) = 0;
        PL(f.tau, i, j) = 0;
    }
}

// ---------------------------------------------------------------------------
// initialize_chi_bicubic, shear_sim.cc:832-913: chi on rows 0..n-1 from the
// mm x nn array and its x, y, xy differences (d = arr, xd, yd, xyd, each
// mm*nn); then the other fields on rows -1..n
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) bicubic_kernel(SimConst c, FieldPtrs f, int mm, int nn, const double *d,
                                                      double lo, double sca) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= m || j >= n) return;
    const int mmnn = mm * nn;
    const double *arr = d, *xd = d + mmnn, *yd = d + 2 * mmnn, *xyd = d + 3 * mmnn;
    const double xsca = c.dx * mm / (c.bx - c.ax), ysca = c.dy * nn / (c.by - c.ay);
    double Y = (j + 0.5) * ysca - 0.5, mY;
    int jj = int(Y);
    if (jj < 0) jj = 0;
    else if (jj > nn - 2) jj = nn - 2;
    Y -= jj;
    mY = 1 - Y;
    const double b0 = Y * mY * mY, b1 = mY * mY * (1 + 2 * Y), b2 = Y * Y * (3 - 2 * Y), b3 = -Y * Y * mY;
    double X = (i + 0.5) * xsca - 0.5, mX;
    if (X < 0) X += mm;
    int ii = int(X);
    if (ii >= mm) ii -= mm;
    const int ij = ii + mm * jj, ijp = (ii == mm - 1 ? ii - mm : ii) + 1 + mm * jj;
    X -= ii;
    mX = 1 - X;
    const double c0 = X * mX * mX, c1 = mX * mX * (1 + 2 * X), c2 = X * X * (3 - 2 * X), c3 = -X * X * mX;
    PL(f.chi, i, j) = lo + sca * (b0 * (xyd[ij] * c0 + yd[ij] * c1 + yd[ijp] * c2 + xyd[ijp] * c3) +
                                  b1 * (xd[ij] * c0 + arr[ij] * c1 + arr[ijp] * c2 + xd[ijp] * c3) +
                                  b2 * (xd[ij + mm] * c0 + arr[ij + mm] * c1 + arr[ijp + mm] * c2 + xd[ijp + mm] * c3) +
                                  b3 * (xyd[ij + mm] * c0 + yd[ij + mm] * c1 + yd[ijp + mm] * c2 + xyd[ijp + mm] * c3));
}

__global__ void __launch_bounds__(256) rest_fields_kernel(SimConst c, FieldPtrs f) {
    const int m = c.m, n = c.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y - 1;
    if (i >= m || j > n) return;
    PL(f.u, i, j) = 0;
    PL(f.v, i, j) = 0;
    PL(f.ddev, i, j) = 0;
    PL(f.p, i, j) = 0;
    PL(f.q, i, j) = 0;
    PL(f.s, i, j) = 0;
    PL(f.tau, i, j) = 0;
    PL(f.X, i, j) = c.ax + i * c.dx;
    PL(f.Y, i, j) = c.ay + j * c.dy;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 g2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

int diag_tractions(const SimConst &c, const FieldPtrs &f, double *out, hipStream_t st) {
    hipLaunchKernelGGL(tractions_kernel, dim3(1), dim3(256), 0, st, c, f, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int diag_partials_needed(const SimConst &c) { return (c.m * (c.n + 1) + 255) / 256; }

int diag_qs(const SimConst &c, const FieldPtrs &f, const double *vel, int staggered, double *cu, double *partial,
            double *out, hipStream_t st) {
    const int nb = (c.m * (c.n - 2) + 255) / 256;
    if (nb <= 0) return STZ_OK;
    hipLaunchKernelGGL(qs_kernel, dim3(nb), dim3(256), 0, st, c, f, vel, staggered, cu, partial);
    hipLaunchKernelGGL(reduce_kernel<2>, dim3(1), dim3(256), 0, st, partial, nb, 1, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int diag_strain(const SimConst &c, const FieldPtrs &f, int corrected, double *cu, hipStream_t st) {
    const dim3 b(64, 4);
    hipLaunchKernelGGL(strain_kernel, g2(c.m, c.n, b), b, 0, st, c, f, corrected, cu);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int diag_l2(const SimConst &c, const FieldPtrs &a, const FieldPtrs &b, double *partial, double *out, hipStream_t st) {
    const int nb = diag_partials_needed(c);
    hipLaunchKernelGGL(l2_kernel, dim3(nb), dim3(256), 0, st, c, a, b, partial);
    hipLaunchKernelGGL(reduce_kernel<4>, dim3(1), dim3(256), 0, st, partial, nb, -1, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int diag_random_fields(const SimConst &c, const FieldPtrs &f, int cut, const double *w, const double *noise,
                       double *tmp, double chi0, double chi1, hipStream_t st) {
    const dim3 b(64, 4);
    const int rows = c.n + 2 * cut;
    hipLaunchKernelGGL(smooth_x_kernel, g2(c.m, rows, b), b, 0, st, c.m, rows, cut, w, noise, tmp);
    hipLaunchKernelGGL(random_fields_kernel, g2(c.m, c.n + 1, b), b, 0, st, c, f, cut, w, tmp, chi0, chi1);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

int diag_bicubic(const SimConst &c, const FieldPtrs &f, int mm, int nn, const double *d, double lo, double sca,
                 hipStream_t st) {
    const dim3 b(64, 4);
    hipLaunchKernelGGL(bicubic_kernel, g2(c.m, c.n, b), b, 0, st, c, f, mm, nn, d, lo, sca);
    hipLaunchKernelGGL(rest_fields_kernel, g2(c.m, c.n + 2, b), b, 0, st, c, f);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}

}  // namespace stz
