// noalloc.hip -- the mean-square residual of the quasi-static projection
// (tgmg_base::l2_error, tgmg.hh:146, on the operator of qs_multi.cc:31-54)
// for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// stz_mg_l2_error computes the same number on a simulation's own vel/src with
// a partials buffer the simulation allocated.  This entry point takes the
// velocity, the source and the output as device pointers and keeps its
// partial sums (and the three distinct row stencils of the operator) in one
// file-scope device array, `scratch`, of fixed size: each GPU holds its own
// copy from the first call until the process ends.  The kernels are the
// residual and the fixed-order sum of stz_mg_l2_error (same per-node
// arithmetic, same 256-thread blocks, same reduction order), so the two
// results are bit-identical.
//
// All calls on one GPU share that GPU's array: calls that overlap on one GPU
// (on different streams, or from different host threads) race.  Nothing has
// to start at zero: every launch writes each scratch word it later reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device.h"
#include "mg_node.h"
#include "sim_internal.h"

namespace stz {

namespace {

constexpr int kNaThreads = 256;
constexpr int kNaMaxBlocks = 1 << 16;                       // partial sums
constexpr long long kNaMaxNodes = (long long)kNaThreads * kNaMaxBlocks;  // 16,777,216 nodes
constexpr int kNaStencils = 3 * 40;                         // bottom wall, interior, top wall
constexpr int kNaScratch = kNaStencils + kNaMaxBlocks;      // doubles

struct Stencils3 {
    double s[kNaStencils];
};

}  // namespace

static __device__ __align__(256) double scratch[kNaScratch];

// the device address of `scratch`, per device, looked up once
static constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(double **out) {
    int dev = 0;
    STZ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d devices this entry point tracks", dev, kMaxDevices);
        return STZ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        STZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("no device address for the scratch array");
            return STZ_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return STZ_OK;
}

// the three row stencils into scratch[0 .. 120)
__global__ void __launch_bounds__(128) na_stencils_kernel(Stencils3 st, double *scr) {
    if (threadIdx.x < kNaStencils) scr[threadIdx.x] = st.s[threadIdx.x];
}

// per-block partial sums of |r|^2 over the nodes of a level whose rows all
// share the interior stencil except the two wall rows
__global__ void __launch_bounds__(kNaThreads) na_resid_kernel(int m, int n, const double *z, const double *b,
                                                              double *scr) {
    const int idx = blockIdx.x * kNaThreads + threadIdx.x;
    double v = 0;
    if (idx < m * n) {
        const int j = idx / m, i = idx - j * m;
        LevelArgs<double> L{};
        L.m = m;
        L.n = n;
        L.z = const_cast<double *>(z);
        const double *s = scr + 40 * (j == 0 ? 0 : (j == n - 1 ? 2 : 1));
        vec2s<double> dl, dc, dr, cl, cr, ul, uc, ur;
        neighbours(L, i, j, dl, dc, dr, cl, cr, ul, uc, ur);
        double ax, ay, rx, ry;
        offdiag9<double>(s, dl, dc, dr, cl, cr, ul, uc, ur, ax, ay);
        const int k = 2 * (i + m * j);
        const vec2s<double> z0 = {z[k], z[k + 1]};
        resid_value<double>(s, b[k], b[k + 1], ax, ay, z0, rx, ry);
        v = rx * rx + ry * ry;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) scr[kNaStencils + blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// fixed-order sum of the partials, divided by the node count
__global__ void __launch_bounds__(kNaThreads) na_sum_kernel(const double *scr, int np, double count, double *out) {
    double v = 0;
    for (int k = threadIdx.x; k < np; k += kNaThreads) v += scr[kNaStencils + k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = ((ws[0] + ws[1]) + (ws[2] + ws[3])) / count;
}

}  // namespace stz

using namespace stz;

extern "C" int stz_qs_l2_error_noalloc(const stz_sim_params *p, double viscdt, const double *vel, const double *src,
                                       double *out, void *hip_stream) {
    if (!p || !vel || !src || !out) {
        set_error("null argument");
        return STZ_ERR_ARG;
    }
    if (p->m < 4 || p->n < 4) {
        set_error("need m >= 4 and n >= 4 (got %d x %d)", p->m, p->n);
        return STZ_ERR_ARG;
    }
    const long long nodes = (long long)p->m * (p->n + 1);
    if (nodes > kNaMaxNodes) {
        set_error("%lld nodes exceed the %lld this entry point's fixed scratch covers", nodes, kNaMaxNodes);
        return STZ_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return STZ_ERR_NO_DEVICE;
    }
    double *scr = nullptr;
    int e;
    if ((e = scratch_ptr(&scr))) return e;
    SimConst c;
    fill_sim_const(p, c, nullptr);
    Stencils3 st;
    top_row_stencil(c, viscdt, true, st.s);
    top_row_stencil(c, viscdt, false, st.s + 40);
    top_row_stencil(c, viscdt, true, st.s + 80);
    const hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    const int m = p->m, n = p->n + 1, nb = (int)((nodes + kNaThreads - 1) / kNaThreads);
    hipLaunchKernelGGL(na_stencils_kernel, dim3(1), dim3(128), 0, stream, st, scr);
    hipLaunchKernelGGL(na_resid_kernel, dim3(nb), dim3(kNaThreads), 0, stream, m, n, vel, src, scr);
    hipLaunchKernelGGL(na_sum_kernel, dim3(1), dim3(kNaThreads), 0, stream, scr, nb, (double)nodes, out);
    STZ_HIP(hipGetLastError());
    return STZ_OK;
}
