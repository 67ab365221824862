// uwshcu_sweep.hip -- the adjoint's backward sweep: each column's tape (recorded by
// uwshcu_adjoint.hip) is swept backwards from the outputs' cotangents to the gradients
// of t0, qv0, u0, v0, tke, pblh and (rpen, rle, rkm).  A translation unit of its own,
// apart from the recording kernel's large one.
// Compiled with -ffp-contract=off: products and sums round as the host build's do.
#include "uwshcu_adjoint.h"
#include "uwshcu_kernels.h"

namespace uwshcu {

// Backward sweep, one column per thread, with a 64-entry window of recent adjoints
// (and its spare entry) per thread in LDS: 520 B per column, 33 KB per workgroup of 64,
// so that the four workgroups a CU needs to give each SIMD one wave fit its LDS.  The
// columns per workgroup (UWSHCU_SWEEP_LANES, 64 or 32) are a build-time parameter.
// Columns whose recording did not complete are skipped.
constexpr int kSweepWindow = 64;
#ifndef UWSHCU_SWEEP_LANES
#define UWSHCU_SWEEP_LANES 64
#endif
constexpr int kSweepLanes = UWSHCU_SWEEP_LANES;
static_assert(kSweepLanes == 64 || kSweepLanes == 32, "a sweep workgroup is one wavefront");
__global__ __launch_bounds__(kSweepLanes) void adjoint_sweep_kernel(int n, int col0, int mkx, int ncnst, long long wld,
                                                                   long long ldc, const TapeCtx* tapes,
                                                                   const int* node_out, double* adj, AdCotangents cot,
                                                                   AdGradients g) {
    __shared__ double s_win[kSweepWindow + 1][kSweepLanes];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const TapeCtx& t = tapes[j];
    if (t.overflow) return;
    const int lane = threadIdx.x;
    auto win = [&](int s) -> double& { return s_win[s][lane]; };
    ad_reverse(t, adj + t.adj, node_out, (size_t)wld, j, mkx, ncnst, (size_t)ldc, col0 + j, cot, g, win,
               kSweepWindow);
}

// Zero n doubles (the sweep's adjoints; a kernel, not a memset, so that a captured
// graph records no memset node).
__global__ __launch_bounds__(256) void zero_kernel(double* p, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0.0;
}

hipError_t launch_adjoint_sweep(const FwdArgs& args, int n, double* adj, long long nadj, const int* node_out,
                                const AdCotangents& cot, const AdGradients& g, hipStream_t stream) {
    if (nadj > 0) {
        long long blocks = (nadj + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, adj, nadj);
    }
    hipLaunchKernelGGL(adjoint_sweep_kernel, dim3((n + kSweepLanes - 1) / kSweepLanes), dim3(kSweepLanes), 0, stream,
                       n, args.col0, args.mkx,
                       args.ncnst, args.wld, (long long)args.ldc, (const TapeCtx*)args.ad.tapes, node_out, adj, cot, g);
    return hipGetLastError();
}

}  // namespace uwshcu
