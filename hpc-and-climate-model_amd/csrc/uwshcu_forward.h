// uwshcu_forward.h -- the forward kernel template, instantiated per precision.
#ifndef UWSHCU_FORWARD_H
#define UWSHCU_FORWARD_H

#include "uwshcu_kernels.h"

namespace uwshcu {

// One thread per column; the saturation table is staged into LDS once per workgroup.
// Thread j of the launch works in workspace slot j on column col0 + j, or, with a
// column map, on column colmap[col0 + j] while col0 + j < *colcount.
template <class R, bool ZW = false>
__device__ __forceinline__ void forward_body(const FwdArgs& args, int n, const uwshcu_consts* __restrict__ consts,
                                             const double* __restrict__ estbl) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    int nvalid = n;
    if (args.colmap) {
        const int left = *args.colcount - args.col0;
        nvalid = left < n ? left : n;
    }
    if ((int)(blockIdx.x * blockDim.x) >= nvalid) return;     // whole workgroup idle
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nvalid) return;
    const int col = args.colmap ? args.colmap[args.col0 + j] : args.col0 + j;
    const int slot = args.resume ? col - args.slot0 : j;
    const Thermo T{consts, s_estbl};
    const Column<R, double, ZW> c(args, T, col, slot);
    c.run_body();
}

// Two register budgets for the same column code.  forward_kernel lets it use all 512
// registers of a lane (one wavefront per SIMD): right for launches with fewer
// wavefronts than SIMDs, where nothing could run beside it.  forward_kernel_occ2
// allows half that (some spilling) so that two wavefronts share a SIMD and hide each
// other's FP64 latency: 1.37x faster at 1 M columns.
template <class R, bool ZW = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1)))
void forward_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts, const double* __restrict__ estbl) {
    forward_body<R, ZW>(args, n, consts, estbl);
}
template <class R, bool ZW = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
void forward_kernel_occ2(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                         const double* __restrict__ estbl) {
    forward_body<R, ZW>(args, n, consts, estbl);
}

// Level-parallel set-up: thread t of the launch handles workspace slot t % n and
// level t / n (+1 for stage 2), so that a wavefront covers consecutive columns.
template <class R, int STAGE, bool ZW = false>
__global__ __launch_bounds__(256) void setup_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                                                    const double* __restrict__ estbl) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    int nvalid = n;
    if (args.colmap) {
        const int left = *args.colcount - args.col0;
        nvalid = left < n ? left : n;
    }
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = (int)(t % n);
    const int k = (int)(t / n) + (STAGE == 2 ? 1 : 0);
    if (k > args.mkx || j >= nvalid) return;
    const int col = args.colmap ? args.colmap[args.col0 + j] : args.col0 + j;
    const Thermo T{consts, s_estbl};
    const Column<R, double, ZW> c(args, T, col, j);
    if (STAGE == 1) c.setup_level1(k);
    else c.setup_level2(k);
}

static __global__ void clear_count_kernel(int* count) {
    if (threadIdx.x == 0) *count = 0;
}

// The forward pass of columns [col0, col0 + n) (or of the mapped columns): the two
// level-parallel set-up stages, then the column kernel.
//
// Large launches (at least two wavefronts per SIMD) also split off the columns whose
// implicit-CIN pass convects a second time (a few per cent of them): the column
// kernel stops each of those where its second convection pass begins, saves its
// scalars and lists it (defer_list, defer_state), and a second column kernel
// continues the listed columns from there in their own workspace slots.  Otherwise
// nearly every wavefront would carry one or two such columns through a second
// convection pass while its other lanes idle.
//
// args0.defer_list / defer_count / defer_state (nullable) give the list's storage
// (n entries, NRESUME x wld doubles).
template <class R, bool ZW = false>
hipError_t launch_column_pipeline(const FwdArgs& args0, int n, const uwshcu_consts* dconsts, const double* destbl,
                                  hipStream_t stream) {
    int* const defer_list = args0.defer_list;
    int* const defer_count = args0.defer_count;
    FwdArgs args = args0;
    args.env_ready = 1;
    args.resume = 0;
    args.defer_list = nullptr;
    args.defer_count = nullptr;
    const long long t1 = (long long)n * (args.mkx + 1), t2 = (long long)n * args.mkx;
    hipLaunchKernelGGL((setup_kernel<R, 1, ZW>), dim3((unsigned)((t1 + 255) / 256)), dim3(256), 0, stream, args, n,
                       dconsts, destbl);
    hipLaunchKernelGGL((setup_kernel<R, 2, ZW>), dim3((unsigned)((t2 + 255) / 256)), dim3(256), 0, stream, args, n,
                       dconsts, destbl);
    const bool large = n >= 2 * 1024 * 64;
    const bool split = large && defer_list && defer_count && args0.defer_state && !args.colmap;
    if (split) {
        hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(64), 0, stream, defer_count);
        args.defer_list = defer_list;
        args.defer_count = defer_count;
    }
    // at least two wavefronts per SIMD of the 256-CU part: the occupancy-2 build
    if (large)
        hipLaunchKernelGGL((forward_kernel_occ2<R, ZW>), dim3((n + 127) / 128), dim3(128), 0, stream, args, n, dconsts, destbl);
    else
        hipLaunchKernelGGL((forward_kernel<R, ZW>), dim3((n + 127) / 128), dim3(128), 0, stream, args, n, dconsts, destbl);
    if (split) {
        // continue the listed columns at their second convection pass
        FwdArgs b = args0;
        b.env_ready = 1;
        b.resume = 1;
        b.slot0 = args0.col0;
        b.defer_list = nullptr;
        b.defer_count = nullptr;
        b.colmap = defer_list;
        b.colcount = defer_count;
        b.col0 = 0;
        hipLaunchKernelGGL((forward_kernel<R, ZW>), dim3((n + 127) / 128), dim3(128), 0, stream, b, n, dconsts, destbl);
    }
    return hipGetLastError();
}

}  // namespace uwshcu
#endif
