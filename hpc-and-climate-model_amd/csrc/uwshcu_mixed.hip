// uwshcu_mixed.hip -- mixed-precision forward: FP32 plume arithmetic and plume
// workspace; FP64 saturation adjustment (conden), LCL inversion (qsinvert), CIN and
// cloud-base closure, environment, wet bulb and positive-moisture fix-up.  After the
// FP32 pass, check_columns_kernel lists the columns that fail the FP64 conservation
// check or took an FP32 branch decision within rounding; the driver recomputes
// exactly those in FP64 (forward_kernel<double> over the list).
#include "uwshcu_check.h"
#include "uwshcu_forward.h"

namespace uwshcu {

hipError_t launch_forward_mixed(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                hipStream_t stream) {
    return launch_column_pipeline<float>(args, n, dconsts, destbl, stream);
}

// One thread per column of [0, ncol): append failing columns to list.
__global__ __launch_bounds__(256) void check_columns_kernel(FwdArgs args, const uwshcu_consts* __restrict__ consts,
                                                            double tol, int* __restrict__ list,
                                                            int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.ncol) return;
    const bool bad = (args.amb_flag && args.amb_flag[i]) || column_fails_check(args, *consts, i, tol);
    if (bad) list[atomicAdd(count, 1)] = i;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

hipError_t launch_check_columns(const FwdArgs& args, const uwshcu_consts* dconsts, double tol, int* list, int* count,
                                hipStream_t stream) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, stream, count, 1, 0);
    const int block = 256;
    const int grid = (args.ncol + block - 1) / block;
    hipLaunchKernelGGL(check_columns_kernel, dim3(grid), dim3(block), 0, stream, args, dconsts, tol, list, count);
    return hipGetLastError();
}

}  // namespace uwshcu
