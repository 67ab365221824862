// uwshcu_fp64.hip -- FP64 forward kernel: one thread per column.
//
// Compiled with -ffp-contract=off so that every product and sum rounds as the
// reference expression is written (no fused multiply-add), which keeps the FP64
// path within round-off of the serial reference.
#include "uwshcu_forward.h"

namespace uwshcu {

hipError_t launch_forward_fp64(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                               hipStream_t stream) {
    return launch_column_pipeline<double>(args, n, dconsts, destbl, stream);
}

__global__ void clear_status_kernel(int* err) {
    if (threadIdx.x < 2) err[threadIdx.x] = 0;
}

hipError_t launch_clear_status(int* err, hipStream_t stream) {
    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(64), 0, stream, err);
    return hipGetLastError();
}

}  // namespace uwshcu
