// uwshcu_mixed.hip -- mixed-precision forward kernel: FP32 bulk arithmetic and
// workspace, FP64 saturation adjustment (conden), LCL inversion (qsinvert), CIN
// integrals, implicit-CIN weight, wet bulb and positive-moisture fix-up.
#include "uwshcu_forward.h"

namespace uwshcu {

hipError_t launch_forward_mixed(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                hipStream_t stream) {
    const int block = 128;
    const int grid = (n + block - 1) / block;
    hipLaunchKernelGGL(forward_kernel<float>, dim3(grid), dim3(block), 0, stream, args, n, dconsts, destbl);
    return hipGetLastError();
}

}  // namespace uwshcu
