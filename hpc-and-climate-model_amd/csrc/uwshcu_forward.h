// uwshcu_forward.h -- the forward kernel template, instantiated per precision.
#ifndef UWSHCU_FORWARD_H
#define UWSHCU_FORWARD_H

#include "uwshcu_kernels.h"

namespace uwshcu {

#ifndef UWSHCU_FWD_WAVES_PER_EU
#define UWSHCU_FWD_WAVES_PER_EU 1
#endif

// One thread per column; the saturation table is staged into LDS once per workgroup.
template <class R>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(UWSHCU_FWD_WAVES_PER_EU)))
void forward_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                                                      const double* __restrict__ estbl) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const Thermo T{consts, s_estbl};
    const Column<R> col(args, T, args.col0 + j);
    col.run();
}

}  // namespace uwshcu
#endif
