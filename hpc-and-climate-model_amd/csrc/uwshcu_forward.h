// uwshcu_forward.h -- the forward kernel template, instantiated per precision.
#ifndef UWSHCU_FORWARD_H
#define UWSHCU_FORWARD_H

#include "uwshcu_kernels.h"

namespace uwshcu {

#ifndef UWSHCU_FWD_WAVES_PER_EU
#define UWSHCU_FWD_WAVES_PER_EU 1
#endif

// One thread per column; the saturation table is staged into LDS once per workgroup.
// Thread j of the launch works in workspace slot j on column col0 + j, or, with a
// column map, on column colmap[col0 + j] while col0 + j < *colcount.
template <class R>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(UWSHCU_FWD_WAVES_PER_EU)))
void forward_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
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
    const Thermo T{consts, s_estbl};
    const Column<R> c(args, T, col, j);
    c.run();
}

}  // namespace uwshcu
#endif
