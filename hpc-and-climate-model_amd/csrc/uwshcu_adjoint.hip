// uwshcu_adjoint.hip -- adjoint kernels: record the FP64 forward pass of each column
// on a tape (Column<AD, AD>) and sum the parameter gradients; the backward sweep is in
// uwshcu_sweep.hip.  One thread per column; tapes live in HBM, paged from a pool.
// Compiled with -ffp-contract=off: the recorded values equal the FP64 forward path's.
#include "uwshcu_adjoint.h"
#include "uwshcu_kernels.h"

namespace uwshcu {

// Recording pass.  Slot j's tape takes its nodes page by page from the pool (its page
// table at pages + j * maxp), so no sizing pass is needed; a column that finds the
// pool empty goes on counting and is counted in pool->retry (the host records it again
// in a launch the pool can hold).  At its end a column takes its n adjoint entries
// from the pool's adjoint array.  The column code is expanded in place: the compiler
// then sees every array pointer come from the kernel arguments and issues global, not
// flat, accesses.
// UWSHCU_AD_LANES columns per 64-lane workgroup (fewer columns per wavefront give more
// wavefronts for a launch of few columns) and UWSHCU_AD_OCC wavefronts per SIMD.
#ifndef UWSHCU_AD_LANES
#define UWSHCU_AD_LANES 64
#endif
#ifndef UWSHCU_AD_OCC
#define UWSHCU_AD_OCC 1
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(UWSHCU_AD_OCC)))
void adjoint_record_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                           const double* __restrict__ estbl, TapePool* pool, int* pages, int maxp, int cap,
                           int* node_out) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    if ((int)threadIdx.x >= UWSHCU_AD_LANES) return;
    const int j = blockIdx.x * UWSHCU_AD_LANES + threadIdx.x;
    if (j >= n) return;
    // the tape header lives in LDS: every recorded operation updates it
    TapeCtx& tape = *ad_tape();
    TapeHot& hot = *ad_hot();
    int* const pt = pages + (size_t)j * maxp;
    const int nsh = pool->nshards;
    ad_begin(tape, hot, nullptr, pt, pool, nsh > 1 ? (int)(blockIdx.x % (unsigned)nsh) : 0, cap,
             ad_inputs(args.mkx));
    const ThermoT<AD> T{consts, s_estbl};
    const Column<AD, AD> col(args, T, args.col0 + j, j, &tape);
    col.template run_body<false>();
    col.ad_export(node_out);
    tape.hot = hot;
    if (tape.overflow == 0) tape.adj = (long long)atomicAdd(&pool->adj_next, (unsigned long long)tape.hot.n);
    else if (tape.overflow == 2) atomicAdd(&pool->retry, 1);
    else record_error(args.err, UWSHCU_ERR_TAPE, args.col0 + j);
    tape.pages = pt;
    args.ad.tapes[j] = tape;     // the sweep reads the header from HBM
}

// Sum of the per-column parameter gradients par[3][ldc] over ncol columns, in FP64:
// one workgroup per parameter, a strided partial sum per thread, then a tree in LDS.
__global__ __launch_bounds__(256) void param_sum_kernel(const double* par, long long ldc, int ncol, double* sum) {
    __shared__ double s[256];
    const int p = blockIdx.x;
    double acc = 0.0;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) acc += par[(size_t)p * ldc + c];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sum[p] = s[0];
}

hipError_t launch_adjoint_record(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                 TapePool* pool, int* pages, int maxp, int cap, int* node_out, hipStream_t stream) {
    hipLaunchKernelGGL(adjoint_record_kernel, dim3((n + UWSHCU_AD_LANES - 1) / UWSHCU_AD_LANES), dim3(64), 0, stream,
                       args, n, dconsts, destbl, pool, pages, maxp, cap, node_out);
    return hipGetLastError();
}

hipError_t launch_param_sum(const double* par, long long ldc, int ncol, double* sum, hipStream_t stream) {
    hipLaunchKernelGGL(param_sum_kernel, dim3(3), dim3(256), 0, stream, par, ldc, ncol, sum);
    return hipGetLastError();
}

}  // namespace uwshcu
