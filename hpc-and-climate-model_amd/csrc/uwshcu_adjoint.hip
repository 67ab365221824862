// uwshcu_adjoint.hip -- adjoint kernels: record the FP64 forward pass of each column
// on a tape (Column<AD, AD>), then sweep the tape backwards from the outputs'
// cotangents to the gradients of t0, qv0, u0, v0, tke, pblh and (rpen, rle, rkm).
// One thread per column; tapes live in HBM, one contiguous region per column slot.
// Compiled with -ffp-contract=off: the recorded values equal the FP64 forward path's.
#include "uwshcu_adjoint.h"
#include "uwshcu_kernels.h"

namespace uwshcu {

// Recording pass: slot j's tape holds cap[j] nodes at tape_base + off[j].  With
// counts != nullptr it is the sizing pass instead: the recording stores nothing and
// counts[col] receives the number of nodes the column needs.  One kernel serves both,
// so that the column code, expanded in place (the compiler then sees every array
// pointer come from the kernel arguments and issues global, not flat, accesses), is
// compiled once.
__global__ __launch_bounds__(64) void adjoint_record_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                                                           const double* __restrict__ estbl, TapeNode* tape_base,
                                                           const long long* off, const int* cap, int* node_out,
                                                           int* counts) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    // the tape header lives in LDS: every recorded operation updates it
    TapeCtx& tape = *ad_tape();
    if (counts) ad_begin(tape, nullptr, 0x7fffffff, ad_inputs(args.mkx));
    else ad_begin(tape, tape_base + off[j], cap[j], ad_inputs(args.mkx));
    const ThermoT<AD> T{consts, s_estbl};
    const Column<AD, AD> col(args, T, args.col0 + j, j, &tape);
    col.template run_body<false>();
    if (counts) {
        counts[args.col0 + j] = tape.n;
        return;
    }
    col.ad_export(node_out);
    args.ad.tapes[j] = tape;     // the sweep reads the header from HBM
    if (tape.overflow) record_error(args.err, UWSHCU_ERR_TAPE, args.col0 + j);
}

// Backward sweep, one column per thread, with a 64-entry window of recent adjoints
// (and its spare entry) per thread in LDS: 33 KB per workgroup of 64, so that the four
// workgroups a CU needs to give each SIMD one wave fit its LDS.
constexpr int kSweepWindow = 64;
__global__ __launch_bounds__(64) void adjoint_sweep_kernel(int n, int col0, int mkx, int ncnst, long long wld,
                                                          long long ldc, const TapeCtx* tapes, const int* node_out,
                                                          AdCotangents cot, AdGradients g) {
    __shared__ double s_win[kSweepWindow + 1][64];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int lane = threadIdx.x;
    auto win = [&](int s) -> double& { return s_win[s][lane]; };
    ad_reverse(tapes[j], node_out, (size_t)wld, j, mkx, ncnst, (size_t)ldc, col0 + j, cot, g, win, kSweepWindow);
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

hipError_t launch_adjoint_count(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                int* counts, hipStream_t stream) {
    hipLaunchKernelGGL(adjoint_record_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, args, n, dconsts, destbl,
                       (TapeNode*)nullptr, (const long long*)nullptr, (const int*)nullptr, (int*)nullptr, counts);
    return hipGetLastError();
}

hipError_t launch_adjoint(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                          TapeNode* tape_base, const long long* off, const int* cap, int* node_out,
                          const AdCotangents& cot, const AdGradients& g, hipStream_t stream) {
    const int block = 64;
    const int grid = (n + block - 1) / block;
    hipLaunchKernelGGL(adjoint_record_kernel, dim3(grid), dim3(block), 0, stream, args, n, dconsts, destbl, tape_base,
                       off, cap, node_out, (int*)nullptr);
    hipLaunchKernelGGL(adjoint_sweep_kernel, dim3(grid), dim3(block), 0, stream, n, args.col0, args.mkx, args.ncnst,
                       args.wld, (long long)args.ldc, (const TapeCtx*)args.ad.tapes, (const int*)node_out, cot, g);
    return hipGetLastError();
}

hipError_t launch_param_sum(const double* par, long long ldc, int ncol, double* sum, hipStream_t stream) {
    hipLaunchKernelGGL(param_sum_kernel, dim3(3), dim3(256), 0, stream, par, ldc, ncol, sum);
    return hipGetLastError();
}

}  // namespace uwshcu
