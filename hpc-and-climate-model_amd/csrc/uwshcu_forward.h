// uwshcu_forward.h -- the forward kernel template, instantiated per precision.
#ifndef UWSHCU_FORWARD_H
#define UWSHCU_FORWARD_H

#include "uwshcu_kernels.h"

namespace uwshcu {

// One thread per column; the saturation table is staged into LDS once per workgroup.
// Thread j of the launch works in workspace slot j on column col0 + j, or, with a
// column map, on column colmap[col0 + j] while col0 + j < *colcount.
template <class R, bool ZW, bool SPLIT, bool FAST, int PART = PART_ALL>
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
    const ThermoT<double, FAST> T{consts, s_estbl};
    const Column<R, double, ZW, FAST> c(args, T, col, slot);
    c.template run_body<SPLIT, PART>();
}

// Two register budgets for the same column code.  forward_kernel lets it use all 512
// registers of a lane (one wavefront per SIMD): right for launches with fewer
// wavefronts than SIMDs, where nothing could run beside it.  forward_kernel_occ2
// allows half that (some spilling) so that two wavefronts share a SIMD and hide each
// other's FP64 latency: 1.37x faster at 1 M columns.
template <class R, bool ZW = false, bool FAST = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1)))
void forward_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts, const double* __restrict__ estbl) {
    forward_body<R, ZW, false, FAST>(args, n, consts, estbl);   // never the first two launches of a split
}
template <class R, bool ZW = false, bool FAST = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
void forward_kernel_occ2(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                         const double* __restrict__ estbl) {
    forward_body<R, ZW, true, FAST>(args, n, consts, estbl);
}

// Register budgets (wavefronts per SIMD) of the parts of a split launch.
#ifndef UWSHCU_OCC_CLOSURE
#define UWSHCU_OCC_CLOSURE 2
#endif
#ifndef UWSHCU_OCC_SORT
#define UWSHCU_OCC_SORT 2
#endif
#ifndef UWSHCU_OCC_REST
#define UWSHCU_OCC_REST 2
#endif
#ifndef UWSHCU_OCC_CIN2
#define UWSHCU_OCC_CIN2 2
#endif

// One part (PART_*) of a split launch, with OCC wavefronts per SIMD as its register budget.
template <class R, bool ZW, bool FAST, int PART, int OCC>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(OCC)))
void forward_kernel_part(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                         const double* __restrict__ estbl) {
    forward_body<R, ZW, true, FAST, PART>(args, n, consts, estbl);
}

// Level-parallel set-up: thread t of the launch handles workspace slot t % n and
// level t / n (+1 for stages 2-4), so that a wavefront covers consecutive columns.
// Stages 1 and 2: the set-A environment and the wet bulb, before the column pass.
// Stages 3 and 4: a listed column's set-B environment (S22) after the first pass of a
// split launch, in the column's own slot (column - slot0), where it continues.
template <class R, int STAGE, bool ZW = false, bool FAST = false>
__global__ __launch_bounds__(256) void setup_kernel(FwdArgs args, int n, const uwshcu_consts* __restrict__ consts,
                                                    const double* __restrict__ estbl) {
    __shared__ double s_estbl[UWSHCU_ESTBL_LEN];
    int nvalid = n;
    if (args.colmap) {
        const int left = *args.colcount - args.col0;
        nvalid = left < n ? left : n;
    }
    {   // whole workgroup past the listed columns (in one level row): nothing to do
        const long long t0 = (long long)blockIdx.x * blockDim.x;
        const long long j0 = t0 % n;
        if (j0 >= nvalid && j0 + blockDim.x <= n) return;
    }
    for (int t = threadIdx.x; t < UWSHCU_ESTBL_LEN; t += blockDim.x) s_estbl[t] = estbl[t];
    __syncthreads();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = (int)(t % n);
    const int k = (int)(t / n) + (STAGE == 1 ? 0 : 1);
    if (k > args.mkx || j >= nvalid) return;
    const int col = args.colmap ? args.colmap[args.col0 + j] : args.col0 + j;
    const ThermoT<double, FAST> T{consts, s_estbl};
    const Column<R, double, ZW, FAST> c(args, T, col, STAGE >= 3 ? col - args.slot0 : j);
    if (STAGE >= 3 && c.rstate(RS_CONT) == 0.0) return;   // it ended in its first pass
    if (STAGE == 1) c.setup_level1(k);
    else if (STAGE == 2) c.setup_level2(k);
    else if (STAGE == 3) c.setup_level_b1(k);
    else c.setup_level_b2(k);
}

static __global__ void clear_count_kernel(int* count) {
    if (threadIdx.x < 2) count[threadIdx.x] = 0;
}

static __global__ __launch_bounds__(256) void fill_zero_kernel(int* p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// Ordered compaction of marks[0, n) into list (column col0 + slot, ascending) and
// *count, in three steps: per-256 block sums, their exclusive scan (one workgroup),
// then each block's entries at its offset.  Keeping the column order keeps the
// listed columns' workspace accesses as coalesced as the marks allow.
static __device__ inline int block_prefix_256(int flag, int* s_wave, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(flag);
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave) before += s_wave[w];
        total += s_wave[w];
    }
    return before + in_wave;
}
static __global__ __launch_bounds__(256) void mark_block_sum_kernel(const int* marks, int n, int* bsum) {
    __shared__ int s_wave[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    int total;
    block_prefix_256(g < n ? marks[g] : 0, s_wave, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}
static __global__ __launch_bounds__(1024) void block_sum_scan_kernel(int* bsum, int nb, int* count) {
    __shared__ int s[1024];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int b = c0 + (int)threadIdx.x;
        const int v = b < nb ? bsum[b] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {   // inclusive Hillis-Steele scan
            const int add = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nb) bsum[b] = carry + s[threadIdx.x] - v;   // exclusive
        carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = carry;
}
static __global__ __launch_bounds__(256) void mark_scatter_kernel(const int* marks, int n, int col0, const int* boff,
                                                                  int* list) {
    __shared__ int s_wave[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int flag = g < n ? marks[g] : 0;
    int total;
    const int pos = block_prefix_256(flag, s_wave, total);
    if (flag) list[boff[blockIdx.x] + pos] = col0 + g;
}

// The forward pass of columns [col0, col0 + n) (or of the mapped columns): the two
// level-parallel set-up stages, then the column kernel.
//
// Large launches (at least two wavefronts per SIMD) split the column pass, so that
// wavefronts are not held by a few of their lanes and each kernel holds only the live
// state of its own stages (the column code as one kernel needs about 500 registers per
// lane and spills 1.7 KB of them at two wavefronts per SIMD):
//  1. every column runs the environment, CIN and trigger stages (PART_CLOSURE); a
//     column that reaches the cloud-base closure saves its scalars and marks itself,
//     and the marks are compacted into list 1 in column order;
//  2. the listed (triggered) columns, packed densely, run the rest of the first pass
//     in four launches: the buoyancy sorting (PART_SORT), the stages from the forcedCu
//     filter through the in-cloud diagnostics (PART_REST), the set-B environment level
//     by level (setup stages 3 and 4), and the implicit-CIN iteration up to its CIN
//     (PART_CIN2), where a column that convects a second time (a few per cent) saves
//     its scalars and marks itself for list 2;
//  3. the columns of list 2 run their second convection pass.
// Each column keeps its workspace slot throughout.  Where a column's working set is
// large (72 levels x 10 constituents), step 1 runs the whole first pass instead and
// step 2 is left out.  Without the split, a wavefront of a launch with 10 % convecting
// columns still runs the whole plume for the few that convect, and nearly every
// wavefront carries one or two second passes.
//
// args0.defer_list / defer_count / defer_state (nullable) give the lists' storage:
// defer_count[0..1]; defer_list[0..defer_list_ints(n)): list 1, list 2, the marks of
// list 1 and the compaction's block sums; and NRESUME x wld doubles.
template <class R, bool ZW = false, bool FAST = false>
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
    hipLaunchKernelGGL((setup_kernel<R, 1, ZW, FAST>), dim3((unsigned)((t1 + 255) / 256)), dim3(256), 0, stream, args, n,
                       dconsts, destbl);
    hipLaunchKernelGGL((setup_kernel<R, 2, ZW, FAST>), dim3((unsigned)((t2 + 255) / 256)), dim3(256), 0, stream, args, n,
                       dconsts, destbl);
    const bool large = n >= 2 * 1024 * 64 && !args0.sparse_map;
    const bool split = large && defer_list && defer_count && args0.defer_state && !args.colmap;
    // The closure split pays where a column's working set is small: measured 1-20 %
    // faster at 35 levels and 3 constituents (18 KB of workspace per column), 11 %
    // slower at 72 levels and 10 constituents (54 KB), where the gaps in the packed
    // list cost the second launch more than the idle lanes it saves.
    const bool closure_split =
        split && ws_bytes_per_column(args.mkx, args.ncnst, sizeof(R), sizeof(double)) <= (size_t)32 * 1024;
    const unsigned nblk = (unsigned)((n + 255) / 256);
    int* const marks = split ? defer_list + 2 * (size_t)n : nullptr;
    int* const bsum = split ? defer_list + 3 * (size_t)n : nullptr;
    if (split) {
        hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(64), 0, stream, defer_count);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(nblk), dim3(256), 0, stream, marks, n);
        args.defer_list = defer_list;
        args.defer_count = defer_count;
        args.defer_mark = marks;
        args.defer_at_closure = closure_split ? 1 : 0;
    }
    // at least two wavefronts per SIMD of the 256-CU part: the occupancy-2 build, or the
    // closure part alone
    if (closure_split)
        hipLaunchKernelGGL((forward_kernel_part<R, ZW, FAST, PART_CLOSURE, UWSHCU_OCC_CLOSURE>), dim3((n + 127) / 128),
                           dim3(128), 0, stream, args, n, dconsts, destbl);
    else if (large)
        hipLaunchKernelGGL((forward_kernel_occ2<R, ZW, FAST>), dim3((n + 127) / 128), dim3(128), 0, stream, args, n, dconsts, destbl);
    else
        hipLaunchKernelGGL((forward_kernel<R, ZW, FAST>), dim3((n + 127) / 128), dim3(128), 0, stream, args, n, dconsts, destbl);
    if (split) {
        auto compact = [&](int* list, int* count) {   // marks -> list in column order
            hipLaunchKernelGGL(mark_block_sum_kernel, dim3(nblk), dim3(256), 0, stream, (const int*)marks, n, bsum);
            hipLaunchKernelGGL(block_sum_scan_kernel, dim3(1), dim3(1024), 0, stream, bsum, (int)nblk, count);
            hipLaunchKernelGGL(mark_scatter_kernel, dim3(nblk), dim3(256), 0, stream, (const int*)marks, n,
                               args0.col0, (const int*)bsum, list);
        };
        FwdArgs b = args0;
        if (closure_split) {
            compact(defer_list, defer_count);
            hipLaunchKernelGGL(fill_zero_kernel, dim3(nblk), dim3(256), 0, stream, marks, n);
            // 2. the triggered columns, packed: (a) from the closure through the buoyancy
            // sorting, (b) from there through the in-cloud diagnostics, (c) the set-B
            // environment, level-parallel, (d) the implicit-CIN iteration's CIN, marking
            // the columns that convect a second time.  Each part holds only its own live
            // state (the saved scalars carry the rest).
            b.env_ready = 1;
            b.resume = 1;
            b.slot0 = args0.col0;
            b.defer_at_closure = 0;
            b.defer_mark = nullptr;
            b.defer_list = nullptr;
            b.defer_count = nullptr;
            b.colmap = defer_list;
            b.colcount = defer_count;
            b.col0 = 0;
            hipLaunchKernelGGL((forward_kernel_part<R, ZW, FAST, PART_SORT, UWSHCU_OCC_SORT>), dim3((n + 127) / 128),
                               dim3(128), 0, stream, b, n, dconsts, destbl);
            b.resume = 3;
            hipLaunchKernelGGL((forward_kernel_part<R, ZW, FAST, PART_REST, UWSHCU_OCC_REST>), dim3((n + 127) / 128),
                               dim3(128), 0, stream, b, n, dconsts, destbl);
            hipLaunchKernelGGL((setup_kernel<R, 3, ZW, FAST>), dim3((unsigned)((t2 + 255) / 256)), dim3(256), 0, stream,
                               b, n, dconsts, destbl);
            hipLaunchKernelGGL((setup_kernel<R, 4, ZW, FAST>), dim3((unsigned)((t2 + 255) / 256)), dim3(256), 0, stream,
                               b, n, dconsts, destbl);
            b.resume = 4;
            b.defer_mark = marks;
            hipLaunchKernelGGL((forward_kernel_part<R, ZW, FAST, PART_CIN2, UWSHCU_OCC_CIN2>), dim3((n + 127) / 128),
                               dim3(128), 0, stream, b, n, dconsts, destbl);
        } else {
            // the first launch marked the columns that need a second convection pass
            b.env_ready = 1;
            b.slot0 = args0.col0;
            b.col0 = 0;
        }
        compact(defer_list + n, defer_count + 1);
        // 3. the second convection passes (a few per cent of the columns), in one launch:
        // as two parts like the first pass they took 7.4 ms against 6.7 ms at 1 M columns
        FwdArgs c = b;
        c.resume = 2;
        c.defer_mark = nullptr;
        c.defer_list = nullptr;
        c.defer_count = nullptr;
        c.defer_at_closure = 0;
        c.colmap = defer_list + n;
        c.colcount = defer_count + 1;
        hipLaunchKernelGGL((forward_kernel<R, ZW, FAST>), dim3((n + 127) / 128), dim3(128), 0, stream, c, n, dconsts, destbl);
    }
    return hipGetLastError();
}

}  // namespace uwshcu
#endif
