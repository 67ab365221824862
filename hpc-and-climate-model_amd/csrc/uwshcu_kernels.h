// uwshcu_kernels.h -- launchers for the column kernels (host-side declarations).
#ifndef UWSHCU_KERNELS_H
#define UWSHCU_KERNELS_H

#include <hip/hip_runtime.h>
#include "uwshcu_adjoint.h"

namespace uwshcu {

// Launch the FP64 forward kernel over columns [args.col0, args.col0 + n) of the chunk.
hipError_t launch_forward_fp64(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                               hipStream_t stream);
// The modified scheme compute_zwshcu (FP64; args.zw, args.omega, args.tpert, args.qpert).
hipError_t launch_forward_zw(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                             hipStream_t stream);
// The fast FP64 pass of UWSHCU_MIXED: saturation adjustment solved to convergence.
hipError_t launch_forward_fast(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                               hipStream_t stream);
// The FP32-plume kernel of UWSHCU_MIXED_UNCHECKED (FP64 saturation/CIN/fix-ups).
hipError_t launch_forward_mixed(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                hipStream_t stream);
// Zero *count, then append to list every column of [0, args.ncol) whose amb_flag is
// set or whose outputs fail the FP64 conservation check at relative tolerance tol.
hipError_t launch_check_columns(const FwdArgs& args, const uwshcu_consts* dconsts, double tol, int* list, int* count,
                                hipStream_t stream);
// Adjoint over columns [args.col0, args.col0 + n), in two launches.  Recording: each
// column records its FP64 forward pass on a tape paged from pool (page table of slot
// j at pages + j * maxp; at most cap nodes), writes the forward outputs and its output
// nodes (node_out), and takes its adjoint entries from the pool's adjoint array
// (args.ad holds the AD output values and the tape headers).  Sweep: zero adj[0, nadj)
// (nadj = the pool's adj_next after the recording), then sweep each recorded tape from
// the cotangents to the gradients.
hipError_t launch_adjoint_record(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                 TapePool* pool, int* pages, int maxp, int cap, int* node_out, hipStream_t stream);
hipError_t launch_adjoint_sweep(const FwdArgs& args, int n, double* adj, long long nadj, const int* node_out,
                                const AdCotangents& cot, const AdGradients& g, hipStream_t stream);
// sum[p] = sum over columns c < ncol of par[p*ldc + c], p = 0..2, in FP64.
hipError_t launch_param_sum(const double* par, long long ldc, int ncol, double* sum, hipStream_t stream);
// Set the two status words to zero (a fill kernel, so that graph capture records no memset node).
hipError_t launch_clear_status(int* err, hipStream_t stream);

// Device-side description of a set of chunks and of their packed column range
// (all pointers device pointers; see uwshcu_chunks.hip).
struct ChunkTable {
    int nchunk, nfield;
    long long n;               // packed columns: off[nchunk]
    const long long* off;      // [nchunk + 1] first packed column of each chunk
    const int* pcols;          // [nchunk] leading dimension of each chunk's arrays
    const int* rows;           // [nfield] rows of each field (levels, levels x constituents, or 1)
    double* const* packed;     // [nfield] packed arrays, [rows][n]
    double* const* ptr;        // [nfield][nchunk] the chunks' arrays (null: zero on gather, skip on scatter)
};
// packed <- chunks (every field of the table)
hipError_t launch_chunks_gather(const ChunkTable& t, hipStream_t stream);
// chunks <- packed (every field of the table with a non-null chunk array)
hipError_t launch_chunks_scatter(const ChunkTable& t, hipStream_t stream);
// The driver-side merge of uwshcu_merge_shallow_device (uwshcu_merge.hip).
hipError_t launch_merge_shallow(int pcols, int ncol, int pver, double latvap, const uwshcu_merge_fields& f,
                                hipStream_t stream);

}  // namespace uwshcu
#endif
