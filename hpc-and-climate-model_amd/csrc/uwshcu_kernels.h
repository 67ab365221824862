// uwshcu_kernels.h -- launchers for the column kernels (host-side declarations).
#ifndef UWSHCU_KERNELS_H
#define UWSHCU_KERNELS_H

#include <hip/hip_runtime.h>
#include "uwshcu_column.h"

namespace uwshcu {

// Launch the FP64 forward kernel over columns [args.col0, args.col0 + n) of the chunk.
hipError_t launch_forward_fp64(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                               hipStream_t stream);
// Same for the mixed-precision kernel (FP32 bulk, FP64 saturation/CIN/fix-ups).
hipError_t launch_forward_mixed(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                                hipStream_t stream);
// Zero *count, then append to list every column of [0, args.ncol) whose amb_flag is
// set or whose outputs fail the FP64 conservation check at relative tolerance tol.
hipError_t launch_check_columns(const FwdArgs& args, const uwshcu_consts* dconsts, double tol, int* list, int* count,
                                hipStream_t stream);
// Set the two status words to zero (a fill kernel, so that graph capture records no memset node).
hipError_t launch_clear_status(int* err, hipStream_t stream);

}  // namespace uwshcu
#endif
