// uwshcu_zw.hip -- FP64 forward kernel of the modified scheme compute_zwshcu
// (zwshcu.F90:630-4969): the column code of the UW scheme with the options of
// uwshcu_zw_params compiled in.  -ffp-contract=off as for the FP64 UW kernel.
#include "uwshcu_forward.h"

namespace uwshcu {

hipError_t launch_forward_zw(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                             hipStream_t stream) {
    return launch_column_pipeline<double, true>(args, n, dconsts, destbl, stream);
}

}  // namespace uwshcu
