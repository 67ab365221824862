// uwshcu_fast.hip -- the fast FP64 forward of the mixed-precision mode: the column
// code with every state in FP64 and the saturation adjustment solved to convergence
// (ThermoT<double, true>) instead of by the reference's ten fixed Newton steps.
// Branch decisions within 1e-8 of their branch point are flagged for the FP64
// recompute.  Compiled with -ffp-contract=off like the FP64 path, so that the two differ
// only where the saturation adjustment's converged root differs from the ten-step
// iterate, and the host build of the same code gives the same bits.
#include "uwshcu_forward.h"

namespace uwshcu {

hipError_t launch_forward_fast(const FwdArgs& args, int n, const uwshcu_consts* dconsts, const double* destbl,
                               hipStream_t stream) {
    return launch_column_pipeline<double, false, true>(args, n, dconsts, destbl, stream);
}

}  // namespace uwshcu
