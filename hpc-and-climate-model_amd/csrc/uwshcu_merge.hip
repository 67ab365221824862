// uwshcu_merge.hip -- the driver-side merge that follows the scheme in
// convect_shallow_tend (convect_shallow.F90:616-660, 790-836), on device arrays.
//
// One thread per column; its loops over levels read and write one double per level
// and field, consecutive columns at consecutive addresses.  Every value is one add,
// one multiply, a copy or a comparison, so the results are the driver's bit for bit.
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "uwshcu.h"

namespace uwshcu {

namespace {

__global__ __launch_bounds__(256) void merge_shallow_kernel(int pcols, int ncol, int pver, double latvap,
                                                            uwshcu_merge_fields f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pcols) return;
    f.massflx_sh[i] = f.cbmf[i];               // pbuf(massflx_sh_idx)(1:pcols) = cbmf(1:pcols)
    f.freqsh[i] = 0.0;                         // :790 freqsh(:) = 0 over all pcols
    if (i >= ncol) return;
    const size_t P = (size_t)pcols;
    // convect_shallow.F90:616-620: in-cloud water, shallow rain + snow production
    for (int k = 0; k < pver; ++k) {
        const size_t x = (size_t)k * P + i;
        f.icwmr[x] = f.qcu[x];
        const double rs = f.qrten[x] + f.qsten[x];
        f.rprdsh[x] = rs;
        f.rprdtot[x] = rs + f.rprddp[x];       // :822
        f.qc[x] = f.qc[x] + f.qc2[x];          // :829
    }
    // :653-654: convective fluxes of sl and qt in energy units
    for (int k = 0; k <= pver; ++k) {
        const size_t x = (size_t)k * P + i;
        f.cmfsl[x] = f.slflx[x];
        f.cmflq[x] = f.qtflx[x] * latvap;
    }
    // :790-795: frequency of shallow convection, as the driver computes it
    double mx = f.umf[i];
    for (int k = 1; k < pver; ++k) {
        const double v = f.umf[(size_t)k * P + i];
        if (v > mx) mx = v;
    }
    f.freqsh[i] = (mx <= 0.0) ? 1.0 : 0.0;
    // :806: cmfmc(:ncol, :pver) += cmfmc2(:ncol, :pver)
    for (int k = 0; k < pver; ++k) {
        const size_t x = (size_t)k * P + i;
        f.cmfmc[x] = f.cmfmc[x] + f.umf[x];
    }
    // :816-821: cloud top and base of either scheme and their pressures
    double cnt = f.cnt[i], cnb = f.cnb[i];
    if (f.cnt2[i] < cnt) cnt = f.cnt2[i];
    if (f.cnb2[i] > cnb) cnb = f.cnb2[i];
    f.cnt[i] = cnt;
    f.cnb[i] = cnb;
    int kt = (int)cnt, kb = (int)cnb;
    kt = kt < 1 ? 1 : (kt > pver ? pver : kt);
    kb = kb < 1 ? 1 : (kb > pver ? pver : kb);
    f.pcnt[i] = f.pmid[(size_t)(kt - 1) * P + i];
    f.pcnb[i] = f.pmid[(size_t)(kb - 1) * P + i];
    f.rliq[i] = f.rliq[i] + f.rliq2[i];         // :830
}

}  // namespace

hipError_t launch_merge_shallow(int pcols, int ncol, int pver, double latvap, const uwshcu_merge_fields& f,
                                hipStream_t stream) {
    hipLaunchKernelGGL(merge_shallow_kernel, dim3((pcols + 255) / 256), dim3(256), 0, stream, pcols, ncol, pver,
                       latvap, f);
    return hipGetLastError();
}

}  // namespace uwshcu
