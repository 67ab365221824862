// checkprobe.hip -- TEST HARNESS: evaluates the product's conservation check
// (column_budget / column_fails_check, uwshcu_check.h) on given column outputs, on
// the GPU or on the host, so that a test can compare its residuals with a
// restatement of CAM's check_energy_chng.
#include <hip/hip_runtime.h>

#include <vector>

#include "uwshcu_check.h"

using namespace uwshcu;

namespace {

// Array order of check_probe: layer fields, interface fields, constituents, column scalars.
enum { A_DP0, A_U0, A_V0, A_QVTEN, A_QLTEN, A_QITEN, A_STEN, A_UTEN, A_VTEN, A_QRTEN, A_QSTEN, A_EVAPC, A_CUFRC,
       A_QCU, A_QLU, A_QIU, A_QC, A_QTFLX, A_SLFLX, A_UMF, A_WU, A_FLXPRC1, A_FLXSNOW1, A_TRTEN, A_PRECIP, A_RLIQ,
       A_SNOW, A_CBMF, A_CUSH, A_N };

size_t rows(int k, int pver, int ncnst) {
    if (k <= A_QC) return (size_t)pver;
    if (k <= A_FLXSNOW1) return (size_t)pver + 1;
    if (k == A_TRTEN) return (size_t)pver * (size_t)(ncnst > 0 ? ncnst : 1);
    return 1;
}

FwdArgs make_args(int ncol, int pver, int ncnst, double dt, double* const* p) {
    FwdArgs a{};
    a.ncol = ncol; a.mkx = pver; a.ncnst = ncnst; a.ldc = ncol; a.dt = dt;
    a.dp0 = p[A_DP0]; a.u0 = p[A_U0]; a.v0 = p[A_V0]; a.qvten = p[A_QVTEN]; a.qlten = p[A_QLTEN];
    a.qiten = p[A_QITEN]; a.sten = p[A_STEN]; a.uten = p[A_UTEN]; a.vten = p[A_VTEN]; a.qrten = p[A_QRTEN];
    a.qsten = p[A_QSTEN]; a.evapc = p[A_EVAPC]; a.cufrc = p[A_CUFRC]; a.qcu = p[A_QCU]; a.qlu = p[A_QLU];
    a.qiu = p[A_QIU]; a.qc = p[A_QC]; a.qtflx = p[A_QTFLX]; a.slflx = p[A_SLFLX]; a.umf = p[A_UMF]; a.wu = p[A_WU];
    a.flxprc1 = p[A_FLXPRC1]; a.flxsnow1 = p[A_FLXSNOW1]; a.trten = p[A_TRTEN]; a.precip = p[A_PRECIP];
    a.rliq = p[A_RLIQ]; a.snow = p[A_SNOW]; a.cbmf = p[A_CBMF]; a.cush = p[A_CUSH];
    return a;
}

UWSHCU_HD inline void probe_col(const FwdArgs& a, const uwshcu_consts& c, int i, double tol, double* res,
                                int* fails) {
    double r[4];
    const bool ok = column_budget(a, c, i, r);
    for (int q = 0; q < 4; ++q) res[4 * (size_t)i + q] = r[q];
    fails[i] = (column_fails_check(a, c, i, tol) ? 1 : 0) | (ok ? 0 : 2);
}

__global__ void probe_kernel(FwdArgs a, uwshcu_consts c, double tol, double* res, int* fails) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.ncol) probe_col(a, c, i, tol, res, fails);
}

}  // namespace

// res[4*i + 0..3] = water residual, energy residual, max |qtflx|, max |slflx| of column i;
// fails[i] = bit 0: column_fails_check(tol), bit 1: a non-finite output.  Host arrays.
extern "C" int check_probe(int on_gpu, const uwshcu_consts* c, int ncol, int pver, int ncnst, double dt,
                           double* const* arr, double tol, double* res, int* fails) {
    if (!on_gpu) {
        const FwdArgs a = make_args(ncol, pver, ncnst, dt, arr);
        for (int i = 0; i < ncol; ++i) probe_col(a, *c, i, tol, res, fails);
        return 0;
    }
    std::vector<double*> d(A_N, nullptr);
    double* dres = nullptr;
    int* dfail = nullptr;
    int rc = 0;
    for (int k = 0; k < A_N && !rc; ++k) {
        const size_t n = rows(k, pver, ncnst) * (size_t)ncol;
        if (hipMalloc(&d[k], n * sizeof(double)) != hipSuccess ||
            hipMemcpy(d[k], arr[k], n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = 1;
    }
    if (!rc && (hipMalloc(&dres, 4 * (size_t)ncol * sizeof(double)) != hipSuccess ||
                hipMalloc(&dfail, (size_t)ncol * sizeof(int)) != hipSuccess))
        rc = 1;
    if (!rc) {
        const FwdArgs a = make_args(ncol, pver, ncnst, dt, d.data());
        hipLaunchKernelGGL(probe_kernel, dim3((ncol + 127) / 128), dim3(128), 0, 0, a, *c, tol, dres, dfail);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(res, dres, 4 * (size_t)ncol * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(fails, dfail, (size_t)ncol * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            rc = 2;
    }
    for (double* p : d) (void)hipFree(p);
    (void)hipFree(dres);
    (void)hipFree(dfail);
    return rc;
}
