// mathprobe.hip -- TEST HARNESS: evaluates the portable FP64 functions of
// uwshcu_math.h, sqrt and division (the operator, div_fast and div_rcp) on the GPU and
// on the host, so that a test can check that both give identical bits for the same
// inputs; and the adjoint sweep's atomic FP64 addition (adj_add) on the GPU.
#include <hip/hip_runtime.h>

#include "uwshcu_adjoint.h"
#include "uwshcu_math.h"

using namespace uwshcu;

constexpr int NOUT = 10;

UWSHCU_HD inline void probe_one(double x, double y, double* o) {
    o[0] = pm_exp(x);
    o[1] = pm_log(y);
    o[2] = pm_pow(y, x);
    o[3] = pm_erfc(x);
    o[4] = sqrt(y);
    o[5] = x / y;
    o[6] = fmax(x, y) - fmin(x, y);
    o[7] = floor(x * 3.7);
    o[8] = div_fast(x, y);
    o[9] = div_rcp(x, y, div_fast(1.0, y));
}

__global__ void probe_kernel(int n, const double* x, const double* y, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) probe_one(x[i], y[i], out + NOUT * (size_t)i);
}

extern "C" int probe_gpu(int n, const double* x, const double* y, double* out) {
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess) return 1;
    if (hipMalloc(&dy, n * sizeof(double)) != hipSuccess) return 1;
    if (hipMalloc(&dout, NOUT * (size_t)n * sizeof(double)) != hipSuccess) return 1;
    if (hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (hipMemcpy(dy, y, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 2;
    hipLaunchKernelGGL(probe_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, dx, dy, dout);
    if (hipDeviceSynchronize() != hipSuccess) return 3;
    if (hipMemcpy(out, dout, NOUT * (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 4;
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return 0;
}

extern "C" int probe_host(int n, const double* x, const double* y, double* out) {
    for (int i = 0; i < n; ++i) probe_one(x[i], y[i], out + NOUT * (size_t)i);
    return 0;
}

// out[i] = x[i], then adj_add(&out[i], y[i]) on the GPU (one thread per element)
__global__ void atomic_add_kernel(int n, const double* y, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adj_add((UWSHCU_GMEM double*)(out + i), y[i]);
}

extern "C" int probe_atomic_add_gpu(int n, const double* x, const double* y, double* out) {
    double *dy = nullptr, *dout = nullptr;
    if (hipMalloc(&dy, n * sizeof(double)) != hipSuccess) return 1;
    if (hipMalloc(&dout, n * sizeof(double)) != hipSuccess) return 1;
    if (hipMemcpy(dy, y, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (hipMemcpy(dout, x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 2;
    hipLaunchKernelGGL(atomic_add_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, dy, dout);
    if (hipDeviceSynchronize() != hipSuccess) return 3;
    if (hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 4;
    (void)hipFree(dy);
    (void)hipFree(dout);
    return 0;
}
