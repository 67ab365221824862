// uwshcu_chunks.hip -- gather many physics chunks into one column range and scatter
// the results back, so that a timestep's chunks run as one launch of the column
// pipeline instead of one launch per chunk (the reference calls the scheme once per
// chunk, physpkg.F90:782-803 -> convect_shallow.F90:605-619).
//
// Each field is a [rows][pcols] array per chunk (Fortran (pcols, rows) layout) and a
// [rows][n] array in the packed range, n = sum of the chunks' ncol.  Thread g of the
// launch owns packed column g: it finds its chunk once (binary search in the column
// offsets) and then copies its column of every field, so consecutive threads touch
// consecutive addresses of both sides.  Only the first ncol columns of a chunk are
// read or written; its padding up to pcols stays untouched.
#include <hip/hip_runtime.h>

#include "uwshcu_kernels.h"

namespace uwshcu {

namespace {

__device__ inline int chunk_of(const ChunkTable& t, long long g) {
    int lo = 0, hi = t.nchunk;   // t.off[lo] <= g < t.off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t.off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void chunks_gather_kernel(ChunkTable t) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < t.n; g += stride) {
        const int c = chunk_of(t, g);
        const long long i = g - t.off[c];
        const long long ld = t.pcols[c];
        for (int f = 0; f < t.nfield; ++f) {
            const double* src = t.ptr[(size_t)f * t.nchunk + c];
            double* dst = t.packed[f];
            const int rows = t.rows[f];
            if (src)
                for (int r = 0; r < rows; ++r) dst[(long long)r * t.n + g] = src[(long long)r * ld + i];
            else
                for (int r = 0; r < rows; ++r) dst[(long long)r * t.n + g] = 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void chunks_scatter_kernel(ChunkTable t) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < t.n; g += stride) {
        const int c = chunk_of(t, g);
        const long long i = g - t.off[c];
        const long long ld = t.pcols[c];
        for (int f = 0; f < t.nfield; ++f) {
            double* dst = t.ptr[(size_t)f * t.nchunk + c];
            if (!dst) continue;
            const double* src = t.packed[f];
            const int rows = t.rows[f];
            for (int r = 0; r < rows; ++r) dst[(long long)r * ld + i] = src[(long long)r * t.n + g];
        }
    }
}

unsigned grid_for(long long n) {
    const long long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

hipError_t launch_chunks_gather(const ChunkTable& t, hipStream_t stream) {
    hipLaunchKernelGGL(chunks_gather_kernel, dim3(grid_for(t.n)), dim3(256), 0, stream, t);
    return hipGetLastError();
}

hipError_t launch_chunks_scatter(const ChunkTable& t, hipStream_t stream) {
    hipLaunchKernelGGL(chunks_scatter_kernel, dim3(grid_for(t.n)), dim3(256), 0, stream, t);
    return hipGetLastError();
}

}  // namespace uwshcu
