// Parameter maps on the device: y = B x (+ c) for a sparse matrix B in CSR, the
// hat filter of the reference's shape-optimisation example (z = B_ij @ p,
// Examples/Shells_Mannheim_Multihalle_Shape.ipynb) and, with B^T in CSR, its
// gradient map (sens @ B_ij).  One wavefront per row: the lanes split the row's
// entries, so rows of any length stay coalesced, and a fixed-order butterfly
// sums them (bitwise-reproducible results).
#include "jsso_internal.hpp"

namespace jsso {

__global__ __launch_bounds__(256) void csr_spmv_kernel(int32_t n_rows, const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colind,
                                                       const double* __restrict__ vals,
                                                       const double* __restrict__ x, const double* __restrict__ c,
                                                       double* __restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n_rows) return;
    double s = 0.0;
    for (int k = rowptr[r] + lane; k < rowptr[r + 1]; k += 64) s += vals[k] * x[colind[k]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[r] = c ? s + c[r] : s;
}

int launch_csr_spmv(int32_t n_rows, const int32_t* rowptr, const int32_t* colind, const double* vals,
                    const double* x, const double* c, double* y, hipStream_t s) {
    if (n_rows <= 0) return JSSO_OK;
    hipLaunchKernelGGL(csr_spmv_kernel, dim3((unsigned)((n_rows + 3) / 4)), dim3(256), 0, s, n_rows, rowptr, colind,
                       vals, x, c, y);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

}  // namespace jsso
