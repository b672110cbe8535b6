// Numeric assembly of K on the 6x6 block-CSR pattern (row gather over the
// element blocks, fixed summation order, no atomics), constraint masking and
// the block-Jacobi scaling the solver uses.
//
//   K_ij       = sum over contributions (element e, local a, b) of Ke_e[a][b]
//   Kc         = K with rows/columns of constrained DOFs replaced by identity
//   Kc_ii      = L_i L_i^T (Cholesky), Linv_i = L_i^-1
//   Ks_ij      = Linv_i Kc_ij Linv_j^T   (i != j; Ks_ii = I implicitly)
//
// The Lagrange-multiplier system of the reference (assemblemodel.py:111-163)
// and this constrained system have the same u: every prescribed displacement
// is zero, so u_c = 0 and K_ff u_f = f_f.
#include "jsso_internal.hpp"

namespace jsso {

// One thread per (block, entry).  Each K value is the sum of its element
// contributions in element order (beams, then quads, element index, a, b).
__global__ __launch_bounds__(256) void gather_blocks_kernel(int64_t nnzb, const int32_t* __restrict__ cptr,
                                                            const int32_t* __restrict__ csrc,
                                                            const int32_t* __restrict__ browof,
                                                            const int32_t* __restrict__ colind,
                                                            const uint8_t* __restrict__ fixed,
                                                            const double* __restrict__ ke,
                                                            double* __restrict__ K, double* __restrict__ Kc) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nnzb * 36) return;
    const int64_t k = tid / 36;
    const int q = (int)(tid % 36);
    double s = 0.0;
    const int c0 = cptr[k], c1 = cptr[k + 1];
    for (int c = c0; c < c1; ++c) s += ke[(int64_t)csrc[c] * 36 + q];
    K[tid] = s;
    const int64_t rdof = 6 * (int64_t)browof[k] + q / 6;
    const int64_t cdof = 6 * (int64_t)colind[k] + q % 6;
    double v = s;
    if (fixed[rdof] | fixed[cdof]) v = (rdof == cdof) ? 1.0 : 0.0;
    Kc[tid] = v;
}

// Per node: Cholesky of the constrained diagonal block, its inverse factor and
// the block-Jacobi inverse Minv_i = Linv_i^T Linv_i.  status[0] |= 1 when a
// block is not positive definite.
__global__ void diag_factor_kernel(int32_t n_nodes, const int32_t* __restrict__ diag,
                                   const double* __restrict__ Kc, double* __restrict__ Linv,
                                   double* __restrict__ Minv, int32_t* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    double A[6][6], L[6][6], Li[6][6];
    const double* src = Kc + (int64_t)diag[i] * 36;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            A[r][c] = 0.5 * (src[r * 6 + c] + src[c * 6 + r]);
            L[r][c] = 0.0;
            Li[r][c] = 0.0;
        }
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = A[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(d > 0.0)) { ok = false; d = 1.0; }
        const double ljj = sqrt(d);
        L[j][j] = ljj;
#pragma unroll
        for (int r = j + 1; r < 6; ++r) {
            double s = A[r][j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[r][k] * L[j][k];
            L[r][j] = s / ljj;
        }
    }
    // inverse of the lower-triangular factor
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        Li[j][j] = 1.0 / L[j][j];
#pragma unroll
        for (int r = j + 1; r < 6; ++r) {
            double s = 0.0;
#pragma unroll
            for (int k = j; k < r; ++k) s -= L[r][k] * Li[k][j];
            Li[r][j] = s / L[r][r];
        }
    }
    double* dst = Linv + (int64_t)i * 36;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) dst[r * 6 + c] = Li[r][c];
    double* m = Minv + (int64_t)i * 36;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += Li[k][r] * Li[k][c];
            m[r * 6 + c] = s;
        }
    if (!ok) atomicOr(status, 1);
}

// One thread per sliced-ELL position (slot-major, lane fastest): the scaled
// block Ks_ij = Linv_i Kc_ij Linv_j^T of that position in FP64, also rounded to
// FP32, written entry by entry so that each store of a wavefront is one
// contiguous run of 64 lanes.  Padding positions (src < 0) keep their zeros.
__global__ __launch_bounds__(256) void scale_blocks_kernel(int64_t n_pos, const int32_t* __restrict__ ell_src,
                                                           const int32_t* __restrict__ browof,
                                                           const int32_t* __restrict__ colind,
                                                           const double* __restrict__ Kc,
                                                           const double* __restrict__ Linv,
                                                           double* __restrict__ Ke64, float* __restrict__ Ke32) {
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n_pos) return;
    const int64_t src = ell_src[pos];
    if (src < 0) return;
    const int i = browof[src], j = colind[src];
    const double* B = Kc + src * 36;
    const double* Li = Linv + (int64_t)i * 36;
    const double* Lj = Linv + (int64_t)j * 36;
    double T[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k2 = 0; k2 <= c; ++k2) s += B[r * 6 + k2] * Lj[c * 6 + k2];  // B Lj^T
            T[r][c] = s;
        }
    const int64_t base = (pos / ROWS_PER_WG) * 36 * ROWS_PER_WG + pos % ROWS_PER_WG;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k2 = 0; k2 <= r; ++k2) s += Li[r * 6 + k2] * T[k2][c];
            const int64_t e = base + (int64_t)(r * 6 + c) * ROWS_PER_WG;
            Ke64[e] = s;
            Ke32[e] = (float)s;
        }
}

int launch_assemble(Plan* p, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    if (p->nnzb > 0) {
        const int64_t total = p->nnzb * 36;
        hipLaunchKernelGGL(gather_blocks_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           p->nnzb, p->d_cptr, p->d_csrc, p->d_browof, p->d_colind, p->d_fixed, p->d_ke,
                           p->d_K, p->d_Kc);
        JSSO_HIP(hipGetLastError());
    }
    if (p->n_nodes > 0) {
        hipLaunchKernelGGL(diag_factor_kernel, dim3((p->n_nodes + 127) / 128), dim3(128), 0, s, p->n_nodes,
                           p->d_diag, p->d_Kc, p->d_Linv, p->d_Minv, p->d_status);
        JSSO_HIP(hipGetLastError());
    }
    if (p->ell_slots > 0) {
        hipLaunchKernelGGL(scale_blocks_kernel, dim3((unsigned)((p->ell_slots + 255) / 256)), dim3(256), 0, s,
                           p->ell_slots, p->d_ell_src, p->d_browof, p->d_colind, p->d_Kc, p->d_Linv, p->d_Ke64,
                           p->d_Ke32);
        JSSO_HIP(hipGetLastError());
    }
    return JSSO_OK;
}

// y = K x on the full (unconstrained) block-CSR matrix, one thread per scalar row.
__global__ __launch_bounds__(256) void spmv_full_kernel(int32_t n_nodes, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const double* __restrict__ K,
                                                        const double* __restrict__ x, double* __restrict__ y) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= 6 * (int64_t)n_nodes) return;
    const int i = (int)(R / 6), r = (int)(R % 6);
    double acc = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const double* b = K + (int64_t)k * 36 + r * 6;
        const double* xj = x + 6 * (int64_t)colind[k];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += b[c] * xj[c];
    }
    y[R] = acc;
}

int launch_spmv_full(Plan* p, const double* K, const double* x, double* y, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    const int64_t n = 6 * (int64_t)p->n_nodes;
    if (n == 0) return JSSO_OK;
    hipLaunchKernelGGL(spmv_full_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p->n_nodes,
                       p->d_rowptr, p->d_colind, K, x, y);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

}  // namespace jsso
