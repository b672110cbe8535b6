// Numeric assembly on the 6x6 block pattern (gathers over the element blocks,
// fixed summation order, no atomics), constraint masking and the block-Jacobi
// scaling the solver uses.  The solver needs only the factors of the diagonal
// blocks and the scaled off-diagonal blocks in sliced ELL; both are gathered
// straight from the element blocks.  K and Kc in block CSR (export, y = K x,
// the plain path) are gathered on demand.
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

// Per node: the constrained diagonal block, summed from its element blocks in
// the same order as K, its Cholesky factor, the inverse factor and the
// block-Jacobi inverse Minv_i = Linv_i^T Linv_i.  status[0] |= 1 when a block
// is not positive definite.
__global__ __launch_bounds__(128) void diag_factor_kernel(int32_t n_nodes, const int32_t* __restrict__ diag,
                                                          const int32_t* __restrict__ cptr,
                                                          const int32_t* __restrict__ csrc,
                                                          const uint8_t* __restrict__ fixed,
                                                          const double* __restrict__ ke,
                                                          double* __restrict__ Linv, double* __restrict__ Minv,
                                                          int32_t* status) {
    const int i = (int)(xcd_wg() * blockDim.x + threadIdx.x);
    if (i >= n_nodes) return;
    double A[6][6], L[6][6], Li[6][6];
    double B[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) B[q] = 0.0;
    const int k = diag[i];
    for (int c = cptr[k]; c < cptr[k + 1]; ++c) {
        const double* src = ke + (int64_t)csrc[c] * 36;
#pragma unroll
        for (int q = 0; q < 36; ++q) B[q] += src[q];
    }
    bool fx[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) fx[r] = fixed[6 * (int64_t)i + r] != 0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double brc = (fx[r] | fx[c]) ? (r == c ? 1.0 : 0.0) : B[r * 6 + c];
            const double bcr = (fx[r] | fx[c]) ? (r == c ? 1.0 : 0.0) : B[c * 6 + r];
            A[r][c] = 0.5 * (brc + bcr);
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

// One workgroup per sliced-ELL slot (64 positions, one per block row of the
// slice): the slot's off-diagonal blocks are summed from their element blocks
// (same order as K), masked (a constrained row or column of an off-diagonal
// block is zero), scaled Ks_ij = Linv_i Kc_ij Linv_j^T and written to the FP64,
// FP32 and FP16 ELL copies (the solver's residual replacement, its iterations,
// and the AMG smoothing products).  Padding positions (src < 0) keep their zeros.
//   phase 1: all 256 threads read the element blocks as (position, entry) pairs,
//            entry fastest (consecutive threads, consecutive values; every load
//            of a thread issued before the first sum), and stage the blocks, the
//            row factors L_i^-1 (64 consecutive rows: contiguous) and the column
//            factors L_j^-1 in LDS;
//   phase 2: wave w computes columns c = w, w + 4 of every position's Ks (lane =
//            position) and stores them into the value-slot layout (16-byte groups of
//            entries, lanes interleaved per group).
constexpr int ASM_WG = 256;
constexpr int MAX_CONTRIB = 2;   // element blocks per off-diagonal block read without a loop (2 on a quad mesh)

__global__ __launch_bounds__(ASM_WG) void ell_assemble_kernel(
    int64_t n_slots, const int32_t* __restrict__ ell_src, const int32_t* __restrict__ browof,
    const int32_t* __restrict__ colind, const int32_t* __restrict__ cptr, const int32_t* __restrict__ csrc,
    const uint8_t* __restrict__ fixed, const double* __restrict__ ke, const double* __restrict__ Linv,
    double* __restrict__ Ke64, float* __restrict__ Ke32, _Float16* __restrict__ Ke16, int32_t n_beam,
    const int32_t* __restrict__ status) {
    constexpr int P = ROWS_PER_WG;
    // K not symmetric (set by this assembly's element kernels): the solver copies
    // hold the symmetric part (K_ij + K_ji^T) / 2, the preconditioning operator of
    // the defect correction; the element blocks of (j, i) are the swapped
    // element-local blocks of the same elements
    const bool sym_part = (*status & STATUS_NONSYM) != 0;
    __shared__ double tile[P][37];
    __shared__ double li[P][37], lj[P][37];
    __shared__ int32_t cidx[P][MAX_CONTRIB], cstart[P], ccount[P], rowi[P], colj[P];
    __shared__ uint8_t rfix[P][6], cfix[P][6];
    const int64_t slot = blockIdx.x;
    if (slot >= n_slots) return;
    const int t = threadIdx.x;
    if (t < P) {
        const int32_t src = ell_src[slot * P + t];
        int n = 0, c0 = 0, i = 0, j = 0;
        int32_t e[MAX_CONTRIB] = {};
        if (src >= 0) {
            c0 = cptr[src];
            n = cptr[src + 1] - c0;
            for (int c = 0; c < n && c < MAX_CONTRIB; ++c) e[c] = csrc[c0 + c];
            i = browof[src];
            j = colind[src];
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                rfix[t][r] = fixed[6 * (int64_t)i + r];
                cfix[t][r] = fixed[6 * (int64_t)j + r];
            }
        }
#pragma unroll
        for (int c = 0; c < MAX_CONTRIB; ++c) cidx[t][c] = e[c];   // 0 (a valid block) when unused
        cstart[t] = c0;
        ccount[t] = n;
        rowi[t] = i;
        colj[t] = j;
    }
    __syncthreads();
    constexpr int NPAIR = P * 36 / ASM_WG;   // 9
    double v[NPAIR][MAX_CONTRIB], lv[NPAIR][2];
#pragma unroll
    for (int it = 0; it < NPAIR; ++it) {
        const int u = t + it * ASM_WG;
        const int p = u / 36, q = u % 36;
#pragma unroll
        for (int c = 0; c < MAX_CONTRIB; ++c) v[it][c] = ke[(int64_t)cidx[p][c] * 36 + q];
        lv[it][0] = Linv[(int64_t)rowi[p] * 36 + q];
        lv[it][1] = Linv[(int64_t)colj[p] * 36 + q];
    }
#pragma unroll
    for (int it = 0; it < NPAIR; ++it) {
        const int u = t + it * ASM_WG;
        const int p = u / 36, q = u % 36;
        const int n = ccount[p];
        double sum = 0.0;
#pragma unroll
        for (int c = 0; c < MAX_CONTRIB; ++c)
            if (c < n) sum += v[it][c];
        for (int c = MAX_CONTRIB; c < n; ++c) sum += ke[(int64_t)csrc[cstart[p] + c] * 36 + q];   // rare meshes
        if (sym_part) {
            const int qt = (q % 6) * 6 + q / 6;
            double st = 0.0;
            for (int c = 0; c < n; ++c) {
                const int32_t src = c < MAX_CONTRIB ? cidx[p][c] : csrc[cstart[p] + c];
                int64_t sw;
                if (src < 4 * n_beam) {
                    const int l = src & 3;
                    sw = (src & ~3) | ((l & 1) << 1) | (l >> 1);
                } else {
                    const int64_t qq = src - 4 * (int64_t)n_beam;
                    const int l = (int)(qq & 15);
                    sw = 4 * (int64_t)n_beam + (qq & ~(int64_t)15) + ((l & 3) << 2) + (l >> 2);
                }
                st += ke[sw * 36 + qt];
            }
            sum = 0.5 * (sum + st);
        }
        tile[p][q] = (n > 0 && !(rfix[p][q / 6] | cfix[p][q % 6])) ? sum : 0.0;
        li[p][q] = lv[it][0];
        lj[p][q] = lv[it][1];
    }
    __syncthreads();
    const int wave = t >> 6, lane = t & 63;
    if (ccount[lane] == 0) return;   // padding (or an empty block): its entries stay zero
    // value-slot layout: 16-byte groups of entries, lanes interleaved per group
    // (2 doubles / 4 floats per group)
    const int64_t base = slot * 36 * P;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = wave + 4 * h;
        if (c >= 6) break;
        // column c of T = B Lj^T, then column c of Ks = Li T (both factors lower triangular)
        double Tc[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
#pragma unroll
            for (int k2 = 0; k2 <= c; ++k2) s += tile[lane][r * 6 + k2] * lj[lane][c * 6 + k2];
            Tc[r] = s;
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
#pragma unroll
            for (int k2 = 0; k2 <= r; ++k2) s += li[lane][r * 6 + k2] * Tc[k2];
            const int e = r * 6 + c;
            Ke64[base + ((int64_t)(e / 2) * P + lane) * 2 + e % 2] = s;
            Ke32[base + ((int64_t)(e / 4) * P + lane) * 4 + e % 4] = (float)s;
            Ke16[base + ((int64_t)(e / 4) * P + lane) * 4 + e % 4] = (_Float16)s;
        }
    }
}

// K and Kc in block CSR, for callers that need them (export, y = K x, the plain path).
int launch_full_blocks(Plan* p, hipStream_t s) {
    if (p->full_valid || p->nnzb == 0) return JSSO_OK;
    const int64_t total = p->nnzb * 36;
    hipLaunchKernelGGL(gather_blocks_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, p->nnzb,
                       p->d_cptr, p->d_csrc, p->d_browof, p->d_colind, p->d_fixed, p->d_ke, p->d_K, p->d_Kc);
    JSSO_HIP(hipGetLastError());
    p->full_valid = true;
    return JSSO_OK;
}

int launch_assemble(Plan* p, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    p->full_valid = false;
    if (p->n_nodes > 0) {
        hipLaunchKernelGGL(diag_factor_kernel, xcd_grid((p->n_nodes + 127) / 128), dim3(128), 0, s, p->n_nodes,
                           p->d_diag, p->d_cptr, p->d_csrc, p->d_fixed, p->d_ke, p->d_Linv, p->d_Minv,
                           p->d_status);
        JSSO_HIP(hipGetLastError());
    }
    const int64_t n_slots = p->ell_slots / ROWS_PER_WG - 1;   // value slots, without the zero slot
    if (n_slots > 0) {
        hipLaunchKernelGGL(ell_assemble_kernel, dim3((unsigned)n_slots), dim3(ASM_WG), 0, s, n_slots, p->d_ell_src,
                           p->d_browof, p->d_colind, p->d_cptr, p->d_csrc, p->d_fixed, p->d_ke, p->d_Linv,
                           p->d_Ke64, p->d_Ke32, (_Float16*)p->d_Ke16, p->n_beam, p->d_status);
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

// r = f - op(K) u with op(K) = K or K^T (the unconstrained block CSR; u is zero
// at constrained DOFs) and r zero at constrained DOFs; vectors [6 n][nrhs].
// Thread per (scalar row, load case); a row's blocks in column order.
__global__ __launch_bounds__(256) void defect_residual_kernel(int32_t n_nodes, int nrhs,
                                                              const int32_t* __restrict__ rowptr,
                                                              const int32_t* __restrict__ colind,
                                                              const int32_t* __restrict__ tpos,
                                                              const uint8_t* __restrict__ fixed,
                                                              const double* __restrict__ K,
                                                              const double* __restrict__ f,
                                                              const double* __restrict__ u, double* __restrict__ r) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= 6 * (int64_t)n_nodes * nrhs) return;
    const int h = (int)(t % nrhs);
    const int64_t R = t / nrhs;
    if (fixed[R]) {
        r[t] = 0.0;
        return;
    }
    const int i = (int)(R / 6), rr = (int)(R % 6);
    double acc = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int64_t j = colind[k];
        // (K^T)_ij = (K_ji)^T: row rr of it is column rr of the block at the transposed position
        const double* b = tpos ? K + (int64_t)tpos[k] * 36 + rr : K + (int64_t)k * 36 + rr * 6;
        const int st = tpos ? 6 : 1;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += b[c * st] * u[(6 * j + c) * nrhs + h];
    }
    r[t] = f[t] - acc;
}

__global__ __launch_bounds__(256) void axpy_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += x[t];
}

int launch_defect_residual(Plan* p, const double* K, const double* f, const double* u, int nrhs, bool transpose,
                           double* r, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    const int64_t n = 6 * (int64_t)p->n_nodes * nrhs;
    if (n == 0) return JSSO_OK;
    hipLaunchKernelGGL(defect_residual_kernel, xcd_grid((n + 255) / 256), dim3(256), 0, s, p->n_nodes, nrhs,
                       p->d_rowptr, p->d_colind, transpose ? (const int32_t*)p->d_tpos : (const int32_t*)nullptr,
                       p->d_fixed, K, f, u, r);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

int launch_axpy(int64_t n, const double* x, double* y, hipStream_t s) {
    if (n <= 0) return JSSO_OK;
    hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, y);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
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
