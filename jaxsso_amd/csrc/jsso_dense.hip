// Dense direct solve of the constrained system -- the counterpart of the
// reference's jax_dense_solve (JaxSSO/solver.py:75-93), which factorises the
// dense Lagrange-augmented matrix with LU.  The augmented system and the
// constrained one have the same u (every prescribed displacement is zero), and
// the constrained matrix is SPD, so this path factorises
//
//   A = Kc  (K with the rows and columns of constrained DOFs replaced by identity)
//
// by a blocked right-looking Cholesky A = L L^T in FP64, then solves L y = f and
// L^T u = y.  A is dense N x N, column-major, N = 6 n_nodes rounded up to a
// multiple of the tile NB = 64 (the padding rows and columns hold the identity),
// and only its lower triangle is referenced.  Per tile column k:
//   dense_potrf   L_kk = chol(A_kk)                     one workgroup, tile in LDS
//   dense_trsm    L_ik = A_ik L_kk^-T        (i > k)     one workgroup per tile
//   dense_update  A_ij -= L_ik L_jk^T        (k < j <= i) one workgroup per tile,
//                                                        4x4 outputs per thread
// then per tile the diagonal solve (one wavefront, a reduction per row) and the
// update of the remaining right-hand side (coalesced over rows, forward; one
// wavefront per row, backward).  Memory: 8 N^2 bytes, owned by the plan.
//
// A plan whose assembly holds members parallel to global Y (the reference's
// non-orthogonal transformation, element.py:92-94) has a non-symmetric K, which
// the reference still solves directly (SuperLU, solver.py:195-209).  Such K is
// factorised here by LU with partial pivoting, P A = L U, blocked like the
// Cholesky: per tile column k
//   lu_pivot_search / lu_pivot_apply / lu_panel_update   one column of the panel
//                                     at a time: the largest |a_ij| below the
//                                     diagonal (first one on ties, as LAPACK),
//                                     the row swap inside the panel, the scaled
//                                     column and the rank-1 update of the panel
//   lu_swap_rest       the panel's row swaps applied to every other column
//   lu_trsm_row        U_kj = L_kk^-1 A_kj  (j > k), L_kk unit lower
//   dense_update<LU>   A_ij -= L_ik U_kj     (i, j > k)
// then P f, L y = P f (unit diagonal) and U u = y.  The transposed system
// (the reference's adjoint solve with K_aug^T, solver.py:236) factorises A^T.
#include <algorithm>

#include "jsso_internal.hpp"

namespace jsso {

constexpr int DNB = 64;   // tile edge

// full = false: the lower triangle only (Cholesky); transpose: A = Kc^T
__global__ __launch_bounds__(256) void dense_fill_kernel(int64_t nnzb, const int32_t* __restrict__ browof,
                                                         const int32_t* __restrict__ colind,
                                                         const uint8_t* __restrict__ fixed,
                                                         const double* __restrict__ K, double* __restrict__ A,
                                                         int64_t N, bool full, bool transpose) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nnzb * 36) return;
    const int64_t k = tid / 36;
    const int q = (int)(tid % 36);
    int64_t R = 6 * (int64_t)browof[k] + q / 6;
    int64_t C = 6 * (int64_t)colind[k] + q % 6;
    if (transpose) {
        const int64_t t = R;
        R = C;
        C = t;
    }
    if (!full && R < C) return;   // lower triangle only
    double v = K[tid];
    if (fixed[R] | fixed[C]) v = (R == C) ? 1.0 : 0.0;
    A[C * N + R] = v;
}

__global__ __launch_bounds__(256) void dense_pad_kernel(int64_t n, int64_t N, double* __restrict__ A) {
    const int64_t i = n + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) A[i * N + i] = 1.0;
}

// Cholesky of the diagonal tile k in LDS; status |= 1 if a pivot is not positive
__global__ __launch_bounds__(256) void dense_potrf_kernel(double* __restrict__ A, int64_t N, int k,
                                                          int32_t* status) {
    __shared__ double T[DNB][DNB + 1];
    const int t = threadIdx.x;
    double* tile = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    for (int e = t; e < DNB * DNB; e += 256) {
        const int c = e / DNB, r = e % DNB;
        T[r][c] = tile[(int64_t)c * N + r];
    }
    __syncthreads();
    bool bad = false;
    for (int j = 0; j < DNB; ++j) {
        double d = T[j][j];
        if (!(d > 0.0)) {
            bad = true;
            d = 1.0;
        }
        const double ljj = sqrt(d);
        __syncthreads();   // every thread has read T[j][j]
        if (t == 0) T[j][j] = ljj;
        for (int r = j + 1 + t; r < DNB; r += 256) T[r][j] /= ljj;
        __syncthreads();
        const int m = DNB - j - 1;
        for (int e = t; e < m * m; e += 256) {
            const int rr = j + 1 + e / m, cc = j + 1 + e % m;
            if (cc <= rr) T[rr][cc] -= T[rr][j] * T[cc][j];
        }
        __syncthreads();
    }
    for (int e = t; e < DNB * DNB; e += 256) {
        const int c = e / DNB, r = e % DNB;
        tile[(int64_t)c * N + r] = r >= c ? T[r][c] : 0.0;
    }
    if (bad && t == 0) atomicOr(status, 1);
}

// L_ik = A_ik L_kk^-T for the tiles below the diagonal of column k; thread = tile row
__global__ __launch_bounds__(DNB) void dense_trsm_kernel(double* __restrict__ A, int64_t N, int k) {
    __shared__ double Lk[DNB][DNB + 1];
    __shared__ double X[DNB][DNB + 1];
    const int i = k + 1 + blockIdx.x;
    const int r = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    double* ti = A + (int64_t)k * DNB * N + (int64_t)i * DNB;
    for (int c = 0; c < DNB; ++c) {
        Lk[r][c] = dk[(int64_t)c * N + r];
        X[r][c] = ti[(int64_t)c * N + r];
    }
    __syncthreads();
    for (int c = 0; c < DNB; ++c) {
        double s = X[r][c];
        for (int m = 0; m < c; ++m) s -= Lk[c][m] * X[r][m];
        X[r][c] = s / Lk[c][c];
    }
    for (int c = 0; c < DNB; ++c) ti[(int64_t)c * N + r] = X[r][c];
}

// Cholesky (LU = false): A_ij -= L_ik L_jk^T for every tile k < j <= i < nt
// (block b -> (i, j) in the lower triangle of the trailing matrix, row by row).
// LU: A_ij -= L_ik U_kj for every tile k < i, j < nt (block b -> (i, j) row-major
// over the m_tiles x m_tiles trailing tiles).  Lj[c][m] holds the
// second factor's column c: L_jk row c (Cholesky) or U_kj column c (LU).
template <bool LU>
__global__ __launch_bounds__(256) void dense_update_kernel(double* __restrict__ A, int64_t N, int k, int m_tiles) {
    __shared__ double Li[DNB][DNB + 1];
    __shared__ double Lj[DNB][DNB + 1];
    const int b = blockIdx.x;
    int ii, jj;
    if (LU) {
        ii = b / m_tiles;
        jj = b % m_tiles;
    } else {
        ii = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
        while ((ii + 1) * (ii + 2) / 2 <= b) ++ii;
        while (ii * (ii + 1) / 2 > b) --ii;
        jj = b - ii * (ii + 1) / 2;
    }
    const int i = k + 1 + ii, j = k + 1 + jj;
    const int t = threadIdx.x;
    const double* pi = A + (int64_t)k * DNB * N + (int64_t)i * DNB;
    if (LU) {
        const double* pj = A + (int64_t)j * DNB * N + (int64_t)k * DNB;   // tile (k, j)
        for (int e = t; e < DNB * DNB; e += 256) {
            const int m = e / DNB, r = e % DNB;
            Li[r][m] = pi[(int64_t)m * N + r];
            Lj[m][r] = pj[(int64_t)m * N + r];                            // U_kj[r][m] -> column m
        }
    } else {
        const double* pj = A + (int64_t)k * DNB * N + (int64_t)j * DNB;
        for (int e = t; e < DNB * DNB; e += 256) {
            const int m = e / DNB, r = e % DNB;
            Li[r][m] = pi[(int64_t)m * N + r];
            Lj[r][m] = pj[(int64_t)m * N + r];
        }
    }
    __syncthreads();
    const int tr = t % 16, tc = t / 16;
    double acc[4][4] = {};
    for (int m = 0; m < DNB; ++m) {
        double a[4], bb[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            a[x] = Li[tr + 16 * x][m];
            bb[x] = Lj[tc + 16 * x][m];
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * bb[y];
    }
    double* cij = A + (int64_t)j * DNB * N + (int64_t)i * DNB;
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int r = tr + 16 * x, c = tc + 16 * y;
            cij[(int64_t)c * N + r] -= acc[x][y];
        }
}

__device__ inline double dense_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// --- LU with partial pivoting --------------------------------------------

// per workgroup: the largest |A[i][j]| over its rows i >= j (first index on ties)
__global__ __launch_bounds__(256) void lu_pivot_search_kernel(const double* __restrict__ A, int64_t N, int64_t j,
                                                              double* __restrict__ pval, int64_t* __restrict__ pidx) {
    __shared__ double sv[256];
    __shared__ int64_t si[256];
    const int t = threadIdx.x;
    double best = -1.0;
    int64_t bi = N;
    const double* col = A + j * N;
    for (int64_t i = j + (int64_t)blockIdx.x * 256 + t; i < N; i += (int64_t)gridDim.x * 256) {
        const double v = fabs(col[i]);
        if (v > best) {   // rows visited in increasing order per thread: keeps the first maximum
            best = v;
            bi = i;
        }
    }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            const double a = sv[t], c = sv[t + o];
            if (c > a || (c == a && si[t + o] < si[t])) {
                sv[t] = c;
                si[t] = si[t + o];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        pval[blockIdx.x] = sv[0];
        pidx[blockIdx.x] = si[0];
    }
}

// pivot of column j from the partial results; swap rows j and p inside the
// panel [c0, c0 + DNB); status |= 1 for a zero pivot
__global__ __launch_bounds__(DNB) void lu_pivot_apply_kernel(double* __restrict__ A, int64_t N, int64_t j, int64_t c0,
                                                             int nparts, const double* __restrict__ pval,
                                                             const int64_t* __restrict__ pidx,
                                                             int64_t* __restrict__ piv, int32_t* status) {
    __shared__ int64_t p_sh;
    const int t = threadIdx.x;
    if (t == 0) {
        double best = -1.0;
        int64_t bi = j;
        for (int b = 0; b < nparts; ++b)
            if (pval[b] > best || (pval[b] == best && pidx[b] < bi)) {
                best = pval[b];
                bi = pidx[b];
            }
        if (!(best > 0.0)) {
            atomicOr(status, 1);
            bi = j;
        }
        piv[j] = bi;
        p_sh = bi;
    }
    __syncthreads();
    const int64_t p = p_sh;
    if (p != j) {
        double* a = A + (c0 + t) * N;
        const double x = a[j];
        a[j] = a[p];
        a[p] = x;
    }
}

// rows i > j: l = a_ij / a_jj stored, and a_ic -= l a_jc for the panel columns c > j
__global__ __launch_bounds__(256) void lu_panel_update_kernel(double* __restrict__ A, int64_t N, int64_t j,
                                                              int64_t c0) {
    const int64_t i = j + 1 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const double d = A[j * N + j];
    const double l = d != 0.0 ? A[j * N + i] / d : 0.0;
    A[j * N + i] = l;
    for (int64_t c = j + 1; c < c0 + DNB; ++c) A[c * N + i] -= l * A[c * N + j];
}

// the panel's row swaps (rows c0 .. c0 + DNB - 1 in order) on every column
// outside the panel; thread = column
__global__ __launch_bounds__(256) void lu_swap_rest_kernel(double* __restrict__ A, int64_t N, int64_t c0,
                                                           const int64_t* __restrict__ piv) {
    int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= N - DNB) return;
    if (c >= c0) c += DNB;
    double* a = A + c * N;
    for (int64_t j = c0; j < c0 + DNB; ++j) {
        const int64_t p = piv[j];
        if (p != j) {
            const double x = a[j];
            a[j] = a[p];
            a[p] = x;
        }
    }
}

// U_kj = L_kk^-1 A_kj for the tiles right of the panel (L_kk unit lower);
// workgroup = one tile, thread = tile column
__global__ __launch_bounds__(DNB) void lu_trsm_row_kernel(double* __restrict__ A, int64_t N, int k) {
    __shared__ double Lk[DNB][DNB + 1];
    __shared__ double X[DNB][DNB + 1];
    const int j = k + 1 + blockIdx.x;
    const int t = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    double* tj = A + (int64_t)j * DNB * N + (int64_t)k * DNB;
    for (int c = 0; c < DNB; ++c) Lk[t][c] = dk[(int64_t)c * N + t];
    for (int c = 0; c < DNB; ++c) X[t][c] = tj[(int64_t)c * N + t];      // X[row][col]
    __syncthreads();
    // thread t solves column t: x_r -= sum_{m<r} L[r][m] x_m
    for (int r = 1; r < DNB; ++r) {
        double s = 0.0;
        for (int m = 0; m < r; ++m) s += Lk[r][m] * X[m][t];
        X[r][t] -= s;
    }
    __syncthreads();
    for (int c = 0; c < DNB; ++c) tj[(int64_t)c * N + t] = X[t][c];
}

// y = P y: the row swaps of the factorisation in order (one thread per rhs)
__global__ void lu_permute_kernel(int64_t N, int nrhs, const int64_t* __restrict__ piv, double* __restrict__ y) {
    const int h = threadIdx.x;
    if (h >= nrhs) return;
    double* v = y + h * N;
    for (int64_t j = 0; j < N; ++j) {
        const int64_t p = piv[j];
        if (p != j) {
            const double x = v[j];
            v[j] = v[p];
            v[p] = x;
        }
    }
}

// forward with the unit lower L_kk (lane m holds y_k[m])
__global__ __launch_bounds__(64) void lu_fwd_diag_kernel(const double* __restrict__ A, int64_t N, int k, int nrhs,
                                                         double* __restrict__ y) {
    const int m = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    for (int h = 0; h < nrhs; ++h) {
        double* yk = y + h * N + (int64_t)k * DNB;
        double ym = yk[m];
        for (int r = 1; r < DNB; ++r) {
            const double lrm = m < r ? dk[(int64_t)m * N + r] : 0.0;
            const double s = dense_wave_sum(lrm * ym);
            if (m == r) ym -= s;
        }
        yk[m] = ym;
    }
}

// backward with the upper U_kk (lane m holds x_k[m])
__global__ __launch_bounds__(64) void lu_bwd_diag_kernel(const double* __restrict__ A, int64_t N, int k, int nrhs,
                                                         double* __restrict__ y) {
    const int m = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    for (int h = 0; h < nrhs; ++h) {
        double* yk = y + h * N + (int64_t)k * DNB;
        double xm = yk[m];
        for (int r = DNB - 1; r >= 0; --r) {
            const double urm = m > r ? dk[(int64_t)m * N + r] : 0.0;   // U[r][m]
            const double s = dense_wave_sum(urm * xm);
            const double xr = (__shfl(xm, r, 64) - s) / dk[(int64_t)r * N + r];
            if (m == r) xm = xr;
        }
        yk[m] = xm;
    }
}

// backward: y_R -= sum_m U[R][kNB + m] x_k[m] for the rows R above tile k (thread = row)
__global__ __launch_bounds__(256) void lu_bwd_update_kernel(const double* __restrict__ A, int64_t N, int k, int nrhs,
                                                            double* __restrict__ y) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= (int64_t)k * DNB) return;
    for (int h = 0; h < nrhs; ++h) {
        const double* xk = y + h * N + (int64_t)k * DNB;
        double s = 0.0;
        for (int m = 0; m < DNB; ++m) s += A[((int64_t)k * DNB + m) * N + R] * xk[m];
        y[h * N + R] -= s;
    }
}


// y (N x nrhs, rhs-major) from f ([dof][rhs]): zero at constrained DOFs and padding
__global__ __launch_bounds__(256) void dense_rhs_kernel(int64_t n, int64_t N, int nrhs,
                                                        const uint8_t* __restrict__ fixed,
                                                        const double* __restrict__ f, double* __restrict__ y) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= N) return;
    for (int h = 0; h < nrhs; ++h) y[h * N + R] = (R < n && !fixed[R]) ? f[R * nrhs + h] : 0.0;
}

__global__ __launch_bounds__(256) void dense_out_kernel(int64_t n, int64_t N, int nrhs,
                                                        const double* __restrict__ y, double* __restrict__ u) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= n) return;
    for (int h = 0; h < nrhs; ++h) u[R * nrhs + h] = y[h * N + R];
}

// forward: y_k = L_kk^-1 y_k (one wavefront; lane m holds y_k[m])
__global__ __launch_bounds__(64) void dense_fwd_diag_kernel(const double* __restrict__ A, int64_t N, int k, int nrhs,
                                                            double* __restrict__ y) {
    const int m = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    for (int h = 0; h < nrhs; ++h) {
        double* yk = y + h * N + (int64_t)k * DNB;
        double ym = yk[m];
        for (int r = 0; r < DNB; ++r) {
            const double lrm = m < r ? dk[(int64_t)m * N + r] : 0.0;
            const double s = dense_wave_sum(lrm * ym);
            const double yr = (__shfl(ym, r, 64) - s) / dk[(int64_t)r * N + r];
            if (m == r) ym = yr;
        }
        yk[m] = ym;
    }
}

// forward: y_R -= sum_m L[R][kNB + m] y_k[m] for the rows R below tile k
__global__ __launch_bounds__(256) void dense_fwd_update_kernel(const double* __restrict__ A, int64_t N, int k,
                                                               int nrhs, double* __restrict__ y) {
    const int64_t R = (int64_t)(k + 1) * DNB + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= N) return;
    for (int h = 0; h < nrhs; ++h) {
        const double* yk = y + h * N + (int64_t)k * DNB;
        double s = 0.0;
        for (int m = 0; m < DNB; ++m) s += A[((int64_t)k * DNB + m) * N + R] * yk[m];
        y[h * N + R] -= s;
    }
}

// backward: x_k = L_kk^-T y_k (one wavefront; lane m holds x_k[m])
__global__ __launch_bounds__(64) void dense_bwd_diag_kernel(const double* __restrict__ A, int64_t N, int k, int nrhs,
                                                            double* __restrict__ y) {
    const int m = threadIdx.x;
    const double* dk = A + (int64_t)k * DNB * N + (int64_t)k * DNB;
    for (int h = 0; h < nrhs; ++h) {
        double* yk = y + h * N + (int64_t)k * DNB;
        double xm = yk[m];
        for (int r = DNB - 1; r >= 0; --r) {
            const double lmr = m > r ? dk[(int64_t)r * N + m] : 0.0;   // L[m][r]
            const double s = dense_wave_sum(lmr * xm);
            const double xr = (__shfl(xm, r, 64) - s) / dk[(int64_t)r * N + r];
            if (m == r) xm = xr;
        }
        yk[m] = xm;
    }
}

// backward: y_R -= sum_m L[kNB + m][R] x_k[m] for the rows R above tile k; one
// wavefront per row (lane m reads column R at row kNB + m)
__global__ __launch_bounds__(256) void dense_bwd_update_kernel(const double* __restrict__ A, int64_t N, int k,
                                                               int nrhs, double* __restrict__ y) {
    const int64_t R = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int m = threadIdx.x % 64;
    if (R >= (int64_t)k * DNB) return;
    const double l = A[R * N + (int64_t)k * DNB + m];
    for (int h = 0; h < nrhs; ++h) {
        const double s = dense_wave_sum(l * y[h * N + (int64_t)k * DNB + m]);
        if (m == 0) y[h * N + R] -= s;
    }
}

int dense_solve(Plan* p, const double* K, const double* f, double* u, int nrhs, bool lu, bool transpose,
                hipStream_t s) {
    const int64_t n = 6 * (int64_t)p->n_nodes;
    if (n == 0) return JSSO_OK;
    const int nt = (int)((n + DNB - 1) / DNB);
    const int64_t N = (int64_t)nt * DNB;
    constexpr int MAXPARTS = 256;
    const size_t need = (size_t)(N * N + N * nrhs + MAXPARTS) * sizeof(double) +
                        (size_t)(N + MAXPARTS + 1) * sizeof(int64_t);
    if (p->dense_bytes < need) {
        if (p->d_dense) JSSO_HIP(hipFree(p->d_dense));
        p->d_dense = nullptr;
        p->dense_bytes = 0;
        hipError_t e = hipMalloc((void**)&p->d_dense, need);
        JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "dense solve needs %zu bytes (N = %lld): %s", need,
                     (long long)N, hipGetErrorString(e));
        p->dense_bytes = need;
    }
    double* A = p->d_dense;
    double* y = A + N * N;
    double* pval = y + N * nrhs;
    int64_t* piv = (int64_t*)(pval + MAXPARTS);
    int64_t* pidx = piv + N;
    int32_t* dstat = (int32_t*)(pidx + MAXPARTS);   // the factorisation's own flag (the plan's status
                                                    // word keeps the assembly's flags)
    JSSO_HIP(hipMemsetAsync(A, 0, (size_t)N * N * sizeof(double), s));
    JSSO_HIP(hipMemsetAsync(dstat, 0, sizeof(int32_t), s));
    const int64_t tot = p->nnzb * 36;
    hipLaunchKernelGGL(dense_fill_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, p->nnzb, p->d_browof,
                       p->d_colind, p->d_fixed, K, A, N, lu, lu && transpose);
    if (N > n)
        hipLaunchKernelGGL(dense_pad_kernel, dim3((unsigned)((N - n + 255) / 256)), dim3(256), 0, s, n, N, A);
    for (int k = 0; k < nt; ++k) {
        const int m = nt - k - 1;
        if (lu) {
            const int64_t c0 = (int64_t)k * DNB;
            for (int64_t j = c0; j < c0 + DNB; ++j) {
                const int parts = (int)std::min<int64_t>(MAXPARTS, (N - j + 255) / 256);
                hipLaunchKernelGGL(lu_pivot_search_kernel, dim3(parts), dim3(256), 0, s, A, N, j, pval, pidx);
                hipLaunchKernelGGL(lu_pivot_apply_kernel, dim3(1), dim3(DNB), 0, s, A, N, j, c0, parts, pval, pidx, piv,
                                   dstat);
                if (j + 1 < N)
                    hipLaunchKernelGGL(lu_panel_update_kernel, dim3((unsigned)((N - j - 1 + 255) / 256)), dim3(256), 0,
                                       s, A, N, j, c0);
            }
            if (nt > 1)
                hipLaunchKernelGGL(lu_swap_rest_kernel, dim3((unsigned)((N - DNB + 255) / 256)), dim3(256), 0, s, A, N,
                                   c0, piv);
            if (m > 0) {
                hipLaunchKernelGGL(lu_trsm_row_kernel, dim3(m), dim3(DNB), 0, s, A, N, k);
                hipLaunchKernelGGL(dense_update_kernel<true>, dim3((unsigned)(m * m)), dim3(256), 0, s, A, N, k, m);
            }
        } else {
            hipLaunchKernelGGL(dense_potrf_kernel, dim3(1), dim3(256), 0, s, A, N, k, dstat);
            if (m > 0) {
                hipLaunchKernelGGL(dense_trsm_kernel, dim3(m), dim3(DNB), 0, s, A, N, k);
                hipLaunchKernelGGL(dense_update_kernel<false>, dim3((unsigned)(m * (m + 1) / 2)), dim3(256), 0, s, A,
                                   N, k, m);
            }
        }
        JSSO_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(dense_rhs_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, n, N, nrhs, p->d_fixed,
                       f, y);
    if (lu) hipLaunchKernelGGL(lu_permute_kernel, dim3(1), dim3(64), 0, s, N, nrhs, piv, y);
    for (int k = 0; k < nt; ++k) {
        if (lu)
            hipLaunchKernelGGL(lu_fwd_diag_kernel, dim3(1), dim3(64), 0, s, A, N, k, nrhs, y);
        else
            hipLaunchKernelGGL(dense_fwd_diag_kernel, dim3(1), dim3(64), 0, s, A, N, k, nrhs, y);
        const int64_t below = N - (int64_t)(k + 1) * DNB;
        if (below > 0)
            hipLaunchKernelGGL(dense_fwd_update_kernel, dim3((unsigned)((below + 255) / 256)), dim3(256), 0, s, A, N,
                               k, nrhs, y);
    }
    for (int k = nt - 1; k >= 0; --k) {
        const int64_t above = (int64_t)k * DNB;
        if (lu) {
            hipLaunchKernelGGL(lu_bwd_diag_kernel, dim3(1), dim3(64), 0, s, A, N, k, nrhs, y);
            if (above > 0)
                hipLaunchKernelGGL(lu_bwd_update_kernel, dim3((unsigned)((above + 255) / 256)), dim3(256), 0, s, A, N,
                                   k, nrhs, y);
        } else {
            hipLaunchKernelGGL(dense_bwd_diag_kernel, dim3(1), dim3(64), 0, s, A, N, k, nrhs, y);
            if (above > 0)
                hipLaunchKernelGGL(dense_bwd_update_kernel, dim3((unsigned)((above + 3) / 4)), dim3(256), 0, s, A, N,
                                   k, nrhs, y);
        }
    }
    hipLaunchKernelGGL(dense_out_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, N, nrhs, y, u);
    JSSO_HIP(hipGetLastError());
    int32_t st = 0;
    JSSO_HIP(hipMemcpyAsync(&st, dstat, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipStreamSynchronize(s));
    JSSO_REQUIRE(st == 0, JSSO_ESINGULAR,
                 lu ? "dense solve: the constrained stiffness is singular (zero pivot)"
                    : "dense solve: the constrained stiffness is not positive definite");
    return JSSO_OK;
}

}  // namespace jsso
