// Conjugate-gradient solvers on the 6x6 block-CSR stiffness.
//
// Batched path (JSSO_PREC_FP64 / JSSO_PREC_MIXED).  Every structure x load
// case of a plan is one CG system; one launch advances all of them.  The
// system is the block-Jacobi-scaled one, A = L^-1 Kc L^-T with unit diagonal
// blocks (split block-Jacobi preconditioning, identical iterates to PCG with
// M = blockdiag(Kc_ii)^-1), so a step needs no preconditioner pass and the
// diagonal blocks are never read.  Two kernels per iteration:
//     cg1:  s = A r; p = r + beta p; q = s + beta q; <p,q>   -> alpha
//     cg2:  x += alpha p; r -= alpha q; <r,r>                 -> beta
// (q = A p is carried by recurrence).  Recurrences run in VT (FP32 for MIXED);
// every `replace_every` iterations the FP32 partial solution is folded into
// the FP64 solution and the residual is replaced by the true FP64 residual
// b - A64 x64 (reliable updates, Sleijpen & van der Vorst 1996), which is also
// where convergence is decided.
//
// Reductions are deterministic: each workgroup writes its partial sums; the
// last workgroup of a system (atomic arrival counter) adds them in a fixed
// order and updates the system's scalars.
//
// Safeguards: the iterate with the smallest true residual is kept and is what
// the solve returns; a system stops when its true residual has not improved
// by 10% for max(2048, it/2) iterations (the attainable-accuracy floor of the
// arithmetic), grows 1000-fold past its best, or is not finite; a
// non-positive or non-finite curvature <p,q> restarts the search direction.
//
// Plain path (JSSO_PREC_PLAIN): textbook block-Jacobi PCG on the constrained
// FP64 K, one structure and one load case per launch sequence.
#include "jsso_internal.hpp"

namespace jsso {

constexpr int NWAVES = SOLVER_WG / 64;
enum { SYS_RUNNING = 0, SYS_CONVERGED = 1, SYS_STALLED = 2 };

// ---------------------------------------------------------------------------
// sampled event timing of cg1 (jsso_timing_enable / jsso_timing_read)
// ---------------------------------------------------------------------------
struct KernelTiming {
    int every = 0;
    int64_t counter = 0, samples = 0;
    double total_ms = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool, pending;
};
static KernelTiming g_timing;

static void timing_begin(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev) {
    ev = nullptr;
    if (g_timing.every <= 0 || (g_timing.counter++ % g_timing.every) != 0) return;
    std::pair<hipEvent_t, hipEvent_t> e;
    if (!g_timing.pool.empty()) {
        e = g_timing.pool.back();
        g_timing.pool.pop_back();
    } else if (hipEventCreate(&e.first) != hipSuccess || hipEventCreate(&e.second) != hipSuccess) {
        return;
    }
    g_timing.pending.push_back(e);
    ev = &g_timing.pending.back();
    (void)hipEventRecord(ev->first, s);
}

static void timing_end(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>* ev) {
    if (ev) (void)hipEventRecord(ev->second, s);
}

// after a stream synchronisation: fold the completed samples
static void timing_collect() {
    for (auto& e : g_timing.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
            g_timing.total_ms += ms;
            g_timing.samples++;
        }
        g_timing.pool.push_back(e);
    }
    g_timing.pending.clear();
}

int timing_enable(int every) {
    g_timing.every = every > 0 ? every : 0;
    g_timing.counter = 0;
    return JSSO_OK;
}

int timing_read(double* avg_ms, int64_t* samples, int reset) {
    if (avg_ms) *avg_ms = g_timing.samples ? g_timing.total_ms / g_timing.samples : 0.0;
    if (samples) *samples = g_timing.samples;
    if (reset) {
        g_timing.total_ms = 0.0;
        g_timing.samples = 0;
    }
    return JSSO_OK;
}

struct SysState {
    double *bnorm2, *rr, *rr_prev, *alpha, *beta, *relres, *best;
    int32_t *active, *counter, *status, *stall, *restart, *save;
    double* partial;
};

struct Rows {
    const int32_t *wg_sys, *wg_row0, *wg_row1, *sys_wg_off;
    int wg_base;  // plain path: launches cover one system's workgroups
};

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) lds[wave * NV + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < NWAVES; ++w) s += lds[w * NV + i];
            v[i] = s;
        }
    }
}

// Writes this workgroup's NV partials; returns true (in every thread) for the
// last workgroup of system `sys`, in which `tot` then holds the system totals
// (valid in thread 0).
template <int NV>
__device__ bool reduce_system(double (&v)[NV], double (&tot)[NV], double* lds, int wg, int sys,
                              const Rows& rows, const SysState& st) {
    __shared__ int am_last;
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) st.partial[(int64_t)wg * NV + i] = v[i];
        __threadfence();
        const int w0 = rows.sys_wg_off[sys], w1 = rows.sys_wg_off[sys + 1];
        const int prev = atomicAdd(&st.counter[sys], 1);
        am_last = (prev == w1 - w0 - 1);
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    const int w0 = rows.sys_wg_off[sys], w1 = rows.sys_wg_off[sys + 1];
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = 0.0;
    for (int w = w0 + (int)threadIdx.x; w < w1; w += blockDim.x) {
#pragma unroll
        for (int i = 0; i < NV; ++i) s[i] += __hip_atomic_load(&st.partial[(int64_t)w * NV + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(s, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) tot[i] = s[i];
        st.counter[sys] = 0;
    }
    return true;
}

// ---------------------------------------------------------------------------
// batched kernels
// ---------------------------------------------------------------------------

// b = L^-1 f (constrained DOFs zeroed), x = 0, r = b.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_init(Rows rows, SysState st, const uint8_t* __restrict__ fixed,
                                                    const double* __restrict__ Linv, const double* __restrict__ f,
                                                    double* __restrict__ b64, double* __restrict__ x64,
                                                    double* __restrict__ r64, VT* __restrict__ xlo,
                                                    VT* __restrict__ rlo, VT* __restrict__ p, VT* __restrict__ q,
                                                    double* __restrict__ xbest) {
    __shared__ double lds[NWAVES * 2 * NR];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[2 * NR];
#pragma unroll
    for (int i = 0; i < 2 * NR; ++i) v[i] = 0.0;
    if (on) {
        const int64_t i = R / 6;
        const int r = (int)(R % 6);
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const int64_t dof = 6 * i + c;
                const double fv = fixed[dof] ? 0.0 : f[dof * NR + h];
                s += Linv[i * 36 + r * 6 + c] * fv;
            }
            const int64_t o = R * NR + h;
            b64[o] = s;
            x64[o] = 0.0;
            xbest[o] = 0.0;
            r64[o] = s;
            xlo[o] = VT(0);
            rlo[o] = VT(s);
            p[o] = VT(0);
            q[o] = VT(0);
            v[h] = s * s;
            const double sl = (double)VT(s);
            v[NR + h] = sl * sl;
        }
    }
    double tot[2 * NR];
    if (reduce_system<2 * NR>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int k = sys * NR + h;
            st.bnorm2[k] = tot[h];
            st.rr[k] = tot[NR + h];
            st.rr_prev[k] = tot[NR + h];
            st.beta[k] = 0.0;
            st.alpha[k] = 0.0;
            st.active[k] = tot[h] > 0.0 ? 1 : 0;
            st.relres[k] = tot[h] > 0.0 ? 1.0 : 0.0;
            st.best[k] = st.relres[k];
            st.status[k] = tot[h] > 0.0 ? SYS_RUNNING : SYS_CONVERGED;
            st.stall[k] = 0;
            st.restart[k] = 0;
            st.save[k] = 0;
        }
    }
}

// s = A r (A = I + off-diagonal scaled blocks); p = r + beta p; q = s + beta q; <p,q>.
template <typename MT, typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_cg1(Rows rows, SysState st, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const MT* __restrict__ A,
                                                   const VT* __restrict__ rlo, VT* __restrict__ p,
                                                   VT* __restrict__ q) {
    __shared__ double lds[NWAVES * NR];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = 0.0;
    if (on) {
        const int64_t i = R / 6;
        const int r = (int)(R % 6);
        VT acc[NR];
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = rlo[R * NR + h];
        const int k0 = rowptr[i], k1 = rowptr[i + 1];
        for (int k = k0; k < k1; ++k) {
            const MT* a = A + (int64_t)k * 36 + r * 6;
            const VT* xj = rlo + (int64_t)6 * col[k] * NR;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const VT av = (VT)a[c];
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[h] += av * xj[c * NR + h];
            }
        }
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int ks = sys * NR + h;
            if (!st.active[ks]) continue;
            const VT beta = (VT)st.beta[ks];
            const int64_t o = R * NR + h;
            const VT pn = rlo[o] + beta * p[o];
            const VT qn = acc[h] + beta * q[o];
            p[o] = pn;
            q[o] = qn;
            v[h] = (double)pn * (double)qn;
        }
    }
    double tot[NR];
    if (reduce_system<NR>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int k = sys * NR + h;
            if (!st.active[k]) continue;
            if (tot[h] > 0.0 && isfinite(tot[h])) {
                st.alpha[k] = st.rr[k] / tot[h];
            } else {
                st.alpha[k] = 0.0;  // breakdown: no step, restart the direction
                st.restart[k] = 1;
            }
        }
    }
}

// x += alpha p; r -= alpha q; <r,r> -> beta.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_cg2(Rows rows, SysState st, const VT* __restrict__ p,
                                                   const VT* __restrict__ q, VT* __restrict__ xlo,
                                                   VT* __restrict__ rlo) {
    __shared__ double lds[NWAVES * NR];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = 0.0;
    if (on) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int ks = sys * NR + h;
            if (!st.active[ks]) continue;
            const VT alpha = (VT)st.alpha[ks];
            const int64_t o = R * NR + h;
            xlo[o] += alpha * p[o];
            const VT rn = rlo[o] - alpha * q[o];
            rlo[o] = rn;
            v[h] = (double)rn * (double)rn;
        }
    }
    double tot[NR];
    if (reduce_system<NR>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int k = sys * NR + h;
            if (!st.active[k]) continue;
            st.rr_prev[k] = st.rr[k];
            st.rr[k] = tot[h];
            st.beta[k] = (st.rr_prev[k] > 0.0 && !st.restart[k]) ? tot[h] / st.rr_prev[k] : 0.0;
            st.restart[k] = 0;
        }
    }
}

// x64 += xlo; xlo = 0.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_fold(Rows rows, const int32_t* __restrict__ active,
                                                    VT* __restrict__ xlo, double* __restrict__ x64) {
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    if (R >= 6 * (int64_t)rows.wg_row1[wg]) return;
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        if (!active[sys * NR + h]) continue;
        const int64_t o = R * NR + h;
        x64[o] += (double)xlo[o];
        xlo[o] = VT(0);
    }
}

// r64 = b - A64 x64; rlo = r64; <r64,r64>, <rlo,rlo>; convergence test.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_replace(Rows rows, SysState st, const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const double* __restrict__ A, const double* __restrict__ b64,
                                                       const double* __restrict__ x64, double* __restrict__ r64,
                                                       VT* __restrict__ rlo, double tol, int it) {
    __shared__ double lds[NWAVES * 2 * NR];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[2 * NR];
#pragma unroll
    for (int i = 0; i < 2 * NR; ++i) v[i] = 0.0;
    if (on) {
        const int64_t i = R / 6;
        const int r = (int)(R % 6);
        double acc[NR];
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = x64[R * NR + h];
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const double* a = A + (int64_t)k * 36 + r * 6;
            const double* xj = x64 + (int64_t)6 * col[k] * NR;
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[h] += a[c] * xj[c * NR + h];
        }
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            if (!st.active[sys * NR + h]) continue;
            const int64_t o = R * NR + h;
            const double rv = b64[o] - acc[h];
            r64[o] = rv;
            const VT rl = VT(rv);
            rlo[o] = rl;
            v[h] = rv * rv;
            v[NR + h] = (double)rl * (double)rl;
        }
    }
    double tot[2 * NR];
    if (reduce_system<2 * NR>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int k = sys * NR + h;
            if (!st.active[k]) continue;
            const double rel = sqrt(tot[h] / st.bnorm2[k]);
            st.relres[k] = rel;
            st.beta[k] = st.rr_prev[k] > 0.0 ? tot[NR + h] / st.rr_prev[k] : 0.0;
            st.rr[k] = tot[NR + h];
            const double best = st.best[k];
            const bool improved = rel < best;
            st.save[k] = improved ? 1 : 0;
            if (improved) st.best[k] = rel;
            if (rel <= 0.9 * best) st.stall[k] = it;   // iteration of the last 10% gain
            const int window = it / 2 > 2048 ? it / 2 : 2048;
            if (rel <= tol || tot[NR + h] == 0.0) {
                st.active[k] = 0;
                st.status[k] = SYS_CONVERGED;
            } else if (!isfinite(rel) || rel > 1e3 * st.best[k] || it - st.stall[k] > window) {
                st.active[k] = 0;
                st.status[k] = SYS_STALLED;
            }
        }
    }
}

// keep the best iterate: xbest = x64 where the last replacement improved it
template <int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_save(Rows rows, const int32_t* __restrict__ save,
                                                    const double* __restrict__ x64, double* __restrict__ xbest) {
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    if (R >= 6 * (int64_t)rows.wg_row1[wg]) return;
#pragma unroll
    for (int h = 0; h < NR; ++h)
        if (save[sys * NR + h]) xbest[R * NR + h] = x64[R * NR + h];
}

// u = L^-T x
template <int NR>
__global__ __launch_bounds__(256) void k_unscale(int32_t n_nodes, const double* __restrict__ Linv,
                                                 const double* __restrict__ x64, double* __restrict__ u) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= 6 * (int64_t)n_nodes) return;
    const int64_t i = R / 6;
    const int r = (int)(R % 6);
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s += Linv[i * 36 + c * 6 + r] * x64[(6 * i + c) * NR + h];
        u[R * NR + h] = s;
    }
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
struct Workspace {
    double *b64, *x64, *r64, *xbest;
    void *xlo, *rlo, *p, *q;
    SysState st;
};

static int get_workspace(Plan* pl, int nrhs, Workspace& w) {
    const size_t n = 6 * (size_t)pl->n_nodes * nrhs;
    const size_t nsys = (size_t)pl->n_sys * nrhs;
    const size_t need[] = {n * 8, n * 8, n * 8, n * 8, n * 8, n * 8, n * 8,
                           (size_t)pl->n_wg * 4 * nrhs * 8, nsys * 8 * 7, nsys * 4 * 5, (size_t)pl->n_sys * 4,
                           0, n * 8};
    const int nb = sizeof(need) / sizeof(need[0]);
    if (pl->ws.size() < 13) pl->ws.resize(13);  // slot 11: adjoint scratch
    for (int i = 0; i < nb; ++i) {
        if (i == 11) continue;
        if (pl->ws[i].bytes < need[i]) {
            if (pl->ws[i].p) JSSO_HIP(hipFree(pl->ws[i].p));
            pl->ws[i].p = nullptr;
            pl->ws[i].bytes = 0;
            hipError_t e = hipMalloc(&pl->ws[i].p, need[i]);
            JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "workspace hipMalloc(%zu): %s", need[i],
                         hipGetErrorString(e));
            pl->ws[i].bytes = need[i];
            if (i == 10) JSSO_HIP(hipMemset(pl->ws[i].p, 0, need[i]));
        }
    }
    w.b64 = (double*)pl->ws[0].p;
    w.x64 = (double*)pl->ws[1].p;
    w.r64 = (double*)pl->ws[2].p;
    w.xlo = pl->ws[3].p;
    w.rlo = pl->ws[4].p;
    w.p = pl->ws[5].p;
    w.q = pl->ws[6].p;
    w.st.partial = (double*)pl->ws[7].p;
    double* sc = (double*)pl->ws[8].p;
    w.st.bnorm2 = sc;
    w.st.rr = sc + nsys;
    w.st.rr_prev = sc + 2 * nsys;
    w.st.alpha = sc + 3 * nsys;
    w.st.beta = sc + 4 * nsys;
    w.st.relres = sc + 5 * nsys;
    w.st.best = sc + 6 * nsys;
    int32_t* ic = (int32_t*)pl->ws[9].p;
    w.st.active = ic;
    w.st.status = ic + nsys;
    w.st.stall = ic + 2 * nsys;
    w.st.restart = ic + 3 * nsys;
    w.st.save = ic + 4 * nsys;
    w.st.counter = (int32_t*)pl->ws[10].p;
    w.xbest = (double*)pl->ws[12].p;
    return JSSO_OK;
}

static int check_status(Plan* p, hipStream_t s) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    int32_t status = 0;
    JSSO_HIP(hipMemcpyAsync(&status, p->d_status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipStreamSynchronize(s));
    JSSO_REQUIRE(!(status & 1), JSSO_ESINGULAR,
                 "a constrained diagonal block of K is not positive definite (unsupported or "
                 "isolated node)");
    JSSO_REQUIRE(!(status & 2), JSSO_EUNSUPPORTED,
                 "a beam-column is parallel to global Y: the reference's transformation "
                 "(element.py:92-94) is not orthogonal and K is not symmetric; CG cannot solve it");
    return JSSO_OK;
}

template <typename MT, typename VT, int NR>
static int run_batched(Plan* pl, const double* f, double* u, double tol, int maxit, int replace_every,
                       jsso_info* info, hipStream_t s) {
    Workspace w;
    int rc = get_workspace(pl, NR, w);
    if (rc) return rc;
    Rows rows{pl->d_wg_sys, pl->d_wg_row0, pl->d_wg_row1, pl->d_sys_wg_off, 0};
    const dim3 grid(pl->n_wg), block(SOLVER_WG);
    const MT* A = nullptr;
    if constexpr (sizeof(MT) == 4) A = (const MT*)pl->d_Ks32; else A = (const MT*)pl->d_Ks64;
    VT *xlo = (VT*)w.xlo, *rlo = (VT*)w.rlo, *p = (VT*)w.p, *q = (VT*)w.q;
    hipLaunchKernelGGL((k_init<VT, NR>), grid, block, 0, s, rows, w.st, pl->d_fixed, pl->d_Linv, f, w.b64,
                       w.x64, w.r64, xlo, rlo, p, q, w.xbest);
    JSSO_HIP(hipGetLastError());
    const int nsys = pl->n_sys * NR;
    std::vector<int32_t> act(nsys);
    int it = 0, repl = 0;
    bool done = false;
    while (!done && it < maxit) {
        const int steps = std::min(replace_every, maxit - it);
        for (int j = 0; j < steps; ++j) {
            std::pair<hipEvent_t, hipEvent_t>* ev;
            timing_begin(s, ev);
            hipLaunchKernelGGL((k_cg1<MT, VT, NR>), grid, block, 0, s, rows, w.st, pl->d_offrowptr,
                               pl->d_offcol, A, (const VT*)rlo, p, q);
            timing_end(s, ev);
            hipLaunchKernelGGL((k_cg2<VT, NR>), grid, block, 0, s, rows, w.st, (const VT*)p, (const VT*)q,
                               xlo, rlo);
        }
        it += steps;
        hipLaunchKernelGGL((k_fold<VT, NR>), grid, block, 0, s, rows, (const int32_t*)w.st.active, xlo, w.x64);
        hipLaunchKernelGGL((k_replace<VT, NR>), grid, block, 0, s, rows, w.st, pl->d_offrowptr, pl->d_offcol,
                           (const double*)pl->d_Ks64, (const double*)w.b64, (const double*)w.x64, w.r64, rlo,
                           tol, it);
        hipLaunchKernelGGL((k_save<NR>), grid, block, 0, s, rows, (const int32_t*)w.st.save,
                           (const double*)w.x64, w.xbest);
        JSSO_HIP(hipGetLastError());
        ++repl;
        JSSO_HIP(hipMemcpyAsync(act.data(), w.st.active, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        JSSO_HIP(hipStreamSynchronize(s));
        timing_collect();
        done = true;
        for (int k = 0; k < nsys; ++k)
            if (act[k]) { done = false; break; }
    }
    const int64_t nr = 6 * (int64_t)pl->n_nodes;
    hipLaunchKernelGGL((k_unscale<NR>), dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, pl->n_nodes,
                       pl->d_Linv, (const double*)w.xbest, u);
    JSSO_HIP(hipGetLastError());
    std::vector<double> rel(nsys);
    JSSO_HIP(hipMemcpyAsync(rel.data(), w.st.best, nsys * sizeof(double), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipMemcpyAsync(act.data(), w.st.status, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipStreamSynchronize(s));
    int conv = 0;
    double mx = 0.0;
    for (int k = 0; k < nsys; ++k) {
        conv += act[k] == SYS_CONVERGED ? 1 : 0;
        mx = std::max(mx, rel[k]);
    }
    if (info) {
        info->iterations = it;
        info->replacements = repl;
        info->converged = conv;
        info->n_systems = nsys;
        info->rel_residual = mx;
    }
    return conv == nsys ? JSSO_OK : JSSO_NOT_CONVERGED;
}

template <typename MT, typename VT>
static int dispatch_nr(Plan* pl, const double* f, double* u, int nrhs, double tol, int maxit, int every,
                       jsso_info* info, hipStream_t s) {
    switch (nrhs) {
        case 1: return run_batched<MT, VT, 1>(pl, f, u, tol, maxit, every, info, s);
        case 2: return run_batched<MT, VT, 2>(pl, f, u, tol, maxit, every, info, s);
        case 3: return run_batched<MT, VT, 3>(pl, f, u, tol, maxit, every, info, s);
        case 4: return run_batched<MT, VT, 4>(pl, f, u, tol, maxit, every, info, s);
        case 8: return run_batched<MT, VT, 8>(pl, f, u, tol, maxit, every, info, s);
        default:
            set_error("nrhs=%d unsupported (1, 2, 3, 4 or 8)", nrhs);
            return JSSO_EINVAL;
    }
}

int solve_batched(Plan* p, const double* f, double* u, int nrhs, bool mixed, double tol, int maxit,
                  jsso_info* info, hipStream_t s) {
    int rc = check_status(p, s);
    if (rc) return rc;
    if (mixed) return dispatch_nr<float, float>(p, f, u, nrhs, tol, maxit, 64, info, s);
    return dispatch_nr<double, double>(p, f, u, nrhs, tol, maxit, 256, info, s);
}

// ---------------------------------------------------------------------------
// plain path: block-Jacobi PCG on Kc (FP64), one system per launch sequence
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SOLVER_WG) void kp_init(Rows rows, SysState st, int sysk, int h, int nrhs,
                                                     const uint8_t* __restrict__ fixed,
                                                     const double* __restrict__ Minv, const double* __restrict__ f,
                                                     double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ z, double* __restrict__ pp) {
    __shared__ double lds[NWAVES];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[1] = {0.0};
    if (on) {
        const int64_t i = R / 6;
        const int rr = (int)(R % 6);
        double zs = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const int64_t dof = 6 * i + c;
            const double fv = fixed[dof] ? 0.0 : f[dof * nrhs + h];
            zs += Minv[i * 36 + rr * 6 + c] * fv;
        }
        const double rv = fixed[R] ? 0.0 : f[R * nrhs + h];
        x[R] = 0.0;
        r[R] = rv;
        z[R] = zs;
        pp[R] = zs;
        v[0] = rv * zs;
    }
    double tot[1];
    if (reduce_system<1>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
        st.bnorm2[sysk] = tot[0];
        st.rr[sysk] = tot[0];
        st.active[sysk] = tot[0] > 0.0 ? 1 : 0;
        st.relres[sysk] = tot[0] > 0.0 ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(SOLVER_WG) void kp_spmv(Rows rows, SysState st, int sysk,
                                                     const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ K,
                                                     const double* __restrict__ pp, double* __restrict__ q) {
    __shared__ double lds[NWAVES];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    double v[1] = {0.0};
    if (on && st.active[sysk]) {
        const int64_t i = R / 6;
        const int r = (int)(R % 6);
        double acc = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const double* a = K + (int64_t)k * 36 + r * 6;
            const double* xj = pp + (int64_t)6 * col[k];
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += a[c] * xj[c];
        }
        q[R] = acc;
        v[0] = pp[R] * acc;
    }
    double tot[1];
    if (reduce_system<1>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
        if (st.active[sysk]) st.alpha[sysk] = tot[0] > 0.0 ? st.rr[sysk] / tot[0] : 0.0;
    }
}

__global__ __launch_bounds__(SOLVER_WG) void kp_update(Rows rows, SysState st, int sysk,
                                                       const double* __restrict__ Minv, const double* __restrict__ pp,
                                                       const double* __restrict__ q, double* __restrict__ x,
                                                       double* __restrict__ r, double* __restrict__ z, double tol) {
    __shared__ double lds[NWAVES];
    const int wg = blockIdx.x + rows.wg_base;
    const int sys = rows.wg_sys[wg];
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    const bool on = R < 6 * (int64_t)rows.wg_row1[wg];
    __shared__ double rs[SOLVER_WG];
    double v[1] = {0.0};
    const bool act = st.active[sysk] != 0;
    const double alpha = st.alpha[sysk];
    double rn = 0.0;
    if (on && act) {
        x[R] += alpha * pp[R];
        rn = r[R] - alpha * q[R];
        r[R] = rn;
    }
    rs[threadIdx.x] = rn;
    __syncthreads();
    if (on && act) {
        const int64_t i = R / 6;
        const int rr = (int)(R % 6);
        const int base = (int)threadIdx.x - rr;
        double zs = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) zs += Minv[i * 36 + rr * 6 + c] * rs[base + c];
        z[R] = zs;
        v[0] = rn * zs;
    }
    double tot[1];
    if (reduce_system<1>(v, tot, lds, wg, sys, rows, st) && threadIdx.x == 0) {
        if (act) {
            const double rz_old = st.rr[sysk];
            st.beta[sysk] = rz_old > 0.0 ? tot[0] / rz_old : 0.0;
            st.rr[sysk] = tot[0];
            const double rel = sqrt(fabs(tot[0]) / st.bnorm2[sysk]);
            st.relres[sysk] = rel;
            if (rel <= tol) st.active[sysk] = 0;
        }
    }
}

__global__ __launch_bounds__(SOLVER_WG) void kp_dir(Rows rows, const SysState st, int sysk,
                                                    const double* __restrict__ z, double* __restrict__ pp) {
    const int wg = blockIdx.x + rows.wg_base;
    const int64_t R = 6 * (int64_t)rows.wg_row0[wg] + threadIdx.x;
    if (R >= 6 * (int64_t)rows.wg_row1[wg]) return;
    if (!st.active[sysk]) return;
    pp[R] = z[R] + st.beta[sysk] * pp[R];
}

__global__ void kp_store(int64_t r0, int64_t r1, int nrhs, int h, const double* __restrict__ x,
                         double* __restrict__ u) {
    const int64_t R = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= r1) return;
    u[R * nrhs + h] = x[R];
}

int solve_plain(Plan* pl, const double* f, double* u, int nrhs, double tol, int maxit, jsso_info* info,
                hipStream_t s) {
    int rc = check_status(pl, s);
    if (rc) return rc;
    Workspace w;
    rc = get_workspace(pl, nrhs, w);
    if (rc) return rc;
    double *x = w.x64, *r = w.r64, *z = w.b64, *pp = (double*)w.p, *q = (double*)w.q;
    const int nsys = pl->n_sys * nrhs;
    int itmax = 0, conv = 0;
    double mx = 0.0;
    std::vector<int32_t> sys_wg(pl->h_sys_wg_off);
    for (int sidx = 0; sidx < pl->n_sys; ++sidx) {
        const int w0 = sys_wg[sidx], w1 = sys_wg[sidx + 1];
        if (w1 == w0) continue;
        Rows rows{pl->d_wg_sys, pl->d_wg_row0, pl->d_wg_row1, pl->d_sys_wg_off, w0};
        const dim3 grid(w1 - w0), block(SOLVER_WG);
        for (int h = 0; h < nrhs; ++h) {
            const int sysk = sidx * nrhs + h;
            hipLaunchKernelGGL(kp_init, grid, block, 0, s, rows, w.st, sysk, h, nrhs, pl->d_fixed, pl->d_Minv, f,
                               x, r, z, pp);
            JSSO_HIP(hipGetLastError());
            int it = 0;
            int32_t act = 1;
            double rel = 0.0;
            while (it < maxit) {
                const int steps = std::min(32, maxit - it);
                for (int j = 0; j < steps; ++j) {
                    hipLaunchKernelGGL(kp_spmv, grid, block, 0, s, rows, w.st, sysk, pl->d_rowptr, pl->d_colind,
                                       (const double*)pl->d_Kc, (const double*)pp, q);
                    hipLaunchKernelGGL(kp_update, grid, block, 0, s, rows, w.st, sysk, (const double*)pl->d_Minv,
                                       (const double*)pp, (const double*)q, x, r, z, tol);
                    hipLaunchKernelGGL(kp_dir, grid, block, 0, s, rows, w.st, sysk, (const double*)z, pp);
                }
                it += steps;
                JSSO_HIP(hipGetLastError());
                JSSO_HIP(hipMemcpyAsync(&act, w.st.active + sysk, sizeof(int32_t), hipMemcpyDeviceToHost, s));
                JSSO_HIP(hipStreamSynchronize(s));
                if (!act) break;
            }
            JSSO_HIP(hipMemcpy(&rel, w.st.relres + sysk, sizeof(double), hipMemcpyDeviceToHost));
            const int64_t r0 = 6 * (int64_t)pl->h_node_off[sidx], r1 = 6 * (int64_t)pl->h_node_off[sidx + 1];
            hipLaunchKernelGGL(kp_store, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), 0, s, r0, r1, nrhs, h,
                               (const double*)x, u);
            JSSO_HIP(hipGetLastError());
            itmax = std::max(itmax, it);
            conv += act ? 0 : 1;
            mx = std::max(mx, rel);
        }
    }
    JSSO_HIP(hipStreamSynchronize(s));
    if (info) {
        info->iterations = itmax;
        info->replacements = 0;
        info->converged = conv;
        info->n_systems = nsys;
        info->rel_residual = mx;
    }
    return conv == nsys ? JSSO_OK : JSSO_NOT_CONVERGED;
}

}  // namespace jsso
