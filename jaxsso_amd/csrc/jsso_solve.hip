// Conjugate-gradient solvers on the 6x6 block stiffness.
//
// Batched path (JSSO_PREC_FP64 / JSSO_PREC_MIXED).  Every structure x load
// case of a plan is one CG system; one launch advances all of them.  The
// system is the block-Jacobi-scaled one, A = L^-1 Kc L^-T with unit diagonal
// blocks (split block-Jacobi preconditioning: the same iterates as PCG with
// M = blockdiag(Kc_ii)^-1), so a step needs no preconditioner pass and the
// diagonal blocks are never read.  One kernel (k_cgf) per iteration k:
//     alpha = rr/<p,q>;  r_k = r_{k-1} - alpha q_{k-1}  (own and gathered rows)
//     x += alpha p;  beta = |r_k|^2 / rr  with |r_k|^2 from <r,r>, <r,q>, <q,q>
//     s = A r_k;  p = r_k + beta p;  q = s + beta q;  partials <p,q> <r,q> <q,q> <r,r>
// (q = A p carried by recurrence).  Recurrences run in VT (FP32 for MIXED);
// every `replace_every` iterations the partial solution is folded into the
// FP64 solution and the residual replaced by the true FP64 residual b - A64 x64
// (reliable updates, Sleijpen & van der Vorst 1996), where convergence is
// decided.
//
// Layout.  A slice is <= 64 block rows of one structure.  The off-diagonal
// blocks of a slice are stored as sliced ELL, values [slot][36 entries][64
// lanes]: in the SpMV one lane owns one block row, so each of the 36 loads of
// a slot is one contiguous 256 B (FP32) / 512 B (FP64) wavefront access.
//
// Reductions without atomics or fences.  Each kernel writes one partial sum
// per slice; every workgroup of the next kernel adds the partials of its own
// structure in a fixed order (so all of them obtain bitwise-identical CG
// scalars) -- the kernel boundary is the only synchronisation.  Partials and
// the r, q vectors alternate between two buffers by iteration parity.
//
// Safeguards: the iterate with the smallest true residual is kept and is what
// the solve returns; a system stops when its true residual has not improved by
// 10% for max(8192, it/2) iterations (the attainable-accuracy floor), grows
// 1000-fold past its best, or is not finite; a non-positive or non-finite
// curvature <p,q> restarts the search direction.
//
// Plain path (JSSO_PREC_PLAIN): textbook block-Jacobi PCG on the constrained
// FP64 K in block CSR, thread per scalar row, one structure and one load case
// per launch sequence -- the baseline the batched paths are measured against.
#include <cstdlib>

#include "solver_kernels.hpp"

namespace jsso {


// ---------------------------------------------------------------------------
// sampled event timing (jsso_timing_enable / jsso_timing_read_slot)
// ---------------------------------------------------------------------------
// Slot JSSO_TIMING_SOLVE holds the batched solver's dominant kernel: k_cgf
// (block-Jacobi; one eager launch in `every`, or a captured chunk of launches
// as that many) or k_pcg_dir (AMG; one chunk in `every` is launched eagerly and
// each of its k_pcg_dir launches is bracketed).  The other slots bracket whole
// phases on every call while timing is on.
struct KernelTiming {
    int every = 0;
    int64_t counter = 0, chunk_counter = 0;
    int64_t samples[JSSO_TIMING_SLOTS] = {};
    double total_ms[JSSO_TIMING_SLOTS] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool, pending;
    std::vector<int> pending_launches;   // launches each pending event pair brackets
    std::vector<int> pending_slot;
};
static KernelTiming g_timing;

static void timing_start(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev, int launches, int slot) {
    std::pair<hipEvent_t, hipEvent_t> e;
    if (!g_timing.pool.empty()) {
        e = g_timing.pool.back();
        g_timing.pool.pop_back();
    } else if (hipEventCreate(&e.first) != hipSuccess || hipEventCreate(&e.second) != hipSuccess) {
        return;
    }
    g_timing.pending.push_back(e);
    g_timing.pending_launches.push_back(launches);
    g_timing.pending_slot.push_back(slot);
    ev = &g_timing.pending.back();
    (void)hipEventRecord(ev->first, s);
}

// one launch in `every` is bracketed by events
void timing_begin(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev) {
    ev = nullptr;
    if (g_timing.every <= 0 || (g_timing.counter++ % g_timing.every) != 0) return;
    timing_start(s, ev, 1, JSSO_TIMING_SOLVE);
}

// a graph of `launches` kernels back to back: bracketed whenever timing is on
void timing_begin_batch(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev, int launches) {
    ev = nullptr;
    if (g_timing.every <= 0) return;
    timing_start(s, ev, launches, JSSO_TIMING_SOLVE);
}

// every call while timing is on (phases: element kernels, gathers, adjoint,
// AMG set-up; the launches of a sampled AMG chunk)
void timing_begin_slot(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev, int slot) {
    ev = nullptr;
    if (g_timing.every <= 0) return;
    timing_start(s, ev, 1, slot);
}

// true for one AMG chunk in `every` while timing is on: that chunk is launched
// eagerly so that its k_pcg_dir launches can be bracketed
bool timing_sample_chunk() {
    if (g_timing.every <= 0) return false;
    return (g_timing.chunk_counter++ % g_timing.every) == 0;
}

// An event pair owned by the caller (a deferred sample): recorded around a
// launch now, added to a slot by timing_settle once the caller knows whether the
// sample counts.
std::pair<hipEvent_t, hipEvent_t> timing_pair() {
    std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
    if (!g_timing.pool.empty()) {
        e = g_timing.pool.back();
        g_timing.pool.pop_back();
    } else if (hipEventCreate(&e.first) != hipSuccess || hipEventCreate(&e.second) != hipSuccess) {
        e = {nullptr, nullptr};
    }
    return e;
}

// the events have completed (the caller synchronised): add to `slot` if
// `count`, then return the pair to the pool
void timing_settle(std::pair<hipEvent_t, hipEvent_t> e, int slot, bool count) {
    if (!e.first) return;
    float ms = 0.f;
    if (count && hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
        g_timing.total_ms[slot] += ms;
        g_timing.samples[slot] += 1;
    }
    (void)hipGetLastError();
    g_timing.pool.push_back(e);
}

void timing_end(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>* ev) {
    if (ev) (void)hipEventRecord(ev->second, s);
}

// fold finished samples into the totals; samples still in flight stay pending
void timing_collect() {
    size_t keep = 0;
    for (size_t i = 0; i < g_timing.pending.size(); ++i) {
        auto e = g_timing.pending[i];
        const int n = g_timing.pending_launches[i];
        const int slot = g_timing.pending_slot[i];
        if (hipEventQuery(e.second) == hipErrorNotReady) {
            g_timing.pending[keep] = e;
            g_timing.pending_launches[keep] = n;
            g_timing.pending_slot[keep] = slot;
            ++keep;
            continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
            g_timing.total_ms[slot] += ms;
            g_timing.samples[slot] += n;
        }
        g_timing.pool.push_back(e);
    }
    g_timing.pending.resize(keep);
    g_timing.pending_launches.resize(keep);
    g_timing.pending_slot.resize(keep);
    (void)hipGetLastError();
}

int timing_enable(int every) {
    g_timing.every = every > 0 ? every : 0;
    g_timing.counter = 0;
    g_timing.chunk_counter = 0;
    return JSSO_OK;
}

int timing_read(int slot, double* avg_ms, int64_t* samples, int reset) {
    JSSO_REQUIRE(slot >= 0 && slot < JSSO_TIMING_SLOTS, JSSO_EINVAL, "timing slot %d out of range", slot);
    if (reset) {
        // wait for what is still in flight, so that a reset leaves nothing behind
        for (auto& e : g_timing.pending) (void)hipEventSynchronize(e.second);
    }
    timing_collect();
    if (avg_ms) *avg_ms = g_timing.samples[slot] ? g_timing.total_ms[slot] / g_timing.samples[slot] : 0.0;
    if (samples) *samples = g_timing.samples[slot];
    if (reset) {
        g_timing.total_ms[slot] = 0.0;
        g_timing.samples[slot] = 0;
    }
    return JSSO_OK;
}


// ---------------------------------------------------------------------------
// batched kernels
// ---------------------------------------------------------------------------

enum { CG_START = 0, CG_STEP = 1, CG_FRESH = 2 };

// Fused CG iteration k (one lane per block row, sliced ELL).
//   CG_STEP : alpha = rr_{k-1}/<p,q>_{k-1}; r_k = r_{k-1} - alpha q_{k-1} (own and
//             gathered rows, on the fly); x += alpha p_{k-1};
//             rr_k = rr_{k-1} - 2 alpha <r,q>_{k-1} + alpha^2 <q,q>_{k-1};
//             beta = rr_k / rr_{k-1}
//   CG_START: r_0 given, alpha = beta = 0
//   CG_FRESH: r_k already in buffer k & 1 (a replacement), alpha = 0,
//             beta = <r_k,r_k> (replacement partials) / rr_{k-1}
// then s = A r_k; p_k = r_k + beta p_{k-1}; q_k = s + beta q_{k-1}; partials
// <r,r>, <p,q>, <r,q>, <q,q> of iteration k.  A non-positive curvature or a
// negative recurrence restarts the direction (alpha = beta = 0).
//
// The slice's metadata and activity come in one dependent load; the own row,
// the first ELL slot and the scalar partials are then all requested before
// any of them is used, so the scalar reduction overlaps the first loads.
//
// WPS waves share a slice: wave v takes ELL slots v, v + WPS, ...; waves
// 1..WPS-1 hand their partial products to wave 0 through LDS, and wave 0
// adds them in wave order and finishes the row.  More waves per slice means
// more loads in flight for structures too small to fill the GPU.
template <typename MT, typename VT, int NR, int MODE, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_cgf(Slices sl, Partials P, SysState st, int k,
                                                  const MT* __restrict__ A, const VT* __restrict__ rprev,
                                                  const VT* __restrict__ qprev, VT* __restrict__ rcur,
                                                  VT* __restrict__ qcur, VT* __restrict__ p,
                                                  VT* __restrict__ xlo) {
    constexpr bool STEP = MODE == CG_STEP;
    __shared__ double lds[4 * NR * WPS];
    __shared__ VT red[WPS > 1 ? WPS - 1 : 1][6 * NR][64];
    __shared__ VT stage[64 * 6 * NR];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int4 ma = sl.meta[3 * w];
    const int4 mb = sl.meta[3 * w + 1];
    const int4 mc = sl.meta[3 * w + 2];
    const int amask = st.slice_act[w];
    // (every metadata field is >= 0: testing them only makes all loads one round trip)
    if ((amask | ((ma.x | ma.y | ma.z | ma.w | mb.x | mb.y | mb.z | mb.w | mc.x | mc.y) < 0)) == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    const int pup = mb.x, nup = mb.y, plo = mb.z, nlo = mb.w, s0 = mc.x, s1 = mc.y;
    const int nslot = nup + nlo;
    const int pk = (k - 1) & 1, ck = k & 1;
    const VT* rsrc = STEP ? rprev : rcur;   // where r_{k-1} (STEP) or r_k lives
    const MT* Aw = A + 36 * (int64_t)ma.w;   // this slice's stored (upper) value slots
    // column position of slot jj: upper slots first, then lower slots
    auto cpos = [&](int jj) -> int64_t {
        return jj < nup ? (int64_t)pup + (int64_t)jj * ROWS_PER_WG + lane
                        : (int64_t)plo + (int64_t)(jj - nup) * ROWS_PER_WG + lane;
    };

    // Straight-line loads, all requested before any is used: the own row (r, q,
    // p and x, needed by the update; every wave loads it so that no branch
    // splits the requests), the first ELL slot's column (a slot past the end
    // reads the metadata word instead and is replaced), and the first scalar
    // partials.
    // (p and x only needed at the end are loaded there for many load cases: registers)
    constexpr bool EARLY = NR * sizeof(VT) <= 8;
    VT rown[6][NR], qown[6][NR], pown[6][NR], xown[6][NR];
    auto load_px = [&]() {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) {
                const int64_t o = (6 * ii + c) * NR + h;
                pown[c][h] = MODE == CG_START ? VT(0) : p[o];
                xown[c][h] = STEP ? xlo[o] : VT(0);
            }
    };
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * ii + c) * NR + h;
            rown[c][h] = rsrc[o];
            qown[c][h] = MODE == CG_START ? VT(0) : qprev[o];
        }
    if (EARLY) load_px();
    int j = wave;
    // (past the last slot, the loads read this slice's metadata record instead: a valid address)
    const int32_t* dummy = (const int32_t*)&sl.meta[3 * w];
    int64_t col = *(j < nslot ? sl.ell_col + cpos(j) : dummy);
    int vp = *(j < nslot ? sl.ell_vpos + cpos(j) : dummy);
    col = j < nslot ? col : ii;

    // CG scalars of this structure (identical in every thread of every slice)
    VT alpha[NR], beta[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) alpha[h] = beta[h] = VT(0);
    if (MODE != CG_START) {
        constexpr int K = STEP ? 4 : 2;
        constexpr int NT = 64 * WPS;   // threads sharing the partial sums
        // slices per thread loaded without a loop (fewer for many load cases: registers)
        constexpr int T0 = WPS == 1 ? 8 : (WPS == 2 ? 6 : 4);
        constexpr int T = T0 / NR > 0 ? T0 / NR : 1;
        double v[K * NR];
#pragma unroll
        for (int a = 0; a < K * NR; ++a) v[a] = 0.0;
        const double* cgp = P.cg[pk];
        auto add_slice = [&](int q, bool valid) {
            const int qq = valid ? q : s0;   // the structure's first slice: always a valid address
            double x[K * NR];
            if (STEP) {
                const double* src = cgp + (int64_t)qq * 4 * NR;
#pragma unroll
                for (int a = 0; a < 4 * NR; ++a) x[a] = src[a];
            } else {
#pragma unroll
                for (int h = 0; h < NR; ++h) {
                    x[h] = P.rtrue[(int64_t)qq * NR + h];
                    x[NR + h] = cgp[((int64_t)qq * 4 + CG_RR) * NR + h];
                }
            }
#pragma unroll
            for (int a = 0; a < K * NR; ++a) v[a] += valid ? x[a] : 0.0;
        };
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int q = s0 + (int)threadIdx.x + t * NT;
            add_slice(q, q < s1);
        }
        for (int q = s0 + (int)threadIdx.x + T * NT; q < s1; q += NT) add_slice(q, true);
        group_sum<K * NR, WPS>(v, lds);
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            if (STEP) {
                const double rr = v[CG_RR * NR + h], pq = v[CG_PQ * NR + h];
                const double rq = v[CG_RQ * NR + h], qq = v[CG_QQ * NR + h];
                if (pq > 0.0 && isfinite(pq) && rr > 0.0) {
                    const double a = rr / pq;
                    const double rrn = rr - 2.0 * a * rq + a * a * qq;
                    alpha[h] = VT(a);
                    beta[h] = rrn > 0.0 ? VT(rrn / rr) : VT(0);
                }
            } else {
                beta[h] = v[NR + h] > 0.0 ? VT(v[h] / v[NR + h]) : VT(0);
            }
        }
    }

    // neighbour products: acc = sum_j A_ij r_k(j)
    VT acc[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[c][h] = VT(0);
    for (; j < nslot; j += WPS) {
        VT xj[6][NR];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) {
                const int64_t o = (6 * col + c) * NR + h;
                VT v = rsrc[o];
                if (STEP) v = v - alpha[h] * qprev[o];
                xj[c][h] = v;
            }
        const bool upper = j < nup;
        // the block: stored here (upper), or the upper block (col, i) of another row, transposed
        MT av[36];
        if (upper)
            load_block(Aw + (int64_t)j * 36 * ROWS_PER_WG, lane, av);
        else
            load_block(A + (int64_t)(vp >> 6) * 36 * ROWS_PER_WG, vp & (ROWS_PER_WG - 1), av);
        // next slot's column before this slot's products
        const int jn = j + WPS;
        int64_t coln = *(jn < nslot ? sl.ell_col + cpos(jn) : dummy);
        const int vpn = *(jn < nslot ? sl.ell_vpos + cpos(jn) : dummy);
        coln = jn < nslot ? coln : ii;
        if (upper) {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e / 6][h] += (VT)av[e] * xj[e % 6][h];
        } else {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e % 6][h] += (VT)av[e] * xj[e / 6][h];
        }
        col = coln;
        vp = vpn;
    }
    if (WPS > 1) {
        if (wave > 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) red[wave - 1][c * NR + h][lane] = acc[c][h];
        }
        __syncthreads();
    }

    // Wave 0 finishes the rows: r_k, s = r_k + acc (unit diagonal blocks), q, p,
    // x, and the partial sums; inactive load cases keep their old values.  The
    // slice's rows of each vector are one contiguous run of nrow * 6 * NR
    // values: they leave through LDS, one vector at a time (values are
    // recomputed rather than kept, to save registers), so that every store
    // instruction writes consecutive addresses (whole cache lines).
    if (WPS > 1 && wave == 0) {
#pragma unroll
        for (int u = 0; u < WPS - 1; ++u)
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[c][h] += red[u][c * NR + h][lane];
    }
    if (!EARLY) load_px();
    const int64_t base = (int64_t)ma.y * 6 * NR;
    const int n = (ma.z - ma.y) * 6 * NR;
    double v[4 * NR];
#pragma unroll
    for (int a = 0; a < 4 * NR; ++a) v[a] = 0.0;
    auto flush = [&](auto&& value, VT* __restrict__ dst) {
        if (wave == 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) stage[lane * 6 * NR + c * NR + h] = value(c, h, on && ((amask >> h) & 1));
        }
        __syncthreads();
        for (int u = (int)threadIdx.x; u < n; u += 64 * WPS) dst[base + u] = stage[u];
        __syncthreads();
    };
    auto r_k = [&](int c, int h) { return STEP ? rown[c][h] - alpha[h] * qown[c][h] : rown[c][h]; };
    auto q_k = [&](int c, int h) { return (r_k(c, h) + acc[c][h]) + beta[h] * qown[c][h]; };
    auto p_k = [&](int c, int h) { return r_k(c, h) + beta[h] * pown[c][h]; };
    if (STEP) flush([&](int c, int h, bool run) { return run ? r_k(c, h) : rown[c][h]; }, rcur);
    flush(
        [&](int c, int h, bool run) {
            const VT rk = r_k(c, h), qn = q_k(c, h);
            if (run) {
                v[CG_RR * NR + h] += (double)rk * (double)rk;
                v[CG_RQ * NR + h] += (double)rk * (double)qn;
                v[CG_QQ * NR + h] += (double)qn * (double)qn;
            }
            return run ? qn : qown[c][h];
        },
        qcur);
    flush(
        [&](int c, int h, bool run) {
            const VT pn = p_k(c, h);
            if (run) v[CG_PQ * NR + h] += (double)pn * (double)q_k(c, h);
            return run ? pn : pown[c][h];
        },
        p);
    if (STEP) flush([&](int c, int h, bool run) { return run ? xown[c][h] + alpha[h] * pown[c][h] : xown[c][h]; }, xlo);
    if (wave == 0) {
#pragma unroll
        for (int a = 0; a < 4 * NR; ++a) v[a] = wave_sum(v[a]);
        if (lane == 0) {
            double* dst = P.cg[ck] + (int64_t)w * 4 * NR;
#pragma unroll
            for (int a = 0; a < 4 * NR; ++a) dst[a] = v[a];
        }
    }
}

// Update half of iteration k before a replacement, folded into the FP64
// solution: alpha from the partials of k - 1; x64 += xlo + alpha p; xlo = 0;
// r_k = r_{k-1} - alpha q_{k-1} into buffer k & 1.
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_upd(Slices sl, Partials P, SysState st, int k,
                                                   const VT* __restrict__ rprev, const VT* __restrict__ qprev,
                                                   VT* __restrict__ rcur, const VT* __restrict__ p,
                                                   VT* __restrict__ xlo, double* __restrict__ x64) {
    __shared__ double lds[NWAVES * 2 * NR];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int sys = sl.sys[w];
    if (structure_done<NR>(st, sys)) return;
    const double* cgp = P.cg[(k - 1) & 1];
    const int s0 = sl.sys_off[sys], s1 = sl.sys_off[sys + 1];
    double v[2 * NR];
#pragma unroll
    for (int a = 0; a < 2 * NR; ++a) v[a] = 0.0;
    for (int q = s0 + (int)threadIdx.x; q < s1; q += SOLVER_WG) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            v[h] += cgp[((int64_t)q * 4 + CG_RR) * NR + h];
            v[NR + h] += cgp[((int64_t)q * 4 + CG_PQ) * NR + h];
        }
    }
    group_sum<2 * NR, NWAVES>(v, lds);
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    if (R >= 6 * (int64_t)sl.row1[w]) return;
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        if (!st.active[sys * NR + h]) continue;
        const double rr = v[h], pq = v[NR + h];
        const bool ok = pq > 0.0 && isfinite(pq) && rr > 0.0;
        const VT alpha = ok ? VT(rr / pq) : VT(0);
        const int64_t o = R * NR + h;
        const VT xn = xlo[o] + alpha * p[o];
        x64[o] += (double)xn;
        xlo[o] = VT(0);
        rcur[o] = rprev[o] - alpha * qprev[o];
    }
}


// Compliance-type reductions: part[slice][rhs] = sum over the slice's DOFs of
// f u; then C[sys][rhs] = scale * (sum of its slices' partials, in order).
template <int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_fu_partial(Slices sl, const double* __restrict__ f,
                                                          const double* __restrict__ u, double* __restrict__ part) {
    __shared__ double lds[NWAVES * NR];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool on = R < 6 * (int64_t)sl.row1[w];
    double v[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = on ? f[R * NR + h] * u[R * NR + h] : 0.0;
    write_partials<NR, NWAVES>(v, lds, part, w);
}

template <int NR>
__global__ __launch_bounds__(64) void k_fu_total(Slices sl, int nsys, const double* __restrict__ part, double scale,
                                                 double* __restrict__ C) {
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    const int s0 = sl.sys_off[sys], s1 = sl.sys_off[sys + 1];
    double v[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = 0.0;
    for (int w = s0 + (int)threadIdx.x; w < s1; w += 64) {
#pragma unroll
        for (int h = 0; h < NR; ++h) v[h] += part[(int64_t)w * NR + h];
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = wave_sum(v[h]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) C[(int64_t)sys * NR + h] = scale * v[h];
    }
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------

int get_workspace(Plan* pl, int nrhs, Workspace& w) {
    const size_t n = 6 * (size_t)pl->n_nodes * nrhs;
    const size_t nsys = (size_t)pl->n_sys * nrhs;
    const size_t npart = (size_t)pl->n_wg * nrhs;
    size_t need[WS_N] = {};
    for (int i = WS_B; i <= WS_P; ++i) need[i] = n * 8;
    need[WS_PART] = npart * 8 * 15;
    need[WS_SD] = nsys * 8 * 3;
    need[WS_SI] = (nsys * 5 + (size_t)pl->n_wg) * 4;
    if (pl->ws.size() < WS_N) pl->ws.resize(WS_N);
    for (int i = 0; i < WS_N; ++i) {
        if (i == WS_ADJ || need[i] == 0) continue;
        if (pl->ws[i].bytes < need[i]) {
            if (pl->ws[i].p) JSSO_HIP(hipFree(pl->ws[i].p));
            pl->ws[i].p = nullptr;
            pl->ws[i].bytes = 0;
            hipError_t e = hipMalloc(&pl->ws[i].p, need[i]);
            JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "workspace hipMalloc(%zu): %s", need[i], hipGetErrorString(e));
            pl->ws[i].bytes = need[i];
        }
    }
    w.b64 = (double*)pl->ws[WS_B].p;
    w.x64 = (double*)pl->ws[WS_X].p;
    w.xbest = (double*)pl->ws[WS_XBEST].p;
    w.xlo = pl->ws[WS_XLO].p;
    w.r[0] = pl->ws[WS_R0].p;
    w.r[1] = pl->ws[WS_R1].p;
    w.q[0] = pl->ws[WS_Q0].p;
    w.q[1] = pl->ws[WS_Q1].p;
    w.p = pl->ws[WS_P].p;
    double* pp = (double*)pl->ws[WS_PART].p;
    w.P.cg[0] = pp;
    w.P.cg[1] = pp + 4 * npart;
    w.P.b = pp + 8 * npart;
    w.P.t = pp + 9 * npart;
    w.P.rtrue = pp + 10 * npart;
    w.P.xx = pp + 14 * npart;
    w.P.rr0 = pp + 11 * npart;
    w.P.rr1 = pp + 12 * npart;
    w.P.pqp = pp + 13 * npart;
    double* sd = (double*)pl->ws[WS_SD].p;
    w.st.bnorm2 = sd;
    w.st.relres = sd + nsys;
    w.st.best = sd + 2 * nsys;
    int32_t* si = (int32_t*)pl->ws[WS_SI].p;
    w.st.active = si;
    w.st.status = si + nsys;
    w.st.stall = si + 2 * nsys;
    w.st.save = si + 3 * nsys;
    w.st.iters = si + 4 * nsys;
    w.st.slice_act = si + 5 * nsys;
    return JSSO_OK;
}

// the assembly's flags: 1 = a constrained diagonal block is not positive
// definite, 2 = a beam-column is parallel to global Y (K not symmetric)
int assembly_status(Plan* p, hipStream_t s, int32_t* status) {
    JSSO_REQUIRE(p->on_device, JSSO_EINVAL, "plan not uploaded to the device");
    if (p->status_cache < 0) {
        int32_t v = 0;
        JSSO_HIP(hipMemcpyAsync(&v, p->d_status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        JSSO_HIP(hipStreamSynchronize(s));
        p->status_cache = v;
    }
    *status = p->status_cache;
    return JSSO_OK;
}

int check_status(Plan* p, hipStream_t s) {
    int32_t status = 0;
    int rc = assembly_status(p, s, &status);
    if (rc) return rc;
    JSSO_REQUIRE(!(status & 1), JSSO_ESINGULAR,
                 "a constrained diagonal block of K is not positive definite (unsupported or "
                 "isolated node)");
    JSSO_REQUIRE(!(status & STATUS_NONSYM) || p->sym_part_solve, JSSO_EUNSUPPORTED,
                 "K is not symmetric (a beam-column parallel to global Y, element.py:92-94, or a quad with "
                 "kx_mod != ky_mod, element.py:852-875); CG cannot solve it (jsso_solve routes such plans "
                 "to the dense LU or the defect correction)");
    return JSSO_OK;
}

Slices make_slices(Plan* pl, int base, int count) {
    return Slices{pl->d_wg_sys,  pl->d_wg_row0, pl->d_wg_row1, pl->d_sys_wg_off,
                  pl->d_ell_vpos, pl->d_ell_col, (const int4*)pl->d_slice_meta, base,
                  count,          slice_run(count, pl->slice_spread && base == 0 && count == pl->n_wg)};
}

// Launch one ordinary CG iteration with `wps` waves per slice (falls back to
// the largest variant whose LDS fits).
template <typename MT, typename VT, int NR>
static void launch_step(int wps, const Slices& sl, const Workspace& w, int it, const MT* A, VT* const (&rb)[2],
                        VT* const (&qb)[2], VT* p, VT* xlo, hipStream_t s) {
    const dim3 grid = slice_grid(sl);
    const VT* rp = rb[(it - 1) & 1];
    const VT* qp = qb[(it - 1) & 1];
    if constexpr (WpsFits<VT, NR, 4>::value) {
        if (wps >= 4) {
            hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_STEP, 4>), grid, dim3(256), 0, s, sl, w.P, w.st, it, A, rp, qp,
                               rb[it & 1], qb[it & 1], p, xlo);
            return;
        }
    }
    if constexpr (WpsFits<VT, NR, 2>::value) {
        if (wps >= 2) {
            hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_STEP, 2>), grid, dim3(128), 0, s, sl, w.P, w.st, it, A, rp, qp,
                               rb[it & 1], qb[it & 1], p, xlo);
            return;
        }
    }
    hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_STEP, 1>), grid, dim3(64), 0, s, sl, w.P, w.st, it, A, rp, qp,
                       rb[it & 1], qb[it & 1], p, xlo);
}

template <typename MT, typename VT, int NR>
static void launch_init(Plan* pl, const Slices& sl, const Workspace& w, const double* f, const MT* A,
                        VT* const (&rb)[2], VT* const (&qb)[2], VT* p, VT* xlo, double tol, hipStream_t s) {
    const dim3 gridS = slice_grid(sl);
    hipLaunchKernelGGL((k_init<VT, NR>), gridS, dim3(SOLVER_WG), 0, s, sl, w.P, pl->d_fixed, pl->d_Linv, f, w.b64,
                       w.x64, w.xbest, xlo, rb[0], p);
    hipLaunchKernelGGL((k_decide<NR>), dim3(pl->n_sys), dim3(64), 0, s, sl, w.P, w.st, pl->n_sys, 0, tol, 0, 0,
                       (const double*)nullptr, 0.0, 0);
    hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_START, 1>), gridS, dim3(64), 0, s, sl, w.P, w.st, 0, A,
                       (const VT*)rb[1], (const VT*)qb[1], rb[0], qb[0], p, xlo);
}

// Autotuner for the waves-per-slice count: time `reps` ordinary iterations of
// each variant (after one untimed launch) on the live problem and keep the
// fastest; a larger count must win by 3% to be chosen.  The caller
// re-initialises the solve afterwards, so the result does not depend on the
// iterations run here.
template <typename MT, typename VT, int NR>
static int tune_wps(Plan* pl, const Slices& sl, const Workspace& w, const MT* A, VT* const (&rb)[2],
                    VT* const (&qb)[2], VT* p, VT* xlo, hipStream_t s, int& chosen) {
    const int reps = 8;
    hipEvent_t e0, e1;
    JSSO_HIP(hipEventCreate(&e0));
    JSSO_HIP(hipEventCreate(&e1));
    int it = 0;
    float best_ms = 0.f;
    chosen = 1;
    for (int wps : {1, 2, 4}) {
        if ((wps == 2 && !WpsFits<VT, NR, 2>::value) || (wps == 4 && !WpsFits<VT, NR, 4>::value)) continue;
        launch_step<MT, VT, NR>(wps, sl, w, ++it, A, rb, qb, p, xlo, s);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch_step<MT, VT, NR>(wps, sl, w, ++it, A, rb, qb, p, xlo, s);
        (void)hipEventRecord(e1, s);
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            set_error("autotune: %s", hipGetErrorString(e));
            return JSSO_EHIP;
        }
        if (wps == 1 || ms < 0.97f * best_ms) {
            best_ms = ms;
            chosen = wps;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

// The replacement step closing a chunk that ends at iteration `it` after
// `steps` iterations: update + fold into FP64, true FP64 residual, convergence
// test, best iterate, fresh direction.
template <typename MT, typename VT, int NR>
static void launch_replace(Plan* pl, int wps, const Slices& sl, const Workspace& w, int it, int steps, double tol,
                           const MT* A, VT* const (&rb)[2], VT* const (&qb)[2], VT* p, VT* xlo, hipStream_t s) {
    const dim3 gridS = slice_grid(sl);
    hipLaunchKernelGGL((k_upd<VT, NR>), gridS, dim3(SOLVER_WG), 0, s, sl, w.P, w.st, it, (const VT*)rb[(it - 1) & 1],
                       (const VT*)qb[(it - 1) & 1], rb[it & 1], (const VT*)p, xlo, w.x64);
    const double* A64 = pl->d_Ke64;
    const double* b64 = w.b64;
    const double* x64 = w.x64;
    bool done = false;
    if constexpr (WpsFits<double, NR, 4>::value) {
        if (!done && wps >= 4) {
            hipLaunchKernelGGL((k_replace<VT, NR, 4>), gridS, dim3(256), 0, s, sl, w.P, w.st, A64, b64, x64,
                               rb[it & 1]);
            done = true;
        }
    }
    if constexpr (WpsFits<double, NR, 2>::value) {
        if (!done && wps >= 2) {
            hipLaunchKernelGGL((k_replace<VT, NR, 2>), gridS, dim3(128), 0, s, sl, w.P, w.st, A64, b64, x64,
                               rb[it & 1]);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL((k_replace<VT, NR, 1>), gridS, dim3(64), 0, s, sl, w.P, w.st, A64, b64, x64, rb[it & 1]);
    hipLaunchKernelGGL((k_decide<NR>), dim3(pl->n_sys), dim3(64), 0, s, sl, w.P, w.st, pl->n_sys, 1, tol, steps, 8192,
                       (const double*)nullptr, 0.0, 0);
    hipLaunchKernelGGL((k_save<NR>), gridS, dim3(SOLVER_WG), 0, s, sl, (const int32_t*)w.st.save,
                       (const double*)w.x64, w.xbest);
    const VT* rp = rb[(it - 1) & 1];
    const VT* qp = qb[(it - 1) & 1];
    done = false;
    if constexpr (WpsFits<VT, NR, 4>::value) {
        if (!done && wps >= 4) {
            hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_FRESH, 4>), gridS, dim3(256), 0, s, sl, w.P, w.st, it, A, rp, qp,
                               rb[it & 1], qb[it & 1], p, xlo);
            done = true;
        }
    }
    if constexpr (WpsFits<VT, NR, 2>::value) {
        if (!done && wps >= 2) {
            hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_FRESH, 2>), gridS, dim3(128), 0, s, sl, w.P, w.st, it, A, rp, qp,
                               rb[it & 1], qb[it & 1], p, xlo);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL((k_cgf<MT, VT, NR, CG_FRESH, 1>), gridS, dim3(64), 0, s, sl, w.P, w.st, it, A, rp, qp,
                           rb[it & 1], qb[it & 1], p, xlo);
}

// HIP graphs of a chunk: its ordinary iterations (one graph launch replaces
// steps - 1 kernel launches, so the host's launch rate no longer bounds short
// iterations) and its replacement step.  Captured once per (buffers, variant,
// chunk length) on a plan-owned stream and launched on the caller's stream.
// Kernels select buffers by the parity of the iteration only, and the
// convergence test counts iterations per system on the device, so with an
// even chunk length the same two graphs serve every chunk.  JSSO_CG_GRAPH=0
// launches eagerly.
bool graphs_enabled() {
    const char* e = getenv("JSSO_CG_GRAPH");
    return !(e && e[0] == '0');
}


int ensure_flag_buffers(Plan* pl, size_t nsys) {
    if (pl->h_flags_count < 2 * nsys) {
        if (pl->h_flags) JSSO_HIP(hipHostFree(pl->h_flags));
        if (pl->h_rel) JSSO_HIP(hipHostFree(pl->h_rel));
        pl->h_flags = nullptr;
        pl->h_rel = nullptr;
        pl->h_flags_count = 0;
        JSSO_HIP(hipHostMalloc((void**)&pl->h_flags, 2 * nsys * sizeof(int32_t), hipHostMallocDefault));
        JSSO_HIP(hipHostMalloc((void**)&pl->h_rel, 2 * nsys * sizeof(double), hipHostMallocDefault));
        pl->h_flags_count = 2 * nsys;
    }
    for (auto& e : pl->flag_ev)
        if (!e) JSSO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return JSSO_OK;
}

template <typename MT, typename VT, int NR>
static int run_batched(Plan* pl, const double* f, double* u, double tol, int maxit, int replace_every,
                       jsso_info* info, hipStream_t s) {
    Workspace w;
    int rc = get_workspace(pl, NR, w);
    if (rc) return rc;
    const Slices sl = make_slices(pl, 0, pl->n_wg);
    const int nsys = pl->n_sys * NR;
    if ((rc = ensure_flag_buffers(pl, nsys))) return rc;
    const MT* A = nullptr;
    if constexpr (sizeof(MT) == 4) A = (const MT*)pl->d_Ke32; else A = (const MT*)pl->d_Ke64;
    VT *xlo = (VT*)w.xlo, *p = (VT*)w.p;
    VT* const rb[2] = {(VT*)w.r[0], (VT*)w.r[1]};
    VT* const qb[2] = {(VT*)w.q[0], (VT*)w.q[1]};
    int& wps = pl->cg_wps[prec_slot<MT, VT>()][NR];
    launch_init<MT, VT, NR>(pl, sl, w, f, A, rb, qb, p, xlo, tol, s);
    JSSO_HIP(hipGetLastError());
    if (wps <= 0) {
        rc = tune_wps<MT, VT, NR>(pl, sl, w, A, rb, qb, p, xlo, s, wps);
        if (rc) return rc;
        launch_init<MT, VT, NR>(pl, sl, w, f, A, rb, qb, p, xlo, tol, s);
        JSSO_HIP(hipGetLastError());
    }
    const int E = replace_every;
    // everything the captured kernels' arguments point at
    const void* key[CgGraph::NKEY] = {A,         rb[0],       rb[1],       qb[0],         qb[1],   p,
                                      xlo,       w.P.cg[0],   w.P.cg[1],   w.st.slice_act, w.b64,  w.x64,
                                      w.xbest,   w.st.best,   w.st.active, pl->d_Ke64};
    CgGraph(&gc)[2] = pl->cg_graph[prec_slot<MT, VT>()][NR];

    // One chunk: iterations it0+1 .. it0+steps-1, then the replacement step at it0+steps.
    auto enqueue_chunk = [&](int it0, int steps) -> int {
        hipGraphExec_t ga = nullptr, gb = nullptr;
        if (steps == E && E % 2 == 0 && it0 % 2 == 0 && pl->graphs_ok && graphs_enabled()) {
            ga = capture_graph(pl, gc[0], key, wps, E, tol, [&](hipStream_t cs) {
                for (int it = 1; it < E; ++it) launch_step<MT, VT, NR>(wps, sl, w, it, A, rb, qb, p, xlo, cs);
            });
            if (ga)
                gb = capture_graph(pl, gc[1], key, wps, E, tol, [&](hipStream_t cs) {
                    launch_replace<MT, VT, NR>(pl, wps, sl, w, E, E, tol, A, rb, qb, p, xlo, cs);
                });
        }
        if (ga && gb) {
            std::pair<hipEvent_t, hipEvent_t>* ev;
            timing_begin_batch(s, ev, steps - 1);
            JSSO_HIP(hipGraphLaunch(ga, s));
            timing_end(s, ev);
            JSSO_HIP(hipGraphLaunch(gb, s));
        } else {
            for (int it = it0 + 1; it < it0 + steps; ++it) {
                std::pair<hipEvent_t, hipEvent_t>* ev;
                timing_begin(s, ev);
                launch_step<MT, VT, NR>(wps, sl, w, it, A, rb, qb, p, xlo, s);
                timing_end(s, ev);
            }
            launch_replace<MT, VT, NR>(pl, wps, sl, w, it0 + steps, steps, tol, A, rb, qb, p, xlo, s);
        }
        JSSO_HIP(hipGetLastError());
        return JSSO_OK;
    };

    // The host reads chunk c's convergence flags while chunk c + 1 runs (one
    // chunk of lookahead): the GPU never waits for the host, and a chunk
    // queued after everything converged finds every slice inactive and exits
    // at once without touching the solution.
    std::vector<int> chunk_end;
    int it = 0, it_done = 0, chunks_done = 0, c = 0;
    for (;;) {
        const bool more = it < maxit;
        if (more) {
            const int steps = std::min(E, maxit - it);
            if ((rc = enqueue_chunk(it, steps))) return rc;
            it += steps;
            chunk_end.push_back(it);
            JSSO_HIP(hipMemcpyAsync(pl->h_flags + (size_t)(c & 1) * nsys, w.st.active, nsys * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
            JSSO_HIP(hipEventRecord(pl->flag_ev[c & 1], s));
            ++c;
        }
        const int chk = more ? c - 2 : c - 1;
        if (chk >= 0) {
            JSSO_HIP(hipEventSynchronize(pl->flag_ev[chk & 1]));
            timing_collect();
            const int32_t* fl = pl->h_flags + (size_t)(chk & 1) * nsys;
            bool any = false;
            for (int k = 0; k < nsys && !any; ++k) any = fl[k] != 0;
            if (!any) {
                it_done = chunk_end[chk];
                chunks_done = chk + 1;
                break;
            }
        }
        if (!more) {
            it_done = it;
            chunks_done = c;
            break;
        }
    }
    const int64_t nr = 6 * (int64_t)pl->n_nodes;
    hipLaunchKernelGGL((k_unscale<NR>), dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, pl->n_nodes,
                       pl->d_Linv, (const double*)w.xbest, u);
    JSSO_HIP(hipGetLastError());
    std::vector<double> rel(nsys);
    std::vector<int32_t> status(nsys);
    JSSO_HIP(hipMemcpyAsync(rel.data(), w.st.best, nsys * sizeof(double), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipMemcpyAsync(status.data(), w.st.status, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    JSSO_HIP(hipStreamSynchronize(s));
    timing_collect();
    int conv = 0;
    double mx = 0.0;
    for (int k = 0; k < nsys; ++k) {
        conv += (status[k] == SYS_CONVERGED || status[k] == SYS_FLOOR) ? 1 : 0;
        mx = std::max(mx, rel[k]);
    }
    if (info) {
        info->iterations = it_done;
        info->replacements = chunks_done;
        info->converged = conv;
        info->n_systems = nsys;
        info->rel_residual = mx;
        info->waves_per_slice = wps;
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

// Iterations between FP64 residual replacements; JSSO_REPLACE_EVERY overrides
// the defaults (64 mixed, 256 FP64) for tuning.  (128 saves 3% on large
// shells but lets the FP32 recurrences of slender beam systems drift too far.)
static int replace_interval(bool mixed) {
    const char* e = getenv("JSSO_REPLACE_EVERY");
    const int v = e ? atoi(e) : 0;
    return v > 0 ? v : (mixed ? 64 : 256);
}

template <int NR>
static int run_compliance(Plan* pl, const double* f, const double* u, double scale, double* C, hipStream_t s) {
    const size_t need = (size_t)pl->n_wg * NR * sizeof(double);
    if (pl->ws.size() <= WS_FU_SLOT) pl->ws.resize(WS_FU_SLOT + 1);
    DBuf& g = pl->ws[WS_FU_SLOT];
    if (g.bytes < need) {
        if (g.p) JSSO_HIP(hipFree(g.p));
        g.p = nullptr;
        g.bytes = 0;
        hipError_t e = hipMalloc(&g.p, need);
        JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "f.u partials hipMalloc(%zu): %s", need, hipGetErrorString(e));
        g.bytes = need;
    }
    const Slices sl = make_slices(pl, 0, pl->n_wg);
    hipLaunchKernelGGL((k_fu_partial<NR>), slice_grid(sl), dim3(SOLVER_WG), 0, s, sl, f, u, (double*)g.p);
    hipLaunchKernelGGL((k_fu_total<NR>), dim3(pl->n_sys), dim3(64), 0, s, sl, pl->n_sys, (const double*)g.p, scale,
                       C);
    JSSO_HIP(hipGetLastError());
    return JSSO_OK;
}

int launch_compliance(Plan* p, const double* f, const double* u, int nrhs, double scale, double* C, hipStream_t s) {
    if (p->n_wg == 0) return JSSO_OK;
    switch (nrhs) {
        case 1: return run_compliance<1>(p, f, u, scale, C, s);
        case 2: return run_compliance<2>(p, f, u, scale, C, s);
        case 3: return run_compliance<3>(p, f, u, scale, C, s);
        case 4: return run_compliance<4>(p, f, u, scale, C, s);
        case 8: return run_compliance<8>(p, f, u, scale, C, s);
        default:
            set_error("nrhs=%d unsupported (1, 2, 3, 4 or 8)", nrhs);
            return JSSO_EINVAL;
    }
}

int solve_batched(Plan* p, const double* f, double* u, int nrhs, int slot, double tol, int maxit,
                  jsso_info* info, hipStream_t s) {
    int rc = check_status(p, s);
    if (rc) return rc;
    if (p->precond == JSSO_PC_AMG) return amg_solve(p, f, u, nrhs, slot, tol, maxit, info, s);
    if (slot == PS_F32) return dispatch_nr<float, float>(p, f, u, nrhs, tol, maxit, replace_interval(true), info, s);
    if (slot == PS_F32_V64)
        return dispatch_nr<float, double>(p, f, u, nrhs, tol, maxit, replace_interval(true), info, s);
    return dispatch_nr<double, double>(p, f, u, nrhs, tol, maxit, replace_interval(false), info, s);
}

// ---------------------------------------------------------------------------
// plain path: block-Jacobi PCG on Kc (FP64, block CSR), one system per launch
// sequence; reductions as above (per-slice partials summed by the consumers).
// <r,z> partials alternate between P.rr0 / P.rr1; <r0,z0> lives in P.b.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SOLVER_WG) void kp_init(Slices sl, Partials P, int h, int nrhs,
                                                     const uint8_t* __restrict__ fixed,
                                                     const double* __restrict__ Minv, const double* __restrict__ f,
                                                     double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ z, double* __restrict__ pp) {
    __shared__ double lds[NWAVES];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool on = R < 6 * (int64_t)sl.row1[w];
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
    double v2[1] = {v[0]};
    write_partials<1, NWAVES>(v, lds, P.rr0, w);
    write_partials<1, NWAVES>(v2, lds, P.b, w);
}

// one thread: anything to solve?
__global__ void kp_start(Slices sl, Partials P, int32_t* active, double* relres) {
    const int sys = sl.sys[sl.base];
    double rz0 = 0.0;
    for (int w = sl.sys_off[sys]; w < sl.sys_off[sys + 1]; ++w) rz0 += P.b[w];
    active[0] = rz0 > 0.0 ? 1 : 0;
    relres[0] = rz0 > 0.0 ? 1.0 : 0.0;
}

// q = Kc p; partial <p,q>.
__global__ __launch_bounds__(SOLVER_WG) void kp_spmv(Slices sl, Partials P, const int32_t* __restrict__ active,
                                                     const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ K,
                                                     const double* __restrict__ pp, double* __restrict__ q) {
    __shared__ double lds[NWAVES];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool on = R < 6 * (int64_t)sl.row1[w];
    double v[1] = {0.0};
    if (on && active[0]) {
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
    write_partials<1, NWAVES>(v, lds, P.pqp, w);
}

// alpha = rz / <p,q>; x += alpha p; r -= alpha q; z = M r; partial <r,z> into rr(it).
__global__ __launch_bounds__(SOLVER_WG) void kp_update(Slices sl, Partials P, const int32_t* __restrict__ active,
                                                       int it, const double* __restrict__ Minv,
                                                       const double* __restrict__ pp, const double* __restrict__ q,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       double* __restrict__ z) {
    __shared__ double lds[NWAVES * 2];
    __shared__ double rs[SOLVER_WG];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    double rz[1], pq[1];
    system_totals<1, NWAVES>(sl, sl.sys[w], P.rr(it - 1), P.pqp, rz, pq, lds);
    const bool act = active[0] != 0;
    const double alpha = (pq[0] > 0.0 && isfinite(pq[0])) ? rz[0] / pq[0] : 0.0;
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool on = R < 6 * (int64_t)sl.row1[w];
    double rn = 0.0;
    if (on) {
        rn = r[R];
        if (act) {
            x[R] += alpha * pp[R];
            rn -= alpha * q[R];
            r[R] = rn;
        }
    }
    rs[threadIdx.x] = rn;
    __syncthreads();
    double v[1] = {0.0};
    if (on) {
        const int64_t i = R / 6;
        const int rr = (int)(R % 6);
        const int base = (int)threadIdx.x - rr;
        double zs = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) zs += Minv[i * 36 + rr * 6 + c] * rs[base + c];
        if (act) z[R] = zs;
        v[0] = rn * zs;
    }
    write_partials<1, NWAVES>(v, lds, P.rr(it), w);
}

// beta = rz_new / rz_old; p = z + beta p.  Convergence on sqrt(rz / rz0); the
// first slice records it for the host and the following launches.
__global__ __launch_bounds__(SOLVER_WG) void kp_dir(Slices sl, Partials P, int32_t* __restrict__ active,
                                                    double* __restrict__ relres, int it, double tol,
                                                    const double* __restrict__ z, double* __restrict__ pp) {
    __shared__ double lds[NWAVES * 2];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    double rzn[1], rzo[1], rz0[1], unused[1];
    system_totals<1, NWAVES>(sl, sl.sys[w], P.rr(it), P.rr(it - 1), rzn, rzo, lds);
    system_totals<1, NWAVES>(sl, sl.sys[w], P.b, nullptr, rz0, unused, lds);
    const bool act = active[0] != 0;
    const double rel = sqrt(fabs(rzn[0]) / rz0[0]);
    const bool conv = !(rel > tol);
    __syncthreads();   // every thread has read active[0] before it may change
    if (threadIdx.x == 0 && w == sl.base && act) {
        relres[0] = rel;
        if (conv) active[0] = 0;
    }
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    if (!act || conv || R >= 6 * (int64_t)sl.row1[w]) return;
    const double beta = rzo[0] > 0.0 ? rzn[0] / rzo[0] : 0.0;
    pp[R] = z[R] + beta * pp[R];
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
    if ((rc = launch_full_blocks(pl, s))) return rc;   // the plain path runs on Kc in block CSR
    Workspace w;
    rc = get_workspace(pl, 1, w);
    if (rc) return rc;
    double *x = w.x64, *r = w.xbest, *z = w.b64, *pp = (double*)w.p, *q = (double*)w.q[0];
    int32_t* active = w.st.active;
    double* relres = w.st.relres;
    const int nsys = pl->n_sys * nrhs;
    int itmax = 0, conv = 0;
    double mx = 0.0;
    for (int sidx = 0; sidx < pl->n_sys; ++sidx) {
        const int w0 = pl->h_sys_wg_off[sidx], w1 = pl->h_sys_wg_off[sidx + 1];
        const Slices sl = make_slices(pl, w0, w1 - w0);
        const dim3 grid = slice_grid(sl), block(SOLVER_WG);
        for (int h = 0; h < nrhs; ++h) {
            hipLaunchKernelGGL(kp_init, grid, block, 0, s, sl, w.P, h, nrhs, pl->d_fixed, pl->d_Minv, f, x, r, z, pp);
            hipLaunchKernelGGL(kp_start, dim3(1), dim3(1), 0, s, sl, w.P, active, relres);
            JSSO_HIP(hipGetLastError());
            int it = 0;
            int32_t act = 1;
            double rel = 0.0;
            while (it < maxit) {
                const int steps = std::min(32, maxit - it);
                for (int j = 0; j < steps; ++j) {
                    ++it;
                    hipLaunchKernelGGL(kp_spmv, grid, block, 0, s, sl, w.P, (const int32_t*)active, pl->d_rowptr,
                                       pl->d_colind, (const double*)pl->d_Kc, (const double*)pp, q);
                    hipLaunchKernelGGL(kp_update, grid, block, 0, s, sl, w.P, (const int32_t*)active, it,
                                       (const double*)pl->d_Minv, (const double*)pp, (const double*)q, x, r, z);
                    hipLaunchKernelGGL(kp_dir, grid, block, 0, s, sl, w.P, active, relres, it, tol,
                                       (const double*)z, pp);
                }
                JSSO_HIP(hipGetLastError());
                JSSO_HIP(hipMemcpyAsync(&act, active, sizeof(int32_t), hipMemcpyDeviceToHost, s));
                JSSO_HIP(hipStreamSynchronize(s));
                if (!act) break;
            }
            JSSO_HIP(hipMemcpy(&rel, relres, sizeof(double), hipMemcpyDeviceToHost));
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
