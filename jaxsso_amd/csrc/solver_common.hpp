// Shared pieces of the batched solvers (jsso_solve.hip: block-Jacobi CG;
// jsso_amg.hip: AMG-preconditioned CG): the slice layout, per-system state,
// per-slice partial sums, deterministic wavefront / workgroup reductions, the
// solver workspace and the host helpers both drivers use.
#pragma once
#include <utility>

#include "jsso_internal.hpp"

namespace jsso {

constexpr int NWAVES = SOLVER_WG / 64;
// SYS_FLOOR: stopped at the attainable accuracy of the FP64 residual (counted as
// converged, see k_decide)
enum { SYS_RUNNING = 0, SYS_CONVERGED = 1, SYS_STALLED = 2, SYS_FLOOR = 3 };

// ---------------------------------------------------------------------------
// slices and reductions
// ---------------------------------------------------------------------------
struct Slices {
    const int32_t *sys, *row0, *row1, *sys_off;  // slice -> structure, rows; structure -> first slice
    const int32_t *ell_vpos, *ell_col;           // symmetric sliced ELL: value position and column per position
    // per slice, three int4: {sys, row0, row1, first value position},
    // {first upper column position, upper slots, first lower column position, lower slots},
    // {first slice of sys, end slice of sys, 0, 0}
    const int4* meta;
    int base;           // first slice of the launch (plain path)
    int count;          // slices in the launch
    int run;            // slices per run of the workgroup -> slice map (slice_run)
};

// Workgroups are dispatched round-robin over the 8 XCDs.  They are mapped to
// slices in runs of `run` consecutive slices, run r going to XCD r mod 8,
// identically in every kernel of a level, so that neighbouring slices (which
// share vector rows, and the stored blocks read through the transpose) stay in
// one XCD's L2 within a run.  One structure, or only small ones (below 64
// slices): one run per XCD (a contiguous eighth of the slices each).  Several
// structures of which one or more is large: runs of 16 slices, shorter
// on small levels so that every XCD gets ~8 or more, which spreads every
// structure over all eight XCDs: as structures converge and their slices go
// idle, the ones still iterating keep every XCD equally busy (with an eighth of
// the plan per XCD, all the slices of a straggler would sit on one XCD).  The
// map is a bijection of [0, grid); the workgroups it sends past `count` return -1.
static inline int slice_run(int count, bool spread) {
    if (!spread) return count > 8 ? (count + 7) / 8 : 1;
    int run = 1;
    while (run < 16 && (count / (2 * run)) >= 64) run *= 2;   // run <= count / 64, at most 16
    return run;
}

__device__ __forceinline__ int slice_of_block(const Slices& sl) {
    const int b = blockIdx.x, bx = b >> 3, run = sl.run;
    const int l = ((bx / run) * 8 + (b & 7)) * run + bx % run;
    return l < sl.count ? sl.base + l : -1;
}

static inline dim3 slice_grid(const Slices& sl) {
    const int u = 8 * sl.run;
    return dim3(u * ((sl.count + u - 1) / u));
}

// per-(structure, rhs) state, written only by k_decide
struct SysState {
    double *bnorm2, *relres, *best;
    int32_t *active, *status, *stall, *save;
    int32_t* slice_act;   // per slice: bit h set while load case h of its structure runs
    int32_t* iters;       // CG iterations run so far, per (structure, rhs)
};

// Partial-sum buffers, per slice.  The fused CG kernel of iteration k writes
// cg[k & 1] and reads cg[(k - 1) & 1]; each slice holds [4][rhs] values.
enum { CG_RR = 0, CG_PQ = 1, CG_RQ = 2, CG_QQ = 3 };
struct Partials {
    double* cg[2];           // fused CG: <r,r>, <p,q>, <r,q>, <q,q>
    double *b, *t, *rtrue;   // <b,b>; true FP64 <r,r> and low-precision <r,r> after a replacement
    double* xx;              // <x,x> of the FP64 solution at a replacement
    double *rr0, *rr1;       // plain path: alternating <r,z>
    double* pqp;             // plain path: <p,q>
    __host__ __device__ double* rr(int it) const { return (it & 1) ? rr1 : rr0; }
};

// true when every load case of structure `sys` has stopped: its slices have
// nothing to do (and nothing will read their partial sums)
template <int NR>
__device__ __forceinline__ bool structure_done(const SysState& st, int sys) {
    bool any = false;
#pragma unroll
    for (int h = 0; h < NR; ++h) any |= st.active[sys * NR + h] != 0;
    return !any;
}

// All-lanes sum over one wavefront, as a butterfly: every step adds a lane's
// partner value, with partners paired by an involution, so paired lanes add
// the same two numbers (IEEE addition is commutative) and after six steps
// every lane holds the bitwise-identical total.  Steps 1-4 pair lanes inside a
// 16-lane row with DPP moves (quad xor 1, quad xor 2, half-row mirror, row
// mirror); steps 5-6 exchange rows with the gfx950 permlane swaps, whose two
// results hold the lane's own value and its partner's (in either order).
// Requires every lane of the wavefront to be active.
template <int CTRL>
__device__ __forceinline__ double dpp_add_partner(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double join_halves(int lo, int hi) {
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_sum(double v) {
    v = dpp_add_partner<0xB1>(v);    // quad_perm [1,0,3,2]
    v = dpp_add_partner<0x4E>(v);    // quad_perm [2,3,0,1]
    v = dpp_add_partner<0x141>(v);   // row_half_mirror
    v = dpp_add_partner<0x140>(v);   // row_mirror
    long long b = __double_as_longlong(v);
    auto l16 = __builtin_amdgcn_permlane16_swap((int)b, (int)b, false, false);
    auto h16 = __builtin_amdgcn_permlane16_swap((int)(b >> 32), (int)(b >> 32), false, false);
    v = join_halves(l16[0], h16[0]) + join_halves(l16[1], h16[1]);
    b = __double_as_longlong(v);
    auto l32 = __builtin_amdgcn_permlane32_swap((int)b, (int)b, false, false);
    auto h32 = __builtin_amdgcn_permlane32_swap((int)(b >> 32), (int)(b >> 32), false, false);
    return join_halves(l32[0], h32[0]) + join_halves(l32[1], h32[1]);
}

// Value slots of the sliced ELL: the 36 entries of lane l's block are stored in
// 16-byte groups of G = 16 / sizeof(T) consecutive entries (FP16: 8-byte groups
// of 4, the same index pattern as FP32), lanes interleaved group by group
// ([slot][36 / G][64 lanes][G]), so one vector load per lane brings G entries
// and a wavefront's load covers 512 B - 1 KB of consecutive bytes.
template <typename T>
__device__ __forceinline__ void load_block(const T* __restrict__ slot, int lane, T (&av)[36]) {
    constexpr int G = sizeof(T) == 2 ? 4 : 16 / sizeof(T);   // FP16: groups of 4 (8-byte loads, 36 = 9 x 4)
    typedef T V __attribute__((ext_vector_type(G)));
#pragma unroll
    for (int g = 0; g < 36 / G; ++g) {
        const V v = *reinterpret_cast<const V*>(slot + ((int64_t)g * ROWS_PER_WG + lane) * G);
#pragma unroll
        for (int t = 0; t < G; ++t) av[g * G + t] = v[t];
    }
}

// Sum over the whole workgroup, result in every thread.  `lds` holds >= NW*NV.
template <int NV, int NW>
__device__ __forceinline__ void group_sum(double (&v)[NV], double* lds) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (NW == 1) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) lds[wave * NV + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += lds[w * NV + i];
        v[i] = s;
    }
}

// Totals over the slices of `sys` of one or two per-slice partial arrays (NR
// values per slice), identical in every thread of every workgroup of the structure.
template <int NR, int NW>
__device__ __forceinline__ void system_totals(const Slices& sl, int sys, const double* a, const double* b,
                                              double (&ta)[NR], double (&tb)[NR], double* lds) {
    const int s0 = sl.sys_off[sys], s1 = sl.sys_off[sys + 1];
    double v[2 * NR];
#pragma unroll
    for (int i = 0; i < 2 * NR; ++i) v[i] = 0.0;
    for (int w = s0 + (int)threadIdx.x; w < s1; w += NW * 64) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            v[h] += a[(int64_t)w * NR + h];
            if (b) v[NR + h] += b[(int64_t)w * NR + h];
        }
    }
    group_sum<2 * NR, NW>(v, lds);
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        ta[h] = v[h];
        tb[h] = v[NR + h];
    }
}

// write this slice's partial sums (thread 0)
template <int NV, int NW>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* lds, double* dst, int slice) {
    group_sum<NV, NW>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) dst[(int64_t)slice * NV + i] = v[i];
    }
}

// Waves-per-slice variants built for VT and NR: the cross-wave LDS buffer stays
// <= 24 KB, and more waves only pay while one slot group fits the registers.
template <typename VT, int NR, int WPS>
struct WpsFits {
    static constexpr bool value =
        WPS == 1 || ((size_t)(WPS - 1) * 6 * NR * 64 * sizeof(VT) <= 24576 && NR * sizeof(VT) <= 16);
};

struct Workspace {
    double *b64, *x64, *xbest;
    void *xlo, *r[2], *q[2], *p;
    SysState st;
    Partials P;
};

enum { WS_B, WS_X, WS_XBEST, WS_XLO, WS_R0, WS_R1, WS_Q0, WS_Q1, WS_P, WS_PART, WS_SD, WS_SI, WS_ADJ, WS_N };
static_assert(WS_ADJ == WS_ADJ_SLOT, "adjoint scratch slot");

// host helpers (jsso_solve.hip)
int get_workspace(Plan* pl, int nrhs, Workspace& w);
int check_status(Plan* p, hipStream_t s);
Slices make_slices(Plan* pl, int base, int count);
bool graphs_enabled();
int ensure_flag_buffers(Plan* pl, size_t nsys);

template <typename F>
hipGraphExec_t capture_graph(Plan* pl, CgGraph& g, const void* const (&key)[CgGraph::NKEY], int wps,
                                    int steps, double tol, F&& enqueue) {
    bool same = g.exec && g.wps == wps && g.steps == steps && g.tol == tol;
    for (int i = 0; same && i < CgGraph::NKEY; ++i) same = g.key[i] == key[i];
    if (same) return g.exec;
    if (g.exec) {
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    if (!pl->capture_stream && hipStreamCreateWithFlags(&pl->capture_stream, hipStreamNonBlocking) != hipSuccess) {
        pl->graphs_ok = false;
        (void)hipGetLastError();
        return nullptr;
    }
    hipStream_t cs = pl->capture_stream;
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        enqueue(cs);
        ok = hipStreamEndCapture(cs, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        g.exec = nullptr;
        pl->graphs_ok = false;
        (void)hipGetLastError();
        return nullptr;
    }
    g.wps = wps;
    g.steps = steps;
    g.tol = tol;
    for (int i = 0; i < CgGraph::NKEY; ++i) g.key[i] = key[i];
    return g.exec;
}

}  // namespace jsso
