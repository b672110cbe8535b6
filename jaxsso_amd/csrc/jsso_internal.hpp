// Internal declarations of libjsso_hip.so: the plan object, error handling and
// the kernel launchers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/jsso.h"

namespace jsso {

void set_error(const char* fmt, ...);

#define JSSO_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::jsso::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call,               \
                              hipGetErrorString(e_));                                   \
            return JSSO_EHIP;                                                           \
        }                                                                               \
    } while (0)

#define JSSO_REQUIRE(cond, code, ...)     \
    do {                                  \
        if (!(cond)) {                    \
            ::jsso::set_error(__VA_ARGS__); \
            return (code);                \
        }                                 \
    } while (0)

constexpr int BDIR = 12;   // beam adjoint directions: 6 coordinates + 6 properties
constexpr int QDIR = 17;   // quad adjoint directions: 12 coordinates + 5 properties
constexpr int WS_ADJ_SLOT = 12;  // Plan::ws slot of the adjoint's gradient scratch (solver slots are below it)
constexpr int WS_FU_SLOT = 13;   // Plan::ws slot of the f.u reduction's per-slice partials
constexpr int WS_DEFECT_SLOT = 14;   // Plan::ws slot of the non-symmetric defect correction's vectors
// status bits of an assembly (d_status): 1 a constrained diagonal block is not
// positive definite; 2 K is not symmetric; 8 because of a quad with kx_mod !=
// ky_mod; 16 because of a beam-column parallel to global Y
constexpr int STATUS_NONSYM = 2, STATUS_ORTHO_QUAD = 8, STATUS_Y_BEAM = 16;
// non-symmetric plans up to this many DOF take the dense LU by default
constexpr int64_t NONSYM_DENSE_MAX_DOF = 16384;
// fixed device scratch of jsso_adjoint_grad_noalloc: 8 Mi doubles = 64 MiB per GPU
// ((12 per beam + 17 per quad) x load cases: e.g. 490k quads with one load case)
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 23;
constexpr int ROWS_PER_WG = 64;   // block rows per solver slice (one wavefront in the SpMV)

// Grids of element- or block-wise kernels: 8 * ceil(n / 8) workgroups, index
// remapped so that each XCD (workgroups go round-robin to the 8 XCDs) runs one
// contiguous eighth of the work.  Outputs next to each other read the same
// inputs (a mesh neighbourhood, a block row), which then stay in one L2 instead
// of being fetched from HBM by every XCD.
static inline dim3 xcd_grid(int64_t nwg) { return dim3((unsigned)(8 * ((nwg + 7) / 8))); }
__device__ __forceinline__ int64_t xcd_wg() {
    const int64_t b = blockIdx.x;
    return (b & 7) * (int64_t)(gridDim.x >> 3) + (b >> 3);
}
// batched-solver variants: matrix type MT x vector type VT
enum { PS_F32 = 0, PS_F64 = 1, PS_F32_V64 = 2 };
template <typename MT, typename VT>
constexpr int prec_slot() { return sizeof(MT) == 8 ? PS_F64 : (sizeof(VT) == 4 ? PS_F32 : PS_F32_V64); }
constexpr int SOLVER_WG = 6 * ROWS_PER_WG;   // threads of the scalar-row solver kernels

// Symmetric sliced ELL of a block pattern (built on the host, see build_ell).
struct EllPattern {
    int32_t n_wg = 0;                                  // slices
    std::vector<int32_t> wg_sys, wg_row0, wg_row1;     // per slice: structure, rows [row0, row1)
    std::vector<int32_t> sys_wg_off;                   // per structure: first slice
    std::vector<int32_t> ell_col, ell_vpos;            // per column position: block column, value position
    std::vector<int32_t> ell_src;                      // per value position: block of the CSR, -1 = padding
    std::vector<int32_t> meta;                         // per slice: 12 ints
    std::vector<int32_t> kref;                         // per block of the CSR: 2*vpos (+1 transposed), -1 diagonal
    int64_t ell_slots = 0;                             // value positions incl. the trailing zero slot
};
int build_ell(int32_t n_sys, const int32_t* node_off, const std::vector<int32_t>& rowptr,
              const std::vector<int32_t>& colind, EllPattern& out);

// Device buffer owned by the plan.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// An instantiated HIP graph of one chunk of ordinary CG iterations, valid for
// the buffers and kernel variant it was captured with.
struct CgGraph {
    static constexpr int NKEY = 16;
    hipGraphExec_t exec = nullptr;
    int wps = 0, steps = 0;
    double tol = 0.0;
    const void* key[NKEY] = {};
};

struct AmgState;   // smoothed-aggregation multigrid hierarchy (jsso_amg.hip)

struct Plan {
    // sizes
    int32_t n_sys = 0, n_nodes = 0, n_beam = 0, n_quad = 0, n_cons = 0;
    int64_t nnzb = 0;      // blocks of K (block CSR incl. diagonal)
    int64_t nnzb_off = 0;  // off-diagonal blocks (solver matrix)
    int64_t n_contrib = 0; // element blocks = 4*n_beam + 16*n_quad
    // host copies (symbolic analysis runs on the host; uploaded on first use)
    std::vector<int32_t> h_node_off, h_rowptr, h_colind;
    std::vector<int32_t> h_browof, h_diag, h_cptr, h_csrc, h_nptr, h_nsrc;
    std::vector<int32_t> h_cnct_beam, h_cnct_quad, h_wg_sys, h_wg_row0, h_wg_row1, h_sys_wg_off;
    // symmetric sliced ELL of the off-diagonal blocks (upper blocks stored once)
    std::vector<int32_t> h_ell_col;    // per column position: block column
    std::vector<int32_t> h_ell_vpos;   // per column position: value position (slot * 64 + lane) of its block
    std::vector<int32_t> h_slice_meta; // per slice: 12 ints, see build_plan
    std::vector<int32_t> h_ell_src;    // per value position: source block in the block CSR, -1 = padding
    std::vector<int32_t> h_kref;       // per block of the CSR: 2 * value position (+1 when transposed), -1 diagonal
    std::vector<uint8_t> h_fixed;
    bool on_device = false;
    // pattern (device)
    int32_t *d_rowptr = nullptr, *d_colind = nullptr, *d_browof = nullptr, *d_diag = nullptr;
    int32_t *d_cptr = nullptr, *d_csrc = nullptr;       // block -> element-block contributions
    int32_t *d_nptr = nullptr, *d_nsrc = nullptr;       // node -> adjoint output offsets
    int32_t *d_cnct_beam = nullptr, *d_cnct_quad = nullptr;
    uint8_t* d_fixed = nullptr;                           // per DOF
    int32_t *d_wg_sys = nullptr, *d_wg_row0 = nullptr, *d_wg_row1 = nullptr;
    int32_t* d_sys_wg_off = nullptr;
    int32_t n_wg = 0;
    bool slice_spread = false;   // several structures, one of >= 64 slices: slice runs spread over XCDs
    // sliced ELL: slice = workgroup above; values [value slot][36][64 lanes]
    int32_t *d_ell_vpos = nullptr, *d_ell_col = nullptr, *d_slice_meta = nullptr;
    int32_t* d_ell_src = nullptr;    // per value position: source block or -1
    int64_t ell_slots = 0;           // value positions (value slots x 64), incl. one all-zero slot
    // waves per slice of the batched CG kernel, per [FP32 | FP64][load cases]:
    // 0 = not tuned yet (the next solve measures the variants and stores the fastest)
    int32_t cg_wps[3][9] = {};
    // waves per slice of the AMG level-0 kernels, per precision slot x load cases
    // (0 = not tuned yet: the next AMG solve times the variants on the live plan)
    int32_t amg_wps[3][9] = {};
    int32_t ke_epw = 0;            // MITC-4 elements per workgroup of the assembly (0 = tune at the next one)
    CgGraph cg_graph[3][9][2];       // captured CG chunks, per precision slot (prec_slot) x load cases x
                                     // [iterations | replacement]
    hipStream_t capture_stream = nullptr;
    bool graphs_ok = true;           // false after a failed capture: launch eagerly
    int32_t* h_flags = nullptr;      // pinned host copies of the convergence flags, two chunks' worth
    double* h_rel = nullptr;         // and of the systems' relative FP64 residuals (AMG chunk schedule)
    size_t h_flags_count = 0;
    hipEvent_t flag_ev[2] = {nullptr, nullptr};
    double* d_Ke64 = nullptr;        // scaled off-diagonal blocks, sliced ELL, FP64
    float* d_Ke32 = nullptr;         // same, FP32
    uint16_t* d_Ke16 = nullptr;      // same, FP16 bits (the AMG smoothing products)
    // numeric state (device)
    double* d_ke = nullptr;       // element blocks, 36 doubles each
    double* d_K = nullptr;        // K, 36 per block (unconstrained)
    double* d_Linv = nullptr;     // inverse Cholesky factor of each constrained diagonal block
    double* d_Minv = nullptr;     // block-Jacobi inverse (plain path)
    double* d_Kc = nullptr;       // constrained K incl. diagonal (plain path)
    int32_t* d_status = nullptr;  // device error flags
    int32_t* d_tpos = nullptr;    // per block of the CSR: position of its transposed block
    std::vector<int32_t> h_tpos;
    int32_t nonsym_route = 0;     // non-symmetric K: 0 auto, 1 dense LU, 2 defect correction
    bool sym_part_solve = false;  // set while the defect correction solves with the symmetric part
    int32_t status_cache = -1;    // host copy of *d_status once read after an assembly (-1: not read)
    bool assembled = false;
    bool full_valid = false;      // d_K / d_Kc hold the current assembly
    double* d_crds = nullptr;     // node coordinates of the current assembly (for the AMG null space)
    // preconditioner of the batched solves (JSSO_PC_*) and the AMG hierarchy
    int32_t precond = JSSO_PC_AMG;
    AmgState* amg = nullptr;
    bool amg_valid = false;       // the AMG numeric set-up matches the current assembly
    int32_t amg_reuse = 0;        // jsso_plan_amg_reuse mode
    int64_t amg_full = 0, amg_light = 0;   // numeric set-ups run: full, and refreshes of a kept hierarchy
    // dense direct solve (jsso_dense_solve): N x N factor + right-hand sides, grown on demand
    double* d_dense = nullptr;
    size_t dense_bytes = 0;
    // solver workspace, grown on demand
    std::vector<DBuf> ws;
    int ws_nrhs = 0;
    std::vector<void*> allocs;    // everything to free
};

template <typename T>
int dalloc(Plan* p, T** ptr, size_t count) {
    *ptr = nullptr;
    if (count == 0) return JSSO_OK;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        return JSSO_ENOMEM;
    }
    p->allocs.push_back(q);
    *ptr = static_cast<T*>(q);
    return JSSO_OK;
}

// kernels launched from the API layer
int launch_ke_beam(int32_t n, const double* crds_e, const int32_t* cnct, const double* crds_nodes,
                   const double* props, double* ke, int32_t* flags, int32_t* status, hipStream_t s);
int launch_ke_quad(int32_t n, const double* crds_e, const int32_t* cnct, const double* crds_nodes,
                   const double* props, double* ke, hipStream_t s, int epw = 16, int32_t* status = nullptr);
int tune_ke_quad(Plan* p, const double* crds, const double* props, double* ke, hipStream_t s, int32_t& out);
int launch_beam_local(int32_t n, const double* crds, const double* props, double* T, double* kl, hipStream_t s);
int launch_quad_local(int32_t n, const double* crds, const double* props, double* xy, double* normals, double* T,
                      double* kl, hipStream_t s);
int launch_quad_gauss(int32_t n, const double* xy, const double* props, double r, double s_, double* J, double* Bk,
                      double* Bg, double* Bm, double* Cb, double* Cs, double* Cm, hipStream_t s);
int launch_assemble(Plan* p, hipStream_t s);
int launch_csr_spmv(int32_t n_rows, const int32_t* rowptr, const int32_t* colind, const double* vals,
                    const double* x, const double* c, double* y, hipStream_t s);
int launch_full_blocks(Plan* p, hipStream_t s);
int launch_adjoint(Plan* p, const double* crds, const double* prop_beam, const double* prop_quad,
                   const double* u, const double* lam, double lam_scale, int nrhs, double* gbuf,
                   double* dprop_beam, double* dprop_quad, hipStream_t s);
int launch_gather_grad(Plan* p, const double* gbuf, int nrhs, double* dcrds, hipStream_t s);
int noalloc_scratch(double** ptr, size_t* bytes);
int launch_spmv_full(Plan* p, const double* K, const double* x, double* y, hipStream_t s);
int solve_batched(Plan* p, const double* f, double* u, int nrhs, int slot, double tol,
                  int maxit, jsso_info* info, hipStream_t s);
int timing_enable(int every);
int timing_read(int slot, double* avg_ms, int64_t* samples, int reset);
void timing_begin(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev);
void timing_begin_batch(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev, int launches);
void timing_begin_slot(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>*& ev, int slot);
bool timing_sample_chunk();
std::pair<hipEvent_t, hipEvent_t> timing_pair();
void timing_settle(std::pair<hipEvent_t, hipEvent_t> e, int slot, bool count);
void timing_end(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>* ev);
void timing_collect();
int launch_compliance(Plan* p, const double* f, const double* u, int nrhs, double scale, double* C, hipStream_t s);
int launch_defect_residual(Plan* p, const double* K, const double* f, const double* u, int nrhs, bool transpose,
                           double* r, hipStream_t s);
int launch_axpy(int64_t n, const double* x, double* y, hipStream_t s);
int amg_solve(Plan* p, const double* f, double* u, int nrhs, int slot, double tol, int maxit, jsso_info* info,
              hipStream_t s);
int amg_levels(Plan* p, int32_t* h_nodes, int32_t cap, int32_t* h_levels);
void amg_destroy(Plan* p);
int dense_solve(Plan* p, const double* K, const double* f, double* u, int nrhs, bool lu, bool transpose,
                hipStream_t s);
int assembly_status(Plan* p, hipStream_t s, int32_t* status);
int solve_plain(Plan* p, const double* f, double* u, int nrhs, double tol, int maxit,
                jsso_info* info, hipStream_t s);

}  // namespace jsso
