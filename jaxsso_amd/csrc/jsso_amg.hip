// Smoothed-aggregation algebraic multigrid (SA-AMG) preconditioned CG for the
// batched solves (no counterpart in the reference, which hands the system to
// SuperLU: this is the "solver robustness beyond block-Jacobi" path).
//
// Every level works on a block-Jacobi-scaled operator with unit 6x6 diagonal
// blocks, stored as the same symmetric sliced ELL as the fine system (only the
// off-diagonal upper blocks; the diagonal is implicit).  Level 0 is the solver
// matrix itself, A0 = L^-1 Kc L^-T.
//
// Hierarchy (Vanek, Mandel & Brezina 1996):
//   symbolic, on the host, once per plan: greedy aggregation of each level's
//   node graph (three passes: root nodes whose neighbours are all free take
//   them; leftovers join a neighbouring aggregate; the rest form new ones),
//   structure by structure; the block patterns of the prolongator P, of
//   W = A P and of the Galerkin operator P^T A P, with the list of products
//   that make each block;
//   numeric, on the device, after every assembly: the near-null space (the six
//   rigid-body modes of every node, in the scaled variables: L_i^T B_i) is
//   orthonormalised per aggregate (tentative prolongator Q, coarse null space
//   R); P = (I - omega A) Q with omega = 4 / (3 rho), rho from a power
//   iteration; A_c = P^T (A P); the coarse diagonal blocks are factored
//   (Cholesky) and A_c and P rescaled so the coarse level again has unit
//   diagonal blocks.  A structure stops coarsening when it has <= AMG_MAXC
//   nodes; its last level is inverted densely.
//
// V-cycle (symmetric, so M is SPD): one damped-Jacobi pre- and post-smoothing
// step z = r / theta with theta = 0.62 rho (the optimal one-step Chebyshev
// smoother on [rho/30, rho] would take 0.517 rho, which needs rho to within a
// few per cent: on heterogeneous SIMP shells a short power iteration reaches
// only 0.83 rho, and the smoother must stay convergent, 2 theta > rho, there),
// restriction P^T, the coarse correction, and prolongation P.  The V-cycle runs in FP32; level 0 runs in the solve's
// precision.  The outer CG keeps the batched path's reliable-update scheme:
// FP32 (or FP64) recurrences, the FP64 true residual every `replace_every`
// iterations, convergence and the best iterate decided there.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <type_traits>

#include <hip/hip_ext.h>

#include "solver_kernels.hpp"

namespace jsso {

constexpr int AMG_MAXC = 26;       // nodes of a structure's directly solved coarsest level (<= 156 DOF)
constexpr int AMG_QR_RPL = 5;      // null-space rows per lane in the per-aggregate QR
constexpr int AMG_MAX_AGG = AMG_QR_RPL * 64 / 6;   // nodes per aggregate (53)
constexpr int AMG_POWER_STEPS = 12;
constexpr int AMG_POWER_STEPS_WARM = 4;   // level-0 steps of a kept hierarchy's refresh (warm start)
constexpr double AMG_THETA = 0.62;        // level-0 smoothing step z = r / (AMG_THETA rho)
// coarse levels: degree-2 Chebyshev smoothing on [AMG_CHEB_LO hi, hi], hi = AMG_CHEB_HI rho
constexpr double AMG_CHEB_HI = 1.1;
constexpr double AMG_CHEB_LO = 0.1;
constexpr int AMG_STALL_WINDOW = 256;     // iterations without a 10% gain that end a solve (floor)
// attainable-accuracy floor (k_decide): within 16 x eps (rho ||x|| + ||b||) / ||b||
// and no 10% gain for 24 iterations (three replacement chunks).  The rounding
// error of a computed FP64 residual is bounded by ~n eps (|b| + |A||x|) per row
// (n ~ 54 terms per row of a shell), typically ~sqrt(n) eps, and || |A||x| || can
// exceed ||A|| ||x||: 16 covers that spread; the 10% window still has to pass.
constexpr double AMG_FLOOR_FACTOR = 16.0;
constexpr int AMG_FLOOR_WINDOW = 24;
constexpr double AMG_RHO_SAFETY = 1.12;   // the power iteration approaches rho from below
constexpr int AMG_MAXLEVELS = 12;
// output blocks up to which the set-up products run one lane per block entry
// (measured crossovers: P and W between 4 k and 25 k blocks, A_c above 75 k)
constexpr int64_t AMG_ENTRY_PW_MAX = 8192;
constexpr int64_t AMG_ENTRY_C_MAX = 131072;
// JSSO_AMG_ENTRY_PRODUCTS=0 / 1 takes the thread-per-block / lane-per-entry form
// on every level (the two give the same bits; the tests compare them)
static bool amg_entry_form(int64_t blocks, int64_t limit) {
    const char* e = getenv("JSSO_AMG_ENTRY_PRODUCTS");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return true;
    return blocks <= limit;
}
constexpr int AMG_CUT_GRAPHS = 7;         // captured chunk lengths below the replacement interval: 2 .. 14

// ---------------------------------------------------------------------------
// host: symbolic hierarchy
// ---------------------------------------------------------------------------
struct AmgHostLevel {
    int32_t n = 0;                                 // nodes (all structures, packed)
    std::vector<int32_t> node_off;                 // per structure: first node (n_sys + 1)
    std::vector<int32_t> rowptr, colind;           // block pattern incl. the diagonal (level >= 1)
    EllPattern ell;                                // level >= 1 (level 0: the plan's)
    std::vector<int32_t> node_sys;                 // per node: structure
    std::vector<uint8_t> cont;                     // per structure: 1 = has a coarser level (P), 0 = not here / terminal
    std::vector<uint8_t> term;                     // per structure: 1 = solved densely at this level
    // to the next level (nodes of continuing structures)
    std::vector<int32_t> agg;                      // per node: aggregate (= coarse node) or -1
    std::vector<int32_t> agg_ptr, agg_nodes;       // per aggregate: member nodes, in node order
    std::vector<int32_t> p_ptr, p_col, p_row;      // P blocks by fine row: coarse column, fine row
    std::vector<int32_t> p_cptr, p_cj, p_caref;    // per P block: products (node j, A block ref or -1 = diagonal)
    std::vector<int32_t> pt_ptr, pt_blk, pt_row;   // per coarse node: the P blocks of its column, their fine rows
    std::vector<int32_t> p_tpos;                   // per P block: its position in the pt lists
    std::vector<int32_t> w_ptr, w_col, w_row;      // W = A P blocks by fine row
    std::vector<int32_t> w_cptr, w_caref, w_cp;    // per W block: products (A block ref or -1, P block)
    std::vector<int32_t> c_cptr, c_p, c_w;         // per block of the next level's CSR: products (P block, W block)
};

struct AmgLevelDev {
    int32_t n = 0, n_wg = 0;
    // pattern of this level (level >= 1)
    int32_t *d_rowptr = nullptr, *d_colind = nullptr, *d_kref = nullptr, *d_browof = nullptr;
    int32_t *d_wg_sys = nullptr, *d_wg_row0 = nullptr, *d_wg_row1 = nullptr, *d_sys_wg_off = nullptr;
    int32_t *d_ell_col = nullptr, *d_ell_vpos = nullptr, *d_meta = nullptr, *d_ell_src = nullptr;
    int64_t ell_slots = 0;
    double* d_A64 = nullptr;
    float* d_A32 = nullptr;
    int32_t* d_node_sys = nullptr;
    uint8_t* d_cont = nullptr;   // per structure
    // smoothing / power iteration
    double *d_rho = nullptr, *d_omega = nullptr;   // per structure
    float* d_theta = nullptr;                      // per structure (coarse levels: the Chebyshev centre)
    float* d_cheb = nullptr;                       // per structure, coarse levels: Chebyshev coefficient c
    double *d_pv[2] = {nullptr, nullptr};          // power iteration vectors (6 n)
    double* d_ppart = nullptr;                     // per slice, per step parity: |y|^2
    // transfer to the next level
    int32_t n_agg = 0;
    int32_t *d_agg = nullptr, *d_agg_ptr = nullptr, *d_agg_nodes = nullptr;
    double *d_B = nullptr, *d_Q = nullptr, *d_R = nullptr;   // 36 per node / per aggregate
    int64_t nP = 0, nW = 0, nPc = 0, nWc = 0, nCc = 0;
    int32_t *d_p_ptr = nullptr, *d_p_col = nullptr, *d_p_row = nullptr;
    int32_t *d_p_cptr = nullptr, *d_p_cj = nullptr, *d_p_caref = nullptr;
    int32_t *d_pt_ptr = nullptr, *d_pt_row = nullptr, *d_p_tpos = nullptr;
    int32_t *d_w_row = nullptr, *d_w_cptr = nullptr, *d_w_caref = nullptr, *d_w_cp = nullptr;
    int32_t *d_c_cptr = nullptr, *d_c_p = nullptr, *d_c_w = nullptr;
    float *d_Pf = nullptr, *d_Wf = nullptr;
    _Float16 *d_P16 = nullptr, *d_PT16 = nullptr;   // FP16, by fine row / transposed by coarse node
    // V-cycle vectors (level >= 1), 6 n x nrhs FP32
    float *d_b = nullptr, *d_t = nullptr, *d_r = nullptr, *d_x = nullptr;
};

struct AmgState {
    int nl = 0;                       // levels
    std::vector<AmgHostLevel> host;
    std::vector<AmgLevelDev> dev;
    // next level's Galerkin storage lives with the coarse level: A_c (CSR order), Lc, Lc^-1
    std::vector<double*> d_Ac, d_Lc, d_Lci;
    std::vector<int32_t*> d_cdiag;
    // dense coarsest solves: per structure {level, first node, nodes, inverse offset}
    std::vector<int32_t> h_term;      // 4 per structure
    int32_t* d_term = nullptr;
    int32_t* d_node_off0 = nullptr;   // level-0 first node of every structure (the null space's origin)
    std::vector<std::vector<int32_t>> term_list;   // per level: structures solved there
    std::vector<int32_t*> d_term_list;
    // A structure whose dense level T is >= 2 runs its level T - 1 and the dense
    // solve in one kernel (k_mg_tail): per level, the structures fused there, the
    // dense solves left to k_mg_direct, and per-structure flags for the level
    // kernels (smoothed here by the level kernels; restriction from here done by
    // the fused kernel)
    std::vector<std::vector<int32_t>> tail_list, direct_list;
    std::vector<int32_t*> d_tail_list, d_direct_list, d_node_off;
    // per right-hand-side count (TAIL_NR) and level: the fused structures whose
    // vectors fit in LDS [0] and the others [1]
    std::vector<int32_t> tail_n[5][2];
    std::vector<int32_t*> d_tail_split[5][2];
    std::vector<uint8_t*> d_smooth, d_rskip;
    float* d_inv = nullptr;           // dense inverses, column-major, FP32
    float* d_Anat = nullptr;         // set-up scratch: a level's blocks in natural layout [position][36], FP32
    int64_t inv_total = 0;
    int vec_nrhs = 0;                 // V-cycle vectors sized for this many load cases
    int wps0 = 4;                     // waves per slice of the level-0 kernels
    // level-0 extra vectors (VT-sized, 8 bytes per value)
    void *d_z = nullptr, *d_t0 = nullptr, *d_p1 = nullptr;
    size_t z_bytes = 0;
    uint64_t gen = 1;                 // bumped whenever a buffer the graphs point at changes
    std::vector<void*> allocs;
    CgGraph graph[3][9][2];       // per precision slot (prec_slot) x load cases x [iterations | replacement]
    // chunks cut short at a predicted convergence: one pair per even length 2, 4, ... 2 * AMG_CUT_GRAPHS
    CgGraph graph_cut[3][9][AMG_CUT_GRAPHS][2];
    bool built = false;           // a full numeric set-up has run (a kept hierarchy can be reused)
    int power_next = AMG_POWER_STEPS;   // level 0: index of the next power-iteration step (parities of
                                        // the vectors and partials continue across refreshes)
};

// Greedy aggregation of one structure's nodes [n0, n1) of a level.  `active`
// excludes nodes with no unknowns (every DOF constrained at level 0).
static void aggregate_structure(const std::vector<int32_t>& rowptr, const std::vector<int32_t>& colind,
                                const std::vector<uint8_t>& active, int n0, int n1, std::vector<int32_t>& agg,
                                int32_t& na) {
    std::vector<int32_t> size;   // nodes per aggregate of this structure (index a - first)
    const int32_t first = na;
    auto nbrs = [&](int i, auto&& f) {
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int j = colind[k];
            if (j != i && active[j]) f(j);
        }
    };
    // pass 1: a node whose active neighbours are all free roots a new aggregate with them
    for (int i = n0; i < n1; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        bool free = true;
        int deg = 0;
        nbrs(i, [&](int j) {
            ++deg;
            free &= agg[j] < 0;
        });
        if (deg == 0 || !free || deg + 1 > AMG_MAX_AGG) continue;
        agg[i] = na;
        nbrs(i, [&](int j) { agg[j] = na; });
        size.push_back(deg + 1);
        ++na;
    }
    // pass 2: leftovers join the aggregate of their first aggregated neighbour
    std::vector<int32_t> pass2(n1 - n0, -1);
    for (int i = n0; i < n1; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        int a = -1;
        nbrs(i, [&](int j) {
            if (a < 0 && agg[j] >= 0 && size[agg[j] - first] < AMG_MAX_AGG) a = agg[j];
        });
        pass2[i - n0] = a;
        if (a >= 0) ++size[a - first];
    }
    for (int i = n0; i < n1; ++i)
        if (pass2[i - n0] >= 0) agg[i] = pass2[i - n0];
    // pass 3: the rest form new aggregates with their free neighbours
    for (int i = n0; i < n1; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        agg[i] = na;
        int cnt = 1;
        nbrs(i, [&](int j) {
            if (agg[j] < 0 && cnt < AMG_MAX_AGG) {
                agg[j] = na;
                ++cnt;
            }
        });
        size.push_back(cnt);
        ++na;
    }
}

// Symbolic hierarchy.  Level 0 is the plan's pattern; each further level is
// built while some structure still has more than AMG_MAXC nodes.
static int amg_build_host(Plan* pl, AmgState& A) {
    const int ns = pl->n_sys;
    A.host.clear();
    A.host.reserve(AMG_MAXLEVELS + 1);   // references to levels stay valid while levels are added
    A.host.emplace_back();
    {
        AmgHostLevel& L0 = A.host[0];
        L0.n = pl->n_nodes;
        L0.node_off = pl->h_node_off;
        L0.rowptr = pl->h_rowptr;
        L0.colind = pl->h_colind;
        L0.ell.kref = pl->h_kref;
    }
    for (int l = 0;; ++l) {
        JSSO_REQUIRE(l < AMG_MAXLEVELS, JSSO_EUNSUPPORTED, "AMG: coarsening did not reach %d nodes in %d levels",
                     AMG_MAXC, AMG_MAXLEVELS);
        AmgHostLevel& F = A.host[l];
        const int n = F.n;
        F.node_sys.assign(n, 0);
        for (int s = 0; s < ns; ++s)
            for (int i = F.node_off[s]; i < F.node_off[s + 1]; ++i) F.node_sys[i] = s;
        F.cont.assign(ns, 0);
        F.term.assign(ns, 0);
        bool any = false;
        for (int s = 0; s < ns; ++s) {
            const int cnt = F.node_off[s + 1] - F.node_off[s];
            if (cnt == 0) continue;   // finished at a finer level
            if (l > 0 && cnt <= AMG_MAXC) {
                F.term[s] = 1;
            } else {
                F.cont[s] = 1;
                any = true;
            }
        }
        if (!any) break;
        // aggregation
        std::vector<uint8_t> active(n, 1);
        if (l == 0)
            for (int i = 0; i < n; ++i) {
                bool all = true;
                for (int c = 0; c < 6; ++c) all &= pl->h_fixed[6 * (size_t)i + c] != 0;
                active[i] = all ? 0 : 1;
            }
        F.agg.assign(n, -1);
        std::vector<int32_t> cnode_off(ns + 1, 0);
        int32_t na = 0;
        for (int s = 0; s < ns; ++s) {
            cnode_off[s] = na;
            if (!F.cont[s]) continue;
            const int32_t before = na;
            aggregate_structure(F.rowptr, F.colind, active, F.node_off[s], F.node_off[s + 1], F.agg, na);
            const int cnt = F.node_off[s + 1] - F.node_off[s];
            JSSO_REQUIRE(l == 0 || na - before < cnt, JSSO_EUNSUPPORTED,
                         "AMG: aggregation of structure %d made no progress at level %d", s, l);
        }
        cnode_off[ns] = na;
        // aggregate members in node order (the first is the origin of its rigid-body modes)
        F.agg_ptr.assign(na + 1, 0);
        for (int i = 0; i < n; ++i)
            if (F.agg[i] >= 0) F.agg_ptr[F.agg[i] + 1]++;
        for (int a = 0; a < na; ++a) F.agg_ptr[a + 1] += F.agg_ptr[a];
        for (int a = 0; a < na; ++a)
            JSSO_REQUIRE(F.agg_ptr[a + 1] - F.agg_ptr[a] <= AMG_MAX_AGG, JSSO_EUNSUPPORTED,
                         "AMG: aggregate larger than %d nodes", AMG_MAX_AGG);
        F.agg_nodes.assign(F.agg_ptr[na], 0);
        {
            std::vector<int32_t> fill(F.agg_ptr.begin(), F.agg_ptr.end() - 1);
            for (int i = 0; i < n; ++i)
                if (F.agg[i] >= 0) F.agg_nodes[fill[F.agg[i]]++] = i;
        }
        // P pattern and products: P(i, J) = sum over j in {i} u N(i), agg(j) = J
        F.p_ptr.assign(n + 1, 0);
        F.p_col.clear();
        F.p_row.clear();
        F.p_cptr.assign(1, 0);
        F.p_cj.clear();
        F.p_caref.clear();
        std::vector<std::pair<int32_t, int32_t>> prod;   // (J, k)
        for (int i = 0; i < n; ++i) {
            prod.clear();
            if (active[i] && F.cont[F.node_sys[i]]) {
                for (int k = F.rowptr[i]; k < F.rowptr[i + 1]; ++k) {
                    const int j = F.colind[k];
                    if (F.agg[j] >= 0) prod.emplace_back(F.agg[j], k);
                }
            }
            std::sort(prod.begin(), prod.end());
            for (size_t t = 0; t < prod.size();) {
                const int J = prod[t].first;
                F.p_col.push_back(J);
                F.p_row.push_back(i);
                for (; t < prod.size() && prod[t].first == J; ++t) {
                    const int k = prod[t].second;
                    const int j = F.colind[k];
                    F.p_cj.push_back(j);
                    F.p_caref.push_back(j == i ? -1 : F.ell.kref[k]);
                }
                F.p_cptr.push_back((int32_t)F.p_cj.size());
            }
            F.p_ptr[i + 1] = (int32_t)F.p_col.size();
        }
        const int64_t nP = F.p_col.size();
        JSSO_REQUIRE(nP < INT32_MAX / 36, JSSO_EUNSUPPORTED, "AMG: prolongator too large");
        // P^T: per coarse node, its P blocks in fine-row order
        F.pt_ptr.assign(na + 1, 0);
        for (int64_t b = 0; b < nP; ++b) F.pt_ptr[F.p_col[b] + 1]++;
        for (int a = 0; a < na; ++a) F.pt_ptr[a + 1] += F.pt_ptr[a];
        F.pt_blk.assign(nP, 0);
        {
            std::vector<int32_t> fill(F.pt_ptr.begin(), F.pt_ptr.end() - 1);
            for (int64_t b = 0; b < nP; ++b) F.pt_blk[fill[F.p_col[b]]++] = (int32_t)b;
        }
        F.pt_row.assign(nP, 0);
        F.p_tpos.assign(nP, 0);
        for (int64_t q = 0; q < nP; ++q) {
            F.pt_row[q] = F.p_row[F.pt_blk[q]];
            F.p_tpos[F.pt_blk[q]] = (int32_t)q;
        }
        // W = A P: W(i, J) = sum over j in {i} u N(i) of A(i, j) P(j, J)
        F.w_ptr.assign(n + 1, 0);
        F.w_col.clear();
        F.w_row.clear();
        F.w_cptr.assign(1, 0);
        F.w_caref.clear();
        F.w_cp.clear();
        std::vector<std::pair<int32_t, std::pair<int32_t, int32_t>>> wprod;   // (J, (aref, P block))
        for (int i = 0; i < n; ++i) {
            wprod.clear();
            if (active[i] && F.cont[F.node_sys[i]]) {
                for (int k = F.rowptr[i]; k < F.rowptr[i + 1]; ++k) {
                    const int j = F.colind[k];
                    const int32_t aref = j == i ? -1 : F.ell.kref[k];
                    for (int b = F.p_ptr[j]; b < F.p_ptr[j + 1]; ++b) wprod.push_back({F.p_col[b], {aref, b}});
                }
            }
            std::sort(wprod.begin(), wprod.end());
            for (size_t t = 0; t < wprod.size();) {
                const int J = wprod[t].first;
                F.w_col.push_back(J);
                F.w_row.push_back(i);
                for (; t < wprod.size() && wprod[t].first == J; ++t) {
                    F.w_caref.push_back(wprod[t].second.first);
                    F.w_cp.push_back(wprod[t].second.second);
                }
                F.w_cptr.push_back((int32_t)F.w_cp.size());
            }
            F.w_ptr[i + 1] = (int32_t)F.w_col.size();
        }
        JSSO_REQUIRE((int64_t)F.w_col.size() < INT32_MAX / 36, JSSO_EUNSUPPORTED, "AMG: A P too large");
        // Galerkin pattern: (I, J) for I in P row i, J in W row i; products for I <= J
        std::vector<std::vector<int32_t>> cols(na);
        std::vector<std::pair<std::pair<int32_t, int32_t>, std::pair<int32_t, int32_t>>> cprod;   // ((I,J),(Pb,Wb))
        for (int i = 0; i < n; ++i)
            for (int pb = F.p_ptr[i]; pb < F.p_ptr[i + 1]; ++pb)
                for (int wb = F.w_ptr[i]; wb < F.w_ptr[i + 1]; ++wb) {
                    const int I = F.p_col[pb], J = F.w_col[wb];
                    cols[I].push_back(J);
                    cols[J].push_back(I);
                    if (I <= J) cprod.push_back({{I, J}, {pb, wb}});
                }
        A.host.emplace_back();
        AmgHostLevel& F0 = A.host[l];   // (emplace_back may have moved it)
        AmgHostLevel& C = A.host[l + 1];
        C.n = na;
        C.node_off = cnode_off;
        C.rowptr.assign(na + 1, 0);
        C.colind.clear();
        for (int I = 0; I < na; ++I) {
            auto& v = cols[I];
            v.push_back(I);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            C.colind.insert(C.colind.end(), v.begin(), v.end());
            C.rowptr[I + 1] = (int32_t)C.colind.size();
            std::vector<int32_t>().swap(v);
        }
        std::sort(cprod.begin(), cprod.end());
        const int64_t nnzc = C.colind.size();
        F0.c_cptr.assign(nnzc + 1, 0);
        F0.c_p.clear();
        F0.c_w.clear();
        {
            size_t t = 0;
            for (int I = 0; I < na; ++I)
                for (int k = C.rowptr[I]; k < C.rowptr[I + 1]; ++k) {
                    const int J = C.colind[k];
                    while (t < cprod.size() && cprod[t].first < std::make_pair(I, J)) ++t;   // (never: every product has its block)
                    for (; t < cprod.size() && cprod[t].first == std::make_pair(I, J); ++t) {
                        F0.c_p.push_back(cprod[t].second.first);
                        F0.c_w.push_back(cprod[t].second.second);
                    }
                    F0.c_cptr[k + 1] = (int32_t)F0.c_p.size();
                }
        }
        int rc = build_ell(ns, C.node_off.data(), C.rowptr, C.colind, C.ell);
        if (rc) return rc;
    }
    A.nl = (int)A.host.size();
    return JSSO_OK;
}


// ---------------------------------------------------------------------------
// device: numeric set-up (FP64), after every assembly
// ---------------------------------------------------------------------------
// entry e (row-major r * 6 + c) of the FP64 block stored at value position vp
// (grouped sliced-ELL layout, two entries per 16-byte group)
__device__ __forceinline__ double ell64(const double* __restrict__ A, int64_t vp, int e) {
    return A[(vp >> 6) * 36 * ROWS_PER_WG + ((int64_t)(e >> 1) * ROWS_PER_WG + (vp & 63)) * 2 + (e & 1)];
}
__device__ __forceinline__ int64_t ell64_index(int64_t vp, int e) {
    return (vp >> 6) * 36 * ROWS_PER_WG + ((int64_t)(e >> 1) * ROWS_PER_WG + (vp & 63)) * 2 + (e & 1);
}
__device__ __forceinline__ int64_t ell32_index(int64_t vp, int e) {
    return (vp >> 6) * 36 * ROWS_PER_WG + ((int64_t)(e >> 2) * ROWS_PER_WG + (vp & 63)) * 4 + (e & 3);
}
// entry (r, c) of the off-diagonal block A(i, j) given its reference
// (2 * value position, + 1 when A(i, j) is the transpose of the stored block);
// -1 is the unit diagonal block
__device__ __forceinline__ double ablk(const double* __restrict__ A, int32_t aref, int r, int c) {
    if (aref < 0) return r == c ? 1.0 : 0.0;
    const int64_t vp = aref >> 1;
    return (aref & 1) ? ell64(A, vp, c * 6 + r) : ell64(A, vp, r * 6 + c);
}

// The same entry from the natural-layout copy (k_amg_unpack): a block row is 48
// consecutive bytes there, three 16-byte pieces 1 KB apart in the ELL.
__device__ __forceinline__ double ablk_nat(const double* __restrict__ An, int32_t aref, int r, int c) {
    if (aref < 0) return r == c ? 1.0 : 0.0;
    const int64_t vp = aref >> 1;
    return (aref & 1) ? An[vp * 36 + c * 6 + r] : An[vp * 36 + r * 6 + c];
}

// FP64 ELL [slot][18][64 positions][2] -> [position][36] (FP32), a workgroup
// per slot of 64 positions transposed through LDS: the reads and the writes
// both cover consecutive addresses.
__global__ __launch_bounds__(256) void k_amg_unpack(int64_t nslot, const double2* __restrict__ ell,
                                                    float2* __restrict__ nat) {
    __shared__ float2 tile[18 * ROWS_PER_WG];
    const int64_t slot = xcd_wg();
    if (slot >= nslot) return;
    const double2* src = ell + slot * 18 * ROWS_PER_WG;
    for (int u = threadIdx.x; u < 18 * ROWS_PER_WG; u += blockDim.x) {
        const double2 v = src[u];   // u = g * 64 + position
        tile[(u & (ROWS_PER_WG - 1)) * 18 + (u >> 6)] = make_float2((float)v.x, (float)v.y);
    }
    __syncthreads();
    float2* dst = nat + slot * ROWS_PER_WG * 18;
    for (int u = threadIdx.x; u < 18 * ROWS_PER_WG; u += blockDim.x) dst[u] = tile[u];
}

// Rigid-body modes of every aggregated fine node relative to its aggregate's
// root, constrained rows zeroed, in the scaled variables: B_i = L_i^T M_i with
// L_i^T = (L_i^-1)^T solved for (upper triangular).  Columns: translations x, y,
// z, rotations about x, y, z; rows: the node's DOFs [ux, uy, uz, rx, ry, rz].
__global__ __launch_bounds__(256) void k_amg_null0(int32_t n, const int32_t* __restrict__ agg,
                                                   const int32_t* __restrict__ agg_ptr,
                                                   const int32_t* __restrict__ agg_nodes,
                                                   const double* __restrict__ crds, const double* __restrict__ Linv,
                                                   const uint8_t* __restrict__ fixed, double* __restrict__ B) {
    const int i = (int)(xcd_wg() * blockDim.x + threadIdx.x);
    if (i >= n) return;
    double* out = B + (int64_t)i * 36;
    const int a = agg[i];
    if (a < 0) {
#pragma unroll
        for (int e = 0; e < 36; ++e) out[e] = 0.0;
        return;
    }
    const int root = agg_nodes[agg_ptr[a]];
    const double dx = crds[3 * (int64_t)i] - crds[3 * (int64_t)root];
    const double dy = crds[3 * (int64_t)i + 1] - crds[3 * (int64_t)root + 1];
    const double dz = crds[3 * (int64_t)i + 2] - crds[3 * (int64_t)root + 2];
    double M[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) M[r][c] = r == c ? 1.0 : 0.0;
    // translations of a rotation e_a x d
    M[1][3] = -dz; M[2][3] = dy;
    M[0][4] = dz;  M[2][4] = -dx;
    M[0][5] = -dy; M[1][5] = dx;
#pragma unroll
    for (int r = 0; r < 6; ++r)
        if (fixed[6 * (int64_t)i + r])
#pragma unroll
            for (int c = 0; c < 6; ++c) M[r][c] = 0.0;
    const double* Li = Linv + (int64_t)i * 36;
    double U[6][6];   // U = Linv^T
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) U[r][c] = Li[c * 6 + r];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double x[6];
#pragma unroll
        for (int r = 5; r >= 0; --r) {
            double v = M[r][c];
#pragma unroll
            for (int t = r + 1; t < 6; ++t) v -= U[r][t] * x[t];
            x[r] = v / U[r][r];
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) out[r * 6 + c] = x[r];
    }
}

// Orthonormalise the stacked null-space rows of each aggregate (modified
// Gram-Schmidt, two passes), one wavefront per aggregate with its 6 m rows
// spread over the lanes (up to AMG_QR_RPL rows per lane): Q (per member node,
// 6 x 6) and R (6 x 6, upper triangular, B = Q R).  A column that is dependent
// on the previous ones (relative norm below 1e-10, e.g. translations of an
// aggregate whose translations are all constrained) is dropped: zero in Q and
// zero pivot in R.
// Level 0 builds the rotations about the aggregate's first node (well
// conditioned); the coarse null space must hold the same global modes in every
// aggregate, so R is then referred to one origin per structure (its first node
// x_o): a rotation about x_o is the local one plus the translation e_a x d,
// d = x_root - x_o, i.e. R <- R T with T = [[I, S], [0, I]], S e_a = e_a x d.
// (crds == nullptr on coarser levels, whose null space is already global.)
__global__ __launch_bounds__(64) void k_amg_qr(int32_t na, const int32_t* __restrict__ agg_ptr,
                                               const int32_t* __restrict__ agg_nodes, const double* __restrict__ B,
                                               const double* __restrict__ crds, const int32_t* __restrict__ node_sys,
                                               const int32_t* __restrict__ node_off, double* __restrict__ Q,
                                               double* __restrict__ R) {
    const int a = (int)xcd_wg();
    if (a >= na) return;
    const int lane = threadIdx.x;
    const int m0 = agg_ptr[a], nrow = 6 * (agg_ptr[a + 1] - m0);
    double v[AMG_QR_RPL][6];
    int64_t off[AMG_QR_RPL];
#pragma unroll
    for (int t = 0; t < AMG_QR_RPL; ++t) {
        const int rho = lane + 64 * t;
        off[t] = rho < nrow ? (int64_t)agg_nodes[m0 + rho / 6] * 36 + (rho % 6) * 6 : -1;
#pragma unroll
        for (int k = 0; k < 6; ++k) v[t][k] = off[t] >= 0 ? B[off[t] + k] : 0.0;
    }
    double Rl[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) Rl[r][c] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double n0 = 0.0;
#pragma unroll
        for (int t = 0; t < AMG_QR_RPL; ++t) n0 += v[t][k] * v[t][k];
        n0 = wave_sum(n0);
#pragma unroll
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int j = 0; j < k; ++j) {
                double d = 0.0;
#pragma unroll
                for (int t = 0; t < AMG_QR_RPL; ++t) d += v[t][j] * v[t][k];
                d = wave_sum(d);
#pragma unroll
                for (int t = 0; t < AMG_QR_RPL; ++t) v[t][k] -= d * v[t][j];
                Rl[j][k] += d;
            }
        double nk = 0.0;
#pragma unroll
        for (int t = 0; t < AMG_QR_RPL; ++t) nk += v[t][k] * v[t][k];
        nk = sqrt(wave_sum(nk));
        const bool drop = !(nk > 1e-10 * sqrt(n0));
        const double sc = drop ? 0.0 : 1.0 / nk;
#pragma unroll
        for (int t = 0; t < AMG_QR_RPL; ++t) v[t][k] *= sc;
        Rl[k][k] = drop ? 0.0 : nk;
    }
    if (crds) {
        const int64_t root = agg_nodes[m0];
        const int64_t o = node_off[node_sys[root]];
        const double dx = crds[3 * root] - crds[3 * o], dy = crds[3 * root + 1] - crds[3 * o + 1],
                     dz = crds[3 * root + 2] - crds[3 * o + 2];
        const double S[3][3] = {{0.0, dz, -dy}, {-dz, 0.0, dx}, {dy, -dx, 0.0}};
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) t += Rl[r][k] * S[k][a];
                Rl[r][3 + a] += t;
            }
    }
#pragma unroll
    for (int t = 0; t < AMG_QR_RPL; ++t)
        if (off[t] >= 0)
#pragma unroll
            for (int k = 0; k < 6; ++k) Q[off[t] + k] = v[t][k];
    if (lane < 36) {
        double val = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                if (r * 6 + c == lane) val = Rl[r][c];
        R[(int64_t)a * 36 + lane] = val;
    }
}

// deterministic start vector in [-1, 1) for the power iteration
__device__ __forceinline__ double amg_hash(int64_t dof) {
    uint32_t u = (uint32_t)dof * 2654435761u;
    u ^= u >> 15;
    u *= 2246822519u;
    u ^= u >> 13;
    return (double)(u >> 8) * (2.0 / 16777216.0) - 1.0;
}

// One power-iteration step y = A v on the slices of a level (lane = block row;
// level 0 reads the FP16 copy that its smoother applies, coarse levels the FP32
// one: only the largest eigenvalue is wanted): step 0
// takes the hashed start vector, later steps v = y_prev / |y_prev| of their
// structure (from the per-slice partials of the previous step).  Writes y and
// this slice's |y|^2.  Structures without a coarser level are skipped.
template <typename MT, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_amg_power(Slices sl, const MT* __restrict__ A, int step,
                                                        const double* __restrict__ vin, double* __restrict__ vout,
                                                        double* __restrict__ part, const uint8_t* __restrict__ cont,
                                                        int32_t n_wg) {
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int4 ma = sl.meta[3 * w], mb = sl.meta[3 * w + 1], mc = sl.meta[3 * w + 2];
    if (!cont[ma.x]) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double sc = 1.0;
    auto v = [&](int64_t node, int c, int) { return step == 0 ? amg_hash(6 * node + c) : vin[6 * node + c] * sc; };
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    // own rows requested first; the scale (from the partial sums) is formed while
    // they and the first slot's indices arrive
    double vo[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) vo[c] = step == 0 ? amg_hash(6 * ii + c) : vin[6 * ii + c];
    double acc[6][1];
    ell_apply_pre<MT, double, 1, WPS>(sl, A, w, ma, mb, acc, v, [&] {
        if (step > 0) {
            const double* pp = part + (int64_t)((step - 1) & 1) * n_wg;
            double t = 0.0;
            for (int q = mc.x + lane; q < mc.y; q += 64) t += pp[q];
            t = wave_sum(t);
            sc = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c][0] = wave == 0 ? (step == 0 ? vo[c] : vo[c] * sc) : 0.0;
    });
    if (wave != 0) return;
    double yy = 0.0;
    if (on) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vout[6 * i + c] = acc[c][0];
            yy += acc[c][0] * acc[c][0];
        }
    }
    yy = wave_sum(yy);
    if (lane == 0) part[(int64_t)(step & 1) * n_wg + w] = yy;
}

// rho = safety * |A v| (|v| = 1) per continuing structure; omega = 4/(3 rho);
// level 0: theta = AMG_THETA rho (one damped step).  Coarse levels (cheb set):
// degree-2 Chebyshev on [lo, hi] = [AMG_CHEB_LO hi, hi], hi = AMG_CHEB_HI rho,
// with centre theta = (hi + lo) / 2, half-width delta = (hi - lo) / 2 and
// sigma = theta / delta; its two steps from x0 are
//   x1 = x0 + r0 / theta,  x2 = x1 + rho1 rho0 r0 / theta + (2 rho1 / delta) r1,
// rho0 = 1 / sigma, rho1 = 1 / (2 sigma - rho0), and both weights equal
//   c = 2 sigma^2 / ((2 sigma^2 - 1) theta),
// so x2 = x0 + c (r0 + r1), r1 = r0 - A r0 / theta (Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., ch. 12: the three-term Chebyshev recurrence).
// A wavefront per structure: the lanes bring the per-slice partials into LDS
// (a structure of a single plan can have hundreds of slices, whose loads one
// lane would wait for in turn), lane 0 adds them in slice order.
constexpr int AMG_THETA_CHUNK = 2048;
__global__ __launch_bounds__(64) void k_amg_theta(int32_t ns, const int32_t* __restrict__ sys_wg_off,
                                                  const double* __restrict__ part, int last,
                                                  const uint8_t* __restrict__ cont, int32_t n_wg,
                                                  double* __restrict__ rho, double* __restrict__ omega,
                                                  float* __restrict__ theta, float* __restrict__ cheb) {
    __shared__ double buf[AMG_THETA_CHUNK];
    const int s = blockIdx.x;
    if (s >= ns || !cont[s]) return;
    const double* pp = part + (int64_t)(last & 1) * n_wg;
    const int q0 = sys_wg_off[s], q1 = sys_wg_off[s + 1];
    double t = 0.0;
    for (int c0 = q0; c0 < q1; c0 += AMG_THETA_CHUNK) {
        const int n = min(AMG_THETA_CHUNK, q1 - c0);
        for (int u = threadIdx.x; u < n; u += 64) buf[u] = pp[c0 + u];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int u = 0; u < n; ++u) t += buf[u];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double r = AMG_RHO_SAFETY * sqrt(t);
    if (!(r > 0.0) || !isfinite(r)) r = 1.0;
    rho[s] = r;
    omega[s] = 4.0 / (3.0 * r);
    if (cheb) {
        const double hi = AMG_CHEB_HI * r, lo = AMG_CHEB_LO * hi;
        const double th = 0.5 * (hi + lo), sg = th / (0.5 * (hi - lo));
        theta[s] = (float)th;
        cheb[s] = (float)(2.0 * sg * sg / ((2.0 * sg * sg - 1.0) * th));
    } else {
        theta[s] = (float)(AMG_THETA * r);
    }
}

// The set-up products (P, W = A P, A_c = P^T W) run one thread per output block
// (36 accumulators): the thread reads each operand block as 16-byte pieces of
// consecutive memory and forms all 36 outputs from it.  (A thread per block row
// issued 6 x 7 8-byte loads per product for the same bytes and ran at half the
// occupancy.)  Their operands -- the natural-layout copy of A, P and W -- are
// stored in FP32 and the products accumulate in FP64: the loads are what bounds
// these kernels, and the V-cycle uses P in FP16 and the coarse operators in
// FP32 anyway (6.8 -> 5.5 ms of set-up per step on the 64-design SIMP batch).
__device__ __forceinline__ void load_blk(const float* __restrict__ p, double (&v)[36]) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const float4 t = q[e];
        v[4 * e] = t.x;
        v[4 * e + 1] = t.y;
        v[4 * e + 2] = t.z;
        v[4 * e + 3] = t.w;
    }
}
__device__ __forceinline__ void store_blk(float* __restrict__ p, const double (&v)[36]) {
    float4* o = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = make_float4((float)v[4 * e], (float)v[4 * e + 1], (float)v[4 * e + 2], (float)v[4 * e + 3]);
}
__device__ __forceinline__ void load_blk(const double* __restrict__ p, double (&v)[36]) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int e = 0; e < 18; ++e) {
        const double2 t = q[e];
        v[2 * e] = t.x;
        v[2 * e + 1] = t.y;
    }
}

// P(i, J) = sum over its products of ((1 - omega) delta_ij I - omega A_off(i, j)) Q_j
__global__ __launch_bounds__(256) void k_amg_psmooth(int64_t nP, const int32_t* __restrict__ p_row,
                                                     const int32_t* __restrict__ p_cptr,
                                                     const int32_t* __restrict__ p_cj,
                                                     const int32_t* __restrict__ p_caref,
                                                     const int32_t* __restrict__ node_sys,
                                                     const double* __restrict__ omega, const float* __restrict__ An,
                                                     const double* __restrict__ Q, float* __restrict__ P) {
    const int64_t b = xcd_wg() * blockDim.x + threadIdx.x;
    if (b >= nP) return;
    const double om = omega[node_sys[p_row[b]]];
    double out[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) out[e] = 0.0;
    // (the next product's indices are requested while this one is formed)
    const int q1 = p_cptr[b + 1];
    int q = p_cptr[b];
    int32_t cjn = q < q1 ? p_cj[q] : 0, arn = q < q1 ? p_caref[q] : -1;
    for (; q < q1; ++q) {
        const int32_t cj = cjn, aref = arn;
        if (q + 1 < q1) {
            cjn = p_cj[q + 1];
            arn = p_caref[q + 1];
        }
        double qv[36];
        load_blk(Q + (int64_t)cj * 36, qv);
        if (aref < 0) {
#pragma unroll
            for (int e = 0; e < 36; ++e) out[e] += (1.0 - om) * qv[e];
        } else {
            double av[36];
            load_blk(An + (int64_t)(aref >> 1) * 36, av);
            const bool tr = aref & 1;
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    const double a = -om * (tr ? av[m * 6 + r] : av[r * 6 + m]);
#pragma unroll
                    for (int c = 0; c < 6; ++c) out[r * 6 + c] += a * qv[m * 6 + c];
                }
        }
    }
    store_blk(P + b * 36, out);
}

// W(i, J) = sum over its products of A(i, j) P(j, J)
__global__ __launch_bounds__(256) void k_amg_w(int64_t nW, const int32_t* __restrict__ w_cptr,
                                               const int32_t* __restrict__ w_caref, const int32_t* __restrict__ w_cp,
                                               const float* __restrict__ An, const float* __restrict__ P,
                                               float* __restrict__ W) {
    const int64_t b = xcd_wg() * blockDim.x + threadIdx.x;
    if (b >= nW) return;
    double out[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) out[e] = 0.0;
    // (the next product's indices are requested while this one is formed)
    const int q1 = w_cptr[b + 1];
    int q = w_cptr[b];
    int32_t cpn = q < q1 ? w_cp[q] : 0, arn = q < q1 ? w_caref[q] : -1;
    for (; q < q1; ++q) {
        const int32_t cp = cpn, aref = arn;
        if (q + 1 < q1) {
            cpn = w_cp[q + 1];
            arn = w_caref[q + 1];
        }
        double pv[36];
        load_blk(P + (int64_t)cp * 36, pv);
        if (aref < 0) {
#pragma unroll
            for (int e = 0; e < 36; ++e) out[e] += pv[e];
        } else {
            double av[36];
            load_blk(An + (int64_t)(aref >> 1) * 36, av);
            const bool tr = aref & 1;
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    const double a = tr ? av[m * 6 + r] : av[r * 6 + m];
#pragma unroll
                    for (int c = 0; c < 6; ++c) out[r * 6 + c] += a * pv[m * 6 + c];
                }
        }
    }
    store_blk(W + b * 36, out);
}

// A_c(I, J) = sum over its products of P(i, I)^T W(i, J) (blocks with I <= J);
// thread per output block, like the two products above
__global__ __launch_bounds__(256) void k_amg_galerkin(int64_t nnzc, const int32_t* __restrict__ c_cptr,
                                                      const int32_t* __restrict__ c_p, const int32_t* __restrict__ c_w,
                                                      const float* __restrict__ P, const float* __restrict__ W,
                                                      double* __restrict__ Ac) {
    const int64_t k = xcd_wg() * blockDim.x + threadIdx.x;
    if (k >= nnzc) return;
    double out[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) out[e] = 0.0;
    // (the next product's indices are requested while this one is formed)
    const int q1 = c_cptr[k + 1];
    int q = c_cptr[k];
    int32_t cpn = q < q1 ? c_p[q] : 0, cwn = q < q1 ? c_w[q] : 0;
    for (; q < q1; ++q) {
        const int32_t cp = cpn, cw = cwn;
        if (q + 1 < q1) {
            cpn = c_p[q + 1];
            cwn = c_w[q + 1];
        }
        double pv[36], wv[36];
        load_blk(P + (int64_t)cp * 36, pv);
        load_blk(W + (int64_t)cw * 36, wv);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int m = 0; m < 6; ++m)
#pragma unroll
                for (int c = 0; c < 6; ++c) out[r * 6 + c] += pv[m * 6 + r] * wv[m * 6 + c];
    }
    double2* o = reinterpret_cast<double2*>(Ac + k * 36);
#pragma unroll
    for (int e = 0; e < 18; ++e) o[e] = make_double2(out[2 * e], out[2 * e + 1]);
}

// The same three products for the small levels of a hierarchy: one lane per
// entry (r, c) of an output block (36 lanes per block, so a workgroup of 256
// lanes spans about seven blocks):
// each lane adds, product by product, the six terms of its entry in the order in
// which a thread holding the whole 6x6 block adds them (the sums are those of
// that formulation, bit for bit), with the operands' block indices requested two
// products ahead and the lane's twelve operand values one product ahead.  On a
// level of few blocks the thread-per-block kernels wait on one chain of
// dependent block loads per lane (~100 us for the Galerkin product at every
// coarse level of the 200x200 shell, 15 us this way); on a large level the
// lane-per-entry form issues 36 times the load instructions and loses (the W
// product at level 0 of the SIMP batch: 0.65 ms per block, 3.4 ms per entry).
//
// P(i, J) = sum over its products of ((1 - omega) delta_ij I - omega A_off(i, j)) Q_j
__global__ __launch_bounds__(256) void k_amg_psmooth_e(int64_t nP, const int32_t* __restrict__ p_row,
                                                     const int32_t* __restrict__ p_cptr,
                                                     const int32_t* __restrict__ p_cj,
                                                     const int32_t* __restrict__ p_caref,
                                                     const int32_t* __restrict__ node_sys,
                                                     const double* __restrict__ omega, const float* __restrict__ An,
                                                     const double* __restrict__ Q, float* __restrict__ P) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= nP * 36) return;
    const int64_t b = t / 36;
    const int e = (int)(t - b * 36), r = e / 6, c = e - 6 * r;
    const double om = omega[node_sys[p_row[b]]];
    const int q0 = p_cptr[b], q1 = p_cptr[b + 1];
    // operands of one product: Q_j column c (and entry e), A row r (column r transposed)
    double qn[6], qd = 0.0;
    float an[6];
    int32_t arn = -1;
    auto fetch = [&](int32_t cj, int32_t aref) {
        const double* qb = Q + (int64_t)cj * 36;
#pragma unroll
        for (int m = 0; m < 6; ++m) qn[m] = qb[m * 6 + c];
        qd = qb[e];
        arn = aref;
        if (aref >= 0) {
            const float* ab = An + (int64_t)(aref >> 1) * 36;
            const bool tr = aref & 1;
#pragma unroll
            for (int m = 0; m < 6; ++m) an[m] = tr ? ab[m * 6 + r] : ab[r * 6 + m];
        }
    };
    int32_t cj2 = 0, ar2 = -1;
    if (q0 < q1) {
        fetch(p_cj[q0], p_caref[q0]);
        if (q0 + 1 < q1) {
            cj2 = p_cj[q0 + 1];
            ar2 = p_caref[q0 + 1];
        }
    }
    double acc = 0.0;
    for (int q = q0; q < q1; ++q) {
        double qc[6];
        float ac[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            qc[m] = qn[m];
            ac[m] = an[m];
        }
        const double qdc = qd;
        const int32_t aref = arn;
        if (q + 1 < q1) {
            fetch(cj2, ar2);
            if (q + 2 < q1) {
                cj2 = p_cj[q + 2];
                ar2 = p_caref[q + 2];
            }
        }
        if (aref < 0) {
            acc += (1.0 - om) * qdc;
        } else {
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const double a = -om * (double)ac[m];
                acc += a * qc[m];
            }
        }
    }
    P[t] = (float)acc;
}

// W(i, J) = sum over its products of A(i, j) P(j, J)
__global__ __launch_bounds__(256) void k_amg_w_e(int64_t nW, const int32_t* __restrict__ w_cptr,
                                               const int32_t* __restrict__ w_caref, const int32_t* __restrict__ w_cp,
                                               const float* __restrict__ An, const float* __restrict__ P,
                                               float* __restrict__ W) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= nW * 36) return;
    const int64_t b = t / 36;
    const int e = (int)(t - b * 36), r = e / 6, c = e - 6 * r;
    const int q0 = w_cptr[b], q1 = w_cptr[b + 1];
    // operands of one product: P column c (and entry e), A row r (column r transposed)
    float pn[6], pd = 0.f, an[6];
    int32_t arn = -1;
    auto fetch = [&](int32_t cp, int32_t aref) {
        const float* pb = P + (int64_t)cp * 36;
#pragma unroll
        for (int m = 0; m < 6; ++m) pn[m] = pb[m * 6 + c];
        pd = pb[e];
        arn = aref;
        if (aref >= 0) {
            const float* ab = An + (int64_t)(aref >> 1) * 36;
            const bool tr = aref & 1;
#pragma unroll
            for (int m = 0; m < 6; ++m) an[m] = tr ? ab[m * 6 + r] : ab[r * 6 + m];
        }
    };
    int32_t cp2 = 0, ar2 = -1;
    if (q0 < q1) {
        fetch(w_cp[q0], w_caref[q0]);
        if (q0 + 1 < q1) {
            cp2 = w_cp[q0 + 1];
            ar2 = w_caref[q0 + 1];
        }
    }
    double acc = 0.0;
    for (int q = q0; q < q1; ++q) {
        float pc[6], ac[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            pc[m] = pn[m];
            ac[m] = an[m];
        }
        const float pdc = pd;
        const int32_t aref = arn;
        if (q + 1 < q1) {
            fetch(cp2, ar2);
            if (q + 2 < q1) {
                cp2 = w_cp[q + 2];
                ar2 = w_caref[q + 2];
            }
        }
        if (aref < 0) {
            acc += (double)pdc;
        } else {
#pragma unroll
            for (int m = 0; m < 6; ++m) acc += (double)ac[m] * (double)pc[m];
        }
    }
    W[t] = (float)acc;
}

// A_c(I, J) = sum over its products of P(i, I)^T W(i, J) (blocks with I <= J)
__global__ __launch_bounds__(256) void k_amg_galerkin_e(int64_t nnzc, const int32_t* __restrict__ c_cptr,
                                                      const int32_t* __restrict__ c_p, const int32_t* __restrict__ c_w,
                                                      const float* __restrict__ P, const float* __restrict__ W,
                                                      double* __restrict__ Ac) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= nnzc * 36) return;
    const int64_t k = t / 36;
    const int e = (int)(t - k * 36), r = e / 6, c = e - 6 * r;
    const int q0 = c_cptr[k], q1 = c_cptr[k + 1];
    // operands of one product: column r of P(i, I), column c of W(i, J)
    float pn[6], wn[6];
    auto fetch = [&](int32_t cp, int32_t cw) {
        const float* pb = P + (int64_t)cp * 36;
        const float* wb = W + (int64_t)cw * 36;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            pn[m] = pb[m * 6 + r];
            wn[m] = wb[m * 6 + c];
        }
    };
    int32_t cp2 = 0, cw2 = 0;
    if (q0 < q1) {
        fetch(c_p[q0], c_w[q0]);
        if (q0 + 1 < q1) {
            cp2 = c_p[q0 + 1];
            cw2 = c_w[q0 + 1];
        }
    }
    double acc = 0.0;
    for (int q = q0; q < q1; ++q) {
        float pc[6], wc[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            pc[m] = pn[m];
            wc[m] = wn[m];
        }
        if (q + 1 < q1) {
            fetch(cp2, cw2);
            if (q + 2 < q1) {
                cp2 = c_p[q + 2];
                cw2 = c_w[q + 2];
            }
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) acc += (double)pc[m] * (double)wc[m];
    }
    Ac[t] = acc;
}

// Cholesky factor of each coarse diagonal block and its inverse.  A pivot that
// is not positive (a null-space column dropped by the QR leaves a zero row and
// column) is taken as a decoupled unknown: unit pivot, zero column.
__global__ __launch_bounds__(128) void k_amg_cfactor(int32_t nc, const int32_t* __restrict__ cdiag,
                                                     const double* __restrict__ Ac, double* __restrict__ Lc,
                                                     double* __restrict__ Lci) {
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= nc) return;
    const double* a = Ac + (int64_t)cdiag[J] * 36;
    double A[6][6], L[6][6], Li[6][6];
    double dmax = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            A[r][c] = 0.5 * (a[r * 6 + c] + a[c * 6 + r]);
            L[r][c] = 0.0;
            Li[r][c] = 0.0;
        }
        dmax = fmax(dmax, fabs(A[r][r]));
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = A[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        const bool zero = !(d > 1e-13 * dmax);
        const double ljj = zero ? 1.0 : sqrt(d);
        L[j][j] = ljj;
#pragma unroll
        for (int r = j + 1; r < 6; ++r) {
            double s = A[r][j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[r][k] * L[j][k];
            L[r][j] = zero ? 0.0 : s / ljj;
        }
    }
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
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            Lc[(int64_t)J * 36 + r * 6 + c] = L[r][c];
            Lci[(int64_t)J * 36 + r * 6 + c] = Li[r][c];
        }
}

// Scaled coarse operator into the coarse level's sliced ELL (FP64 and FP32):
// As(I, J) = Lc_I^-1 A_c(I, J) Lc_J^-T for every stored (upper) block.
__global__ __launch_bounds__(256) void k_amg_cscale(int64_t nval, const int32_t* __restrict__ ell_src,
                                                    const int32_t* __restrict__ browof,
                                                    const int32_t* __restrict__ colind, const double* __restrict__ Ac,
                                                    const double* __restrict__ Lci, double* __restrict__ A64,
                                                    float* __restrict__ A32) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= nval * 36) return;
    const int64_t vp = t / 36;
    const int e = (int)(t % 36), r = e / 6, c = e % 6;
    const int32_t src = ell_src[vp];
    double v = 0.0;
    if (src >= 0) {
        const double* a = Ac + (int64_t)src * 36;
        const double* li = Lci + (int64_t)browof[src] * 36;
        const double* lj = Lci + (int64_t)colind[src] * 36;
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) s += a[p * 6 + q] * lj[c * 6 + q];
            v += li[r * 6 + p] * s;
        }
    }
    A64[ell64_index(vp, e)] = v;
    A32[ell32_index(vp, e)] = (float)v;
}

// P scaled to the coarse variables, FP16 for the V-cycle (|entries| stay O(1): a
// few units at most; the smallest underflow harmlessly): Ps(i, J) = P(i, J) Lc_J^-T,
// by fine row (prolongation) and transposed by coarse node (restriction)
__global__ __launch_bounds__(256) void k_amg_pscale(int64_t nP, const int32_t* __restrict__ p_col,
                                                    const int32_t* __restrict__ p_tpos,
                                                    const float* __restrict__ P, const double* __restrict__ Lci,
                                                    _Float16* __restrict__ P16, _Float16* __restrict__ PT16) {
    // thread per block row r: the row of P once, the six entries of the row
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= nP * 6) return;
    const int64_t b = t / 6;
    const int r = (int)(t % 6);
    const float* p = P + b * 36 + r * 6;
    const double* l = Lci + (int64_t)p_col[b] * 36;
    const int64_t tp = (int64_t)p_tpos[b] * 36;
    float pr[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) pr[m] = p[m];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) s += pr[m] * l[c * 6 + m];
        P16[b * 36 + r * 6 + c] = (_Float16)s;
        PT16[tp + c * 6 + r] = (_Float16)s;   // per coarse node, block transposed
    }
}

// coarse null space in the scaled coarse variables: B_J = Lc_J^T R_J
__global__ __launch_bounds__(256) void k_amg_nullc(int32_t nc, const double* __restrict__ Lc,
                                                   const double* __restrict__ R, double* __restrict__ B) {
    const int64_t t = xcd_wg() * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nc * 36) return;
    const int64_t J = t / 36;
    const int e = (int)(t % 36), r = e / 6, c = e % 6;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) s += Lc[J * 36 + m * 6 + r] * R[J * 36 + m * 6 + c];
    B[t] = s;
}

// Dense inverse of a structure's coarsest scaled operator (N = 6 nodes <= 156),
// one workgroup of 1024 threads per structure, all in LDS with the lower
// triangle packed row by row (entry (r, c), c <= r, at r (r + 1) / 2 + c; 98 KB
// for N = 156):
//  1. the matrix (unit diagonal blocks plus the stored off-diagonal blocks) is
//     factored by a right-looking Cholesky A = L L^T in blocks of four columns,
//     two barriers per block (the panel, then the rank-4 update of the trailing
//     triangle by a 32 x 32 thread tile);
//  2. X = L^-1 replaces L in place by right-looking forward substitution on
//     L X = I in blocks of four rows, three barriers per block: the block rows of
//     X are finished (thread per column), then every later row i takes
//     S_ij -= L_i,block X_block,j, where S_ij (held at (i, j) since the block of
//     column j) becomes X_ij when its row is finished.
//     (the column-by-column form, one column per barrier pair, took 330 us for
//     128 structures of 138 DOF, this one 302 us: the FP64 LDS traffic of the
//     trailing updates, not the barriers, sets the cost);
//  3. A^-1 = X^T X by 4 x 4 register blocks of the lower block triangle, each
//     written to both triangles (the result is exactly symmetric), in FP32,
//     row-major.
constexpr int AMG_DENSE_MAX = 6 * AMG_MAXC;
__device__ __forceinline__ int tri(int r, int c) { return r * (r + 1) / 2 + c; }
__global__ __launch_bounds__(1024) void k_amg_dense(const int32_t* __restrict__ list, const int32_t* __restrict__ term,
                                                   const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ colind,
                                                   const int32_t* __restrict__ browof,
                                                   const int32_t* __restrict__ kref, const double* __restrict__ A,
                                                   float* __restrict__ inv) {
    constexpr int DB = 4;   // columns (rows) per block step
    __shared__ double M[AMG_DENSE_MAX * (AMG_DENSE_MAX + 1) / 2];
    __shared__ double pan[AMG_DENSE_MAX][DB];
    const int s = list[blockIdx.x];
    const int node0 = term[4 * s + 1], nt = term[4 * s + 2];
    const int64_t off = term[4 * s + 3];
    const int N = 6 * nt;
    const int NT = N * (N + 1) / 2;
    const int tid = threadIdx.x, ty = tid >> 5, tx = tid & 31;
    for (int u = tid; u < NT; u += blockDim.x) M[u] = 0.0;
    __syncthreads();
    // scatter the blocks of the structure's rows (lower triangle only): thread per (block, entry)
    const int k0 = rowptr[node0], k1 = rowptr[node0 + nt];
    for (int u = tid; u < (k1 - k0) * 36; u += blockDim.x) {
        const int k = k0 + u / 36, e = u % 36, r = e / 6, c = e % 6;
        const int R = 6 * (browof[k] - node0) + r, C = 6 * (colind[k] - node0) + c;
        if (C > R) continue;
        M[tri(R, C)] = kref[k] < 0 ? (r == c ? 1.0 : 0.0) : ablk(A, kref[k], r, c);
    }
    __syncthreads();
    // 1. Cholesky, blocks of DB columns: every thread factors the diagonal block
    // in registers and the row threads below it form their panel row
    // (L21 = A21 L11^-T) into `pan`; after the barrier the block and the panel
    // are written back and the trailing triangle takes the rank-DB update.
    for (int k0 = 0; k0 < N; k0 += DB) {
        const int kb = N - k0 < DB ? N - k0 : DB;
        double L11[DB][DB];
#pragma unroll
        for (int a = 0; a < DB; ++a)
#pragma unroll
            for (int c = 0; c < DB; ++c) L11[a][c] = 0.0;
#pragma unroll
        for (int c = 0; c < DB; ++c) {
            if (c < kb) {
                double d = M[tri(k0 + c, k0 + c)];
#pragma unroll
                for (int m = 0; m < c; ++m) d -= L11[c][m] * L11[c][m];
                const double lcc = d > 0.0 ? sqrt(d) : 1.0;
                L11[c][c] = lcc;
#pragma unroll
                for (int a = c + 1; a < DB; ++a) {
                    if (a < kb) {
                        double v = M[tri(k0 + a, k0 + c)];
#pragma unroll
                        for (int m = 0; m < c; ++m) v -= L11[a][m] * L11[c][m];
                        L11[a][c] = v / lcc;
                    }
                }
            }
        }
        for (int i = k0 + kb + tid; i < N; i += blockDim.x) {
            double pv[DB];
#pragma unroll
            for (int c = 0; c < DB; ++c) {
                pv[c] = 0.0;
                if (c < kb) {
                    double v = M[tri(i, k0 + c)];
#pragma unroll
                    for (int m = 0; m < c; ++m) v -= pv[m] * L11[c][m];
                    pv[c] = v / L11[c][c];
                    pan[i][c] = pv[c];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < DB; ++a)
#pragma unroll
            for (int c = 0; c <= a; ++c)
                if (tid == a * DB + c && a < kb) M[tri(k0 + a, k0 + c)] = L11[a][c];
        for (int i = k0 + kb + tid; i < N; i += blockDim.x)
            for (int c = 0; c < kb; ++c) M[tri(i, k0 + c)] = pan[i][c];
        for (int r = k0 + kb + ty; r < N; r += 32)
            for (int c = k0 + kb + tx; c <= r; c += 32) {
                double v = 0.0;
                for (int m = 0; m < kb; ++m) v += pan[r][m] * pan[c][m];
                M[tri(r, c)] -= v;
            }
        __syncthreads();
    }
    // 2. X = L^-1 in place by forward substitution on L X = I, blocks of DB rows:
    // every thread reads the block's L into registers and the row threads below
    // copy their panel of L into `pan`; then thread j finishes column j of the
    // block rows (S_kj, held at (k, j) since earlier blocks, minus the block's
    // own terms, over L_kk); then every later row i takes S_ij -= L_i,block X_block,j
    // (for j inside the block S_ij starts at 0: (i, j) still held L_ij).
    for (int k0 = 0; k0 < N; k0 += DB) {
        const int kb = N - k0 < DB ? N - k0 : DB;
        double Lb[DB][DB];
#pragma unroll
        for (int a = 0; a < DB; ++a)
#pragma unroll
            for (int c = 0; c < DB; ++c) Lb[a][c] = (a < kb && c <= a) ? M[tri(k0 + a, k0 + c)] : 0.0;
        for (int i = k0 + kb + tid; i < N; i += blockDim.x)
            for (int c = 0; c < kb; ++c) pan[i][c] = M[tri(i, k0 + c)];
        __syncthreads();
        for (int j = tid; j < k0 + kb; j += blockDim.x) {
            double x[DB];
#pragma unroll
            for (int a = 0; a < DB; ++a) {
                x[a] = 0.0;
                const int k = k0 + a;
                if (a < kb && j <= k) {
                    double v = j < k0 ? M[tri(k, j)] : (j == k ? 1.0 : 0.0);
#pragma unroll
                    for (int m = 0; m < a; ++m) v -= Lb[a][m] * x[m];   // x[m] = 0 where j > k0 + m
                    x[a] = v / Lb[a][a];
                    M[tri(k, j)] = x[a];
                }
            }
        }
        __syncthreads();
        for (int i = k0 + kb + ty; i < N; i += 32)
            for (int j = tx; j < k0 + kb; j += 32) {
                double v = j < k0 ? M[tri(i, j)] : 0.0;
                for (int m = 0; m < kb; ++m)
                    if (j <= k0 + m) v -= pan[i][m] * M[tri(k0 + m, j)];
                M[tri(i, j)] = v;
            }
        __syncthreads();
    }
    // 3. A^-1 = X^T X: (A^-1)_rc = sum_{k >= max(r, c)} X_kr X_kc
    const int nb = (N + 3) / 4;
    for (int t = tid; t < nb * (nb + 1) / 2; t += blockDim.x) {
        int R = (int)((sqrtf(8.f * t + 1.f) - 1.f) * 0.5f);
        while (R * (R + 1) / 2 > t) --R;
        while ((R + 1) * (R + 2) / 2 <= t) ++R;
        const int C = t - R * (R + 1) / 2;
        const int r0 = 4 * R, c0 = 4 * C;
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int k = r0; k < N; ++k) {
            double xr[4], xc[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                xr[a] = r0 + a <= k ? M[tri(k, r0 + a)] : 0.0;
                xc[a] = c0 + a <= k ? M[tri(k, c0 + a)] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += xr[a] * xc[b];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int r = r0 + a, c = c0 + b;
                if (r < N && c < N) {
                    inv[off + (int64_t)r * N + c] = (float)acc[a][b];
                    inv[off + (int64_t)c * N + r] = (float)acc[a][b];
                }
            }
    }
}

// ---------------------------------------------------------------------------
// device: the preconditioned CG iteration and the V-cycle
// ---------------------------------------------------------------------------
// Partial slots of the AMG iteration k (per slice, in P.cg[k & 1]): CG_RR holds
// <r_k, z_k>, CG_RQ holds <r_{k-1}, z_k>, CG_PQ holds <p_k, q_k>.  The
// direction uses the Polak-Ribiere (flexible CG) coefficient
//     beta_k = (<r_{k-1}, z_{k-1}> - <r_{k-2}, z_{k-1}>) / <r_{k-2}, z_{k-2}>, clipped at 0,
// which keeps the iteration convergent with the FP32 V-cycle (a preconditioner
// that is symmetric only to FP32 rounding) and across residual replacements.

// Waves-per-slice variants whose cross-wave LDS buffer stays <= 24 KB.
template <typename VT, int NR, int WPS>
struct AmgWpsFits {
    static constexpr bool value = WPS == 1 || (size_t)(WPS - 1) * 6 * NR * 64 * sizeof(VT) <= 24576;
};

// Totals over the slices [s0, s1) of one structure of M (buffer, slot) pairs, in
// one wavefront (every lane holds the same butterfly totals).  Each lane issues
// the loads of 4 slices at once, then adds them in slice order.
template <int NR, int M>
__device__ __forceinline__ void wave_totals_m(const double* const (&cg)[M], const int (&slot)[M], int s0, int s1,
                                              double (&t)[M][NR]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int h = 0; h < NR; ++h) t[m][h] = 0.0;
    for (int q0 = s0 + lane; q0 < s1; q0 += 256) {
        double v[4][M][NR];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = min(q0 + 64 * u, s1 - 1);
#pragma unroll
            for (int m = 0; m < M; ++m)
#pragma unroll
                for (int h = 0; h < NR; ++h) v[u][m][h] = cg[m][((int64_t)q * 4 + slot[m]) * NR + h];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q0 + 64 * u < s1) {
#pragma unroll
                for (int m = 0; m < M; ++m)
#pragma unroll
                    for (int h = 0; h < NR; ++h) t[m][h] += v[u][m][h];
            }
    }
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int h = 0; h < NR; ++h) t[m][h] = wave_sum(t[m][h]);
}

// acc[c][h] += sum over the slice's ELL slots of A_off(i, j) x(j), x(j) given
// by `xj(col, c, h)` (lane = block row).  With WPS wavefronts per slice, wave v
// takes slots v, v + WPS, ... and waves 1..WPS-1 hand their sums to wave 0
// through LDS, which adds them in wave order: only wave 0's acc is the result
// (the caller puts the own-row term into wave 0's acc beforehand).  The next
// slot's column and value position are requested before the current slot's
// products; a slot past the end reads the slice's metadata word instead.
template <typename MT, typename VT, int NR, int WPS, int G = 1, typename F, typename Pre>
__device__ __forceinline__ void ell_apply_pre(const Slices& sl, const MT* __restrict__ A, int w, const int4& ma,
                                              const int4& mb, VT (&acc)[6][NR], F&& xj, Pre&& pre) {
    // G groups of WPS waves, each group on a slice of its own
    __shared__ VT red[WPS > 1 ? G : 1][WPS > 1 ? WPS - 1 : 1][WPS > 1 ? 6 * NR : 1][64];
    const int lane = threadIdx.x & 63, wave = WPS > 1 ? (threadIdx.x >> 6) % WPS : 0;
    const int grp = G > 1 ? (threadIdx.x >> 6) / WPS : 0;
    const int nup = mb.y, nslot = mb.y + mb.w;
    const MT* Aw = A + 36 * (int64_t)ma.w;
    const int32_t* dummy = (const int32_t*)&sl.meta[3 * w];
    auto cpos = [&](int jj) -> int64_t {
        return jj < nup ? (int64_t)mb.x + (int64_t)jj * ROWS_PER_WG + lane
                        : (int64_t)mb.z + (int64_t)(jj - nup) * ROWS_PER_WG + lane;
    };
    int j = wave;
    int64_t col = *(j < nslot ? sl.ell_col + cpos(j) : dummy);
    int vp = *(j < nslot ? sl.ell_vpos + cpos(j) : dummy);
    pre();   // (work that xj needs, run while the first slot's indices arrive)
    for (; j < nslot; j += WPS) {
        VT x[6][NR];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) x[c][h] = xj(col, c, h);
        const bool upper = j < nup;
        MT av[36];
        if (upper)
            load_block(Aw + (int64_t)j * 36 * ROWS_PER_WG, lane, av);
        else
            load_block(A + (int64_t)(vp >> 6) * 36 * ROWS_PER_WG, vp & (ROWS_PER_WG - 1), av);
        const int jn = j + WPS;
        const int64_t coln = *(jn < nslot ? sl.ell_col + cpos(jn) : dummy);
        const int vpn = *(jn < nslot ? sl.ell_vpos + cpos(jn) : dummy);
        if (upper) {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e / 6][h] += (VT)av[e] * x[e % 6][h];
        } else {
#pragma unroll
            for (int e = 0; e < 36; ++e)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[e % 6][h] += (VT)av[e] * x[e / 6][h];
        }
        col = coln;
        vp = vpn;
    }
    if (WPS > 1) {
        if (wave > 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) red[grp][wave - 1][c * NR + h][lane] = acc[c][h];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int u = 0; u < WPS - 1; ++u)
#pragma unroll
                for (int c = 0; c < 6; ++c)
#pragma unroll
                    for (int h = 0; h < NR; ++h) acc[c][h] += red[grp][u][c * NR + h][lane];
        }
    }
}

template <typename MT, typename VT, int NR, int WPS, int G = 1, typename F>
__device__ __forceinline__ void ell_apply(const Slices& sl, const MT* __restrict__ A, int w, const int4& ma,
                                          const int4& mb, VT (&acc)[6][NR], F&& xj) {
    ell_apply_pre<MT, VT, NR, WPS, G>(sl, A, w, ma, mb, acc, xj, [] {});
}

// The rows of a slice form one contiguous run of (row1 - row0) * 6 * NR values
// per vector: wave 0 puts value(c, h) of its lane's row into LDS, then every
// thread of the workgroup copies the run, so each store instruction writes
// consecutive addresses (whole cache lines rather than 4-byte pieces 24 bytes
// apart).  Called by all threads.
template <typename VT, int NR, int WPS, typename F, typename DT>
__device__ __forceinline__ void flush_rows(VT* stage, const int4& ma, F&& value, DT* __restrict__ dst) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) stage[lane * 6 * NR + c * NR + h] = value(c, h);
    }
    __syncthreads();
    const int64_t base = (int64_t)ma.y * 6 * NR;
    const int n = (ma.z - ma.y) * 6 * NR;
    for (int u = threadIdx.x; u < n; u += 64 * WPS) dst[base + u] = (DT)stage[u];
    __syncthreads();
}

// M such runs (one per vector) through one staging round: the stores of M
// flush_rows calls with one pair of barriers.  `stage` holds M * 64 * 6 * NR
// values; run 0 goes to the FP32 vector dst0, runs 1.. to dst[1..].
template <typename VT, int NR, int WPS, int M, typename F>
__device__ __forceinline__ void flush_rows_m(VT* stage, const int4& ma, F&& value, float* __restrict__ dst0,
                                             VT* const (&dst)[M]) {
    constexpr int S = 64 * 6 * NR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h)
#pragma unroll
                for (int m = 0; m < M; ++m) stage[m * S + lane * 6 * NR + c * NR + h] = value(m, c, h);
    }
    __syncthreads();
    const int64_t base = (int64_t)ma.y * 6 * NR;
    const int n = (ma.z - ma.y) * 6 * NR;
    for (int u = threadIdx.x; u < n; u += 64 * WPS) {
        dst0[base + u] = (float)stage[u];
#pragma unroll
        for (int m = 1; m < M; ++m) dst[m][base + u] = stage[m * S + u];
    }
    __syncthreads();
}

// Direction and product of iteration k: beta from the partials of k - 1 and
// k - 2; p_k = z + beta p_{k-1} (own and gathered rows); q = A p_k; partials <p_k, q>.
template <typename MT, typename VT, int NR, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_pcg_dir(Slices sl, Partials P, SysState st, int k,
                                                      const MT* __restrict__ A, const VT* __restrict__ z,
                                                      const VT* __restrict__ pprev, VT* __restrict__ pcur,
                                                      VT* __restrict__ q) {
    const int w = slice_of_block(sl);
    if (w < 0) return;
    // (metadata requested together with the slice's flags, not after them)
    const int4 ma = sl.meta[3 * w], mb = sl.meta[3 * w + 1], mc = sl.meta[3 * w + 2];
    const int amask = st.slice_act[w];
    if (!amask) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    // own rows requested first; beta (from the partial sums) is formed while
    // they and the first slot's indices arrive
    VT zo[6][NR], po[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * ii + c) * NR + h;
            zo[c][h] = z[o];
            po[c][h] = pprev[o];
        }
    VT beta[NR];
    auto pk = [&](int64_t node, int c, int h) {
        const int64_t o = (6 * node + c) * NR + h;
        return z[o] + beta[h] * pprev[o];
    };
    VT acc[6][NR];
    ell_apply_pre<MT, VT, NR, WPS>(sl, A, w, ma, mb, acc, pk, [&] {
        const double* const cgp = (k & 1) ? P.cg[0] : P.cg[1];   // P.cg[(k - 1) & 1]
        const double* const cgc = (k & 1) ? P.cg[1] : P.cg[0];   // P.cg[k & 1]
        const double* const bufs[3] = {cgp, cgp, cgc};
        const int slots[3] = {CG_RR, CG_RQ, CG_RR};
        double tt[3][NR];
        wave_totals_m<NR, 3>(bufs, slots, mc.x, mc.y, tt);
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const double b = tt[2][h] > 0.0 ? (tt[0][h] - tt[1][h]) / tt[2][h] : 0.0;
            beta[h] = (isfinite(b) && b > 0.0) ? VT(b) : VT(0);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[c][h] = wave == 0 ? zo[c][h] + beta[h] * po[c][h] : VT(0);
    });
    __shared__ VT stage[64 * 6 * NR];
    double v[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = 0.0;
    // (inactive load cases keep p; their q is never used; one staging round for
    // both vectors measured slower here)
    flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) {
        const VT pn = zo[c][h] + beta[h] * po[c][h];
        if (on && ((amask >> h) & 1)) v[h] += (double)pn * (double)acc[c][h];
        return ((amask >> h) & 1) ? pn : po[c][h];
    }, pcur);
    flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) { return acc[c][h]; }, q);
    if (wave != 0) return;
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = wave_sum(v[h]);
    if (lane == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) P.cg[k & 1][((int64_t)w * 4 + CG_PQ) * NR + h] = v[h];
    }
}

// Update of iteration k and the V-cycle's first smoothing step on level 0:
// alpha = <r,z>_{k-1} / <p,q>_k; r_k = r_{k-1} - alpha q (own and gathered
// rows); x += alpha p_k; t0 = r_k - A (r_k / theta).  FRESH: r_k is given (a
// replacement or the start), no update.  The smoothing product uses the FP16
// copy of A (a preconditioner needs no more; halves the kernel's matrix bytes).
template <typename VT, int NR, bool FRESH, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_pcg_pre(Slices sl, Partials P, SysState st, int k,
                                                      const _Float16* __restrict__ A, const float* __restrict__ theta,
                                                      const VT* __restrict__ rprev, VT* __restrict__ rcur,
                                                      const VT* __restrict__ q, const VT* __restrict__ p,
                                                      VT* __restrict__ xlo, float* __restrict__ t0) {
    const int w = slice_of_block(sl);
    if (w < 0) return;
    // (metadata requested together with the slice's flags, not after them)
    const int4 ma = sl.meta[3 * w], mb = sl.meta[3 * w + 1], mc = sl.meta[3 * w + 2];
    const int amask = st.slice_act[w];
    if (!amask) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    // own rows requested first; alpha (from the partial sums) is formed while
    // they and the first slot's indices arrive
    VT ro[6][NR], qo[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * ii + c) * NR + h;
            ro[c][h] = FRESH ? rcur[o] : rprev[o];
            qo[c][h] = FRESH ? VT(0) : q[o];
        }
    // own rows of x and p for the update at the end: requested now when the
    // registers allow it (one right-hand side), not after the product
    constexpr bool PF = !FRESH && NR == 1;
    constexpr int NP = PF ? NR : 1;
    VT xo[6][NP], pko[6][NP];
    if constexpr (PF) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NP; ++h) {
                const int64_t o = (6 * ii + c) * NR + h;
                xo[c][h] = xlo[o];
                pko[c][h] = p[o];
            }
    }
    VT alpha[NR];
    auto rk = [&](int64_t node, int c, int h) {
        const int64_t o = (6 * node + c) * NR + h;
        return FRESH ? rcur[o] : rprev[o] - alpha[h] * q[o];
    };
    const VT th = (VT)theta[ma.x];
    VT acc[6][NR];
    ell_apply_pre<_Float16, VT, NR, WPS>(sl, A, w, ma, mb, acc, rk, [&] {
#pragma unroll
        for (int h = 0; h < NR; ++h) alpha[h] = VT(0);
        if (!FRESH) {
            const double* const bufs[2] = {(k & 1) ? P.cg[0] : P.cg[1], (k & 1) ? P.cg[1] : P.cg[0]};
            const int slots[2] = {CG_RR, CG_PQ};
            double tt[2][NR];
            wave_totals_m<NR, 2>(bufs, slots, mc.x, mc.y, tt);
#pragma unroll
            for (int h = 0; h < NR; ++h)
                alpha[h] = (tt[1][h] > 0.0 && isfinite(tt[1][h]) && tt[0][h] > 0.0) ? VT(tt[0][h] / tt[1][h]) : VT(0);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h)
                acc[c][h] = wave == 0 ? (FRESH ? ro[c][h] : ro[c][h] - alpha[h] * qo[c][h]) : VT(0);
    });
    // t0, r and x through one staging round when that stays small (<= 12 KB)
    constexpr bool JOIN = !FRESH && 3 * 64 * 6 * NR * sizeof(VT) <= 12288;
    __shared__ VT stage[(JOIN ? 3 : 1) * 64 * 6 * NR];
    // (inactive load cases keep r and x)
    auto val = [&](int m, int c, int h) {
        const int64_t o = (6 * ii + c) * NR + h;
        const VT rn = FRESH ? ro[c][h] : ro[c][h] - alpha[h] * qo[c][h];
        if (m == 0) return rn - acc[c][h] / th;
        if (m == 1) return ((amask >> h) & 1) ? rn : ro[c][h];
        const VT xv = PF ? xo[c][PF ? h : 0] : xlo[o];
        return ((amask >> h) & 1) ? xv + alpha[h] * (PF ? pko[c][PF ? h : 0] : p[o]) : xv;
    };
    if constexpr (JOIN) {
        VT* const dst[3] = {nullptr, rcur, xlo};
        flush_rows_m<VT, NR, WPS, 3>(stage, ma, val, t0, dst);
    } else {
        flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) { return val(0, c, h); }, t0);
        if (!FRESH) {
            flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) { return val(1, c, h); }, rcur);
            flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) { return val(2, c, h); }, xlo);
        }
    }
}

// Last smoothing step on level 0: z = z1 + (r - A z1) / theta, with z1 (in
// t0) the prolongated correction (FP16 copy of A, as in k_pcg_pre); partials
// <r_k, z> and <r_{k-1}, z>.
template <typename VT, int NR, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_pcg_post(Slices sl, Partials P, SysState st, int k,
                                                       const _Float16* __restrict__ A, const float* __restrict__ theta,
                                                       const VT* __restrict__ r, const VT* __restrict__ rprev,
                                                       const float* __restrict__ z1, VT* __restrict__ z) {
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int4 ma = sl.meta[3 * w], mb = sl.meta[3 * w + 1];
    const int amask = st.slice_act[w];
    if (!amask) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + lane;
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    auto zz = [&](int64_t node, int c, int h) { return (VT)z1[(6 * node + c) * NR + h]; };
    const VT th = (VT)theta[ma.x];
    // own rows of z1, r and r_{k-1} requested before the product (the row update
    // after it then waits for no load)
    VT acc[6][NR], ro[6][NR], rpo[6][NR], zo[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * ii + c) * NR + h;
            zo[c][h] = zz(ii, c, h);
            ro[c][h] = r[o];
            rpo[c][h] = k > 0 ? rprev[o] : VT(0);
            acc[c][h] = wave == 0 ? zo[c][h] : VT(0);
        }
    ell_apply<_Float16, VT, NR, WPS>(sl, A, w, ma, mb, acc, zz);
    __shared__ VT stage[64 * 6 * NR];
    double v[NR], vp[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) v[h] = vp[h] = 0.0;
    // (z of inactive load cases is never used)
    flush_rows<VT, NR, WPS>(stage, ma, [&](int c, int h) {
        const VT rv = ro[c][h];
        const VT zn = zo[c][h] + (rv - acc[c][h]) / th;
        if (on && ((amask >> h) & 1)) {
            v[h] += (double)rv * (double)zn;
            if (k > 0) vp[h] += (double)rpo[c][h] * (double)zn;   // (no r_{-1} at the start)
        }
        return zn;
    }, z);
    if (wave != 0) return;
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        v[h] = wave_sum(v[h]);
        vp[h] = wave_sum(vp[h]);
    }
    if (lane == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            P.cg[k & 1][((int64_t)w * 4 + CG_RR) * NR + h] = v[h];
            P.cg[k & 1][((int64_t)w * 4 + CG_RQ) * NR + h] = vp[h];
        }
    }
}

// Update of the iteration that closes a chunk, folded into the FP64 solution:
// alpha as above; x64 += xlo + alpha p_k; xlo = 0.  (r is replaced next.)
template <typename VT, int NR>
__global__ __launch_bounds__(SOLVER_WG) void k_pcg_fold(Slices sl, Partials P, SysState st, int k,
                                                        const VT* __restrict__ p, VT* __restrict__ xlo,
                                                        double* __restrict__ x64) {
    __shared__ double lds[NWAVES * 2 * NR];
    const int w = slice_of_block(sl);
    if (w < 0) return;
    const int sys = sl.sys[w];
    if (structure_done<NR>(st, sys)) return;
    const int s0 = sl.sys_off[sys], s1 = sl.sys_off[sys + 1];
    const double* rzp = P.cg[(k - 1) & 1];
    const double* pqp = P.cg[k & 1];
    // this thread's row, requested before the reduction
    const int64_t R = 6 * (int64_t)sl.row0[w] + threadIdx.x;
    const bool row = R < 6 * (int64_t)sl.row1[w];
    const int64_t Rr = row ? R : 6 * (int64_t)sl.row0[w];
    VT xl[NR], pv[NR];
    double xv[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        const int64_t o = Rr * NR + h;
        xl[h] = xlo[o];
        pv[h] = p[o];
        xv[h] = x64[o];
    }
    double v[2 * NR];
#pragma unroll
    for (int a = 0; a < 2 * NR; ++a) v[a] = 0.0;
    for (int q = s0 + (int)threadIdx.x; q < s1; q += SOLVER_WG) {
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            v[h] += rzp[((int64_t)q * 4 + CG_RR) * NR + h];
            v[NR + h] += pqp[((int64_t)q * 4 + CG_PQ) * NR + h];
        }
    }
    group_sum<2 * NR, NWAVES>(v, lds);
    if (!row) return;
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        if (!st.active[sys * NR + h]) continue;
        const double rz = v[h], pq = v[NR + h];
        const VT alpha = (pq > 0.0 && isfinite(pq) && rz > 0.0) ? VT(rz / pq) : VT(0);
        const int64_t o = R * NR + h;
        x64[o] = xv[h] + (double)(xl[h] + alpha * pv[h]);
        xlo[o] = VT(0);
    }
}

// Coarse-level smoothing products (degree-2 Chebyshev, centre theta, weight c):
//   SM_PRE1:  t = b - A b / theta                       (v = b)
//   SM_PRE2:  r = b - c (theta (b - t) + A t)           (v = t): the residual of x = c (b + t)
//   SM_POST1: r0 = b - A z1                             (v = z1)
//   SM_POST2: x = z1 + c (2 r0 - A r0 / theta)          (v = r0, w = z1)
// written as out = cb b + cv v + ca (A v) + w.
enum { SM_PRE1 = 0, SM_PRE2 = 1, SM_POST1 = 2, SM_POST2 = 3 };
struct SmoothCoef {
    float cb, cv, ca;
    bool use_b, use_w;
};
__device__ __forceinline__ SmoothCoef smooth_coef(int mode, float th, float c) {
    switch (mode) {
        case SM_PRE1: return {1.f, 0.f, -1.f / th, true, false};
        case SM_PRE2: return {1.f - c * th, c * th, -c, true, false};
        case SM_POST1: return {1.f, 0.f, -1.f, true, false};
        default: return {0.f, 2.f * c, -c / th, false, true};
    }
}

template <int NR, int WPS>
__global__ __launch_bounds__(64 * WPS) void k_mg_smooth(Slices sl, const float* __restrict__ A,
                                                        const float* __restrict__ theta,
                                                        const float* __restrict__ cheb,
                                                        const uint8_t* __restrict__ smooth, SysState st, int mode,
                                                        const float* __restrict__ b, const float* __restrict__ v,
                                                        const float* __restrict__ w, float* __restrict__ out) {
    const int q = slice_of_block(sl);
    if (q < 0) return;
    const int4 ma = sl.meta[3 * q], mb = sl.meta[3 * q + 1];
    if (!smooth[ma.x] || structure_done<NR>(st, ma.x)) return;
    const int wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)ma.y + (threadIdx.x & 63);
    const bool on = i < ma.z;
    const int64_t ii = on ? i : (int64_t)ma.y;
    auto vv = [&](int64_t node, int c, int h) { return v[(6 * node + c) * NR + h]; };
    const SmoothCoef k = smooth_coef(mode, theta[ma.x], cheb[ma.x]);
    float acc[6][NR];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[c][h] = wave == 0 ? vv(ii, c, h) : 0.f;
    ell_apply<float, float, NR, WPS>(sl, A, q, ma, mb, acc, vv);
    if (wave != 0 || !on) return;
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            const int64_t o = (6 * i + c) * NR + h;
            float y = k.cv * vv(i, c, h) + k.ca * acc[c][h];
            if (k.use_b) y += k.cb * b[o];
            if (k.use_w) y += w[o];
            out[o] = y;
        }
}

// sum over the 4 lanes of an aligned quad (butterfly: all four hold the same bits)
__device__ __forceinline__ float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}

// Restriction b_c = P^T t: four lanes per coarse DOF row (J, c), each taking
// every fourth block of the column (stored transposed and contiguous per
// coarse node: PT[q][c][r], FP16), then a quad sum.  Structures flagged in
// `skip` restrict inside their fused kernel.
template <typename VT, int NR>
__global__ __launch_bounds__(256) void k_mg_restrict(int32_t nc, const int32_t* __restrict__ node_sys_c,
                                                     const uint8_t* __restrict__ skip,
                                                     const int32_t* __restrict__ pt_ptr,
                                                     const int32_t* __restrict__ pt_row,
                                                     const _Float16* __restrict__ PT, SysState st,
                                                     const VT* __restrict__ t, float* __restrict__ bc) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(u & 3);
    const int64_t jc = u >> 2;
    const bool valid = jc < 6 * (int64_t)nc;
    const int J = valid ? (int)(jc / 6) : 0, c = (int)(jc % 6);
    // (the column's range is requested together with the structure's flags)
    const int qa = pt_ptr[J] + g, q1 = pt_ptr[J + 1];
    const bool run = valid && !skip[node_sys_c[J]] && !structure_done<NR>(st, node_sys_c[J]);
    float acc[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) acc[h] = 0.f;
    if (run) {
#pragma unroll 4
        for (int q = qa; q < q1; q += 4) {
            const int64_t i = pt_row[q];
            const _Float16* pr = PT + (int64_t)q * 36 + c * 6;
            float pv[6];
#pragma unroll
            for (int r = 0; r < 6; ++r) pv[r] = (float)pr[r];
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[h] += pv[r] * (float)t[(6 * i + r) * NR + h];
        }
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) acc[h] = quad_sum(acc[h]);
    if (run && g == 0) {
#pragma unroll
        for (int h = 0; h < NR; ++h) bc[jc * NR + h] = acc[h];
    }
}

// Prolongation and the pre-smoothing result before it, P in FP16, thread per
// fine DOF row (i, r) of the structures that have a coarser level:
//   level 0 (cheb null):  z1 = b / theta + P x_c
//   coarse levels:        z1 = c (b + t) + P x_c   (t = b - A b / theta from SM_PRE1)
template <typename VT, typename ZT, int NR>
__global__ __launch_bounds__(256) void k_mg_prolong(int32_t n, const int32_t* __restrict__ node_sys,
                                                    const uint8_t* __restrict__ cont, const int32_t* __restrict__ p_ptr,
                                                    const int32_t* __restrict__ p_col, const _Float16* __restrict__ P,
                                                    const float* __restrict__ theta, const float* __restrict__ cheb,
                                                    SysState st, const VT* __restrict__ b,
                                                    const float* __restrict__ t, const float* __restrict__ xc,
                                                    ZT* __restrict__ z1) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= 6 * (int64_t)n) return;
    const int i = (int)(u / 6), r = (int)(u % 6);
    // (the row's range and vector entries are requested together with the structure's flags)
    const int qa = p_ptr[i], qb = p_ptr[i + 1];
    float bv[NR], tv[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        bv[h] = (float)b[u * NR + h];
        tv[h] = cheb ? t[u * NR + h] : 0.f;
    }
    const int s = node_sys[i];
    if (!cont[s] || structure_done<NR>(st, s)) return;
    float acc[NR];
    if (cheb) {
        const float c = cheb[s];
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = c * (bv[h] + tv[h]);
    } else {
        const float th = theta[s];
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = bv[h] / th;
    }
    // four blocks per round, their loads issued together (a block past the end
    // re-reads the row's first one and is not added)
    for (int q0 = qa; q0 < qb; q0 += 4) {
        int64_t J[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) J[v] = p_col[q0 + v < qb ? q0 + v : qa];
        float pv[4][6], xv[4][6][NR];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const _Float16* pb = P + (int64_t)(q0 + v < qb ? q0 + v : qa) * 36 + r * 6;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                pv[v][c] = (float)pb[c];
#pragma unroll
                for (int h = 0; h < NR; ++h) xv[v][c][h] = xc[(6 * J[v] + c) * NR + h];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (q0 + v < qb) {
#pragma unroll
                for (int c = 0; c < 6; ++c)
#pragma unroll
                    for (int h = 0; h < NR; ++h) acc[h] += pv[v][c] * xv[v][c][h];
            }
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) z1[u * NR + h] = (ZT)acc[h];
}

// Coarsest levels: x = A^-1 b with the dense (symmetric) inverse of each
// structure solved at this level: workgroup per structure, b staged in LDS,
// thread per row r reading A^-1 by columns (entry (c, r) = (r, c): coalesced
// across the threads) with 8 independent partial sums
template <int NR>
__global__ __launch_bounds__(256) void k_mg_direct(const int32_t* __restrict__ list, const int32_t* __restrict__ term,
                                                   const float* __restrict__ inv, SysState st,
                                                   const float* __restrict__ b, float* __restrict__ x) {
    __shared__ float bs[AMG_DENSE_MAX * NR];
    const int s = list[blockIdx.x];
    if (structure_done<NR>(st, s)) return;
    const int64_t node0 = term[4 * s + 1];
    const int N = 6 * term[4 * s + 2];
    const float* Ai = inv + term[4 * s + 3];
    for (int u = threadIdx.x; u < N * NR; u += blockDim.x) bs[u] = b[6 * node0 * NR + u];
    __syncthreads();
    constexpr int U = 8;
    for (int r = threadIdx.x; r < N; r += blockDim.x) {
        float acc[U][NR];
#pragma unroll
        for (int q = 0; q < U; ++q)
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[q][h] = 0.f;
        int c = 0;
        for (; c + U <= N; c += U) {
            float a[U];
#pragma unroll
            for (int q = 0; q < U; ++q) a[q] = Ai[(int64_t)(c + q) * N + r];
#pragma unroll
            for (int q = 0; q < U; ++q)
#pragma unroll
                for (int h = 0; h < NR; ++h) acc[q][h] += a[q] * bs[(c + q) * NR + h];
        }
        for (; c < N; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[0][h] += Ai[(int64_t)c * N + r] * bs[c * NR + h];
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            float t = 0.f;
#pragma unroll
            for (int q = 0; q < U; ++q) t += acc[q][h];
            x[(6 * node0 + r) * NR + h] = t;
        }
    }
}

// A structure's last sparse level l and its dense level l + 1 in one kernel
// (workgroup per structure; every step ends at a barrier, and the level-l
// vectors, a few KB per structure, pass between the steps through global
// memory inside the workgroup), with the coarse levels' degree-2 Chebyshev
// smoothing (k_mg_smooth's modes; TG slices at a time, TW waves per slice):
//   t = b - A b / theta,  r = b - c (theta (b - t) + A t)    (pre-smoothing)
//   b_c = P^T r                  (four lanes per coarse row, into LDS)
//   x_c = A_c^-1 b_c             (dense inverse, columns split over the workgroup)
//   z1 = c (b + t) + P x_c       (into r)
//   r0 = b - A z1 (into t),  x = z1 + c (2 r0 - A r0 / theta)  (post-smoothing;
//   x is the level's result, read by the prolongation above)
// It replaces seven launches per iteration on a small, latency-bound level.
template <int NR>
struct TailShape {
    static constexpr int TW = NR <= 2 ? 4 : (NR <= 4 ? 2 : 1);  // waves per slice
    static constexpr int TG = 4;                                 // slices at a time
    static constexpr int THREADS = 64 * TW * TG;
    static constexpr int ROWS = 160;                             // >= AMG_DENSE_MAX
    static constexpr int CHUNKS = THREADS / ROWS;                // column chunks of the dense product
    // entries of the inverse each thread loads at the start (its whole column
    // chunk of its row) when that fits the registers
    static constexpr int PREFETCH = CHUNKS >= 6 ? (AMG_DENSE_MAX + CHUNKS - 1) / CHUNKS : 0;
    // structures of at most VEC_NODES nodes keep the level's four vectors in LDS
    // (48 KB); larger ones pass them through global memory
    static constexpr int VEC_NODES = 512 / NR;
};
constexpr int TAIL_NR[5] = {1, 2, 3, 4, 8};   // right-hand-side counts of the solver
static_assert(TailShape<1>::ROWS >= AMG_DENSE_MAX, "dense rows exceed the tail kernel's row span");

// (vectors indexed from node nb: the structure's first node when they live in LDS, 0 in global memory)
template <int NR>
__device__ __forceinline__ void tail_smooth(const Slices& sl, const float* __restrict__ A, int q0, int q1,
                                            const SmoothCoef& k, int64_t nb, const float* __restrict__ b,
                                            const float* __restrict__ src, const float* __restrict__ w,
                                            float* __restrict__ dst) {
    using S = TailShape<NR>;
    const int lane = threadIdx.x & 63, grp = (threadIdx.x >> 6) / S::TW, sub = (threadIdx.x >> 6) % S::TW;
    auto xs = [&](int64_t node, int c, int h) { return src[(6 * (node - nb) + c) * NR + h]; };
    if (q1 - q0 == 1) {
        // a single slice: every wavefront of the workgroup takes a share of its
        // ELL slots (a coarse row has many), rather than one group of TW
        constexpr int W = S::TW * S::TG;
        const int wave = threadIdx.x >> 6;
        const int4 ma = sl.meta[3 * q0], mb = sl.meta[3 * q0 + 1];
        const int64_t i = (int64_t)ma.y + lane;
        const bool on = i < ma.z;
        const int64_t ii = on ? i : (int64_t)ma.y;
        float acc[6][NR];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[c][h] = wave == 0 ? xs(ii, c, h) : 0.f;
        ell_apply<float, float, NR, W, 1>(sl, A, q0, ma, mb, acc, xs);
        if (wave == 0 && on) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) {
                    const int64_t o = (6 * (i - nb) + c) * NR + h;
                    float y = k.cv * src[o] + k.ca * acc[c][h];
                    if (k.use_b) y += k.cb * b[o];
                    if (k.use_w) y += w[o];
                    dst[o] = y;
                }
        }
        __syncthreads();
        return;
    }
    for (int base = q0; base < q1; base += S::TG) {
        // every group runs a slice (groups past the end repeat the last one) so that
        // all waves meet the same barriers; only owners store
        const bool own = base + grp < q1;
        const int q = own ? base + grp : q1 - 1;
        const int4 ma = sl.meta[3 * q], mb = sl.meta[3 * q + 1];
        const int64_t i = (int64_t)ma.y + lane;
        const bool on = i < ma.z;
        const int64_t ii = on ? i : (int64_t)ma.y;
        float acc[6][NR];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[c][h] = sub == 0 ? xs(ii, c, h) : 0.f;
        ell_apply<float, float, NR, S::TW, S::TG>(sl, A, q, ma, mb, acc, xs);
        if (own && sub == 0 && on) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int h = 0; h < NR; ++h) {
                    const int64_t o = (6 * (i - nb) + c) * NR + h;
                    float y = k.cv * src[o] + k.ca * acc[c][h];
                    if (k.use_b) y += k.cb * b[o];
                    if (k.use_w) y += w[o];
                    dst[o] = y;
                }
        }
        __syncthreads();  // the cross-wave buffer is reused by the next round
    }
}

template <int NR, bool LV>
__global__ __launch_bounds__(TailShape<NR>::THREADS) void k_mg_tail(
    const int32_t* __restrict__ list, Slices sl, const float* __restrict__ A, const float* __restrict__ theta,
    const float* __restrict__ cheb, const int32_t* __restrict__ node_off, const int32_t* __restrict__ pt_ptr, const int32_t* __restrict__ pt_row,
    const _Float16* __restrict__ PT, const int32_t* __restrict__ p_ptr, const int32_t* __restrict__ p_col,
    const _Float16* __restrict__ P, const int32_t* __restrict__ term, const float* __restrict__ inv, SysState st,
    const float* __restrict__ bg, float* __restrict__ tg, float* __restrict__ rg, float* __restrict__ xg) {
    using S = TailShape<NR>;
    __shared__ float bc[AMG_DENSE_MAX * NR], xc[AMG_DENSE_MAX * NR];
    __shared__ float part[S::CHUNKS][S::ROWS * NR];
    __shared__ float vec[LV ? 4 : 1][LV ? S::VEC_NODES * 6 * NR : 1];
    const int s = list[blockIdx.x];
    if (structure_done<NR>(st, s)) return;
    const int tid = threadIdx.x;
    const int q0 = sl.sys_off[s], q1 = sl.sys_off[s + 1];
    const float th = theta[s], ch = cheb[s];
    const int64_t n0 = node_off[s], n1 = node_off[s + 1];
    // the level's vectors b, t, r, x: in LDS from node n0 (LV), else in global memory
    const int64_t nb = LV ? n0 : 0;
    const float* b = LV ? vec[0] : bg;
    float* t = LV ? vec[1] : tg;
    float* r = LV ? vec[2] : rg;
    float* x = LV ? vec[3] : xg;
    const int64_t c0 = term[4 * s + 1];
    const int N = 6 * term[4 * s + 2];
    const float* Ai = inv + term[4 * s + 3];
    // the dense inverse does not depend on anything computed here: its loads are
    // issued first and arrive while b is staged and the smoothing runs
    const int dr = tid % S::ROWS, dk = tid / S::ROWS;
    const bool dense_lane = dk < S::CHUNKS && dr < N;
    float ainv[S::PREFETCH > 0 ? S::PREFETCH : 1];
#pragma unroll
    for (int m = 0; m < S::PREFETCH; ++m) {
        const int c = dk + m * S::CHUNKS;
        ainv[m] = dense_lane && c < N ? Ai[(int64_t)c * N + dr] : 0.f;
    }
    if constexpr (LV) {
        for (int64_t u = tid; u < 6 * (n1 - n0) * NR; u += S::THREADS) vec[0][u] = bg[6 * n0 * NR + u];
        __syncthreads();
    }
    // 1. pre-smoothing: t, then the residual r of x = c (b + t)
    tail_smooth<NR>(sl, A, q0, q1, smooth_coef(SM_PRE1, th, ch), nb, b, b, nullptr, t);
    tail_smooth<NR>(sl, A, q0, q1, smooth_coef(SM_PRE2, th, ch), nb, b, t, nullptr, r);
    // 2. restriction of r into LDS: lanes 4u..4u+3 take every fourth block of coarse row u
    for (int u0 = 0; u0 < 4 * N; u0 += S::THREADS) {
        const int v = u0 + tid;
        const bool valid = v < 4 * N;
        const int u = valid ? v >> 2 : 0, g = v & 3;
        const int64_t J = c0 + u / 6;
        const int c = u % 6;
        float acc[NR];
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = 0.f;
        if (valid) {
            const int e1 = pt_ptr[J + 1];
#pragma unroll 4
            for (int q = pt_ptr[J] + g; q < e1; q += 4) {
                const int64_t i = pt_row[q];
                const _Float16* pr = PT + (int64_t)q * 36 + c * 6;
                float pv[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) pv[m] = (float)pr[m];
#pragma unroll
                for (int m = 0; m < 6; ++m)
#pragma unroll
                    for (int h = 0; h < NR; ++h) acc[h] += pv[m] * r[(6 * (i - nb) + m) * NR + h];
            }
        }
#pragma unroll
        for (int h = 0; h < NR; ++h) acc[h] = quad_sum(acc[h]);
        if (valid && g == 0) {
#pragma unroll
            for (int h = 0; h < NR; ++h) bc[u * NR + h] = acc[h];
        }
    }
    __syncthreads();
    // 3. dense solve: thread (chunk k, row r) sums columns k, k + CHUNKS, ... of
    // row r (the inverse is symmetric: read by columns, coalesced over r)
    if constexpr (S::PREFETCH > 0) {
        if (dense_lane) {
            float a0[NR], a1[NR];
#pragma unroll
            for (int h = 0; h < NR; ++h) a0[h] = a1[h] = 0.f;
#pragma unroll
            for (int m = 0; m < S::PREFETCH; m += 2) {
                const int c0 = dk + m * S::CHUNKS, c1 = c0 + S::CHUNKS;
#pragma unroll
                for (int h = 0; h < NR; ++h) {
                    if (c0 < N) a0[h] += ainv[m] * bc[c0 * NR + h];
                    if (m + 1 < S::PREFETCH && c1 < N) a1[h] += ainv[m + 1] * bc[c1 * NR + h];
                }
            }
#pragma unroll
            for (int h = 0; h < NR; ++h) part[dk][dr * NR + h] = a0[h] + a1[h];
        }
    } else {
        const int r = tid % S::ROWS, k = tid / S::ROWS;
        if (k < S::CHUNKS && r < N) {
            float a0[NR], a1[NR];
#pragma unroll
            for (int h = 0; h < NR; ++h) a0[h] = a1[h] = 0.f;
            int c = k;
            for (; c + S::CHUNKS < N; c += 2 * S::CHUNKS) {
                const float v0 = Ai[(int64_t)c * N + r], v1 = Ai[(int64_t)(c + S::CHUNKS) * N + r];
#pragma unroll
                for (int h = 0; h < NR; ++h) {
                    a0[h] += v0 * bc[c * NR + h];
                    a1[h] += v1 * bc[(c + S::CHUNKS) * NR + h];
                }
            }
            if (c < N) {
                const float v0 = Ai[(int64_t)c * N + r];
#pragma unroll
                for (int h = 0; h < NR; ++h) a0[h] += v0 * bc[c * NR + h];
            }
#pragma unroll
            for (int h = 0; h < NR; ++h) part[k][r * NR + h] = a0[h] + a1[h];
        }
    }
    __syncthreads();
    for (int u = tid; u < N * NR; u += S::THREADS) {
        float v = 0.f;
#pragma unroll
        for (int k = 0; k < S::CHUNKS; ++k) v += part[k][u];
        xc[u] = v;
    }
    __syncthreads();
    // 4. prolongation with the pre-smoothing result, z1 = c (b + t) + P x_c, into r
    // (rows u and u + THREADS per thread, block m of both in the same step, so that
    // their index and value loads are in flight together)
    for (int64_t ub = 6 * n0 + tid; ub < 6 * n1; ub += 2 * S::THREADS) {
        bool ok[2];
        int qs[2], qe[2], row[2];
        int64_t ul[2];
        float acc[2][NR];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int64_t u = ub + a * S::THREADS;
            ok[a] = u < 6 * n1;
            const int64_t uc = ok[a] ? u : ub;
            const int64_t i = uc / 6;
            row[a] = (int)(uc % 6);
            ul[a] = uc - 6 * nb;
            qs[a] = p_ptr[i];
            qe[a] = ok[a] ? p_ptr[i + 1] : qs[a];
#pragma unroll
            for (int h = 0; h < NR; ++h) acc[a][h] = ch * (b[ul[a] * NR + h] + t[ul[a] * NR + h]);
        }
        const int len = qe[0] - qs[0] > qe[1] - qs[1] ? qe[0] - qs[0] : qe[1] - qs[1];
        for (int m = 0; m < len; ++m) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (qs[a] + m >= qe[a]) continue;
                const int q = qs[a] + m;
                const int J = p_col[q] - (int)c0;
                const _Float16* pb = P + (int64_t)q * 36 + row[a] * 6;
#pragma unroll
                for (int c = 0; c < 6; ++c)
#pragma unroll
                    for (int h = 0; h < NR; ++h) acc[a][h] += (float)pb[c] * xc[(6 * J + c) * NR + h];
            }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
            if (ok[a]) {
#pragma unroll
                for (int h = 0; h < NR; ++h) r[ul[a] * NR + h] = acc[a][h];
            }
    }
    __syncthreads();
    // 5. post-smoothing: r0 = b - A z1 into t, then x
    tail_smooth<NR>(sl, A, q0, q1, smooth_coef(SM_POST1, th, ch), nb, b, r, nullptr, t);
    tail_smooth<NR>(sl, A, q0, q1, smooth_coef(SM_POST2, th, ch), nb, b, t, r, x);
    if constexpr (LV) {   // tail_smooth ended at a barrier
        for (int64_t u = tid; u < 6 * (n1 - n0) * NR; u += S::THREADS) xg[6 * n0 * NR + u] = vec[3][u];
    }
}

// ---------------------------------------------------------------------------
// host: device copy of the hierarchy, numeric set-up, V-cycle, driver
// ---------------------------------------------------------------------------
template <typename T>
static int amg_alloc(AmgState& A, T** dst, size_t count) {
    *dst = nullptr;
    if (count == 0) return JSSO_OK;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "AMG hipMalloc(%zu bytes): %s", count * sizeof(T),
                 hipGetErrorString(e));
    A.allocs.push_back(q);
    JSSO_HIP(hipMemset(q, 0, count * sizeof(T)));
    *dst = (T*)q;
    return JSSO_OK;
}

template <typename T>
static int amg_upload(AmgState& A, T** dst, const std::vector<T>& src) {
    int rc = amg_alloc(A, dst, src.size());
    if (rc || src.empty()) return rc;
    JSSO_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return JSSO_OK;
}

static Slices level_slices(Plan* pl, AmgState& A, int l) {
    if (l == 0) return make_slices(pl, 0, pl->n_wg);
    const AmgLevelDev& D = A.dev[l];
    return Slices{D.d_wg_sys, D.d_wg_row0, D.d_wg_row1, D.d_sys_wg_off, D.d_ell_vpos, D.d_ell_col,
                  (const int4*)D.d_meta, 0, D.n_wg, slice_run(D.n_wg, pl->slice_spread)};
}

static int amg_build_device(Plan* pl, AmgState& A) {
    const int nl = A.nl, ns = pl->n_sys;
    A.dev.assign(nl, AmgLevelDev{});
    A.d_Ac.assign(nl, nullptr);
    A.d_Lc.assign(nl, nullptr);
    A.d_Lci.assign(nl, nullptr);
    A.d_cdiag.assign(nl, nullptr);
    int rc = 0;
#define AMG_TRY(x) \
    if ((rc = (x))) return rc
    for (int l = 0; l < nl; ++l) {
        AmgHostLevel& H = A.host[l];
        AmgLevelDev& D = A.dev[l];
        D.n = H.n;
        AMG_TRY(amg_upload(A, &D.d_node_sys, H.node_sys));
        AMG_TRY(amg_upload(A, &D.d_cont, H.cont));
        if (l == 0) {
            D.n_wg = pl->n_wg;
        } else {
            const EllPattern& E = H.ell;
            D.n_wg = E.n_wg;
            D.ell_slots = E.ell_slots;
            std::vector<int32_t> browof(H.colind.size()), cdiag(H.n);
            for (int i = 0; i < H.n; ++i)
                for (int k = H.rowptr[i]; k < H.rowptr[i + 1]; ++k) {
                    browof[k] = i;
                    if (H.colind[k] == i) cdiag[i] = k;
                }
            AMG_TRY(amg_upload(A, &D.d_rowptr, H.rowptr));
            AMG_TRY(amg_upload(A, &D.d_colind, H.colind));
            AMG_TRY(amg_upload(A, &D.d_kref, E.kref));
            AMG_TRY(amg_upload(A, &D.d_browof, browof));
            AMG_TRY(amg_upload(A, &A.d_cdiag[l], cdiag));
            AMG_TRY(amg_upload(A, &D.d_wg_sys, E.wg_sys));
            AMG_TRY(amg_upload(A, &D.d_wg_row0, E.wg_row0));
            AMG_TRY(amg_upload(A, &D.d_wg_row1, E.wg_row1));
            AMG_TRY(amg_upload(A, &D.d_sys_wg_off, E.sys_wg_off));
            AMG_TRY(amg_upload(A, &D.d_ell_col, E.ell_col));
            AMG_TRY(amg_upload(A, &D.d_ell_vpos, E.ell_vpos));
            AMG_TRY(amg_upload(A, &D.d_meta, E.meta));
            AMG_TRY(amg_upload(A, &D.d_ell_src, E.ell_src));
            AMG_TRY(amg_alloc(A, &D.d_A64, (size_t)E.ell_slots * 36));
            AMG_TRY(amg_alloc(A, &D.d_A32, (size_t)E.ell_slots * 36));
            AMG_TRY(amg_alloc(A, &A.d_Ac[l], H.colind.size() * 36));
            AMG_TRY(amg_alloc(A, &A.d_Lc[l], (size_t)H.n * 36));
            AMG_TRY(amg_alloc(A, &A.d_Lci[l], (size_t)H.n * 36));
        }
        const bool has_p = l + 1 < nl;
        if (!has_p) continue;
        D.n_agg = (int32_t)H.agg_ptr.size() - 1;
        D.nP = H.p_col.size();
        D.nW = H.w_col.size();
        AMG_TRY(amg_upload(A, &D.d_agg, H.agg));
        AMG_TRY(amg_upload(A, &D.d_agg_ptr, H.agg_ptr));
        AMG_TRY(amg_upload(A, &D.d_agg_nodes, H.agg_nodes));
        AMG_TRY(amg_upload(A, &D.d_p_ptr, H.p_ptr));
        AMG_TRY(amg_upload(A, &D.d_p_col, H.p_col));
        AMG_TRY(amg_upload(A, &D.d_p_row, H.p_row));
        AMG_TRY(amg_upload(A, &D.d_p_cptr, H.p_cptr));
        AMG_TRY(amg_upload(A, &D.d_p_cj, H.p_cj));
        AMG_TRY(amg_upload(A, &D.d_p_caref, H.p_caref));
        AMG_TRY(amg_upload(A, &D.d_pt_ptr, H.pt_ptr));
        AMG_TRY(amg_upload(A, &D.d_pt_row, H.pt_row));
        AMG_TRY(amg_upload(A, &D.d_p_tpos, H.p_tpos));
        AMG_TRY(amg_upload(A, &D.d_w_cptr, H.w_cptr));
        AMG_TRY(amg_upload(A, &D.d_w_caref, H.w_caref));
        AMG_TRY(amg_upload(A, &D.d_w_cp, H.w_cp));
        AMG_TRY(amg_upload(A, &D.d_c_cptr, H.c_cptr));
        AMG_TRY(amg_upload(A, &D.d_c_p, H.c_p));
        AMG_TRY(amg_upload(A, &D.d_c_w, H.c_w));
        AMG_TRY(amg_alloc(A, &D.d_B, (size_t)H.n * 36));
        AMG_TRY(amg_alloc(A, &D.d_Q, (size_t)H.n * 36));
        AMG_TRY(amg_alloc(A, &D.d_R, (size_t)D.n_agg * 36));
        AMG_TRY(amg_alloc(A, &D.d_Pf, (size_t)D.nP * 36));
        AMG_TRY(amg_alloc(A, &D.d_P16, (size_t)D.nP * 36));
        AMG_TRY(amg_alloc(A, &D.d_PT16, (size_t)D.nP * 36));
        AMG_TRY(amg_alloc(A, &D.d_Wf, (size_t)D.nW * 36));
        AMG_TRY(amg_alloc(A, &D.d_rho, (size_t)ns));
        AMG_TRY(amg_alloc(A, &D.d_omega, (size_t)ns));
        AMG_TRY(amg_alloc(A, &D.d_theta, (size_t)ns));
        AMG_TRY(amg_alloc(A, &D.d_cheb, (size_t)ns));
        AMG_TRY(amg_alloc(A, &D.d_pv[0], 6 * (size_t)H.n));
        AMG_TRY(amg_alloc(A, &D.d_pv[1], 6 * (size_t)H.n));
        AMG_TRY(amg_alloc(A, &D.d_ppart, 2 * (size_t)D.n_wg));
    }
    AMG_TRY(amg_upload(A, &A.d_node_off0, pl->h_node_off));
    // dense coarsest solves
    A.h_term.assign(4 * (size_t)ns, -1);
    A.term_list.assign(nl, {});
    int64_t off = 0;
    for (int l = 0; l < nl; ++l)
        for (int s = 0; s < ns; ++s)
            if (A.host[l].term[s]) {
                const int cnt = A.host[l].node_off[s + 1] - A.host[l].node_off[s];
                A.h_term[4 * s] = l;
                A.h_term[4 * s + 1] = A.host[l].node_off[s];
                A.h_term[4 * s + 2] = cnt;
                A.h_term[4 * s + 3] = (int32_t)off;
                A.term_list[l].push_back(s);
                off += 36 * (int64_t)cnt * cnt;
                JSSO_REQUIRE(off < INT32_MAX, JSSO_EUNSUPPORTED, "AMG: dense coarse solves too large");
            }
    A.inv_total = off;
    AMG_TRY(amg_upload(A, &A.d_term, A.h_term));
    A.d_term_list.assign(nl, nullptr);
    for (int l = 0; l < nl; ++l) AMG_TRY(amg_upload(A, &A.d_term_list[l], A.term_list[l]));
    AMG_TRY(amg_alloc(A, &A.d_inv, (size_t)off));
    int64_t max_pos = pl->ell_slots;
    for (int l = 1; l < nl; ++l) max_pos = std::max<int64_t>(max_pos, A.host[l].ell.ell_slots);
    AMG_TRY(amg_alloc(A, &A.d_Anat, (size_t)max_pos * 36));
    // fused last sparse levels (JSSO_AMG_FUSE=0 keeps every level on the per-level kernels)
    const char* fe = getenv("JSSO_AMG_FUSE");
    const bool fuse = !(fe && fe[0] == '0');
    std::vector<int> fl(ns, -1);
    for (int s = 0; s < ns; ++s)
        if (fuse && A.h_term[4 * s] >= 2) fl[s] = A.h_term[4 * s] - 1;
    A.tail_list.assign(nl, {});
    A.direct_list.assign(nl, {});
    A.d_tail_list.assign(nl, nullptr);
    A.d_direct_list.assign(nl, nullptr);
    A.d_node_off.assign(nl, nullptr);
    A.d_smooth.assign(nl, nullptr);
    A.d_rskip.assign(nl, nullptr);
    for (int l = 0; l < nl; ++l) {
        std::vector<uint8_t> smooth(ns, 0), rskip(ns, 0);
        for (int s = 0; s < ns; ++s) {
            smooth[s] = A.host[l].cont[s] && fl[s] != l ? 1 : 0;
            rskip[s] = fl[s] == l ? 1 : 0;
            if (fl[s] == l) A.tail_list[l].push_back(s);
            if (A.host[l].term[s] && fl[s] != l - 1) A.direct_list[l].push_back(s);
        }
        AMG_TRY(amg_upload(A, &A.d_smooth[l], smooth));
        AMG_TRY(amg_upload(A, &A.d_rskip[l], rskip));
        AMG_TRY(amg_upload(A, &A.d_tail_list[l], A.tail_list[l]));
        for (int v = 0; v < 5; ++v) {
            const int cap = 512 / TAIL_NR[v];   // TailShape<NR>::VEC_NODES
            std::vector<int32_t> fit, big;
            for (int32_t st : A.tail_list[l])
                (A.host[l].node_off[st + 1] - A.host[l].node_off[st] <= cap ? fit : big).push_back(st);
            if (l == 0)
                for (int f = 0; f < 2; ++f) {
                    A.tail_n[v][f].assign(nl, 0);
                    A.d_tail_split[v][f].assign(nl, nullptr);
                }
            A.tail_n[v][0][l] = (int32_t)fit.size();
            A.tail_n[v][1][l] = (int32_t)big.size();
            AMG_TRY(amg_upload(A, &A.d_tail_split[v][0][l], fit));
            AMG_TRY(amg_upload(A, &A.d_tail_split[v][1][l], big));
        }
        AMG_TRY(amg_upload(A, &A.d_direct_list[l], A.direct_list[l]));
        AMG_TRY(amg_upload(A, &A.d_node_off[l], A.host[l].node_off));
    }
#undef AMG_TRY
    return JSSO_OK;
}

static inline unsigned nblk(int64_t threads, int wg) { return (unsigned)((threads + wg - 1) / wg); }

// Call f with the largest waves-per-slice count <= wps whose LDS buffer fits.
template <typename VT, int NR, typename F>
static void wps_dispatch(int wps, F&& f) {
    if constexpr (AmgWpsFits<VT, NR, 4>::value) {
        if (wps >= 4) {
            f(std::integral_constant<int, 4>{});
            return;
        }
    }
    if constexpr (AmgWpsFits<VT, NR, 2>::value) {
        if (wps >= 2) {
            f(std::integral_constant<int, 2>{});
            return;
        }
    }
    f(std::integral_constant<int, 1>{});
}

// Numeric set-up on the current assembly (stream-ordered).
static int amg_setup_numeric(Plan* pl, AmgState& A, hipStream_t s) {
    const int nl = A.nl, ns = pl->n_sys;
    for (int l = 0; l + 1 < nl; ++l) {
        AmgLevelDev& D = A.dev[l];
        AmgLevelDev& C = A.dev[l + 1];
        const AmgHostLevel& H = A.host[l];
        const double* A64 = l == 0 ? pl->d_Ke64 : D.d_A64;
        const Slices sl = level_slices(pl, A, l);
        const int32_t* sys_wg_off = l == 0 ? pl->d_sys_wg_off : D.d_sys_wg_off;
        if (l == 0)
            hipLaunchKernelGGL(k_amg_null0, xcd_grid(nblk(D.n, 256)), dim3(256), 0, s, D.n, D.d_agg, D.d_agg_ptr,
                               D.d_agg_nodes, (const double*)pl->d_crds, (const double*)pl->d_Linv,
                               (const uint8_t*)pl->d_fixed, D.d_B);
        if (D.n_agg > 0)
            hipLaunchKernelGGL(k_amg_qr, xcd_grid(D.n_agg), dim3(64), 0, s, D.n_agg, D.d_agg_ptr,
                               D.d_agg_nodes, (const double*)D.d_B, l == 0 ? (const double*)pl->d_crds : nullptr,
                               D.d_node_sys, A.d_node_off0, D.d_Q, D.d_R);
        for (int step = 0; step < AMG_POWER_STEPS; ++step) {
            if (l == 0)
                hipLaunchKernelGGL((k_amg_power<_Float16, 4>), slice_grid(sl), dim3(256), 0, s, sl,
                                   (const _Float16*)pl->d_Ke16, step, (const double*)D.d_pv[(step + 1) & 1],
                                   D.d_pv[step & 1], D.d_ppart, (const uint8_t*)D.d_cont, D.n_wg);
            else
                hipLaunchKernelGGL((k_amg_power<float, 4>), slice_grid(sl), dim3(256), 0, s, sl,
                                   (const float*)D.d_A32, step, (const double*)D.d_pv[(step + 1) & 1],
                                   D.d_pv[step & 1], D.d_ppart, (const uint8_t*)D.d_cont, D.n_wg);
        }
        hipLaunchKernelGGL(k_amg_theta, dim3(ns), dim3(64), 0, s, ns, sys_wg_off,
                           (const double*)D.d_ppart, AMG_POWER_STEPS - 1, (const uint8_t*)D.d_cont, D.n_wg, D.d_rho,
                           D.d_omega, D.d_theta, l == 0 ? nullptr : D.d_cheb);
        if (D.nP > 0) {
            const int64_t npos = l == 0 ? pl->ell_slots : D.ell_slots;
            hipLaunchKernelGGL(k_amg_unpack, xcd_grid(npos / ROWS_PER_WG), dim3(256), 0, s, npos / ROWS_PER_WG,
                               (const double2*)A64, (float2*)A.d_Anat);
            // (the lane-per-entry forms on levels of few blocks: the same sums, bit for bit)
            const bool pe = amg_entry_form(D.nP, AMG_ENTRY_PW_MAX), we = amg_entry_form(D.nW, AMG_ENTRY_PW_MAX);
            hipLaunchKernelGGL(pe ? k_amg_psmooth_e : k_amg_psmooth, xcd_grid(nblk(D.nP * (pe ? 36 : 1), 256)),
                               dim3(256), 0, s, D.nP, D.d_p_row, D.d_p_cptr, D.d_p_cj, D.d_p_caref, D.d_node_sys,
                               (const double*)D.d_omega, (const float*)A.d_Anat, (const double*)D.d_Q, D.d_Pf);
            hipLaunchKernelGGL(we ? k_amg_w_e : k_amg_w, xcd_grid(nblk(D.nW * (we ? 36 : 1), 256)), dim3(256), 0, s,
                               D.nW, D.d_w_cptr, D.d_w_caref, D.d_w_cp, (const float*)A.d_Anat, (const float*)D.d_Pf,
                               D.d_Wf);
        }
        const int64_t nnzc = A.host[l + 1].colind.size();
        if (nnzc > 0) {
            const bool ge = amg_entry_form(nnzc, AMG_ENTRY_C_MAX);
            hipLaunchKernelGGL(ge ? k_amg_galerkin_e : k_amg_galerkin, xcd_grid(nblk(nnzc * (ge ? 36 : 1), 256)),
                               dim3(256), 0, s, nnzc, D.d_c_cptr, D.d_c_p, D.d_c_w, (const float*)D.d_Pf,
                               (const float*)D.d_Wf, A.d_Ac[l + 1]);
            hipLaunchKernelGGL(k_amg_cfactor, dim3(nblk(C.n, 128)), dim3(128), 0, s, C.n, A.d_cdiag[l + 1],
                               (const double*)A.d_Ac[l + 1], A.d_Lc[l + 1], A.d_Lci[l + 1]);
            const int64_t nval = C.ell_slots - ROWS_PER_WG;
            if (nval > 0)
                hipLaunchKernelGGL(k_amg_cscale, xcd_grid(nblk(nval * 36, 256)), dim3(256), 0, s, nval, C.d_ell_src,
                                   C.d_browof, C.d_colind, (const double*)A.d_Ac[l + 1],
                                   (const double*)A.d_Lci[l + 1], C.d_A64, C.d_A32);
        }
        if (D.nP > 0)
            hipLaunchKernelGGL(k_amg_pscale, xcd_grid(nblk(D.nP * 6, 256)), dim3(256), 0, s, D.nP, D.d_p_col,
                               D.d_p_tpos, (const float*)D.d_Pf, (const double*)A.d_Lci[l + 1], D.d_P16,
                               D.d_PT16);
        if (C.d_B && C.n > 0)
            hipLaunchKernelGGL(k_amg_nullc, xcd_grid(nblk((int64_t)C.n * 36, 256)), dim3(256), 0, s, C.n,
                               (const double*)A.d_Lc[l + 1], (const double*)D.d_R, C.d_B);
        (void)H;
    }
    for (int l = 1; l < nl; ++l) {
        if (A.term_list[l].empty()) continue;
        const AmgLevelDev& D = A.dev[l];
        hipLaunchKernelGGL(k_amg_dense, dim3((unsigned)A.term_list[l].size()), dim3(1024), 0, s, A.d_term_list[l],
                           A.d_term, D.d_rowptr, D.d_colind, D.d_browof, D.d_kref, (const double*)D.d_A64, A.d_inv);
    }
    JSSO_HIP(hipGetLastError());
    A.built = true;
    A.power_next = AMG_POWER_STEPS;
    return JSSO_OK;
}

// Refresh of a kept hierarchy after an assembly (jsso_plan_amg_reuse): the
// coarse levels and the prolongators stay as built for an earlier design (the
// V-cycle remains a fixed symmetric positive definite preconditioner, only a
// less exact one as the design moves away); the level-0 smoothing follows the
// new matrix through its spectral bound, a few power-iteration steps started
// from the last eigenvector estimate.
static int amg_setup_light(Plan* pl, AmgState& A, hipStream_t s) {
    AmgLevelDev& D = A.dev[0];
    const Slices sl = level_slices(pl, A, 0);
    const int ns = pl->n_sys;
    const int s0 = A.power_next;
    for (int step = s0; step < s0 + AMG_POWER_STEPS_WARM; ++step)
        hipLaunchKernelGGL((k_amg_power<_Float16, 4>), slice_grid(sl), dim3(256), 0, s, sl,
                           (const _Float16*)pl->d_Ke16, step, (const double*)D.d_pv[(step + 1) & 1],
                           D.d_pv[step & 1], D.d_ppart, (const uint8_t*)D.d_cont, D.n_wg);
    const int last = s0 + AMG_POWER_STEPS_WARM - 1;
    hipLaunchKernelGGL(k_amg_theta, dim3(ns), dim3(64), 0, s, ns, (const int32_t*)pl->d_sys_wg_off,
                       (const double*)D.d_ppart, last, (const uint8_t*)D.d_cont, D.n_wg, D.d_rho, D.d_omega,
                       D.d_theta, (float*)nullptr);
    JSSO_HIP(hipGetLastError());
    A.power_next = last + 1;
    return JSSO_OK;
}

// Level-0 extra vectors (8 bytes per value) and the coarse V-cycle vectors for nrhs load cases.
static int amg_vectors(Plan* pl, AmgState& A, int nrhs) {
    const size_t zb = 6 * (size_t)pl->n_nodes * nrhs * sizeof(double);
    if (A.z_bytes < zb) {
        for (void** v : {&A.d_z, &A.d_t0, &A.d_p1}) {
            if (*v) JSSO_HIP(hipFree(*v));
            *v = nullptr;
            hipError_t e = hipMalloc(v, zb);
            JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "AMG vectors hipMalloc(%zu): %s", zb, hipGetErrorString(e));
        }
        A.z_bytes = zb;
        ++A.gen;
    }
    if (A.vec_nrhs < nrhs) {
        for (int l = 1; l < A.nl; ++l) {
            AmgLevelDev& D = A.dev[l];
            for (float** v : {&D.d_b, &D.d_t, &D.d_r, &D.d_x}) {
                if (*v) JSSO_HIP(hipFree(*v));
                *v = nullptr;
                const size_t bytes = 6 * (size_t)std::max(D.n, 1) * nrhs * sizeof(float);
                hipError_t e = hipMalloc((void**)v, bytes);
                JSSO_REQUIRE(e == hipSuccess, JSSO_ENOMEM, "AMG vectors hipMalloc(%zu): %s", bytes,
                             hipGetErrorString(e));
                JSSO_HIP(hipMemset(*v, 0, bytes));
            }
        }
        A.vec_nrhs = nrhs;
        ++A.gen;
    }
    return JSSO_OK;
}

// Level-0 buffers of one AMG solve
template <typename VT>
struct AmgBufs {
    VT* r[2];
    VT* p[2];
    VT *q, *z, *xlo;
    float* t0;   // level-0 preconditioner vector (pre-smoothing residual, then z1): FP32 in every precision
};

// The V-cycle after the level-0 pre-smoothing of iteration k (t0 holds
// r_k - A r_k / theta): restriction and pre-smoothing down the levels, the
// dense coarsest solves, prolongation and post-smoothing up; ends with z = M r_k
// and the partials <r_k, z>.
template <typename MT, typename VT, int NR>
static void amg_vcycle_rest(Plan* pl, AmgState& A, const Workspace& w, const AmgBufs<VT>& bf, const MT* A0, int k,
                            hipStream_t s) {
    const int nl = A.nl;
    for (int l = 0; l + 1 < nl; ++l) {
        const AmgLevelDev& D = A.dev[l];
        const AmgLevelDev& C = A.dev[l + 1];
        if (C.n == 0) continue;
        const dim3 gr(nblk(24 * (int64_t)C.n, 256));
        const bool any_restrict = A.tail_list[l].size() < (size_t)std::count(A.host[l].cont.begin(),
                                                                               A.host[l].cont.end(), 1);
        if (any_restrict) {
            if (l == 0)
                hipLaunchKernelGGL((k_mg_restrict<float, NR>), gr, dim3(256), 0, s, C.n, C.d_node_sys,
                                   (const uint8_t*)A.d_rskip[l], D.d_pt_ptr, D.d_pt_row, (const _Float16*)D.d_PT16,
                                   w.st, (const float*)bf.t0, C.d_b);
            else
                hipLaunchKernelGGL((k_mg_restrict<float, NR>), gr, dim3(256), 0, s, C.n, C.d_node_sys,
                                   (const uint8_t*)A.d_rskip[l], D.d_pt_ptr, D.d_pt_row, (const _Float16*)D.d_PT16,
                                   w.st, (const float*)D.d_r, C.d_b);
        }
        if (l + 2 < nl && A.tail_list[l + 1].size() < (size_t)std::count(A.host[l + 1].cont.begin(),
                                                                          A.host[l + 1].cont.end(), 1))
            wps_dispatch<float, NR>(4, [&](auto W) {
                constexpr int WP = decltype(W)::value;
                const Slices cs = level_slices(pl, A, l + 1);
                // t = b - A b / theta, then r = b - c (theta (b - t) + A t)
                hipLaunchKernelGGL((k_mg_smooth<NR, WP>), slice_grid(cs), dim3(64 * WP), 0, s, cs,
                                   (const float*)C.d_A32, (const float*)C.d_theta, (const float*)C.d_cheb,
                                   (const uint8_t*)A.d_smooth[l + 1], w.st, (int)SM_PRE1, (const float*)C.d_b,
                                   (const float*)C.d_b, (const float*)nullptr, C.d_t);
                hipLaunchKernelGGL((k_mg_smooth<NR, WP>), slice_grid(cs), dim3(64 * WP), 0, s, cs,
                                   (const float*)C.d_A32, (const float*)C.d_theta, (const float*)C.d_cheb,
                                   (const uint8_t*)A.d_smooth[l + 1], w.st, (int)SM_PRE2, (const float*)C.d_b,
                                   (const float*)C.d_t, (const float*)nullptr, C.d_r);
            });
        if (!A.tail_list[l + 1].empty()) {
            const AmgLevelDev& T = A.dev[l + 1];
            constexpr int v = NR == 1 ? 0 : NR == 2 ? 1 : NR == 3 ? 2 : NR == 4 ? 3 : 4;
            static_assert(TAIL_NR[v] == NR, "right-hand-side count without tail lists");
            auto launch = [&](auto LVc) {
                constexpr bool LV = decltype(LVc)::value;
                const int cnt = A.tail_n[v][LV ? 0 : 1][l + 1];
                if (cnt == 0) return;
                hipLaunchKernelGGL((k_mg_tail<NR, LV>), dim3((unsigned)cnt), dim3(TailShape<NR>::THREADS), 0, s,
                                   A.d_tail_split[v][LV ? 0 : 1][l + 1], level_slices(pl, A, l + 1),
                                   (const float*)T.d_A32, (const float*)T.d_theta, (const float*)T.d_cheb,
                                   A.d_node_off[l + 1], T.d_pt_ptr, T.d_pt_row, (const _Float16*)T.d_PT16, T.d_p_ptr,
                                   T.d_p_col, (const _Float16*)T.d_P16, A.d_term, (const float*)A.d_inv, w.st,
                                   (const float*)T.d_b, T.d_t, T.d_r, T.d_x);
            };
            launch(std::true_type{});
            launch(std::false_type{});
        }
    }
    for (int l = 1; l < nl; ++l) {
        if (A.direct_list[l].empty()) continue;
        const AmgLevelDev& D = A.dev[l];
        hipLaunchKernelGGL((k_mg_direct<NR>), dim3((unsigned)A.direct_list[l].size()), dim3(256), 0, s,
                           A.d_direct_list[l], A.d_term, (const float*)A.d_inv, w.st, (const float*)D.d_b, D.d_x);
    }
    for (int l = nl - 2; l >= 0; --l) {
        const AmgLevelDev& D = A.dev[l];
        const AmgLevelDev& C = A.dev[l + 1];
        if (C.n == 0) continue;
        if (l == 0) {
            hipLaunchKernelGGL((k_mg_prolong<VT, float, NR>), dim3(nblk(6 * (int64_t)D.n, 256)), dim3(256), 0, s, D.n,
                               D.d_node_sys, (const uint8_t*)A.d_smooth[0], D.d_p_ptr, D.d_p_col, (const _Float16*)D.d_P16,
                               (const float*)D.d_theta, (const float*)nullptr, w.st, (const VT*)bf.r[k & 1],
                               (const float*)nullptr, (const float*)C.d_x, bf.t0);
            wps_dispatch<VT, NR>(A.wps0, [&](auto W) {
                constexpr int WP = decltype(W)::value;
                const Slices s0 = level_slices(pl, A, 0);
                hipLaunchKernelGGL((k_pcg_post<VT, NR, WP>), slice_grid(s0), dim3(64 * WP), 0, s, s0, w.P, w.st, k,
                                   (const _Float16*)pl->d_Ke16, (const float*)D.d_theta,
                                   (const VT*)bf.r[k & 1], (const VT*)bf.r[(k - 1) & 1], (const float*)bf.t0, bf.z);
            });
        } else if (A.tail_list[l].size() < (size_t)std::count(A.host[l].cont.begin(), A.host[l].cont.end(), 1)) {
            // z1 = c (b + t) + P x_c into r, r0 = b - A z1 into t, x = z1 + c (2 r0 - A r0 / theta)
            hipLaunchKernelGGL((k_mg_prolong<float, float, NR>), dim3(nblk(6 * (int64_t)D.n, 256)), dim3(256), 0, s, D.n,
                               D.d_node_sys, (const uint8_t*)A.d_smooth[l], D.d_p_ptr, D.d_p_col,
                               (const _Float16*)D.d_P16, (const float*)D.d_theta, (const float*)D.d_cheb, w.st,
                               (const float*)D.d_b, (const float*)D.d_t, (const float*)C.d_x, D.d_r);
            wps_dispatch<float, NR>(4, [&](auto W) {
                constexpr int WP = decltype(W)::value;
                const Slices fs = level_slices(pl, A, l);
                hipLaunchKernelGGL((k_mg_smooth<NR, WP>), slice_grid(fs), dim3(64 * WP), 0, s, fs,
                                   (const float*)D.d_A32, (const float*)D.d_theta, (const float*)D.d_cheb,
                                   (const uint8_t*)A.d_smooth[l], w.st, (int)SM_POST1, (const float*)D.d_b,
                                   (const float*)D.d_r, (const float*)nullptr, D.d_t);
                hipLaunchKernelGGL((k_mg_smooth<NR, WP>), slice_grid(fs), dim3(64 * WP), 0, s, fs,
                                   (const float*)D.d_A32, (const float*)D.d_theta, (const float*)D.d_cheb,
                                   (const uint8_t*)A.d_smooth[l], w.st, (int)SM_POST2, (const float*)nullptr,
                                   (const float*)D.d_t, (const float*)D.d_r, D.d_x);
            });
        }
    }
}

// one ordinary PCG iteration k; `timed` (eager launches only) brackets the
// direction and CG-product kernel (the iteration's largest) with an event pair
// appended to `timed`, settled by the caller (jsso_timing_enable)
template <typename MT, typename VT, int NR>
static void amg_step(Plan* pl, AmgState& A, const Workspace& w, const AmgBufs<VT>& bf, const MT* A0, int k,
                     hipStream_t s, std::vector<std::pair<hipEvent_t, hipEvent_t>>* timed = nullptr) {
    const Slices sl = level_slices(pl, A, 0);
    const dim3 g = slice_grid(sl);
    const _Float16* A16 = (const _Float16*)pl->d_Ke16;
    wps_dispatch<VT, NR>(A.wps0, [&](auto W) {
        constexpr int WP = decltype(W)::value;
        std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
        if (timed) ev = timing_pair();
        if (ev.first) {
            // the dispatch itself stamps the events (kernel start and end, as a
            // profiler sees them), not the stream around it
            hipExtLaunchKernelGGL((k_pcg_dir<MT, VT, NR, WP>), g, dim3(64 * WP), 0, s, ev.first, ev.second, 0, sl,
                                  w.P, w.st, k, A0, (const VT*)bf.z, (const VT*)bf.p[(k - 1) & 1], bf.p[k & 1],
                                  bf.q);
            timed->push_back(ev);
        } else {
            hipLaunchKernelGGL((k_pcg_dir<MT, VT, NR, WP>), g, dim3(64 * WP), 0, s, sl, w.P, w.st, k, A0,
                               (const VT*)bf.z, (const VT*)bf.p[(k - 1) & 1], bf.p[k & 1], bf.q);
        }
        hipLaunchKernelGGL((k_pcg_pre<VT, NR, false, WP>), g, dim3(64 * WP), 0, s, sl, w.P, w.st, k, A16,
                           (const float*)A.dev[0].d_theta, (const VT*)bf.r[(k - 1) & 1], bf.r[k & 1],
                           (const VT*)bf.q, (const VT*)bf.p[k & 1], bf.xlo, bf.t0);
    });
    amg_vcycle_rest<MT, VT, NR>(pl, A, w, bf, A0, k, s);
}

// The iteration k that closes a chunk of `steps` iterations: direction and
// product, update folded into the FP64 solution, FP64 residual replacement,
// convergence test, best iterate, and the V-cycle on the new residual.
template <typename MT, typename VT, int NR>
static void amg_replace(Plan* pl, AmgState& A, const Workspace& w, const AmgBufs<VT>& bf, const MT* A0, int k,
                        int steps, double tol, hipStream_t s) {
    const Slices sl = level_slices(pl, A, 0);
    const dim3 g = slice_grid(sl);
    const _Float16* A16 = (const _Float16*)pl->d_Ke16;
    wps_dispatch<VT, NR>(A.wps0, [&](auto W) {
        constexpr int WP = decltype(W)::value;
        hipLaunchKernelGGL((k_pcg_dir<MT, VT, NR, WP>), g, dim3(64 * WP), 0, s, sl, w.P, w.st, k, A0,
                           (const VT*)bf.z, (const VT*)bf.p[(k - 1) & 1], bf.p[k & 1], bf.q);
    });
    hipLaunchKernelGGL((k_pcg_fold<VT, NR>), g, dim3(SOLVER_WG), 0, s, sl, w.P, w.st, k, (const VT*)bf.p[k & 1],
                       bf.xlo, w.x64);
    constexpr int RW = WpsFits<double, NR, 4>::value ? 4 : 1;
    hipLaunchKernelGGL((k_replace<VT, NR, RW>), g, dim3(64 * RW), 0, s, sl, w.P, w.st, (const double*)pl->d_Ke64,
                       (const double*)w.b64, (const double*)w.x64, bf.r[k & 1]);
    hipLaunchKernelGGL((k_decide<NR>), dim3(pl->n_sys), dim3(64), 0, s, sl, w.P, w.st, pl->n_sys, 1, tol, steps,
                       AMG_STALL_WINDOW, (const double*)A.dev[0].d_rho, AMG_FLOOR_FACTOR, AMG_FLOOR_WINDOW);
    hipLaunchKernelGGL((k_save<NR>), g, dim3(SOLVER_WG), 0, s, sl, (const int32_t*)w.st.save, (const double*)w.x64,
                       w.xbest);
    wps_dispatch<VT, NR>(A.wps0, [&](auto W) {
        constexpr int WP = decltype(W)::value;
        hipLaunchKernelGGL((k_pcg_pre<VT, NR, true, WP>), g, dim3(64 * WP), 0, s, sl, w.P, w.st, k, A16,
                           (const float*)A.dev[0].d_theta, (const VT*)bf.r[(k - 1) & 1], bf.r[k & 1],
                           (const VT*)bf.q, (const VT*)bf.p[k & 1], bf.xlo, bf.t0);
    });
    amg_vcycle_rest<MT, VT, NR>(pl, A, w, bf, A0, k, s);
}


// Symbolic hierarchy on the host (no device work), then its device copy.
static int amg_ensure_host(Plan* pl) {
    if (pl->amg) return JSSO_OK;
    AmgState* A = new AmgState();
    int rc = amg_build_host(pl, *A);
    if (rc) {
        delete A;
        return rc;
    }
    pl->amg = A;
    return JSSO_OK;
}

static int amg_prepare(Plan* pl, hipStream_t s) {
    int rc = amg_ensure_host(pl);
    if (rc) return rc;
    AmgState& A = *pl->amg;
    if (A.dev.empty()) {
        if ((rc = amg_build_device(pl, A))) {
            A.dev.clear();
            return rc;
        }
        ++A.gen;
    }
    if (!pl->amg_valid) {
        JSSO_REQUIRE(pl->d_crds, JSSO_EINVAL, "AMG needs the node coordinates of an assembly");
        std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
        timing_begin_slot(s, ev, JSSO_TIMING_SETUP);
        if (pl->amg_reuse == 1 && A.built) {
            if ((rc = amg_setup_light(pl, A, s))) return rc;
            ++pl->amg_light;
        } else {
            if ((rc = amg_setup_numeric(pl, A, s))) return rc;
            ++pl->amg_full;
            if (pl->amg_reuse == 2) pl->amg_reuse = 1;
        }
        timing_end(s, ev);
        pl->amg_valid = true;
    }
    return JSSO_OK;
}

// Autotuning of the level-0 kernels' waves per slice (1, 2 or 4 wavefronts share
// a 64-row slice's ELL slots): the direction and CG-product kernel, the
// iteration's longest, is timed for every variant that fits the register budget
// on the live plan, right after the start-up kernels have set the slices'
// activity (its outputs are rewritten by the first iteration, so the timing
// leaves no trace), and the fastest variant serves every level-0 kernel of this
// (precision, load cases) from then on.
template <typename MT, typename VT, int NR>
static int tune_amg_wps(Plan* pl, AmgState& A, const Workspace& w, const AmgBufs<VT>& bf, const MT* A0,
                        hipStream_t s, int32_t& out) {
    const Slices sl = level_slices(pl, A, 0);
    const dim3 g = slice_grid(sl);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    JSSO_HIP(hipEventCreate(&e0));
    JSSO_HIP(hipEventCreate(&e1));
    float best = 0.f;
    int pick = 0;
    for (int cand : {4, 2, 1}) {
        int used = 0;
        wps_dispatch<VT, NR>(cand, [&](auto W) { used = decltype(W)::value; });
        if (used != cand) continue;   // does not fit the register budget
        float ms = 0.f;
        for (int rep = 0; rep < 4; ++rep) {
            wps_dispatch<VT, NR>(cand, [&](auto W) {
                constexpr int WP = decltype(W)::value;
                if (rep == 1) (void)hipEventRecord(e0, s);
                hipLaunchKernelGGL((k_pcg_dir<MT, VT, NR, WP>), g, dim3(64 * WP), 0, s, sl, w.P, w.st, 1, A0,
                                   (const VT*)bf.z, (const VT*)bf.p[0], bf.p[1], bf.q);
            });
        }
        (void)hipEventRecord(e1, s);
        JSSO_HIP(hipEventSynchronize(e1));
        JSSO_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (pick == 0 || ms < best) {
            best = ms;
            pick = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    JSSO_HIP(hipGetLastError());
    out = pick > 0 ? pick : 1;
    return JSSO_OK;
}

template <typename MT, typename VT, int NR>
static int run_amg(Plan* pl, const double* f, double* u, double tol, int maxit, int replace_every, jsso_info* info,
                   hipStream_t s) {
    Workspace w;
    int rc = get_workspace(pl, NR, w);
    if (rc) return rc;
    AmgState& A = *pl->amg;
    if ((rc = amg_vectors(pl, A, NR))) return rc;
    const Slices sl = level_slices(pl, A, 0);
    const int nsys = pl->n_sys * NR;
    if ((rc = ensure_flag_buffers(pl, nsys))) return rc;
    const MT* A0 = nullptr;
    if constexpr (sizeof(MT) == 4) A0 = (const MT*)pl->d_Ke32; else A0 = (const MT*)pl->d_Ke64;
    AmgBufs<VT> bf;
    bf.r[0] = (VT*)w.r[0];
    bf.r[1] = (VT*)w.r[1];
    bf.p[0] = (VT*)w.p;
    bf.p[1] = (VT*)A.d_p1;
    bf.q = (VT*)w.q[0];
    bf.z = (VT*)A.d_z;
    bf.t0 = (float*)A.d_t0;
    bf.xlo = (VT*)w.xlo;
    const dim3 gridS = slice_grid(sl);
    // start: b, x = 0, r_0 = b, p = 0; both partial buffers cleared (so the
    // first direction has beta = 0); z_0 = M r_0
    JSSO_HIP(hipMemsetAsync(w.P.cg[0], 0, 2 * 4 * (size_t)pl->n_wg * NR * sizeof(double), s));
    hipLaunchKernelGGL((k_init<VT, NR>), gridS, dim3(SOLVER_WG), 0, s, sl, w.P, pl->d_fixed, pl->d_Linv, f, w.b64,
                       w.x64, w.xbest, bf.xlo, bf.r[0], bf.p[0]);
    hipLaunchKernelGGL((k_decide<NR>), dim3(pl->n_sys), dim3(64), 0, s, sl, w.P, w.st, pl->n_sys, 0, tol, 0, 0,
                       (const double*)nullptr, 0.0, 0);
    int32_t& wsel = pl->amg_wps[prec_slot<MT, VT>()][NR];
    if (wsel <= 0 && (rc = tune_amg_wps<MT, VT, NR>(pl, A, w, bf, A0, s, wsel))) return rc;
    A.wps0 = wsel;
    wps_dispatch<VT, NR>(A.wps0, [&](auto W) {
        constexpr int WP = decltype(W)::value;
        hipLaunchKernelGGL((k_pcg_pre<VT, NR, true, WP>), gridS, dim3(64 * WP), 0, s, sl, w.P, w.st, 0,
                           (const _Float16*)pl->d_Ke16,
                           (const float*)A.dev[0].d_theta, (const VT*)bf.r[1], bf.r[0], (const VT*)bf.q,
                           (const VT*)bf.p[0], bf.xlo, bf.t0);
    });
    amg_vcycle_rest<MT, VT, NR>(pl, A, w, bf, A0, 0, s);
    JSSO_HIP(hipGetLastError());

    const int E = replace_every;
    const void* key[CgGraph::NKEY] = {A0,        w.r[0],   w.r[1],   w.q[0],         w.p,         A.d_p1,
                                      A.d_z,     A.d_t0,   w.xlo,    w.P.cg[0],      w.st.slice_act, w.b64,
                                      w.x64,     w.xbest,  w.st.active, (const void*)(uintptr_t)A.gen};
    CgGraph(&gc)[2] = A.graph[prec_slot<MT, VT>()][NR];
    // timing samples: the k_pcg_dir launches of sampled chunks, kept per chunk and
    // counted only for chunks that started with every system still running (all
    // slices active, so the launch processes the whole plan)
    std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> samples;
    int chunk_no = 0;
    // `may_sample`: every system was still running at the last chunk end read
    // (only such chunks can count as samples, so no other one is launched eagerly
    // for timing)
    auto enqueue_chunk = [&](int it0, int steps, bool may_sample) -> int {
        hipGraphExec_t ga = nullptr, gb = nullptr;
        // a chunk sampled for timing runs eagerly with its k_pcg_dir launches bracketed
        const bool sampled = may_sample && timing_sample_chunk();
        samples.emplace_back();
        auto* tv = sampled ? &samples[chunk_no] : nullptr;
        ++chunk_no;
        // the full-length chunk, or one cut short (even lengths; k only matters
        // through its parity, so a graph serves every even start)
        CgGraph* gp = nullptr;
        if (steps == E)
            gp = gc;
        else if (steps < E && steps >= 2 && steps % 2 == 0 && steps / 2 - 1 < AMG_CUT_GRAPHS)
            gp = A.graph_cut[prec_slot<MT, VT>()][NR][steps / 2 - 1];
        if (!sampled && gp && steps % 2 == 0 && it0 % 2 == 0 && pl->graphs_ok && graphs_enabled()) {
            ga = capture_graph(pl, gp[0], key, A.wps0, steps, tol, [&](hipStream_t cs) {
                for (int it = 1; it < steps; ++it) amg_step<MT, VT, NR>(pl, A, w, bf, A0, it, cs);
            });
            if (ga)
                gb = capture_graph(pl, gp[1], key, A.wps0, steps, tol, [&](hipStream_t cs) {
                    amg_replace<MT, VT, NR>(pl, A, w, bf, A0, steps, steps, tol, cs);
                });
        }
        if (ga && gb) {
            JSSO_HIP(hipGraphLaunch(ga, s));
            JSSO_HIP(hipGraphLaunch(gb, s));
        } else {
            for (int it = it0 + 1; it < it0 + steps; ++it) amg_step<MT, VT, NR>(pl, A, w, bf, A0, it, s, tv);
            amg_replace<MT, VT, NR>(pl, A, w, bf, A0, it0 + steps, steps, tol, s);
        }
        JSSO_HIP(hipGetLastError());
        return JSSO_OK;
    };

    // Convergence flags (and the FP64 relative residuals) are read one chunk
    // behind the GPU, as in the block-Jacobi solver.  The residual is only tested
    // at a chunk's end; the slowest running system's residuals at the last two
    // known chunk ends give its convergence factor per iteration, and so the
    // iteration at which it is predicted to reach tol.  A chunk that would
    // straddle that iteration is cut to end just after it (an even length: its
    // graphs are captured once per length); a wrong prediction costs one more
    // replacement.  A chunk in flight that is predicted to end the solve is read
    // before the next one is enqueued: the GPU then waits for the host's
    // turnaround instead of running a whole chunk after convergence.  After two
    // such predictions that failed, the solve keeps reading one chunk behind.
    std::vector<int> chunk_end;
    std::vector<char> ends_solve;                // per chunk: read early as predicted to end the solve
    std::vector<std::pair<int, double>> known;   // (iteration, slowest running system's residual)
    std::vector<char> all_running_after;         // per chunk read: every system still running after it
    int it = 0, it_done = 0, chunks_done = 0, c = 0, next_read = 0, misses = 0;
    // flags of chunks next_read..upto in order: 1 while a system is still running
    // (0 once none is, with it_done / chunks_done set), or an error code
    auto read_upto = [&](int upto) -> int {
        for (; next_read <= upto; ++next_read) {
            const int chk = next_read;
            JSSO_HIP(hipEventSynchronize(pl->flag_ev[chk & 1]));
            timing_collect();
            const int32_t* fl = pl->h_flags + (size_t)(chk & 1) * nsys;
            const double* rl = pl->h_rel + (size_t)(chk & 1) * nsys;
            int running = 0;
            double worst = 0.0;
            for (int k = 0; k < nsys; ++k)
                if (fl[k] != 0) {
                    ++running;
                    worst = std::max(worst, rl[k]);
                }
            all_running_after.push_back(running == nsys);
            if (running == 0) {
                it_done = chunk_end[chk];
                chunks_done = chk + 1;
                return 0;
            }
            if (ends_solve[chk]) ++misses;
            known.push_back({chunk_end[chk], worst});
        }
        return 1;
    };
    // predicted iteration at which the slowest running system reaches tol (0: no prediction)
    auto predict = [&]() -> double {
        const size_t nk = known.size();
        if (nk < 2 || !(known[nk - 1].first > known[nk - 2].first && known[nk - 1].second > tol &&
                        known[nk - 1].second < known[nk - 2].second))
            return 0.0;
        const double per_it = std::log(known[nk - 1].second / known[nk - 2].second) /
                              (known[nk - 1].first - known[nk - 2].first);
        return known[nk - 1].first + std::log(tol / known[nk - 1].second) / per_it;
    };
    for (;;) {
        const bool more = it < maxit;
        if (more && next_read == c - 1 && misses < 2) {
            const double kstar = predict();
            if (kstar > 0.0 && kstar <= chunk_end[c - 1]) {
                ends_solve[c - 1] = 1;
                const int r = read_upto(c - 1);
                if (r < 0) return r;
                if (r == 0) break;
            }
        }
        if (more) {
            int steps = std::min(E, maxit - it);
            const double kstar = predict();
            if (kstar > it && kstar < it + steps - 1) {
                int cut = (int)std::ceil(kstar - it) + 1;
                cut += cut & 1;
                steps = std::max(2, std::min(steps, cut));
            }
            // a chunk can only count as a timing sample if every system was still
            // running at the last chunk end read
            const bool may_sample = all_running_after.empty() || all_running_after.back();
            if ((rc = enqueue_chunk(it, steps, may_sample))) return rc;
            it += steps;
            chunk_end.push_back(it);
            ends_solve.push_back(0);
            JSSO_HIP(hipMemcpyAsync(pl->h_flags + (size_t)(c & 1) * nsys, w.st.active, nsys * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
            JSSO_HIP(hipMemcpyAsync(pl->h_rel + (size_t)(c & 1) * nsys, w.st.relres, nsys * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
            JSSO_HIP(hipEventRecord(pl->flag_ev[c & 1], s));
            ++c;
        }
        // one chunk behind (the host buffers hold two chunks), all of them at the end
        const int upto = more ? c - 2 : c - 1;
        if (upto >= 0) {
            const int r = read_upto(upto);
            if (r < 0) return r;
            if (r == 0) break;
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
    for (size_t c = 0; c < samples.size(); ++c) {
        const bool full = c == 0 || (c - 1 < all_running_after.size() && all_running_after[c - 1]);
        for (auto& e : samples[c]) timing_settle(e, JSSO_TIMING_SOLVE, full);
    }
    int conv = 0, floor = 0;
    double mx = 0.0;
    for (int k = 0; k < nsys; ++k) {
        conv += (status[k] == SYS_CONVERGED || status[k] == SYS_FLOOR) ? 1 : 0;
        floor += status[k] == SYS_FLOOR ? 1 : 0;
        mx = std::max(mx, rel[k]);
    }
    if (info) {
        info->iterations = it_done;
        info->replacements = chunks_done;
        info->converged = conv;
        info->n_systems = nsys;
        info->rel_residual = mx;
        info->waves_per_slice = A.wps0;
        info->levels = A.nl;
        info->dir_launches = it;
        info->at_floor = floor;
        int full = 0;
        while (full < (int)all_running_after.size() && all_running_after[full]) ++full;
        info->full_iterations = full < (int)chunk_end.size() ? chunk_end[full] : it_done;
    }
    return conv == nsys ? JSSO_OK : JSSO_NOT_CONVERGED;
}

// Iterations between FP64 residual replacements (and convergence tests) of the
// AMG solve; JSSO_AMG_REPLACE_EVERY overrides (an even count, so the chunk
// graphs are reused).  8: the solve needs ~100 iterations, so the test that
// ends it comes soon after convergence, and the FP32 recurrences stay close
// to the FP64 residual.
static int amg_replace_interval() {
    const char* e = getenv("JSSO_AMG_REPLACE_EVERY");
    const int v = e ? atoi(e) : 0;
    return v > 0 ? v : 8;
}

template <typename MT, typename VT>
static int amg_dispatch(Plan* pl, const double* f, double* u, int nrhs, double tol, int maxit, jsso_info* info,
                        hipStream_t s) {
    const int E = amg_replace_interval();
    switch (nrhs) {
        case 1: return run_amg<MT, VT, 1>(pl, f, u, tol, maxit, E, info, s);
        case 2: return run_amg<MT, VT, 2>(pl, f, u, tol, maxit, E, info, s);
        case 3: return run_amg<MT, VT, 3>(pl, f, u, tol, maxit, E, info, s);
        case 4: return run_amg<MT, VT, 4>(pl, f, u, tol, maxit, E, info, s);
        case 8: return run_amg<MT, VT, 8>(pl, f, u, tol, maxit, E, info, s);
        default:
            set_error("nrhs=%d unsupported (1, 2, 3, 4 or 8)", nrhs);
            return JSSO_EINVAL;
    }
}

int amg_solve(Plan* pl, const double* f, double* u, int nrhs, int slot, double tol, int maxit, jsso_info* info,
              hipStream_t s) {
    int rc = amg_prepare(pl, s);
    if (rc) return rc;
    if (slot == PS_F32) return amg_dispatch<float, float>(pl, f, u, nrhs, tol, maxit, info, s);
    if (slot == PS_F32_V64) return amg_dispatch<float, double>(pl, f, u, nrhs, tol, maxit, info, s);
    return amg_dispatch<double, double>(pl, f, u, nrhs, tol, maxit, info, s);
}

int amg_levels(Plan* pl, int32_t* h_nodes, int32_t cap, int32_t* h_levels) {
    int rc = amg_ensure_host(pl);
    if (rc) return rc;
    const AmgState& A = *pl->amg;
    if (h_levels) *h_levels = A.nl;
    for (int l = 0; l < A.nl && l < cap; ++l) h_nodes[l] = A.host[l].n;
    return JSSO_OK;
}

void amg_destroy(Plan* pl) {
    AmgState* A = pl->amg;
    if (!A) return;
    for (auto& row : A->graph)
        for (auto& pair : row)
            for (auto& g : pair)
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& row : A->graph_cut)
        for (auto& lens : row)
            for (auto& pair : lens)
                for (auto& g : pair)
                    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void* a : A->allocs) (void)hipFree(a);
    for (void* v : {A->d_z, A->d_t0, A->d_p1})
        if (v) (void)hipFree(v);
    for (auto& D : A->dev)
        for (float* v : {D.d_b, D.d_t, D.d_r, D.d_x})
            if (v) (void)hipFree(v);
    delete A;
    pl->amg = nullptr;
}

}  // namespace jsso
