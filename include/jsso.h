/*
 * jsso.h -- C ABI of libjsso_hip.so, the MI355X (gfx950) implementation of the
 * JaxSSO finite-element forward + adjoint path.
 *
 * The reference (GaoyuanWu/JaxSSO, pure Python/JAX) has no native boundary; the
 * entry points below replace the device work behind its Python API:
 *
 *   jsso_element_ke_beam   <- BeamCol.element_K_beamcol        JaxSSO/element.py:130-139
 *   jsso_element_ke_quad   <- Quad.element_K_quad              JaxSSO/element.py:1073-1084
 *   jsso_plan_create       <- element_K_*_indices + sort_indices (symbolic part of
 *                             K_beamcol/K_quad/K_aug)          JaxSSO/element.py:141-149,1097-1106;
 *                                                              JaxSSO/assemblemodel.py:111-163
 *   jsso_assemble          <- K_func / K_aug_func (numeric: vmap(element_K_*) +
 *                             bcoo_sum_duplicates)             JaxSSO/assemblemodel.py:196-229
 *   jsso_solve             <- sci_sparse_solve / jax_sparse_solve (forward)
 *                                                              JaxSSO/solver.py:102-125,176-218
 *   jsso_adjoint_grad      <- the custom VJP of the solve plus reverse mode through
 *                             vmap(element_K_*)               JaxSSO/solver.py:221-250;
 *                                                              JaxSSO/SSO_model.py:329-339
 *   jsso_spmv              <- K_aug @ u inside the VJP residual JaxSSO/solver.py:239-241
 *
 * Conventions
 *  - Every array argument named d_* is a DEVICE pointer (allocated by the caller,
 *    e.g. torch); h_* are host pointers.  FP64 unless stated.  Indices int32.
 *  - Per-node DOF order [ux, uy, uz, rx, ry, rz]; global DOF = 6*node + k
 *    (JaxSSO/model.py:183-214).
 *  - A plan packs n_sys independent structures: structure s owns nodes
 *    [node_off[s], node_off[s+1]); element connectivities hold global node ids and
 *    every element must lie inside one structure.
 *  - Vectors with several right-hand sides are laid out [dof][rhs].
 *  - Return value: 0 ok; 1 solver not converged (info holds details); negative on
 *    error (jsso_last_error gives the message).  No exception crosses the ABI.
 *  - Work is ordered on the caller's stream (hipStream_t passed as void*, NULL =
 *    default stream).  jsso_solve synchronises that stream to test convergence.
 *  - A plan owns its pattern arrays and its device workspace; calls that use the
 *    same plan must not overlap.
 */
#ifndef JSSO_H
#define JSSO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jsso_plan jsso_plan;

enum {
    JSSO_OK = 0,
    JSSO_NOT_CONVERGED = 1,
    JSSO_EINVAL = -1,
    JSSO_EHIP = -2,
    JSSO_ENOMEM = -3,
    JSSO_EUNSUPPORTED = -4,
    JSSO_ESINGULAR = -5
};

/* solver precision */
enum {
    JSSO_PREC_FP64 = 0,  /* FP64 matrix and vectors, batched, reliable FP64 residual */
    JSSO_PREC_MIXED = 1, /* FP32 matrix + FP32 CG recurrences, FP64 reliable updates  */
    JSSO_PREC_PLAIN = 2  /* baseline: FP64 block-Jacobi PCG, one structure per launch */
};

/* preconditioner of the batched precisions (jsso_plan_set_preconditioner) */
enum {
    JSSO_PC_BLOCK_JACOBI = 0, /* split block-Jacobi scaling                             */
    JSSO_PC_AMG = 1           /* smoothed-aggregation AMG V-cycle on the scaled system
                                 (the default)                                          */
};

typedef struct {
    int32_t iterations;   /* CG iterations performed (max over systems)            */
    int32_t replacements; /* FP64 true-residual replacements performed             */
    int32_t converged;    /* systems (structure x rhs) that met the tolerance       */
    int32_t n_systems;    /* structures x rhs                                       */
    double rel_residual;  /* max over systems of ||L^-1 (f - K u)|| / ||L^-1 f||    */
    int32_t waves_per_slice; /* batched CG kernel variant used (1, 2 or 4 wavefronts
                                per 64-block-row slice); 0 on the plain path      */
    int32_t levels;          /* AMG levels (JSSO_PC_AMG), 0 otherwise                 */
} jsso_info;

/* plan size indices for jsso_plan_sizes */
enum {
    JSSO_SZ_NODES = 0, JSSO_SZ_NNZB = 1, JSSO_SZ_BEAMS = 2, JSSO_SZ_QUADS = 3,
    JSSO_SZ_SYSTEMS = 4, JSSO_SZ_CONTRIB = 5, JSSO_SZ_COUNT = 6
};

int jsso_version(void);
int jsso_last_error(char* buf, size_t n);

/* Symbolic plan: 6x6 block-CSR pattern of K (the reference's scalar COO pattern,
 * bit-exact), the block -> element contribution map, node -> element adjacency
 * and constrained DOFs.  h_cons_dofs: global DOF ids with zero prescribed
 * displacement (duplicates rejected, as they make the reference's K_aug singular). */
int jsso_plan_create(int32_t n_sys, const int32_t* h_node_off,
                     int32_t n_beam, const int32_t* h_cnct_beam,
                     int32_t n_quad, const int32_t* h_cnct_quad,
                     int32_t n_cons, const int32_t* h_cons_dofs,
                     jsso_plan** out);
int jsso_plan_destroy(jsso_plan* plan);
int jsso_plan_sizes(const jsso_plan* plan, int64_t* h_sizes /* JSSO_SZ_COUNT */);
/* host copy of the block pattern: h_rowptr (n_nodes+1), h_colind (nnzb) */
int jsso_plan_pattern(const jsso_plan* plan, int32_t* h_rowptr, int32_t* h_colind);

/* Per-element global stiffness, row-major.  d_crds: beams n*6 [x1,y1,z1,x2,y2,z2],
 * quads n*12; d_props: beams n*6 [E,G,Iy,Iz,J,A], quads n*5 [t,E,nu,kx_mod,ky_mod].
 * d_ke: n*144 / n*576.  d_flags (beams, may be NULL): 1 where the member is parallel
 * to global Y (the reference's non-orthogonal Cxz==0 branch, element.py:92-94). */
int jsso_element_ke_beam(int32_t n, const double* d_crds, const double* d_props,
                         double* d_ke, int32_t* d_flags, void* stream);
int jsso_element_ke_quad(int32_t n, const double* d_crds, const double* d_props,
                         double* d_ke, void* stream);

/* Numeric assembly of K from nodal coordinates d_crds (n_nodes*3) and element
 * properties (d_prop_beam n_beam*6, d_prop_quad n_quad*5; NULL when absent).
 * Builds the plan's FP64 K, and the constrained, block-Jacobi-scaled FP64 and
 * FP32 copies the solver uses.  d_K_out (36*nnzb, may be NULL) receives K
 * (unconstrained, the reference's K) in block-CSR order, each block row-major. */
int jsso_assemble(jsso_plan* plan, const double* d_crds, const double* d_prop_beam,
                  const double* d_prop_quad, double* d_K_out, void* stream);

/* Solve K u = f with u = 0 at constrained DOFs, for nrhs load cases per structure.
 * d_f, d_u: (6*n_nodes) x nrhs.  tol: on the block-Jacobi-scaled residual. */
int jsso_solve(jsso_plan* plan, const double* d_f, double* d_u, int32_t nrhs,
               int32_t precision, double tol, int32_t maxit, jsso_info* h_info,
               void* stream);

/* Adjoint gradient of the objective through the solve:
 *     g = - sum_e lam_e^T (dK_e/dp) u_e
 * for p = nodal coordinates and every element property.  d_lam == NULL means
 * lam = lam_scale * u (compliance C = 1/2 f.u has lam = u/2, no second solve).
 * Outputs (each may be NULL): d_dcrds (n_nodes*3) x nrhs,
 * d_dprop_beam (n_beam*6) x nrhs, d_dprop_quad (n_quad*5) x nrhs, layout [..][rhs].
 * Members parallel to global Y give NaN coordinate derivatives (as the reference). */
int jsso_adjoint_grad(jsso_plan* plan, const double* d_crds, const double* d_prop_beam,
                      const double* d_prop_quad, const double* d_u, const double* d_lam,
                      double lam_scale, int32_t nrhs, double* d_dcrds,
                      double* d_dprop_beam, double* d_dprop_quad, void* stream);

/* y = K x with the assembled, unconstrained K (nrhs = 1). */
int jsso_spmv(jsso_plan* plan, const double* d_x, double* d_y, void* stream);

/* Compliance-type reduction per structure and load case:
 *     d_C[s][h] = scale * sum over the DOFs of structure s of f[dof][h] u[dof][h]
 * (d_C: n_sys x nrhs; scale 1/2 gives the strain-energy objective C = 1/2 f.u of
 * JaxSSO/SSO_model.py:275-306, per packed structure).  Deterministic: per-slice
 * partial sums added in a fixed order, no atomics. */
int jsso_compliance(jsso_plan* plan, const double* d_f, const double* d_u, int32_t nrhs,
                    double scale, double* d_C, void* stream);

/* Tuning of the batched CG kernel (no counterpart in the reference: the
 * reference hands the solve to SuperLU).  The first batched solve on a plan for a
 * (precision, nrhs) pair times the 1-, 2- and 4-wavefront-per-slice variants on
 * the live problem and keeps the fastest for the plan's later solves.
 * set_wps >= 0 overrides the stored choice (0 = tune again on the next solve,
 * 1/2/4 = force); h_wps, if not NULL, receives the stored choice (0 = untuned). */
int jsso_plan_cg_tuning(jsso_plan* plan, int32_t precision, int32_t nrhs, int32_t set_wps,
                        int32_t* h_wps);

/* Preconditioner of the batched solves (JSSO_PREC_FP64 / JSSO_PREC_MIXED); no
 * counterpart in the reference, which factorises with SuperLU (JaxSSO/solver.py:
 * 176-218).  JSSO_PC_AMG builds a smoothed-aggregation hierarchy on the plan's
 * pattern (host, once) and its numeric operators after each jsso_assemble (device,
 * at the next solve); the plain path always uses block-Jacobi. */
int jsso_plan_set_preconditioner(jsso_plan* plan, int32_t pc);
/* Symbolic AMG hierarchy of a plan (host only, no device work): *h_levels receives
 * the number of levels, h_nodes[l] (up to cap entries) the block rows of level l. */
int jsso_plan_amg_levels(jsso_plan* plan, int32_t* h_nodes, int32_t cap, int32_t* h_levels);

/* Measurement hook (no counterpart in the reference): sampled HIP-event timing
 * of the batched solver's dominant kernel: with block-Jacobi the fused CG
 * iteration that applies the scaled stiffness (a captured chunk of iterations
 * counts as that many launches); with AMG the direction and CG-product kernel
 * (k_pcg_dir), sampled only where it is launched eagerly (JSSO_CG_GRAPH=0).  every > 0
 * brackets one such launch in `every` with events on the solver's stream; 0 disables.  jsso_timing_read returns the mean duration
 * (ms) and the number of sampled launches, optionally resetting the counters. */
int jsso_timing_enable(int32_t every);
int jsso_timing_read(double* h_avg_ms, int64_t* h_samples, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* JSSO_H */
