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
    JSSO_PREC_PLAIN = 2, /* baseline: FP64 block-Jacobi PCG, one structure per launch */
    JSSO_PREC_MIXED_V64 = 3 /* FP32 matrix in the CG product, FP64 vectors and CG
                               recurrences, FP64 reliable updates                     */
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
    int32_t dir_launches;    /* AMG: launches of the direction/product kernel, incl.
                                the chunk queued ahead of the convergence test       */
    int32_t full_iterations; /* AMG: iterations run while every system was running   */
    int32_t at_floor;        /* AMG: systems (counted in `converged`) that stopped at
                                the attainable accuracy of the FP64 residual,
                                ~eps (||A|| ||x|| + ||b||) / ||b||, above tol      */
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

/* Element-level parts of the reference's element API, one element per thread,
 * row-major outputs; every output may be NULL.
 * Beams (d_crds n*6, d_props n*6 as above): d_T n*144 <- BeamCol.T (element.py:
 * 66-105; 4 diagonal copies of the direction cosines), d_klocal n*144 <-
 * BeamCol.K_local (element.py:107-128). */
int jsso_element_beam_local(int32_t n, const double* d_crds, const double* d_props,
                            double* d_T, double* d_klocal, void* stream);
/* Quads (d_crds n*12, d_props n*5 as above): d_xy n*8 local in-plane coordinates
 * [x1,y1,..,x4,y4] (node 3 the origin) and d_normals n*9 (the element normal and
 * the normals of triangles 1-2-3 and 1-3-4) <- Quad.loc_crds (element.py:488-539);
 * d_T n*576 <- Quad.T (element.py:643-696); d_klocal n*576 (k_m + k_b before the
 * rotation) <- Quad.element_K_quad_local (element.py:1086-1095). */
int jsso_element_quad_local(int32_t n, const double* d_crds, const double* d_props,
                            double* d_xy, double* d_normals, double* d_T, double* d_klocal,
                            void* stream);
/* Quads at natural coordinates (r, s), from d_xy (n*8, as jsso_element_quad_local
 * writes): d_J n*4 <- Quad.J, d_Bk n*36 <- Quad.B_kappa, d_Bg n*24 <-
 * Quad.B_gamma_MITC4, d_Bm n*24 <- Quad.B_m (element.py:698-818); from d_props
 * (n*5): d_Cb n*9, d_Cs n*4, d_Cm n*9 <- Quad.Cb / Cs / Cm (element.py:820-875). */
int jsso_element_quad_gauss(int32_t n, const double* d_xy, const double* d_props, double r,
                            double s, double* d_J, double* d_Bk, double* d_Bg, double* d_Bm,
                            double* d_Cb, double* d_Cs, double* d_Cm, void* stream);

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

/* jsso_adjoint_grad without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as jsso_adjoint_grad; the per-element gradient scratch is a fixed
 * 64 MiB device array per GPU instead of plan workspace, so (12 per beam + 17 per
 * quad) x nrhs doubles must fit in 8 Mi doubles (JSSO_EINVAL otherwise, with the
 * sizes in the message).  The plan must already be on the device (any
 * jsso_assemble puts it there).  Calls on one GPU share the array: calls that
 * overlap on one GPU race. */
int jsso_adjoint_grad_noalloc(jsso_plan* plan, const double* d_crds, const double* d_prop_beam,
                              const double* d_prop_quad, const double* d_u, const double* d_lam,
                              double lam_scale, int32_t nrhs, double* d_dcrds,
                              double* d_dprop_beam, double* d_dprop_quad, void* stream);

/* y = K x with the assembled, unconstrained K (nrhs = 1). */
int jsso_spmv(jsso_plan* plan, const double* d_x, double* d_y, void* stream);

/* y = K x with K given by its block values d_K (36*nnzb, block-CSR order of the
 * plan's pattern, unconstrained, as jsso_assemble's d_K_out), so that a K kept
 * by the caller is applied as itself whatever the plan assembled since; d_K NULL
 * uses the plan's current assembly (nrhs = 1).
 *   <- BCOO K @ x / K_aug @ u_aug             JaxSSO/assemblemodel.py:111-163,196-213 */
int jsso_spmv_k(jsso_plan* plan, const double* d_K, const double* d_x, double* d_y, void* stream);

/* Dense direct solve: the constrained K (rows/columns of constrained DOFs replaced
 * by identity, u = 0 there) factorised in FP64 on the device -- by a blocked
 * Cholesky, or, when the plan's last assembly makes K non-symmetric (members
 * parallel to global Y, element.py:92-94; quads with kx_mod != ky_mod and nu != 0,
 * whose Cm is not symmetric, element.py:852-875), by a blocked LU with partial pivoting --
 * then triangular solves.  transpose != 0 solves Kc^T u = f (the adjoint system).
 * d_K as for jsso_spmv_k (NULL = the plan's current assembly); d_f, d_u:
 * (6*n_nodes) x nrhs, nrhs <= 64.  Needs 8 N^2 bytes of device memory (N =
 * 6*n_nodes rounded up to a multiple of 64).  Returns JSSO_ESINGULAR for a
 * stiffness that is not positive definite (Cholesky) or singular (LU).
 *   <- jax_dense_solve (dense LU of K_aug)      JaxSSO/solver.py:75-93
 *      and, for non-symmetric K, SuperLU on K_aug / K_aug^T  JaxSSO/solver.py:195-209,236
 * jsso_solve routes such plans to the LU path itself up to 16 384 DOF (see
 * jsso_solve_transpose for larger ones), whatever the precision asked for
 * (info->iterations = 0). */
int jsso_dense_solve(jsso_plan* plan, const double* d_K, const double* d_f, double* d_u, int32_t nrhs,
                     int32_t transpose, void* stream);
/* K^T u = f with u = 0 at supports: the adjoint system of a non-symmetric plan
 * (same arguments and results as jsso_solve; a symmetric plan is solved as by
 * jsso_solve).
 *   <- the custom VJP's solve with K_aug^T  JaxSSO/solver.py:236
 * Non-symmetric plans (both entry points): the dense LU (jsso_dense_solve) up to
 * 16 384 DOF or when members parallel to global Y are present; above that a
 * defect correction: r = f - K u (K^T u), S du = r with S = (K + K^T)/2 (the
 * assembly writes S into the solver copies of such a plan; inner batched solve
 * at relative tolerance 1e-3 with the plan's preconditioner), u += du, until
 * ||r|| / ||f|| <= tol for every system.  h_info->iterations counts the inner
 * iterations, h_info->replacements the outer steps.  Returns JSSO_NOT_CONVERGED
 * when a step reduces the residual by less than 10% before tol is reached. */
int jsso_solve_transpose(jsso_plan* plan, const double* d_f, double* d_u, int32_t nrhs, int32_t precision,
                         double tol, int32_t maxit, jsso_info* h_info, void* stream);
/* Route of non-symmetric plans: mode 0 automatic (above), 1 always the dense LU,
 * 2 always the defect correction (JSSO_EUNSUPPORTED with members parallel to
 * global Y, whose transformation, element.py:92-94, leaves S indefinite);
 * -1 only reads it.  *h_mode (may be NULL) receives the mode in force. */
int jsso_plan_nonsymmetric_route(jsso_plan* plan, int32_t mode, int32_t* h_mode);

/* *h_symmetric = 0 when the plan's last assembly holds members parallel to global
 * Y or quads with kx_mod != ky_mod (its K is not symmetric, so an adjoint needs
 * the transposed solve), else 1.
 * Synchronises the stream. */
int jsso_plan_symmetric(jsso_plan* plan, int32_t* h_symmetric, void* stream);

/* Compliance-type reduction per structure and load case:
 *     d_C[s][h] = scale * sum over the DOFs of structure s of f[dof][h] u[dof][h]
 * (d_C: n_sys x nrhs; scale 1/2 gives the strain-energy objective C = 1/2 f.u of
 * JaxSSO/SSO_model.py:275-306, per packed structure).  Deterministic: per-slice
 * partial sums added in a fixed order, no atomics. */
int jsso_compliance(jsso_plan* plan, const double* d_f, const double* d_u, int32_t nrhs,
                    double scale, double* d_C, void* stream);

/* y = B x (+ c) for a sparse B in CSR on the device (n_rows rows, int32 row
 * pointers and columns, FP64 values; d_c may be NULL): the design-parameter maps
 * of the reference's examples, e.g. the hat filter z = B_ij @ p and, with B^T
 * passed in CSR, its gradient map sens @ B_ij
 *   <- Examples/Shells_Mannheim_Multihalle_Shape.ipynb:375-378,454-461 (the
 *      reference builds and applies B_ij as a dense numpy matrix) */
int jsso_csr_spmv(int32_t n_rows, const int32_t* d_rowptr, const int32_t* d_colind, const double* d_vals,
                  const double* d_x, const double* d_c, double* d_y, void* stream);

/* Tuning of the batched CG kernels (no counterpart in the reference: the
 * reference hands the solve to SuperLU).  The first batched solve on a plan for a
 * (precision, nrhs) pair times the 1-, 2- and 4-wavefront-per-slice variants on
 * the live problem and keeps the fastest for the plan's later solves: with the
 * block-Jacobi preconditioner the fused iteration kernel k_cgf, with AMG the
 * level-0 kernels (timed on k_pcg_dir, the direction and CG product).  The
 * choice is kept per preconditioner; this call reads or sets the one of the
 * plan's current preconditioner.  set_wps >= 0 overrides the stored choice
 * (0 = tune again on the next solve, 1/2/4 = force); h_wps, if not NULL,
 * receives the stored choice (0 = untuned). */
int jsso_plan_cg_tuning(jsso_plan* plan, int32_t precision, int32_t nrhs, int32_t set_wps,
                        int32_t* h_wps);

/* Tuning of the assembly's MITC-4 element kernel (no counterpart in the
 * reference): elements per workgroup (4, 8 or 16; 16 lanes per element), timed
 * on the live plan at its first jsso_assemble (every variant writes the same
 * blocks) and kept for the plan's later assemblies.  set_epw >= 0 overrides
 * (0 = tune again at the next assembly, 4/8/16 = force); h_epw, if not NULL,
 * receives the stored choice (0 = untuned). */
int jsso_plan_element_tuning(jsso_plan* plan, int32_t set_epw, int32_t* h_epw);

/* Preconditioner of the batched solves (JSSO_PREC_FP64 / JSSO_PREC_MIXED); no
 * counterpart in the reference, which factorises with SuperLU (JaxSSO/solver.py:
 * 176-218).  JSSO_PC_AMG builds a smoothed-aggregation hierarchy on the plan's
 * pattern (host, once) and its numeric operators after each jsso_assemble (device,
 * at the next solve); the plain path always uses block-Jacobi. */
int jsso_plan_set_preconditioner(jsso_plan* plan, int32_t pc);
/* Reuse of the AMG hierarchy across assemblies, for optimisation loops (no
 * counterpart in the reference, which refactorises with SuperLU every time).
 * mode 0 (default): every solve after a jsso_assemble rebuilds the numeric
 * hierarchy.  mode 1: keep the hierarchy built for an earlier design; after an
 * assembly only the level-0 smoothing bound follows the new matrix (a few
 * warm-started power-iteration steps), the coarse operators and prolongators
 * stay (the V-cycle is still a fixed SPD preconditioner, iterations grow as the
 * design moves away).  mode 2: rebuild fully at the next solve, then keep (mode
 * 1).  mode -1 only queries.  h_counts (3 entries, may be NULL) receives the
 * mode, the full set-ups and the refreshes run so far. */
int jsso_plan_amg_reuse(jsso_plan* plan, int32_t mode, int64_t* h_counts);

/* Symbolic AMG hierarchy of a plan (host only, no device work): *h_levels receives
 * the number of levels, h_nodes[l] (up to cap entries) the block rows of level l. */
int jsso_plan_amg_levels(jsso_plan* plan, int32_t* h_nodes, int32_t cap, int32_t* h_levels);

/* Measurement hook (no counterpart in the reference): sampled HIP-event timing
 * of the batched solver's dominant kernel: with block-Jacobi the fused CG
 * iteration that applies the scaled stiffness (a captured chunk of iterations
 * counts as that many launches); with AMG the direction and CG-product kernel
 * (k_pcg_dir) of one iteration chunk in `every` (that chunk is launched eagerly).
 * With block-Jacobi, every > 0 brackets one eager launch in `every` (a captured
 * chunk counts as its launches) with events on the solver's stream; 0 disables.  jsso_timing_read returns the mean duration
 * (ms) and the number of sampled launches, optionally resetting the counters. */
int jsso_timing_enable(int32_t every);
int jsso_timing_read(double* h_avg_ms, int64_t* h_samples, int32_t reset);

/* Timing slots of jsso_timing_read_slot (no counterpart in the reference).
 * SOLVE is what jsso_timing_read returns: the block-Jacobi k_cgf launches as
 * above, or with AMG the k_pcg_dir launches of one iteration chunk in `every`,
 * which then runs with eager launches (the other chunks keep their graphs).
 * The phase slots bracket every call while timing is on: ELEMENT the element
 * stiffness kernels of jsso_assemble, GATHER its gather/scale kernels (diagonal
 * factors and the solver's ELL copies), ADJOINT the kernels of
 * jsso_adjoint_grad, SETUP the AMG numeric set-up after an assembly. */
enum {
    JSSO_TIMING_SOLVE = 0, JSSO_TIMING_ELEMENT = 1, JSSO_TIMING_GATHER = 2,
    JSSO_TIMING_ADJOINT = 3, JSSO_TIMING_SETUP = 4, JSSO_TIMING_SLOTS = 5
};
int jsso_timing_read_slot(int32_t slot, double* h_avg_ms, int64_t* h_samples, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* JSSO_H */
