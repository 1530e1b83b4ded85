/*
 * mycel_hip.h -- C ABI of libmycel_hip.so, the MI355X (gfx950) implementation of Mycel.Core's
 * matrix-free Hex8 elasticity path.
 *
 * Each entry point replaces one managed member of Mycel.Core; the reference interface is cited
 * beside it (paths relative to the reference repository root). A P/Invoke shim re-implements the
 * managed class bodies on top of these calls.
 *
 * Conventions
 *   - Every function returns an int32 status (MYCEL_OK = 0). The codes map to the exceptions the
 *     reference throws; mycel_last_error() returns a thread-local message for the last failure.
 *   - Host arrays are borrowed for the duration of the call only; nothing is retained.
 *   - Vectors over DOFs are box-sized, interleaved [ux,uy,uz] per node, node index
 *     n = i + (nx+1)(j + (ny+1)k), DOF = 3n + c (Voxelization/GridSpec.cs:109-153).
 *     Design-variable arrays (rho, rho_bar, dc, s) are active-compact in ascending element order.
 *   - Functions whose name ends in _d take device pointers (HBM-resident data) and a hipStream_t
 *     passed as void* (NULL = the handle's own stream); they return when the work is enqueued
 *     unless stated otherwise.
 *   - One thread per handle at a time; distinct handles are independent, each on its own stream.
 *     PCG breakdowns are not errors: they return converged = 0 with a message code, as PcgResult.
 */
#ifndef MYCEL_HIP_H
#define MYCEL_HIP_H

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define MYCEL_API __attribute__((visibility("default")))
#else
#define MYCEL_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define MYCEL_OK 0
#define MYCEL_E_ARGUMENT_NULL 1     /* ArgumentNullException */
#define MYCEL_E_ARGUMENT 2          /* ArgumentException (length mismatch, bad warm start) */
#define MYCEL_E_OUT_OF_RANGE 3      /* ArgumentOutOfRangeException (fixed DOF, h <= 0) */
#define MYCEL_E_INVALID_OPERATION 4 /* InvalidOperationException (<6 fixed DOFs, empty patch) */
#define MYCEL_E_NOT_CONVERGED 5     /* InvalidOperationException thrown on PCG non-convergence */
#define MYCEL_E_CANCELED 6          /* OperationCanceledException */
#define MYCEL_E_DEVICE 7            /* HIP runtime failure */
#define MYCEL_E_OUT_OF_MEMORY 8     /* device allocation or HBM budget */

/* PcgResult.Message codes (Fem/PcgSolver.cs:139-368). */
#define MYCEL_PCG_CONVERGED 0
#define MYCEL_PCG_RHS_ZERO 1
#define MYCEL_PCG_INITIAL_CONVERGED 2
#define MYCEL_PCG_CURVATURE_BREAKDOWN 3
#define MYCEL_PCG_STEP_BREAKDOWN 4
#define MYCEL_PCG_STAGNATION 5
#define MYCEL_PCG_DIRECTION_BREAKDOWN 6
#define MYCEL_PCG_MAX_ITERATIONS 7
#define MYCEL_PCG_CANCELED 8

/* Precision / preconditioner of the solve.
 * FP64: the reference's Jacobi-PCG arithmetic throughout (Fem/PcgSolver.cs:80-370).
 * MIXED_FP32: Jacobi-PCG iterating on FP32 vectors with the FP32 operator and FP64 dot products, with FP64
 *   reliable updates (the true residual r = P(b - K u) in FP64) and convergence on that true residual.
 * MIXED_BF16: MIXED_FP32 with every search direction rounded to bfloat16 where it is formed (the first, each p
 *   update, the one after a reliable update), so each K*p contraction takes BF16 inputs and accumulates in FP32;
 *   the FP64 reliable updates and the convergence test are MIXED_FP32's. A precision study: it converges to
 *   the same FP64 residual in more iterations.
 * FP64_ORDERED: the reference's Jacobi-PCG with every sum formed in the reference's order (colour-ordered K*u with
 *   4-lane Ke rows, 8,192-DOF blocked dot products with 4 lane chains, sequential ||Pb||) and no contracted
 *   multiply-adds: bit-identical to the reference restatement, for parity runs. In mycel_simp_run it also takes
 *   the OC volumes, the compliance and the snapshot sums sequentially.
 * MIXED_GMG: FP64 flexible PCG (FP64 operator, residual, dot products, convergence test and update)
 *   preconditioned by one FP32 geometric-multigrid V-cycle per iteration in place of Jacobi: a second
 *   IPreconditioner (Fem/IPreconditioner.cs:9-20). On a grid that does not coarsen by 2 per axis down to
 *   <= 140 DOFs (an axis reaches one element first) the solve runs the FP64 Jacobi iteration instead and reports
 *   all its iterations as fallback_iterations. */
#define MYCEL_PRECISION_FP64 0
#define MYCEL_PRECISION_MIXED_FP32 1
#define MYCEL_PRECISION_MIXED_BF16 2
#define MYCEL_PRECISION_MIXED_GMG 3
#define MYCEL_PRECISION_FP64_ORDERED 4

/* K*u forms (mycel_op_set_matvec_form). */
#define MYCEL_MATVEC_AUTO 0        /* by grid size (the default) */
#define MYCEL_MATVEC_NODE_GATHER 1 /* one thread per node (columns 0) or per 2/4/8-node z-column */
#define MYCEL_MATVEC_ELEMENT 2     /* element-centric: 8x8 element-column tiles, columns = 2/3/4/8 node layers */
#define MYCEL_MATVEC_REFERENCE_ORDER 3 /* the reference's colour order and 4-lane row sums (bit-identical) */

typedef struct mycel_op mycel_op;
typedef struct mycel_filter mycel_filter;

/* GridSpec (Voxelization/GridSpec.cs:11-44). h is the float32 voxel pitch widened to double. */
typedef struct {
    int32_t nx, ny, nz;
    double h;
} mycel_grid;

/* FeModel counters (Fem/FeModel.cs:21-29). */
typedef struct {
    int32_t total_elements, total_nodes, total_dofs;
    int32_t active_elements, active_nodes, active_fixed_dofs;
} mycel_op_info;

/* PcgResult (Fem/PcgResult.cs:7-14); message is a MYCEL_PCG_* code. */
typedef struct {
    int32_t iterations;
    double final_residual;
    int32_t converged;
    int32_t message;
    double elapsed_ms;
    int32_t outer_corrections;   /* reliable updates with the FP64 operator (mixed precision only) */
    int32_t fallback_iterations; /* FP64 iterations after an FP32 breakdown (mixed precision only) */
    double message_value;        /* the quantity a breakdown message quotes: pAp (CURVATURE), alpha (STEP) or
                                    beta (DIRECTION) */
} mycel_pcg_result;

/* IProgress<PcgProgress> (Fem/PcgResult.cs:5). Called on the calling thread between iterations, once per batch
 * of iterations the device ran without a host round trip (32 Jacobi, 4 multigrid or 16 reference-order
 * iterations), with the latest iteration and relative residual. A solve given a progress callback or a cancel
 * flag reads the device state after every batch; one given neither runs the first batches of a solve blind
 * (covering up to 3/4 of the previous converged Jacobi solve's count, 1/2 of the multigrid one's). */
typedef void (*mycel_pcg_progress_cb)(int32_t iteration, double relative_residual, void *user);

/* ---- library ------------------------------------------------------------------------------ */
MYCEL_API const char *mycel_last_error(void);
MYCEL_API int32_t mycel_version(void);
/* PcgResult.Message (Fem/PcgSolver.cs:149-367) of a result, as the reference words it (UTF-8, NUL-terminated,
 * truncated to capacity - 1 bytes); an empty string for a converged solve, whose Message is null. */
MYCEL_API int32_t mycel_pcg_message(const mycel_pcg_result *result, char *buf, int32_t capacity);
/* Number of visible HIP devices; selects device `ordinal` for handles created afterwards. */
MYCEL_API int32_t mycel_device_count(int32_t *count);
MYCEL_API int32_t mycel_set_device(int32_t ordinal);

/* ---- Hex8 element (Fem/Hex8Element.cs) ---------------------------------------------------- */
/* Hex8Element.CreateKe(h, nu) (Fem/Hex8Element.cs:98-178): 24x24 row-major, E = 1. */
MYCEL_API int32_t mycel_hex8_ke(double h, double nu, double *ke576);
/* Hex8Element.CreateBc(h) (Fem/Hex8Element.cs:220-241): 6x24 row-major. */
MYCEL_API int32_t mycel_hex8_bc(double h, double *bc144);

/* ---- FeModel helpers (Fem/FeModel.cs) ----------------------------------------------------- */
/* FeModel.AssembleForces (Fem/FeModel.cs:263-321). Load case c owns nodes/areas
 * [case_offsets[c], case_offsets[c+1]); force_xyz holds 3 float32 per case (System.Numerics.Vector3). */
MYCEL_API int32_t mycel_assemble_forces(int32_t total_nodes, int32_t n_cases, const int32_t *case_offsets,
                              const int32_t *nodes, const double *areas, const float *force_xyz,
                              double *f_out);
/* PassiveRegion.Build(grid, null, nodes, radius) (Voxelization/PassiveRegion.cs:13-62). */
MYCEL_API int32_t mycel_passive_region(const mycel_grid *grid, const int32_t *nodes, int32_t n_nodes,
                             int32_t radius, uint8_t *mask_box);

/* ---- FemOperator (Fem/FemOperator.cs) ----------------------------------------------------- */
/* new FeModel(grid, domain, FeMaterial(e0, nu, emin_ratio), fixedDofs, elementDensity, penal)
 * followed by new FemOperator(model) (Fem/FeModel.cs:43-184, Fem/FemOperator.cs:33-108).
 * density_box may be NULL (all 1.0, FeModel.cs:74-78). */
MYCEL_API int32_t mycel_op_create(const mycel_grid *grid, const uint8_t *domain, const int32_t *fixed_dofs,
                        int32_t n_fixed, double e0, double nu, double emin_ratio, double penal,
                        const double *density_box, mycel_op **out);
MYCEL_API int32_t mycel_op_destroy(mycel_op *op);
MYCEL_API int32_t mycel_op_info_get(const mycel_op *op, mycel_op_info *info);
/* FemOperator.UpdateDensities(densities, penal) (Fem/FemOperator.cs:114-138): box-sized densities. */
MYCEL_API int32_t mycel_op_update_densities(mycel_op *op, const double *density_box, double penal);
MYCEL_API int32_t mycel_op_update_densities_d(mycel_op *op, const double *d_density_box, double penal, void *stream);
/* FemOperator.ApplyK(x, y) (Fem/FemOperator.cs:196-243). Synchronous. */
MYCEL_API int32_t mycel_op_apply_k(mycel_op *op, const double *x, double *y);
/* FemOperator.ApplyA(x, y, xIsAlreadyProjected) (Fem/FemOperator.cs:250-294). Synchronous. */
MYCEL_API int32_t mycel_op_apply_a(mycel_op *op, const double *x, double *y, int32_t x_is_projected);
/* Device-pointer forms of ApplyK / ApplyA. */
MYCEL_API int32_t mycel_op_apply_k_d(mycel_op *op, const double *d_x, double *d_y, void *stream);
MYCEL_API int32_t mycel_op_apply_a_d(mycel_op *op, const double *d_x, double *d_y, int32_t x_is_projected, void *stream);
/* The FP32 operator of the mixed-precision solver (FP32 Ke and E_e), device pointers, FP32 DOF vectors. */
MYCEL_API int32_t mycel_op_apply_k_f32_d(mycel_op *op, const float *d_x, float *d_y, void *stream);
/* K*u kernel form (a tuning option with no counterpart in the reference) for ApplyK / ApplyA and the FP32
 * operator; the solvers' fused-dot launches always use the automatic form. MYCEL_MATVEC_NODE_GATHER: columns
 * = 0 runs one node per thread, 2 / 4 / 8 let each thread march that many nodes along z (all bit-identical to
 * each other). MYCEL_MATVEC_ELEMENT: each element's product once, in the monomial basis, columns = 2 / 3 / 4 /
 * 8 node layers per tile (all bit-identical to each other). MYCEL_MATVEC_AUTO (columns ignored): by grid size.
 * The two families agree to rounding (both within 1e-12 of the FP64 reference). */
MYCEL_API int32_t mycel_op_set_matvec_form(mycel_op *op, int32_t form, int32_t columns);
/* Island deflation of the multigrid preconditioner (no counterpart in the reference): with solid_ratio > 0, every
 * face-connected cluster of elements with E >= solid_ratio * E0 that touches no fixed DOF (a floating island, held
 * only by the void around it) adds its six rigid-body modes to the V-cycle as an additive coarse correction
 * z = V(r) + Z (Z^T A Z)^-1 Z^T r, formed in FP64 (up to 64 islands, the largest). Those modes are near-null
 * modes of A at high stiffness contrast that the smoother and the geometric coarse levels do not represent. The
 * islands are found again whenever the densities change. 0 (the default) turns it off. */
MYCEL_API int32_t mycel_op_set_island_deflation(mycel_op *op, double solid_ratio);
/* The gather form with columns 0 / 2 / 4 / 8, or -1 for MYCEL_MATVEC_AUTO. */
MYCEL_API int32_t mycel_op_set_matvec_columns(mycel_op *op, int32_t columns);
/* FemOperator.DiagA / Ke / ConstrainedDofs / ActiveElements properties. */
MYCEL_API int32_t mycel_op_get_diag_a(mycel_op *op, double *diag);
MYCEL_API int32_t mycel_op_get_ke(mycel_op *op, double *ke576);
MYCEL_API int32_t mycel_op_get_constrained(mycel_op *op, uint8_t *mask_dofs);
MYCEL_API int32_t mycel_op_get_active_elements(mycel_op *op, int32_t *active);
/* Stream the handle enqueues on (hipStream_t as void*). */
MYCEL_API void *mycel_op_stream(mycel_op *op);

/* ---- PcgSolver (Fem/PcgSolver.cs) + JacobiPreconditioner (Fem/JacobiPreconditioner.cs) ---- */
/* PcgSolver.Solve(op, b, x, new JacobiPreconditioner(), tol, maxIterations, progress, ct)
 * (Fem/PcgSolver.cs:80-370). x is the warm start on entry and the solution on return.
 * cancel: optional host flag polled after every device batch (nonzero = cancel -> MYCEL_E_CANCELED).
 * Returns MYCEL_E_INVALID_OPERATION when fewer than 6 fixed DOFs sit on active nodes. */
MYCEL_API int32_t mycel_pcg_solve(mycel_op *op, const double *b, double *x, double tol, int32_t max_iter,
                        int32_t precision, const volatile int32_t *cancel,
                        mycel_pcg_progress_cb progress, void *user, mycel_pcg_result *result);
/* Device-pointer form; synchronous on the handle's stream. */
MYCEL_API int32_t mycel_pcg_solve_d(mycel_op *op, const double *d_b, double *d_x, double tol, int32_t max_iter,
                          int32_t precision, mycel_pcg_result *result);

/* IPreconditioner.Setup + Apply (Fem/IPreconditioner.cs:9-20) on device vectors: z = M^-1 r with the Jacobi
 * preconditioner (precision FP64 or FP64_ORDERED; JacobiPreconditioner.cs:15-35) or one multigrid V-cycle
 * (MIXED_GMG; its hierarchy is set up or refreshed first). r is a projected residual (zero on constrained DOFs).
 * Enqueued on `stream` (NULL = the handle's stream). */
MYCEL_API int32_t mycel_precond_apply_d(mycel_op *op, int32_t precision, const double *d_r, double *d_z, void *stream);

/* Several right-hand sides of one operator solved in lockstep (the multi-RHS form of PcgSolver.Solve; no
 * counterpart in the reference, which solves one RHS at a time): d_b and d_x hold n_rhs box-DOF vectors
 * back to back (x: warm starts in, solutions out); every RHS keeps its own stopping test and result. Kernels
 * of all RHS share each launch. Host pointers (mycel_pcg_solve_batch) or device pointers (_d, synchronous on
 * the handle's stream); precision must be MYCEL_PRECISION_MIXED_GMG. results: n_rhs entries. */
MYCEL_API int32_t mycel_pcg_solve_batch(mycel_op *op, const double *b, double *x, int32_t n_rhs, double tol,
                                        int32_t max_iter, int32_t precision, mycel_pcg_result *results);
MYCEL_API int32_t mycel_pcg_solve_batch_d(mycel_op *op, const double *d_b, double *d_x, int32_t n_rhs, double tol,
                                          int32_t max_iter, int32_t precision, mycel_pcg_result *results);
/* Independent systems on different operators (each its own grid and handle) solved together: every multigrid
 * iteration is one launch per kernel for all systems (no counterpart in the reference, which solves one model at
 * a time; the job server packs its jobs' solves the same way). count <= 16 distinct handles; d_b[i], d_x[i]:
 * device box-DOF vectors of ops[i] (x: warm start in, solution out). Precision MYCEL_PRECISION_MIXED_GMG; each
 * result and solution is bit-identical to mycel_pcg_solve_d of that system alone. Synchronous. */
MYCEL_API int32_t mycel_pcg_solve_packed_d(mycel_op *const *ops, const double *const *d_b, double *const *d_x,
                                           int32_t count, double tol, int32_t max_iter, mycel_pcg_result *results);

/* ---- StaticAnalysis (Fem/StaticAnalysis.cs) ----------------------------------------------- */
/* StaticAnalysis.Run(model, options) (Fem/StaticAnalysis.cs:35-142). f: assembled forces (box DOFs),
 * u: warm start in (zeros for a cold start), displacements out. reactions (box DOFs) and
 * von_mises (box elements) may be NULL. throw_on_nonconvergence -> MYCEL_E_NOT_CONVERGED. */
MYCEL_API int32_t mycel_static_analysis(mycel_op *op, const double *f, double *u, double tol, int32_t max_iter,
                              int32_t throw_on_nonconvergence, int32_t precision, double *compliance,
                              double *max_displacement, double *reactions, double *von_mises,
                              mycel_pcg_result *result);
/* StaticAnalysis.ComputeElementVonMises (Fem/StaticAnalysis.cs:144-216). */
MYCEL_API int32_t mycel_von_mises(mycel_op *op, const double *u, double *von_mises_box);
/* StaticAnalysis.AverageElementToNode (Fem/StaticAnalysis.cs:222-272). */
MYCEL_API int32_t mycel_average_element_to_node(mycel_op *op, const double *element_box, double *nodal);
/* StaticAnalysis.SampleNodalScalar (Fem/StaticAnalysis.cs:278-366) at n_points points at once: points_xyz holds
 * float32 (x, y, z) triples (System.Numerics.Vector3), origin_xyz the GridSpec.Origin; out[q] is the trilinear
 * interpolation over the active nodes of the point's cell, or the value at the nearest active node. */
MYCEL_API int32_t mycel_sample_nodal_scalar(mycel_op *op, const double *nodal, int32_t n_points,
                                            const float *points_xyz, const float *origin_xyz, double *out);

/* ---- DensityFilter (Optimization/DensityFilter.cs) ---------------------------------------- */
/* new DensityFilter(grid, domain, activeElements, rmin) (Optimization/DensityFilter.cs:34-199). */
MYCEL_API int32_t mycel_filter_create(const mycel_grid *grid, const uint8_t *domain, const int32_t *active,
                            int32_t n_active, double rmin, mycel_filter **out);
MYCEL_API int32_t mycel_filter_destroy(mycel_filter *f);
/* DensityFilter.S (Optimization/DensityFilter.cs:20,187-198). */
MYCEL_API int32_t mycel_filter_get_s(mycel_filter *f, double *s);
MYCEL_API int32_t mycel_filter_nnz(mycel_filter *f, int64_t *nnz);
/* ApplyFilter (:205-251) and ApplySensitivityFilter (:257-303). */
MYCEL_API int32_t mycel_filter_apply(mycel_filter *f, const double *in, double *out);
MYCEL_API int32_t mycel_filter_apply_adjoint(mycel_filter *f, const double *in, double *out);

/* ---- SensitivityAnalysis (Optimization/SensitivityAnalysis.cs:13-146) --------------------- */
MYCEL_API int32_t mycel_sensitivities(mycel_op *op, const double *u, const double *rho_phys_active,
                            double *dc_active, double penal);

/* ---- OcOptimizer.Update (Optimization/OcOptimizer.cs:12-175) ------------------------------ */
/* passive may be NULL. rho_new may alias rho. lambda_out receives the selected multiplier. */
MYCEL_API int32_t mycel_oc_update(const double *rho, const double *dc, const double *s, const uint8_t *passive,
                        double *rho_new, int32_t n, double volume_fraction, double move_limit,
                        double rho_min, double lambda_guess, int32_t max_bisect, double volume_tol,
                        double *lambda_out);

/* No counterpart in the reference: OcOptimizer.Update on device pointers for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It allocates
 * nothing: its scratch (the bisection state and the partial volume sums) is one fixed device array of
 * 640 KiB + 256 B on each device the library runs on, held until the process ends, which bounds n to
 * MYCEL_OC_NOALLOC_MAX_ELEMENTS (larger: MYCEL_E_OUT_OF_RANGE). Every call on a device shares that array, so
 * calls that overlap on one device race. d_passive may be NULL; d_rho_new may alias d_rho. Synchronous on
 * `stream` (a hipStream_t as void*, NULL = the null stream). Results are identical to mycel_oc_update's. */
#define MYCEL_OC_NOALLOC_MAX_ELEMENTS 16777216
MYCEL_API int32_t mycel_oc_update_noalloc_d(const double *d_rho, const double *d_dc, const double *d_s,
                                            const uint8_t *d_passive, double *d_rho_new, int32_t n,
                                            double volume_fraction, double move_limit, double rho_min,
                                            double lambda_guess, int32_t max_bisect, double volume_tol, void *stream,
                                            double *lambda_out);

/* ---- TopologyOptimizer (Optimization/TopologyOptimizer.cs) -------------------------------- */
/* OptimizerOptions (Optimization/OptimizerOptions.cs:5-26): the fields the SIMP loop reads.
 * The material's Emin comes from the operator (FeMaterial.EminRatio), as in the reference. */
typedef struct {
    double volume_fraction;     /* 0.4 */
    double filter_radius_mm;    /* 2.0 */
    double penal;               /* 3.0 */
    double move_limit;          /* 0.2 */
    int32_t max_iterations;     /* 200 */
    double solver_tol_initial;  /* 1e-4 */
    double solver_tol_final;    /* 1e-6 */
    int32_t solver_max_iterations; /* 5000 */
    double rho_min;             /* 1e-3 */
    int32_t protect_boundary_hinges; /* 1 */
    int32_t start_iteration;    /* 1 */
    int32_t enforce_min_filter_radius; /* 1 */
    int32_t continuation;       /* 0 */
    int32_t precision;          /* MYCEL_PRECISION_* for the linear solves */
    int32_t enable_heaviside;   /* 0: smooth Heaviside projection of the filtered field, beta 1 -> 8
                                   (Optimization/HeavisideProjection.cs, TopologyOptimizer.cs:193-201,281-289,315-318) */
    double initial_lambda;      /* 1.0: the OC multiplier's first guess (TopologyOptimizer.cs:154); a resumed run
                                   passes the checkpointed one */
    const double *initial_design; /* NULL: optional FP64 initial design density, box- or active-sized; when set it
                                     replaces the float32 initial_density of mycel_simp_run (an exact resume from
                                     a checkpoint's FP64 design; the reference's InitialDensity is float32) */
    int32_t initial_design_len;
    /* The loop state a resumed run continues from (no counterpart in the reference, whose StartIteration restarts
     * the schedule: TopologyOptimizer.cs:154-162,212-224). resume_state = 0 (the default, and what a zeroed struct
     * holds): none, the start iteration solves to solver_tol_initial as the reference does. resume_state = 1: the
     * start iteration's tolerance follows the schedule from resume_max_change (the previous iteration's max
     * change), the Heaviside beta is resume_beta, and the stop test starts from resume_history, the last
     * resume_history_len (<= 6) compliances, oldest first. */
    int32_t resume_state;
    double resume_max_change;
    double resume_beta;
    int32_t resume_history_len;
    double resume_history[6];
    double island_deflation; /* 0: the multigrid solver's island deflation (mycel_op_set_island_deflation) for this
                                run's solves, with this solid ratio; 0 keeps the operator's setting */
} mycel_simp_options;

/* Fills the reference defaults (OptimizerOptions.cs:5-26). */
MYCEL_API void mycel_simp_options_default(mycel_simp_options *o);

/* IterationSnapshot + IterationTimings (Optimization/IterationSnapshot.cs:5-31), FP64 scalars. */
typedef struct {
    int32_t iteration;
    double compliance;
    double volume_fraction;
    double max_change;
    double grey_level;
    int32_t solver_iterations;
    double solver_residual;
    double max_displacement;
    double lambda;
    double solve_ms, sensitivity_ms, filter_ms, update_ms, total_ms;
    double beta_next; /* the Heaviside beta of the next iteration (1 without projection): resume state */
} mycel_snapshot;

/* progress.Report(snapshot) (TopologyOptimizer.cs:341-372). rho_bar (the snapshot's PhysicalDensity) and rho
 * (its DesignDensity, after this iteration's OC update): active-compact FP64, valid during the call only.
 * Return nonzero to stop after this iteration (cooperative cancel). */
typedef int32_t (*mycel_snapshot_cb)(const mycel_snapshot *snap, const double *rho_bar, const double *rho,
                                     void *user);

typedef struct {
    int32_t converged;
    int32_t total_iterations;
    double final_compliance;
    double final_volume_fraction;
    int32_t canceled;
} mycel_simp_result;

/* TopologyOptimizer.Run (Optimization/TopologyOptimizer.cs:16-407); Heaviside projection and p-continuation
 * as the options select them.
 * f: assembled forces; passive_mask_box: optional OptimizerOptions.PassiveMask; patch_nodes: the
 * load-patch nodes (for ProtectBoundaryHinges); initial_density: optional float32, box- or
 * active-sized; warm_u: optional box DOFs. rho_out / rho_bar_out (active) and u_out (box DOFs)
 * may be NULL. Returns MYCEL_E_NOT_CONVERGED when a PCG solve fails (the reference throws). */
MYCEL_API int32_t mycel_simp_run(mycel_op *op, const double *f, const mycel_simp_options *opt,
                       const uint8_t *passive_mask_box, const int32_t *patch_nodes, int32_t n_patch_nodes,
                       const float *initial_density, int32_t initial_density_len, const double *warm_u,
                       mycel_snapshot_cb cb, void *user, mycel_simp_result *result,
                       mycel_snapshot *snapshots, int32_t snapshots_capacity,
                       double *rho_out, double *rho_bar_out, double *u_out);
/* Several independent load cases on one mesh (the multi-RHS configuration): f, warm_u and u_out hold
 * n_cases box-DOF vectors back to back, and the objective is the summed compliance sum_k f_k.u_k with
 * dc = sum_k dc_k. Each case is solved on its own, warm-started from its previous solution; a snapshot's
 * solver_iterations is the sum over the cases, solver_residual and max_displacement the largest. No
 * counterpart in the reference, whose FeModel.AssembleForces sums load cases into one RHS
 * (Fem/FeModel.cs:267-318); n_cases = 1 is mycel_simp_run. */
MYCEL_API int32_t mycel_simp_run_cases(mycel_op *op, const double *f, int32_t n_cases, const mycel_simp_options *opt,
                                       const uint8_t *passive_mask_box, const int32_t *patch_nodes,
                                       int32_t n_patch_nodes, const float *initial_density, int32_t initial_density_len,
                                       const double *warm_u, mycel_snapshot_cb cb, void *user,
                                       mycel_simp_result *result, mycel_snapshot *snapshots,
                                       int32_t snapshots_capacity, double *rho_out, double *rho_bar_out,
                                       double *u_out);


/* ---- Run persistence (Persistence/RunWriter.cs, Persistence/RunReader.cs) -------------------- */
/* The 8-bit frame of an iteration file and its checksum (RunWriter.cs:215-229,376-388): q[i] =
 * round-half-away(clamp(phys[i], 0, 1) * 255) in float32, checksum = FNV-1a (32-bit) over the bytes q.
 * q may be NULL (checksum only). */
MYCEL_API int32_t mycel_run_quantize(const float *phys, int64_t n, uint8_t *q, uint32_t *checksum);
/* FNV-1a (32-bit) over n bytes, as RunReader.ReadIteration verifies a frame (RunReader.cs:213-223). */
MYCEL_API int32_t mycel_run_checksum(const uint8_t *q, int64_t n, uint32_t *checksum);

/* ---- Job server (continuous batching; no counterpart in the reference, whose only job machinery is
 * one run at a time behind a RAM guard: Optimization/OptimizationCoordinator.cs:214-353,
 * Diagnostics/MemoryGuard.cs:434-495) ------------------------------------------------------- */
typedef struct mycel_server mycel_server;

#define MYCEL_JOB_FEA 0  /* StaticAnalysis.Run */
#define MYCEL_JOB_SIMP 1 /* TopologyOptimizer.Run */

#define MYCEL_JOB_QUEUED 0
#define MYCEL_JOB_RUNNING 1
#define MYCEL_JOB_DONE 2
#define MYCEL_JOB_FAILED 3

/* One job. Arrays are copied at submit time. forces: assembled box DOFs; patch_nodes: the load-patch
 * nodes (ProtectBoundaryHinges); density_box: optional (FEA). tol/max_iter/precision drive FEA jobs,
 * simp drives SIMP jobs (its precision field selects the solver precision). */
typedef struct {
    int32_t kind;
    mycel_grid grid;
    const uint8_t *domain;
    const int32_t *fixed_dofs;
    int32_t n_fixed;
    const double *forces;
    const int32_t *patch_nodes;
    int32_t n_patch_nodes;
    double e0, nu, emin_ratio, penal;
    const double *density_box;
    double tol;
    int32_t max_iter;
    int32_t precision;
    mycel_simp_options simp;
} mycel_job_desc;

typedef struct {
    int32_t state;          /* MYCEL_JOB_* */
    int32_t status;         /* MYCEL_OK or the failure status */
    double compliance;
    double max_displacement;
    int32_t pcg_iterations; /* total over the job */
    int32_t simp_iterations;
    int32_t converged;
    double footprint_bytes; /* HBM charged against the budget */
    double queue_ms, run_ms, latency_ms;
} mycel_job_info;

/* hbm_budget_bytes <= 0 uses 95% of the device's free memory; workers = jobs resident at once
 * (each on its own HIP stream). */
MYCEL_API int32_t mycel_server_create(double hbm_budget_bytes, int32_t workers, mycel_server **out);
/* Waits for every submitted job to finish, then frees the server. */
MYCEL_API int32_t mycel_server_destroy(mycel_server *server);
MYCEL_API int32_t mycel_server_submit(mycel_server *server, const mycel_job_desc *job, int64_t *job_id);
MYCEL_API int32_t mycel_server_poll(mycel_server *server, int64_t job_id, mycel_job_info *info);
/* timeout_s < 0 waits without limit. */
MYCEL_API int32_t mycel_server_wait(mycel_server *server, int64_t job_id, double timeout_s, mycel_job_info *info);
/* u_out: box DOFs; rho_bar_out: active elements (SIMP jobs). Either may be NULL. */
MYCEL_API int32_t mycel_server_fetch(mycel_server *server, int64_t job_id, double *u_out, double *rho_bar_out);
MYCEL_API int32_t mycel_server_release(mycel_server *server, int64_t job_id);
MYCEL_API int32_t mycel_server_stats(mycel_server *server, double *budget_bytes, double *in_use_bytes,
                                     int32_t *queued, int32_t *active);

/* Packing (no counterpart in the reference): the multigrid solves of the resident jobs (FEA jobs and every SIMP
 * iteration's solve, MYCEL_PRECISION_MIXED_GMG) advance together, each iteration one launch per kernel for all
 * of them (up to 16), 4 iterations per batch; solves join and leave between batches, and a membership that
 * stays the same for a batch is captured as a hipGraph. A job's result is bit-identical packed or not.
 * On by default; set before the first submit. Jobs whose solve cannot be packed (island deflation, grids
 * without a multigrid hierarchy, other precisions) run on their worker's stream as before. */
MYCEL_API int32_t mycel_server_set_packing(mycel_server *server, int32_t enable);
typedef struct {
    int64_t batches;        /* packed batches run */
    int64_t member_batches; /* sum over batches of the solves in the batch */
    int64_t launches;       /* kernel launches of those batches */
    int64_t graph_batches;  /* batches replayed from a captured graph */
    int64_t rebuilds;       /* membership changes (tables rebuilt) */
    int64_t solves;         /* solves that left the pack */
    double build_ms;        /* broker host time: rebuilding tables and capturing graphs */
    double enqueue_ms;      /* broker host time: enqueuing batches */
    double wait_ms;         /* broker host time: waiting for batches to finish on the device */
} mycel_pack_stats;
MYCEL_API int32_t mycel_server_pack_stats(mycel_server *server, mycel_pack_stats *out);

#ifdef __cplusplus
}
#endif

#endif
