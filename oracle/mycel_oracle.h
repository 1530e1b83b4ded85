/*
 * mycel_oracle.h -- FP64 CPU restatement of Mycel.Core's matrix-free Hex8 elasticity path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle for the HIP product library
 * (libmycel_hip.so) and the timed CPU baseline in bench.py. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it. Nothing in mycel_amd/ links, loads or calls it.
 *
 * Every function follows the reference file:line given beside it (paths relative to the
 * reference repository root, src/Mycel.Core/...). Compile with -ffp-contract=off: .NET does not
 * fuse a*b+c, so neither does the oracle.
 */
#ifndef MYCEL_ORACLE_H
#define MYCEL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror the exception types the reference throws. */
enum {
    ORC_OK = 0,
    ORC_E_ARGUMENT_NULL = 1,
    ORC_E_ARGUMENT = 2,
    ORC_E_OUT_OF_RANGE = 3,
    ORC_E_INVALID_OPERATION = 4,
    ORC_E_NOT_CONVERGED = 5
};

/* PcgResult message codes (Fem/PcgSolver.cs:139-368). */
enum {
    ORC_PCG_CONVERGED = 0,
    ORC_PCG_RHS_ZERO = 1,
    ORC_PCG_INITIAL_CONVERGED = 2,
    ORC_PCG_CURVATURE_BREAKDOWN = 3,
    ORC_PCG_STEP_BREAKDOWN = 4,
    ORC_PCG_STAGNATION = 5,
    ORC_PCG_DIRECTION_BREAKDOWN = 6,
    ORC_PCG_MAX_ITERATIONS = 7
};

typedef struct orc_model orc_model;

/* Hex8Element.CreateKe / CreateBc / CreateConstitutiveMatrix (Fem/Hex8Element.cs:77-241). */
int orc_hex8_ke(double h, double nu, double *ke576);
int orc_hex8_bc(double h, double *bc144);
void orc_constitutive(double e, double nu, double *d36);

/* FeModel ctor + FemOperator ctor (Fem/FeModel.cs:43-184, Fem/FemOperator.cs:33-108).
 * domain: uint8[nx*ny*nz]; fixed_dofs: int32[n_fixed]; density: box-sized or NULL (all 1.0). */
int orc_model_create(int nx, int ny, int nz, double h, double e0, double nu, double emin_ratio,
                     double penal, const uint8_t *domain, const int32_t *fixed_dofs, int n_fixed,
                     const double *density, orc_model **out);
void orc_model_destroy(orc_model *m);

int orc_model_active_count(const orc_model *m);
int orc_model_active_fixed_count(const orc_model *m);
void orc_model_active_elements(const orc_model *m, int32_t *out);
void orc_model_constrained(const orc_model *m, uint8_t *out);
void orc_model_active_nodes(const orc_model *m, uint8_t *out);
const double *orc_model_ke(const orc_model *m);
void orc_model_diag_a(const orc_model *m, double *out);

/* FeModel.AssembleForces (Fem/FeModel.cs:263-321). Load cases are given as a flat node/area list
 * with case offsets; forces are float32 (System.Numerics.Vector3). */
int orc_assemble_forces(int n_nodes_total, int n_cases, const int32_t *case_offsets,
                        const int32_t *nodes, const double *areas, const float *force_xyz,
                        double *f_out);

/* rho^p as the SIMP stiffness and sensitivities take it (Fem/FemOperator.cs:100,133,
 * Optimization/SensitivityAnalysis.cs:94,141): rho^3 and rho^2 correctly rounded (error-free products),
 * other exponents through the C runtime's pow. out[i] = r[i]^p. */
void orc_pow_penal_n(const double *r, int n, double p, double *out);

/* FemOperator.UpdateDensities + AssembleDiagonal (Fem/FemOperator.cs:114-172). */
void orc_update_densities(orc_model *m, const double *density_box, double penal);

/* FemOperator.ApplyK: 8-colour, 256-element chunks, Vector256 row sums (Fem/FemOperator.cs:196-243,379-439).
 * n_threads <= 0 uses all OpenMP threads; the result is bit-identical for any thread count. */
void orc_apply_k(const orc_model *m, const double *x, double *y, int n_threads);
/* FemOperator.ApplyA (Fem/FemOperator.cs:250-294). */
void orc_apply_a(const orc_model *m, const double *x, double *y, int projected, int n_threads);
/* FemOperator.ApplyKScalarReference (Fem/FemOperator.cs:299-366). */
void orc_apply_k_scalar(const orc_model *m, const double *x, double *y);

typedef struct {
    int iterations;
    double final_residual;
    int converged;
    int message;
    double elapsed_ms;
} orc_pcg_result;

/* PcgSolver.Solve with JacobiPreconditioner (Fem/PcgSolver.cs:80-553, Fem/JacobiPreconditioner.cs:15-35).
 * x is in/out (warm start). Returns ORC_E_INVALID_OPERATION on the <6 fixed DOF preflight. */
int orc_pcg_solve(const orc_model *m, const double *b, double *x, double tol, int max_iter,
                  int n_threads, orc_pcg_result *res);

/* StaticAnalysis.Run (Fem/StaticAnalysis.cs:35-142). f: assembled forces (box-sized).
 * u_inout: warm start in, displacements out. reactions, von_mises (box elements) may be NULL. */
int orc_static_analysis(orc_model *m, const double *f, double *u_inout, double tol, int max_iter,
                        int throw_on_nonconv, int n_threads, double *compliance, double *max_disp,
                        double *reactions, double *von_mises, orc_pcg_result *res);
/* StaticAnalysis.ComputeElementVonMises (Fem/StaticAnalysis.cs:144-216). */
void orc_von_mises(const orc_model *m, const double *u, double *vm_box);
/* StaticAnalysis.AverageElementToNode (Fem/StaticAnalysis.cs:222-272). */
void orc_average_element_to_node(const orc_model *m, const double *elem_box, double *nodal);
double orc_sample_nodal_scalar(const orc_model *m, const double *nodal, float px, float py, float pz, float ox,
                               float oy, float oz);

/* DensityFilter (Optimization/DensityFilter.cs:34-303). */
typedef struct orc_filter orc_filter;
int orc_filter_create(int nx, int ny, int nz, const uint8_t *domain, const int32_t *active,
                      int n_active, double rmin, orc_filter **out);
void orc_filter_destroy(orc_filter *f);
int orc_filter_nnz(const orc_filter *f);
void orc_filter_s(const orc_filter *f, double *s_out);
void orc_filter_apply(const orc_filter *f, const double *in, double *out);
void orc_filter_apply_adjoint(const orc_filter *f, const double *in, double *out);

/* SensitivityAnalysis.ComputeComplianceSensitivities (Optimization/SensitivityAnalysis.cs:13-146). */
void orc_sensitivities(const orc_model *m, const double *u, const double *rho_phys_active,
                       double *dc_active, double penal);

/* OcOptimizer.Update (Optimization/OcOptimizer.cs:12-175). passive may be NULL. rho_new may alias rho.
 * Returns the selected lambda; *n_evals receives the number of EvaluateVolume calls. */
double orc_oc_update(const double *rho, const double *dc, const double *s, const uint8_t *passive,
                     double *rho_new, int n, double vf, double move, double rho_min,
                     double lambda_guess, int max_bisect, double vol_tol, int *n_evals);

/* PassiveRegion.Build (Voxelization/PassiveRegion.cs:13-62). */
void orc_passive_region(int nx, int ny, int nz, const int32_t *nodes, int n_nodes, int radius,
                        uint8_t *mask_box);

/* OptimizerOptions (Optimization/OptimizerOptions.cs:5-26), fields the SIMP loop reads. */
typedef struct {
    double volume_fraction;
    double filter_radius_mm;
    double penal;
    double move_limit;
    int max_iterations;
    double solver_tol_initial;
    double solver_tol_final;
    int solver_max_iterations;
    double rho_min;
    int protect_boundary_hinges;
    int start_iteration;
    int enforce_min_filter_radius;
    int continuation;
    int enable_heaviside;
} orc_simp_options;

/* HeavisideProjection.Project / Derivative (Optimization/HeavisideProjection.cs:12-37). */
void orc_heaviside_project(const double *rho_tilde, double *rho_bar, int n, double beta, double eta);
void orc_heaviside_derivative(const double *rho_tilde, double *deriv, int n, double beta, double eta);

typedef struct {
    int iteration;
    double compliance;
    double volume_fraction;
    double max_change;
    double grey_level;
    int solver_iterations;
    double solver_residual;
    double max_displacement;
    double lambda;
    double solver_tolerance; /* the relative tolerance this iteration's solves used (TopologyOptimizer.cs:212-224) */
} orc_snapshot;

/* TopologyOptimizer.Run (Optimization/TopologyOptimizer.cs:16-407), Heaviside off.
 * passive_mask_box: optional uint8[n_elem]; patch_nodes: union of load-patch nodes (for
 * ProtectBoundaryHinges); initial_density: optional (active- or box-sized, float32 as the reference);
 * warm_u: optional box DOFs. Outputs: snapshots[max_iterations], rho/rho_bar active-sized FP64,
 * u box-sized. Returns ORC_E_NOT_CONVERGED if a PCG solve fails (the reference throws). */
int orc_simp_run_cases(orc_model *m, const double *f, int n_cases, const orc_simp_options *opt,
                       const uint8_t *passive_mask_box, const int32_t *patch_nodes, int n_patch_nodes,
                       const float *initial_density, int initial_density_len, const double *warm_u,
                       int n_threads, orc_snapshot *snapshots, int *n_snapshots, int *converged,
                       double *rho_out, double *rho_bar_out, double *u_out);
int orc_simp_run(orc_model *m, const double *f, const orc_simp_options *opt,
                 const uint8_t *passive_mask_box, const int32_t *patch_nodes, int n_patch_nodes,
                 const float *initial_density, int initial_density_len, const double *warm_u,
                 int n_threads, orc_snapshot *snapshots, int *n_snapshots, int *converged,
                 double *rho_out, double *rho_bar_out, double *u_out);

#ifdef __cplusplus
}
#endif

#endif
