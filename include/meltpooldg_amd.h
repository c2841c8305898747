/*
 * meltpooldg_amd.h -- C ABI of the MI355X (gfx950) matrix-free operator library.
 *
 * The library replaces the deal.II MatrixFree/FEEvaluation cell loop behind MeltPoolDG's
 * heat-transfer Jacobian and level-set advection-diffusion operators, plus the Jacobi
 * preconditioned CG/GMRES solves that call them.  Every entry point names the reference
 * interface it stands in for.  All vector arguments are DEVICE pointers (hipMalloc'ed or torch
 * device tensors) of n_dofs entries in the library's lexicographic node numbering; work is
 * enqueued on the stream set with mf_set_stream (default: the null stream).  Functions return
 * an MF_* status; mf_last_error() gives the message of the last failure on the calling thread.
 * Contexts are not re-entrant (as the reference operators, which mutate caches in const
 * methods): one thread and one stream per context.
 */
#ifndef MELTPOOLDG_AMD_H
#define MELTPOOLDG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
  MF_OK = 0,
  MF_ERR_ARG = 1,          /* invalid argument (reference: AssertThrow / ExcMessage) */
  MF_ERR_HIP = 2,          /* HIP runtime error */
  MF_NO_CONVERGENCE = 3,   /* reference: SolverControl::NoConvergence */
  MF_BREAKDOWN = 4,        /* Krylov breakdown (zero or NaN denominator) */
  MF_ERR_STATE = 5         /* call order violated (e.g. vmult before set_linearization) */
};

const char *mf_last_error(void);
int mf_version(void);
/* Tuning knob without reference counterpart: which device kernel applies the tangent
 * (0 = pencil kernel, one LDS round trip per 1D sweep, the only variant at present).
 * Process-wide setting. */
int mf_set_tangent_kernel(int variant);

/* ------------------------------------------------------------------------------------------ */
/* Mesh / DoF / constraint setup.  Replaces ScratchData::build + MatrixFree::reinit and the     */
/* DoF-index/constraint setup (source/core/scratch_data.cpp:194-293,                            */
/* include/meltpooldg/utilities/constraints.hpp:254-269, include/meltpooldg/utilities/          */
/* fe_util.hpp:21-81).  The mesh is a structured hex mesh of [lo,hi] with ncell cells per       */
/* direction, FE_Q(degree) on Gauss-Lobatto points and QGauss(degree+1); DoFs are numbered      */
/* lexicographically on the global node lattice (x fastest).                                    */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int dim;          /* must be 3 */
  int degree;       /* 1, 2 or 3 */
  int ncell[3];
  double lo[3];
  double hi[3];
} mf_mesh_desc;

typedef struct mf_scratch mf_scratch;

int mf_scratch_create(const mf_mesh_desc *desc, mf_scratch **out);
int mf_scratch_destroy(mf_scratch *s);
int64_t mf_scratch_n_dofs(const mf_scratch *s);
int64_t mf_scratch_n_cells(const mf_scratch *s);
/* Stream on which all work of this scratch and its operators is enqueued (hipStream_t). */
int mf_set_stream(mf_scratch *s, void *hip_stream);
/* Dirichlet constraints (the heat_dof_idx / advec_diff_dof_idx constraint set): a host array of
 * n_dofs bytes, 1 = constrained.  NULL clears them.  Replaces
 * Constraints::make_DBC_and_HNC_plus_PBC_and_merge_HNC_plus_PBC_into_DBC (constraints.hpp:254-269)
 * for meshes without hanging nodes. */
int mf_scratch_set_dirichlet_mask(mf_scratch *s, const uint8_t *mask_host);
/* 1D tables the kernels use (K = degree+1): S[K][K] = l_j(x_q), DS[K][K] = l_j'(x_q),
 * Dq[K][K] collocation derivative on the Gauss points, w[K] Gauss weights, gl[K] support points */
int mf_scratch_shape_info(const mf_scratch *s, double *S, double *DS, double *Dq, double *w,
                          double *gl);

/* --- adaptive meshes with hanging nodes --------------------------------------------------- */
/* Band-refined octree mesh over a base box mesh: cells meeting the spherical shell
 * | |x - center| - radius | < half_width are refined `levels` times, level by level, then the
 * mesh is 2:1 balanced across faces, edges and vertices.  Output: cells[n][4] = (level, ix, iy,
 * iz) in Morton order.  With cells == NULL only *n_cells is set.  Replaces the refinement loop
 * of the mp-heat-transfer application (applications/mp-heat-transfer/
 * heat_transfer_application.cpp:484-527, source/utilities/amr.cpp) for the benchmark mesh. */
int mf_mesh_refine_band(const int ncell[3], const double lo[3], const double hi[3], int levels,
                        const double center[3], double radius, double half_width, int32_t *cells,
                        int64_t capacity, int64_t *n_cells);
/* Where the adaptive-mesh setup runs: 1 (default) on the device when one is present (band
 * refinement and 2:1 balance, DoF numbering, support points, boundary flags, hanging-node
 * constraint lines, their transposition and the hanging-cell tables), with the same cells,
 * numbering and constraints as the host path; 0 everything on the host.  Chain resolution,
 * colouring and the cell lists run on the host either way.  No reference counterpart (deal.II
 * builds all of it on the host). */
int mf_set_adaptive_setup(int device);
/* Scratch on an adaptive mesh (cells as above, 2:1 balanced): DoF numbering in order of first
 * appearance over the cells, hanging-node constraints with chains resolved, and the default
 * Dirichlet set = boundary nodes that are not hanging (HNC merged into DBC with the hanging line
 * winning).  Replaces FiniteElementUtils::distribute_dofs (fe_util.hpp:21-81) and
 * Constraints::make_DBC_and_HNC_plus_PBC_and_merge_HNC_plus_PBC_into_DBC (constraints.hpp:
 * 117-151, 254-269).  k_cell fields are indexed in the order of `cells`. */
int mf_scratch_create_adaptive(int degree, const int ncell[3], const double lo[3],
                               const double hi[3], const int32_t *cells, int64_t n_cells,
                               mf_scratch **out);
int mf_scratch_is_adaptive(const mf_scratch *s);
/* support points [n_dofs][3] (host); DoFHandler support points (fe_util.hpp) */
int mf_scratch_support_points(const mf_scratch *s, double *xyz);
/* default Dirichlet rows (host, n_dofs bytes): all boundary nodes of a structured mesh,
 * non-hanging boundary nodes of an adaptive mesh */
int mf_scratch_default_dirichlet_mask(const mf_scratch *s, uint8_t *mask);
/* number of hanging nodes, constraint entries and cell colours of the scratch */
int mf_scratch_constraint_info(const mf_scratch *s, int64_t *n_hanging, int64_t *n_entries,
                               int *n_colours);
/* hanging-node constraint lines (host copies): node_line[n_dofs] (-1 = free), line_ptr
 * [n_hanging+1], line_col / line_val [n_entries] (AffineConstraints lines after close()) */
int mf_scratch_constraint_lines(const mf_scratch *s, int32_t *node_line, uint32_t *line_ptr,
                                uint32_t *line_col, double *line_val);
/* AffineConstraints::distribute on a device vector: x_h = sum_m w_hm x_m on hanging nodes */
int mf_distribute_constraints(mf_scratch *s, double *x);
/* Field transfer between two adaptive meshes of the same base mesh and degree (the remesh step
 * of AMR::refine_grid, source/utilities/amr.cpp:59-112, with dealii::SolutionTransfer for
 * FE_Q): every node of the new mesh takes the value of the old field (its hanging values
 * first distributed) at its support point, found in the old leaf that contains it; then the
 * new mesh's hanging constraints are distributed.  Exact for fields of the new space that the
 * old mesh represents (refinement); interpolation from the finer old cells on coarsening.
 * old_vec and new_vec are device vectors of the two meshes' sizes. */
int mf_adaptive_transfer(mf_scratch *old_s, const double *old_vec, mf_scratch *new_s,
                         double *new_vec);

/* ------------------------------------------------------------------------------------------ */
/* Operators.  Replaces OperatorMatrixFree<dim,double> (include/meltpooldg/core/              */
/* operator_base.hpp:189-283) for HeatDiffuseMultiPhaseOperator (source/heat/                   */
/* heat_diffuse_operator.cpp) and AdvectionDiffusionOperator (source/level_set/                 */
/* advection_diffusion_operator.cpp).                                                           */
/* ------------------------------------------------------------------------------------------ */
enum { MF_LAW_CONSTANT = 0, MF_LAW_TANH = 1, MF_LAW_MUSHY = 2 };

typedef struct {
  int law;            /* MF_LAW_CONSTANT, MF_LAW_TANH or MF_LAW_MUSHY */
  int t_dependent;    /* include the dk/dT and d(rho cp)/dT Jacobian terms (reference
                       * do_solidification) */
  double rho_cp;      /* volumetric heat capacity (constant and tanh laws) */
  double k;           /* constant conductivity, or base value of the tanh law */
  double tanh_amp;    /* k(T) = k + tanh_amp * tanh((T - tanh_Tc) / tanh_width) */
  double tanh_Tc;
  double tanh_width;
  /* MF_LAW_MUSHY: MaterialTypes::liquid_solid with the mushy_zone transition
   * (include/meltpooldg/core/material.templates.hpp:296-365,546-600,748-780): solid fraction
   * f = clamp((liquidus - T) / (liquidus - solidus), 0, 1); conductivity, density and specific
   * heat each l + (s - l)(3 f^2 - 2 f^3) (level_set_tools.hpp:72-89), rho cp = rho * cp
   * (get_material_parameters_and_derivatives, heat_diffuse_operator.cpp:1167-1210).  With
   * t_dependent the tangent adds d(rho cp)/dT ((T - T_old)/dt + grad T . w) u, which reads the
   * old temperature (mf_operator_set_temperature_old).  No latent heat (apparent capacity). */
  double liquid_k, liquid_rho, liquid_cp;
  double solid_k, solid_rho, solid_cp;
  double solidus, liquidus;
} mf_heat_params;

typedef struct {
  double theta;        /* 1 implicit Euler, 0.5 Crank-Nicolson */
  double diffusivity;
  double mass_factor;  /* 1 for theta schemes, 1.5 for BDF2 */
  /* ConvectionStabilizationData (include/meltpooldg/level_set/advection_diffusion_data.hpp:
   * 34-38): supg = 1 selects ConvectionStabilizationType::SUPG
   * (advection_diffusion_operator.cpp:455-530).  The tangent's gradient term gains
   * tau w (mass_factor/dt u + theta w . grad u - theta D lap u), create_rhs's
   * tau w (u_old/dt - (1-theta) w . grad u_old + (1-theta) D lap u_old), with per cell
   * tau = (diameter / sqrt(3)) * (supg_coefficient < 0 ? (max_q |w| >= 1e-6 ? 1 / max_q |w| : 0)
   * : supg_coefficient).  The reference's default coefficient is -1.  The transposed operator
   * (mf_operator_set_transpose) is not available with SUPG. */
  int supg;
  double supg_coefficient;
} mf_advec_params;

typedef struct mf_operator mf_operator;

int mf_heat_operator_create(mf_scratch *s, const mf_heat_params *p, mf_operator **out);
int mf_advection_operator_create(mf_scratch *s, const mf_advec_params *p, mf_operator **out);
int mf_operator_destroy(mf_operator *op);

/* OperatorBase::reset_time_increment (operator_base.hpp:31-46) */
int mf_operator_reset_time_increment(mf_operator *op, double dt);
/* OperatorBase::reset_dof_index: 0 = constrained index (Dirichlet rows skipped, entries read as
 * zero), 1 = no-BC index (used for the Dirichlet lifting, utilities_matrixfree.hpp:33-100) */
int mf_operator_reset_dof_index(mf_operator *op, int dof_index);
/* Linearisation / coefficient fields (device pointers, n_dofs doubles each, kept by reference
 * like the reference operator keeps const references to its vectors; NULL = absent):
 * T = temperature (heat) , vel = nodal velocity, component-major [3][n_dofs],
 * k_cell = per-cell conductivity factor [n_cells].  Replaces the vectors wired in at
 * heat_diffuse_operation.cpp:55-70 and the pre() ghost update (heat_diffuse_operator.cpp:137-157).
 * Must be called again whenever the content of T changes (the FP32 copy is refreshed here). */
int mf_operator_set_fields(mf_operator *op, const double *T, const double *vel,
                           const double *k_cell);

/* The old temperature of the heat operator (HeatDiffuseMultiPhaseOperator::temperature_old,
 * wired at heat_diffuse_operation.cpp:55-70; read plain in tangent_local_cell_operation,
 * heat_diffuse_operator.cpp:1004-1009): needed by the mushy-zone law's d(rho cp)/dT term.
 * Device pointer of n_dofs doubles kept by reference; call again after its content changes.
 * mf_heat_newton_step sets it to its T_old argument. */
int mf_operator_set_temperature_old(mf_operator *op, const double *T_old);

/* Robin / Neumann boundary conditions of the heat operator on the sides of a structured box mesh:
 * side s = 2*direction + (0 lower, 1 upper), the boundary ids of a colourised
 * hyper_rectangle.  Convection alpha (T - T_inf) and radiation eps sigma (T^4 - T_inf^4)
 * (HeatData::convection / ::radiation, include/meltpooldg/heat/heat_data.hpp:46-56) enter the
 * tangent (alpha u + 4 eps sigma T^3 u, tangent_boundary_loop, heat_diffuse_operator.cpp:243-269,
 * 1088-1129), its diagonal and create_rhs (rhs_boundary_loop, :719-781); a Neumann side adds
 * -(q.n) to create_rhs, here with a constant normal flux per side.  A side may not carry Neumann
 * together with radiation or convection (the reference's AssertThrow at :743-746). */
typedef struct {
  unsigned convection_sides; /* bit s: side s has convection */
  unsigned radiation_sides;
  unsigned neumann_sides;
  double convection_coefficient;
  double convection_temperature_infinity;
  double emissivity;
  double radiation_temperature_infinity;
  double neumann_flux[6]; /* q.n on each Neumann side */
} mf_heat_boundary;

/* NULL removes all boundary terms.  Heat operators on structured meshes only. */
int mf_operator_set_boundary(mf_operator *op, const mf_heat_boundary *b);

/* OperatorMatrixFree::vmult (operator_base.hpp:239-243; heat_diffuse_operator.cpp:179-197,
 * advection_diffusion_operator.cpp:236-272): dst = A src, dst zeroed first, constrained rows 0 */
int mf_operator_vmult(mf_operator *op, double *dst, const double *src);
/* dst += A src without zeroing dst first (deal.II MatrixFree::loop with zero_dst = false, the
 * mode the reference uses for create_rhs, heat_diffuse_operator.cpp:801) */
int mf_operator_vmult_add(mf_operator *op, double *dst, const double *src);
/* FP32 instance of the same operator (no reference counterpart: new precision form) */
int mf_operator_vmult_f32(mf_operator *op, float *dst, const float *src);
/* OperatorMatrixFree::compute_inverse_diagonal_from_matrixfree (operator_base.hpp:268-272;
 * heat_diffuse_operator.cpp:271-284, advection_diffusion_operator.cpp:399-446) */
int mf_operator_compute_inverse_diagonal(mf_operator *op, double *inv_diag);
/* The same inverse diagonal without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The reduction partials of the relative threshold live in a static
 * 8 KiB device array per GPU (1024 doubles, the maximum; the need does not grow with the
 * vector length), looked up on the first call on each device with that device current.
 * Calls that overlap on one GPU race on that array. */
int mf_operator_compute_inverse_diagonal_noalloc(mf_operator *op, double *inv_diag);
/* OperatorMatrixFree::create_rhs (operator_base.hpp:208-212; heat_diffuse_operator.cpp:783-806,
 * advection_diffusion_operator.cpp:274-358): ADDS into dst.  Heat: src = T_old, q_s = heat
 * source (may be NULL), result -R(T); on an adaptive mesh T, src and q_s are read as stored
 * (pass constraint-consistent vectors) and the hanging rows of dst are moved to their masters.
 * Advection: src = old field, q_s must be NULL. */
int mf_operator_create_rhs(mf_operator *op, double *dst, const double *src, const double *q_s);

/* Inflow/outflow boundary of the advection operator (AdvectionDiffusionOperator::
 * set_inflow_outflow_bc and do_pre_vmult / do_post_vmult, advection_diffusion_operator.cpp:
 * 529-578; the DoF set of AdvectionDiffusionOperation::create_inflow_outflow_constraints,
 * advection_diffusion_operation.cpp:368-499).  mask_host[n_dofs]: 1 on the inflow/outflow DoFs
 * (NULL removes them).  With the treatment on (mf_operator_set_inflow_outflow_pre_post, the
 * default), vmult reads their src entries as 0 and returns dst = src on their rows, the inverse
 * diagonal is 1 there, and mf_advection_step lifts their values from dirichlet_values like
 * Dirichlet values and sets them after the solve. */
int mf_advection_set_inflow_outflow(mf_operator *op, const uint8_t *mask_host);
/* set_apply_inflow_outflow_constraints_pre_post_matrix_free (advection_diffusion_operator.hpp) */
int mf_operator_set_inflow_outflow_pre_post(mf_operator *op, int enable);
/* The inflow DoFs of a structured box mesh: boundary nodes on the sides in `sides` (bit
 * 2*direction + upper, the colourised boundary ids) whose nodal velocity (device, [3][n_dofs])
 * has n . u <= 0 against the side's outer normal n; mask_host[n_dofs] gets 1 there, 0 elsewhere. */
int mf_inflow_outflow_mask(mf_scratch *s, const double *vel, unsigned sides, uint8_t *mask_host);

/* ------------------------------------------------------------------------------------------ */
/* Level-set normal vector and curvature projections.  Replaces NormalVectorOperator           */
/* (source/level_set/normal_vector_operator.cpp:117-366) and CurvatureOperator                  */
/* (source/level_set/curvature_operator.cpp:160-371) in their matrix-free form.  Tangent:       */
/* a(u,v) = sum_q m_q (u v + eta grad u . grad v) JxW with eta = filter_parameter *             */
/* (cell diameter / sqrt(3))^2 per cell (compute_cell_size_dependent_filter_parameter_mf,       */
/* normal_vector_operator.hpp:95-112) and the narrow-band indicator m_q = |phi_q| < threshold   */
/* (Tools::compute_mask_narrow_band, level_set_tools.hpp:359-368), else 1.  The normal vector  */
/* operator acts on block vectors [3][n_dofs] (BlockVectorType), the same scalar operator on    */
/* every block; the curvature operator on [n_dofs].  No Dirichlet rows (the normal and          */
/* curvature DoF indices carry hanging-node constraints only).  vmult, vmult_f32,               */
/* compute_inverse_diagonal (|d| > 1e-14 max(1, |d|_inf) ? 1/d : 1), create_rhs, mf_solve and   */
/* the Jacobi object take these operators like the others; no time increment is needed.        */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  double filter_parameter;        /* NormalVectorData / CurvatureData::filter_parameter (2.0) */
  int narrow_band;                /* narrow_band.enable (0) */
  double narrow_band_threshold;   /* narrow_band.level_set_threshold (0.9999999) */
  int compute_normalized_vector;  /* normal vector rhs submits grad phi / |grad phi| (0) */
} mf_levelset_projection_params;

/* NormalVectorOperator: create_rhs(dst [3][n_dofs], src = level set [n_dofs]) adds
 * sum_q v m_q grad phi JxW (normalised when asked, normalize<dim>(.., 1e-16),
 * normal_vector_operator.cpp:158-237) */
int mf_normal_vector_operator_create(mf_scratch *s, const mf_levelset_projection_params *p,
                                     mf_operator **out);
/* CurvatureOperator: create_rhs(dst [n_dofs], src = normal vector [3][n_dofs]) adds
 * -sum_q v m_q div(n_hat) JxW, n_hat = the nodal normal normalised with the tolerance of
 * compute_numerical_zero_of_norm of the domain volume (curvature_operator.cpp:185-255) */
int mf_curvature_operator_create(mf_scratch *s, const mf_levelset_projection_params *p,
                                 mf_operator **out);
/* the level set the narrow band reads (device pointer [n_dofs], kept by reference; NULL =
 * none; required when narrow_band is set).  Call again when its content changes. */
int mf_operator_set_level_set(mf_operator *op, const double *level_set);
/* entries of the operator's vectors: 3 n_dofs for the normal vector operator, else n_dofs */
int64_t mf_operator_n_rows(const mf_operator *op);
/* the curvature operator's nodal normalisation tolerance (UtilityFunctions::
 * compute_numerical_zero_of_norm, utility_functions.hpp:133-143) */
double mf_operator_normal_tolerance(const mf_operator *op);

/* ------------------------------------------------------------------------------------------ */
/* Olsson (hyperbolic, continuous Galerkin) reinitialization.  Replaces OlssonOperator          */
/* (source/level_set/olsson_CG_operator.cpp) in its matrix-free form without interface         */
/* constraints and wetting faces: tangent a(u,v) = sum_q [u v / dt + (D grad u) . grad v] JxW,  */
/* D = eps_t I + (eps_n - eps_t) n n^T, n = the interpolated normal vector normalised with the  */
/* tolerance of compute_numerical_zero_of_norm; create_rhs(dst, psi) adds                       */
/* sum_q JxW [(1 - psi^2)/2 n - eps_n (grad psi . n) n - eps_t (grad psi - (grad psi . n) n)] . */
/* grad v (the reference zeroes dst first).  No Dirichlet rows.                                 */
/* ------------------------------------------------------------------------------------------ */
enum { MF_THICKNESS_PROPORTIONAL = 0, MF_THICKNESS_ABSOLUTE = 1, MF_THICKNESS_CELLS = 2 };

typedef struct {
  int thickness_type;                 /* interface_thickness_parameter.type: eps_n = value * h,
                                         value, or value * h / 6 (h = cell diameter / sqrt(3)) */
  double thickness_value;             /* interface_thickness_parameter.value (0.5) */
  double tangential_diffusion_factor; /* hyperbolic.cg.tangential_diffusion_factor (0): eps_t */
} mf_reinit_params;

int mf_reinit_operator_create(mf_scratch *s, const mf_reinit_params *p, mf_operator **out);
/* the normal vector [3][n_dofs] (device pointer, kept by reference); call again after it
 * changes (its FP32 copy is refreshed here) */
int mf_operator_set_normal_vector(mf_operator *op, const double *normal);
/* eps_n of the smallest cell: the pseudo time step of the reference is
 * pseudo_time_step_factor * this value (level_set_operation.cpp:528-532) */
double mf_reinit_min_epsilon(const mf_operator *op);

/* ------------------------------------------------------------------------------------------ */
/* Formats around the path.                                                                     */
/* ------------------------------------------------------------------------------------------ */
/* VTU output of device-resident fields (replaces GenericDataOut / Postprocessor with DataOut::  */
/* write_vtu, source/post_processing/generic_data_out.cpp, postprocessor.cpp): one point per     */
/* DoF (its support point), every cell written as degree^3 trilinear hexahedra on its nodes,    */
/* appended raw binary.  fields[f]: device pointer, n_components[f] = 1 ([n_dofs]) or 3          */
/* ([3][n_dofs], written as a vector). */
int mf_write_vtu(const mf_scratch *s, const char *filename, int n_fields, const char *const *names,
                 const int *n_components, const double *const *fields);
/* Restart (replaces Restart::serialize_internal / deserialize_internal, source/utilities/       */
/* restart.cpp:115-186): one file with the mesh description (box, degree, adaptive cells, the   */
/* Dirichlet mask) and the given device vectors.  Loading rebuilds the scratch, and with it the  */
/* same DoF numbering, then reads the vectors back (lengths must match). */
int mf_restart_save(const mf_scratch *s, const char *path, int n_vectors,
                    const double *const *vectors, const int64_t *lengths);
int mf_restart_load_scratch(const char *path, mf_scratch **out);
int mf_restart_load_vectors(const mf_scratch *s, const char *path, int n_vectors,
                            double *const *vectors, const int64_t *lengths);

/* ------------------------------------------------------------------------------------------ */
/* Solvers.  Replaces LinearSolver::solve (include/meltpooldg/linear_algebra/                   */
/* linear_solver.hpp:22-134) with JacobiPreconditioner (preconditioner_jacobi.hpp:29-89):       */
/* ReductionControl(max_it, abs_tol, rel_tol), success when |r| <= max(abs_tol, rel_tol |r_0|). */
/* ------------------------------------------------------------------------------------------ */
enum { MF_SOLVER_CG = 0, MF_SOLVER_GMRES = 1 };
enum { MF_PRECOND_IDENTITY = 0, MF_PRECOND_JACOBI = 1 };
enum { MF_PRECISION_FP64 = 0, MF_PRECISION_FP32_REFINED = 1 };

typedef struct {
  int solver;         /* MF_SOLVER_CG or MF_SOLVER_GMRES (right preconditioned) */
  int restart;        /* GMRES basis size, 1..30 (0 = 30; larger values: MF_ERR_ARG) */
  int max_iterations; /* 0 = 10000 */
  double abs_tol;
  double rel_tol;
  int preconditioner; /* MF_PRECOND_* */
  int precision;      /* MF_PRECISION_FP64 or MF_PRECISION_FP32_REFINED */
  double inner_rel_tol; /* FP32-refined: relative tolerance of each inner FP32 solve (0 = auto:
                           sqrt(rel_tol) clamped to [1e-5, 1e-2], 2 passes for 1e-10) */
  int max_outer;      /* FP32-refined: maximum defect-correction passes (0 = 50) */
  int force_re_orthogonalization; /* GMRES: second Gram-Schmidt pass for every basis vector
                           (deal.II SolverGMRES::AdditionalData::force_re_orthogonalization; the
                           reference keeps the default, 0: a second pass only once a pass has
                           cancelled the vector below 10 sqrt(eps) of its norm,
                           linear_solver.hpp:112-114) */
} mf_solver_desc;

typedef struct {
  int iterations;        /* FP64: Krylov iterations; refined: total inner iterations */
  int outer_iterations;  /* refined: defect-correction passes (FP64: 0) */
  double initial_residual;
  double final_residual; /* |b - A x| in FP64 */
  int status;            /* MF_OK / MF_NO_CONVERGENCE / MF_BREAKDOWN */
} mf_solve_report;

/* Solve A x = b with x holding the initial guess (device pointers).  The Jacobi diagonal is
 * recomputed from the operator (Preconditioner::update) when precond = JACOBI. */
int mf_solve(mf_operator *op, const mf_solver_desc *desc, double *x, const double *b,
             mf_solve_report *report);

/* JacobiPreconditioner (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:29-89):
 * create = constructor + reinit (device storage for the inverse diagonal), update = update()
 * (compute_inverse_diagonal_from_matrixfree of the operator it was created for), vmult =
 * vmult() (dst = D^-1 src).  The object keeps an FP32 copy for the refined solver.  Not
 * re-entrant; it must be updated again when the operator's fields, dt or constraints change. */
typedef struct mf_jacobi mf_jacobi;
int mf_jacobi_create(mf_operator *op, mf_jacobi **out);
int mf_jacobi_update(mf_jacobi *pc);
int mf_jacobi_vmult(mf_jacobi *pc, double *dst, const double *src);
/* copy of the current inverse diagonal (device pointer dst of n_dofs entries) */
int mf_jacobi_inverse_diagonal(mf_jacobi *pc, double *dst);
int mf_jacobi_destroy(mf_jacobi *pc);

/* LinearSolver::solve(op, x, b, data, preconditioner) (linear_solver.hpp:22-134) with a
 * preconditioner updated beforehand (pc = NULL: identity); desc->preconditioner is ignored. */
int mf_solve_preconditioned(mf_operator *op, const mf_solver_desc *desc, mf_jacobi *pc,
                            double *x, const double *b, mf_solve_report *report);

/* ------------------------------------------------------------------------------------------ */
/* Callers on either side of the operator                                                      */
/* ------------------------------------------------------------------------------------------ */
/* NonlinearSolverData (include/meltpooldg/linear_algebra/nonlinear_solver_data.hpp:13-20).
 * Every field is used as given (a tolerance of 0 cannot be met); mf_newton_desc_init fills the
 * reference defaults (10, 1e-9, 1e-10, 0, 1e-8, 1e-9, relative off).  Convergence: |r| < tol_r
 * and |d| < tol_d, checked before d is added; with use_relative_tolerances |r| is divided by
 * the first iteration's |r| and |d| by the current |x| (each only when that reference norm is
 * positive).  After max_iterations the *_alt tolerances apply for max_iterations_alt more. */
typedef struct {
  int max_iterations;              /* max_nonlinear_iterations */
  double residual_tolerance;
  double correction_tolerance;     /* field_correction_tolerance */
  int max_iterations_alt;          /* max_nonlinear_iterations_alt */
  double residual_tolerance_alt;
  double correction_tolerance_alt; /* field_correction_tolerance_alt */
  int use_relative_tolerances;
} mf_newton_desc;

int mf_newton_desc_init(mf_newton_desc *d);

typedef struct {
  int iterations;          /* Newton iterations (linear solves) */
  int linear_iterations;   /* accumulated Krylov iterations */
  double residual_norm;    /* |r| of the last iteration */
  double correction_norm;  /* |d| of the last iteration */
  int status;
} mf_newton_report;

/* The Newton-Raphson driver itself (NewtonRaphsonSolver::solve, source/linear_algebra/
 * newton_raphson_solver.cpp:22-175), on the host, over the caller's state through callbacks
 * (the reference's std::function members residual / solve_with_jacobian /
 * distribute_constraints / norm_of_solution_vector, newton_raphson_solver.hpp).  Each callback
 * returns an MF_* status; a non-zero status ends the solve with that status.
 *   residual(user): rhs = -R(x) at the current iterate
 *   solve_with_jacobian(user, &linear_iterations): d = J^-1 rhs
 *   norms(user, &|rhs|, &|d|, &|x|): l2 norms (|x| only needed when need_x is non-zero)
 *   update(user): x += d, then distribute the constraints
 * The report's residual_norm / correction_norm are the last absolute norms. */
typedef struct {
  int (*residual)(void *user);
  int (*solve_with_jacobian)(void *user, int *linear_iterations);
  int (*norms)(void *user, int need_x, double *residual_norm, double *update_norm,
               double *solution_norm);
  int (*update)(void *user);
} mf_newton_callbacks;

int mf_newton_solve(const mf_newton_desc *desc, const mf_newton_callbacks *cb, void *user,
                    mf_newton_report *report);

/* One implicit heat step: NewtonRaphsonSolver::solve (source/linear_algebra/
 * newton_raphson_solver.cpp:22-80) with the residual / Jacobian wiring of HeatDiffuseOperation
 * (source/heat/heat_diffuse_operation.cpp:86-116, 261-289).  T: predictor in, solution out (the
 * operator's linearisation field is set to T); T_old: previous step; q_s: heat source (NULL =
 * none); dirichlet_values: values of constrained DoFs (NULL = homogeneous).
 * Failure: MF_NO_CONVERGENCE (reference ExcNewtonDidNotConverge). */
int mf_heat_newton_step(mf_operator *op, double *T, const double *T_old, const double *q_s,
                        const double *dirichlet_values, const mf_solver_desc *desc,
                        const mf_newton_desc *newton, mf_newton_report *report);
/* One advection-diffusion step: AdvectionDiffusionOperation::solve (source/level_set/
 * advection_diffusion_operation.cpp:227-367) with the Dirichlet lifting of
 * create_rhs_and_apply_dirichlet_matrixfree (include/meltpooldg/linear_algebra/
 * utilities_matrixfree.hpp:33-100).  phi: initial guess in, solution out; phi_old: field the
 * right-hand side is built from (the old solution, or the BDF2 combination 2 phi_n - phi_{n-1}/2
 * with mass_factor 1.5 and theta 1); dirichlet_values: NULL = homogeneous.  On an adaptive
 * mesh phi_old must be constraint-consistent; phi comes back with its hanging values
 * distributed. */
int mf_advection_step(mf_operator *op, double *phi, const double *phi_old,
                      const double *dirichlet_values, const mf_solver_desc *desc,
                      mf_solve_report *report);
/* Predictor::vmult with linear extrapolation (include/meltpooldg/linear_algebra/
 * predictor.hpp:88-161, source/linear_algebra/predictor_linear.cpp:74-96):
 * dst = (1 + tau) x_n - tau x_{n-1}, tau = dt_n / dt_{n-1}. */
int mf_predictor_linear_extrapolation(mf_scratch *s, double *dst, const double *x_n,
                                      const double *x_nm1, double tau);
/* LaserHeatSourceVolumetric::compute_volumetric_heat_source with the Gaussian volumetric
 * profile (source/heat/laser_heat_source_volumetric.cpp:20-58,
 * include/meltpooldg/heat/laser_intensity_profiles.hpp:175-218): nodal values of
 * P / (r sqrt(pi/2))^3 exp(-2 |x - x_L|^2 / r^2); on an adaptive mesh at the support points,
 * then the hanging values are distributed (as the reference's constraints.distribute). */
int mf_laser_heat_source(mf_scratch *s, double power, double radius, const double position[3],
                         double *q_s);
/* VectorTools::compute_norm, FE L2 norm (source/utilities/vector_tools.cpp:137-201) with the
 * per-cell values rounded to single precision as in the reference; result on the host.  On an
 * adaptive mesh u is read as stored (pass a constraint-consistent vector). */
int mf_compute_norm_l2(mf_scratch *s, const double *u, double *result);

/* Derivative of the Gaussian volumetric source with respect to the radius (adjoint). */
int mf_laser_heat_source_radius_derivative(mf_scratch *s, double power, double radius,
                                           const double position[3], double *dq_dr);

/* LeastSquaresProjection::vmult (include/meltpooldg/linear_algebra/predictor.hpp:17-81): the
 * combination sum_s c_s x_s of the old solutions that minimises |A sum_s c_s x_s - rhs|, with
 * A = the operator's vmult (constraints as in mf_operator_vmult).  The images A x_s are
 * orthogonalised by modified Gram-Schmidt (Vector::add_and_dot), stopping at the first s whose
 * squared residual norm falls below 1e-16 times |A x_0|^2 (that x_s and later ones are not
 * used); the triangular system is solved by back substitution.  old_solutions: n_old (1..8)
 * device vectors of mf_operator_n_rows entries; n_used (may be NULL) returns how many entered
 * the combination. */
int mf_predictor_least_squares(mf_operator *op, int n_old, const double *const *old_solutions,
                               const double *rhs, double *solution, int *n_used);

/* PredictorType (include/meltpooldg/linear_algebra/predictor_data.hpp:10-19). */
enum { MF_PREDICTOR_NONE = 0, MF_PREDICTOR_ZERO = 1, MF_PREDICTOR_LINEAR_EXTRAPOLATION = 2,
       MF_PREDICTOR_LEAST_SQUARES_PROJECTION = 3 };

/* One heat time step of the mp-heat-transfer application on the device: the loop body of
 * HeatTransferApplication::run (applications/mp-heat-transfer/heat_transfer_application.cpp:
 * 52-120) for the heat-diffuse operator with a Gaussian volumetric laser.  A step advances the
 * time (TimeIterator::compute_next_time_increment, source/time_integration/time_iterator.cpp:
 * 44-68), moves the laser by scan_speed * dt along the normalised scan direction
 * (LaserOperation::move_laser, source/heat/laser_operation.cpp:236-250), sets the heat source
 * to the volumetric profile at the new position (compute_heat_source, :316-333), resets the
 * operator's time increment, applies the predictor and commits the history (Predictor::vmult,
 * include/meltpooldg/linear_algebra/predictor.hpp:119-161) and runs the Newton solve
 * (HeatDiffuseOperation::solve, source/heat/heat_diffuse_operation.cpp:214-290).  The linear
 * predictor uses the reference's increments: the old increment is 0 at the first step (the
 * predictor is then the current solution) and equal to the new one afterwards. */
typedef struct {
  int laser;                 /* 1: Gaussian volumetric laser source, 0: no heat source */
  double power;              /* LaserData::power */
  double radius;             /* LaserData::radius */
  double start[3];           /* LaserData::starting_position */
  double scan_direction[3];  /* LaserData::scan_direction (normalised by the library) */
  double scan_speed;         /* LaserData::scan_speed */
  double start_time;         /* TimeSteppingData::start_time */
  int predictor;             /* MF_PREDICTOR_* (HeatData default: linear extrapolation) */
  mf_solver_desc solver;     /* linear solver of each Newton iteration */
  mf_newton_desc newton;
} mf_heat_stepper_desc;

typedef struct mf_heat_stepper mf_heat_stepper;

/* The stepper keeps the current and recent old temperature and the heat source on the device;
 * T0 (device, n_dofs) is copied.  The operator's linearisation field points to the stepper's
 * current temperature after each step. */
int mf_heat_stepper_create(mf_operator *op, const mf_heat_stepper_desc *d, const double *T0,
                           mf_heat_stepper **out);
int mf_heat_stepper_destroy(mf_heat_stepper *h);
/* One step of size dt; dirichlet_values: values of the constrained DoFs (NULL = homogeneous).
 * Failure: MF_NO_CONVERGENCE (reference ExcHeatTransferNoConvergence). */
int mf_heat_stepper_step(mf_heat_stepper *h, double dt, const double *dirichlet_values,
                         mf_newton_report *report);
/* Device pointers to the stepper's vectors (valid until the next step or destroy), the current
 * time and the laser position; any output may be NULL. */
int mf_heat_stepper_state(const mf_heat_stepper *h, const double **T, const double **T_old,
                          const double **heat_source, double *time, double position[3]);
/* Copies the current temperature (which = 0), the recent old temperature (1) or the heat
 * source (2) into dst (device, n_dofs) on the scratch's stream. */
int mf_heat_stepper_copy(const mf_heat_stepper *h, int which, double *dst);

/* ------------------------------------------------------------------------------------------ */
/* Differentiation of the implicit heat transient (no reference counterpart: new).             */
/* ------------------------------------------------------------------------------------------ */
/* vmult / solves apply A^T instead of A when transpose = 1 (a distinct device code path for
 * the non-symmetric convection and dk/dT terms). The diagonal is the same for both. */
int mf_operator_set_transpose(mf_operator *op, int transpose);
/* dst = M src (mass matrix; src read without constraints, as the heat residual reads its
 * source term; Dirichlet rows of dst zero). */
int mf_mass_vmult(mf_scratch *s, double *dst, const double *src);
/* result = a . b (deterministic device reduction; result on the host). */
int mf_vector_dot(mf_scratch *s, const double *a, const double *b, double *result);
/* grad_k_cell[c] += scale * sum_q JxW grad(lambda) . grad(T) over cell c (unit conductivity):
 * the parameter-gradient kernel for per-cell conductivity factors. */
int mf_heat_conductivity_gradient(mf_operator *op, const double *lambda, const double *T,
                                  double scale, double *grad_k_cell);

/* Moving-laser implicit-Euler transient, T_n = A^{-1}((rho cp/dt) M T_{n-1} + M q_n), laser at
 * start + scan_velocity * n dt, with checkpointed adjoint (states stored every
 * checkpoint_interval steps, recomputed segment-wise in the backward sweep). */
typedef struct {
  int n_steps;
  int checkpoint_interval;
  double dt;
  double power;
  double radius;
  double start[3];
  double scan_velocity[3];
  mf_solver_desc solver;   /* for the forward and the adjoint solves */
} mf_transient_desc;

typedef struct mf_transient mf_transient;

int mf_transient_create(mf_operator *op, const mf_transient_desc *d, mf_transient **out);
int mf_transient_destroy(mf_transient *tr);
/* T_final = T_N from T_0 (device vectors) */
int mf_transient_forward(mf_transient *tr, const double *T0, double *T_final);
/* Given dJ/dT_N (dual vector, e.g. M (T_N - T_target)): accumulates dJ/dk_cell into
 * grad_k_cell[n_cells] (may be NULL) and returns dJ/dP, dJ/dr on the host. */
int mf_transient_backward(mf_transient *tr, const double *dJ_dTN, double *grad_k_cell,
                          double *grad_power, double *grad_radius);

/* ------------------------------------------------------------------------------------------ */
/* One-phase CutFEM heat operator.  Replaces HeatCutOperator without two_phase                 */
/* (source/heat/heat_cut_operator.cpp) and the matching parts of HeatCutOperation              */
/* (source/heat/heat_cut_operation.cpp): FE_Q1 only (the reference's AssertThrow at :498) on a */
/* structured box mesh.  Tangent (tangent_cell_loop + tangent_inner_face_loop, :587-607,      */
/* 1028-1388): sum over the liquid part of each cell of                                        */
/*   (cv + dcv/dT T) u v + dt theta (k grad u + dk/dT (grad T) u) . grad v                    */
/* (tangent_domain_integral, :157-190) plus the ghost penalty on the faces between an          */
/* intersected cell and an intersected or liquid cell                                          */
/*   gamma [d_n u][d_n v],  gamma = dt theta gamma_A k_l h / 3 + gamma_M h^3 / 3               */
/* (ghost_penalty_face_integral, :425-450) over the whole face (QGauss(2)).                    */
/* Cells are classified from the nodal level set (liquid where phi > 0; MeshClassifier with the */
/* CutUtil mapping, source/cut/util.cpp:100-128): phi >= 0 on the whole cell: liquid; phi < 0  */
/* on the whole cell: gas (FE_Nothing: its nodes that touch no other cell carry no DoF);       */
/* otherwise intersected.  The liquid part of an intersected cell must be a sub-box: phi must  */
/* vary along one axis only on that cell (axis-aligned planar interfaces, which is what the    */
/* reference's cut test cases use); the quadrature there is QGauss(2) per direction on the     */
/* sub-box, the rule dealii::NonMatching::QuadratureGenerator builds for such cells.  Other    */
/* interfaces give MF_ERR_ARG.  DoFs without a liquid or intersected cell are kept in the      */
/* library's node numbering as identity rows (vmult 0, diagonal 1, src read as 0).             */
/* No velocity, evaporative cooling or interface laser flux (HeatCutOperation's optional      */
/* terms), no FP32 form.                                                                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  /* law MF_LAW_CONSTANT: the liquid phase, rho_cp = rho * cp and k (do_solidification false);
   * MF_LAW_MUSHY: the liquid_solid law of mf_heat_params with its T-derivatives in the tangent
   * (do_solidification true).  t_dependent and the tanh fields are ignored. */
  mf_heat_params material;
  double theta;              /* HeatData::cut.theta (include/meltpooldg/heat/heat_data.hpp:34),
                                0.5 in the reference */
  double gamma_A;            /* GhostPenaltyData::gamma_A_degree_1 (cut_data.hpp:12-32) */
  double gamma_M;            /* GhostPenaltyData::gamma_M_degree_1 */
  double cell_side_length;   /* h of the ghost penalty; <= 0: min cell diameter / sqrt(3)
                                (ScratchData::get_min_cell_size, scratch_data.cpp:170) */
} mf_heat_cut_params;

/* HeatCutOperator constructor + reinit (:454-559).  The scratch must be a structured FE_Q1
 * mesh.  The operator takes mf_operator_reset_time_increment (init_time_advance: the ost
 * factors dt theta and dt (1 - theta)), mf_operator_set_fields (T: the linearisation of the
 * mushy law and the current temperature of create_rhs), mf_operator_vmult / vmult_add,
 * mf_operator_compute_inverse_diagonal (compute_inverse_diagonal_from_matrixfree, :2141-2153,
 * 1/d where |d| > 1e-14 max(1, |d|_inf)), mf_operator_create_rhs (create_rhs, :609-628: src =
 * T_old, q_s must be NULL; adds -R(T)), mf_solve (FP64) and the Jacobi object. */
int mf_heat_cut_operator_create(mf_scratch *s, const mf_heat_cut_params *p, mf_operator **out);
/* The level set (host array of n_dofs nodal values on the scratch's Q1 nodes): classification
 * of the cells, liquid sub-boxes of the intersected cells, ghost-penalty faces and the DoFs
 * that exist (HeatCutOperation::classify_cells + distribute_dofs + compute_intersected_
 * quadrature, heat_cut_operation.cpp:172-330).  MF_ERR_ARG for an intersected cell whose
 * level set varies along more than one axis. */
int mf_heat_cut_set_level_set(mf_operator *op, const double *phi_host);
/* counts of the classification: liquid, intersected and gas cells, ghost-penalty faces and
 * DoFs that exist (pointers may be NULL) */
int mf_heat_cut_info(const mf_operator *op, int64_t *n_liquid, int64_t *n_intersected,
                     int64_t *n_gas, int64_t *n_faces, int64_t *n_active_dofs);
/* host copies: category[n_cells] (0 liquid, 1 intersected, 2 gas: CutUtil::CellCategory),
 * box[n_cells][3][2] (liquid part in reference coordinates), active[n_dofs] (any may be NULL) */
int mf_heat_cut_cells(const mf_operator *op, int8_t *category, double *box, uint8_t *active);
/* CutUtil::compute_cut_norm, L2 (source/cut/cut_norm.cpp:20-217): sqrt of the integral of u^2
 * over the liquid part of the domain (u device, read plain); result on the host */
int mf_heat_cut_norm_l2(mf_operator *op, const double *u, double *result);
/* One time step of HeatCutOperation::solve for a fixed interface (heat_cut_operation.cpp:
 * 431-482): NewtonRaphsonSolver::solve with r = create_rhs(T_old) at T, J d = r by the linear
 * solver, the constraints distributed after each update.  T: the predictor on entry (the
 * reference copies the old solution), the solution on exit; dirichlet_values: values of the
 * scratch's constrained DoFs (NULL = homogeneous). */
int mf_heat_cut_step(mf_operator *op, double *T, const double *T_old,
                     const double *dirichlet_values, const mf_solver_desc *desc,
                     const mf_newton_desc *newton, mf_newton_report *report);

#ifdef __cplusplus
}
#endif
#endif
