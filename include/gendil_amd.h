/*
 * gendil_amd.h - C-ABI of the MI355X (gfx950) matrix-free DG operator library.
 *
 * This is the drop-in boundary for GenDiL's matrix-free operator apply on
 * Cartesian hypercube meshes (mass + upwind advection, 1-6D).  Everything
 * crossing it is a plain pointer, size or POD descriptor; vectors are
 * device-resident (caller-owned device pointers), applies are stream-ordered
 * on the caller's hipStream_t (passed as gd_stream, NULL = default stream).
 *
 * Every entry point returns an int status (GD_OK = 0); gd_last_error() gives
 * a thread-local message for the last failure.  Nothing aborts across the ABI
 * (GenDiL itself aborts via GENDIL_VERIFY, include/gendil/Utilities/debug.hpp:59).
 *
 * References below are paths in the GenDiL tree (include/gendil/...).
 */
#ifndef GENDIL_AMD_H
#define GENDIL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GD_MAX_DIM 6
#define GD_MAX_TERMS 8

typedef struct gd_plan gd_plan; /* opaque operator plan */
typedef void* gd_stream;        /* hipStream_t */

enum gd_status {
  GD_OK = 0,
  GD_ERR_INVALID = 1,     /* bad argument */
  GD_ERR_UNSUPPORTED = 2, /* valid but not compiled in (dim/order/rule combo) */
  GD_ERR_HIP = 3,         /* HIP runtime error */
  GD_ERR_ALLOC = 4,       /* device allocation failed */
  GD_ERR_NOT_CONVERGED = 5
};

enum gd_basis { GD_BASIS_GLL = 0, GD_BASIS_GL = 1 };
enum gd_beta_kind { GD_BETA_SEPARABLE = 0, GD_BETA_POINTWISE = 1, GD_BETA_AFFINE = 2 };
enum gd_bc_mode {
  GD_BC_NONE = 0,  /* boundary faces skipped: MakeAdvectionOperator(fes, ir, adv) */
  GD_BC_INFLOW = 1 /* upwind with u+ = g(X): MakeAdvectionOperator(fes, ir, adv, g) */
};

/* Scalar field f(X) = sum_r scale[r] * prod_d T_{r,d}(X_d).  Each factor is a
 * 1D table sampled at the positions gd_plan_axis_points returns for axis d;
 * offset[r][d] indexes `data`, -1 means the factor is 1.  Replaces GenDiL's
 * device lambdas sigma(X), adv(X, v), g(X) (massoperator.hpp:270,
 * advectionoperator.hpp:1416-1528), which cannot cross a C-ABI. */
typedef struct gd_separable {
  int32_t rank;
  double scale[GD_MAX_TERMS];
  int64_t offset[GD_MAX_TERMS][GD_MAX_DIM];
  const double* data; /* host pointer, copied at set time */
  int64_t data_len;
} gd_separable;

/* Mesh + finite element + integration rule: CartesianMesh<D>(sizes, h, origin,
 * periodic) (Meshes/MeshDataStructures/CartesianMesh/cartesianndmesh.hpp:16-80),
 * MakeLobattoFiniteElement / MakeLegendreFiniteElement (finiteelement.hpp:103-109),
 * MakeIntegrationRule(IntegrationRuleNumPoints<q...>) (integrationrule.hpp:80). */
typedef struct gd_desc {
  int32_t dim;
  int64_t sizes[GD_MAX_DIM];
  double h[GD_MAX_DIM];
  double origin[GD_MAX_DIM];
  int32_t periodic[GD_MAX_DIM];
  int32_t order;    /* uniform polynomial order p (n = p+1 nodes per dim) */
  int32_t num_quad; /* uniform Gauss-Legendre points per dim */
  int32_t basis;    /* gd_basis */
  /* Optional p-adaptive description (host pointers, length prod(sizes)):
   * per-element order and rule; NULL = uniform.  No GenDiL counterpart: GenDiL
   * fixes orders at compile time per space (tensorshapefunctions.hpp:21). */
  const int8_t* elem_order;
  const int8_t* elem_quad;
} gd_desc;

/* ---- plan lifetime ---------------------------------------------------- */
/* Builds the 1D tables GenDiL builds at operator construction
 * (bilinearoperator.hpp:20-106, doftoquad.hpp:116-151) and uploads them. */
int gd_plan_create(const gd_desc* desc, gd_plan** plan);
int gd_plan_destroy(gd_plan* plan);
const char* gd_last_error(void);
int gd_version(void);

int64_t gd_plan_num_dofs(const gd_plan* plan);
int64_t gd_plan_num_elements(const gd_plan* plan);
/* Device pointer to the per-element DoF offsets (int64, E+1 entries) of a
 * p-adaptive plan; NULL for a uniform plan. */
const int64_t* gd_plan_dof_offsets(const gd_plan* plan);
/* Sample positions for separable tables on axis d: sizes[d]*(num_quad+2)
 * values; per element e_d: the num_quad rule points, then xi=0, then xi=1,
 * each X = (origin + h*e_d) + h*xi (hypercubecell.hpp:41-52). */
int gd_plan_axis_points(const gd_plan* plan, int32_t axis, double* out);
/* 1D rule on [0,1]: GL (basis=1) or GLL (basis=0) points and weights. */
int gd_rule_points(int32_t basis, int32_t npts, double* pts, double* wts);

/* ---- coefficients -------------------------------------------------------
 * Each setter first waits for all work on the device (hipDeviceSynchronize),
 * so applies queued earlier on any stream use the previous coefficients, then
 * rebuilds every device table the kernels read (synchronous copies).  Applies
 * queued after it use the new coefficients.  Setters are set-up calls: the
 * apply entry points themselves never allocate and never copy, and they are
 * ordered only on the stream they are given. */
/* mass density sigma(X) (massoperator.hpp:65-83) */
int gd_plan_set_sigma(gd_plan* plan, const gd_separable* sigma);
/* advection field beta(X), one separable per component (dim entries) */
int gd_plan_set_beta_separable(gd_plan* plan, const gd_separable* beta);
/* beta(X) = c + A X, A row-major dim x dim */
int gd_plan_set_beta_affine(gd_plan* plan, const double* A, const double* c);
/* per-point beta (device pointers, caller-owned, read at apply time; the
 * arrays must hold exactly these sizes, which the C-ABI cannot check):
 * cell [E][dim][q^dim]; face normal component for axis a:
 * [F_a][q^(dim-1)], F_a = prod_{d!=a} sizes[d] * (sizes[a] + !periodic[a]),
 * global face f_a between element f_a-1 (+ side) and f_a (- side). */
int gd_plan_set_beta_pointwise_f64(gd_plan* plan, const double* beta_cell,
                                   const double* const* beta_face);
/* inflow boundary: mode gd_bc_mode, g separable (may be NULL for g = 0) */
int gd_plan_set_bc(gd_plan* plan, int32_t mode, const gd_separable* g);
/* SIPDG diffusion coefficients (DiffusionOperator constructor and
 * MakeDiffusionOperator, diffusionoperator.hpp:441-455, 519-537):
 * diffusivity mu(X) separable (NULL = 1; GenDiL's `velocity` argument),
 * sigma the symmetry parameter (-1 symmetric, 0 incomplete, +1
 * non-symmetric interior penalty), kappa the penalty factor.  Uniform-order
 * plans only (GD_ERR_UNSUPPORTED on p-adaptive ones). */
int gd_plan_set_diffusion(gd_plan* plan, const gd_separable* mu, double sigma, double kappa);

/* ---- operator apply ---------------------------------------------------- */
/* y = mass_coef * M(sigma) x + adv_coef * A(beta, g) x, y overwritten (L2
 * output semantics, massoperator.hpp:209-216, advectionoperator.hpp:1070-1076).
 * x and y must not alias. */
int gd_apply_f64(const gd_plan* plan, double mass_coef, double adv_coef,
                 const double* x, double* y, gd_stream stream);
int gd_apply_f32(const gd_plan* plan, float mass_coef, float adv_coef,
                 const float* x, float* y, gd_stream stream);
/* MassFiniteElementOperator::operator() (massoperator.hpp:209) */
int gd_mass_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream);
int gd_mass_apply_f32(const gd_plan* plan, const float* x, float* y, gd_stream stream);
/* AdvectionOperator::operator() (advectionoperator.hpp:1070) */
int gd_adv_apply_f64(const gd_plan* plan, const double* u, double* y, gd_stream stream);
int gd_adv_apply_f32(const gd_plan* plan, const float* u, float* y, gd_stream stream);
/* y = A_diff x, overwritten (DiffusionOperator::operator(),
 * diffusionoperator.hpp:469-476; volume DiffusionElementOperator :44-101,
 * faces DiffusionFaceOperator :130-331): every element's volume term plus
 * the symmetric-interior-penalty terms of its 2*dim faces; faces on a
 * non-periodic boundary carry the weak homogeneous Dirichlet terms.  Needs
 * gd_plan_set_diffusion; x and y must not overlap.  Compiled (nodes,
 * points): 1D (2,3) (2,4) (3,4) (4,4) (4,5) (5,6); 2D (2,3) (2,4) (3,4)
 * (3,5) (4,4) (4,5) (4,6); 3D (2,3) (2,4) (3,3) (3,4) (3,5) (4,4) (4,5)
 * (4,6); 6D (2,3); others return GD_ERR_UNSUPPORTED. */
int gd_diff_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream);
int gd_diff_apply_f32(const gd_plan* plan, const float* x, float* y, gd_stream stream);
/* Adjoint of the linear part, z = (mass_coef M + adv_coef A)^T r.  No GenDiL
 * counterpart (GenDiL has no operator adjoint). */
int gd_apply_T_f64(const gd_plan* plan, double mass_coef, double adv_coef,
                   const double* r, double* z, gd_stream stream);
int gd_adv_apply_T_f64(const gd_plan* plan, const double* r, double* z, gd_stream stream);
/* dL/dbeta for L = <r, adv_coef * A(beta) u> with pointwise beta; writes the
 * cell and face arrays in the layout of gd_plan_set_beta_pointwise_f64.  No
 * GenDiL counterpart. */
int gd_adv_grad_beta_f64(const gd_plan* plan, double adv_coef, const double* u,
                         const double* r, double* dbeta_cell, double* const* dbeta_face,
                         gd_stream stream);

/* ---- face-centric advection (accumulating) ------------------------------ */
/* y += the upwind face terms of every interior face of the plan's Cartesian
 * mesh, each face computed once: + on the minus (upper) cell, - on the plus
 * (lower) cell.  No cell term, no boundary or periodic-wrap faces.  Replaces
 * AdvectionFaceOperator over CartesianInteriorFaceConnectivity
 * (advectionoperator.hpp:1105-1180, 478-548; cartesianglobalfaceconnectivity.hpp:
 * 15-95), built by MakeAdvectionFaceOperator(fes, face_mesh, ir, adv)
 * (:1546-1565).  Uniform plans, separable or affine beta; x and y must not
 * overlap. */
int gd_adv_face_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream);
/* Face `face` (0..2*dim-1, GenDiL's local face numbering) of the minus plan's
 * mesh glued to the opposite face of the plus plan's mesh: y_minus += and
 * y_plus -= the upwind fluxes on the interface, both sides interpolated to
 * the shared rule (both plans must have the same num_quad; their orders may
 * differ), geometry and beta from the minus plan.  No input side may share
 * storage with an output side (GD_ERR_INVALID).  Replaces
 * AdvectionNonconformingFaceOperator over CartesianIntermeshFaceConnectivity
 * <dim, face> (advectionoperator.hpp:578-676, 1210-1396; cartesianintermeshface
 * connectivity.hpp:36-180), built by MakeAdvectionFaceOperator(fes_minus,
 * fes_plus, face_mesh, ir, adv) (:1586-1607).  The four vectors may be parts
 * of one vector (GenDiL's L2Restriction shift). */
int gd_adv_intermesh_apply_f64(const gd_plan* minus, const gd_plan* plus, int32_t face,
                               const double* x_minus, const double* x_plus, double* y_minus,
                               double* y_plus, gd_stream stream);
/* Face `face` of a coarse mesh (minus) glued to a fine mesh (plus) with an
 * integer number r_d of fine cells per coarse cell along each tangential
 * axis: one face per fine cell on the interface, the coarse side interpolated
 * at the subface's mapped points with weights scaled by 1/r_d, geometry and
 * beta at the coarse face's own rule points, as GenDiL's operator does.
 * Replaces AdvectionNonconformingFaceOperator over
 * NonconformingCartesianIntermeshFaceConnectivity<dim, face>
 * (nonconformingcartesianintermeshfaceconnectivity.hpp:73-243;
 * advectionoperator.hpp:596-676; facetquaddata.hpp:86-155).  Same rule on
 * both plans; accumulates over one launch per subface position, so it never
 * runs in place: no input side may share storage with an output side
 * (GD_ERR_INVALID).  The first call for a new refinement ratio builds the
 * subface basis tables on the coarse plan (synchronous allocation and copy). */
int gd_adv_hmesh_apply_f64(const gd_plan* coarse, const gd_plan* fine, int32_t face,
                           const double* x_coarse, const double* x_fine, double* y_coarse,
                           double* y_fine, gd_stream stream);

/* ---- kernel variant selection (autotuner) ------------------------------ */
/* Times every compiled variant of the apply for this plan on scratch
 * vectors, keeps the fastest for later applies of that precision (4 or 8
 * bytes), and reports it.  Uniform plans: the generic LDS-staged kernel (0),
 * the register-resident 3D kernel with 1, 2 or 4 lanes per element, and the
 * Kronecker-factored 3D kernel (16 + lanes per element: 17, 18, 20) and its
 * LDS-staged form (48..55: 48 + 1 outputs staged through LDS, + 2 blocks of
 * 128 threads, + 4 one term at a time) and its streaming form (one wave per
 * block, persistent over tiles of 16 elements of one x-row; 56: one tile
 * buffer, refilled as soon as a tile's outputs are in registers, two waves
 * per SIMD; 57: two tile buffers, the next tile's copies into LDS in flight
 * while a tile computes, one wave per SIMD; rows must be a multiple of 16
 * elements) where they apply; 4D-6D plans of order 1 with constant
 * coefficients: 64 = the two-node Kronecker kernel (one lane per node).
 * p-adaptive plans:
 * the thread-group layouts of the quadrature kernel (0 = one wave per n=3
 * element, 1 = one block per n=3 element; n=4 elements take a block, n=2
 * elements a wave) and, in 6D for a field with paired affine components of
 * one sign each, 2 = the Kronecker-factored kernel.  No GenDiL counterpart:
 * GenDiL fixes its thread layout at compile time (ThreadBlockLayout,
 * threadlayout.hpp:46). */
int gd_plan_autotune(gd_plan* plan, int32_t precision_bytes, double mass_coef, double adv_coef,
                     int32_t reps, int32_t* chosen, double* best_ms);
/* Pin a variant: -1 (default); uniform plans 0 (generic), 1, 2 or 4 lanes
 * (register kernel), 17, 18 or 20 (Kronecker kernel), 48..55 (LDS-staged
 * Kronecker kernel), 56 or 57 (streaming Kronecker kernel), 64 (two-node
 * Kronecker kernel, 4D-6D); p-adaptive plans the variant index as reported by
 * gd_plan_autotune.  A pinned variant that does not serve a later apply
 * falls back to the default for that apply. */
int gd_plan_set_variant(gd_plan* plan, int32_t precision_bytes, int32_t lanes);
/* The variant an apply with these coefficients would run now (after falling
 * back from a pinned variant that does not serve the plan), in the codes above. */
int gd_plan_query_variant(const gd_plan* plan, int32_t precision_bytes, double mass_coef, double adv_coef,
                          int32_t* variant);

/* ---- device memory and vector kernels for the Krylov loops ------------ */
/* Vector ops used by GMRES_no_alloc / ConjugateGradient (gmres.hpp:74-288,
 * conjugategradient.hpp:217; vector.hpp:536-628).  GenDiL's Dot is host-only
 * (Utilities/MathHelperFunctions/dot.hpp:117-130); these stay on the device
 * and accumulate in FP64. */
int gd_malloc(void** ptr, size_t bytes);
int gd_free(void* ptr);
int gd_memcpy_h2d(void* dst, const void* src, size_t bytes);
int gd_memcpy_d2h(void* dst, const void* src, size_t bytes);
int gd_memset(void* dst, int value, size_t bytes, gd_stream stream);
int gd_stream_sync(gd_stream stream);
/* *done = 1 once every operation queued on the stream has finished, 0 while
 * some is pending (hipStreamQuery); an error code if the stream failed. */
int gd_stream_query(gd_stream stream, int32_t* done);
int gd_device_sync(void);

/* result[j] = sum_i V_j[i] * w[i] for j < k; V_j = V + j*ldv (FP64 accumulate),
 * written to device memory `result` (k doubles). */
int gd_multidot_f64(int64_t n, int32_t k, const double* V, int64_t ldv, const double* w,
                    double* result, gd_stream stream);
int gd_multidot_f32(int64_t n, int32_t k, const float* V, int64_t ldv, const float* w,
                    double* result, gd_stream stream);
/* gd_multidot_f64 without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same kernels and the
 * same result bit for bit; the block partials live in a fixed array of the
 * library (256 KB per GPU, held until the process ends) instead of a buffer
 * grown with hipMalloc.  At most 32 vectors per call (k > 32 returns
 * GD_ERR_INVALID).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (on different streams, or from different threads) race
 * and must be ordered by the caller. */
int gd_multidot_f64_noalloc(int64_t n, int32_t k, const double* V, int64_t ldv, const double* w,
                            double* result, gd_stream stream);
/* w -= sum_j h[j] * V_j  (h device, FP64) */
int gd_multiaxpy_f64(int64_t n, int32_t k, const double* h, const double* V, int64_t ldv,
                     double* w, gd_stream stream);
int gd_multiaxpy_f32(int64_t n, int32_t k, const double* h, const float* V, int64_t ldv,
                     float* w, gd_stream stream);
/* y = a*x + b*y */
int gd_axpby_f64(int64_t n, double a, const double* x, double b, double* y, gd_stream stream);
int gd_axpby_f32(int64_t n, float a, const float* x, float b, float* y, gd_stream stream);
/* precision conversion: y_f32 = scale * x_f64 ; y_f64 = scale * x_f32 */
int gd_convert_f64_to_f32(int64_t n, double scale, const double* x, float* y, gd_stream stream);
int gd_convert_f32_to_f64(int64_t n, double scale, const float* x, double* y, gd_stream stream);
/* y += scale * x (x f32, y f64) */
int gd_add_f32_to_f64(int64_t n, double scale, const float* x, double* y, gd_stream stream);
/* ---- device-resident Arnoldi process for GMRES(m), m <= 32 ------------ */
/* Replaces the host-driven modified Gram-Schmidt and Givens loop of
 * GMRES_no_alloc (gmres.hpp:150-215, m+1 dot/axpy pairs per iteration with a
 * blocking device-to-host copy per dot) by three fused passes over the basis
 * per iteration (classical Gram-Schmidt applied twice, FP64 accumulation)
 * whose last block also forms the Hessenberg column and its Givens rotation.
 * The workspace (gd_gmres_workspace_bytes(), device memory) holds H, the
 * rotations, g, the basis scales and the reduction partials.
 *   gd_gmres_begin:  g = beta e_0, scale of V_0 = s0; starts every cycle
 *     (it also resets the workspace's pass bookkeeping).
 *   gd_arnoldi_step: V + (k+1)*ldv holds K V_k on entry and the next basis
 *     vector on exit; if resid_out is not null, |g_{k+1}| (the GMRES residual
 *     estimate) is stored there with a system-scope store (pinned host memory
 *     from gd_host_alloc is allowed and may be polled) by the first block of
 *     the next step or update on this workspace, early in that kernel, so no
 *     kernel's completion waits for the host write.  A 16-byte aligned FP32
 *     V with ldv a multiple of 4 runs 16-byte accesses.
 *   gd_gmres_update: x += V_{0..k-1} y with y from the k x k triangular
 *     system (gmres.hpp:240-265); stores the last step's residual estimate. */
size_t gd_gmres_workspace_bytes(void);
int gd_gmres_begin(void* ws, double beta, double s0, gd_stream stream);
int gd_arnoldi_step_f64(int64_t n, int32_t m, int32_t k, double* V, int64_t ldv, void* ws, double* resid_out,
                        gd_stream stream);
int gd_arnoldi_step_f32(int64_t n, int32_t m, int32_t k, float* V, int64_t ldv, void* ws, double* resid_out,
                        gd_stream stream);
int gd_gmres_update_f64(int64_t n, int32_t m, int32_t k, const double* V, int64_t ldv, void* ws, double* x,
                        gd_stream stream);
int gd_gmres_update_f32(int64_t n, int32_t m, int32_t k, const float* V, int64_t ldv, void* ws, double* x,
                        gd_stream stream);

/* Device-resident CG and BiCGSTAB iterations (FP64).  The recurrence
 * scalars, the convergence test and the breakdown guards live on the device
 * (workspace of gd_ksolve_workspace_bytes(), device memory), so a host loop
 * issues iterations without waiting for any dot product; once the iteration
 * converges or stops, every later call on the workspace does nothing.  The
 * relative residual ||r|| / ||b|| of each iteration goes to resid_out (NaN
 * once stopped without converging) from the first block of the next kernel
 * on the workspace.  gd_ksolve_status copies the state (0 running,
 * 1 converged, 2 stopped), the completed iterations and the residual to the
 * host (synchronous).
 *   CG (conjugategradient.hpp:125-235): p = r on entry; per iteration the
 *     caller forms Ap = K p, then gd_cg_step_f64 does alpha, x, r, the test,
 *     beta and p.
 *   BiCGSTAB (bicgstab.hpp:44-157): rhat = r on entry; per iteration
 *     gd_bicgstab_p_f64 (p), v = K p, gd_bicgstab_s_f64 (alpha, s, the
 *     early-exit test), t = K s, gd_bicgstab_xr_f64 (omega, x, r, the test,
 *     rho); the reference's breakdown guards (|.| < 1e-30) and fallbacks. */
size_t gd_ksolve_workspace_bytes(void);
int gd_cg_begin_f64(int64_t n, const double* r, double bnorm, double rel_tol, void* ws, gd_stream stream);
int gd_cg_step_f64(int64_t n, double* x, double* r, double* p, const double* ap, void* ws, double* resid_out,
                   gd_stream stream);
/* The same CG passes on FP32 vectors (inner products, alpha and beta in
 * FP64): the inner solve of gd_cg_mixed.  No GenDiL counterpart. */
int gd_cg_begin_f32(int64_t n, const float* r, double bnorm, double rel_tol, void* ws, gd_stream stream);
int gd_cg_step_f32(int64_t n, float* x, float* r, float* p, const float* ap, void* ws, double* resid_out,
                   gd_stream stream);
int gd_bicgstab_begin_f64(int64_t n, const double* rhat, const double* r, double bnorm, double rel_tol, void* ws,
                          gd_stream stream);
int gd_bicgstab_p_f64(int64_t n, const double* r, double* p, const double* v, void* ws, gd_stream stream);
int gd_bicgstab_s_f64(int64_t n, const double* r, const double* rhat, const double* v, double* s, void* ws,
                      gd_stream stream);
int gd_bicgstab_xr_f64(int64_t n, double* x, double* r, const double* rhat, const double* p, const double* s,
                       const double* t, void* ws, double* resid_out, gd_stream stream);
int gd_ksolve_status(const void* ws, gd_stream stream, int32_t* state, int32_t* iterations, double* rel_residual);

/* Pinned host memory (coherent: a host loop can poll values the device
 * stores there, e.g. gd_arnoldi_step's resid_out, without draining the
 * stream) and events. */
typedef void* gd_event;
int gd_host_alloc(void** ptr, size_t bytes);
int gd_host_free(void* ptr);
int gd_event_create(gd_event* ev);
int gd_event_destroy(gd_event ev);
int gd_event_record(gd_event ev, gd_stream stream);
int gd_event_sync(gd_event ev);

/* ---- Krylov solvers on the operator mass_coef*M + adv_coef*A ------------ */
typedef struct gd_solve_params {
  int32_t restart;        /* GMRES restart length m (gmres.hpp:74) */
  int32_t max_iters;      /* total inner iterations */
  double rel_tol;         /* on ||b - K x|| / ||b|| */
  int32_t max_refinements;/* mixed precision: outer FP64 refinement steps */
  double inner_rel_tol;   /* mixed precision: inner FP32 GMRES tolerance */
} gd_solve_params;

typedef struct gd_solve_result {
  int32_t success;
  int32_t iterations;      /* inner iterations (operator applies in the Arnoldi loop) */
  int32_t refinements;
  double final_rel_residual; /* true FP64 residual ||b - K x|| / ||b|| */
  double seconds;
} gd_solve_result;

/* Restarted GMRES, all FP64 (GMRES_no_alloc semantics, gmres.hpp:74-288,
 * with the device-resident Arnoldi process above instead of host MGS).
 * x holds x0. */
int gd_gmres_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b,
                 double* x, const gd_solve_params* params, gd_solve_result* result,
                 gd_stream stream);
/* FP32 operator apply + FP32 Krylov basis inside GMRES; FP64 residual
 * recompute, FP64 inner products / orthogonalization coefficients and FP64
 * iterative-refinement update.  No GenDiL counterpart. */
int gd_gmres_mixed(const gd_plan* plan, double mass_coef, double adv_coef, const double* b,
                   double* x, const gd_solve_params* params, gd_solve_result* result,
                   gd_stream stream);
/* Unpreconditioned CG (conjugategradient.hpp:217) for SPD operators (mass),
 * device-resident (gd_cg_step_f64).  Stops at ||r|| / ||b|| <= rel_tol; the
 * result reports the true residual. */
int gd_cg_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b,
              double* x, const gd_solve_params* params, gd_solve_result* result,
              gd_stream stream);
/* CG with FP32 inner solves refined in FP64: every refinement step forms the
 * true residual r = b - K x with the FP64 apply, runs CG on K d = r / ||r||
 * with FP32 vectors and the FP32 apply until its residual falls below
 * max(inner_rel_tol, rel_tol ||b|| / ||r||) (or max_iters in all), and adds
 * ||r|| d to x in FP64; stops once the true FP64 residual is at most rel_tol
 * (or after max_refinements steps).  result->iterations counts the inner
 * iterations, result->refinements the steps.  No GenDiL counterpart. */
int gd_cg_mixed(const gd_plan* plan, double mass_coef, double adv_coef, const double* b,
                double* x, const gd_solve_params* params, gd_solve_result* result,
                gd_stream stream);
/* Unpreconditioned BiCGSTAB (BiCGSTAB_no_alloc, bicgstab.hpp:44-157),
 * device-resident (gd_bicgstab_*_f64): the reference's stopping rule
 * (||r|| / ||b|| < rel_tol), early exit, fallback and breakdown guards.
 * x holds x0; the result reports the true residual. */
int gd_bicgstab_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b,
                    double* x, const gd_solve_params* params, gd_solve_result* result,
                    gd_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* GENDIL_AMD_H */
