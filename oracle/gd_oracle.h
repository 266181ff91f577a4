/* CPU restatement of GenDiL's matrix-free DG operator apply (mass + upwind
 * advection on Cartesian hypercube meshes).  TEST INFRASTRUCTURE ONLY: this
 * library is the checker for the HIP kernels and the "port" CPU baseline.  It
 * is never linked into, loaded by, or called from the product library.
 *
 * Every function states the reference file:line it follows (paths relative to
 * the GenDiL source tree).
 *
 * Layout (GenDiL's L2 layout, include/gendil/FiniteElementMethod/
 * MatrixFreeOperators/KernelOperators/DoFIO/elementtensorview.hpp:42-64):
 * DoF (i0..i_{D-1}) of element e at i0 + n*(i1 + n*(...)) + n^D*e, and
 * e = e0 + M0*(e1 + M1*(...)) (getstructuredsubindex.hpp:87).
 */
#ifndef GD_ORACLE_H
#define GD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDO_MAX_DIM 6
#define GDO_MAX_TERMS 8

/* A scalar field f(X) = sum_r scale[r] * prod_d T_{r,d}(X_d), each factor a 1D
 * table sampled at the per-axis positions returned by gdo_axis_points
 * (offset -1 = factor 1). */
typedef struct gdo_separable {
  int32_t rank;
  double scale[GDO_MAX_TERMS];
  int64_t offset[GDO_MAX_TERMS][GDO_MAX_DIM];
  const double* data;
} gdo_separable;

enum { GDO_BASIS_GLL = 0, GDO_BASIS_GL = 1 };
enum { GDO_BETA_SEPARABLE = 0, GDO_BETA_POINTWISE = 1, GDO_BETA_AFFINE = 2 };
enum { GDO_BC_NONE = 0, GDO_BC_INFLOW = 1 };

typedef struct gdo_problem {
  int32_t dim;
  int64_t sizes[GDO_MAX_DIM];
  double h[GDO_MAX_DIM];
  double origin[GDO_MAX_DIM];
  int32_t periodic[GDO_MAX_DIM];
  int32_t order;      /* p: n = p+1 nodes per dim (uniform order) */
  int32_t nq;         /* quadrature points per dim (uniform rule) */
  int32_t basis;      /* GDO_BASIS_* */
  /* y = mass_coef * M(sigma) x + adv_coef * A(beta, g) x */
  double mass_coef;
  double adv_coef;
  gdo_separable sigma;
  int32_t beta_kind;
  gdo_separable beta[GDO_MAX_DIM];
  /* pointwise beta: cell [E][D][nq^D]; face normal component per axis a:
   * [F_a][nq^(D-1)] with F_a = prod_{d!=a} M_d * (M_a + !periodic_a). */
  const double* beta_cell;
  const double* beta_face[GDO_MAX_DIM];
  /* affine beta: beta_c(X) = c[c] + sum_d A[c][d] X_d */
  double beta_affine_A[GDO_MAX_DIM][GDO_MAX_DIM];
  double beta_affine_c[GDO_MAX_DIM];
  int32_t bc_mode;    /* GDO_BC_NONE: boundary faces skipped; INFLOW: u+ = g */
  gdo_separable bc;
  /* p-adaptive: per-element order/rule (NULL = uniform).  DoFs of element e
   * start at dof_offset[e]; tables for order p are built on demand. */
  const int8_t* elem_order;
  const int8_t* elem_nq;
  const int64_t* dof_offset;
  int32_t num_threads;  /* 0 = OpenMP default */
  /* Planted mistakes, for tests that show the GenDiL fixtures catch them
   * (0 = none): 1 = a face between elements of different orders integrated
   * with the smaller of the two cell rules instead of the larger; 2 = the
   * neighbour's trace on such a face taken from its node slab with two
   * tangential node axes exchanged (D >= 3).  The first is the same on both
   * sides of a face, so conservation cannot see it.  3 and 4: see
   * gdo_diffusion_apply. */
  int32_t plant;
} gdo_problem;

int gdo_version(void);
/* GL (basis=1) or GLL (basis=0) points and weights on [0,1]. */
int gdo_points(int32_t basis, int32_t npts, double* pts, double* wts);
/* 1D sampling positions for axis d: sizes[d]*(nq+2) values, per element
 * e_d: nq rule points, then xi=0, then xi=1 (X = (o + h e) + h xi). */
int gdo_axis_points(const gdo_problem* p, int32_t axis, double* out);
int64_t gdo_num_dofs(const gdo_problem* p);
/* y = mass_coef*M x + adv_coef*A x (element-centric, overwrite). */
int gdo_apply(const gdo_problem* p, const double* x, double* y);
/* The forward apply on a subset of elements: writes y only at the DoFs of
 * elems[0..count) (for checks of a large apply on a sample of its elements). */
int gdo_apply_elements(const gdo_problem* p, const double* x, double* y, const int64_t* elems, int64_t count);
/* z = (linear part of mass_coef*M + adv_coef*A)^T r. */
int gdo_apply_transpose(const gdo_problem* p, const double* r, double* z);
/* For pointwise beta: dL/dbeta_cell [E][D][nq^D] and dL/dbeta_face[a]
 * for L = <r, A(beta) u> (adv_coef included). */
int gdo_grad_beta(const gdo_problem* p, const double* u, const double* r,
                  double* dbeta_cell, double* const* dbeta_face);
/* SIPDG diffusion (DiffusionOperator, diffusionoperator.hpp:44-331), y
 * overwritten: diffusivity mu, symmetry parameter sigma (-1 symmetric, 0
 * incomplete, 1 non-symmetric interior penalty), penalty kappa; faces on a
 * non-periodic boundary carry the weak homogeneous Dirichlet terms.  Uses the
 * mesh, basis and rule of p (uniform order; mass/advection fields ignored).
 * p->plant 3: interior symmetry term without its factor 1/2; 4: the
 * neighbour's normal derivative taken at the wrong end of its element. */
int gdo_diffusion_apply(const gdo_problem* p, const gdo_separable* mu, double sigma, double kappa,
                        const double* x, double* y);

#ifdef __cplusplus
}
#endif

#endif
