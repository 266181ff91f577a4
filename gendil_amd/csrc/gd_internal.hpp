// Internal types shared by the plan code and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gendil_amd.h"

namespace gd {

constexpr int kMaxDim = GD_MAX_DIM;
constexpr int kMaxTerms = GD_MAX_TERMS;
constexpr int kMaxOrderNodes = 8;  // n = p+1 <= 8
constexpr int kMaxQuad = 8;

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define GD_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return ::gd::hip_fail(e_, #call); \
  } while (0)

// Separable field on the device: value = sum_r scale[r] * prod_d tab[r][d][...]
template <typename T>
struct FieldDev {
  int rank;
  T scale[kMaxTerms];
  const T* tab[kMaxTerms][kMaxDim];  // nullptr = factor 1
};

// 1D tables for one (basis, n, q) pair, laid out contiguously on the device:
// B[q][n] values, Gq[q][q] collocated derivative, b0[n], b1[n] face traces,
// w[q] weights, xq[q] points, Gb[q][n] basis derivative, d0[n], d1[n] basis
// derivative at the element ends.
template <typename T>
struct Tables1D {
  const T* B;
  const T* Gq;
  const T* b0;
  const T* b1;
  const T* w;
  const T* xq;
  const T* Gb;  // basis derivative at quadrature points Gb[q][i] = l_i'(x_q)
  const T* d0;  // l_i'(0)
  const T* d1;  // l_i'(1)
};

// Kernel parameters (passed by value as kernel arguments).
template <typename T>
struct Params {
  int64_t M[kMaxDim];       // elements per axis
  int64_t E;                // number of elements
  int periodic_mask;        // bit d set = periodic axis d
  T h[kMaxDim];
  T invh[kMaxDim];
  T origin[kMaxDim];
  T detJ;
  T mass_coef;
  T adv_coef;
  Tables1D<T> tab;
  // coefficients
  int beta_kind;            // gd_beta_kind
  const FieldDev<T>* beta;  // [dim] (separable)
  const FieldDev<T>* sigma;
  const FieldDev<T>* bc;
  int bc_mode;
  T affA[kMaxDim][kMaxDim];
  T affc[kMaxDim];
  const T* beta_cell;       // pointwise
  const T* beta_face[kMaxDim];
  // grad-beta outputs
  T* dbeta_cell;
  T* dbeta_face[kMaxDim];
  // SIPDG diffusion: diffusivity, symmetry parameter, penalty
  const FieldDev<T>* mu;
  T ip_sigma;
  T ip_kappa;
  // 1: the basis values at the element ends are unit vectors (b0 = e_0,
  // b1 = e_{n-1}: GLL), so a face trace is the node layer at that end
  int node_traces;
  // XCD-aware block remap
  int num_blocks;
};

struct TableSet {
  int n = 0, q = 0;
  void* dev64 = nullptr;  // contiguous device block (double)
  void* dev32 = nullptr;  // contiguous device block (float)
  Tables1D<double> t64{};
  Tables1D<float> t32{};
  std::vector<double> B, Gq, b0, b1, w, xq, Gb, d0, d1;  // host copies
};

struct HostField {
  int rank = 0;
  std::vector<double> scale;
  std::vector<std::vector<int64_t>> offset;  // [rank][dim]
  std::vector<double> data;
};

namespace pa {
struct State;
void destroy_state(State* s);
}  // namespace pa
namespace k3k {
struct State;
void destroy_state(State* s);
}  // namespace k3k

}  // namespace gd

struct gd_plan {
  gd_desc desc{};
  int dim = 0;
  int64_t E = 0;
  int64_t ndofs = 0;
  bool padapt = false;
  std::vector<int8_t> elem_order, elem_quad;
  std::vector<int64_t> dof_offset_host;
  int64_t* dof_offset_dev = nullptr;
  int8_t* elem_order_dev = nullptr;
  int8_t* elem_quad_dev = nullptr;
  std::vector<gd::TableSet*> tables;  // one per (n, q) in use
  // coefficients
  int beta_kind = GD_BETA_SEPARABLE;
  bool beta_set = false;
  gd::HostField beta_host[gd::kMaxDim];
  gd::HostField sigma_host;
  gd::HostField bc_host;
  int bc_mode = GD_BC_NONE;
  // SIPDG diffusion (gd_plan_set_diffusion)
  bool diff_set = false;
  gd::HostField mu_host;
  double ip_sigma = 0.0, ip_kappa = 0.0;
  double affA[gd::kMaxDim][gd::kMaxDim] = {};
  double affc[gd::kMaxDim] = {};
  const double* beta_cell = nullptr;
  const double* beta_face[gd::kMaxDim] = {};
  // device copies of separable fields: [0..dim) beta, dim sigma, dim+1 bc,
  // dim+2 diffusivity
  void* fields64 = nullptr;
  void* fields32 = nullptr;
  std::vector<gd::FieldDev<double>> fields64_host;  // host copies of the device descriptors
  std::vector<gd::FieldDev<float>> fields32_host;
  std::vector<void*> field_tables;  // device table allocations
  // autotuned 3D kernel choice ([0] FP64, [1] FP32): -1 = not tuned,
  // 0 = generic kernel, 1/2/4 = register kernel with that many lanes per
  // element, 16 + lanes = Kronecker-factored kernel
  int tuned_k3_lanes[2] = {-1, -1};
  // p-adaptive path: kernel variant per precision (-1 = default), lazily
  // built element buckets and device parameter block, and a counter bumped by
  // every coefficient change so the device copy is refreshed
  int tuned_pa[2] = {-1, -1};
  mutable gd::pa::State* pa_state = nullptr;
  // factor tables of the Kronecker-factored 3D kernel (lazily built)
  mutable gd::k3k::State* k3k_state = nullptr;
  // basis tables at the mapped points of the subfaces of an h-nonconforming
  // interface (face operators), lazily built for the refinement ratios in
  // hmap_key
  mutable void* hmap_dev = nullptr;
  mutable std::vector<int64_t> hmap_key;
  int coef_version = 0;
  ~gd_plan();
};

namespace gd {
const TableSet* find_tables(const gd_plan* p, int n, int q);
template <typename T>
Params<T> make_params(const gd_plan* p, const TableSet* ts, T mass_coef, T adv_coef);
}  // namespace gd
