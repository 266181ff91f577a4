// Host side of the p-adaptive apply: eligibility, element buckets per order,
// device parameter block, launch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "gd_internal.hpp"
#include "gd_padapt.hpp"
#include "gd_padapt_api.hpp"

namespace gd {
namespace pa {

struct State {
  int D = 0;
  const char* elems_why = nullptr;  // why the element orders/rules are not served (null: served)
  int32_t* lists = nullptr;  // device: elements of n=4 | n=3 | n=2, ascending
  int count[3] = {0, 0, 0};
  void* params[2] = {nullptr, nullptr};  // device PParams<double>, PParams<float>
  int version[2] = {-1, -1};
  bool kron_ok = false;  // the current field suits the Kronecker 6D kernel
};

void destroy_state(State* s)
{
  if (!s) return;
  if (s->lists) (void)hipFree(s->lists);
  for (void* p : s->params)
    if (p) (void)hipFree(p);
  delete s;
}

// 6D: 0, 1 = quadrature kernel layouts, 2 = Kronecker kernel (phase-space
// fields, falls back to 0 for other fields)
int num_variants(int D) { return D == 6 ? 3 : 1; }

Variant variant(int D, int v)
{
  if (D == 6) return v == 1 ? Variant{256, 256, 64} : (v == 2 ? Variant{256, 64, 64} : Variant{256, 64, 64});
  return Variant{64, 64, 64};
}

namespace {

bool constant_field(const HostField& f, double* value)
{
  double v = 0.0;
  for (int r = 0; r < f.rank; ++r) {
    for (int64_t off : f.offset[r])
      if (off >= 0) return false;
    v += f.scale[r];
  }
  *value = v;
  return true;
}

template <typename T>
PParams<T> host_params(const gd_plan* plan, const State& st)
{
  PParams<T> k;
  std::memset(&k, 0, sizeof(k));
  const int D = plan->dim;
  k.E = plan->E;
  double detJ = 1.0;
  for (int d = 0; d < D; ++d) {
    k.M[d] = plan->desc.sizes[d];
    k.h[d] = (T)plan->desc.h[d];
    k.invh[d] = (T)(1.0 / plan->desc.h[d]);
    k.origin[d] = (T)plan->desc.origin[d];
    if (plan->desc.periodic[d]) k.periodic_mask |= 1 << d;
    detJ *= plan->desc.h[d];
  }
  k.detJ = (T)detJ;
  k.beta_affine = plan->beta_kind == GD_BETA_AFFINE;
  for (int c = 0; c < D; ++c) {
    bool zero = true;
    if (plan->beta_kind == GD_BETA_AFFINE) {
      k.affc[c] = (T)plan->affc[c];
      zero = plan->affc[c] == 0.0;
      for (int d = 0; d < D; ++d) {
        k.affA[c][d] = (T)plan->affA[c][d];
        zero = zero && plan->affA[c][d] == 0.0;
      }
    } else {
      double v = 0.0;
      constant_field(plan->beta_host[c], &v);
      k.affc[c] = (T)v;
      zero = v == 0.0;
    }
    if (zero) k.zero_axis_mask |= 1 << c;
  }
  // Kronecker 6D kernel: beta_c = a_c X_pc + b_c, pc = (c + 3) mod 6, of one
  // sign over the mesh (its extremes are at the ends of the pc-range)
  k.kron_ok = D == 6 ? 1 : 0;
  for (int c = 0; c < D && k.kron_ok; ++c) {
    const int pc = (c + 3) % 6;
    double a = 0.0, b = 0.0;
    if (plan->beta_kind == GD_BETA_AFFINE) {
      for (int d = 0; d < D; ++d)
        if (d != pc && plan->affA[c][d] != 0.0) k.kron_ok = 0;
      a = plan->affA[c][pc];
      b = plan->affc[c];
    } else {
      constant_field(plan->beta_host[c], &b);
    }
    const double lo = plan->desc.origin[pc], hi = lo + plan->desc.sizes[pc] * plan->desc.h[pc];
    const double v0 = a * lo + b, v1 = a * hi + b;
    if (v0 >= 0.0 && v1 >= 0.0) continue;
    if (v0 <= 0.0 && v1 <= 0.0) k.neg_mask |= 1 << c;
    else k.kron_ok = 0;
  }
  double sigma = 1.0, g = 0.0;
  constant_field(plan->sigma_host, &sigma);
  constant_field(plan->bc_host, &g);
  k.sigma = (T)sigma;
  k.bc_mode = plan->bc_mode;
  k.bc_value = plan->bc_mode == GD_BC_INFLOW ? (T)g : T(0);
  k.order = plan->elem_order_dev;
  k.dof_offset = plan->dof_offset_dev;
  int64_t off = 0;
  for (int c = 0; c < 3; ++c) {
    k.list[c] = st.lists + off;
    k.count[c] = st.count[c];
    off += st.count[c];
  }
  return k;
}

int ensure_state(const gd_plan* plan)
{
  if (plan->pa_state) return GD_OK;
  auto* st = new State();
  st->D = plan->dim;
  for (int64_t e = 0; e < plan->E && !st->elems_why; ++e) {
    const int n = plan->elem_order[e] + 1;
    if (n < 2 || n > 4) st->elems_why = "p-adaptive kernels need orders 1..3";
    else if (plan->elem_quad[e] != n) st->elems_why = "p-adaptive kernels need q = p + 1 per element";
  }
  if (st->elems_why) {
    plan->pa_state = st;
    return GD_OK;
  }
  std::vector<int32_t> buckets[3];
  for (int64_t e = 0; e < plan->E; ++e) buckets[4 - (plan->elem_order[e] + 1)].push_back((int32_t)e);
  std::vector<int32_t> all;
  all.reserve(plan->E);
  for (int c = 0; c < 3; ++c) {
    st->count[c] = (int)buckets[c].size();
    all.insert(all.end(), buckets[c].begin(), buckets[c].end());
  }
  if (hipMalloc(&st->lists, all.size() * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&st->params[0], sizeof(PParams<double>)) != hipSuccess ||
      hipMalloc(&st->params[1], sizeof(PParams<float>)) != hipSuccess) {
    destroy_state(st);
    return fail(GD_ERR_ALLOC, "p-adaptive apply: device allocation failed");
  }
  if (hipMemcpy(st->lists, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    destroy_state(st);
    return fail(GD_ERR_HIP, "p-adaptive apply: upload of element lists failed");
  }
  plan->pa_state = st;
  return GD_OK;
}

}  // namespace

bool eligible(const gd_plan* plan, const char** why)
{
  auto no = [&](const char* w) {
    if (why) *why = w;
    return false;
  };
  if (!plan->padapt) return no("not a p-adaptive plan");
  if (plan->dim != 3 && plan->dim != 6) return no("p-adaptive kernels are compiled for 3D and 6D");
  if (plan->desc.basis != GD_BASIS_GLL) return no("p-adaptive kernels need the GLL basis");
  if (plan->E >= ((int64_t)1 << 31)) return no("too many elements for the p-adaptive kernels");
  if (ensure_state(plan) != GD_OK) return no("p-adaptive set-up failed");
  if (plan->pa_state->elems_why) return no(plan->pa_state->elems_why);
  double v;
  if (plan->beta_kind == GD_BETA_POINTWISE) return no("p-adaptive kernels take affine or constant beta");
  if (plan->beta_kind == GD_BETA_SEPARABLE)
    for (int c = 0; c < plan->dim; ++c)
      if (!constant_field(plan->beta_host[c], &v)) return no("p-adaptive kernels take affine or constant beta");
  if (!constant_field(plan->sigma_host, &v)) return no("p-adaptive kernels take a constant sigma");
  if (plan->bc_mode == GD_BC_INFLOW && !constant_field(plan->bc_host, &v))
    return no("p-adaptive kernels take constant inflow data");
  return true;
}

int refresh(const gd_plan* plan)
{
  if (!plan->padapt) return GD_OK;
  if (int rc = ensure_state(plan)) return rc;
  if (!eligible(plan, nullptr)) return GD_OK;
  State& st = *plan->pa_state;
  const PParams<double> h64 = host_params<double>(plan, st);
  const PParams<float> h32 = host_params<float>(plan, st);
  GD_HIP(hipMemcpy(st.params[0], &h64, sizeof(h64), hipMemcpyHostToDevice));
  GD_HIP(hipMemcpy(st.params[1], &h32, sizeof(h32), hipMemcpyHostToDevice));
  st.version[0] = st.version[1] = plan->coef_version;
  st.kron_ok = h64.kron_ok != 0;
  return GD_OK;
}

template <typename T>
int apply(const gd_plan* plan, int v, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  const char* why = nullptr;
  if (!eligible(plan, &why)) return fail(GD_ERR_UNSUPPORTED, why);
  State& st = *plan->pa_state;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  // the parameter block is rebuilt by the coefficient setters (refresh)
  if (st.version[slot] != plan->coef_version)
    return fail(GD_ERR_INVALID, "p-adaptive apply: device parameters out of date");
  const bool kron = plan->dim == 6 && st.kron_ok;
  if (v < 0) v = kron ? 2 : 0;
  if (v >= num_variants(plan->dim) || (v == 2 && !kron)) v = 0;
  if (v == 2) {
    const int it0 = st.count[0], it1 = (st.count[1] + 3) / 4, it2 = (st.count[2] + 3) / 4;
    return launch_kron<T>(adv_coef != T(0), it0 + it1 + it2, it0, it1, st.params[slot], mass_coef, adv_coef, x, y, s);
  }
  const Variant var = variant(plan->dim, v);
  const int tpe[3] = {var.tpe4, var.tpe3, var.tpe2};
  int items[3];
  for (int c = 0; c < 3; ++c) {
    const int epi = kNT / tpe[c];
    items[c] = (st.count[c] + epi - 1) / epi;
  }
  const int grid = items[0] + items[1] + items[2];
  return launch<T>(plan->dim, adv_coef != T(0), v, grid, items[0], items[1], st.params[slot], mass_coef, adv_coef, x,
                   y, s);
}

int resolved_variant(const gd_plan* plan, int v)
{
  const char* why = nullptr;
  if (!eligible(plan, &why)) return 0;
  if (plan->pa_state->version[0] != plan->coef_version && plan->pa_state->version[1] != plan->coef_version) {
    // parameters not built yet for this field: decide from a host copy
    const PParams<double> hp = host_params<double>(plan, *plan->pa_state);
    plan->pa_state->kron_ok = hp.kron_ok != 0;
  }
  const bool kron = plan->dim == 6 && plan->pa_state->kron_ok;
  if (v < 0) v = kron ? 2 : 0;
  if (v >= num_variants(plan->dim) || (v == 2 && !kron)) v = 0;
  return v;
}

template int apply<double>(const gd_plan*, int, double, double, const double*, double*, hipStream_t);
template int apply<float>(const gd_plan*, int, float, float, const float*, float*, hipStream_t);

}  // namespace pa
}  // namespace gd
