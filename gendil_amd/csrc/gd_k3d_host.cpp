// Host side of the register-resident 3D kernel: eligibility, one-axis field
// signatures, kernel-argument packing and instance lookup.
#include <hip/hip_runtime.h>

#include "gd_k3d.hpp"
#include "gd_k3d_api.hpp"

namespace gd {
namespace k3 {

namespace {

// One-axis description of a separable field: constant (axis -1) or a single
// tabulated factor; returns false for anything richer.
template <typename T>
bool one_axis(const HostField& h, const FieldDev<T>& dev, int dim, Field1<T>& out)
{
  out.scale = T(0);
  out.axis = -1;
  out.tab = nullptr;
  if (h.rank == 0) return true;
  if (h.rank != 1) return false;
  out.scale = (T)h.scale[0];
  for (int d = 0; d < dim; ++d) {
    if (h.offset[0][d] < 0) continue;
    if (out.axis >= 0) return false;  // two tabulated factors
    out.axis = d;
    out.tab = dev.tab[0][d];
  }
  return true;
}

template <typename T>
const std::vector<FieldDev<T>>& host_fields(const gd_plan* p)
{
  if constexpr (std::is_same_v<T, double>) return p->fields64_host;
  else return p->fields32_host;
}

template <typename T>
K3List<T> list_for(int N)
{
  if constexpr (std::is_same_v<T, double>) {
    if (N == 2) return instances_f64_n2();
    if (N == 3) return instances_f64_n3();
    if (N == 4) return instances_f64_n4();
  } else {
    if (N == 2) return instances_f32_n2();
    if (N == 3) return instances_f32_n3();
    if (N == 4) return instances_f32_n4();
  }
  return K3List<T>{nullptr, 0};
}

template <typename T>
bool prepare(const gd_plan* plan, int lanes, bool adv, bool mass, K3Params<T>& prm, K3Launch<T>* fn)
{
  const int N = plan->desc.order + 1;
  if (plan->dim != 3 || plan->padapt || plan->desc.basis != GD_BASIS_GLL || plan->desc.num_quad != N) return false;
  if (plan->E >= (int64_t)1 << 31) return false;
  for (int d = 0; d < 3; ++d)
    if (plan->desc.sizes[d] >= (int64_t)1 << 31) return false;
  const auto& hf = host_fields<T>(plan);
  if (hf.size() < 5) return false;
  int ax[4] = {-1, -1, -1, -1};
  if (adv) {
    if (plan->beta_kind != GD_BETA_SEPARABLE) return false;
    for (int d = 0; d < 3; ++d) {
      if (!one_axis<T>(plan->beta_host[d], hf[d], 3, prm.beta[d])) return false;
      ax[d] = prm.beta[d].axis;
    }
  } else {
    for (int d = 0; d < 3; ++d) prm.beta[d] = Field1<T>{T(0), -1, nullptr};
  }
  if (mass) {
    if (!one_axis<T>(plan->sigma_host, hf[3], 3, prm.sigma)) return false;
    ax[3] = prm.sigma.axis;
  } else {
    prm.sigma = Field1<T>{T(0), -1, nullptr};
  }
  const int want_sig = sig(ax[0], ax[1], ax[2], ax[3]);
  const bool bcg = adv && plan->bc_mode == GD_BC_INFLOW && plan->bc_host.rank > 0;
  const K3List<T> l = list_for<T>(N);
  // mass-only instances ignore the beta axes
  const int s = adv ? want_sig : sig(-1, -1, -1, ax[3]);
  // the instance compiled for this signature, else the run-time-axis one
  const K3Entry<T>* found = nullptr;
  for (int i = 0; i < l.n && !found; ++i) {
    const K3Entry<T>& k = l.e[i];
    if (k.lanes == lanes && k.adv == (int)adv && k.mass == (int)mass && k.bcg == (int)bcg && k.sig == s) found = &k;
  }
  for (int i = 0; i < l.n && !found; ++i) {
    const K3Entry<T>& k = l.e[i];
    if (k.lanes == lanes && k.adv == (int)adv && k.mass == (int)mass && k.bcg == (int)bcg && k.sig == kSigAny)
      found = &k;
  }
  if (!found) return false;
  *fn = found->fn;
  prm.E = (int)plan->E;
  prm.M0 = (int)plan->desc.sizes[0];
  prm.M1 = (int)plan->desc.sizes[1];
  prm.M2 = (int)plan->desc.sizes[2];
  prm.periodic = 0;
  double detJ = 1.0;
  for (int d = 0; d < 3; ++d) {
    if (plan->desc.periodic[d]) prm.periodic |= 1 << d;
    prm.invh[d] = (T)(1.0 / plan->desc.h[d]);
    detJ *= plan->desc.h[d];
  }
  prm.detJ = (T)detJ;
  prm.bc_inflow = plan->bc_mode == GD_BC_INFLOW ? 1 : 0;
  if constexpr (std::is_same_v<T, double>)
    prm.bc = static_cast<const FieldDev<double>*>(plan->fields64) + 4;
  else
    prm.bc = static_cast<const FieldDev<float>*>(plan->fields32) + 4;
  prm.num_blocks = 0;
  return true;
}

}  // namespace

template <typename T>
bool available(const gd_plan* plan, int lanes, bool adv, bool mass)
{
  K3Params<T> prm{};
  K3Launch<T> fn = nullptr;
  return prepare<T>(plan, lanes, adv, mass, prm, &fn);
}

template <typename T>
int apply(const gd_plan* plan, int lanes, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  K3Params<T> prm{};
  K3Launch<T> fn = nullptr;
  const bool adv = adv_coef != T(0), mass = mass_coef != T(0);
  if (!prepare<T>(plan, lanes, adv, mass, prm, &fn)) return GD_ERR_UNSUPPORTED;
  prm.adv_coef = adv_coef;
  prm.mass_coef = mass_coef;
  return fn(prm, x, y, s);
}

template bool available<double>(const gd_plan*, int, bool, bool);
template bool available<float>(const gd_plan*, int, bool, bool);
template int apply<double>(const gd_plan*, int, double, double, const double*, double*, hipStream_t);
template int apply<float>(const gd_plan*, int, float, float, const float*, float*, hipStream_t);

}  // namespace k3
}  // namespace gd
