// Host side and instances of the two-node Kronecker apply (gd_n2k.hpp):
// eligibility, the 2 x 2 factor matrices from the plan's 1D tables, launch.
#include <hip/hip_runtime.h>

#include <cmath>

#include "gd_internal.hpp"
#include "gd_n2k.hpp"
#include "gd_n2k_api.hpp"

namespace gd {
namespace n2k {
namespace {

// constant value of a field with no tabulated factor; false if it has one
bool constant_value(const HostField& f, double* v)
{
  double s = 0.0;
  for (int r = 0; r < f.rank; ++r) {
    for (int64_t off : f.offset[r])
      if (off >= 0) return false;
    s += f.scale[r];
  }
  *v = s;
  return true;
}

struct Geometry {
  double beta[6] = {};
  double sigma = 0.0;
};

bool eligible(const gd_plan* plan, bool adv, bool mass, Geometry* g)
{
  if (!plan || plan->padapt || plan->dim < 4 || plan->dim > 6 || plan->desc.order != 1) return false;
  if (plan->E >= ((int64_t)1 << 31) / 64) return false;
  if (mass && !constant_value(plan->sigma_host, &g->sigma)) return false;
  if (adv) {
    if (plan->beta_kind != GD_BETA_SEPARABLE) return false;
    for (int d = 0; d < plan->dim; ++d)
      if (!constant_value(plan->beta_host[d], &g->beta[d])) return false;
    if (plan->bc_mode == GD_BC_INFLOW) {
      double gv = 0.0;
      if (!constant_value(plan->bc_host, &gv) || gv != 0.0) return false;
    }
  }
  return find_tables(plan, 2, plan->desc.num_quad) != nullptr;
}

using M2 = double[4];

void mul(const M2 a, const M2 b, M2 c)
{
  const double r[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                       a[2] * b[1] + a[3] * b[3]};
  for (int i = 0; i < 4; ++i) c[i] = r[i];
}

template <typename T>
bool params(const gd_plan* plan, double mass_coef, double adv_coef, N2Params<T>& p)
{
  Geometry g;
  const bool adv = adv_coef != 0.0, mass = mass_coef != 0.0;
  if (!eligible(plan, adv, mass, &g)) return false;
  const TableSet* ts = find_tables(plan, 2, plan->desc.num_quad);
  const int Q = plan->desc.num_quad, D = plan->dim;
  // M = B^T W B, Dm = Gb^T W B (derivative on the test function)
  M2 Mm = {0, 0, 0, 0}, Dm = {0, 0, 0, 0};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < Q; ++q) {
        Mm[i * 2 + j] += ts->B[q * 2 + i] * ts->w[q] * ts->B[q * 2 + j];
        Dm[i * 2 + j] += ts->Gb[q * 2 + i] * ts->w[q] * ts->B[q * 2 + j];
      }
  const double det = Mm[0] * Mm[3] - Mm[1] * Mm[2];
  if (!(std::fabs(det) > 0.0)) return false;
  const M2 Mi = {Mm[3] / det, -Mm[1] / det, -Mm[2] / det, Mm[0] / det};
  const double* b0 = ts->b0.data();
  const double* b1 = ts->b1.data();
  double detJ = 1.0;
  for (int d = 0; d < D; ++d) detJ *= plan->desc.h[d];
  p = N2Params<T>{};
  p.E = (int)plan->E;
  p.periodic = 0;
  p.negmask = 0;
  p.bc_zero = plan->bc_mode == GD_BC_INFLOW ? 1 : 0;
  p.tab[76] = (T)(mass_coef * detJ * g.sigma);
  int st = 1;
  for (int d = 0; d < 6; ++d) {
    p.M[d] = d < D ? (int)plan->desc.sizes[d] : 1;
    p.stride[d] = st;
    st *= p.M[d];
    if (d < D && plan->desc.periodic[d]) p.periodic |= 1 << d;
  }
  // row (b, b), (b, 1 - b) of a row-major 2 x 2 matrix at dst + 2 b
  auto rows = [&](const M2 a, T* dst) {
    dst[0] = (T)a[0];
    dst[1] = (T)a[1];
    dst[2] = (T)a[3];
    dst[3] = (T)a[2];
  };
  rows(Mm, p.tab + 72);
  for (int d = 0; d < D; ++d) {
    const double kt = adv ? adv_coef * g.beta[d] * detJ / plan->desc.h[d] : 0.0;
    const bool neg = g.beta[d] < 0.0;
    if (neg) p.negmask |= 1 << d;
    M2 cb, ci, f;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        cb[i * 2 + j] = -kt * Dm[i * 2 + j];
        // own outflow face: beta.n u v with beta.n = |beta_d| detJ / h_d
        ci[i * 2 + j] = cb[i * 2 + j] + (neg ? -kt * b0[i] * b0[j] : kt * b1[i] * b1[j]);
        // inflow face: beta.n u_nb v, beta.n = -|beta_d| detJ / h_d, the
        // neighbour's trace on its opposite face
        f[i * 2 + j] = neg ? kt * b1[i] * b0[j] : -kt * b0[i] * b1[j];
      }
    mul(Mi, cb, cb);
    mul(Mi, ci, ci);
    mul(Mi, f, f);
    rows(cb, p.tab + d * 12);
    rows(ci, p.tab + d * 12 + 4);
    rows(f, p.tab + d * 12 + 8);
  }
  return true;
}

}  // namespace

template <typename T>
bool available(const gd_plan* plan, bool adv, bool mass)
{
  Geometry g;
  return eligible(plan, adv, mass, &g);
}

template <typename T>
int apply(const gd_plan* plan, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  N2Params<T> p;
  if (!params<T>(plan, (double)mass_coef, (double)adv_coef, p)) return GD_ERR_UNSUPPORTED;
  int rc = 0;
  switch (plan->dim) {
    case 4: rc = launch<T, 4>(p, x, y, s); break;
    case 5: rc = launch<T, 5>(p, x, y, s); break;
    default: rc = launch<T, 6>(p, x, y, s); break;
  }
  if (rc) return fail(GD_ERR_HIP, "two-node Kronecker apply: launch failed");
  return GD_OK;
}

template bool available<double>(const gd_plan*, bool, bool);
template bool available<float>(const gd_plan*, bool, bool);
template int apply<double>(const gd_plan*, double, double, const double*, double*, hipStream_t);
template int apply<float>(const gd_plan*, float, float, const float*, float*, hipStream_t);

}  // namespace n2k
}  // namespace gd
