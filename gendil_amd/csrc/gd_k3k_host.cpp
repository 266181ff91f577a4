// Host side of the Kronecker-factored 3D kernel: eligibility, per-position
// factor tables, kernel-argument packing and instance lookup.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gd_k3k.hpp"
#include "gd_k3k_api.hpp"

namespace gd {
namespace k3k {

struct State {
  int version = -1;
  bool ok = false;        // coefficients fit this kernel (one-axis, one sign per component on faces)
  int negmask = 0;        // bit d: beta_d <= 0 on every face point
  int axis[3] = {-1, -1, -1};
  double cval[3] = {0.0, 0.0, 0.0};  // value of a constant component
  void* tab[2][3] = {};   // [FP64, FP32][component] device tables
  void* fs[2][3] = {};
};

void destroy_state(State* s)
{
  if (!s) return;
  for (auto& pr : s->tab)
    for (void* q : pr)
      if (q) (void)hipFree(q);
  for (auto& pr : s->fs)
    for (void* q : pr)
      if (q) (void)hipFree(q);
  delete s;
}

int default_lanes(int n, size_t elem_bytes)
{
  (void)elem_bytes;
  if (n == 4) return 4;
  if (n == 2) return 2;
  return 1;
}

int default_staged(int n, size_t elem_bytes)
{
  if (n != 4 && n != 2) return -1;
  // measured on MI355X at 64^3, p = 3: one-pass z stage, outputs staged
  // through LDS; FP64 in 256-thread blocks, FP32 in 128-thread blocks
  return elem_bytes == 8 ? 32 + 1 : 32 + 2 + 1;
}

namespace {

template <typename T>
KList<T> list_for(int N)
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
  return KList<T>{nullptr, 0};
}

// constant value of a field with no tabulated factor; false if it has one
bool constant_of(const HostField& f, double* v)
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

template <typename T>
int upload(const std::vector<double>& h, void** dst)
{
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  if (h.empty()) return GD_OK;
  std::vector<T> c(h.begin(), h.end());
  GD_HIP(hipMalloc(dst, c.size() * sizeof(T)));
  GD_HIP(hipMemcpy(*dst, c.data(), c.size() * sizeof(T), hipMemcpyHostToDevice));
  return GD_OK;
}

}  // namespace

int refresh(const gd_plan* plan)
{
  if (!plan->k3k_state) plan->k3k_state = new State();
  State& st = *plan->k3k_state;
  st.version = plan->coef_version;
  st.ok = false;
  st.negmask = 0;
  if (plan->dim != 3 || plan->padapt || plan->desc.basis != GD_BASIS_GLL) return GD_OK;
  if (plan->beta_kind != GD_BETA_SEPARABLE) return GD_OK;
  const int N = plan->desc.order + 1, Q = N;
  const TableSet* ts = find_tables(plan, N, Q);
  if (!ts) return GD_OK;
  for (int d = 0; d < 3; ++d) {
    const HostField& f = plan->beta_host[d];
    std::vector<double> tab, fs;
    st.axis[d] = -1;
    st.cval[d] = 0.0;
    if (constant_of(f, &st.cval[d])) {
      if (st.cval[d] < 0.0) st.negmask |= 1 << d;
    } else {
      if (f.rank != 1) return GD_OK;
      int a = -1;
      for (int dd = 0; dd < 3; ++dd)
        if (f.offset[0][dd] >= 0) {
          if (a >= 0) return GD_OK;  // two tabulated factors
          a = dd;
        }
      st.axis[d] = a;
      const int64_t off = f.offset[0][a];
      const int64_t Ma = plan->desc.sizes[a];
      const double s = f.scale[0];
      tab.assign((size_t)Ma * N * N, 0.0);
      if (a == d) fs.assign((size_t)Ma * 2, 0.0);
      bool any_pos = false, any_neg = false;
      for (int64_t pos = 0; pos < Ma; ++pos) {
        const double* v = f.data.data() + off + pos * (Q + 2);
        double beta[16];
        for (int q = 0; q < Q + 2; ++q) beta[q] = s * v[q];
        // face points: the rule points along a (a != d) or xi = 0, 1 (a == d)
        const int q0 = a != d ? 0 : Q, q1 = a != d ? Q : Q + 2;
        for (int q = q0; q < q1; ++q) {
          any_pos |= beta[q] > 0.0;
          any_neg |= beta[q] < 0.0;
        }
        if (a == d) {
          fs[pos * 2] = beta[Q];
          fs[pos * 2 + 1] = beta[Q + 1];
        }
        const std::vector<double>& L = a == d ? ts->Gb : ts->B;
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < N; ++j) {
            double acc = 0.0;
            for (int q = 0; q < Q; ++q) acc += L[q * N + i] * ts->w[q] * beta[q] * ts->B[q * N + j];
            tab[(pos * N + i) * N + j] = acc;
          }
      }
      if (any_pos && any_neg) return GD_OK;  // the upwind side changes along a face
      if (any_neg) st.negmask |= 1 << d;
    }
    int rc = upload<double>(tab, &st.tab[0][d]);
    if (!rc) rc = upload<float>(tab, &st.tab[1][d]);
    if (!rc) rc = upload<double>(fs, &st.fs[0][d]);
    if (!rc) rc = upload<float>(fs, &st.fs[1][d]);
    if (rc) return rc;
  }
  st.ok = true;
  return GD_OK;
}

namespace {

template <typename T>
bool prepare(const gd_plan* plan, int lanes, bool adv, bool mass, KParams<T>& prm, KLaunch<T>* fn)
{
  const int N = plan->desc.order + 1;
  if (plan->dim != 3 || plan->padapt || plan->desc.basis != GD_BASIS_GLL || plan->desc.num_quad != N) return false;
  if (plan->E >= (int64_t)1 << 31) return false;
  // the streaming kernel deals tiles of 64 / N elements of one x-row
  if ((lanes == 40 || lanes == 41 || lanes == 42) && (64 % N != 0 || plan->desc.sizes[0] % (64 / N) != 0))
    return false;
  double sigma = 0.0;
  if (mass && !constant_of(plan->sigma_host, &sigma)) return false;
  int sigc = ksig(-1, -1, -1);
  int slot = sizeof(T) == 8 ? 0 : 1;
  const State* st = nullptr;
  if (adv) {
    if (plan->bc_mode == GD_BC_INFLOW) {
      double g = 0.0;
      if (!constant_of(plan->bc_host, &g) || g != 0.0) return false;
    }
    // the factor tables are rebuilt by the coefficient setters (refresh)
    st = plan->k3k_state;
    if (!st || st->version != plan->coef_version || !st->ok) return false;
    sigc = ksig(st->axis[0], st->axis[1], st->axis[2]);
  }
  const KList<T> l = list_for<T>(N);
  for (int i = 0; i < l.n; ++i) {
    const KEntry<T>& k = l.e[i];
    if (k.lanes != lanes || k.adv != (int)adv || k.mass != (int)mass || k.sig != sigc) continue;
    *fn = k.fn;
    prm.E = (int)plan->E;
    prm.M0 = (int)plan->desc.sizes[0];
    prm.M1 = (int)plan->desc.sizes[1];
    prm.M2 = (int)plan->desc.sizes[2];
    prm.periodic = 0;
    double detJ = 1.0;
    for (int d = 0; d < 3; ++d) {
      if (plan->desc.periodic[d]) prm.periodic |= 1 << d;
      detJ *= plan->desc.h[d];
    }
    prm.bc_zero = plan->bc_mode == GD_BC_INFLOW ? 1 : 0;
    prm.negmask = st ? st->negmask : 0;
    prm.num_blocks = 0;
    for (int d = 0; d < 3; ++d) {
      // adv_coef is folded in by apply()
      double K = detJ / plan->desc.h[d];
      if (st && st->axis[d] < 0) K *= st->cval[d];
      prm.Kt[d] = (T)K;
      prm.Kc[d] = (T)K;
      prm.tab[d] = st ? static_cast<const T*>(st->tab[slot][d]) : nullptr;
      prm.fsamp[d] = st ? static_cast<const T*>(st->fs[slot][d]) : nullptr;
    }
    prm.Km = (T)(detJ * sigma);
    return true;
  }
  return false;
}

}  // namespace

template <typename T>
bool available(const gd_plan* plan, int lanes, bool adv, bool mass)
{
  KParams<T> prm{};
  KLaunch<T> fn = nullptr;
  return prepare<T>(plan, lanes, adv, mass, prm, &fn);
}

template <typename T>
int apply(const gd_plan* plan, int lanes, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  KParams<T> prm{};
  KLaunch<T> fn = nullptr;
  const bool adv = adv_coef != T(0), mass = mass_coef != T(0);
  if (!prepare<T>(plan, lanes, adv, mass, prm, &fn)) return GD_ERR_UNSUPPORTED;
  // the LDS-staged variants move 16-byte chunks
  if (lanes >= 32 && ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) != 0)
    return GD_ERR_UNSUPPORTED;
  for (int d = 0; d < 3; ++d) {
    prm.Kt[d] *= adv_coef;
    prm.Kc[d] *= adv_coef;
  }
  prm.Km *= mass_coef;
  return fn(prm, x, y, s);
}

template <typename T>
int apply_faces(const gd_plan* plan, const T* x, T* y, hipStream_t s)
{
  KParams<T> prm{};
  KLaunch<T> fn = nullptr;
  if (!prepare<T>(plan, 42, true, false, prm, &fn)) return GD_ERR_UNSUPPORTED;
  if (((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) != 0) return GD_ERR_UNSUPPORTED;
  // the interior faces alone: no cell terms, no faces across a periodic
  // wrap, no boundary faces
  prm.periodic = 0;
  prm.bc_zero = 0;
  for (int d = 0; d < 3; ++d) prm.Kc[d] = T(0);
  return fn(prm, x, y, s);
}

template bool available<double>(const gd_plan*, int, bool, bool);
template int apply_faces<double>(const gd_plan*, const double*, double*, hipStream_t);
template bool available<float>(const gd_plan*, int, bool, bool);
template int apply<double>(const gd_plan*, int, double, double, const double*, double*, hipStream_t);
template int apply<float>(const gd_plan*, int, float, float, const float*, float*, hipStream_t);

}  // namespace k3k
}  // namespace gd
