// C-ABI operator applies: argument checks, kernel lookup and launch.
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "gd_dispatch.hpp"
#include "gd_k3d_api.hpp"
#include "gd_k3k_api.hpp"
#include "gd_n2k_api.hpp"
#include "gd_padapt_api.hpp"

namespace gd {

const KernelSet* find_kernels(int D, int N, int Q)
{
  KernelList l{nullptr, 0};
  switch (D) {
    case 1: l = kernels_d1(); break;
    case 2: l = kernels_d2(); break;
    case 3: l = kernels_d3(); break;
    case 4: l = kernels_d4(); break;
    case 5: l = kernels_d5(); break;
    case 6: l = kernels_d6(); break;
    default: break;
  }
  for (int i = 0; i < l.count; ++i)
    if (l.sets[i].N == N && l.sets[i].Q == Q) return &l.sets[i];
  return nullptr;
}

namespace {

int tuned_lanes(const gd_plan* plan, int /*N*/, size_t bytes) { return plan->tuned_k3_lanes[bytes == 8 ? 0 : 1]; }

int lookup(const gd_plan* plan, const KernelSet** ks)
{
  if (!plan) return fail(GD_ERR_INVALID, "null plan");
  if (plan->padapt) return fail(GD_ERR_UNSUPPORTED, "p-adaptive plans support the forward apply only");
  const int N = plan->desc.order + 1, Q = plan->desc.num_quad;
  *ks = find_kernels(plan->dim, N, Q);
  if (!*ks)
    return fail(GD_ERR_UNSUPPORTED, "no kernel compiled for dim=" + std::to_string(plan->dim) +
                                        " nodes=" + std::to_string(N) + " quad=" + std::to_string(Q));
  return GD_OK;
}

int check_adv(const gd_plan* plan, bool need_pointwise)
{
  if (!plan->beta_set) return fail(GD_ERR_INVALID, "advection field not set");
  if (need_pointwise && plan->beta_kind != GD_BETA_POINTWISE)
    return fail(GD_ERR_INVALID, "gradient w.r.t. beta needs pointwise beta");
  return GD_OK;
}

// 3D kernel choice: from the tuning table (autotuner or gd_plan_set_variant).
// 0 = generic kernel; 1, 2, 4 = register-resident kernel
// with that many lanes per element; 16 + lanes = Kronecker-factored kernel
// (lanes 1, 2, 4: register-resident; 32..39: LDS-staged, bit 0 = outputs
// staged through LDS, bit 1 = 128-thread blocks, bit 2 = one term at a time;
// 40, 41: streaming with one or two tile buffers).  Default: the streaming
// Kronecker kernel with one buffer where it applies
// (n = 4, rows a multiple of 16 elements; measured on MI355X at 64^3 p=3:
// FP64 67 us against 97 us for the staged kernel and 77-85 us with two
// buffers, FP32 34 against 45 us),
// else the staged Kronecker kernel, else the register Kronecker kernel, else
// the register kernel.
template <typename T>
int k3_choice(const gd_plan* plan, int N, bool adv, bool mass)
{
  const int tuned = tuned_lanes(plan, N, sizeof(T));
  if (tuned >= 0 && tuned != n2k::kVariant) return tuned;
  if (N == 4 && k3k::available<T>(plan, 40, adv, mass)) return 16 + 40;
  const int ks = k3k::default_staged(N, sizeof(T));
  if (ks >= 0 && k3k::available<T>(plan, ks, adv, mass)) return 16 + ks;
  const int kl = k3k::default_lanes(N, sizeof(T));
  if (k3k::available<T>(plan, kl, adv, mass)) return 16 + kl;
  // measured on MI355X at 64^3 p=3 (a one-axis field the Kronecker kernels
  // do not take): 2 lanes 171 us FP64 / 88 us FP32, 4 lanes 209 / 94 us
  if (N == 4) return 2;
  return 1;
}

// The register-resident kernel covers 3D GLL elements with q = p+1 and
// one-axis separable (or constant) coefficients; see gd_k3d_host.cpp.
template <typename T>
bool k3_eligible(const gd_plan* plan, int choice, bool adv, bool mass)
{
  if (choice == n2k::kVariant) return n2k::available<T>(plan, adv, mass);
  if (choice >= 16) return k3k::available<T>(plan, choice - 16, adv, mass);
  return k3::available<T>(plan, choice, adv, mass);
}

template <typename T>
int padapt_apply(const gd_plan* plan, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  if (!x || !y) return fail(GD_ERR_INVALID, "null vector");
  if ((const void*)x == (const void*)y) return fail(GD_ERR_INVALID, "input and output must not alias");
  if (adv_coef != T(0)) {
    const int rc = check_adv(plan, false);
    if (rc) return rc;
  }
  return pa::apply<T>(plan, plan->tuned_pa[sizeof(T) == 8 ? 0 : 1], mass_coef, adv_coef, x, y, s);
}

template <typename T>
int apply_impl(const gd_plan* plan, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  if (plan && plan->padapt) return padapt_apply<T>(plan, mass_coef, adv_coef, x, y, s);
  const KernelSet* ks = nullptr;
  int rc = lookup(plan, &ks);
  if (rc) return rc;
  if (!x || !y) return fail(GD_ERR_INVALID, "null vector");
  if ((const void*)x == (const void*)y) return fail(GD_ERR_INVALID, "input and output must not alias");
  const bool adv = adv_coef != T(0);
  if (adv) {
    rc = check_adv(plan, false);
    if (rc) return rc;
  }
  if constexpr (std::is_same_v<T, float>) {
    if (adv && plan->beta_kind == GD_BETA_POINTWISE)
      return fail(GD_ERR_UNSUPPORTED, "pointwise beta is FP64 only");
  }
  // 4D-6D, order 1, constant coefficients: the two-node Kronecker kernel
  // unless the generic kernel (0) is pinned (measured on MI355X for GenDiL's
  // 6D benchmark operator, 5^6 elements: see n2k)
  if (plan->dim >= 4 && tuned_lanes(plan, 2, sizeof(T)) != 0 && n2k::available<T>(plan, adv, mass_coef != T(0))) {
    const int r = n2k::apply<T>(plan, mass_coef, adv_coef, x, y, s);
    if (r != GD_ERR_UNSUPPORTED) return r;
  }
  {
    const int N = plan->desc.order + 1;
    const int choice = k3_choice<T>(plan, N, adv, mass_coef != T(0));
    if (choice >= 16) {
      int r = k3k::apply<T>(plan, choice - 16, mass_coef, adv_coef, x, y, s);
      if (r != GD_ERR_UNSUPPORTED) return r;
      if (choice >= 48) {  // staged variant refused (e.g. unaligned vectors): register variant
        r = k3k::apply<T>(plan, k3k::default_lanes(N, sizeof(T)), mass_coef, adv_coef, x, y, s);
        if (r != GD_ERR_UNSUPPORTED) return r;
      }
    }
    const int lanes = choice >= 16 ? (N == 4 ? 2 : 1) : choice;
    if (lanes > 0) {
      const int r = k3::apply<T>(plan, lanes, mass_coef, adv_coef, x, y, s);
      if (r != GD_ERR_UNSUPPORTED) return r;
    }
  }
  // the generic kernel (its parameter block is built only here: the kernels
  // above take their own)
  const TableSet* ts = find_tables(plan, plan->desc.order + 1, plan->desc.num_quad);
  const Params<T> prm = make_params<T>(plan, ts, mass_coef, adv_coef);
  if constexpr (std::is_same_v<T, double>) return adv ? ks->fwd64_adv(prm, x, y, s) : ks->fwd64_mass(prm, x, y, s);
  else return adv ? ks->fwd32_adv(prm, x, y, s) : ks->fwd32_mass(prm, x, y, s);
}

}  // namespace
}  // namespace gd

using gd::fail;

extern "C" {

int gd_apply_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* x, double* y,
                 gd_stream stream)
{
  return gd::apply_impl<double>(plan, mass_coef, adv_coef, x, y, (hipStream_t)stream);
}

int gd_apply_f32(const gd_plan* plan, float mass_coef, float adv_coef, const float* x, float* y,
                 gd_stream stream)
{
  return gd::apply_impl<float>(plan, mass_coef, adv_coef, x, y, (hipStream_t)stream);
}

int gd_mass_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream)
{
  return gd_apply_f64(plan, 1.0, 0.0, x, y, stream);
}

int gd_mass_apply_f32(const gd_plan* plan, const float* x, float* y, gd_stream stream)
{
  return gd_apply_f32(plan, 1.0f, 0.0f, x, y, stream);
}

int gd_adv_apply_f64(const gd_plan* plan, const double* u, double* y, gd_stream stream)
{
  return gd_apply_f64(plan, 0.0, 1.0, u, y, stream);
}

int gd_adv_apply_f32(const gd_plan* plan, const float* u, float* y, gd_stream stream)
{
  return gd_apply_f32(plan, 0.0f, 1.0f, u, y, stream);
}

int gd_apply_T_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* r, double* z,
                   gd_stream stream)
{
  const gd::KernelSet* ks = nullptr;
  int rc = gd::lookup(plan, &ks);
  if (rc) return rc;
  if (!r || !z) return fail(GD_ERR_INVALID, "null vector");
  if ((const void*)r == (const void*)z) return fail(GD_ERR_INVALID, "input and output must not alias");
  const gd::TableSet* ts = gd::find_tables(plan, plan->desc.order + 1, plan->desc.num_quad);
  const gd::Params<double> prm = gd::make_params<double>(plan, ts, mass_coef, adv_coef);
  if (adv_coef == 0.0) return ks->fwd64_mass(prm, r, z, (hipStream_t)stream);  // M is symmetric
  rc = gd::check_adv(plan, false);
  if (rc) return rc;
  return ks->trn64(prm, r, z, (hipStream_t)stream);
}

int gd_plan_autotune(gd_plan* plan, int32_t precision_bytes, double mass_coef, double adv_coef, int32_t reps,
                     int32_t* chosen, double* best_ms)
{
  if (!plan || (precision_bytes != 4 && precision_bytes != 8)) return fail(GD_ERR_INVALID, "autotune: bad arguments");
  if (reps < 1) reps = 5;
  const int slot = precision_bytes == 8 ? 0 : 1;
  if (plan->padapt) {
    const char* why = nullptr;
    if (!gd::pa::eligible(plan, &why)) return fail(GD_ERR_UNSUPPORTED, why);
  }
  const size_t bytes = (size_t)plan->ndofs * precision_bytes;
  void *x = nullptr, *y = nullptr;
  if (hipMalloc(&x, bytes) != hipSuccess || hipMalloc(&y, bytes) != hipSuccess) {
    if (x) (void)hipFree(x);
    return fail(GD_ERR_ALLOC, "autotune: scratch allocation failed");
  }
  (void)hipMemset(x, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int* knob = plan->padapt ? &plan->tuned_pa[slot] : &plan->tuned_k3_lanes[slot];
  std::vector<int> choices;
  if (plan->padapt)
    for (int v = 0; v < gd::pa::num_variants(plan->dim); ++v) choices.push_back(v);
  else
    choices = {0, 1, 2, 4, 17, 18, 20, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, gd::n2k::kVariant};
  const int saved = *knob;
  double best = 1e300;
  int best_lanes = 0;
  int rc = GD_OK;
  for (int lanes : choices) {
    *knob = lanes;
    auto run = [&]() -> int {
      if (precision_bytes == 8)
        return gd_apply_f64(plan, mass_coef, adv_coef, (const double*)x, (double*)y, nullptr);
      return gd_apply_f32(plan, (float)mass_coef, (float)adv_coef, (const float*)x, (float*)y, nullptr);
    };
    if (lanes > 0 && !plan->padapt) {
      const bool ok = precision_bytes == 8
                          ? gd::k3_eligible<double>(plan, lanes, adv_coef != 0.0, mass_coef != 0.0)
                          : gd::k3_eligible<float>(plan, lanes, adv_coef != 0.0, mass_coef != 0.0);
      if (!ok) continue;
    }
    rc = run();  // warm-up
    if (rc) break;
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && rc == GD_OK; ++r) rc = run();
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    if (rc) break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / reps;
    if (per < best) {
      best = per;
      best_lanes = lanes;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(x);
  (void)hipFree(y);
  if (rc) {
    *knob = saved;
    return rc;
  }
  *knob = best_lanes;
  if (chosen) *chosen = best_lanes;
  if (best_ms) *best_ms = best;
  return GD_OK;
}

int gd_plan_query_variant(const gd_plan* plan, int32_t precision_bytes, double mass_coef, double adv_coef,
                          int32_t* variant)
{
  if (!plan || !variant || (precision_bytes != 4 && precision_bytes != 8))
    return fail(GD_ERR_INVALID, "gd_plan_query_variant: bad arguments");
  if (plan->padapt) {
    *variant = gd::pa::resolved_variant(plan, plan->tuned_pa[precision_bytes == 8 ? 0 : 1]);
    return GD_OK;
  }
  const bool adv = adv_coef != 0.0, mass = mass_coef != 0.0;
  const int N = plan->desc.order + 1;
  auto resolve = [&](auto tag) -> int {
    using T = decltype(tag);
    if (plan->dim >= 4 && plan->tuned_k3_lanes[sizeof(T) == 8 ? 0 : 1] != 0 &&
        gd::n2k::available<T>(plan, adv, mass))
      return gd::n2k::kVariant;
    const int choice = gd::k3_choice<T>(plan, N, adv, mass);
    if (choice >= 16 && gd::k3k::available<T>(plan, choice - 16, adv, mass)) return choice;
    const int lanes = choice >= 16 ? (N == 4 ? 2 : 1) : choice;
    if (lanes > 0 && gd::k3::available<T>(plan, lanes, adv, mass)) return lanes;
    return 0;
  };
  *variant = precision_bytes == 8 ? resolve(0.0) : resolve(0.0f);
  return GD_OK;
}

int gd_plan_set_variant(gd_plan* plan, int32_t precision_bytes, int32_t lanes)
{
  if (!plan || (precision_bytes != 4 && precision_bytes != 8)) return fail(GD_ERR_INVALID, "gd_plan_set_variant: bad arguments");
  const int slot = precision_bytes == 8 ? 0 : 1;
  if (plan->padapt) {
    if (lanes < -1 || lanes >= gd::pa::num_variants(plan->dim))
      return fail(GD_ERR_INVALID, "gd_plan_set_variant: no such p-adaptive variant");
    plan->tuned_pa[slot] = lanes;
    return GD_OK;
  }
  if (lanes != -1 && lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 17 && lanes != 18 &&
      lanes != 20 && !(lanes >= 48 && lanes <= 57) && lanes != gd::n2k::kVariant)
    return fail(GD_ERR_INVALID, "gd_plan_set_variant: bad arguments");
  plan->tuned_k3_lanes[slot] = lanes;
  return GD_OK;
}

int gd_adv_apply_T_f64(const gd_plan* plan, const double* r, double* z, gd_stream stream)
{
  return gd_apply_T_f64(plan, 0.0, 1.0, r, z, stream);
}

int gd_adv_grad_beta_f64(const gd_plan* plan, double adv_coef, const double* u, const double* r,
                         double* dbeta_cell, double* const* dbeta_face, gd_stream stream)
{
  const gd::KernelSet* ks = nullptr;
  int rc = gd::lookup(plan, &ks);
  if (rc) return rc;
  rc = gd::check_adv(plan, true);
  if (rc) return rc;
  if (!u || !r || !dbeta_cell || !dbeta_face) return fail(GD_ERR_INVALID, "null argument");
  const gd::TableSet* ts = gd::find_tables(plan, plan->desc.order + 1, plan->desc.num_quad);
  gd::Params<double> prm = gd::make_params<double>(plan, ts, 0.0, adv_coef);
  prm.dbeta_cell = dbeta_cell;
  for (int d = 0; d < plan->dim; ++d) {
    if (!dbeta_face[d]) return fail(GD_ERR_INVALID, "null face gradient array");
    prm.dbeta_face[d] = dbeta_face[d];
  }
  // the kernel reads r through its second vector argument
  return ks->grad64(prm, u, const_cast<double*>(r), (hipStream_t)stream);
}

}  // extern "C"
