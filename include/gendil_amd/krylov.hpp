// Device-resident Krylov solvers over the C-ABI (header-only C++20).
//
// gmres_solve<TV>: restarted GMRES(m) in the shape of GenDiL's GMRES_no_alloc
// (include/gendil/Solvers/gmres.hpp:74-288): every cycle starts from the true
// residual r = b - K x computed in FP64, builds an Arnoldi basis of at most m
// vectors, applies Givens rotations to the Hessenberg column on the host
// (gmres.hpp:170-210), and ends with x += V y (gmres.hpp:265).  Differences
// from the reference, all on the device side:
//   * orthogonalisation is one fused CGS2 step per iteration (three passes
//     over the basis) instead of m+1 sequential MGS dot/axpy pairs
//     (gmres.hpp:160-165); every inner product accumulates in FP64;
//   * basis vectors are stored unnormalised with a host-side scale;
//   * TV = float stores the basis in FP32 and applies the operator in FP32,
//     the cycle's starting residual, the inner products, the Hessenberg /
//     Givens arithmetic and the update of x stay FP64.  Each cycle is then
//     one step of iterative refinement: x_{k+1} = x_k + d, d from an FP32
//     Krylov space built on r_k = b - K x_k (FP64).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gendil_amd.h"

namespace gendil_amd {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what)
{
  if (rc != GD_OK) throw Error(rc, std::string(what) + ": " + gd_last_error());
}

template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) : n_(n)
  {
    if (n) check(gd_malloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "gd_malloc");
  }
  ~DeviceArray()
  {
    if (p_) (void)gd_free(p_);
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_)
  {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept
  {
    if (this != &o) {
      if (p_) (void)gd_free(p_);
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// The operator K = mass_coef * M + adv_coef * A of one plan, applied in
// precision T (y overwritten).
struct CombinedOperator {
  const gd_plan* plan;
  double mass_coef;
  double adv_coef;
  gd_stream stream;
  void apply(double scale, const double* x, double* y) const
  {
    check(gd_apply_f64(plan, scale * mass_coef, scale * adv_coef, x, y, stream), "gd_apply_f64");
  }
  void apply(double scale, const float* x, float* y) const
  {
    check(gd_apply_f32(plan, (float)(scale * mass_coef), (float)(scale * adv_coef), x, y, stream),
          "gd_apply_f32");
  }
};

struct SolveOptions {
  int restart = 30;
  int max_iters = 10000;
  double rel_tol = 1e-10;
  int max_refinements = 1000;
  // FP32 basis: an Arnoldi cycle stops once its estimated residual falls
  // below max(inner_rel_tol * ||r_k||, rel_tol * ||b||).
  double inner_rel_tol = 1e-6;
};

struct SolveReport {
  bool success = false;
  int iterations = 0;
  int cycles = 0;
  double final_rel_residual = 0.0;
  double seconds = 0.0;
};

namespace detail {

inline double dev_dot(int64_t n, const double* a, const double* b, double* scratch, gd_stream s)
{
  check(gd_multidot_f64(n, 1, a, n, b, scratch, s), "gd_multidot_f64");
  double h = 0.0;
  check(gd_memcpy_d2h(&h, scratch, sizeof(double)), "gd_memcpy_d2h");
  return h;
}

template <typename TV>
int cgs2(int64_t n, int k, const TV* V, int64_t ldv, TV* w, const double* s, double* c, double* hcol,
         double* norm2, gd_stream st)
{
  if constexpr (std::is_same_v<TV, double>) return gd_cgs2_step_f64(n, k, V, ldv, w, s, c, hcol, norm2, st);
  else return gd_cgs2_step_f32(n, k, V, ldv, w, s, c, hcol, norm2, st);
}

template <typename TV>
int basis_update(int64_t n, int k, const double* c, const TV* V, int64_t ldv, double* x, gd_stream st)
{
  if constexpr (std::is_same_v<TV, double>) return gd_basis_update_f64(n, k, c, V, ldv, x, st);
  else return gd_basis_update_f32(n, k, c, V, ldv, x, st);
}

}  // namespace detail

// Solve K x = b.  b, x: FP64 device vectors of length n; x holds x0.
template <typename TV>
SolveReport gmres_solve(const CombinedOperator& K, int64_t n, const double* b, double* x, const SolveOptions& opt)
{
  if (opt.restart < 1 || opt.restart > 32) throw Error(GD_ERR_INVALID, "gmres: restart must be 1..32");
  const auto t0 = std::chrono::steady_clock::now();
  const int m = opt.restart;
  const gd_stream st = K.stream;
  DeviceArray<TV> V((size_t)(m + 1) * n);
  DeviceArray<double> r(n);
  DeviceArray<double> small(4 * (m + 2) + 8);  // s | c | hcol | norm2 | dot scratch
  double* d_s = small.data();
  double* d_c = d_s + (m + 2);
  double* d_h = d_c + (m + 2);
  double* d_n2 = d_h + (m + 2);
  double* d_dot = d_n2 + 2;

  std::vector<double> s(m + 1, 1.0), H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), hcol(m + 2);
  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, d_dot, st));
  if (bnorm == 0.0) {  // zero right-hand side: x = 0 (gmres.hpp:106-113)
    check(gd_axpby_f64(n, 0.0, b, 0.0, x, st), "gd_axpby_f64");
    check(gd_stream_sync(st), "sync");
    rep.success = true;
    return rep;
  }
  while (true) {
    // true residual in FP64 (gmres.hpp:98-104, 268-275)
    K.apply(1.0, x, r.data());
    check(gd_axpby_f64(n, 1.0, b, -1.0, r.data(), st), "gd_axpby_f64");
    const double beta = std::sqrt(detail::dev_dot(n, r.data(), r.data(), d_dot, st));
    rep.final_rel_residual = beta / bnorm;
    if (rep.final_rel_residual <= opt.rel_tol) {
      rep.success = true;
      break;
    }
    if (rep.iterations >= opt.max_iters || rep.cycles >= opt.max_refinements) break;
    ++rep.cycles;
    // V_0 = r (stored unnormalised, scale 1/beta)
    if constexpr (std::is_same_v<TV, double>) check(gd_axpby_f64(n, 1.0, r.data(), 0.0, V.data(), st), "copy");
    else check(gd_convert_f64_to_f32(n, 1.0 / beta, r.data(), V.data(), st), "convert");
    s[0] = std::is_same_v<TV, double> ? 1.0 / beta : 1.0;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    double target = opt.rel_tol * bnorm;
    if constexpr (!std::is_same_v<TV, double>) target = std::max(target, opt.inner_rel_tol * beta);
    int k = 0;
    for (; k < m && rep.iterations < opt.max_iters; ++k) {
      ++rep.iterations;
      TV* w = V.data() + (size_t)(k + 1) * n;
      K.apply(s[k], V.data() + (size_t)k * n, w);
      check(gd_memcpy_h2d(d_s, s.data(), (k + 1) * sizeof(double)), "h2d");
      check(detail::cgs2<TV>(n, k + 1, V.data(), n, w, d_s, d_c, d_h, d_n2, st), "cgs2");
      check(gd_memcpy_d2h(hcol.data(), d_h, (k + 1) * sizeof(double)), "d2h");
      double n2 = 0.0;
      check(gd_memcpy_d2h(&n2, d_n2, sizeof(double)), "d2h");
      const double hk1 = std::sqrt(n2);
      for (int i = 0; i <= k; ++i) H[(size_t)i * m + k] = hcol[i];
      H[(size_t)(k + 1) * m + k] = hk1;
      s[k + 1] = hk1 > 0.0 ? 1.0 / hk1 : 0.0;
      // apply previous rotations, then form the new one (gmres.hpp:170-210)
      for (int i = 0; i < k; ++i) {
        const double a = H[(size_t)i * m + k], c2 = H[(size_t)(i + 1) * m + k];
        H[(size_t)i * m + k] = cs[i] * a + sn[i] * c2;
        H[(size_t)(i + 1) * m + k] = -sn[i] * a + cs[i] * c2;
      }
      const double a = H[(size_t)k * m + k], c2 = H[(size_t)(k + 1) * m + k];
      const double rho = std::hypot(a, c2);
      cs[k] = rho > 0.0 ? a / rho : 1.0;
      sn[k] = rho > 0.0 ? c2 / rho : 0.0;
      H[(size_t)k * m + k] = rho;
      H[(size_t)(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) <= target || hk1 == 0.0) {  // converged or happy breakdown
        ++k;
        break;
      }
    }
    // back substitution H y = g, then x += V (s y)  (gmres.hpp:240-265)
    for (int i = k - 1; i >= 0; --i) {
      double acc = g[i];
      for (int j = i + 1; j < k; ++j) acc -= H[(size_t)i * m + j] * y[j];
      y[i] = acc / H[(size_t)i * m + i];
    }
    std::vector<double> coef(k);
    for (int j = 0; j < k; ++j) coef[j] = s[j] * y[j];
    check(gd_memcpy_h2d(d_c, coef.data(), k * sizeof(double)), "h2d");
    check(detail::basis_update<TV>(n, k, d_c, V.data(), n, x, st), "basis_update");
  }
  check(gd_stream_sync(st), "sync");
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

// Unpreconditioned CG (conjugategradient.hpp:217) for SPD K, FP64.
inline SolveReport cg_solve(const CombinedOperator& K, int64_t n, const double* b, double* x,
                            const SolveOptions& opt)
{
  const auto t0 = std::chrono::steady_clock::now();
  const gd_stream st = K.stream;
  DeviceArray<double> r(n), p(n), Ap(n), tmp(4);
  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, tmp.data(), st));
  K.apply(1.0, x, r.data());
  check(gd_axpby_f64(n, 1.0, b, -1.0, r.data(), st), "axpby");
  check(gd_axpby_f64(n, 1.0, r.data(), 0.0, p.data(), st), "copy");
  double rr = detail::dev_dot(n, r.data(), r.data(), tmp.data(), st);
  if (bnorm == 0.0) {
    rep.success = true;
    return rep;
  }
  while (rep.iterations < opt.max_iters) {
    rep.final_rel_residual = std::sqrt(rr) / bnorm;
    if (rep.final_rel_residual <= opt.rel_tol) {
      rep.success = true;
      break;
    }
    ++rep.iterations;
    K.apply(1.0, p.data(), Ap.data());
    const double pAp = detail::dev_dot(n, p.data(), Ap.data(), tmp.data(), st);
    const double alpha = rr / pAp;
    check(gd_axpby_f64(n, alpha, p.data(), 1.0, x, st), "axpby");
    check(gd_axpby_f64(n, -alpha, Ap.data(), 1.0, r.data(), st), "axpby");
    const double rr_new = detail::dev_dot(n, r.data(), r.data(), tmp.data(), st);
    check(gd_axpby_f64(n, 1.0, r.data(), rr_new / rr, p.data(), st), "axpby");
    rr = rr_new;
  }
  // report the true residual
  K.apply(1.0, x, r.data());
  check(gd_axpby_f64(n, 1.0, b, -1.0, r.data(), st), "axpby");
  rep.final_rel_residual = std::sqrt(detail::dev_dot(n, r.data(), r.data(), tmp.data(), st)) / bnorm;
  rep.success = rep.final_rel_residual <= opt.rel_tol * 10.0;
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

}  // namespace gendil_amd
