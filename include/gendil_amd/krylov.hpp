// Device-resident Krylov solvers over the C-ABI (header-only C++20).
//
// gmres_solve<TV>: restarted GMRES(m) in the shape of GenDiL's GMRES_no_alloc
// (include/gendil/Solvers/gmres.hpp:74-288): every cycle starts from the true
// residual r = b - K x computed in FP64, builds an Arnoldi basis of at most m
// vectors, applies Givens rotations to the Hessenberg column
// (gmres.hpp:170-210), and ends with x += V y (gmres.hpp:265).  Differences
// from the reference:
//   * the Arnoldi process runs on the device (gd_arnoldi_step_*): one fused
//     CGS2 orthogonalisation per iteration (three passes over the basis)
//     instead of m+1 sequential MGS dot/axpy pairs (gmres.hpp:160-165), every
//     inner product accumulated in FP64, and the Hessenberg column, Givens
//     rotation and triangular solve done on the device too;
//   * the host reads one residual estimate per iteration from pinned memory,
//     one iteration behind the device, so the stream never drains;
//   * TV = float stores the basis in FP32 and applies the operator in FP32,
//     the cycle's starting residual, the inner products, the Hessenberg /
//     Givens arithmetic and the update of x stay FP64.  Each cycle is then
//     one step of iterative refinement: x_{k+1} = x_k + d, d from an FP32
//     Krylov space built on r_k = b - K x_k (FP64).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
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

// Pinned host array (the device writes per-iteration results into it).
template <typename T>
class PinnedArray {
 public:
  explicit PinnedArray(size_t n) { check(gd_host_alloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "gd_host_alloc"); }
  ~PinnedArray()
  {
    if (p_) (void)gd_host_free(p_);
  }
  PinnedArray(const PinnedArray&) = delete;
  PinnedArray& operator=(const PinnedArray&) = delete;
  T& operator[](size_t i) { return p_[i]; }
  T* data() { return p_; }

 private:
  T* p_ = nullptr;
};

class Event {
 public:
  Event() { check(gd_event_create(&e_), "gd_event_create"); }
  ~Event()
  {
    if (e_) (void)gd_event_destroy(e_);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  void record(gd_stream s) { check(gd_event_record(e_, s), "gd_event_record"); }
  void sync() { check(gd_event_sync(e_), "gd_event_sync"); }

 private:
  gd_event e_ = nullptr;
};

namespace detail {

inline double dev_dot(int64_t n, const double* a, const double* b, double* scratch, gd_stream s)
{
  check(gd_multidot_f64(n, 1, a, n, b, scratch, s), "gd_multidot_f64");
  // gd_memcpy_d2h runs on the null stream, which is not ordered after a
  // non-blocking stream: wait for the dot product on its own stream first
  check(gd_stream_sync(s), "gd_stream_sync");
  double h = 0.0;
  check(gd_memcpy_d2h(&h, scratch, sizeof(double)), "gd_memcpy_d2h");
  return h;
}

template <typename TV>
int arnoldi_step(int64_t n, int m, int k, TV* V, int64_t ldv, void* ws, double* resid, gd_stream st)
{
  if constexpr (std::is_same_v<TV, double>) return gd_arnoldi_step_f64(n, m, k, V, ldv, ws, resid, st);
  else return gd_arnoldi_step_f32(n, m, k, V, ldv, ws, resid, st);
}

template <typename TV>
int update(int64_t n, int m, int k, const TV* V, int64_t ldv, void* ws, double* x, gd_stream st)
{
  if constexpr (std::is_same_v<TV, double>) return gd_gmres_update_f64(n, m, k, V, ldv, ws, x, st);
  else return gd_gmres_update_f32(n, m, k, V, ldv, ws, x, st);
}

// Waits until the device has stored a value (>= 0, or NaN) over the negative
// marker at *p, checking the stream for errors now and then.
inline double wait_for_value(const double* p, gd_stream st)
{
  for (unsigned spin = 1;; ++spin) {
    const double v = *static_cast<const volatile double*>(p);
    if (!(v < 0.0)) return v;
    if (spin % 4096 == 0) {
      int32_t done = 0;
      check(gd_stream_query(st, &done), "gd_stream_query");
      if (done) {
        const double w = *static_cast<const volatile double*>(p);
        if (!(w < 0.0)) return w;
        throw Error(GD_ERR_HIP, "gmres: the stream finished without storing the residual estimate");
      }
    }
  }
}

}  // namespace detail

// Solve K x = b.  b, x: FP64 device vectors of length n; x holds x0.
//
// The host never waits inside a cycle for the iteration it has just queued:
// iteration k is enqueued before the residual estimate of iteration k-1 is
// read (polled in pinned memory, where the device stores it; no event or copy
// enters the stream), so the stream always holds work.  When iteration k-1
// meets the tolerance, the speculative iteration k is discarded (its column is
// not used by the update).  The basis leading dimension is rounded up to 4
// values so every vector starts 16-byte aligned.
template <typename TV>
SolveReport gmres_solve(const CombinedOperator& K, int64_t n, const double* b, double* x, const SolveOptions& opt)
{
  if (opt.restart < 1 || opt.restart > 32) throw Error(GD_ERR_INVALID, "gmres: restart must be 1..32");
  const auto t0 = std::chrono::steady_clock::now();
  const int m = opt.restart;
  const gd_stream st = K.stream;
  const int64_t ldv = (n + 3) / 4 * 4;
  DeviceArray<TV> V((size_t)(m + 1) * ldv);
  DeviceArray<double> r(n), dot(2);
  DeviceArray<unsigned char> ws(gd_gmres_workspace_bytes());
  PinnedArray<double> resid(m);

  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, dot.data(), st));
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
    const double beta = std::sqrt(detail::dev_dot(n, r.data(), r.data(), dot.data(), st));
    rep.final_rel_residual = beta / bnorm;
    if (rep.final_rel_residual <= opt.rel_tol) {
      rep.success = true;
      break;
    }
    if (rep.iterations >= opt.max_iters || rep.cycles >= opt.max_refinements) break;
    ++rep.cycles;
    // V_0 = r / beta
    if constexpr (std::is_same_v<TV, double>) check(gd_axpby_f64(n, 1.0 / beta, r.data(), 0.0, V.data(), st), "scale");
    else check(gd_convert_f64_to_f32(n, 1.0 / beta, r.data(), V.data(), st), "convert");
    check(gd_gmres_begin(ws.data(), beta, 1.0, st), "gd_gmres_begin");
    double target = opt.rel_tol * bnorm;
    if constexpr (!std::is_same_v<TV, double>) target = std::max(target, opt.inner_rel_tol * beta);
    const int budget = std::min(m, opt.max_iters - rep.iterations);
    int used = budget;  // columns that enter the update
    for (int k = 0; k < budget; ++k) {
      TV* vk = V.data() + (size_t)k * ldv;
      resid[k] = -1.0;  // overwritten by the device when iteration k ends
      K.apply(1.0, vk, vk + ldv);
      check(detail::arnoldi_step<TV>(n, m, k, V.data(), ldv, ws.data(), resid.data() + k, st), "gd_arnoldi_step");
      if (k >= 1) {
        const double e = detail::wait_for_value(resid.data() + k - 1, st);
        if (e <= target) {  // converged (or happy breakdown) at iteration k-1
          used = k;
          break;
        }
      }
    }
    rep.iterations += used;
    check(detail::update<TV>(n, m, used, V.data(), ldv, ws.data(), x, st), "gd_gmres_update");
  }
  check(gd_stream_sync(st), "sync");
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

namespace detail {

// true relative residual ||b - K x|| / ||b|| (synchronous)
inline double true_residual(const CombinedOperator& K, int64_t n, const double* b, const double* x, double* r,
                            double* scratch, double bnorm)
{
  K.apply(1.0, x, r);
  check(gd_axpby_f64(n, 1.0, b, -1.0, r, K.stream), "gd_axpby_f64");
  return std::sqrt(dev_dot(n, r, r, scratch, K.stream)) / bnorm;
}

// The device-resident iterations: the host issues iteration k, then waits for
// the residual estimate of iteration k-1 (so the stream always holds the next
// iteration); the device itself stops updating once converged or stopped.
template <typename Issue>
void drive(int max_iters, double rel_tol, bool strict, double* resid, gd_stream st, Issue&& issue)
{
  for (int k = 0; k < max_iters; ++k) {
    resid[k % 2] = -1.0;
    issue(k, resid + k % 2);
    if (k >= 1) {
      const double e = wait_for_value(resid + (k - 1) % 2, st);
      if (std::isnan(e) || (strict ? e < rel_tol : e <= rel_tol)) break;
    }
  }
}

}  // namespace detail

namespace detail {

inline int cg_begin(int64_t n, const double* r, double bnorm, double tol, void* ws, gd_stream st)
{
  return gd_cg_begin_f64(n, r, bnorm, tol, ws, st);
}
inline int cg_begin(int64_t n, const float* r, double bnorm, double tol, void* ws, gd_stream st)
{
  return gd_cg_begin_f32(n, r, bnorm, tol, ws, st);
}
inline int cg_step(int64_t n, double* x, double* r, double* p, const double* ap, void* ws, double* out, gd_stream st)
{
  return gd_cg_step_f64(n, x, r, p, ap, ws, out, st);
}
inline int cg_step(int64_t n, float* x, float* r, float* p, const float* ap, void* ws, double* out, gd_stream st)
{
  return gd_cg_step_f32(n, x, r, p, ap, ws, out, st);
}
inline int copy(int64_t n, const double* a, double* b, gd_stream st) { return gd_axpby_f64(n, 1.0, a, 0.0, b, st); }
inline int copy(int64_t n, const float* a, float* b, gd_stream st) { return gd_axpby_f32(n, 1.0f, a, 0.0f, b, st); }

// The device-resident CG iterations on K x = rhs from the residual r = rhs -
// K x (r is overwritten, p and ap are scratch), vectors in precision T, until
// ||r|| <= rel_tol * bnorm or `max_iters`; returns (state, iterations).
template <typename T>
std::pair<int32_t, int32_t> cg_iterate(const CombinedOperator& K, int64_t n, T* x, T* r, T* p, T* ap, void* ws,
                                       double* resid, double bnorm, double rel_tol, int max_iters)
{
  const gd_stream st = K.stream;
  check(copy(n, r, p, st), "copy");
  check(cg_begin(n, r, bnorm, rel_tol, ws, st), "gd_cg_begin");
  drive(max_iters, rel_tol, false, resid, st, [&](int, double* out) {
    K.apply(1.0, p, ap);
    check(cg_step(n, x, r, p, ap, ws, out, st), "gd_cg_step");
  });
  int32_t state = 0, iters = 0;
  double rel = 0.0;
  check(gd_ksolve_status(ws, st, &state, &iters, &rel), "gd_ksolve_status");
  return {state, iters};
}

}  // namespace detail

// Unpreconditioned CG (conjugategradient.hpp:217) for SPD K, FP64,
// device-resident (gd_cg_step_f64): no dot product synchronises the stream.
inline SolveReport cg_solve(const CombinedOperator& K, int64_t n, const double* b, double* x,
                            const SolveOptions& opt)
{
  const auto t0 = std::chrono::steady_clock::now();
  const gd_stream st = K.stream;
  DeviceArray<double> r(n), p(n), Ap(n), tmp(4);
  DeviceArray<unsigned char> ws(gd_ksolve_workspace_bytes());
  PinnedArray<double> resid(2);
  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, tmp.data(), st));
  if (bnorm == 0.0) {  // zero right-hand side: x = 0
    check(gd_axpby_f64(n, 0.0, b, 0.0, x, st), "gd_axpby_f64");
    check(gd_stream_sync(st), "sync");
    rep.success = true;
    return rep;
  }
  K.apply(1.0, x, r.data());
  check(gd_axpby_f64(n, 1.0, b, -1.0, r.data(), st), "axpby");
  const auto [state, iters] = detail::cg_iterate<double>(K, n, x, r.data(), p.data(), Ap.data(), ws.data(),
                                                         resid.data(), bnorm, opt.rel_tol, opt.max_iters);
  rep.iterations = iters;
  rep.success = state == 1;
  rep.final_rel_residual = detail::true_residual(K, n, b, x, r.data(), tmp.data(), bnorm);
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

// CG with FP32 inner solves and FP64 iterative refinement: each step forms
// the true residual r = b - K x in FP64, solves K d = r / ||r|| by CG with
// FP32 vectors and the FP32 apply to max(inner_rel_tol, rel_tol ||b|| /
// ||r||), and updates x += ||r|| d in FP64.
inline SolveReport cg_mixed_solve(const CombinedOperator& K, int64_t n, const double* b, double* x,
                                  const SolveOptions& opt)
{
  const auto t0 = std::chrono::steady_clock::now();
  const gd_stream st = K.stream;
  DeviceArray<double> r(n), tmp(4);
  DeviceArray<float> d(n), r32(n), p(n), Ap(n);
  DeviceArray<unsigned char> ws(gd_ksolve_workspace_bytes());
  PinnedArray<double> resid(2);
  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, tmp.data(), st));
  if (bnorm == 0.0) {
    check(gd_axpby_f64(n, 0.0, b, 0.0, x, st), "gd_axpby_f64");
    check(gd_stream_sync(st), "sync");
    rep.success = true;
    return rep;
  }
  double rel = detail::true_residual(K, n, b, x, r.data(), tmp.data(), bnorm);
  while (rel > opt.rel_tol && rep.cycles < opt.max_refinements && rep.iterations < opt.max_iters) {
    const double rnorm = rel * bnorm;
    // r (FP64, already b - K x) -> r32 = r / ||r||, d = 0
    check(gd_convert_f64_to_f32(n, 1.0 / rnorm, r.data(), r32.data(), st), "gd_convert_f64_to_f32");
    check(gd_axpby_f32(n, 0.0f, r32.data(), 0.0f, d.data(), st), "gd_axpby_f32");
    const double inner = std::max(opt.inner_rel_tol, opt.rel_tol * bnorm / rnorm);
    const auto [state, iters] =
        detail::cg_iterate<float>(K, n, d.data(), r32.data(), p.data(), Ap.data(), ws.data(), resid.data(), 1.0,
                                  inner, opt.max_iters - rep.iterations);
    rep.iterations += iters;
    rep.cycles += 1;
    if (iters == 0 && state != 1) break;  // no progress possible in FP32
    check(gd_add_f32_to_f64(n, rnorm, d.data(), x, st), "gd_add_f32_to_f64");
    rel = detail::true_residual(K, n, b, x, r.data(), tmp.data(), bnorm);
  }
  rep.success = rel <= opt.rel_tol;
  rep.final_rel_residual = rel;
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

// Unpreconditioned BiCGSTAB (BiCGSTAB_no_alloc, bicgstab.hpp:44-157), FP64,
// device-resident (gd_bicgstab_*_f64).
inline SolveReport bicgstab_solve(const CombinedOperator& K, int64_t n, const double* b, double* x,
                                  const SolveOptions& opt)
{
  const auto t0 = std::chrono::steady_clock::now();
  const gd_stream st = K.stream;
  DeviceArray<double> r(n), rhat(n), p(n), v(n), s(n), t(n), tmp(4);
  DeviceArray<unsigned char> ws(gd_ksolve_workspace_bytes());
  PinnedArray<double> resid(2);
  SolveReport rep;
  const double bnorm = std::sqrt(detail::dev_dot(n, b, b, tmp.data(), st));
  if (bnorm == 0.0) {  // zero right-hand side: x = 0 (bicgstab.hpp:74-78)
    check(gd_axpby_f64(n, 0.0, b, 0.0, x, st), "gd_axpby_f64");
    check(gd_stream_sync(st), "sync");
    rep.success = true;
    return rep;
  }
  K.apply(1.0, x, r.data());
  check(gd_axpby_f64(n, 1.0, b, -1.0, r.data(), st), "axpby");
  check(gd_axpby_f64(n, 1.0, r.data(), 0.0, rhat.data(), st), "copy");
  check(gd_bicgstab_begin_f64(n, rhat.data(), r.data(), bnorm, opt.rel_tol, ws.data(), st), "gd_bicgstab_begin_f64");
  detail::drive(opt.max_iters, opt.rel_tol, true, resid.data(), st, [&](int, double* out) {
    check(gd_bicgstab_p_f64(n, r.data(), p.data(), v.data(), ws.data(), st), "gd_bicgstab_p_f64");
    K.apply(1.0, p.data(), v.data());
    check(gd_bicgstab_s_f64(n, r.data(), rhat.data(), v.data(), s.data(), ws.data(), st), "gd_bicgstab_s_f64");
    K.apply(1.0, s.data(), t.data());
    check(gd_bicgstab_xr_f64(n, x, r.data(), rhat.data(), p.data(), s.data(), t.data(), ws.data(), out, st),
          "gd_bicgstab_xr_f64");
  });
  int32_t state = 0, iters = 0;
  double rel = 0.0;
  check(gd_ksolve_status(ws.data(), st, &state, &iters, &rel), "gd_ksolve_status");
  rep.iterations = iters;
  rep.success = state == 1;
  rep.final_rel_residual = detail::true_residual(K, n, b, x, r.data(), tmp.data(), bnorm);
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

}  // namespace gendil_amd
