// Device-resident CG and BiCGSTAB iterations (FP64).  Every scalar of the
// recurrences (rho, alpha, omega, beta, the residual norm) lives in a small
// device workspace: each vector pass reads the scalars it needs at its start,
// reductions end in a one-block finishing kernel (gd_reduce.hpp, fixed order),
// and the finishing kernels also take the control decisions (convergence,
// breakdown), setting a state that makes every later pass of the workspace a
// no-op.  The host enqueues one iteration ahead and polls the residual
// estimate in pinned memory, so no dot product synchronises the stream.
//
// CG (conjugategradient.hpp:125-235): Ap = A p (caller), then
//   pass 1: <p, Ap>                 -> alpha = rr / <p, Ap>
//   pass 2: x += alpha p, r -= alpha Ap, <r, r> -> rel, beta = rr'/rr
//   pass 3: p = r + beta p          (publishes the residual estimate)
// BiCGSTAB (bicgstab.hpp:44-157), per iteration with the caller's applies:
//   p-pass: p = r (first) or r + beta (p - omega v), beta = (rho/rho')(alpha/omega)
//   v = A p;  s-passes: <rhat, v> -> alpha = rho / <rhat, v>; s = r - alpha v, <s, s>
//   t = A s;  xr-passes: <t, s>, <t, t> -> omega; x += alpha p + omega s,
//             r = s - omega t, <r, r>, <rhat, r> -> rel, rho
// with the reference's early exit (||s|| below the tolerance: x += alpha p),
// its <t, t> fallback (alpha step only) and its breakdown guards (|.| < 1e-30,
// absolute, as there).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "gd_internal.hpp"
#include "gd_reduce.hpp"

namespace gd {
namespace ksolve {

using red::kFinThreads;
using red::kMaxBlocks;
using red::kThreads;

constexpr double kTiny = 1e-30;  // breakdown / stagnation guard (bicgstab.hpp:58)
enum State : int { kRunning = 0, kConverged = 1, kStopped = 2 };

struct Workspace {
  double bnorm, tol;
  double rr;                  // CG: <r, r>
  double rho, rho_prev;       // BiCGSTAB: <rhat, r> now and one iteration before
  double alpha, omega, beta;
  double rel;                 // ||r|| / ||b|| after the last completed iteration
  double rel_s;               // BiCGSTAB: ||s|| / ||b||
  int state;                  // State
  int iters;                  // completed iterations
  int early;                  // BiCGSTAB: s met the tolerance (alpha step only, then converged)
  int fallback;               // BiCGSTAB: <t, t> vanished (alpha step only, then stop)
  double* pending;            // caller pointer owed the residual estimate
  double pending_value;
  double partial[2][kMaxBlocks];
};

__device__ __forceinline__ void publish(Workspace* ws) { red::publish_pending(&ws->pending, &ws->pending_value); }

// residual estimate handed to the host: the relative residual, NaN when the
// iteration stopped without converging
__device__ __forceinline__ void owe(Workspace* ws, double* resid_out)
{
  ws->pending = resid_out;
  ws->pending_value = ws->state == kStopped ? __builtin_nan("") : ws->rel;
}

#define GD_KSOLVE_LOOP(i) \
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)

// ---- shared passes ---------------------------------------------------------

// partials of <a, b> (and <c, d> when c is not null), accumulated in FP64;
// skipped once the iteration has ended, except when starting one (begin)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_dots(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                   const T* __restrict__ c, const T* __restrict__ d,
                                                   Workspace* __restrict__ ws, int begin)
{
  if (!begin && ws->state != kRunning) return;
  double acc[2] = {0.0, 0.0};
  if (c) {
    GD_KSOLVE_LOOP(i)
    {
      acc[0] += (double)a[i] * (double)b[i];
      acc[1] += (double)c[i] * (double)d[i];
    }
    red::store_block_partials<2>(acc, ws->partial);
  } else {
    double a1[1] = {0.0};
    GD_KSOLVE_LOOP(i) a1[0] += (double)a[i] * (double)b[i];
    red::store_block_partials<1>(a1, ws->partial);
  }
}

// ---- CG --------------------------------------------------------------------

// begin: rr = <r, r>, rel, state
__global__ void __launch_bounds__(kFinThreads) k_cg_fin0(int nblk, double bnorm, double tol, Workspace* __restrict__ ws)
{
  __shared__ double col[1];
  red::sum_partials<1>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    ws->bnorm = bnorm;
    ws->tol = tol;
    ws->rr = col[0];
    ws->rel = sqrt(col[0]) / bnorm;
    ws->iters = 0;
    ws->state = ws->rel <= tol ? kConverged : kRunning;
    ws->pending = nullptr;
  }
}

__global__ void __launch_bounds__(kFinThreads) k_cg_fin1(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[1];
  if (ws->state != kRunning) return;
  red::sum_partials<1>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    // the reference CG has no breakdown guard (conjugategradient.hpp:175-182);
    // only an exact zero, which would divide by zero, stops it here
    const double pap = col[0];
    if (pap == 0.0) {
      ws->state = kStopped;
    } else {
      ws->alpha = ws->rr / pap;
    }
  }
}

// vectors in precision T, the scalars and <r, r> in FP64
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cg_xr(int64_t n, T* __restrict__ x, T* __restrict__ r,
                                                    const T* __restrict__ p, const T* __restrict__ ap,
                                                    Workspace* __restrict__ ws)
{
  if (ws->state != kRunning) return;
  const T alpha = (T)ws->alpha;
  double acc[1] = {0.0};
  GD_KSOLVE_LOOP(i)
  {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * ap[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  red::store_block_partials<1>(acc, ws->partial);
}

__global__ void __launch_bounds__(kFinThreads) k_cg_fin2(int nblk, Workspace* __restrict__ ws, double* resid_out)
{
  __shared__ double col[1];
  const bool run = ws->state == kRunning;
  if (run) red::sum_partials<1>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    if (run) {
      const double rr = col[0];
      ws->beta = rr / ws->rr;
      ws->rr = rr;
      ws->rel = sqrt(rr) / ws->bnorm;
      ws->iters += 1;
      if (ws->rel <= ws->tol) ws->state = kConverged;
    }
    owe(ws, resid_out);
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_cg_p(int64_t n, const T* __restrict__ r, T* __restrict__ p,
                                                   Workspace* __restrict__ ws)
{
  publish(ws);
  if (ws->state != kRunning) return;
  const T beta = (T)ws->beta;
  GD_KSOLVE_LOOP(i) p[i] = r[i] + beta * p[i];
}

// ---- BiCGSTAB ----------------------------------------------------------------

// begin: rho = <rhat, r>, rel = ||r|| / ||b||, state
__global__ void __launch_bounds__(kFinThreads) k_bi_fin0(int nblk, double bnorm, double tol, Workspace* __restrict__ ws)
{
  __shared__ double col[2];
  red::sum_partials<2>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    ws->bnorm = bnorm;
    ws->tol = tol;
    ws->rho = col[0];
    ws->rho_prev = 1.0;
    ws->alpha = 1.0;
    ws->omega = 1.0;
    ws->rel = sqrt(col[1]) / bnorm;
    ws->iters = 0;
    ws->early = 0;
    ws->fallback = 0;
    ws->pending = nullptr;
    ws->state = ws->rel < tol ? kConverged : (fabs(ws->rho) < kTiny ? kStopped : kRunning);
  }
}

__global__ void __launch_bounds__(kThreads) k_bi_p(int64_t n, const double* __restrict__ r, double* __restrict__ p,
                                                   const double* __restrict__ v, Workspace* __restrict__ ws)
{
  publish(ws);
  if (ws->state != kRunning) return;
  if (ws->iters == 0) {
    GD_KSOLVE_LOOP(i) p[i] = r[i];
  } else {
    const double beta = (ws->rho / ws->rho_prev) * (ws->alpha / ws->omega), omega = ws->omega;
    GD_KSOLVE_LOOP(i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
  }
}

__global__ void __launch_bounds__(kFinThreads) k_bi_alpha(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[1];
  if (ws->state != kRunning) return;
  red::sum_partials<1>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    if (fabs(col[0]) < kTiny) ws->state = kStopped;
    else ws->alpha = ws->rho / col[0];
  }
}

__global__ void __launch_bounds__(kThreads) k_bi_s(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                                   double* __restrict__ s, Workspace* __restrict__ ws)
{
  if (ws->state != kRunning) return;
  const double alpha = ws->alpha;
  double acc[1] = {0.0};
  GD_KSOLVE_LOOP(i)
  {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    acc[0] += si * si;
  }
  red::store_block_partials<1>(acc, ws->partial);
}

__global__ void __launch_bounds__(kFinThreads) k_bi_snorm(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[1];
  if (ws->state != kRunning) return;
  red::sum_partials<1>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    ws->rel_s = sqrt(col[0]) / ws->bnorm;
    ws->early = ws->rel_s < ws->tol ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kFinThreads) k_bi_omega(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[2];
  if (ws->state != kRunning || ws->early) return;
  red::sum_partials<2>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    const double ts = col[0], tt = col[1];
    if (fabs(tt) < kTiny) {
      ws->fallback = 1;
    } else {
      const double omega = ts / tt;
      if (fabs(omega) < kTiny) ws->state = kStopped;
      else ws->omega = omega;
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_bi_xr(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ rhat, const double* __restrict__ p,
                                                    const double* __restrict__ s, const double* __restrict__ t,
                                                    Workspace* __restrict__ ws)
{
  if (ws->state != kRunning) return;
  const double alpha = ws->alpha;
  double acc[2] = {0.0, 0.0};
  if (ws->early || ws->fallback) {
    GD_KSOLVE_LOOP(i)
    {
      x[i] += alpha * p[i];
      const double ri = s[i];
      r[i] = ri;
      acc[0] += ri * ri;
      acc[1] += rhat[i] * ri;
    }
  } else {
    const double omega = ws->omega;
    GD_KSOLVE_LOOP(i)
    {
      x[i] += alpha * p[i] + omega * s[i];
      const double ri = s[i] - omega * t[i];
      r[i] = ri;
      acc[0] += ri * ri;
      acc[1] += rhat[i] * ri;
    }
  }
  red::store_block_partials<2>(acc, ws->partial);
}

__global__ void __launch_bounds__(kFinThreads) k_bi_fin(int nblk, Workspace* __restrict__ ws, double* resid_out)
{
  __shared__ double col[2];
  const bool run = ws->state == kRunning;
  if (run) red::sum_partials<2>(ws->partial, nblk, col);
  if (threadIdx.x == 0) {
    if (run) {
      ws->iters += 1;
      ws->rel = sqrt(col[0]) / ws->bnorm;
      if (ws->early) {
        ws->rel = ws->rel_s;
        ws->state = kConverged;
      } else if (ws->fallback) {
        ws->state = ws->rel < ws->tol ? kConverged : kStopped;
      } else if (ws->rel < ws->tol) {
        ws->state = kConverged;
      } else {
        ws->rho_prev = ws->rho;
        ws->rho = col[1];
        if (fabs(ws->rho) < kTiny) ws->state = kStopped;
      }
    }
    owe(ws, resid_out);
  }
}

#undef GD_KSOLVE_LOOP

inline int blocks(int64_t n) { return red::grid_for(n, 1); }

}  // namespace ksolve
}  // namespace gd

using gd::fail;
namespace ks = gd::ksolve;

extern "C" {

size_t gd_ksolve_workspace_bytes(void) { return sizeof(ks::Workspace); }

}  // extern "C"

namespace gd {
namespace ksolve {

template <typename T>
int cg_begin(int64_t n, const T* r, double bnorm, double rel_tol, void* ws, gd_stream stream)
{
  if (n < 1 || !r || !ws || !(bnorm > 0.0)) return fail(GD_ERR_INVALID, "gd_cg_begin: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  auto* w = static_cast<Workspace*>(ws);
  const int nb = blocks(n);
  hipLaunchKernelGGL(k_dots<T>, dim3(nb), dim3(kThreads), 0, st, n, r, r, nullptr, nullptr, w, 1);
  hipLaunchKernelGGL(k_cg_fin0, dim3(1), dim3(kFinThreads), 0, st, nb, bnorm, rel_tol, w);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template <typename T>
int cg_step(int64_t n, T* x, T* r, T* p, const T* ap, void* ws, double* resid_out, gd_stream stream)
{
  if (n < 1 || !x || !r || !p || !ap || !ws) return fail(GD_ERR_INVALID, "gd_cg_step: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  auto* w = static_cast<Workspace*>(ws);
  const int nb = blocks(n);
  hipLaunchKernelGGL(k_dots<T>, dim3(nb), dim3(kThreads), 0, st, n, p, ap, nullptr, nullptr, w, 0);
  hipLaunchKernelGGL(k_cg_fin1, dim3(1), dim3(kFinThreads), 0, st, nb, w);
  hipLaunchKernelGGL(k_cg_xr<T>, dim3(nb), dim3(kThreads), 0, st, n, x, r, p, ap, w);
  hipLaunchKernelGGL(k_cg_fin2, dim3(1), dim3(kFinThreads), 0, st, nb, w, resid_out);
  hipLaunchKernelGGL(k_cg_p<T>, dim3(nb), dim3(kThreads), 0, st, n, r, p, w);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // namespace ksolve
}  // namespace gd

extern "C" {

int gd_cg_begin_f64(int64_t n, const double* r, double bnorm, double rel_tol, void* ws, gd_stream stream)
{
  return gd::ksolve::cg_begin<double>(n, r, bnorm, rel_tol, ws, stream);
}

int gd_cg_step_f64(int64_t n, double* x, double* r, double* p, const double* ap, void* ws, double* resid_out,
                   gd_stream stream)
{
  return gd::ksolve::cg_step<double>(n, x, r, p, ap, ws, resid_out, stream);
}

int gd_cg_begin_f32(int64_t n, const float* r, double bnorm, double rel_tol, void* ws, gd_stream stream)
{
  return gd::ksolve::cg_begin<float>(n, r, bnorm, rel_tol, ws, stream);
}

int gd_cg_step_f32(int64_t n, float* x, float* r, float* p, const float* ap, void* ws, double* resid_out,
                   gd_stream stream)
{
  return gd::ksolve::cg_step<float>(n, x, r, p, ap, ws, resid_out, stream);
}

int gd_bicgstab_begin_f64(int64_t n, const double* rhat, const double* r, double bnorm, double rel_tol, void* ws,
                          gd_stream stream)
{
  if (n < 1 || !rhat || !r || !ws || !(bnorm > 0.0)) return fail(GD_ERR_INVALID, "gd_bicgstab_begin_f64: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  auto* w = static_cast<ks::Workspace*>(ws);
  const int nb = ks::blocks(n);
  hipLaunchKernelGGL(ks::k_dots<double>, dim3(nb), dim3(ks::kThreads), 0, st, n, rhat, r, r, r, w, 1);
  hipLaunchKernelGGL(ks::k_bi_fin0, dim3(1), dim3(ks::kFinThreads), 0, st, nb, bnorm, rel_tol, w);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_bicgstab_p_f64(int64_t n, const double* r, double* p, const double* v, void* ws, gd_stream stream)
{
  if (n < 1 || !r || !p || !v || !ws) return fail(GD_ERR_INVALID, "gd_bicgstab_p_f64: bad arguments");
  hipLaunchKernelGGL(ks::k_bi_p, dim3(ks::blocks(n)), dim3(ks::kThreads), 0, (hipStream_t)stream, n, r, p, v,
                     static_cast<ks::Workspace*>(ws));
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_bicgstab_s_f64(int64_t n, const double* r, const double* rhat, const double* v, double* s, void* ws,
                      gd_stream stream)
{
  if (n < 1 || !r || !rhat || !v || !s || !ws) return fail(GD_ERR_INVALID, "gd_bicgstab_s_f64: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  auto* w = static_cast<ks::Workspace*>(ws);
  const int nb = ks::blocks(n);
  hipLaunchKernelGGL(ks::k_dots<double>, dim3(nb), dim3(ks::kThreads), 0, st, n, rhat, v, nullptr, nullptr, w, 0);
  hipLaunchKernelGGL(ks::k_bi_alpha, dim3(1), dim3(ks::kFinThreads), 0, st, nb, w);
  hipLaunchKernelGGL(ks::k_bi_s, dim3(nb), dim3(ks::kThreads), 0, st, n, r, v, s, w);
  hipLaunchKernelGGL(ks::k_bi_snorm, dim3(1), dim3(ks::kFinThreads), 0, st, nb, w);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_bicgstab_xr_f64(int64_t n, double* x, double* r, const double* rhat, const double* p, const double* s,
                       const double* t, void* ws, double* resid_out, gd_stream stream)
{
  if (n < 1 || !x || !r || !rhat || !p || !s || !t || !ws)
    return fail(GD_ERR_INVALID, "gd_bicgstab_xr_f64: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  auto* w = static_cast<ks::Workspace*>(ws);
  const int nb = ks::blocks(n);
  hipLaunchKernelGGL(ks::k_dots<double>, dim3(nb), dim3(ks::kThreads), 0, st, n, t, s, t, t, w, 0);
  hipLaunchKernelGGL(ks::k_bi_omega, dim3(1), dim3(ks::kFinThreads), 0, st, nb, w);
  hipLaunchKernelGGL(ks::k_bi_xr, dim3(nb), dim3(ks::kThreads), 0, st, n, x, r, rhat, p, s, t, w);
  hipLaunchKernelGGL(ks::k_bi_fin, dim3(1), dim3(ks::kFinThreads), 0, st, nb, w, resid_out);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_ksolve_status(const void* ws, gd_stream stream, int32_t* state, int32_t* iterations, double* rel_residual)
{
  if (!ws) return fail(GD_ERR_INVALID, "gd_ksolve_status: null workspace");
  ks::Workspace h;
  GD_HIP(hipMemcpyAsync(&h, ws, offsetof(ks::Workspace, partial), hipMemcpyDeviceToHost, (hipStream_t)stream));
  GD_HIP(hipStreamSynchronize((hipStream_t)stream));
  if (state) *state = h.state;
  if (iterations) *iterations = h.iters;
  if (rel_residual) *rel_residual = h.rel;
  return GD_OK;
}

}  // extern "C"
