// Vector kernels for the device-resident Krylov loops.  All reductions
// accumulate in FP64 and are deterministic (fixed two-stage order).
//
// The Gram-Schmidt kernels work on a Krylov basis V stored as k vectors of
// length n with leading dimension ldv, in FP64 or FP32, each stored vector
// carrying a scale s_j kept by the host (V_j,true = s_j V_j,stored), so the
// basis never needs a separate normalisation pass.
#include <hip/hip_runtime.h>

#include <cmath>

#include "gd_internal.hpp"

namespace gd {
namespace blas {

constexpr int kThreads = 256;
constexpr int kMaxK = 32;       // vectors per pass
constexpr int kMaxBlocks = 1024;

inline int grid_for(int64_t n)
{
  const int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
  return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

// Block-wide sum of K per-thread accumulators into out[0..K) (thread 0 writes).
template <int K>
__device__ __forceinline__ void block_reduce_store(double (&acc)[K], int k, double* out)
{
  __shared__ double red[kThreads / 64][K];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j >= k) break;
    double v = acc[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += kThreads) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][j];
    out[j] = s;
  }
}

// partial[blk][j] = sum over this block's rows of V_j[i] * w[i]
template <typename TV, typename TW>
__global__ void __launch_bounds__(kThreads) k_multidot(int64_t n, int k, const TV* __restrict__ V, int64_t ldv,
                                                       const TW* __restrict__ w, double* __restrict__ partial)
{
  double acc[kMaxK];
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < kMaxK; ++j)
      if (j < k) acc[j] += (double)V[j * ldv + i] * wi;
  }
  block_reduce_store<kMaxK>(acc, k, partial + (int64_t)blockIdx.x * kMaxK);
}

// w[i] -= sum_j c[j] V_j[i]; then (mode 0) partial = V^T w_new, (mode 1)
// partial[0] = ||w_new||^2, (mode 2) nothing.
template <typename TV, typename TW>
__global__ void __launch_bounds__(kThreads) k_update(int64_t n, int k, const double* __restrict__ c,
                                                     const TV* __restrict__ V, int64_t ldv, TW* __restrict__ w,
                                                     int mode, double sign, double* __restrict__ partial)
{
  double cj[kMaxK];
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) cj[j] = j < k ? c[j] : 0.0;
  double acc[kMaxK];
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    double v[kMaxK];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxK; ++j) {
      v[j] = j < k ? (double)V[j * ldv + i] : 0.0;
      s += cj[j] * v[j];
    }
    const double wn = (double)w[i] - sign * s;
    const TW wt = (TW)wn;
    w[i] = wt;
    const double wr = (double)wt;
    if (mode == 0) {
#pragma unroll
      for (int j = 0; j < kMaxK; ++j) acc[j] += v[j] * wr;
    } else if (mode == 1) {
      acc[0] += wr * wr;
    }
  }
  if (mode == 0) block_reduce_store<kMaxK>(acc, k, partial + (int64_t)blockIdx.x * kMaxK);
  else if (mode == 1) block_reduce_store<kMaxK>(acc, 1, partial + (int64_t)blockIdx.x * kMaxK);
}

// result[j] = sum_b partial[b][j]   (single block, fixed order)
__global__ void __launch_bounds__(kThreads) k_finalize(int nblk, int k, const double* __restrict__ partial,
                                                       double* __restrict__ result)
{
  for (int j = threadIdx.x; j < k; j += kThreads) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * kMaxK + j];
    result[j] = s;
  }
}

// Gram-Schmidt finalize: h[j] = s[j] * sum_b partial[b][j] (true coefficient),
// c[j] = s[j] * h[j] (coefficient on the stored vector), hcol[j] += h[j].
__global__ void __launch_bounds__(kThreads) k_gs_finalize(int nblk, int k, const double* __restrict__ partial,
                                                          const double* __restrict__ s, double* __restrict__ c,
                                                          double* __restrict__ hcol, int accumulate)
{
  for (int j = threadIdx.x; j < k; j += kThreads) {
    double sum = 0.0;
    for (int b = 0; b < nblk; ++b) sum += partial[(int64_t)b * kMaxK + j];
    const double h = s[j] * sum;
    c[j] = s[j] * h;
    hcol[j] = accumulate ? hcol[j] + h : h;
  }
}

template <typename TX, typename TY>
__global__ void k_axpby(int64_t n, double a, const TX* __restrict__ x, double b, TY* __restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    y[i] = (TY)(a * (double)x[i] + (b == 0.0 ? 0.0 : b * (double)y[i]));
}

template <typename TX, typename TY>
__global__ void k_scale_convert(int64_t n, double a, const TX* __restrict__ x, TY* __restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    y[i] = (TY)(a * (double)x[i]);
}

// Scratch for the two-stage reductions: one buffer per device, grown on
// demand.  Calls on different streams of one device must not overlap.
double* scratch(int k_blocks)
{
  static double* buf = nullptr;
  static int cap = 0;
  if (k_blocks > cap) {
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    if (hipMalloc(&buf, (size_t)k_blocks * kMaxK * sizeof(double)) != hipSuccess) {
      cap = 0;
      return nullptr;
    }
    cap = k_blocks;
  }
  return buf;
}

template <typename TV, typename TW>
int multidot(int64_t n, int k, const TV* V, int64_t ldv, const TW* w, double* result, hipStream_t s)
{
  if (n < 0 || k < 0 || (k > 0 && (!V || !w || !result))) return fail(GD_ERR_INVALID, "multidot: bad arguments");
  for (int j0 = 0; j0 < k; j0 += kMaxK) {
    const int kk = k - j0 < kMaxK ? k - j0 : kMaxK;
    const int nb = grid_for(n);
    double* part = scratch(nb);
    if (!part) return fail(GD_ERR_ALLOC, "multidot: scratch allocation failed");
    hipLaunchKernelGGL((k_multidot<TV, TW>), dim3(nb), dim3(kThreads), 0, s, n, kk, V + j0 * ldv, ldv, w, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, s, nb, kk, part, result + j0);
  }
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template <typename TV, typename TW>
int multiaxpy(int64_t n, int k, const double* h, const TV* V, int64_t ldv, TW* w, double sign, hipStream_t s)
{
  if (n < 0 || k < 0 || (k > 0 && (!V || !w || !h))) return fail(GD_ERR_INVALID, "multiaxpy: bad arguments");
  for (int j0 = 0; j0 < k; j0 += kMaxK) {
    const int kk = k - j0 < kMaxK ? k - j0 : kMaxK;
    hipLaunchKernelGGL((k_update<TV, TW>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, kk, h + j0, V + j0 * ldv,
                       ldv, w, 2, sign, nullptr);
  }
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // namespace blas

// --- Gram-Schmidt step used by the solvers (declared in gd_krylov.hpp) ----
// One classical Gram-Schmidt pass against k <= 32 stored vectors, fused:
//   pass 1: hcol = S V^T w                     (S = diag(s))
//   pass 2: w -= V (S hcol); h2 = S V^T w; hcol += h2   (re-orthogonalisation)
//   pass 3: w -= V (S h2);  norm2 = ||w||^2
// Device arrays: s[k], c[k] (scratch), hcol[k], norm2[1].
template <typename TV>
int cgs2_step(int64_t n, int k, const TV* V, int64_t ldv, TV* w, const double* s, double* c, double* hcol,
              double* norm2, hipStream_t st)
{
  using namespace blas;
  if (k > kMaxK) return fail(GD_ERR_INVALID, "cgs2_step: at most 32 basis vectors per pass");
  const int nb = grid_for(n);
  double* part = scratch(nb);
  if (!part) return fail(GD_ERR_ALLOC, "cgs2_step: scratch allocation failed");
  hipLaunchKernelGGL((k_multidot<TV, TV>), dim3(nb), dim3(kThreads), 0, st, n, k, V, ldv, w, part);
  hipLaunchKernelGGL(k_gs_finalize, dim3(1), dim3(kThreads), 0, st, nb, k, part, s, c, hcol, 0);
  hipLaunchKernelGGL((k_update<TV, TV>), dim3(nb), dim3(kThreads), 0, st, n, k, c, V, ldv, w, 0, 1.0, part);
  hipLaunchKernelGGL(k_gs_finalize, dim3(1), dim3(kThreads), 0, st, nb, k, part, s, c, hcol, 1);
  hipLaunchKernelGGL((k_update<TV, TV>), dim3(nb), dim3(kThreads), 0, st, n, k, c, V, ldv, w, 1, 1.0, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, st, nb, 1, part, norm2);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template int cgs2_step<double>(int64_t, int, const double*, int64_t, double*, const double*, double*, double*,
                               double*, hipStream_t);
template int cgs2_step<float>(int64_t, int, const float*, int64_t, float*, const double*, double*, double*,
                              double*, hipStream_t);

// x += sum_j c_j V_j (x FP64, V FP64 or FP32)
template <typename TV>
int basis_update(int64_t n, int k, const double* c, const TV* V, int64_t ldv, double* x, hipStream_t st)
{
  return blas::multiaxpy<TV, double>(n, k, c, V, ldv, x, -1.0, st);
}
template int basis_update<double>(int64_t, int, const double*, const double*, int64_t, double*, hipStream_t);
template int basis_update<float>(int64_t, int, const double*, const float*, int64_t, double*, hipStream_t);

}  // namespace gd

using gd::fail;

extern "C" {

int gd_multidot_f64(int64_t n, int32_t k, const double* V, int64_t ldv, const double* w, double* result,
                    gd_stream stream)
{
  return gd::blas::multidot<double, double>(n, k, V, ldv, w, result, (hipStream_t)stream);
}

int gd_multidot_f32(int64_t n, int32_t k, const float* V, int64_t ldv, const float* w, double* result,
                    gd_stream stream)
{
  return gd::blas::multidot<float, float>(n, k, V, ldv, w, result, (hipStream_t)stream);
}

int gd_multiaxpy_f64(int64_t n, int32_t k, const double* h, const double* V, int64_t ldv, double* w,
                     gd_stream stream)
{
  return gd::blas::multiaxpy<double, double>(n, k, h, V, ldv, w, 1.0, (hipStream_t)stream);
}

int gd_multiaxpy_f32(int64_t n, int32_t k, const double* h, const float* V, int64_t ldv, float* w,
                     gd_stream stream)
{
  return gd::blas::multiaxpy<float, float>(n, k, h, V, ldv, w, 1.0, (hipStream_t)stream);
}

int gd_axpby_f64(int64_t n, double a, const double* x, double b, double* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "axpby: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_axpby<double, double>), dim3(gd::blas::grid_for(n)), dim3(gd::blas::kThreads), 0,
                     (hipStream_t)stream, n, a, x, b, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_axpby_f32(int64_t n, float a, const float* x, float b, float* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "axpby: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_axpby<float, float>), dim3(gd::blas::grid_for(n)), dim3(gd::blas::kThreads), 0,
                     (hipStream_t)stream, n, (double)a, x, (double)b, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_convert_f64_to_f32(int64_t n, double scale, const double* x, float* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "convert: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_scale_convert<double, float>), dim3(gd::blas::grid_for(n)),
                     dim3(gd::blas::kThreads), 0, (hipStream_t)stream, n, scale, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_convert_f32_to_f64(int64_t n, double scale, const float* x, double* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "convert: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_scale_convert<float, double>), dim3(gd::blas::grid_for(n)),
                     dim3(gd::blas::kThreads), 0, (hipStream_t)stream, n, scale, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_cgs2_step_f64(int64_t n, int32_t k, const double* V, int64_t ldv, double* w, const double* s, double* c,
                     double* hcol, double* norm2, gd_stream stream)
{
  return gd::cgs2_step<double>(n, k, V, ldv, w, s, c, hcol, norm2, (hipStream_t)stream);
}

int gd_cgs2_step_f32(int64_t n, int32_t k, const float* V, int64_t ldv, float* w, const double* s, double* c,
                     double* hcol, double* norm2, gd_stream stream)
{
  return gd::cgs2_step<float>(n, k, V, ldv, w, s, c, hcol, norm2, (hipStream_t)stream);
}

int gd_basis_update_f64(int64_t n, int32_t k, const double* c, const double* V, int64_t ldv, double* x,
                        gd_stream stream)
{
  return gd::basis_update<double>(n, k, c, V, ldv, x, (hipStream_t)stream);
}

int gd_basis_update_f32(int64_t n, int32_t k, const double* c, const float* V, int64_t ldv, double* x,
                        gd_stream stream)
{
  return gd::basis_update<float>(n, k, c, V, ldv, x, (hipStream_t)stream);
}

int gd_add_f32_to_f64(int64_t n, double scale, const float* x, double* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "add: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_axpby<float, double>), dim3(gd::blas::grid_for(n)), dim3(gd::blas::kThreads), 0,
                     (hipStream_t)stream, n, scale, x, 1.0, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // extern "C"
