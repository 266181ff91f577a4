// Vector kernels for the device-resident Krylov loops.  All reductions
// accumulate in FP64 and are deterministic (fixed two-stage order).
//
// The multi-vector kernels work on k vectors of length n stored with leading
// dimension ldv, in FP64 or FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

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
    if (j < k) {
      double v = acc[j];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[wv][j] = v;
    }
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

// Sum of the nblk block partials of column j, by the whole block in a fixed
// order (strided partial sums, then a fixed-shape tree), result in thread 0.
__device__ __forceinline__ double block_sum_column(int nblk, int j, const double* __restrict__ partial)
{
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kThreads) s += partial[(int64_t)b * kMaxK + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// result[j] = sum_b partial[b][j]   (single block, fixed order)
__global__ void __launch_bounds__(kThreads) k_finalize(int nblk, int k, const double* __restrict__ partial,
                                                       double* __restrict__ result)
{
  for (int j = 0; j < k; ++j) {
    const double s = block_sum_column(nblk, j, partial);
    if (threadIdx.x == 0) result[j] = s;
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

// ---- allocation-free multi-dot --------------------------------------------
// Block partials of gd_multidot_f64_noalloc: one fixed array per device,
// holding at most kMaxBlocks x kMaxK doubles (256 KB), kept until the process
// ends.  Its name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch_partials[kMaxBlocks * kMaxK];

constexpr int kMaxDevices = 64;

// The device address of scratch_partials on the current device, looked up
// once per device (the array has its own address on every GPU).
int noalloc_partials(double** out)
{
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  GD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GD_ERR_UNSUPPORTED, "multidot_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    GD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch_partials)));
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return GD_OK;
}

}  // namespace blas

}  // namespace gd

using gd::fail;

extern "C" {

int gd_multidot_f64(int64_t n, int32_t k, const double* V, int64_t ldv, const double* w, double* result,
                    gd_stream stream)
{
  return gd::blas::multidot<double, double>(n, k, V, ldv, w, result, (hipStream_t)stream);
}

int gd_multidot_f64_noalloc(int64_t n, int32_t k, const double* V, int64_t ldv, const double* w, double* result,
                            gd_stream stream)
{
  using namespace gd::blas;
  if (k > kMaxK)
    return fail(GD_ERR_INVALID, "gd_multidot_f64_noalloc: at most 32 vectors per call (its fixed scratch holds the "
                                "partials of 32)");
  if (n < 0 || k < 0 || (k > 0 && (!V || !w || !result)))
    return fail(GD_ERR_INVALID, "gd_multidot_f64_noalloc: bad arguments");
  if (k == 0) return GD_OK;
  double* part = nullptr;
  if (int rc = noalloc_partials(&part)) return rc;
  // the same kernels as gd_multidot_f64; every partial is written before the
  // finishing kernel reads it, so nothing needs clearing between calls
  const int nb = grid_for(n);
  hipLaunchKernelGGL((k_multidot<double, double>), dim3(nb), dim3(kThreads), 0, (hipStream_t)stream, n, k, V, ldv, w,
                     part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, (hipStream_t)stream, nb, k, part, result);
  GD_HIP(hipGetLastError());
  return GD_OK;
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

int gd_add_f32_to_f64(int64_t n, double scale, const float* x, double* y, gd_stream stream)
{
  if (n < 0 || (n > 0 && (!x || !y))) return fail(GD_ERR_INVALID, "add: bad arguments");
  hipLaunchKernelGGL((gd::blas::k_axpby<float, double>), dim3(gd::blas::grid_for(n)), dim3(gd::blas::kThreads), 0,
                     (hipStream_t)stream, n, scale, x, 1.0, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // extern "C"
