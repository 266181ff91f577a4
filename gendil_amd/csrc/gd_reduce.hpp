// Deterministic two-stage reductions shared by the Krylov kernels: every
// block of a pass stores its column sums as partials (column-major, one
// contiguous column per reduced quantity), and a one-block finishing kernel
// sums them in a fixed order.  The kernel boundary between the two orders the
// partials; a device-scope fence in every block, the alternative, measured far
// slower on the eight-XCD part.
#pragma once

#include <hip/hip_runtime.h>

namespace gd {
namespace red {

constexpr int kThreads = 256;       // pass kernels
constexpr int kMaxBlocks = 1024;    // partials per column
constexpr int kFinThreads = 1024;   // finishing kernels

inline int grid_for(int64_t n, int w)
{
  const int64_t per = (int64_t)kThreads * 4 * w;
  const int64_t b = (n + per - 1) / per;
  return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// out[j] = sum over the block's NT threads of acc[j] (waves in order); out is
// visible to every thread on return.
template <int NT, int K>
__device__ __forceinline__ void block_sum(const double (&acc)[K], double* out)
{
  constexpr int NW = NT / 64;
  __shared__ double red[NW][K];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double v = wave_sum(acc[j]);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < K; j += NT) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w][j];
    out[j] = s;
  }
  __syncthreads();
}

// The block's column sums as its partials, partial[j][blockIdx.x].
template <int K>
__device__ __forceinline__ void store_block_partials(const double (&acc)[K], double (*partial)[kMaxBlocks])
{
  __shared__ double row[K];
  block_sum<kThreads, K>(acc, row);
  for (int j = threadIdx.x; j < K; j += kThreads) partial[j][blockIdx.x] = row[j];
}

// col[j] = sum over the nblk blocks' partials of column j, j < K: 32 threads
// per column, each summing every 32nd block (consecutive threads on
// consecutive blocks, all loads in flight together), then a 5-step shuffle
// sum over the 32 threads.  A fixed order for a given grid; visible to every
// thread on return.
template <int K>
__device__ __forceinline__ void sum_partials(const double (*partial)[kMaxBlocks], int nblk, double* col)
{
  constexpr int TPC = 32;                  // threads per column
  constexpr int CPP = kFinThreads / TPC;   // columns per round
  constexpr int PER = kMaxBlocks / TPC;    // blocks per thread
  const int lane = threadIdx.x % TPC, c0 = threadIdx.x / TPC;
#pragma unroll
  for (int r = 0; r < (K + CPP - 1) / CPP; ++r) {
    const int j = c0 + r * CPP;
    if (j < K) {  // uniform over each 32-thread group; the shuffles stay inside it
      const double* cp = partial[j];
      double v[PER];
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int bb = lane + TPC * i;
        const double x = cp[bb < nblk ? bb : nblk - 1];
        v[i] = bb < nblk ? x : 0.0;
      }
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < PER; ++i) sum += v[i];
#pragma unroll
      for (int off = TPC / 2; off > 0; off >>= 1) sum += __shfl_down(sum, off, TPC);
      if (lane == 0) col[j] = sum;
    }
  }
  __syncthreads();
}

// First block, first thread: hand a value left by an earlier kernel to the
// caller's pointer (system scope: the pointer may be pinned host memory), early
// in a kernel that runs on, so no kernel's completion waits for the host write.
__device__ __forceinline__ void publish_pending(double** pending, const double* value)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* p = *pending;
    if (p) {
      __hip_atomic_store(p, *value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      *pending = nullptr;
    }
  }
}

}  // namespace red
}  // namespace gd
