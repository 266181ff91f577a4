// Device-resident Arnoldi process for restarted GMRES(m), m <= 32.
//
// One iteration k (basis vectors V_0..V_k stored with scales s_j, true vector
// v_j = s_j V_j, and W = V_{k+1} holding K V_k) is three fused passes over the
// basis, each followed by a one-block kernel that sums the block partials in
// block order (deterministic):
//   pass A: h1_j = s_j <V_j, W>                         (classical GS)
//   pass B: W -= sum_j s_j h1_j V_j, then h2_j = s_j <V_j, W>, ||W||^2
//           (re-orthogonalisation, the "twice is enough" CGS2 step)
//   pass C: Z = (W - sum_j s_j h2_j V_j) / p with p^2 = ||W||^2 - sum_j h2_j^2
//           (Pythagoras), stored as V_{k+1}; nu = ||Z|| exactly; s_{k+1} = 1/nu
// and the finalize kernel of pass C finishes the Hessenberg column
// H_{jk} = s_k (h1_j + h2_j), H_{k+1,k} = s_k p nu, applies the previous Givens
// rotations and forms the new one (gmres.hpp:170-210), and writes the residual
// estimate |g_{k+1}| to an optional caller pointer (pinned host memory, so the
// host can read it after an event with no copy in the stream).  Stored vectors
// are kept at unit norm (up to rounding), so the operator is applied with
// scale 1 and FP32 bases cannot drift out of range.
//
// The cycle end solves the triangular system H y = g in every block (k <= 32,
// cheap) and adds x += sum_j s_j y_j V_j in FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "gd_internal.hpp"

namespace gd {
namespace arnoldi {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxM = 32;          // restart length
constexpr int kMaxBlocks = 1024;
constexpr int kStride = kMaxM + 2; // partial columns per block: k+1 dots + norm
constexpr int kFinThreads = 1024;  // finalize kernels: one wave per partial column

struct Workspace {
  double H[kMaxM][kMaxM + 1];  // column k, rows 0..k+1 (rotated in place)
  double cs[kMaxM], sn[kMaxM];
  double g[kMaxM + 1];
  double s[kMaxM + 1];
  double c[kMaxM];   // coefficients on the stored vectors for the next update pass
  double h1[kMaxM];  // accumulated true GS coefficients of the current column
  double pnorm, pinv;
  double partial[kStride][kMaxBlocks];  // [column][block]: a column is contiguous
};

inline int grid_for(int64_t n)
{
  const int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
  return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sums of the first `cols` accumulators into partial[:][blockIdx.x].
template <int K>
__device__ __forceinline__ void store_block_partials(const double (&acc)[K], int cols, Workspace* ws)
{
  __shared__ double red[kWaves][K];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < cols) {
      const double v = wave_sum(acc[j]);
      if (lane == 0) red[wv][j] = v;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < cols; j += kThreads) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += red[w][j];
    ws->partial[j][blockIdx.x] = s;
  }
}

// out[j] = sum over blocks of partial[j][b], j < cols: one wave per column
// (contiguous, all loads in flight together), fixed order.  Results land in
// shared memory `col` (visible to all threads after the trailing barrier).
__device__ __forceinline__ void sum_columns(const Workspace* ws, int nblk, int cols, double* col)
{
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  for (int j = wv; j < cols; j += kFinThreads / 64) {
    const double* c = ws->partial[j];
    double v[kMaxBlocks / 64];
#pragma unroll
    for (int i = 0; i < kMaxBlocks / 64; ++i) {
      const int b = lane + 64 * i;
      v[i] = b < nblk ? c[b] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxBlocks / 64; ++i) s += v[i];
    s = wave_sum(s);
    if (lane == 0) col[j] = s;
  }
  __syncthreads();
}

// The basis passes take the number of stored vectors KK as a template
// parameter: every load of a pass is then unconditional and independent, so
// all of them are in flight together (a runtime count made each load a
// predicated branch followed by a wait).  Reductions end in a single-block
// finalize kernel; the kernel boundary orders the partials.

// Pass A: partial[b][j] = <V_j, W> over block b's rows.
template <typename TV, int KK>
__global__ void __launch_bounds__(kThreads) k_pass_a(int64_t n, const TV* __restrict__ V, int64_t ldv,
                                                     Workspace* __restrict__ ws)
{
  const TV* __restrict__ w = V + (int64_t)KK * ldv;
  double acc[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < KK; ++j) acc[j] += (double)V[j * ldv + i] * wi;
  }
  store_block_partials<KK>(acc, KK, ws);
}

__global__ void __launch_bounds__(kFinThreads) k_fin_a(int nblk, int kk, Workspace* __restrict__ ws)
{
  __shared__ double col[kStride];
  sum_columns(ws, nblk, kk, col);
  for (int j = threadIdx.x; j < kk; j += kFinThreads) {
    const double h = ws->s[j] * col[j];
    ws->h1[j] = h;
    ws->c[j] = ws->s[j] * h;
  }
}

// Pass B: W -= sum_j c_j V_j; partial = <V_j, W_new>, ||W_new||^2 (column KK).
template <typename TV, int KK>
__global__ void __launch_bounds__(kThreads) k_pass_b(int64_t n, TV* __restrict__ V, int64_t ldv,
                                                     Workspace* __restrict__ ws)
{
  TV* __restrict__ w = V + (int64_t)KK * ldv;
  double cj[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) cj[j] = ws->c[j];
  double acc[KK + 1];
#pragma unroll
  for (int j = 0; j <= KK; ++j) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    double v[KK];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) {
      v[j] = (double)V[j * ldv + i];
      s += cj[j] * v[j];
    }
    const TV wt = (TV)((double)w[i] - s);
    w[i] = wt;
    const double wr = (double)wt;
#pragma unroll
    for (int j = 0; j < KK; ++j) acc[j] += v[j] * wr;
    acc[KK] += wr * wr;
  }
  store_block_partials<KK + 1>(acc, KK + 1, ws);
}

__global__ void __launch_bounds__(kFinThreads) k_fin_b(int nblk, int kk, Workspace* __restrict__ ws)
{
  __shared__ double col[kStride];
  sum_columns(ws, nblk, kk + 1, col);
  if (threadIdx.x == 0) {
    double ss = 0.0;
    for (int j = 0; j < kk; ++j) {
      const double h2 = ws->s[j] * col[j];
      ws->h1[j] += h2;
      ws->c[j] = ws->s[j] * h2;
      ss += h2 * h2;
    }
    const double p2 = col[kk] - ss;
    const double p = p2 > 0.0 ? sqrt(p2) : 0.0;
    ws->pnorm = p;
    ws->pinv = p > 0.0 ? 1.0 / p : 0.0;
  }
}

// Pass C: Z = (W - sum_j c_j V_j) / p stored as V_KK; partial = ||Z||^2.
template <typename TV, int KK>
__global__ void __launch_bounds__(kThreads) k_pass_c(int64_t n, TV* __restrict__ V, int64_t ldv,
                                                     Workspace* __restrict__ ws)
{
  TV* __restrict__ w = V + (int64_t)KK * ldv;
  const double pinv = ws->pinv;
  double cj[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) cj[j] = ws->c[j];
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) s += cj[j] * (double)V[j * ldv + i];
    const TV zt = (TV)(((double)w[i] - s) * pinv);
    w[i] = zt;
    const double zr = (double)zt;
    acc[0] += zr * zr;
  }
  store_block_partials<1>(acc, 1, ws);
}

// Finishes column k of the Hessenberg matrix and its Givens rotation.
__global__ void __launch_bounds__(kFinThreads) k_fin_c(int nblk, int k, Workspace* __restrict__ ws,
                                                    double* __restrict__ resid_out)
{
  __shared__ double col[kStride];
  sum_columns(ws, nblk, 1, col);
  if (threadIdx.x == 0) {
    const int kk = k + 1;
    const double nu = sqrt(col[0]);
    const double sk = ws->s[k];
    ws->s[k + 1] = nu > 0.0 ? 1.0 / nu : 0.0;
    double* h = ws->H[k];
    for (int j = 0; j < kk; ++j) h[j] = sk * ws->h1[j];
    h[k + 1] = sk * ws->pnorm * nu;
    for (int i = 0; i < k; ++i) {
      const double a = h[i], b = h[i + 1];
      h[i] = ws->cs[i] * a + ws->sn[i] * b;
      h[i + 1] = -ws->sn[i] * a + ws->cs[i] * b;
    }
    const double a = h[k], b = h[k + 1];
    const double rho = hypot(a, b);
    const double cs = rho > 0.0 ? a / rho : 1.0, sn = rho > 0.0 ? b / rho : 0.0;
    ws->cs[k] = cs;
    ws->sn[k] = sn;
    h[k] = rho;
    h[k + 1] = 0.0;
    const double gk = ws->g[k];
    ws->g[k + 1] = -sn * gk;
    ws->g[k] = cs * gk;
    if (resid_out) *resid_out = fabs(ws->g[k + 1]);
  }
}

// Cycle start: g = beta e_0, s_0 = s0.
__global__ void k_begin(double beta, double s0, Workspace* __restrict__ ws)
{
  const int t = threadIdx.x;
  if (t <= kMaxM) ws->g[t] = t == 0 ? beta : 0.0;
  if (t == 0) ws->s[0] = s0;
}

// Cycle end: y = H(0:k,0:k)^{-1} g(0:k) in every block, x += sum_j s_j y_j V_j.
template <typename TV, int KK>
__global__ void __launch_bounds__(kThreads) k_update_x(int64_t n, const TV* __restrict__ V, int64_t ldv,
                                                       const Workspace* __restrict__ ws, double* __restrict__ x)
{
  __shared__ double coef[KK];
  if (threadIdx.x == 0) {
    double y[KK];
#pragma unroll
    for (int i = KK - 1; i >= 0; --i) {
      double acc = ws->g[i];
#pragma unroll
      for (int j = i + 1; j < KK; ++j) acc -= ws->H[j][i] * y[j];
      y[i] = acc / ws->H[i][i];
    }
#pragma unroll
    for (int j = 0; j < KK; ++j) coef[j] = ws->s[j] * y[j];
  }
  __syncthreads();
  double cj[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) cj[j] = coef[j];
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) s += cj[j] * (double)V[j * ldv + i];
    x[i] += s;
  }
}

// f(std::integral_constant<int, k>) for a runtime 1 <= k <= kMaxM
template <int LO, typename F>
void with_count(int k, F&& f)
{
  if constexpr (LO <= kMaxM) {
    if (k == LO) f(std::integral_constant<int, LO>{});
    else with_count<LO + 1>(k, f);
  }
}

template <typename TV>
int step(int64_t n, int m, int k, TV* V, int64_t ldv, void* wsp, double* resid_out, hipStream_t st)
{
  if (n < 1 || m < 1 || m > kMaxM || k < 0 || k >= m || !V || !wsp || ldv < n)
    return fail(GD_ERR_INVALID, "gd_arnoldi_step: bad arguments");
  Workspace* ws = static_cast<Workspace*>(wsp);
  const int nb = grid_for(n);
  with_count<1>(k + 1, [&](auto KK_) {
    constexpr int KK = decltype(KK_)::value;
    hipLaunchKernelGGL((k_pass_a<TV, KK>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
    hipLaunchKernelGGL(k_fin_a, dim3(1), dim3(kFinThreads), 0, st, nb, KK, ws);
    hipLaunchKernelGGL((k_pass_b<TV, KK>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
    hipLaunchKernelGGL(k_fin_b, dim3(1), dim3(kFinThreads), 0, st, nb, KK, ws);
    hipLaunchKernelGGL((k_pass_c<TV, KK>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
    hipLaunchKernelGGL(k_fin_c, dim3(1), dim3(kFinThreads), 0, st, nb, k, ws, resid_out);
  });
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template <typename TV>
int update(int64_t n, int m, int k, const TV* V, int64_t ldv, const void* wsp, double* x, hipStream_t st)
{
  if (n < 1 || m < 1 || m > kMaxM || k < 0 || k > m || !V || !wsp || !x || ldv < n)
    return fail(GD_ERR_INVALID, "gd_gmres_update: bad arguments");
  if (k == 0) return GD_OK;
  const Workspace* ws = static_cast<const Workspace*>(wsp);
  with_count<1>(k, [&](auto KK_) {
    constexpr int KK = decltype(KK_)::value;
    hipLaunchKernelGGL((k_update_x<TV, KK>), dim3(grid_for(n)), dim3(kThreads), 0, st, n, V, ldv, ws, x);
  });
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // namespace arnoldi
}  // namespace gd

using gd::fail;

extern "C" {

size_t gd_gmres_workspace_bytes(void) { return sizeof(gd::arnoldi::Workspace); }

int gd_gmres_begin(void* ws, double beta, double s0, gd_stream stream)
{
  if (!ws) return fail(GD_ERR_INVALID, "gd_gmres_begin: null workspace");
  hipLaunchKernelGGL(gd::arnoldi::k_begin, dim3(1), dim3(64), 0, (hipStream_t)stream, beta, s0,
                     static_cast<gd::arnoldi::Workspace*>(ws));
  GD_HIP(hipGetLastError());
  return GD_OK;
}

int gd_arnoldi_step_f64(int64_t n, int32_t m, int32_t k, double* V, int64_t ldv, void* ws, double* resid_out,
                        gd_stream stream)
{
  return gd::arnoldi::step<double>(n, m, k, V, ldv, ws, resid_out, (hipStream_t)stream);
}

int gd_arnoldi_step_f32(int64_t n, int32_t m, int32_t k, float* V, int64_t ldv, void* ws, double* resid_out,
                        gd_stream stream)
{
  return gd::arnoldi::step<float>(n, m, k, V, ldv, ws, resid_out, (hipStream_t)stream);
}

int gd_gmres_update_f64(int64_t n, int32_t m, int32_t k, const double* V, int64_t ldv, const void* ws, double* x,
                        gd_stream stream)
{
  return gd::arnoldi::update<double>(n, m, k, V, ldv, ws, x, (hipStream_t)stream);
}

int gd_gmres_update_f32(int64_t n, int32_t m, int32_t k, const float* V, int64_t ldv, const void* ws, double* x,
                        gd_stream stream)
{
  return gd::arnoldi::update<float>(n, m, k, V, ldv, ws, x, (hipStream_t)stream);
}

}  // extern "C"
