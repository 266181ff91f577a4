// Device-resident Arnoldi process for restarted GMRES(m), m <= 32.
//
// One iteration k (basis vectors V_0..V_k stored with scales s_j, true vector
// v_j = s_j V_j, and W = V_{k+1} holding K V_k) is three fused passes over the
// basis, each followed by a one-block kernel that sums the block partials in a
// fixed order (deterministic) and finishes the pass's scalar work:
//   pass A: h1_j = s_j <V_j, W>                         (classical GS)
//   pass B: W -= sum_j s_j h1_j V_j, then h2_j = s_j <V_j, W>, ||W||^2
//           (re-orthogonalisation, the "twice is enough" CGS2 step)
//   pass C: Z = (W - sum_j s_j h2_j V_j) / p with p^2 = ||W||^2 - sum_j h2_j^2
//           (Pythagoras), stored as V_{k+1}; nu = ||Z|| exactly; s_{k+1} = 1/nu
// and the finishing kernel of pass C forms the Hessenberg column
// H_{jk} = s_k (h1_j + h2_j), H_{k+1,k} = s_k p nu, applies the previous Givens
// rotations and forms the new one (gmres.hpp:170-210).  Stored vectors are
// kept at unit norm (up to rounding), so the operator is applied with scale 1
// and FP32 bases cannot drift out of range.
//
// The residual estimate |g_{k+1}| goes to the caller's pointer (pinned host
// memory, so a host loop can poll it with no copy or event in the stream) from
// the first block of the next kernel on the workspace (the next step's pass A
// or the cycle's update), early in a kernel that still runs for a while: a
// store to host memory at the end of a kernel holds the kernel's completion
// for the bus round trip.
//
// The reductions end in separate one-block kernels rather than in the last
// block of each pass: a device-scope fence per block (needed to publish the
// partials across the XCDs' L2 caches) cost more than the extra launch.
//
// FP32 bases are read 16 bytes per lane (4 values) where the basis and its
// leading dimension are 16-byte aligned (pass B, which keeps every loaded
// value live, 8 bytes); FP64 bases one value per lane, which measured as fast.
//
// The cycle end solves the triangular system H y = g in every block (k <= 32,
// cheap) and adds x += sum_j s_j y_j V_j in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gd_internal.hpp"
#include "gd_reduce.hpp"

namespace gd {
namespace arnoldi {

using red::kFinThreads;
using red::kMaxBlocks;
using red::kThreads;
constexpr int kMaxM = 32;           // restart length
constexpr int kStride = kMaxM + 2;  // partial columns per block: k+1 dots + norm
static_assert(kFinThreads == kMaxBlocks, "one block's partials per finishing thread");

struct Workspace {
  double H[kMaxM][kMaxM + 1];  // column k, rows 0..k+1 (rotated)
  double cs[kMaxM], sn[kMaxM];
  double g[kMaxM + 1];
  double s[kMaxM + 1];
  double c[kMaxM];   // coefficients on the stored vectors for the next update pass
  double h1[kMaxM];  // accumulated true GS coefficients of the current column
  double pnorm, pinv;
  double* pending;   // caller pointer still owed the last residual estimate
  double pending_value;
  double partial[kStride][kMaxBlocks];  // [column][block]: a column is contiguous
};

// W values of one vector, loaded or stored as one access of up to 16 bytes
// alignment (a 32-byte pack is two 16-byte accesses)
template <typename T, int W>
struct alignas(W * sizeof(T) > 16 ? 16 : W * sizeof(T)) Pack {
  T v[W];
};

template <typename T, int W>
__device__ __forceinline__ Pack<T, W> ld(const T* p)
{
  return *reinterpret_cast<const Pack<T, W>*>(p);
}

template <typename T, int W>
__device__ __forceinline__ void st(T* p, const Pack<T, W>& v)
{
  *reinterpret_cast<Pack<T, W>*>(p) = v;
}

using red::grid_for;

template <int K>
__device__ __forceinline__ void store_block_partials(const double (&acc)[K], Workspace* ws)
{
  red::store_block_partials<K>(acc, ws->partial);
}

template <int K>
__device__ __forceinline__ void sum_partials(const Workspace* ws, int nblk, double* col)
{
  red::sum_partials<K>(ws->partial, nblk, col);
}

__device__ __forceinline__ void publish_pending(Workspace* ws) { red::publish_pending(&ws->pending, &ws->pending_value); }

// The basis passes take the number of stored vectors KK as a template
// parameter: every load of a pass is then unconditional and independent, so
// all of them are in flight together (a runtime count made each load a
// predicated branch followed by a wait).  W values per lane and access; the
// n % W values past the last whole pack are taken one per lane afterwards.

// Pass A: partial row = <V_j, W>, j < KK.
template <typename TV, int KK, int W>
__global__ void __launch_bounds__(kThreads) k_pass_a(int64_t n, const TV* __restrict__ V, int64_t ldv,
                                                     Workspace* __restrict__ ws)
{
  publish_pending(ws);
  const TV* __restrict__ w = V + (int64_t)KK * ldv;
  double acc[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) acc[j] = 0.0;
  const int64_t nv = n / W, stride = (int64_t)gridDim.x * kThreads;
  const int64_t t0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (int64_t i = t0; i < nv; i += stride) {
    const Pack<TV, W> wp = ld<TV, W>(w + i * W);
#pragma unroll
    for (int j = 0; j < KK; ++j) {
      const Pack<TV, W> vp = ld<TV, W>(V + j * ldv + i * W);
#pragma unroll
      for (int l = 0; l < W; ++l) acc[j] += (double)vp.v[l] * (double)wp.v[l];
    }
  }
  for (int64_t i = nv * W + t0; i < n; i += stride) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < KK; ++j) acc[j] += (double)V[j * ldv + i] * wi;
  }
  store_block_partials<KK>(acc, ws);
}

// h1_j = s_j <V_j, W>, c_j = s_j h1_j.
template <int KK>
__global__ void __launch_bounds__(kFinThreads) k_fin_a(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[KK];
  sum_partials<KK>(ws, nblk, col);
  for (int j = threadIdx.x; j < KK; j += kFinThreads) {
    const double h = ws->s[j] * col[j];
    ws->h1[j] = h;
    ws->c[j] = ws->s[j] * h;
  }
}

// Pass B: W -= sum_j c_j V_j; partial row = <V_j, W_new>, ||W_new||^2.
template <typename TV, int KK, int W>
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
  const int64_t nv = n / W, stride = (int64_t)gridDim.x * kThreads;
  const int64_t t0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (int64_t i = t0; i < nv; i += stride) {
    Pack<TV, W> vp[KK];
    double s[W];
#pragma unroll
    for (int l = 0; l < W; ++l) s[l] = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) {
      vp[j] = ld<TV, W>(V + j * ldv + i * W);
#pragma unroll
      for (int l = 0; l < W; ++l) s[l] += cj[j] * (double)vp[j].v[l];
    }
    const Pack<TV, W> wp = ld<TV, W>(w + i * W);
    Pack<TV, W> wt;
#pragma unroll
    for (int l = 0; l < W; ++l) wt.v[l] = (TV)((double)wp.v[l] - s[l]);
    st<TV, W>(w + i * W, wt);
#pragma unroll
    for (int l = 0; l < W; ++l) {
      const double wr = (double)wt.v[l];
#pragma unroll
      for (int j = 0; j < KK; ++j) acc[j] += (double)vp[j].v[l] * wr;
      acc[KK] += wr * wr;
    }
  }
  for (int64_t i = nv * W + t0; i < n; i += stride) {
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
  store_block_partials<KK + 1>(acc, ws);
}

// h2_j = s_j <V_j, W>, h1_j += h2_j, c_j = s_j h2_j,
// p = sqrt(||W||^2 - sum_j h2_j^2).
template <int KK>
__global__ void __launch_bounds__(kFinThreads) k_fin_b(int nblk, Workspace* __restrict__ ws)
{
  __shared__ double col[KK + 1];
  __shared__ double h2sq[KK];
  sum_partials<KK + 1>(ws, nblk, col);
  for (int j = threadIdx.x; j < KK; j += kFinThreads) {
    const double h2 = ws->s[j] * col[j];
    ws->h1[j] += h2;
    ws->c[j] = ws->s[j] * h2;
    h2sq[j] = h2 * h2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) ss += h2sq[j];
    const double p2 = col[KK] - ss;
    const double p = p2 > 0.0 ? sqrt(p2) : 0.0;
    ws->pnorm = p;
    ws->pinv = p > 0.0 ? 1.0 / p : 0.0;
  }
}

// Pass C: Z = (W - sum_j c_j V_j) / p stored as V_KK; partial = ||Z||^2.
template <typename TV, int KK, int W>
__global__ void __launch_bounds__(kThreads) k_pass_c(int64_t n, TV* __restrict__ V, int64_t ldv,
                                                     Workspace* __restrict__ ws)
{
  TV* __restrict__ w = V + (int64_t)KK * ldv;
  const double pinv = ws->pinv;
  double cj[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) cj[j] = ws->c[j];
  double acc[1] = {0.0};
  const int64_t nv = n / W, stride = (int64_t)gridDim.x * kThreads;
  const int64_t t0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (int64_t i = t0; i < nv; i += stride) {
    double s[W];
#pragma unroll
    for (int l = 0; l < W; ++l) s[l] = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) {
      const Pack<TV, W> vp = ld<TV, W>(V + j * ldv + i * W);
#pragma unroll
      for (int l = 0; l < W; ++l) s[l] += cj[j] * (double)vp.v[l];
    }
    const Pack<TV, W> wp = ld<TV, W>(w + i * W);
    Pack<TV, W> zt;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      zt.v[l] = (TV)(((double)wp.v[l] - s[l]) * pinv);
      const double zr = (double)zt.v[l];
      acc[0] += zr * zr;
    }
    st<TV, W>(w + i * W, zt);
  }
  for (int64_t i = nv * W + t0; i < n; i += stride) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) s += cj[j] * (double)V[j * ldv + i];
    const TV zt = (TV)(((double)w[i] - s) * pinv);
    w[i] = zt;
    const double zr = (double)zt;
    acc[0] += zr * zr;
  }
  store_block_partials<1>(acc, ws);
}

// Column k = KK - 1 of the Hessenberg matrix and its Givens rotation; the
// residual estimate |g_{k+1}| is left for the next kernel to publish.
template <int KK>
__global__ void __launch_bounds__(kFinThreads) k_fin_c(int nblk, Workspace* __restrict__ ws,
                                                       double* __restrict__ resid_out)
{
  constexpr int k = KK - 1;
  __shared__ double col[1];
  __shared__ double hc[KK + 1];
  sum_partials<1>(ws, nblk, col);
  const double sk = ws->s[k];
  for (int j = threadIdx.x; j <= k; j += kFinThreads) hc[j] = sk * ws->h1[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nu = sqrt(col[0]);
    ws->s[k + 1] = nu > 0.0 ? 1.0 / nu : 0.0;
    double h[KK + 1];
#pragma unroll
    for (int j = 0; j <= k; ++j) h[j] = hc[j];
    h[k + 1] = sk * ws->pnorm * nu;
#pragma unroll
    for (int i = 0; i < k; ++i) {
      const double a = h[i], b = h[i + 1], c = ws->cs[i], sn = ws->sn[i];
      h[i] = c * a + sn * b;
      h[i + 1] = -sn * a + c * b;
    }
    const double a = h[k], b = h[k + 1];
    const double rho = hypot(a, b);
    const double cs = rho > 0.0 ? a / rho : 1.0, sn = rho > 0.0 ? b / rho : 0.0;
    ws->cs[k] = cs;
    ws->sn[k] = sn;
    h[k] = rho;
    h[k + 1] = 0.0;
#pragma unroll
    for (int j = 0; j <= k + 1; ++j) ws->H[k][j] = h[j];
    const double gk = ws->g[k];
    const double g1 = -sn * gk;
    ws->g[k + 1] = g1;
    ws->g[k] = cs * gk;
    ws->pending = resid_out;
    ws->pending_value = fabs(g1);
  }
}

// Cycle start: g = beta e_0, s_0 = s0, no residual estimate owed (the
// workspace may hold anything before the first cycle).
__global__ void k_begin(double beta, double s0, Workspace* __restrict__ ws)
{
  const int t = threadIdx.x;
  if (t <= kMaxM) ws->g[t] = t == 0 ? beta : 0.0;
  if (t == 0) {
    ws->s[0] = s0;
    ws->pending = nullptr;
  }
}

// Cycle end: y = H(0:k,0:k)^{-1} g(0:k) in every block, x += sum_j s_j y_j V_j.
template <typename TV, int KK, int W>
__global__ void __launch_bounds__(kThreads) k_update_x(int64_t n, const TV* __restrict__ V, int64_t ldv,
                                                       Workspace* __restrict__ ws, double* __restrict__ x)
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
  publish_pending(ws);
  __syncthreads();
  double cj[KK];
#pragma unroll
  for (int j = 0; j < KK; ++j) cj[j] = coef[j];
  const int64_t nv = n / W, stride = (int64_t)gridDim.x * kThreads;
  const int64_t t0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (int64_t i = t0; i < nv; i += stride) {
    double s[W];
#pragma unroll
    for (int l = 0; l < W; ++l) s[l] = 0.0;
#pragma unroll
    for (int j = 0; j < KK; ++j) {
      const Pack<TV, W> vp = ld<TV, W>(V + j * ldv + i * W);
#pragma unroll
      for (int l = 0; l < W; ++l) s[l] += cj[j] * (double)vp.v[l];
    }
    Pack<double, W> xp = ld<double, W>(x + i * W);
#pragma unroll
    for (int l = 0; l < W; ++l) xp.v[l] += s[l];
    st<double, W>(x + i * W, xp);
  }
  for (int64_t i = nv * W + t0; i < n; i += stride) {
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

// values per lane and access: 4 for a 16-byte aligned FP32 basis whose
// leading dimension is a whole number of 16-byte packs (and 16-byte aligned x
// for the update), else 1
template <typename TV>
bool packed(const void* V, int64_t ldv, const void* x = nullptr)
{
  return sizeof(TV) == 4 && reinterpret_cast<uintptr_t>(V) % 16 == 0 && (ldv * (int64_t)sizeof(TV)) % 16 == 0 &&
         reinterpret_cast<uintptr_t>(x) % 16 == 0;
}

template <typename TV>
int step(int64_t n, int m, int k, TV* V, int64_t ldv, void* wsp, double* resid_out, hipStream_t st)
{
  if (n < 1 || m < 1 || m > kMaxM || k < 0 || k >= m || !V || !wsp || ldv < n)
    return fail(GD_ERR_INVALID, "gd_arnoldi_step: bad arguments");
  Workspace* ws = static_cast<Workspace*>(wsp);
  auto launch = [&](auto W_) {
    constexpr int W = decltype(W_)::value;
    // pass B keeps every loaded value until the updated W is known: at most
    // two values per lane, so 32 stored vectors stay in registers
    constexpr int WB = W < 2 ? W : 2;
    const int nb = grid_for(n, W);
    with_count<1>(k + 1, [&](auto KK_) {
      constexpr int KK = decltype(KK_)::value;
      hipLaunchKernelGGL((k_pass_a<TV, KK, W>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
      hipLaunchKernelGGL((k_fin_a<KK>), dim3(1), dim3(kFinThreads), 0, st, nb, ws);
      hipLaunchKernelGGL((k_pass_b<TV, KK, WB>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
      hipLaunchKernelGGL((k_fin_b<KK>), dim3(1), dim3(kFinThreads), 0, st, nb, ws);
      hipLaunchKernelGGL((k_pass_c<TV, KK, W>), dim3(nb), dim3(kThreads), 0, st, n, V, ldv, ws);
      hipLaunchKernelGGL((k_fin_c<KK>), dim3(1), dim3(kFinThreads), 0, st, nb, ws, resid_out);
    });
  };
  if (packed<TV>(V, ldv)) launch(std::integral_constant<int, 4>{});
  else launch(std::integral_constant<int, 1>{});
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template <typename TV>
int update(int64_t n, int m, int k, const TV* V, int64_t ldv, void* wsp, double* x, hipStream_t st)
{
  if (n < 1 || m < 1 || m > kMaxM || k < 0 || k > m || !V || !wsp || !x || ldv < n)
    return fail(GD_ERR_INVALID, "gd_gmres_update: bad arguments");
  if (k == 0) return GD_OK;
  Workspace* ws = static_cast<Workspace*>(wsp);
  auto launch = [&](auto W_) {
    constexpr int W = decltype(W_)::value;
    with_count<1>(k, [&](auto KK_) {
      constexpr int KK = decltype(KK_)::value;
      hipLaunchKernelGGL((k_update_x<TV, KK, W>), dim3(grid_for(n, W)), dim3(kThreads), 0, st, n, V, ldv, ws, x);
    });
  };
  if (packed<TV>(V, ldv, x)) launch(std::integral_constant<int, 4>{});
  else launch(std::integral_constant<int, 1>{});
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

int gd_gmres_update_f64(int64_t n, int32_t m, int32_t k, const double* V, int64_t ldv, void* ws, double* x,
                        gd_stream stream)
{
  return gd::arnoldi::update<double>(n, m, k, V, ldv, ws, x, (hipStream_t)stream);
}

int gd_gmres_update_f32(int64_t n, int32_t m, int32_t k, const float* V, int64_t ldv, void* ws, double* x,
                        gd_stream stream)
{
  return gd::arnoldi::update<float>(n, m, k, V, ldv, ws, x, (hipStream_t)stream);
}

}  // extern "C"
