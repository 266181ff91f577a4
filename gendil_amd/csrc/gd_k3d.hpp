// Register-resident 3D kernel for GLL elements with Q == N (the p / q = p+1
// configurations of the benchmarks): fused mass + upwind advection apply,
// cell + interior-facet + boundary-facet terms in one pass.
//
// Mapping: each element is split over LANES adjacent lanes along z; lane h
// holds the PL = N / LANES z-planes [h*PL, h*PL+PL) of every element tensor
// in registers.  x- and y-contractions are in-register; z-contractions
// accumulate the other lanes' planes through DPP quad broadcasts (no LDS, no
// barriers).  Global loads/stores are element-contiguous.
//
// The 1D tables are compile-time constants (gd_consttab.hpp), so every
// contraction is a chain of FMAs against literals.  Coefficient fields are
// "one-axis" separable: component f(X) = scale_f * T_f(X_{axis_f}) (or a
// constant, axis -1); the axes are template parameters (AX) so the per-point
// factor is a register, not a select.
//
// What is computed is GenDiL's element-centric operator
// (massoperator.hpp:39-89, advectionoperator.hpp:106-441), re-factored:
//   * the cell gradient/test contraction uses the basis-derivative table
//     Gb[q][i] = l_i'(x_q) (equal to GenDiL's collocated Gq times B):
//     y = Bz^T (Gbx^T By^T D0 + Bx^T Gby^T D1 + Bx^T By^T Dm) + Gbz^T Bx^T By^T D2
//   * GLL face traces are slab selections; the z-face own trace is the
//     boundary plane of the x-y interpolation of the cell pass;
//   * when beta.n has one sign at every point of a face across the whole
//     wave, the trace the upwind flux does not select is not computed.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "gd_consttab.hpp"
#include "gd_internal.hpp"

namespace gd {
namespace k3 {

template <typename T>
struct Field1 {
  T scale;
  int axis;      // -1: constant; 0..2: table axis
  const T* tab;  // [M_axis][Q+2] samples (rule points, xi=0, xi=1)
};

template <typename T>
struct K3Params {
  int E, M0, M1, M2;
  int periodic;  // bit d = periodic axis d
  int bc_inflow;  // boundary faces: 0 skipped, 1 inflow with u+ = g
  int num_blocks;
  T invh[3];
  T detJ;
  T adv_coef;
  T mass_coef;
  Field1<T> beta[3];
  Field1<T> sigma;
  const FieldDev<T>* bc;  // general separable g (BCG instances only)
};

// axis signature: 4 digits base 4 (axis+1) for beta0, beta1, beta2, sigma
constexpr int sig(int b0, int b1, int b2, int s) { return (b0 + 1) + 4 * (b1 + 1) + 16 * (b2 + 1) + 64 * (s + 1); }
constexpr int sig_axis(int S, int f) { return ((S >> (2 * f)) & 3) - 1; }
// every one-axis signature, the axes read from the kernel arguments
constexpr int kSigAny = 256;
// the sample axis of field f in an instance: -1..2 compiled in, 3 run time
constexpr int field_axis(int S, int f) { return S == kSigAny ? 3 : sig_axis(S, f); }

// ---- cross-lane broadcast within groups of LANES lanes -------------------
template <int LANES, int SRC>
__device__ __forceinline__ float bcast(float v)
{
  if constexpr (LANES == 1) return v;
  else {
    constexpr int ctrl = LANES == 2 ? (SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6))
                                    : (SRC | (SRC << 2) | (SRC << 4) | (SRC << 6));
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
  }
}

template <int LANES, int SRC>
__device__ __forceinline__ double bcast(double v)
{
  if constexpr (LANES == 1) return v;
  else {
    constexpr int ctrl = LANES == 2 ? (SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6))
                                    : (SRC | (SRC << 2) | (SRC << 4) | (SRC << 6));
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f)
{
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

// Lane-dependent coefficient A[row = h*PL + cc][col] of a constexpr N x N
// matrix: select among the LANES candidate rows.
template <int N, int LANES, int PL, typename T, typename Mat>
__device__ __forceinline__ T lane_row(const Mat& A, int h, int cc, int col)
{
  T c = (T)A(cc, col);
#pragma unroll
  for (int g = 1; g < LANES; ++g) c = (h == g) ? (T)A(g * PL + cc, col) : c;
  return c;
}

// Plane-at-a-time z contraction: every lane has computed plane P (its local
// plane cc, global plane c = h*PL + cc);  out[kk][m] += A(k_out, c) * P_c[m]
// for the lane's output planes k_out = h*PL + kk, summing over the LANES
// source lanes.  A(k, c) is constexpr.
template <int N, int LANES, int PL, int M, typename T, typename Mat>
__device__ __forceinline__ void zacc_plane(const Mat& A, int h, int cc, const T (&P)[M], T (&out)[PL][M])
{
  sfor<0, LANES>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
    const int c = S * PL + cc;  // global source plane
    T coef[PL];
#pragma unroll
    for (int kk = 0; kk < PL; ++kk) coef[kk] = lane_row<N, LANES, PL, T>([&](int k, int cidx) { return A(k, cidx); }, h, kk, c);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const T t = bcast<LANES, S>(P[m]);
#pragma unroll
      for (int kk = 0; kk < PL; ++kk) out[kk][m] += coef[kk] * t;
    }
  });
}

// Full z-interpolation of own planes in[kk][m] (node planes) into out[cc][m]
// (quadrature planes): out = sum_k A(c_out, k) in_k.
template <int N, int LANES, int PL, int M, typename T, typename Mat>
__device__ __forceinline__ void zcontract(const Mat& A, int h, const T (&in)[PL][M], T (&out)[PL][M])
{
#pragma unroll
  for (int cc = 0; cc < PL; ++cc)
#pragma unroll
    for (int m = 0; m < M; ++m) out[cc][m] = T(0);
#pragma unroll
  for (int kk = 0; kk < PL; ++kk) zacc_plane<N, LANES, PL, M>(A, h, kk, in[kk], out);
}

template <typename T, int NV>
struct alignas(16) Chunk {
  T v[NV];
};

template <typename T, int COUNT>
__device__ __forceinline__ void copy_block(const T* __restrict__ src, T* __restrict__ dst)
{
  constexpr int VW = 16 / sizeof(T);
  if constexpr ((COUNT * sizeof(T)) % 16 == 0) {
    const Chunk<T, VW>* s = reinterpret_cast<const Chunk<T, VW>*>(src);
    Chunk<T, VW>* d = reinterpret_cast<Chunk<T, VW>*>(dst);
#pragma unroll
    for (int v = 0; v < COUNT / VW; ++v) d[v] = s[v];
  } else {
#pragma unroll
    for (int v = 0; v < COUNT; ++v) dst[v] = src[v];
  }
}

__device__ __forceinline__ int k3_remap(int pb, int nb)
{
  // blocks are dispatched round-robin over the 8 XCDs: give each XCD one
  // contiguous range of elements so face neighbours hit its own L2
  const int per = nb / 8, rem = nb % 8;
  const int xcd = pb % 8, idx = pb / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

template <typename T, int Q>
__device__ __forceinline__ T sep_general(const FieldDev<T>& f, const int (&sub)[3], int kx, int ky, int kz)
{
  T sum = T(0);
  const int k[3] = {kx, ky, kz};
  for (int r = 0; r < f.rank; ++r) {
    T t = f.scale[r];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* tb = f.tab[r][d];
      if (tb) t *= tb[sub[d] * (Q + 2) + k[d]];
    }
    sum += t;
  }
  return sum;
}

// Per-element samples of a one-axis field along its axis (Q+2 values).
template <typename T, int Q, int AX>
struct FieldSamples {
  T v[AX >= 0 ? Q + 2 : 1];
  __device__ __forceinline__ void load(const Field1<T>& f, const int (&sub)[3])
  {
    if constexpr (AX >= 0) {
      const T* tb = f.tab + sub[AX] * (Q + 2);
#pragma unroll
      for (int k = 0; k < Q + 2; ++k) v[k] = f.scale * tb[k];
    } else {
      v[0] = f.scale;
    }
  }
  // value at compile-time sample indices (kx, ky, kz)
  __device__ __forceinline__ T at(int kx, int ky, int kz) const
  {
    if constexpr (AX == 0) return v[kx];
    else if constexpr (AX == 1) return v[ky];
    else if constexpr (AX == 2) return v[kz];
    else return v[0];
  }
  // z sample index h*PL + cc, with h lane dependent
  template <int LANES, int PL>
  __device__ __forceinline__ T atc(int kx, int ky, int h, int cc) const
  {
    if constexpr (AX == 2) {
      T r = v[cc];
#pragma unroll
      for (int g = 1; g < LANES; ++g) r = (h == g) ? v[g * PL + cc] : r;
      return r;
    } else {
      return at(kx, ky, 0);
    }
  }
  // z sample index Q (plus = false) or Q+1, plus lane dependent
  __device__ __forceinline__ T atzs(int kx, int ky, bool plus) const
  {
    if constexpr (AX == 2) return plus ? v[Q + 1] : v[Q];
    else return at(kx, ky, 0);
  }
};

// Run-time axis (one instance serves every one-axis signature): a constant
// field holds its value at every sample index, and a lookup is the one-hot
// weighted sum wx v[kx] + wy v[ky] + wz v[kz] with the wave-uniform weights of
// the field's axis.  Products with 0 and 1 and sums with 0 are exact for the
// finite samples, and unlike a select between elements of one array (which
// the optimiser turns into a load at a run-time index, putting the samples on
// the stack) the sum leaves every sample in a register.
template <typename T, int Q>
struct FieldSamples<T, Q, 3> {
  T v[Q + 2];
  T wx, wy, wz;
  __device__ __forceinline__ void load(const Field1<T>& f, const int (&sub)[3])
  {
    wx = f.axis <= 0 ? T(1) : T(0);  // a constant is read at the x index
    wy = f.axis == 1 ? T(1) : T(0);
    wz = f.axis == 2 ? T(1) : T(0);
    if (f.axis >= 0) {
      const int s = f.axis == 0 ? sub[0] : (f.axis == 1 ? sub[1] : sub[2]);
      const T* tb = f.tab + s * (Q + 2);
#pragma unroll
      for (int k = 0; k < Q + 2; ++k) v[k] = f.scale * tb[k];
    } else {
#pragma unroll
      for (int k = 0; k < Q + 2; ++k) v[k] = f.scale;
    }
  }
  __device__ __forceinline__ T at(int kx, int ky, int kz) const { return wx * v[kx] + wy * v[ky] + wz * v[kz]; }
  template <int LANES, int PL>
  __device__ __forceinline__ T atc(int kx, int ky, int h, int cc) const
  {
    // z sample h * PL + cc: the lane's own, again as a one-hot sum
    T vz = T(0);
#pragma unroll
    for (int g = 0; g < LANES; ++g) vz += (h == g ? T(1) : T(0)) * v[g * PL + cc];
    return wx * v[kx] + wy * v[ky] + wz * vz;
  }
  __device__ __forceinline__ T atzs(int kx, int ky, bool plus) const
  {
    const T vz = (plus ? T(0) : T(1)) * v[Q] + (plus ? T(1) : T(0)) * v[Q + 1];
    return wx * v[kx] + wy * v[ky] + wz * vz;
  }
};

template <typename T, int N, int LANES, bool ADV, bool MASS, bool BCG, int SIG, int NT>
__global__ void __launch_bounds__(NT) k3d_apply(K3Params<T> p, const T* __restrict__ x, T* __restrict__ y)
{
  constexpr int Q = N;
  constexpr int PL = N / LANES;
  constexpr int N2 = N * N, N3 = N * N * N;
  constexpr int EPB = NT / LANES;
  constexpr auto TB = ct::make_gll_gauss<N, Q>();
  auto Bm = [](int q, int i) { return TB.B[q][i]; };
  auto BmT = [](int i, int q) { return TB.B[q][i]; };
  auto GbT = [](int i, int q) { return TB.Gb[q][i]; };

  const int h = LANES == 1 ? 0 : (int)(threadIdx.x % LANES);
  const int blk = k3_remap(blockIdx.x, p.num_blocks);
  const int e_raw = blk * EPB + (int)(threadIdx.x / LANES);
  const bool valid = e_raw < p.E;
  const int e = valid ? e_raw : p.E - 1;

  // ---- geometry (32-bit: the host guarantees E < 2^31) ----
  int sub[3];
  {
    const unsigned r = (unsigned)e / (unsigned)p.M0;
    sub[0] = e - (int)r * p.M0;
    sub[2] = (int)(r / (unsigned)p.M1);
    sub[1] = (int)r - sub[2] * p.M1;
  }
  auto neighbor = [&](int a, bool plus) -> int {
    const int Ma = a == 0 ? p.M0 : (a == 1 ? p.M1 : p.M2);
    const int st = a == 0 ? 1 : (a == 1 ? p.M0 : p.M0 * p.M1);
    const bool per = (p.periodic >> a) & 1;
    if (Ma == 1) return per ? e : -1;
    if (!plus) return sub[a] > 0 ? e - st : (per ? e + (Ma - 1) * st : -1);
    return sub[a] < Ma - 1 ? e + st : (per ? e - (Ma - 1) * st : -1);
  };

  constexpr int AX0 = field_axis(SIG, 0), AX1 = field_axis(SIG, 1), AX2 = field_axis(SIG, 2), AXS = field_axis(SIG, 3);
  FieldSamples<T, Q, AX0> f0;
  FieldSamples<T, Q, AX1> f1;
  FieldSamples<T, Q, AX2> f2;
  FieldSamples<T, Q, AXS> fs;
  if constexpr (ADV) {
    f0.load(p.beta[0], sub);
    f1.load(p.beta[1], sub);
    f2.load(p.beta[2], sub);
  }
  if constexpr (MASS) fs.load(p.sigma, sub);

  // ---- own planes ----
  T U[PL][N][N];
  copy_block<T, PL * N2>(x + (long long)e * N3 + h * PL * N2, &U[0][0][0]);

  // output accumulator Y[kk][j*N + i]
  T Y[PL][N * N];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int m = 0; m < N * N; ++m) Y[kk][m] = T(0);

  const T detJ = p.detJ;
  // Phase boundaries keep the scheduler from hoisting every face's neighbour
  // loads and tables to the top of the kernel, which would inflate register
  // pressure far beyond the live data.
#define GD_K3_PHASE() __builtin_amdgcn_sched_barrier(0)
  GD_K3_PHASE();

  // boundary value at a face point (sample indices)
  auto gval = [&](int kx, int ky, int kz) -> T {
    if constexpr (BCG) return p.bc ? sep_general<T, Q>(*p.bc, sub, kx, ky, kz) : T(0);
    else return T(0);
  };

  if constexpr (ADV) {
    // ================= x faces =================
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      GD_K3_PHASE();
      const int i_f = s ? N - 1 : 0;
      const int kn = s ? Q + 1 : Q;
      const int nb = neighbor(0, s);
      const bool boundary = nb < 0;
      const bool active = !(boundary && !p.bc_inflow);
      const T nrm = (s ? T(1) : T(-1)) * p.invh[0];
      T dot[PL][Q];
      bool any_pos = false, any_np = false;
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          dot[cc][b] = f0.template atc<LANES, PL>(kn, b, h, cc) * nrm;
          any_pos |= dot[cc][b] > T(0);
          any_np |= !(dot[cc][b] > T(0));
        }
      const bool need_own = __any(active && any_pos);
      const bool need_nb = __any(active && any_np);
      T own[PL][Q] = {}, nbt[PL][Q] = {};
      if (need_own) {
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += (T)TB.B[b][j] * U[kk][j][i_f];
            t1[kk][b] = acc;
          }
        zcontract<N, LANES, PL, Q>(Bm, h, t1, own);
      }
      if (need_nb) {
        const T* src = x + (long long)(boundary ? e : nb) * N3 + h * PL * N2 + (N - 1 - i_f);
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk) {
          T nu[N];
#pragma unroll
          for (int j = 0; j < N; ++j) nu[j] = src[kk * N2 + j * N];
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += (T)TB.B[b][j] * nu[j];
            t1[kk][b] = acc;
          }
        }
        zcontract<N, LANES, PL, Q>(Bm, h, t1, nbt);
      }
      if (need_own || need_nb) {
        T G[PL][N];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int j = 0; j < N; ++j) G[kk][j] = T(0);
#pragma unroll
        for (int cc = 0; cc < PL; ++cc) {
          const int c = h * PL + cc;
          T F[Q];
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            const T ext = boundary ? gval(kn, b, c) : nbt[cc][b];
            const T d = dot[cc][b];
            const T up = d > T(0) ? d * own[cc][b] : d * ext;
            F[b] = active ? p.adv_coef * ((T)TB.w[b] * detJ * up) : T(0);
          }
          // weight of the c direction is lane dependent: apply per plane
          T wc = (T)TB.w[cc];
#pragma unroll
          for (int g = 1; g < LANES; ++g) wc = (h == g) ? (T)TB.w[g * PL + cc] : wc;
          T G1[N];
#pragma unroll
          for (int j = 0; j < N; ++j) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < Q; ++b) acc += (T)TB.B[b][j] * F[b];
            G1[j] = wc * acc;
          }
          zacc_plane<N, LANES, PL, N>(BmT, h, cc, G1, G);
        }
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int j = 0; j < N; ++j) Y[kk][j * N + i_f] += G[kk][j];
      }
    }
  }

  GD_K3_PHASE();
  // x-interpolation Tx[kk][j][a]  (U is not needed after this)
  T Tx[PL][N][Q];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T s = T(0);
#pragma unroll
        for (int i = 0; i < N; ++i) s += (T)TB.B[a][i] * U[kk][j][i];
        Tx[kk][j][a] = s;
      }

  if constexpr (ADV) {
    // ================= y faces =================
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      GD_K3_PHASE();
      const int j_f = s ? N - 1 : 0;
      const int kn = s ? Q + 1 : Q;
      const int nb = neighbor(1, s);
      const bool boundary = nb < 0;
      const bool active = !(boundary && !p.bc_inflow);
      const T nrm = (s ? T(1) : T(-1)) * p.invh[1];
      T dot[PL][Q];
      bool any_pos = false, any_np = false;
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          dot[cc][a] = f1.template atc<LANES, PL>(a, kn, h, cc) * nrm;
          any_pos |= dot[cc][a] > T(0);
          any_np |= !(dot[cc][a] > T(0));
        }
      const bool need_own = __any(active && any_pos);
      const bool need_nb = __any(active && any_np);
      T own[PL][Q] = {}, nbt[PL][Q] = {};
      if (need_own) {
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int a = 0; a < Q; ++a) t1[kk][a] = Tx[kk][j_f][a];
        zcontract<N, LANES, PL, Q>(Bm, h, t1, own);
      }
      if (need_nb) {
        const T* src = x + (long long)(boundary ? e : nb) * N3 + h * PL * N2 + (N - 1 - j_f) * N;
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk) {
          T nu[N];
#pragma unroll
          for (int i = 0; i < N; ++i) nu[i] = src[kk * N2 + i];
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < N; ++i) acc += (T)TB.B[a][i] * nu[i];
            t1[kk][a] = acc;
          }
        }
        zcontract<N, LANES, PL, Q>(Bm, h, t1, nbt);
      }
      if (need_own || need_nb) {
        T G[PL][N];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int i = 0; i < N; ++i) G[kk][i] = T(0);
#pragma unroll
        for (int cc = 0; cc < PL; ++cc) {
          const int c = h * PL + cc;
          T wc = (T)TB.w[cc];
#pragma unroll
          for (int g = 1; g < LANES; ++g) wc = (h == g) ? (T)TB.w[g * PL + cc] : wc;
          T F[Q];
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            const T ext = boundary ? gval(a, kn, c) : nbt[cc][a];
            const T d = dot[cc][a];
            const T up = d > T(0) ? d * own[cc][a] : d * ext;
            F[a] = active ? p.adv_coef * ((T)TB.w[a] * detJ * up) : T(0);
          }
          T G1[N];
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int a = 0; a < Q; ++a) acc += (T)TB.B[a][i] * F[a];
            G1[i] = wc * acc;
          }
          zacc_plane<N, LANES, PL, N>(BmT, h, cc, G1, G);
        }
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int i = 0; i < N; ++i) Y[kk][j_f * N + i] += G[kk][i];
      }
    }
  }

  GD_K3_PHASE();
  // y-interpolation Txy[kk][b][a]
  T Txy[PL][Q * Q];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int b = 0; b < Q; ++b)
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < N; ++j) s += (T)TB.B[b][j] * Tx[kk][j][a];
        Txy[kk][b * Q + a] = s;
      }

  if constexpr (ADV) {
    // ================= z faces: lane holding the boundary plane =================
    constexpr int NZ = LANES == 1 ? 2 : 1;
#pragma unroll
    for (int zz = 0; zz < NZ; ++zz) {
      GD_K3_PHASE();
      const int s = LANES == 1 ? zz : (h == LANES - 1 ? 1 : 0);
      const bool mine = LANES == 1 || h == 0 || h == LANES - 1;
      const int kk_f = s ? PL - 1 : 0;
      const int kn = s ? Q + 1 : Q;
      const int nb = neighbor(2, s);
      const bool boundary = nb < 0;
      const bool active = mine && !(boundary && !p.bc_inflow);
      const T nrm = (s ? T(1) : T(-1)) * p.invh[2];
      T dot[Q * Q];
      bool any_np = false;
#pragma unroll
      for (int b = 0; b < Q; ++b)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          dot[b * Q + a] = f2.atzs(a, b, s != 0) * nrm;
          any_np |= !(dot[b * Q + a] > T(0));
        }
      const bool need_nb = __any(active && any_np);
      T nbt[Q * Q] = {};
      if (need_nb) {
        const T* src = x + (long long)(boundary ? e : nb) * N3 + (s ? 0 : N - 1) * N2;
        T t1[N][Q];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T nu[N];
#pragma unroll
          for (int i = 0; i < N; ++i) nu[i] = src[j * N + i];
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < N; ++i) acc += (T)TB.B[a][i] * nu[i];
            t1[j][a] = acc;
          }
        }
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += (T)TB.B[b][j] * t1[j][a];
            nbt[b * Q + a] = acc;
          }
      }
      if (__any(active)) {
        T G1[Q][N];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          T F[Q];
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            const T ext = boundary ? gval(a, b, kn) : nbt[b * Q + a];
            const T d = dot[b * Q + a];
            // own trace: boundary plane of the x-y interpolation.  When the
            // plane depends on the lane both candidates are first pinned in
            // registers: a select of two elements of one array would
            // otherwise become one load at a lane-dependent index, which puts
            // the whole array on the stack.
            T ownv;
            if constexpr (PL == 1 || LANES == 1) {
              ownv = Txy[kk_f][b * Q + a];
            } else {
              T lo = Txy[0][b * Q + a], hi = Txy[PL - 1][b * Q + a];
              __asm__ volatile("" : "+v"(lo), "+v"(hi));
              ownv = s ? hi : lo;
            }
            const T up = d > T(0) ? d * ownv : d * ext;
            F[a] = active ? p.adv_coef * ((T)(TB.w[a] * TB.w[b]) * detJ * up) : T(0);
          }
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int a = 0; a < Q; ++a) acc += (T)TB.B[a][i] * F[a];
            G1[b][i] = acc;
          }
        }
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < Q; ++b) acc += (T)TB.B[b][j] * G1[b][i];
            if constexpr (PL == 1 || LANES == 1) {
              Y[kk_f][j * N + i] += acc;
            } else {
              Y[0][j * N + i] += s ? T(0) : acc;
              Y[PL - 1][j * N + i] += s ? acc : T(0);
            }
          }
      }
    }
  }

  GD_K3_PHASE();
  // z-interpolation: Bu[cc][b*Q+a] for own quadrature planes
  T Bu[PL][Q * Q];
  zcontract<N, LANES, PL, Q * Q>(Bm, h, Txy, Bu);

  // ---- cell terms, one quadrature plane at a time, straight into Y ----
  // c_d(a,b,c) = K_d w_a w_b w_c beta_d,  K_d = -adv_coef detJ / h_d
  //   S2  = Bx^T By^T (c2 Bu)                          -> Y += Gbz^T S2
  //   S01 = Gbx^T By^T (c0 Bu) + Bx^T Gby^T (c1 Bu) + Bx^T By^T (m Bu)  -> Y += Bz^T S01
  const T K0 = -p.adv_coef * detJ * p.invh[0];
  const T K1 = -p.adv_coef * detJ * p.invh[1];
  const T K2 = -p.adv_coef * detJ * p.invh[2];
  const T Km = p.mass_coef * detJ;
#pragma unroll
  for (int cc = 0; cc < PL; ++cc) {
    GD_K3_PHASE();
    T wc = (T)TB.w[cc];
#pragma unroll
    for (int g = 1; g < LANES; ++g) wc = (h == g) ? (T)TB.w[g * PL + cc] : wc;
    if constexpr (ADV) {
      T E3[N][Q];
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T v[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b)
          v[b] = (K2 * (T)(TB.w[a] * TB.w[b]) * wc) * f2.template atc<LANES, PL>(a, b, h, cc) * Bu[cc][b * Q + a];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T acc = T(0);
#pragma unroll
          for (int b = 0; b < Q; ++b) acc += (T)TB.B[b][j] * v[b];
          E3[j][a] = acc;
        }
      }
      T S2[N * N];
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T acc = T(0);
#pragma unroll
          for (int a = 0; a < Q; ++a) acc += (T)TB.B[a][i] * E3[j][a];
          S2[j * N + i] = acc;
        }
      zacc_plane<N, LANES, PL, N * N>(GbT, h, cc, S2, Y);
    }
    {
      T E0[N][Q], E12[N][Q];
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T v0[Q], v1[Q], vm[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          const T bu = Bu[cc][b * Q + a];
          const T wab = (T)(TB.w[a] * TB.w[b]) * wc;
          v0[b] = ADV ? (K0 * wab) * f0.template atc<LANES, PL>(a, b, h, cc) * bu : T(0);
          v1[b] = ADV ? (K1 * wab) * f1.template atc<LANES, PL>(a, b, h, cc) * bu : T(0);
          vm[b] = MASS ? (Km * wab) * fs.template atc<LANES, PL>(a, b, h, cc) * bu : T(0);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T a0 = T(0), a12 = T(0);
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            if constexpr (ADV) {
              a0 += (T)TB.B[b][j] * v0[b];
              a12 += (T)TB.Gb[b][j] * v1[b];
            }
            if constexpr (MASS) a12 += (T)TB.B[b][j] * vm[b];
          }
          E0[j][a] = a0;
          E12[j][a] = a12;
        }
      }
      T S01[N * N];
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T acc = T(0);
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            if constexpr (ADV) acc += (T)TB.Gb[a][i] * E0[j][a];
            acc += (T)TB.B[a][i] * E12[j][a];
          }
          S01[j * N + i] = acc;
        }
      zacc_plane<N, LANES, PL, N * N>(BmT, h, cc, S01, Y);
    }
  }

  if (valid) copy_block<T, PL * N2>(&Y[0][0], y + (long long)e * N3 + h * PL * N2);
}

}  // namespace k3
}  // namespace gd
