// Register-resident 3D kernel for GLL elements with Q == N (the p / q = p+1
// configurations of the benchmarks): fused mass + upwind advection apply,
// cell + interior-facet + boundary-facet terms in one pass.
//
// Mapping: each element is split over LANES adjacent lanes along z; lane h
// holds the PL = N / LANES z-planes [h*PL, h*PL+PL) of every element tensor
// in registers.  x- and y-contractions are in-register; z-contractions gather
// the other lanes' planes with DPP quad broadcasts (no LDS, no barriers).
// Global loads/stores are element-contiguous (an element's lanes are
// adjacent, so a wave covers 64/LANES consecutive elements).
//
// What is computed is GenDiL's element-centric operator
// (massoperator.hpp:39-89, advectionoperator.hpp:106-441), re-factored:
//   * the cell test-function/gradient contraction uses the basis-derivative
//     table Gb[q][i] = l_i'(x_q) (equal to GenDiL's collocated Gq * B),
//     y = Bz^T (Gbx^T By^T D0 + Bx^T Gby^T D1 + Bx^T By^T Dm) + Gbz^T Bx^T By^T D2;
//   * GLL face traces are slab selections, and the z-face own trace is the
//     x-y interpolated boundary plane of the cell interpolation;
//   * when the sign of beta.n is the same at every point of a face across the
//     whole wave, the trace the upwind flux does not select is skipped.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {
namespace k3 {

// ---- cross-lane plane gather ------------------------------------------------
// Value of v held by lane `src` of this lane's group of LANES lanes.
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

// full[k][m] for all N planes from own[kk][m] (PL planes per lane)
template <int LANES, int PL, int M, typename T>
__device__ __forceinline__ void gather(const T (&own)[PL][M], T (&full)[PL * LANES][M])
{
  if constexpr (LANES == 1) {
#pragma unroll
    for (int kk = 0; kk < PL; ++kk)
#pragma unroll
      for (int m = 0; m < M; ++m) full[kk][m] = own[kk][m];
  } else {
    [&]<int... S>(std::integer_sequence<int, S...>) {
      (
          [&] {
#pragma unroll
            for (int kk = 0; kk < PL; ++kk)
#pragma unroll
              for (int m = 0; m < M; ++m) full[S * PL + kk][m] = bcast<LANES, S>(own[kk][m]);
          }(),
          ...);
    }(std::make_integer_sequence<int, LANES>{});
  }
}

// Cross-lane z-contraction: out[cc][m] = sum_k A(h*PL+cc, k) * in_full[k][m]
// where A(c, k) = TR ? Mat[k*N + c] : Mat[c*N + k]   (Mat is N x N, uniform)
template <int N, int LANES, int PL, int M, bool TR, bool ACC, typename T>
__device__ __forceinline__ void zcontract(const T* __restrict__ Mat, int h, const T (&in)[PL][M], T (&out)[PL][M])
{
  if constexpr (!ACC) {
#pragma unroll
    for (int cc = 0; cc < PL; ++cc)
#pragma unroll
      for (int m = 0; m < M; ++m) out[cc][m] = T(0);
  }
  // source lanes one at a time: broadcast each of its planes and accumulate,
  // so no copy of all N planes is ever held
  [&]<int... S>(std::integer_sequence<int, S...>) {
    (
        [&] {
#pragma unroll
          for (int kk = 0; kk < PL; ++kk) {
            const int k = S * PL + kk;
            T coef[PL];
#pragma unroll
            for (int cc = 0; cc < PL; ++cc) {
              // lane-dependent row h*PL+cc: select among the LANES candidates
              T c = TR ? Mat[k * N + cc] : Mat[cc * N + k];
#pragma unroll
              for (int g = 1; g < LANES; ++g) {
                const int row = g * PL + cc;
                const T cg = TR ? Mat[k * N + row] : Mat[row * N + k];
                c = (h == g) ? cg : c;
              }
              coef[cc] = c;
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
              const T t = bcast<LANES, S>(in[kk][m]);
#pragma unroll
              for (int cc = 0; cc < PL; ++cc) out[cc][m] += coef[cc] * t;
            }
          }
        }(),
        ...);
  }(std::make_integer_sequence<int, LANES>{});
}

struct Geo {
  int e0, e1, e2;   // element sub-indices
  long long nb[6];  // neighbour per local face (0..2: -x,-y,-z; 3..5: +x,+y,+z), -1 = boundary
};

template <typename T>
__device__ __forceinline__ Geo geometry(const Params<T>& p, long long e)
{
  Geo g;
  const long long M0 = p.M[0], M1 = p.M[1], M2 = p.M[2];
  const long long s1 = M0, s2 = M0 * M1;
  const long long r = e / M0;
  g.e0 = (int)(e - r * M0);
  g.e2 = (int)(r / M1);
  g.e1 = (int)(r - (long long)g.e2 * M1);
  const int sub[3] = {g.e0, g.e1, g.e2};
  const long long Ms[3] = {M0, M1, M2};
  const long long st[3] = {1, s1, s2};
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int a = f % 3;
    const bool plus = f >= 3;
    const bool per = (p.periodic_mask >> a) & 1;
    long long nb;
    if (Ms[a] == 1) nb = per ? e : -1;
    else if (!plus) nb = sub[a] > 0 ? e - st[a] : (per ? e + (Ms[a] - 1) * st[a] : -1);
    else nb = sub[a] < Ms[a] - 1 ? e + st[a] : (per ? e - (Ms[a] - 1) * st[a] : -1);
    g.nb[f] = nb;
  }
  return g;
}

// rank-1 separable factor of a field along one axis at sample index k
template <typename T, int Q>
__device__ __forceinline__ T factor(const FieldDev<T>& f, int axis, int sub, int k)
{
  const T* tb = f.tab[0][axis];
  return tb ? tb[sub * (Q + 2) + k] : T(1);
}

template <typename T, int Q>
__device__ __forceinline__ T field_at(const FieldDev<T>& f, const int (&sub)[3], int kx, int ky, int kz)
{
  if (f.rank == 0) return T(0);
  return f.scale[0] * factor<T, Q>(f, 0, sub[0], kx) * factor<T, Q>(f, 1, sub[1], ky) * factor<T, Q>(f, 2, sub[2], kz);
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

template <typename T, int NV>
struct alignas(16) Chunk {
  T v[NV];
};

// copy COUNT contiguous values, with 16-byte accesses when the block allows it
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
  const int per = nb / 8, rem = nb % 8;
  const int xcd = pb % 8, idx = pb / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

template <typename T, int N, int LANES, bool ADV, int NT>
__global__ void __launch_bounds__(NT) k3d_apply(Params<T> p, const T* __restrict__ x, T* __restrict__ y)
{
  constexpr int Q = N;
  constexpr int PL = N / LANES;
  constexpr int N2 = N * N, N3 = N * N * N;
  constexpr int EPB = NT / LANES;
  const int h = LANES == 1 ? 0 : (int)(threadIdx.x % LANES);
  const int blk = k3_remap(blockIdx.x, p.num_blocks);
  const long long e_raw = (long long)blk * EPB + threadIdx.x / LANES;
  const bool valid = e_raw < p.E;
  const long long e = valid ? e_raw : p.E - 1;

  const T* __restrict__ B = p.tab.B;    // [q][i]
  const T* __restrict__ Gb = p.tab.Gb;  // basis derivative [q][i]
  const T* __restrict__ w = p.tab.w;

  const Geo geo = geometry(p, e);
  const int sub[3] = {geo.e0, geo.e1, geo.e2};

  // ---- load own planes: U[kk][j][i] (16-byte vector loads when aligned) ----
  T U[PL][N][N];
  copy_block<T, PL * N2>(x + e * N3 + (long long)h * PL * N2, &U[0][0][0]);

  T Y[PL][N][N];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i) Y[kk][j][i] = T(0);

  const T detJ = p.detJ;
  const bool sep1 = p.beta_kind == GD_BETA_SEPARABLE;  // rank <= 1 checked on the host

  // ---- x-interpolation: Tx[kk][j][a] ----
  T Tx[PL][N][Q];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T s = T(0);
#pragma unroll
        for (int i = 0; i < N; ++i) s += B[a * N + i] * U[kk][j][i];
        Tx[kk][j][a] = s;
      }

  if constexpr (ADV) {
    // ================= x faces (normal along i) =================
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int f = s ? 3 : 0;
      const int i_f = s ? N - 1 : 0;
      const long long nb = geo.nb[f];
      const bool boundary = nb < 0;
      const bool active = !(boundary && p.bc_mode == GD_BC_NONE);
      // dot = beta_x * (+-1/h_x) at points (c = h*PL+cc, b)
      T dot[PL][Q];
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          const int c = h * PL + cc;
          const T beta = field_at<T, Q>(p.beta[0], sub, s ? Q + 1 : Q, b, c);
          dot[cc][b] = beta * ((s ? T(1) : T(-1)) * p.invh[0]);
        }
      bool any_pos = false, any_nonpos = false;
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          any_pos |= dot[cc][b] > T(0);
          any_nonpos |= !(dot[cc][b] > T(0));
        }
      const bool need_own = __any(active && any_pos);
      const bool need_nb = __any(active && any_nonpos);
      T own[PL][Q] = {}, nbt[PL][Q] = {};
      if (need_own) {
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += B[b * N + j] * U[kk][j][i_f];
            t1[kk][b] = acc;
          }
        zcontract<N, LANES, PL, Q, false, false>(B, h, t1, own);
      }
      if (need_nb) {
        T nu[PL][N];
        const long long nbl = boundary ? e : nb;
        const T* src = x + nbl * N3 + (long long)h * PL * N2 + (N - 1 - i_f);
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int j = 0; j < N; ++j) nu[kk][j] = src[kk * N2 + j * N];
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += B[b * N + j] * nu[kk][j];
            t1[kk][b] = acc;
          }
        zcontract<N, LANES, PL, Q, false, false>(B, h, t1, nbt);
      }
      if (need_own || need_nb) {
        T F[PL][Q];
#pragma unroll
        for (int cc = 0; cc < PL; ++cc)
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            const int c = h * PL + cc;
            T ext;
            if (boundary) ext = p.bc ? sep_general<T, Q>(*p.bc, sub, s ? Q + 1 : Q, b, c) : T(0);
            else ext = nbt[cc][b];
            const T d = dot[cc][b];
            const T up = d > T(0) ? d * own[cc][b] : d * ext;
            F[cc][b] = active ? p.adv_coef * ((w[b] * w[c] * detJ) * up) : T(0);
          }
        // test functions: y[kk][j][i_f] += sum_c Bz[c][k] sum_b By[b][j] F[c][b]
        T G1[PL][N];
#pragma unroll
        for (int cc = 0; cc < PL; ++cc)
#pragma unroll
          for (int j = 0; j < N; ++j) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < Q; ++b) acc += B[b * N + j] * F[cc][b];
            G1[cc][j] = acc;
          }
        T G[PL][N];
        zcontract<N, LANES, PL, N, true, false>(B, h, G1, G);
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int j = 0; j < N; ++j) Y[kk][j][i_f] += G[kk][j];
      }
    }

    // ================= y faces (normal along j) =================
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int f = s ? 4 : 1;
      const int j_f = s ? N - 1 : 0;
      const long long nb = geo.nb[f];
      const bool boundary = nb < 0;
      const bool active = !(boundary && p.bc_mode == GD_BC_NONE);
      T dot[PL][Q];
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          const int c = h * PL + cc;
          const T beta = field_at<T, Q>(p.beta[1], sub, a, s ? Q + 1 : Q, c);
          dot[cc][a] = beta * ((s ? T(1) : T(-1)) * p.invh[1]);
        }
      bool any_pos = false, any_nonpos = false;
#pragma unroll
      for (int cc = 0; cc < PL; ++cc)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          any_pos |= dot[cc][a] > T(0);
          any_nonpos |= !(dot[cc][a] > T(0));
        }
      const bool need_own = __any(active && any_pos);
      const bool need_nb = __any(active && any_nonpos);
      T own[PL][Q] = {}, nbt[PL][Q] = {};
      if (need_own) {
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int a = 0; a < Q; ++a) t1[kk][a] = Tx[kk][j_f][a];
        zcontract<N, LANES, PL, Q, false, false>(B, h, t1, own);
      }
      if (need_nb) {
        T nu[PL][N];
        const long long nbl = boundary ? e : nb;
        const T* src = x + nbl * N3 + (long long)h * PL * N2 + (N - 1 - j_f) * N;
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int i = 0; i < N; ++i) nu[kk][i] = src[kk * N2 + i];
        T t1[PL][Q];
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < N; ++i) acc += B[a * N + i] * nu[kk][i];
            t1[kk][a] = acc;
          }
        zcontract<N, LANES, PL, Q, false, false>(B, h, t1, nbt);
      }
      if (need_own || need_nb) {
        T F[PL][Q];
#pragma unroll
        for (int cc = 0; cc < PL; ++cc)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            const int c = h * PL + cc;
            T ext;
            if (boundary) ext = p.bc ? sep_general<T, Q>(*p.bc, sub, a, s ? Q + 1 : Q, c) : T(0);
            else ext = nbt[cc][a];
            const T d = dot[cc][a];
            const T up = d > T(0) ? d * own[cc][a] : d * ext;
            F[cc][a] = active ? p.adv_coef * ((w[a] * w[c] * detJ) * up) : T(0);
          }
        T G1[PL][N];
#pragma unroll
        for (int cc = 0; cc < PL; ++cc)
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int a = 0; a < Q; ++a) acc += B[a * N + i] * F[cc][a];
            G1[cc][i] = acc;
          }
        T G[PL][N];
        zcontract<N, LANES, PL, N, true, false>(B, h, G1, G);
#pragma unroll
        for (int kk = 0; kk < PL; ++kk)
#pragma unroll
          for (int i = 0; i < N; ++i) Y[kk][j_f][i] += G[kk][i];
      }
    }
  }

  // ---- y-interpolation: Txy[kk][b][a] ----
  T Txy[PL][Q][Q];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int b = 0; b < Q; ++b)
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < N; ++j) s += B[b * N + j] * Tx[kk][j][a];
        Txy[kk][b][a] = s;
      }

  if constexpr (ADV) {
    // ================= z faces: the lane holding the boundary plane =================
    constexpr int NZ = LANES == 1 ? 2 : 1;
#pragma unroll
    for (int zz = 0; zz < NZ; ++zz) {
      // LANES == 1: both faces; otherwise lane 0 takes z-, lane LANES-1 takes z+
      const int s = LANES == 1 ? zz : (h == LANES - 1 ? 1 : 0);
      const bool mine = LANES == 1 || h == 0 || h == LANES - 1;
      const int f = s ? 5 : 2;
      const int kk_f = s ? PL - 1 : 0;
      const long long nb = geo.nb[f];
      const bool boundary = nb < 0;
      const bool active = mine && !(boundary && p.bc_mode == GD_BC_NONE);
      T dot[Q][Q];
#pragma unroll
      for (int b = 0; b < Q; ++b)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          const T beta = field_at<T, Q>(p.beta[2], sub, a, b, s ? Q + 1 : Q);
          dot[b][a] = beta * ((s ? T(1) : T(-1)) * p.invh[2]);
        }
      bool any_pos = false, any_nonpos = false;
#pragma unroll
      for (int b = 0; b < Q; ++b)
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          any_pos |= dot[b][a] > T(0);
          any_nonpos |= !(dot[b][a] > T(0));
        }
      const bool need_nb = __any(active && any_nonpos);
      T nbt[Q][Q] = {};
      if (need_nb) {
        const long long nbl = boundary ? e : nb;
        const T* src = x + nbl * N3 + (long long)(s ? 0 : N - 1) * N2;
        T nu[N][N];
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int i = 0; i < N; ++i) nu[j][i] = src[j * N + i];
        T t1[N][Q];
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < N; ++i) acc += B[a * N + i] * nu[j][i];
            t1[j][a] = acc;
          }
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) acc += B[b * N + j] * t1[j][a];
            nbt[b][a] = acc;
          }
      }
      if (__any(active)) {
        T F[Q][Q];
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            T ext;
            if (boundary) ext = p.bc ? sep_general<T, Q>(*p.bc, sub, a, b, s ? Q + 1 : Q) : T(0);
            else ext = nbt[b][a];
            const T d = dot[b][a];
            const T up = d > T(0) ? d * Txy[kk_f][b][a] : d * ext;
            F[b][a] = active ? p.adv_coef * ((w[a] * w[b] * detJ) * up) : T(0);
          }
        T G1[Q][N];
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int a = 0; a < Q; ++a) acc += B[a * N + i] * F[b][a];
            G1[b][i] = acc;
          }
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
          for (int i = 0; i < N; ++i) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < Q; ++b) acc += B[b * N + j] * G1[b][i];
            Y[kk_f][j][i] += acc;
          }
      }
    }
  }

  // ---- z-interpolation: Bu[cc][b][a] for own quadrature planes ----
  T Bu[PL][Q * Q];
  {
    T in[PL][Q * Q];
#pragma unroll
    for (int kk = 0; kk < PL; ++kk)
#pragma unroll
      for (int b = 0; b < Q; ++b)
#pragma unroll
        for (int a = 0; a < Q; ++a) in[kk][b * Q + a] = Txy[kk][b][a];
    zcontract<N, LANES, PL, Q * Q, false, false>(B, h, in, Bu);
  }

  // quadrature-point coefficient of the cell terms at (a, b, c)
  auto coef_beta = [&](int d, int a, int b, int c) -> T {
    const T beta = sep1 ? field_at<T, Q>(p.beta[d], sub, a, b, c) : T(0);
    return p.adv_coef * (-(w[a] * w[b] * w[c]) * detJ * (beta * p.invh[d]));
  };
  auto coef_mass = [&](int a, int b, int c) -> T {
    return p.mass_coef * ((w[a] * w[b] * w[c]) * detJ * field_at<T, Q>(*p.sigma, sub, a, b, c));
  };

  // ---- pass A: S2 = Bx^T By^T (c2 Bu);  Y += Gbz^T S2 ----
  if constexpr (ADV) {
    T S2[PL][N * N];
#pragma unroll
    for (int cc = 0; cc < PL; ++cc) {
      const int c = h * PL + cc;
      T E3[N][Q];  // [j][a]
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T v[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) v[b] = coef_beta(2, a, b, c) * Bu[cc][b * Q + a];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T acc = T(0);
#pragma unroll
          for (int b = 0; b < Q; ++b) acc += B[b * N + j] * v[b];
          E3[j][a] = acc;
        }
      }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T acc = T(0);
#pragma unroll
          for (int a = 0; a < Q; ++a) acc += B[a * N + i] * E3[j][a];
          S2[cc][j * N + i] = acc;
        }
    }
    T Z[PL][N * N];
    zcontract<N, LANES, PL, N * N, true, false>(Gb, h, S2, Z);
#pragma unroll
    for (int kk = 0; kk < PL; ++kk)
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) Y[kk][j][i] += Z[kk][j * N + i];
  }

  // ---- pass B: S01 = Gbx^T By^T (c0 Bu) + Bx^T (Gby^T (c1 Bu) + By^T (m Bu));  Y += Bz^T S01 ----
  {
    const bool mass = p.mass_coef != T(0);
    T S01[PL][N * N];
#pragma unroll
    for (int cc = 0; cc < PL; ++cc) {
      const int c = h * PL + cc;
      T E0[N][Q], E12[N][Q];
#pragma unroll
      for (int a = 0; a < Q; ++a) {
        T v0[Q], v1[Q], vm[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) {
          const T bu = Bu[cc][b * Q + a];
          v0[b] = ADV ? coef_beta(0, a, b, c) * bu : T(0);
          v1[b] = ADV ? coef_beta(1, a, b, c) * bu : T(0);
          vm[b] = mass ? coef_mass(a, b, c) * bu : T(0);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T a0 = T(0), a12 = T(0);
#pragma unroll
          for (int b = 0; b < Q; ++b) {
            if constexpr (ADV) {
              a0 += B[b * N + j] * v0[b];
              a12 += Gb[b * N + j] * v1[b];
            }
            a12 += B[b * N + j] * vm[b];
          }
          E0[j][a] = a0;
          E12[j][a] = a12;
        }
      }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T acc = T(0);
#pragma unroll
          for (int a = 0; a < Q; ++a) {
            if constexpr (ADV) acc += Gb[a * N + i] * E0[j][a];
            acc += B[a * N + i] * E12[j][a];
          }
          S01[cc][j * N + i] = acc;
        }
    }
    T Z[PL][N * N];
    zcontract<N, LANES, PL, N * N, true, false>(B, h, S01, Z);
#pragma unroll
    for (int kk = 0; kk < PL; ++kk)
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) Y[kk][j][i] += Z[kk][j * N + i];
  }

  // ---- store own planes ----
  if (valid) copy_block<T, PL * N2>(&Y[0][0][0], y + e * N3 + (long long)h * PL * N2);
}

}  // namespace k3
}  // namespace gd
