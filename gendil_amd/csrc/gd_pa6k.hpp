// Kronecker-factored p-adaptive 6D apply for phase-space fields (GLL basis,
// orders n_e = p_e + 1 in {2, 3, 4}, Gauss rule q_e = n_e), one launch.
//
// Fields served: every component beta_c depends at most on its paired
// coordinate, beta_c = a_c X_pc + b_c with pc = (c + 3) mod 6 (phase-space
// advection: beta = (X3, X4, X5, v) and the like), and keeps one sign over
// the whole mesh; sigma and inflow data constant.  Then every integral of the
// element operator is a Kronecker product of exact 1D integrals (the Gauss
// rule of max(q_e, q_nb) points integrates each of them exactly, including
// the mixed-order face products and the linear beta factor), so this is the
// operator of the quadrature kernel (gd_padapt.hpp; advectionoperator.hpp:
// 137-227 with the shared face rule of :596-676) regrouped:
//   cell c : -Kt_c [D]_c (x) [beta0 M + beta1 M1]_pc (x) [M]_others
//   own face (outflow side) folded into [D]_c: +-Kt_c E at the outflow node
//   inflow face: -+Kt_c e_inflow-node (x) mixed [beta0 Mx + beta1 Mx1]_pc
//                (x) [Mx]_others applied to the neighbour's face-node slab
//   mass   : Km M (x) ... (x) M
// with Kt_c = adv_coef detJ / h_c, D = Gb^T W B, M = B^T W B, M1 = B^T W xi B,
// Mx = B_e^T W B_nb on the shared rule, Km = mass_coef detJ sigma.
//
// Cell schedule: the axes are paired (0,3), (1,4), (2,5); with MM_p = M (x) M
// and G_p the sum of the two pair-local operators of components p_0, p_1:
//   A (pair 2,5):  P = MM P25 U,  Q = G_25 U
//   B (pair 1,4):  R = MM_14 P,   W = G_14 P + MM_14 Q
//   C (pair 0,3):  Y = G_03 R + MM_03 (W + Km R)
// A thread holds one N x N slice of the current pair for fixed other indices
// (N^4 slices) and applies the pair operators in registers; the phases
// exchange through LDS.  The inflow-face slabs are converted by the wave that
// owns the face with in-place pencil sweeps (fixed stride 4 per tangential
// axis) and added at the face's node layer.
// Schedule: n = 4 elements take a 256-thread block each; n = 3 and n = 2
// elements one wave each (four per block), so those need wave barriers only.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_consttab.hpp"
#include "gd_padapt.hpp"

#include <utility>

namespace gd {
namespace pa6 {

using pa::PParams;

constexpr int ipw(int b, int e) { return e == 0 ? 1 : b * ipw(b, e - 1); }

// LDS position of element entry idx.  For N = 4 the low six bits (the bank)
// are XOR-ed with (i4, i3, i3) in their three 2-bit fields: the 64 lanes of a
// wave, which in the three pair phases run over (i0,i1,i3), (i0,i2,i3) and
// (i1,i2,i4) with the rest fixed, then hit 64 different banks in every phase
// (with the plain layout phase (0,3) is 4-way conflicted).  A bijection, and
// linear over XOR: swz(a ^ b) = swz(a) ^ swz(b), so an index assembled from
// disjoint bit fields swizzles field by field (compile-time fields fold).
template <int N>
__device__ __forceinline__ int swz(int idx)
{
  if constexpr (N == 4) {
    const int i3 = (idx >> 6) & 3, i4 = (idx >> 8) & 3;
    return idx ^ (i4 | (i3 << 2) | (i3 << 4));
  } else {
    return idx;
  }
}

// Face scratch: the five tangential axes with strides 1, 4, 16, 64, 256
// (1024 values); with PAD 1, 5, 20, 80, 320 (1280 values: the first axis
// padded to 5 so a sweep along it, whose lanes step the second axis, has no
// bank conflicts; measured slower overall, as the larger LDS footprint costs
// a resident block per CU).
template <bool PAD>
constexpr int scr_words() { return PAD ? 1280 : 1024; }

// global loads of a face slab in flight per lane (8 measured fastest; all 16
// of an n = 4 slab at once spill registers)
constexpr int SLAB_CHUNK = 8;
template <bool PAD>
constexpr int sstride(int k) { return k == 0 ? 1 : (PAD ? 5 : 4) * ipw(4, k - 1); }

// 1D matrices between an NO-node and an NI-node GLL basis on the Gauss rule of
// max(NO, NI) points: mass Mx, xi-weighted mass Mx1; for NO == NI also the
// derivative matrix D[i][j] = sum_q w_q l_i'(x_q) l_j(x_q).
template <int NO, int NI>
struct Mix {
  static constexpr int Q = NO > NI ? NO : NI;
  struct Mats {
    double M[NO][NI];
    double M1[NO][NI];
    double D[NO][NI];
  };
  static constexpr Mats make()
  {
    const ct::Tables<NO, Q> to = ct::make_gll_gauss<NO, Q>();
    const ct::Tables<NI, Q> ti = ct::make_gll_gauss<NI, Q>();
    Mats m{};
    for (int i = 0; i < NO; ++i)
      for (int j = 0; j < NI; ++j) {
        double s = 0.0, s1 = 0.0, sd = 0.0;
        for (int q = 0; q < Q; ++q) {
          s += to.w[q] * to.B[q][i] * ti.B[q][j];
          s1 += to.w[q] * to.xq[q] * to.B[q][i] * ti.B[q][j];
          sd += to.w[q] * to.Gb[q][i] * ti.B[q][j];
        }
        m.M[i][j] = s;
        m.M1[i][j] = s1;
        m.D[i][j] = sd;
      }
    return m;
  }
  static constexpr Mats m = make();
};

template <int I, int NN, typename F>
__device__ __forceinline__ void kfor6(F&& f)
{
  if constexpr (I < NN) {
    f(std::integral_constant<int, I>{});
    kfor6<I + 1, NN>(f);
  }
}

template <int G>
__device__ __forceinline__ void gsync()
{
  if constexpr (G == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Per-element data of one component
template <typename T>
struct Comp {
  T kt;     // adv_coef detJ / h_c
  T b0, b1; // beta_c = b0 + b1 xi along the partner axis
  bool own; // the outflow face has an upwind term (a neighbour or inflow data)
  bool neg; // beta_c <= 0: outflow on the - side, inflow from the + side
};

// S <- A S (rows: the pair's first axis), A given entry-wise
template <typename T, int N, typename F>
__device__ __forceinline__ void lmul(T (&S)[N][N], F A)
{
  T R[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += A(i, k) * S[k][j];
      R[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) S[i][j] = R[i][j];
}

// S <- S B^T (columns: the pair's second axis)
template <typename T, int N, typename F>
__device__ __forceinline__ void rmul(T (&S)[N][N], F B)
{
  T R[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s += S[i][k] * B(j, k);
      R[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) S[i][j] = R[i][j];
}

// Y += MM S = M S M^T
template <typename T, int N>
__device__ __forceinline__ void add_mm(T (&Y)[N][N], const T (&S)[N][N], T scale)
{
  using MT = Mix<N, N>;
  T R[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) R[i][j] = S[i][j];
  lmul<T, N>(R, [](int i, int k) { return (T)MT::m.M[i][k]; });
  rmul<T, N>(R, [](int j, int k) { return (T)MT::m.M[j][k]; });
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) Y[i][j] += scale * R[i][j];
}

// Y += G_p S for the pair (a, a+3): component a (derivative on rows, beta on
// columns) and component a+3 (beta on rows, derivative on columns)
template <typename T, int N>
__device__ __forceinline__ void add_g(T (&Y)[N][N], const T (&S)[N][N], const Comp<T>& ca, bool act_a,
                                      const Comp<T>& cb, bool act_b)
{
  using MT = Mix<N, N>;
  if (act_a) {
    T R[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) R[i][j] = S[i][j];
    rmul<T, N>(R, [&](int j, int k) { return ca.b0 * (T)MT::m.M[j][k] + ca.b1 * (T)MT::m.M1[j][k]; });
    // D' = -Kt D + own-face term at the outflow node
    T L[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) L[i][j] = R[i][j];
    lmul<T, N>(L, [](int i, int k) { return (T)MT::m.D[i][k]; });
    const int io = ca.neg ? 0 : N - 1;
    const T fo = ca.own ? (ca.neg ? -ca.kt : ca.kt) : T(0);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Y[i][j] += -ca.kt * L[i][j] + (i == io ? fo * R[i][j] : T(0));
  }
  if (act_b) {
    T R[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) R[i][j] = S[i][j];
    lmul<T, N>(R, [&](int i, int k) { return cb.b0 * (T)MT::m.M[i][k] + cb.b1 * (T)MT::m.M1[i][k]; });
    T L[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) L[i][j] = R[i][j];
    rmul<T, N>(L, [](int j, int k) { return (T)MT::m.D[j][k]; });
    const int jo = cb.neg ? 0 : N - 1;
    const T fo = cb.own ? (cb.neg ? -cb.kt : cb.kt) : T(0);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Y[i][j] += -cb.kt * L[i][j] + (j == jo ? fo * R[i][j] : T(0));
  }
}

// offset of slice t of pair (A, A+3): the other four axes, ascending, from t
template <int N, int A>
__device__ __forceinline__ int slice_base(int t)
{
  int base = 0, r = t;
#pragma unroll
  for (int d = 0; d < 6; ++d)
    if (d != A && d != A + 3) {
      base += (r % N) * ipw(N, d);
      r /= N;
    }
  return base;
}

// LDS position of entry (i, j) of the slice whose swizzled base is sb.  For
// N = 4 the slice's axes are zero bit fields of the base, so b + i s_A + j s_A3
// = b ^ (i s_A) ^ (j s_A3) and the swizzle of the constant part folds: one XOR
// per entry.  Other N are not swizzled and add.
template <int N, int A>
__device__ __forceinline__ int slice_pos(int sb, int i, int j)
{
  if constexpr (N == 4) return sb ^ swz<N>(i * ipw(N, A) + j * ipw(N, A + 3));
  else return sb + i * ipw(N, A) + j * ipw(N, A + 3);
}

template <typename T, int N, int SPT, int A>
__device__ __forceinline__ void read_slices(const T* X, int tid, int G, T (&S)[SPT][N][N])
{
  constexpr int N4 = ipw(N, 4);
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
    const int t = min(tid + s * G, N4 - 1);
    const int sb = swz<N>(slice_base<N, A>(t));
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) S[s][i][j] = X[slice_pos<N, A>(sb, i, j)];
  }
}

template <typename T, int N, int SPT, int A>
__device__ __forceinline__ void write_slices(T* X, int tid, int G, const T (&S)[SPT][N][N])
{
  constexpr int N4 = ipw(N, 4);
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
    const int t = tid + s * G;
    if (t < N4) {
      const int sb = swz<N>(slice_base<N, A>(t));
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) X[slice_pos<N, A>(sb, i, j)] = S[s][i][j];
    }
  }
}

// The five tangential axes of a face normal to C, ascending.
template <int C>
struct FaceAxes {
  static constexpr int ax(int k) { return k < C ? k : k + 1; }
};

// One inflow face (normal to C) of an N-node element from an NB-node
// neighbour (or inflow data g): convert the slab in the wave's scratch; one
// wave (64 lanes) takes part.
template <typename T, int N, int NB, int C>
__device__ __forceinline__ void face_convert(const T* __restrict__ xnb, bool data, T g, const Comp<T>& cc, T* scr,
                                             int lane)
{
  constexpr int WL = 64;
  using MX = Mix<N, NB>;
  using FA = FaceAxes<C>;
  constexpr int PC = (C + 3) % 6;
  // the neighbour's face-node slab (or the constant inflow data)
  {
    // the slab's loads issued SLAB_CHUNK at a time per lane, then stored to
    // the scratch (one load per loop trip left each load's latency exposed)
    constexpr int NS = ipw(NB, 5);
    constexpr int LPS = (NS + WL - 1) / WL;
    constexpr int CH = LPS < SLAB_CHUNK ? LPS : SLAB_CHUNK;  // loads in flight per lane
    const int ns = cc.neg ? 0 : NB - 1;  // neighbour node at the shared face
    const T* __restrict__ xs = xnb + (data ? ns * ipw(NB, C) : 0);
#pragma unroll 1
    for (int c0 = 0; c0 < LPS; c0 += CH) {
      T val[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        int r = min(lane + (c0 + u) * WL, NS - 1), src = 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          src += (r % NB) * ipw(NB, FA::ax(k));
          r /= NB;
        }
        val[u] = data ? xs[src] : g;
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int v = lane + (c0 + u) * WL;
        if (c0 + u < LPS && (NS % WL == 0 || v < NS)) {
          int r = v, dst = 0;
#pragma unroll
          for (int k = 0; k < 5; ++k) {
            dst += (r % NB) * sstride<false>(k);
            r /= NB;
          }
          scr[dst] = val[u];
        }
      }
    }
  }
  gsync<64>();
  // five in-place sweeps: slot k goes from NB to N values; the face
  // coefficient enters with the last one
  const T coef = cc.neg ? cc.kt : -cc.kt;
  kfor6<0, 5>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    constexpr bool BETA = FA::ax(k) == PC;
    constexpr int P = ipw(N, k) * ipw(NB, 4 - k);
    const T fin = k == 4 ? coef : T(1);
    T mat[N][NB];
#pragma unroll
    for (int ii = 0; ii < N; ++ii)
#pragma unroll
      for (int jj = 0; jj < NB; ++jj)
        mat[ii][jj] = fin * (BETA ? cc.b0 * (T)MX::m.M[ii][jj] + cc.b1 * (T)MX::m.M1[ii][jj] : (T)MX::m.M[ii][jj]);
#pragma unroll 1
    for (int q = lane; q < P; q += WL) {
      int r = q, base = 0;
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        if (j == k) continue;
        const int ext = j < k ? N : NB;
        base += (r % ext) * sstride<false>(j);
        r /= ext;
      }
      T in[NB];
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) in[jj] = scr[base + jj * sstride<false>(k)];
#pragma unroll
      for (int ii = 0; ii < N; ++ii) {
        T sum = T(0);
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) sum += mat[ii][jj] * in[jj];
        scr[base + ii * sstride<false>(k)] = sum;
      }
    }
    gsync<64>();
  });
}

// Add a converted face (scratch) at the element's inflow node layer.
template <typename T, int N, int C>
__device__ __forceinline__ void face_add(const Comp<T>& cc, const T* scr, T* Y, int lane)
{
  constexpr int WL = 64;
  constexpr int NS = ipw(N, 5);
  using FA = FaceAxes<C>;
  constexpr int LPS = (NS + WL - 1) / WL;
  constexpr int CH = LPS < 4 ? LPS : 4;  // entries in flight per lane
  const int ni = cc.neg ? N - 1 : 0;
#pragma unroll 1
  for (int c0 = 0; c0 < LPS; c0 += CH) {
    int dst[CH];
    T sv[CH], yv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      int r = min(lane + (c0 + u) * WL, NS - 1), d = 0, src = 0;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int dg = r % N;
        r /= N;
        d += dg * ipw(N, FA::ax(k));
        src += dg * sstride<false>(k);
      }
      dst[u] = swz<N>(d + ni * ipw(N, C));
      sv[u] = scr[src];
      yv[u] = Y[dst[u]];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u)
      if (c0 + u < LPS && (NS % WL == 0 || lane + (c0 + u) * WL < NS)) Y[dst[u]] = yv[u] + sv[u];
  }
}

// Convert inflow face C into scr; false when the face has no term (boundary
// without inflow data).
template <typename T, int N, int C>
__device__ __forceinline__ bool face_dispatch_c(const PParams<T>& p, const T* __restrict__ x, int nbn,
                                                long long off, bool data, const Comp<T>& cc, T* scr, int lane)
{
  if (nbn == 0) return false;
  if (!data) {
    face_convert<T, N, N, C>(nullptr, false, p.bc_value, cc, scr, lane);
    return true;
  }
  const T* xn = x + off;
  if (nbn == 2) face_convert<T, N, 2, C>(xn, true, T(0), cc, scr, lane);
  else if (nbn == 3) face_convert<T, N, 3, C>(xn, true, T(0), cc, scr, lane);
  else face_convert<T, N, 4, C>(xn, true, T(0), cc, scr, lane);
  return true;
}

template <typename T, int N>
__device__ __forceinline__ bool face_dispatch(const PParams<T>& p, const T* __restrict__ x, int c, int nbn,
                                              long long off, bool data, const Comp<T>& cc, T* scr, int lane)
{
  switch (c) {
    case 0: return face_dispatch_c<T, N, 0>(p, x, nbn, off, data, cc, scr, lane);
    case 1: return face_dispatch_c<T, N, 1>(p, x, nbn, off, data, cc, scr, lane);
    case 2: return face_dispatch_c<T, N, 2>(p, x, nbn, off, data, cc, scr, lane);
    case 3: return face_dispatch_c<T, N, 3>(p, x, nbn, off, data, cc, scr, lane);
    case 4: return face_dispatch_c<T, N, 4>(p, x, nbn, off, data, cc, scr, lane);
    default: return face_dispatch_c<T, N, 5>(p, x, nbn, off, data, cc, scr, lane);
  }
}

template <typename T, int N>
__device__ __forceinline__ void face_add_any(int c, const Comp<T>& cc, const T* scr, T* Y, int lane)
{
  switch (c) {
    case 0: face_add<T, N, 0>(cc, scr, Y, lane); break;
    case 1: face_add<T, N, 1>(cc, scr, Y, lane); break;
    case 2: face_add<T, N, 2>(cc, scr, Y, lane); break;
    case 3: face_add<T, N, 3>(cc, scr, Y, lane); break;
    case 4: face_add<T, N, 4>(cc, scr, Y, lane); break;
    default: face_add<T, N, 5>(cc, scr, Y, lane); break;
  }
}

// LDS words per element slot: two element buffers, plus a face scratch for
// the wave-per-element classes (n = 4 uses the second buffer and what follows
// it as four scratches)
// Per-element component record (LDS, at the end of a slot): coefficients,
// flags, and the inflow neighbour (order, DoF offset).
template <typename T>
struct alignas(8) CompRec {
  T kt, b0, b1;
  int flags;      // bit 0 own (outflow face term), bit 1 neg, bit 3 inflow data instead of a neighbour
  int nbn;        // neighbour nodes per dimension (0: no face term)
  long long off;  // neighbour DoF offset
};
constexpr int kCompWords = 6 * 48 / 4;  // 4-byte words, enough for six records of either precision

template <int N>
constexpr int slot_words()
{
  // the second buffer is free once the last phase has read it, and takes the
  // face scratch (four of them for the block-per-element class)
  const int scr = N == 4 ? 4 * scr_words<false>() : scr_words<false>();
  return ipw(N, 6) + (ipw(N, 6) > scr ? ipw(N, 6) : scr) + kCompWords;
}
constexpr int block_words()
{
  const int a = slot_words<4>(), b = 4 * slot_words<3>(), c = 4 * slot_words<2>();
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

template <typename T, int N, int G, bool ADV>
__device__ __forceinline__ void element(const PParams<T>& p, T mass_coef, T adv_coef, int64_t e,
                                        const T* __restrict__ x, T* __restrict__ y, T* slot, int tid)
{
  constexpr int N4 = ipw(N, 4), N6 = ipw(N, 6);
  constexpr int SPT = (N4 + G - 1) / G;
  T* X0 = slot;
  T* X1 = slot + N6;
  const int64_t off = p.dof_offset[e];
  int sub[6];
  pa::decode<T, 6>(p, e, sub);
  constexpr int LPT = (N6 + G - 1) / G;  // entries per thread
  {
    T v[LPT];
#pragma unroll
    for (int k = 0; k < LPT; ++k) v[k] = x[off + min(tid + k * G, N6 - 1)];
#pragma unroll
    for (int k = 0; k < LPT; ++k)
      if (N6 % G == 0 || tid + k * G < N6) X0[swz<N>(tid + k * G)] = v[k];
  }

  // per-element component data
  Comp<T> cp[6];
  int act = 0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    cp[c] = Comp<T>{T(0), T(0), T(0), false, false};
    if (!ADV || ((p.zero_axis_mask >> c) & 1)) continue;
    act |= 1 << c;
    const int pc = (c + 3) % 6;
    const T a = p.affA[c][pc];
    cp[c].kt = adv_coef * p.detJ * p.invh[c];
    cp[c].b0 = a * (p.origin[pc] + p.h[pc] * (T)sub[pc]) + p.affc[c];
    cp[c].b1 = a * p.h[pc];
    cp[c].neg = (p.neg_mask >> c) & 1;
    const int64_t dn = pa::neighbour<T, 6>(p, e, sub, cp[c].neg ? c : c + 6);  // outflow side
    cp[c].own = dn >= 0 || p.bc_mode == GD_BC_INFLOW;
  }
  const T km = mass_coef * p.detJ * p.sigma;
  // component records for the face pass: lane c computes component c from
  // the parameters (global memory) alone, and the neighbour's order and
  // offset loads complete during the cell phases
  CompRec<T>* ctab = reinterpret_cast<CompRec<T>*>(slot + slot_words<N>() - kCompWords * 4 / (int)sizeof(T));
  if (ADV && tid < 6) {
    const int c = tid, pc = (c + 3) % 6;
    CompRec<T> r{T(0), T(0), T(0), 0, 0, 0};
    if (!((p.zero_axis_mask >> c) & 1)) {
      int64_t st = 1;
      for (int d = 0; d < 6; ++d)
        if (d < pc) st *= p.M[d];
      const int spc = (int)((e / st) % p.M[pc]);
      const T a = p.affA[c][pc];
      r.kt = adv_coef * p.detJ * p.invh[c];
      r.b0 = a * (p.origin[pc] + p.h[pc] * (T)spc) + p.affc[c];
      r.b1 = a * p.h[pc];
      const bool neg = (p.neg_mask >> c) & 1;
      const int64_t dn = pa::neighbour<T, 6>(p, e, sub, neg ? c : c + 6);  // outflow side
      const int64_t un = pa::neighbour<T, 6>(p, e, sub, neg ? c + 6 : c);  // inflow side
      const bool own = dn >= 0 || p.bc_mode == GD_BC_INFLOW;
      r.flags = (own ? 1 : 0) | (neg ? 2 : 0);
      if (un >= 0) {
        r.nbn = p.order[un] + 1;
        r.off = p.dof_offset[un];
      } else if (p.bc_mode == GD_BC_INFLOW) {
        r.nbn = N;
        r.flags |= 8;
      }
    }
    ctab[c] = r;
  }
  gsync<G>();

  // ---- A: pair (2,5) ----
  T S[SPT][N][N], U1[SPT][N][N];
  read_slices<T, N, SPT, 2>(X0, tid, G, S);
  gsync<G>();
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) U1[s][i][j] = T(0);
    add_g<T, N>(U1[s], S[s], cp[2], (act >> 2) & 1, cp[5], (act >> 5) & 1);
    T P[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) P[i][j] = T(0);
    add_mm<T, N>(P, S[s], T(1));
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) S[s][i][j] = P[i][j];
  }
  write_slices<T, N, SPT, 2>(X0, tid, G, S);   // P
  write_slices<T, N, SPT, 2>(X1, tid, G, U1);  // Q
  gsync<G>();

  // ---- B: pair (1,4) ----
  read_slices<T, N, SPT, 1>(X0, tid, G, S);   // P
  read_slices<T, N, SPT, 1>(X1, tid, G, U1);  // Q
  gsync<G>();
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
    T R[N][N], Wv[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) R[i][j] = Wv[i][j] = T(0);
    add_mm<T, N>(R, S[s], T(1));
    add_g<T, N>(Wv, S[s], cp[1], (act >> 1) & 1, cp[4], (act >> 4) & 1);
    add_mm<T, N>(Wv, U1[s], T(1));
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        S[s][i][j] = R[i][j];
        U1[s][i][j] = Wv[i][j];
      }
  }
  write_slices<T, N, SPT, 1>(X0, tid, G, S);   // R
  write_slices<T, N, SPT, 1>(X1, tid, G, U1);  // W
  gsync<G>();

  // ---- C: pair (0,3) ----
  read_slices<T, N, SPT, 0>(X0, tid, G, S);   // R
  read_slices<T, N, SPT, 0>(X1, tid, G, U1);  // W
  gsync<G>();
#pragma unroll
  for (int s = 0; s < SPT; ++s) {
    T Yv[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        Yv[i][j] = T(0);
        U1[s][i][j] += km * S[s][i][j];
      }
    add_g<T, N>(Yv, S[s], cp[0], act & 1, cp[3], (act >> 3) & 1);
    add_mm<T, N>(Yv, U1[s], T(1));
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) S[s][i][j] = Yv[i][j];
  }
  write_slices<T, N, SPT, 0>(X0, tid, G, S);  // Y without the inflow faces
  gsync<G>();

  // ---- inflow faces ----
  // (faces index the components at run time: they read them from the LDS
  // records, so the per-component registers are never indexed dynamically,
  // which would put them on the stack)
  auto pick = [&](int c) -> Comp<T> {
    const CompRec<T> r = ctab[c];
    return Comp<T>{r.kt, r.b0, r.b1, (r.flags & 1) != 0, (r.flags & 2) != 0};
  };
  if constexpr (ADV) {
    if constexpr (G == 64) {
      T* scr = X1;
#pragma unroll 1
      for (int c = 0; c < 6; ++c) {
        if (!((act >> c) & 1)) continue;
        const Comp<T> cc = pick(c);
        const int nbn = ctab[c].nbn, fl = ctab[c].flags;
        const long long noff = ctab[c].off;
        if (face_dispatch<T, N>(p, x, c, nbn, noff, (fl & 8) == 0, cc, scr, tid)) {
          face_add_any<T, N>(c, cc, scr, X0, tid);
          gsync<64>();
        }
      }
    } else {
      // four waves convert one face each per round (the second buffer is
      // free now: four scratches), then add them in face order
      const int w = tid / 64, lane = tid % 64;
      T* scr = X1 + w * scr_words<false>();
      const int nf = __builtin_popcount(act);
#pragma unroll 1
      for (int r0 = 0; r0 < nf; r0 += 4) {
        // this wave's face: the (r0 + w)-th active component
        int c = -1;
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (((act >> k) & 1) && __builtin_popcount(act & ((1 << k) - 1)) == r0 + w) c = k;
        const int cs = c < 0 ? 0 : c;
        const Comp<T> cc = pick(cs);
        const int nbn = ctab[cs].nbn, fl = ctab[cs].flags;
        const long long noff = ctab[cs].off;
        const bool mine = c >= 0 && face_dispatch<T, N>(p, x, c, nbn, noff, (fl & 8) == 0, cc, scr, lane);
        __syncthreads();
#pragma unroll 1
        for (int turn = 0; turn < 4; ++turn) {
          if (turn == w && mine) face_add_any<T, N>(c, cc, scr, X0, lane);
          __syncthreads();
        }
      }
    }
  }
  {
    T v[LPT];
#pragma unroll
    for (int k = 0; k < LPT; ++k) v[k] = X0[swz<N>(min(tid + k * G, N6 - 1))];
#pragma unroll
    for (int k = 0; k < LPT; ++k)
      if (N6 % G == 0 || tid + k * G < N6) y[off + tid + k * G] = v[k];
  }
}

// One launch: it0 blocks of one n = 4 element each, then it1 blocks of four
// n = 3 elements (a wave each), then blocks of four n = 2 elements.
template <typename T, bool ADV>
__global__ void __launch_bounds__(256, 4) pa6k_kernel(const PParams<T>* __restrict__ pp, int it0, int it1, T mass_coef,
                                                   T adv_coef, const T* __restrict__ x, T* __restrict__ y)
{
  const PParams<T>& p = *pp;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int it = blockIdx.x;
  if (it < it0) {
    element<T, 4, 256, ADV>(p, mass_coef, adv_coef, p.list[0][it], x, y, lds, threadIdx.x);
  } else if ((it -= it0) < it1) {
    const int w = threadIdx.x / 64, idx = it * 4 + w;
    if (idx < p.count[1])
      element<T, 3, 64, ADV>(p, mass_coef, adv_coef, p.list[1][idx], x, y, lds + w * slot_words<3>(), threadIdx.x % 64);
  } else {
    it -= it1;
    const int w = threadIdx.x / 64, idx = it * 4 + w;
    if (idx < p.count[2])
      element<T, 2, 64, ADV>(p, mass_coef, adv_coef, p.list[2][idx], x, y, lds + w * slot_words<2>(), threadIdx.x % 64);
  }
}

}  // namespace pa6
}  // namespace gd
