// p-adaptive Cartesian DG apply (GLL basis, Gauss rule with q_e = p_e + 1
// points per dimension, n_e = p_e + 1 in {2, 3, 4}) in one launch, one block
// per work item.
//
// What is computed per element e is the element-centric operator of the
// uniform kernels with per-element orders:
//   y_e = mass_coef * M_e u_e + adv_coef * (cell + faces)
// cell  = sum_d B^T Gq_d^T (c_d . B u_e), c_d = -w detJ beta_d / h_d
//         (advectionoperator.hpp:137-167; applygradienttestfunctionsthreaded.hpp:164-254)
// faces = for every face with a neighbour (or inflow data): the upwind flux
//         F = w detJ (dot > 0 ? dot u^- : dot u^+), dot = beta_a sign / h_a,
//         evaluated on the Gauss rule of qf = max(q_e, q_nb) points per
//         tangential dimension, both traces interpolated to those points from
//         their own GLL face slabs, and tested with the own basis
//         (advectionoperator.hpp:175-227, 391-438).  The face rule is shared
//         by both sides, as for GenDiL's non-conforming face operator
//         (advectionoperator.hpp:596-676, 1210-1229), so the flux is
//         conservative: sum_i (A u)_i = 0 on a periodic mesh.
//
// Schedule: elements are bucketed by order on the host, most expensive
// first; a work item is one n=4 element (a 256-thread block, or one wave per
// element), or one element per wave of the lower orders (4 per block), or one
// element per block; the thread group per order is a template choice picked
// by the autotuner.  Every order runs in the same launch.
//
// Every 1D contraction is an in-place pencil sweep on LDS: a thread reads a
// whole pencil into registers, then writes the contracted pencil back to the
// same positions (tensors are stored with a fixed per-dimension stride, so
// pencils never overlap).  Tables are compile-time constants.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "gd_consttab.hpp"
#include "gd_internal.hpp"

namespace gd {
namespace pa {

constexpr int kNT = 256;
constexpr int kNMax = 4;

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
constexpr int imax(int a, int b) { return a > b ? a : b; }

template <typename T>
struct Coefs {
  T mass_coef, adv_coef;
};

template <typename T>
struct PParams {
  int64_t M[kMaxDim];
  int64_t E;
  int periodic_mask;
  T h[kMaxDim], invh[kMaxDim], origin[kMaxDim];
  T detJ;
  // beta: affine (A X + c) or constant per component
  int beta_affine;
  T affA[kMaxDim][kMaxDim];
  T affc[kMaxDim];
  int zero_axis_mask;  // bit a: beta_a == 0 identically (face flux vanishes)
  int neg_mask;        // bit a: beta_a <= 0 on the whole mesh (Kronecker 6D kernel)
  int kron_ok;         // the field suits the Kronecker 6D kernel (gd_pa6k.hpp)
  T sigma;             // constant
  int bc_mode;
  T bc_value;          // constant inflow value
  const int8_t* order;
  const int64_t* dof_offset;
  // element lists per order class: [0] n=4, [1] n=3, [2] n=2
  const int32_t* list[3];
  int count[3];
};

// compile-time tables for n GLL nodes and q Gauss points
template <int N, int Q>
struct Tab {
  static constexpr ct::Tables<N, Q> t = ct::make_gll_gauss<N, Q>();
};
template <int Q>
struct TabG {
  static constexpr ct::Colloc<Q> c = ct::make_gauss_colloc<Q>();
};

// t[k] for a runtime k < Q by selects (a dynamic index into a local copy of a
// constant table would go through scratch memory)
template <typename T, int Q>
__device__ __forceinline__ T pick(const double (&t)[Q], int k)
{
  T v = (T)t[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) v = k == i ? (T)t[i] : v;
  return v;
}

// arr[i] for a runtime (wave-uniform) i < D by selects, so the kernel
// arguments are never indexed dynamically (that would copy them to scratch)
template <int D, typename V>
__device__ __forceinline__ V sel(const V (&arr)[kMaxDim], int i)
{
  V v = arr[0];
#pragma unroll
  for (int d = 1; d < D; ++d) v = i == d ? arr[d] : v;
  return v;
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for_pa(F&& f)
{
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_pa<I + 1, N>(f);
  }
}

template <int I, typename F>
__device__ __forceinline__ void static_for_down_pa(F&& f)
{
  if constexpr (I >= 0) {
    f(std::integral_constant<int, I>{});
    static_for_down_pa<I - 1>(f);
  }
}

// synchronisation of the threads that share one element
template <int TPE>
__device__ __forceinline__ void group_sync()
{
  if constexpr (TPE == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Offset of a pencil in a tensor of DIMS dims stored with stride S^d: pencil
// p enumerates dims < J with extent ELO and dims > J with extent EHI.
template <int DIMS, int S, int J, int ELO, int EHI>
__device__ __forceinline__ int pencil_offset(int p)
{
  constexpr int LO = ipow(ELO, J);
  const int lo = p % LO;
  int hi = p / LO;
  int off = 0;
  if constexpr (ELO == S) {
    off = lo;
  } else {
    int r = lo;
#pragma unroll
    for (int d = 0; d < J; ++d) {
      off += (r % ELO) * ipow(S, d);
      r /= ELO;
    }
  }
  if constexpr (EHI == S) {
    off += hi * ipow(S, J + 1);
  } else {
#pragma unroll
    for (int d = J + 1; d < DIMS; ++d) {
      off += (hi % EHI) * ipow(S, d);
      hi /= EHI;
    }
  }
  return off;
}

// In-place contraction along dim J: NIN values in, NOUT out per pencil,
// out[o] = sum_i mat(o, i) in[i].
template <int TPE, int DIMS, int S, int J, int ELO, int EHI, int NIN, int NOUT, typename T, typename Mat>
__device__ __forceinline__ void sweep(T* buf, int tid, Mat mat)
{
  constexpr int P = ipow(ELO, J) * ipow(EHI, DIMS - 1 - J);
  constexpr int ST = ipow(S, J);
#pragma unroll 1
  for (int p = tid; p < P; p += TPE) {
    T* b = buf + pencil_offset<DIMS, S, J, ELO, EHI>(p);
    T v[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) v[i] = b[i * ST];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
      T s = T(0);
#pragma unroll
      for (int i = 0; i < NIN; ++i) s += mat(o, i) * v[i];
      b[o * ST] = s;
    }
  }
}

template <typename T, int D>
__device__ __forceinline__ void decode(const PParams<T>& p, int64_t e, int* sub)
{
#pragma unroll
  for (int d = 0; d < D; ++d) {
    sub[d] = (int)(e % p.M[d]);
    e /= p.M[d];
  }
}

// neighbour across face f (axis f % D, + side if f >= D), -1 on a boundary
template <typename T, int D>
__device__ __forceinline__ int64_t neighbour(const PParams<T>& p, int64_t e, const int* sub, int f)
{
  const int a = f % D;
  const bool plus = f >= D;
  const bool per = (p.periodic_mask >> a) & 1;
  int64_t stride = 1;
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (d < a) stride *= p.M[d];
  const int64_t Ma = sel<D>(p.M, a);
  int sa = 0;
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (d == a) sa = sub[d];
  if (Ma == 1) return per ? e : -1;
  if (!plus) return sa > 0 ? e - stride : (per ? e + (Ma - 1) * stride : -1);
  return sa < Ma - 1 ? e + stride : (per ? e - (Ma - 1) * stride : -1);
}

// beta component c at physical point X (c a compile-time or wave-uniform index)
template <typename T, int D>
__device__ __forceinline__ T beta_at(const PParams<T>& p, int c, const T* X)
{
  T v = sel<D>(p.affc, c);
  if (p.beta_affine) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      T a = p.affA[0][d];
#pragma unroll
      for (int cc = 1; cc < D; ++cc) a = c == cc ? p.affA[cc][d] : a;
      v += a * X[d];
    }
  }
  return v;
}

// Face term of one element for a neighbour with NB nodes (NB = N on the
// boundary).  F0 and F1 hold QF^(D-1) values each.
template <typename T, int D, int N, int NB, int TPE>
__device__ __forceinline__ void face(const PParams<T>& p, const Coefs<T> cf, const T* __restrict__ x, const T* U, T* Y, T* F0, T* F1,
                                     int tid, const int* sub, int f, int64_t nb)
{
  constexpr int QF = imax(N, NB);
  constexpr int DT = D - 1;  // tangential dims
  constexpr int QF5 = ipow(QF, DT);
  const int a = f % D;
  const bool plus = f >= D;
  const bool boundary = nb < 0;
  // ---- slabs (GLL: the trace is the node slab at the face) ----
  {
    constexpr int NS = ipow(N, DT);
    const int ia = plus ? N - 1 : 0;
#pragma unroll 1
    for (int t = tid; t < NS; t += TPE) {
      int r = t, full = 0, fo = 0, mul = 1, fmul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d == a) {
          full += ia * mul;
        } else {
          const int id = r % N;
          r /= N;
          full += id * mul;
          fo += id * fmul;
          fmul *= QF;
        }
        mul *= N;
      }
      F0[fo] = U[full];
    }
    if (!boundary) {
      constexpr int NSB = ipow(NB, DT);
      const int ib = plus ? 0 : NB - 1;
      const T* xn = x + p.dof_offset[nb];
#pragma unroll 1
      for (int t = tid; t < NSB; t += TPE) {
        int r = t, full = 0, fo = 0, mul = 1, fmul = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if (d == a) {
            full += ib * mul;
          } else {
            const int id = r % NB;
            r /= NB;
            full += id * mul;
            fo += id * fmul;
            fmul *= QF;
          }
          mul *= NB;
        }
        F1[fo] = xn[full];
      }
    }
  }
  group_sync<TPE>();
  // ---- tangential interpolation, last dim first ----
  static_for_down_pa<DT - 1>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    sweep<TPE, DT, QF, j, N, QF, N, QF, T>(F0, tid, [&](int o, int i) { return (T)Tab<N, QF>::t.B[o][i]; });
    if (!boundary) sweep<TPE, DT, QF, j, NB, QF, NB, QF, T>(F1, tid, [&](int o, int i) { return (T)Tab<NB, QF>::t.B[o][i]; });
    group_sync<TPE>();
  });
  // ---- upwind flux at the face points (into F0) ----
  {
    const T nrm = (plus ? T(1) : T(-1)) * sel<D>(p.invh, a);
#pragma unroll 1
    for (int t = tid; t < QF5; t += TPE) {
      T X[D];
      T w = T(1);
      int r = t;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d == a) {
          X[d] = (p.origin[d] + p.h[d] * (T)sub[d]) + p.h[d] * (plus ? T(1) : T(0));
        } else {
          const int k = r % QF;
          r /= QF;
          X[d] = (p.origin[d] + p.h[d] * (T)sub[d]) + p.h[d] * pick<T, QF>(Tab<1, QF>::t.xq, k);
          w *= pick<T, QF>(Tab<1, QF>::t.w, k);
        }
      }
      const T dot = beta_at<T, D>(p, a, X) * nrm;
      const T own = F0[t];
      const T ext = boundary ? p.bc_value : F1[t];
      const T up = dot > T(0) ? dot * own : dot * ext;
      F0[t] = cf.adv_coef * ((w * p.detJ) * up);
    }
  }
  group_sync<TPE>();
  // ---- own test functions, first dim first ----
  static_for_pa<0, DT>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    sweep<TPE, DT, QF, j, N, QF, QF, N, T>(F0, tid, [&](int o, int i) { return (T)Tab<N, QF>::t.B[i][o]; });
    group_sync<TPE>();
  });
  // ---- add into the own face slab ----
  {
    constexpr int NS = ipow(N, DT);
    const int ia = plus ? N - 1 : 0;
#pragma unroll 1
    for (int t = tid; t < NS; t += TPE) {
      int r = t, full = 0, fo = 0, mul = 1, fmul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d == a) {
          full += ia * mul;
        } else {
          const int id = r % N;
          r /= N;
          full += id * mul;
          fo += id * fmul;
          fmul *= QF;
        }
        mul *= N;
      }
      Y[full] += F0[fo];
    }
  }
  group_sync<TPE>();
}

// LDS words per element slot: U, Y, and a region that holds the cell
// accumulator or the two face buffers.
template <int D, int N>
constexpr int slot_words()
{
  return 2 * ipow(N, D) + imax(ipow(N, D), 2 * ipow(kNMax, D - 1));
}

template <typename T, int D, int N, bool ADV, int TPE>
__device__ __forceinline__ void element(const PParams<T>& p, const Coefs<T> cf, int64_t e, const T* __restrict__ x, T* __restrict__ y,
                                        T* slot, int tid)
{
  constexpr int ND = ipow(N, D);
  constexpr int Q = N;
  T* U = slot;
  T* Y = slot + ND;
  T* A = slot + 2 * ND;
  T* F0 = A;
  T* F1 = A + ipow(kNMax, D - 1);
  const int64_t dof0 = p.dof_offset[e];
#pragma unroll 1
  for (int i = tid; i < ND; i += TPE) {
    U[i] = x[dof0 + i];
    Y[i] = T(0);
  }
  int sub[D];
  decode<T, D>(p, e, sub);
  group_sync<TPE>();
  if constexpr (ADV) {
    for (int f = 0; f < 2 * D; ++f) {
      const int a = f % D;
      if ((p.zero_axis_mask >> a) & 1) continue;  // beta_a == 0: zero flux
      const int64_t nb = neighbour<T, D>(p, e, sub, f);
      if (nb < 0 && p.bc_mode == GD_BC_NONE) continue;
      const int nbn = nb < 0 ? N : p.order[nb] + 1;
      if (nbn == 2) face<T, D, N, 2, TPE>(p, cf, x, U, Y, F0, F1, tid, sub, f, nb);
      else if (nbn == 3) face<T, D, N, 3, TPE>(p, cf, x, U, Y, F0, F1, tid, sub, f, nb);
      else face<T, D, N, 4, TPE>(p, cf, x, U, Y, F0, F1, tid, sub, f, nb);
    }
  }
  // ---- cell: B u in place, last dim first ----
  static_for_down_pa<D - 1>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    sweep<TPE, D, N, d, N, N, N, Q, T>(U, tid, [&](int o, int i) { return (T)Tab<N, Q>::t.B[o][i]; });
    group_sync<TPE>();
  });
  // ---- quadrature-point terms: A = sum_d Gq_d^T (c_d . Bu) + mass ----
  static_for_pa<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    constexpr int P = ND / Q;
    constexpr int ST = ipow(Q, d);
#pragma unroll 1
    for (int pp = tid; pp < P; pp += TPE) {
      const int off = pencil_offset<D, Q, d, Q, Q>(pp);
      int k[D];
      {
        int r = off;
#pragma unroll
        for (int dd = 0; dd < D; ++dd) {
          k[dd] = r % Q;
          r /= Q;
        }
      }
      T X[D];
      T wo = T(1);  // weight of the other dims
#pragma unroll
      for (int dd = 0; dd < D; ++dd) {
        X[dd] = (p.origin[dd] + p.h[dd] * (T)sub[dd]) + p.h[dd] * pick<T, Q>(Tab<N, Q>::t.xq, k[dd]);
        if (dd != d) wo *= pick<T, Q>(Tab<N, Q>::t.w, k[dd]);
      }
      T v[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) v[q] = U[off + q * ST];
      T c[Q];
      if constexpr (ADV) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          X[d] = (p.origin[d] + p.h[d] * (T)sub[d]) + p.h[d] * (T)Tab<N, Q>::t.xq[q];
          const T w = wo * (T)Tab<N, Q>::t.w[q];
          const T cq = cf.adv_coef * (-w * p.detJ * (beta_at<T, D>(p, d, X) * p.invh[d]));
          c[q] = cq * v[q];
        }
      }
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        T s = T(0);
        if constexpr (ADV) {
#pragma unroll
          for (int q = 0; q < Q; ++q) s += (T)TabG<Q>::c.Gq[q][i] * c[q];
        }
        T* dst = A + off + i * ST;
        if constexpr (d == 0) {
          if (cf.mass_coef != T(0)) {
            const T w = wo * (T)Tab<N, Q>::t.w[i];
            const T mv = cf.mass_coef * (w * p.detJ * p.sigma * v[i]);
            s = ADV ? s + mv : mv;
          }
          *dst = s;
        } else {
          *dst += s;
        }
      }
    }
    group_sync<TPE>();
  });
  // ---- test functions, first dim first; the last sweep adds into Y ----
  static_for_pa<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    if constexpr (d < D - 1) {
      sweep<TPE, D, N, d, N, Q, Q, N, T>(A, tid, [&](int o, int i) { return (T)Tab<N, Q>::t.B[i][o]; });
    } else {
      constexpr int P = ND / N;
      constexpr int ST = ipow(N, d);
#pragma unroll 1
      for (int pp = tid; pp < P; pp += TPE) {
        const int off = pencil_offset<D, N, d, N, Q>(pp);
        T v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = A[off + q * ST];
#pragma unroll
        for (int o = 0; o < N; ++o) {
          T s = T(0);
#pragma unroll
          for (int q = 0; q < Q; ++q) s += (T)Tab<N, Q>::t.B[q][o] * v[q];
          Y[off + o * ST] += s;
        }
      }
    }
    group_sync<TPE>();
  });
#pragma unroll 1
  for (int i = tid; i < ND; i += TPE) y[dof0 + i] = Y[i];
}

template <int D, int TPE4, int TPE3, int TPE2>
constexpr int lds_words()
{
  return imax(imax((kNT / TPE4) * slot_words<D, 4>(), (kNT / TPE3) * slot_words<D, 3>()),
              (kNT / TPE2) * slot_words<D, 2>());
}

// One block per work item.  The hardware dispatcher hands out the blocks as
// slots free up, which balances the uneven per-item cost; a grid-stride loop
// over items would keep every loop-invariant table constant live across the
// loop and spill scalar registers.
// The parameters live in device memory (one copy per plan): read through a
// pointer they become scalar loads, whereas a large by-value kernel argument
// that is indexed through references gets copied to scratch.
template <typename T, int D, bool ADV, int TPE4, int TPE3, int TPE2>
__global__ void __launch_bounds__(kNT) padapt_kernel(const PParams<T>* __restrict__ pp, int it0, int it1,
                                                     T mass_coef, T adv_coef, const T* __restrict__ x,
                                                     T* __restrict__ y)
{
  const PParams<T>& p = *pp;
  const Coefs<T> cf{mass_coef, adv_coef};
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int it = blockIdx.x;
  if (it < it0) {
    constexpr int EPI = kNT / TPE4;
    const int g = threadIdx.x / TPE4, tid = threadIdx.x % TPE4;
    const int idx = it * EPI + g;
    if (idx < p.count[0]) element<T, D, 4, ADV, TPE4>(p, cf, p.list[0][idx], x, y, lds + g * slot_words<D, 4>(), tid);
  } else if ((it -= it0) < it1) {
    constexpr int EPI = kNT / TPE3;
    const int g = threadIdx.x / TPE3, tid = threadIdx.x % TPE3;
    const int idx = it * EPI + g;
    if (idx < p.count[1]) element<T, D, 3, ADV, TPE3>(p, cf, p.list[1][idx], x, y, lds + g * slot_words<D, 3>(), tid);
  } else {
    it -= it1;
    constexpr int EPI = kNT / TPE2;
    const int g = threadIdx.x / TPE2, tid = threadIdx.x % TPE2;
    const int idx = it * EPI + g;
    if (idx < p.count[2]) element<T, D, 2, ADV, TPE2>(p, cf, p.list[2][idx], x, y, lds + g * slot_words<D, 2>(), tid);
  }
}

}  // namespace pa
}  // namespace gd
