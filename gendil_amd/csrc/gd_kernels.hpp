// gfx950 kernels for the element-centric matrix-free DG apply on Cartesian
// meshes: fused cell + interior-facet + boundary-facet terms, one workgroup
// per batch of EPB consecutive elements.
//
// Element DoFs are staged in LDS once (coalesced, element-contiguous loads);
// every 1D contraction is a "pencil sweep": each thread owns one pencil along
// the contracted dimension, reads it from LDS once, multiplies by the 1D
// table (uniform addresses -> scalar loads) and writes the pencil back.  Face
// traces of the neighbour are read straight from HBM/L2 (only the face slab
// for GLL nodes) inside the same pass, so the face terms need no extra
// kernel and no atomics: every element writes only its own DoFs.
//
// Algorithm (what is computed) follows GenDiL's element-centric operators:
//   MassElementOperator              massoperator.hpp:39-89
//   AdvectionFusedOperatorWith[out]BC advectionoperator.hpp:106-441
// with the interpolation / test-function contractions of
// interpolatevaluesserial.hpp:25-103, applytestfunctionsserial.hpp:26-111 and
// the collocated gradient transpose of applygradienttestfunctionsthreaded.hpp:164-254.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "gd_internal.hpp"

namespace gd {
namespace dev {

template <int B, int E>
struct Pow {
  static constexpr int v = B * Pow<B, E - 1>::v;
};
template <int B>
struct Pow<B, 0> {
  static constexpr int v = 1;
};

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

template <int I, typename F>
__device__ __forceinline__ void static_for_down(F&& f)
{
  if constexpr (I >= 0) {
    f(std::integral_constant<int, I>{});
    static_for_down<I - 1>(f);
  }
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }

enum Mode { kForward = 0, kTranspose = 1, kGradBeta = 2 };

// Compile-time configuration of one kernel instance.
template <typename T_, int D_, int N_, int Q_, int MODE_>
struct Cfg {
  using T = T_;
  static constexpr int D = D_, N = N_, Q = Q_, MODE = MODE_;
  static constexpr int ND = Pow<N, D>::v;
  static constexpr int QD = Pow<Q, D>::v;
  static constexpr int MX = cmax(N, Q);
  static constexpr int MXD = Pow<MX, D>::v;
  static constexpr int NF = Pow<N, D - 1>::v;
  static constexpr int QF = Pow<Q, D - 1>::v;
  static constexpr int MF = Pow<MX, D - 1>::v;
  // traces per face: forward/transpose own+neighbour; grad: u and r on both sides
  static constexpr int TPF = MODE == kGradBeta ? 4 : 2;
  static constexpr int CELL_SCR = 3 * MXD;
  // faces per group: largest divisor of 2D whose trace buffers fit in the cell scratch
  static constexpr int pick_fg()
  {
    int best = 1;
    for (int g = 1; g <= 2 * D; ++g)
      if ((2 * D) % g == 0 && 2 * TPF * g * MF <= CELL_SCR) best = g;
    return best;
  }
  static constexpr int FG = pick_fg();
  static constexpr int FACE_SCR = 2 * TPF * FG * MF;
  static constexpr int SCR = cmax(CELL_SCR, FACE_SCR);
  // per-element LDS: U (input), Y (output or second input), scratch
  static constexpr int PER_ELEM = 2 * ND + SCR;
  // elements per block: as many as fit in 32 KB of LDS (at most 16).  Smaller
  // blocks leave more of them resident per CU to hide the sweeps' barriers and
  // the neighbour-trace loads; measured on MI355X against 16, 24, 48, 64 and
  // 96 KB (and 128 / 512 threads): 48^3 p=4 q=5 per-point advection 1.07 ->
  // 0.63 ms, its beta contraction 1.15 -> 0.84 ms, the 6D p=1 q=3 benchmark
  // operator 1.19 -> 0.89 ms, all relative to 64 KB.  6D blocks take 128
  // threads (the 6D operator 0.89 -> 0.72 ms; the 3D ones lose with 128).
  static constexpr int LDS_BUDGET = 32 * 1024;
  static constexpr int EPB_RAW = LDS_BUDGET / (int)(PER_ELEM * sizeof(T));
  static constexpr int EPB = EPB_RAW < 1 ? 1 : (EPB_RAW > 16 ? 16 : EPB_RAW);
  static constexpr int NT = D >= 6 ? 128 : 256;
};

// ---- pencil sweep --------------------------------------------------------
// Batch of NB tensors viewed as [OUTER][NIN][INNER] -> [OUTER][NOUT][INNER].
template <int NT, int NB, int OUTER, int NIN, int NOUT, int INNER, bool ACC, typename T, typename AIn,
          typename AOut, typename Mat>
__device__ __forceinline__ void sweep(AIn ain, AOut aout, Mat m)
{
  constexpr int PEN = OUTER * INNER;
  constexpr int TOTAL = NB * PEN;
  for (int t = threadIdx.x; t < TOTAL; t += NT) {
    const int b = t / PEN;
    const int r = t - b * PEN;
    const int o = r / INNER;
    const int a = r - o * INNER;
    const T* src = ain(b) + o * NIN * INNER + a;
    T v[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) v[i] = src[i * INNER];
    T* dst = aout(b) + o * NOUT * INNER + a;
#pragma unroll
    for (int j = 0; j < NOUT; ++j) {
      T s = T(0);
#pragma unroll
      for (int i = 0; i < NIN; ++i) s += m(j, i) * v[i];
      if constexpr (ACC) dst[j * INNER] += s;
      else dst[j * INNER] = s;
    }
  }
}

// ---- coefficient evaluation -----------------------------------------------
template <typename T, int D, int Q>
__device__ __forceinline__ T eval_sep(const FieldDev<T>& f, const int* sub, const int* k)
{
  T sum = T(0);
  for (int r = 0; r < f.rank; ++r) {
    T t = f.scale[r];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T* tb = f.tab[r][d];
      if (tb) t *= tb[sub[d] * (Q + 2) + k[d]];
    }
    sum += t;
  }
  return sum;
}

template <typename T, int D, int Q>
__device__ __forceinline__ T xi_of(const Params<T>& p, int k)
{
  return k < Q ? p.tab.xq[k] : (k == Q ? T(0) : T(1));
}

// beta component c at sample indices k (k[d] < Q rule point, Q: xi=0, Q+1: xi=1)
template <typename T, int D, int Q>
__device__ __forceinline__ T beta_at(const Params<T>& p, int c, const int* sub, const int* k)
{
  if (p.beta_kind == GD_BETA_AFFINE) {
    T v = p.affc[c];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T X = (p.origin[d] + p.h[d] * (T)sub[d]) + p.h[d] * xi_of<T, D, Q>(p, k[d]);
      v += p.affA[c][d] * X;
    }
    return v;
  }
  return eval_sep<T, D, Q>(p.beta[c], sub, k);
}

// ---- per-element info in LDS ----------------------------------------------
template <int D>
struct ElemInfo {
  long long e;            // global element, -1 = padding
  int sub[D];
  long long nb[2 * D];    // neighbour element per local face, -1 = boundary
  long long fidx[2 * D];  // global face index (pointwise beta layout)
};

template <typename T, int D>
__device__ __forceinline__ void make_info(const Params<T>& p, long long e, ElemInfo<D>& inf)
{
  inf.e = e;
  if (e < 0) return;
  long long r = e;
  long long stride[D];
  long long s = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    inf.sub[d] = (int)(r % p.M[d]);
    r /= p.M[d];
    stride[d] = s;
    s *= p.M[d];
  }
#pragma unroll
  for (int f = 0; f < 2 * D; ++f) {
    const int a = f % D;
    const bool plus = f >= D;
    const bool per = (p.periodic_mask >> a) & 1;
    const long long Ma = p.M[a];
    const int sa = inf.sub[a];
    long long nb;
    if (Ma == 1) nb = per ? e : -1;
    else if (!plus) nb = sa > 0 ? e - stride[a] : (per ? e + (Ma - 1) * stride[a] : -1);
    else nb = sa < Ma - 1 ? e + stride[a] : (per ? e - (Ma - 1) * stride[a] : -1);
    inf.nb[f] = nb;
    // global face index, extents M_d (d != a), M_a + !periodic (d == a)
    long long fi = 0;
    for (int d = D - 1; d >= 0; --d) {
      long long ext = p.M[d];
      long long sd = inf.sub[d];
      if (d == a) {
        ext = p.M[a] + (per ? 0 : 1);
        if (plus) {
          sd = sa + 1;
          if (per && sd == Ma) sd = 0;
        }
      }
      fi = fi * ext + sd;
    }
    inf.fidx[f] = fi;
  }
}

__device__ __forceinline__ int xcd_remap(int pb, int nb)
{
  // Blocks are dispatched round-robin over the 8 XCDs; give each XCD one
  // contiguous range of elements so face neighbours hit its own L2.
  const int per = nb / 8, rem = nb % 8;
  const int xcd = pb % 8, idx = pb / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Map a tangential multi-index t (extent EXT per tangential dim, dims != a in
// increasing order) and normal index i (extent EXTN) to the linear index of a
// D-dim tensor with extent EXT in every tangential dim and EXTN in dim a.
template <int D, int EXT, int EXTN>
__device__ __forceinline__ int face_to_full(int t, int i, int a)
{
  int idx = 0, mul = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (d == a) {
      idx += i * mul;
      mul *= EXTN;
    } else {
      idx += (t % EXT) * mul;
      t /= EXT;
      mul *= EXT;
    }
  }
  return idx;
}

// Tangential index of full index idx (extent EXT in all dims) for face axis a.
template <int D, int EXT>
__device__ __forceinline__ int full_to_face(int idx, int a, int& i_normal)
{
  int t = 0, mul = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int id = idx % EXT;
    idx /= EXT;
    if (d == a) i_normal = id;
    else {
      t += id * mul;
      mul *= EXT;
    }
  }
  return t;
}

// ---------------------------------------------------------------------------
// The fused kernel.  MODE = forward apply, transpose apply, or d/dbeta.
//   forward:   y = mass_coef*M x + adv_coef*A x
//   transpose: y = (mass_coef*M + adv_coef*A_lin)^T x
//   gradbeta:  dbeta arrays for L = <r, adv_coef*A(beta) x>, r = `y` input
// ---------------------------------------------------------------------------
template <class C, bool ADV>
__global__ void __launch_bounds__(C::NT) dg_kernel(Params<typename C::T> p, const typename C::T* __restrict__ x,
                                                   typename C::T* __restrict__ y)
{
  using T = typename C::T;
  constexpr int D = C::D, N = C::N, Q = C::Q, ND = C::ND, QD = C::QD, NT = C::NT, EPB = C::EPB;
  constexpr int MXD = C::MXD, MF = C::MF, NF = C::NF, QF = C::QF, FG = C::FG, TPF = C::TPF;
  constexpr int PE = C::PER_ELEM;
  constexpr int MODE = C::MODE;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  __shared__ ElemInfo<D> info[EPB];

  const int blk = xcd_remap(blockIdx.x, p.num_blocks);
  const long long e0 = (long long)blk * EPB;

  if (threadIdx.x < EPB) {
    const long long e = e0 + threadIdx.x;
    make_info<T, D>(p, e < p.E ? e : -1, info[threadIdx.x]);
  }
  // U = x (and for gradbeta Y = r), element-contiguous coalesced loads
  for (int i = threadIdx.x; i < EPB * ND; i += NT) {
    const int el = i / ND, j = i - el * ND;
    const bool valid = e0 + el < p.E;
    const long long g = e0 * ND + i;
    lds[el * PE + j] = valid ? x[g] : T(0);
    if constexpr (MODE == kGradBeta) lds[el * PE + ND + j] = valid ? y[g] : T(0);
    else lds[el * PE + ND + j] = T(0);
  }
  __syncthreads();

  auto U = [&](int el) { return lds + el * PE; };
  auto Y = [&](int el) { return lds + el * PE + ND; };
  auto S = [&](int el, int k) { return lds + el * PE + 2 * ND + k * MXD; };
  auto R = [&](int el, int reg) { return lds + el * PE + 2 * ND + reg * (TPF * FG * MF); };

  const T* __restrict__ Bt = p.tab.B;
  const T* __restrict__ Gq = p.tab.Gq;
  const T* __restrict__ b0 = p.tab.b0;
  const T* __restrict__ b1 = p.tab.b1;
  const T* __restrict__ wq = p.tab.w;

  // ======================= face terms =======================
  if constexpr (ADV) {
    for (int g = 0; g < 2 * D / FG; ++g) {
      // ---- slabs: contract the normal dimension (GLL: a selection) ----
      // trace slot layout in region 0: [TPF][FG][MF]
      for (int i = threadIdx.x; i < EPB * FG * NF; i += NT) {
        const int el = i / (FG * NF);
        const int rr = i - el * FG * NF;
        const int j = rr / NF;
        const int t = rr - j * NF;
        const int f = g * FG + j;
        const int a = f % D;
        const bool plus = f >= D;
        const ElemInfo<D>& inf = info[el];
        T own = T(0), own2 = T(0), nbv = T(0), nbv2 = T(0);
        const long long nb = inf.e >= 0 ? inf.nb[f] : -1;
        if (p.node_traces) {
          // the traces are node layers: this side's at its face, the
          // neighbour's at the opposite one
          const int io = face_to_full<D, N, N>(t, plus ? N - 1 : 0, a);
          own = U(el)[io];
          if constexpr (MODE == kGradBeta) own2 = Y(el)[io];
          if (nb >= 0) {
            const int in = face_to_full<D, N, N>(t, plus ? 0 : N - 1, a);
            nbv = x[nb * ND + in];
            if constexpr (MODE == kGradBeta) nbv2 = y[nb * ND + in];
          }
        } else {
          // contract the normal dimension with the basis at the face (the
          // tables are read at uniform addresses and picked per lane: a
          // per-lane table pointer would load them with vector loads)
#pragma unroll
          for (int q = 0; q < N; ++q) {
            const T c = plus ? b1[q] : b0[q];
            const int idx = face_to_full<D, N, N>(t, q, a);
            own += c * U(el)[idx];
            if constexpr (MODE == kGradBeta) own2 += c * Y(el)[idx];
          }
          if (nb >= 0) {
#pragma unroll
            for (int q = 0; q < N; ++q) {
              const T c = plus ? b0[q] : b1[q];
              const int idx = face_to_full<D, N, N>(t, q, a);
              nbv += c * x[nb * ND + idx];
              if constexpr (MODE == kGradBeta) nbv2 += c * y[nb * ND + idx];
            }
          }
        }
        T* r0 = R(el, 0);
        r0[(0 * FG + j) * MF + t] = own;
        r0[(1 * FG + j) * MF + t] = nbv;
        if constexpr (MODE == kGradBeta) {
          r0[(2 * FG + j) * MF + t] = own2;
          r0[(3 * FG + j) * MF + t] = nbv2;
        }
      }
      __syncthreads();
      // ---- tangential interpolation: D-1 sweeps N -> Q, last dim first ----
      static_for_down<D - 2>([&](auto jt_) {
        constexpr int jt = decltype(jt_)::value;
        constexpr int step = (D - 2) - jt;  // 0,1,...
        constexpr int src = step % 2, dst = 1 - src;
        constexpr int INNER = Pow<N, jt>::v;
        constexpr int OUTER = Pow<Q, D - 2 - jt>::v;
        sweep<NT, EPB * TPF * FG, OUTER, N, Q, INNER, false, T>(
            [&](int b) { return R(b / (TPF * FG), src) + (b % (TPF * FG)) * MF; },
            [&](int b) { return R(b / (TPF * FG), dst) + (b % (TPF * FG)) * MF; },
            [&](int j, int i) { return Bt[j * N + i]; });
        __syncthreads();
      });
      constexpr int TRREG = (D - 1) % 2;  // region holding the traces at face points
      constexpr int FLREG = 1 - TRREG;
      // ---- flux at face quadrature points ----
      for (int i = threadIdx.x; i < EPB * FG * QF; i += NT) {
        const int el = i / (FG * QF);
        const int rr = i - el * FG * QF;
        const int j = rr / QF;
        const int t = rr - j * QF;
        const int f = g * FG + j;
        const int a = f % D;
        const bool plus = f >= D;
        const ElemInfo<D>& inf = info[el];
        const T* tr = R(el, TRREG);
        const T own = tr[(0 * FG + j) * MF + t];
        const T nbv = tr[(1 * FG + j) * MF + t];
        T flux = T(0);
        const long long nb = inf.e >= 0 ? inf.nb[f] : -1;
        const bool boundary = nb < 0;
        if (inf.e >= 0 && !(boundary && p.bc_mode == GD_BC_NONE)) {
          int k[D];
          {
            int tt = t;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              if (d == a) k[d] = plus ? Q + 1 : Q;
              else {
                k[d] = tt % Q;
                tt /= Q;
              }
            }
          }
          T w = T(1);
#pragma unroll
          for (int d = 0; d < D; ++d) w *= (d == a) ? T(1) : wq[k[d]];
          T beta;
          if (p.beta_kind == GD_BETA_POINTWISE) beta = p.beta_face[a][inf.fidx[f] * QF + t];
          else beta = beta_at<T, D, Q>(p, a, inf.sub, k);
          const T nrm = (plus ? T(1) : T(-1)) * p.invh[a];
          const T dot = beta * nrm;
          const T wdj = w * p.detJ;
          if constexpr (MODE == kForward) {
            T ext;
            if (boundary) ext = p.bc ? eval_sep<T, D, Q>(*p.bc, inf.sub, k) : T(0);
            else ext = nbv;
            const T upwind = dot > T(0) ? dot * own : dot * ext;
            flux = p.adv_coef * (wdj * upwind);
          } else if constexpr (MODE == kTranspose) {
            T v = dot > T(0) ? wdj * (dot * own) : T(0);
            if (!boundary) {
              T nbeta;
              if (p.beta_kind == GD_BETA_POINTWISE) nbeta = beta;
              else {
                int nsub[D];
#pragma unroll
                for (int d = 0; d < D; ++d) nsub[d] = inf.sub[d];
                const int Ma = (int)p.M[a];
                nsub[a] = plus ? (inf.sub[a] + 1 == Ma ? 0 : inf.sub[a] + 1)
                               : (inf.sub[a] == 0 ? Ma - 1 : inf.sub[a] - 1);
                int nk[D];
#pragma unroll
                for (int d = 0; d < D; ++d) nk[d] = k[d];
                nk[a] = plus ? Q : Q + 1;
                nbeta = beta_at<T, D, Q>(p, a, nsub, nk);
              }
              const T ndot = nbeta * (-nrm);
              v += ndot > T(0) ? T(0) : wdj * (ndot * nbv);
            }
            flux = p.adv_coef * v;
          } else {  // gradbeta: write dbeta for faces this element owns
            const T rown = tr[(2 * FG + j) * MF + t];
            const T rnb = tr[(3 * FG + j) * MF + t];
            if (!plus || boundary) {
              T gsum;
              if (boundary) {
                const T ext = p.bc ? eval_sep<T, D, Q>(*p.bc, inf.sub, k) : T(0);
                const T up = dot > T(0) ? own : ext;
                gsum = rown * wdj * nrm * up;
              } else {
                const T up_own = dot > T(0) ? own : nbv;
                const T ndot = beta * (-nrm);
                const T up_nb = ndot > T(0) ? nbv : own;
                gsum = rown * wdj * nrm * up_own + rnb * wdj * (-nrm) * up_nb;
              }
              p.dbeta_face[a][inf.fidx[f] * QF + t] = p.adv_coef * gsum;
            }
          }
        } else if constexpr (MODE == kGradBeta) {
          if (inf.e >= 0 && boundary && !plus) p.dbeta_face[a][inf.fidx[f] * QF + t] = T(0);
          if (inf.e >= 0 && boundary && plus) p.dbeta_face[a][inf.fidx[f] * QF + t] = T(0);
        }
        if constexpr (MODE != kGradBeta) R(el, FLREG)[j * MF + t] = flux;
      }
      __syncthreads();
      if constexpr (MODE != kGradBeta) {
        // ---- tangential test functions: D-1 sweeps Q -> N ----
        static_for<0, D - 1>([&](auto jt_) {
          constexpr int jt = decltype(jt_)::value;
          constexpr int src = (FLREG + jt) % 2, dst = 1 - src;
          constexpr int INNER = Pow<N, jt>::v;
          constexpr int OUTER = Pow<Q, D - 2 - jt>::v;
          sweep<NT, EPB * FG, OUTER, Q, N, INNER, false, T>(
              [&](int b) { return R(b / FG, src) + (b % FG) * MF; },
              [&](int b) { return R(b / FG, dst) + (b % FG) * MF; },
              [&](int j, int i) { return Bt[i * N + j]; });
          __syncthreads();
        });
        constexpr int GREG = (FLREG + (D - 1)) % 2;
        // ---- add face contributions into Y ----
        for (int i = threadIdx.x; i < EPB * ND; i += NT) {
          const int el = i / ND;
          const int idx = i - el * ND;
          T acc = T(0);
#pragma unroll
          for (int j = 0; j < FG; ++j) {
            const int f = g * FG + j;
            const int a = f % D;
            const bool plus = f >= D;
            int in;
            const int t = full_to_face<D, N>(idx, a, in);
            if (p.node_traces) {
              if (in == (plus ? N - 1 : 0)) acc += R(el, GREG)[j * MF + t];
            } else {
              const T c = plus ? b1[in] : b0[in];
              if (c != T(0)) acc += c * R(el, GREG)[j * MF + t];
            }
          }
          Y(el)[idx] += acc;
        }
        __syncthreads();
      }
    }
  }

  // ======================= cell terms =======================
  // interpolation to quadrature points, last dim first: U -> S0 -> S1 -> ...
  static_for_down<D - 1>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    constexpr int step = (D - 1) - d;
    constexpr int INNER = Pow<N, d>::v;
    constexpr int OUTER = Pow<Q, D - 1 - d>::v;
    sweep<NT, EPB, OUTER, N, Q, INNER, false, T>(
        [&](int b) { return step == 0 ? U(b) : S(b, (step - 1) % 2); },
        [&](int b) { return S(b, step % 2); }, [&](int j, int i) { return Bt[j * N + i]; });
    __syncthreads();
  });
  constexpr int BU = (D - 1) % 2;  // scratch slot holding B x
  constexpr int FREE = 1 - BU;
  constexpr int ACCS = 2;

  auto point_k = [&](int i, int* k) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      k[d] = i % Q;
      i /= Q;
    }
  };
  auto cell_weight = [&](const int* k) {
    T w = T(1);
#pragma unroll
    for (int d = 0; d < D; ++d) w *= wq[k[d]];
    return w;
  };
  auto cell_beta = [&](int el, int d, int i, const int* k) -> T {
    if (p.beta_kind == GD_BETA_POINTWISE) return p.beta_cell[(info[el].e * D + d) * QD + i];
    return beta_at<T, D, Q>(p, d, info[el].sub, k);
  };

  if constexpr (MODE == kGradBeta) {
    // B r into the free slot: Y(r) -> FREE? Use ping-pong through ACCS.
    static_for_down<D - 1>([&](auto d_) {
      constexpr int d = decltype(d_)::value;
      constexpr int step = (D - 1) - d;
      constexpr int INNER = Pow<N, d>::v;
      constexpr int OUTER = Pow<Q, D - 1 - d>::v;
      // slots alternate FREE <-> ACCS ending in FREE when D odd, ACCS when D even
      constexpr int a_slot = (step % 2 == 0) ? ((D % 2 == 1) ? FREE : ACCS) : ((D % 2 == 1) ? ACCS : FREE);
      constexpr int p_slot = (a_slot == FREE) ? ACCS : FREE;
      sweep<NT, EPB, OUTER, N, Q, INNER, false, T>(
          [&](int b) { return step == 0 ? Y(b) : S(b, p_slot); }, [&](int b) { return S(b, a_slot); },
          [&](int j, int i) { return Bt[j * N + i]; });
      __syncthreads();
    });
    constexpr int BR = FREE;     // B r
    constexpr int GS = ACCS;     // gradient scratch
    static_for<0, D>([&](auto d_) {
      constexpr int d = decltype(d_)::value;
      constexpr int INNER = Pow<Q, d>::v;
      constexpr int OUTER = Pow<Q, D - 1 - d>::v;
      sweep<NT, EPB, OUTER, Q, Q, INNER, false, T>([&](int b) { return S(b, BR); },
                                                  [&](int b) { return S(b, GS); },
                                                  [&](int j, int i) { return Gq[j * Q + i]; });
      __syncthreads();
      for (int i = threadIdx.x; i < EPB * QD; i += NT) {
        const int el = i / QD, qi = i - el * QD;
        if (info[el].e < 0) continue;
        int k[D];
        point_k(qi, k);
        const T w = cell_weight(k);
        const T wb = -w * p.detJ * S(el, BU)[qi];
        p.dbeta_cell[(info[el].e * D + d) * QD + qi] = p.adv_coef * (S(el, GS)[qi] * (wb * p.invh[d]));
      }
      __syncthreads();
    });
    return;
  } else {
    // ---- quadrature-point terms accumulated into ACCS ----
    if constexpr (ADV) {
      static_for<0, D>([&](auto d_) {
        constexpr int d = decltype(d_)::value;
        constexpr int INNER = Pow<Q, d>::v;
        constexpr int OUTER = Pow<Q, D - 1 - d>::v;
        if constexpr (MODE == kForward) {
          // FREE = c_d * Bx, then ACCS (+)= Gq_d^T FREE
          for (int i = threadIdx.x; i < EPB * QD; i += NT) {
            const int el = i / QD, qi = i - el * QD;
            int k[D];
            point_k(qi, k);
            const T w = cell_weight(k);
            const T c = (info[el].e >= 0) ? p.adv_coef * (-w * p.detJ * (cell_beta(el, d, qi, k) * p.invh[d])) : T(0);
            S(el, FREE)[qi] = c * S(el, BU)[qi];
          }
          __syncthreads();
          sweep<NT, EPB, OUTER, Q, Q, INNER, (d > 0), T>([&](int b) { return S(b, FREE); },
                                                       [&](int b) { return S(b, ACCS); },
                                                       [&](int j, int i) { return Gq[i * Q + j]; });
          __syncthreads();
        } else {
          // FREE = Gq_d Bx, then ACCS (+)= c_d * FREE
          sweep<NT, EPB, OUTER, Q, Q, INNER, false, T>([&](int b) { return S(b, BU); },
                                                      [&](int b) { return S(b, FREE); },
                                                      [&](int j, int i) { return Gq[j * Q + i]; });
          __syncthreads();
          for (int i = threadIdx.x; i < EPB * QD; i += NT) {
            const int el = i / QD, qi = i - el * QD;
            int k[D];
            point_k(qi, k);
            const T w = cell_weight(k);
            const T c = (info[el].e >= 0) ? p.adv_coef * (-w * p.detJ * (cell_beta(el, d, qi, k) * p.invh[d])) : T(0);
            const T v = c * S(el, FREE)[qi];
            if (d == 0) S(el, ACCS)[qi] = v;
            else S(el, ACCS)[qi] += v;
          }
          __syncthreads();
        }
      });
    }
    if (p.mass_coef != T(0)) {
      for (int i = threadIdx.x; i < EPB * QD; i += NT) {
        const int el = i / QD, qi = i - el * QD;
        int k[D];
        point_k(qi, k);
        const T w = cell_weight(k);
        const T sigma = (info[el].e >= 0) ? eval_sep<T, D, Q>(*p.sigma, info[el].sub, k) : T(0);
        const T v = p.mass_coef * (w * p.detJ * sigma * S(el, BU)[qi]);
        if constexpr (ADV) S(el, ACCS)[qi] += v;
        else S(el, ACCS)[qi] = v;
      }
      __syncthreads();
    }
    // ---- test functions: D sweeps Q -> N, first dim first, last into Y ----
    static_for<0, D>([&](auto d_) {
      constexpr int d = decltype(d_)::value;
      constexpr int INNER = Pow<N, d>::v;
      constexpr int OUTER = Pow<Q, D - 1 - d>::v;
      // slots: ACCS -> FREE -> BU -> FREE ... ; last -> Y (accumulate)
      constexpr int src = d == 0 ? ACCS : ((d % 2 == 1) ? FREE : BU);
      constexpr int dst = (d % 2 == 0) ? FREE : BU;
      if constexpr (d == D - 1) {
        sweep<NT, EPB, OUTER, Q, N, INNER, true, T>([&](int b) { return S(b, src); },
                                                   [&](int b) { return Y(b); },
                                                   [&](int j, int i) { return Bt[i * N + j]; });
      } else {
        sweep<NT, EPB, OUTER, Q, N, INNER, false, T>([&](int b) { return S(b, src); },
                                                    [&](int b) { return S(b, dst); },
                                                    [&](int j, int i) { return Bt[i * N + j]; });
      }
      __syncthreads();
    });
    // ---- store ----
    for (int i = threadIdx.x; i < EPB * ND; i += NT) {
      const int el = i / ND;
      if (e0 + el < p.E) y[e0 * ND + i] = Y(el)[i - el * ND];
    }
  }
}

template <class C>
constexpr size_t lds_bytes()
{
  return (size_t)C::EPB * C::PER_ELEM * sizeof(typename C::T);
}

}  // namespace dev
}  // namespace gd
