// gfx950 kernel of the SIPDG diffusion apply on Cartesian meshes: volume and
// face terms of one element fused in one pass, one workgroup per batch of EPB
// consecutive elements, every element writing only its own DoFs (no atomics).
//
// What is computed follows GenDiL's DiffusionOperator:
//   volume  DiffusionElementOperator  diffusionoperator.hpp:44-101
//           y = B^T sum_d Gq_d^T (w detJ mu / h_d^2) Gq_d B u
//   faces   DiffusionFaceOperator     diffusionoperator.hpp:130-331
//           per face point, with u, g the trace and the reference normal
//           derivative of this side (-) and of the neighbour (+), n = s/h_a:
//             interior: avg = mu/2 (g- + g+)/h_a, jump = u- - u+, c = 1/2
//             boundary: avg = mu g-/h_a,          jump = u-,      c = 1
//             Du = -w detJ avg n + w detJ kappa mu jump / h_a^2
//             DG = sigma w detJ c mu jump n / h_a
//           tested with the basis values (Du) and the basis normal
//           derivatives (DG) of this side's face.
//
// Layout per element in LDS: U (input), Y (output), scratch.  Every 1D
// contraction is a pencil sweep (gd_kernels.hpp): a thread reads one pencil
// from LDS, contracts it with the 1D table and writes the result pencil.  The
// neighbour's traces are read straight from HBM/L2 inside the same pass; the
// normal derivative needs the neighbour's whole pencil along the face normal.
#pragma once

#include "gd_kernels.hpp"

namespace gd {
namespace dev {

template <typename T_, int D_, int N_, int Q_>
struct DCfg {
  using T = T_;
  static constexpr int D = D_, N = N_, Q = Q_;
  static constexpr int ND = Pow<N, D>::v;
  static constexpr int QD = Pow<Q, D>::v;
  static constexpr int MX = cmax(N, Q);
  static constexpr int MXD = Pow<MX, D>::v;
  static constexpr int NF = Pow<N, D - 1>::v;
  static constexpr int QF = Pow<Q, D - 1>::v;
  static constexpr int MF = Pow<MX, D - 1>::v;
  // traces per face: own value, own normal derivative, neighbour value,
  // neighbour normal derivative
  static constexpr int TPF = 4;
  // cell scratch: two ping-pong tensors, the accumulator, w detJ mu
  static constexpr int CELL_SCR = 4 * MXD;
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
  // the diffusivity's 1D factor samples of the element (Q rule points, then
  // xi = 0, xi = 1) per term and axis, 1 where a term has no factor
  static constexpr int MUS = kMaxTerms * D * (Q + 2);
  static constexpr int PER_ELEM = 2 * ND + SCR + MUS;
  // measured on MI355X for 3D n = 4, q = 5 FP64 (64^3): 128 threads over
  // 16 KB of elements 945 us, 256 threads over 32 KB 1000-1050 us, 64
  // threads 1010-2270 us, 64 KB budgets 1470-2300 us
  static constexpr int LDS_BUDGET = D == 3 ? 16 * 1024 : 32 * 1024;
  static constexpr int EPB_RAW = LDS_BUDGET / (int)(PER_ELEM * sizeof(T));
  static constexpr int EPB = EPB_RAW < 1 ? 1 : (EPB_RAW > 16 ? 16 : EPB_RAW);
  static constexpr int NT = D >= 6 || D == 3 ? 128 : 256;
};

template <class C>
__global__ void __launch_bounds__(C::NT) diff_kernel(Params<typename C::T> p, FieldDev<typename C::T> mu,
                                                     const typename C::T* __restrict__ x,
                                                     typename C::T* __restrict__ y)
{
  using T = typename C::T;
  constexpr int D = C::D, N = C::N, Q = C::Q, ND = C::ND, QD = C::QD, NT = C::NT, EPB = C::EPB;
  constexpr int MXD = C::MXD, MF = C::MF, NF = C::NF, QF = C::QF, FG = C::FG, TPF = C::TPF;
  constexpr int PE = C::PER_ELEM;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  __shared__ ElemInfo<D> info[EPB];

  const int blk = xcd_remap(blockIdx.x, p.num_blocks);
  const long long e0 = (long long)blk * EPB;

  if (threadIdx.x < EPB) {
    const long long e = e0 + threadIdx.x;
    make_info<T, D>(p, e < p.E ? e : -1, info[threadIdx.x]);
  }
  for (int i = threadIdx.x; i < EPB * ND; i += NT) {
    const int el = i / ND, j = i - el * ND;
    const bool valid = e0 + el < p.E;
    lds[el * PE + j] = valid ? x[e0 * ND + i] : T(0);
    lds[el * PE + ND + j] = T(0);
  }
  __syncthreads();

  auto U = [&](int el) { return lds + el * PE; };
  auto Y = [&](int el) { return lds + el * PE + ND; };
  auto S = [&](int el, int k) { return lds + el * PE + 2 * ND + k * MXD; };
  auto R = [&](int el, int reg) { return lds + el * PE + 2 * ND + reg * (TPF * FG * MF); };
  auto MU = [&](int el) { return lds + el * PE + 2 * ND + C::SCR; };

  // the element's diffusivity factor samples, (term, axis) runs of Q + 2
  constexpr int QS = Q + 2;
  for (int i = threadIdx.x; i < EPB * mu.rank * D * QS; i += NT) {
    const int el = i / (mu.rank * D * QS);
    const int rr = i - el * (mu.rank * D * QS);
    const int rd = rr / QS, k = rr - rd * QS;
    const int r = rd / D, d = rd - r * D;
    const T* tb = mu.tab[r][d];
    T v = T(1);
    if (tb && info[el].e >= 0) v = tb[info[el].sub[d] * QS + k];
    MU(el)[rd * QS + k] = v;
  }
  __syncthreads();
  // mu at per-axis sample indices k (k[d] < Q rule point, Q: xi = 0, Q + 1: xi = 1)
  auto mu_at = [&](int el, const int* k) {
    const T* m = MU(el);
    T sum = T(0);
    for (int r = 0; r < mu.rank; ++r) {
      T t = mu.scale[r];
#pragma unroll
      for (int d = 0; d < D; ++d) t *= m[(r * D + d) * QS + k[d]];
      sum += t;
    }
    return sum;
  };

  const T* __restrict__ Bt = p.tab.B;
  const T* __restrict__ Gq = p.tab.Gq;
  const T* __restrict__ b0 = p.tab.b0;
  const T* __restrict__ b1 = p.tab.b1;
  const T* __restrict__ d0 = p.tab.d0;
  const T* __restrict__ d1 = p.tab.d1;
  const T* __restrict__ wq = p.tab.w;

  // ======================= face terms =======================
  for (int g = 0; g < 2 * D / FG; ++g) {
    // ---- normal dimension: values and derivatives at the face, both sides ----
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
      const T* bo = plus ? b1 : b0;  // this side's face
      const T* go = plus ? d1 : d0;
      const T* bn = plus ? b0 : b1;  // the neighbour's face is its opposite one
      const T* gn = plus ? d0 : d1;
      T ov = T(0), og = T(0), nv = T(0), ng = T(0);
#pragma unroll
      for (int q = 0; q < N; ++q) {
        const T u = U(el)[face_to_full<D, N, N>(t, q, a)];
        ov += bo[q] * u;
        og += go[q] * u;
      }
      const long long nb = inf.e >= 0 ? inf.nb[f] : -1;
      if (nb >= 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const T u = x[nb * ND + face_to_full<D, N, N>(t, q, a)];
          nv += bn[q] * u;
          ng += gn[q] * u;
        }
      }
      T* r0 = R(el, 0);
      r0[(0 * FG + j) * MF + t] = ov;
      r0[(1 * FG + j) * MF + t] = og;
      r0[(2 * FG + j) * MF + t] = nv;
      r0[(3 * FG + j) * MF + t] = ng;
    }
    __syncthreads();
    // ---- tangential interpolation: D-1 sweeps N -> Q, last dim first ----
    static_for_down<D - 2>([&](auto jt_) {
      constexpr int jt = decltype(jt_)::value;
      constexpr int step = (D - 2) - jt;
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
    // ---- fluxes at the face quadrature points ----
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
      T du = T(0), dg = T(0);
      if (inf.e >= 0) {
        const bool boundary = inf.nb[f] < 0;
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
        const T m = mu_at(el, k);
        const T ov = tr[(0 * FG + j) * MF + t];
        const T g_own = tr[(1 * FG + j) * MF + t] * p.invh[a];
        T avg, jump, half;
        if (!boundary) {
          const T g_nb = tr[(3 * FG + j) * MF + t] * p.invh[a];
          avg = T(0.5) * m * (g_own + g_nb);
          jump = ov - tr[(2 * FG + j) * MF + t];
          half = T(0.5);
        } else {
          avg = m * g_own;
          jump = ov;
          half = T(1);
        }
        const T nrm = (plus ? T(1) : T(-1)) * p.invh[a];
        const T dot = avg * nrm;
        const T norm = p.invh[a];
        const T wdj = w * p.detJ;
        du = -wdj * dot + wdj * norm * p.ip_kappa * norm * m * jump;
        dg = (p.ip_sigma * wdj * jump * m * half * nrm) * p.invh[a];
      }
      T* fl = R(el, FLREG);
      fl[(0 * FG + j) * MF + t] = du;
      fl[(1 * FG + j) * MF + t] = dg;
    }
    __syncthreads();
    // ---- tangential test functions: D-1 sweeps Q -> N over both fluxes ----
    static_for<0, D - 1>([&](auto jt_) {
      constexpr int jt = decltype(jt_)::value;
      constexpr int src = (FLREG + jt) % 2, dst = 1 - src;
      constexpr int INNER = Pow<N, jt>::v;
      constexpr int OUTER = Pow<Q, D - 2 - jt>::v;
      sweep<NT, EPB * 2 * FG, OUTER, Q, N, INNER, false, T>(
          [&](int b) { return R(b / (2 * FG), src) + (b % (2 * FG)) * MF; },
          [&](int b) { return R(b / (2 * FG), dst) + (b % (2 * FG)) * MF; },
          [&](int j, int i) { return Bt[i * N + j]; });
      __syncthreads();
    });
    constexpr int GREG = (FLREG + (D - 1)) % 2;
    // ---- normal dimension: basis values and derivatives of this side ----
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
        const T cb = plus ? b1[in] : b0[in];
        const T cg = plus ? d1[in] : d0[in];
        acc += cb * R(el, GREG)[(0 * FG + j) * MF + t] + cg * R(el, GREG)[(1 * FG + j) * MF + t];
      }
      Y(el)[idx] += acc;
    }
    __syncthreads();
  }

  // ======================= volume term =======================
  // interpolation to the rule points, last dim first: U -> S0 -> S1 -> ...
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
  constexpr int BU = (D - 1) % 2;  // B u
  constexpr int FREE = 1 - BU;
  constexpr int ACCS = 2;
  constexpr int CW = 3;  // w detJ mu at the rule points
  for (int i = threadIdx.x; i < EPB * QD; i += NT) {
    const int el = i / QD, qi = i - el * QD;
    int k[D];
    int r = qi;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      k[d] = r % Q;
      r /= Q;
    }
    T w = T(1);
#pragma unroll
    for (int d = 0; d < D; ++d) w *= wq[k[d]];
    S(el, CW)[qi] = info[el].e >= 0 ? w * p.detJ * mu_at(el, k) : T(0);
  }
  __syncthreads();  // every pencil below reads w detJ mu at points other threads wrote
  static_for<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    constexpr int INNER = Pow<Q, d>::v;
    constexpr int OUTER = Pow<Q, D - 1 - d>::v;
    constexpr int PEN = OUTER * INNER;
    // G_d^T diag(w detJ mu / h_d^2) G_d acts within each pencil along d: one
    // thread per pencil forms the reference gradient at the pencil's rule
    // points (Gq_d B u), maps it with inv_J twice (ApplyMapping,
    // ApplyMappingTranspose), scales it by w detJ mu and adds Gq_d^T of it
    // into the accumulator
    const T ih = p.invh[d];
    for (int t = threadIdx.x; t < EPB * PEN; t += NT) {
      const int b = t / PEN;
      const int r = t - b * PEN;
      const int o = r / INNER;
      const int a = r - o * INNER;
      const int base = o * Q * INNER + a;
      const T* src = S(b, BU) + base;
      const T* cw = S(b, CW) + base;
      T* acc = S(b, ACCS) + base;
      T v[Q], g[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) v[i] = src[i * INNER];
#pragma unroll
      for (int j = 0; j < Q; ++j) {
        T s = T(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) s += Gq[j * Q + i] * v[i];
        s *= ih;
        s *= ih;
        g[j] = cw[j * INNER] * s;
      }
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < Q; ++j) s += Gq[j * Q + i] * g[j];
        if constexpr (d == 0) acc[i * INNER] = s;
        else acc[i * INNER] += s;
      }
    }
    __syncthreads();
  });
  // ---- test functions: D sweeps Q -> N, first dim first, last into Y ----
  static_for<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    constexpr int INNER = Pow<N, d>::v;
    constexpr int OUTER = Pow<Q, D - 1 - d>::v;
    constexpr int src = d == 0 ? ACCS : ((d % 2 == 1) ? FREE : BU);
    constexpr int dst = (d % 2 == 0) ? FREE : BU;
    if constexpr (d == D - 1) {
      sweep<NT, EPB, OUTER, Q, N, INNER, true, T>([&](int b) { return S(b, src); }, [&](int b) { return Y(b); },
                                                 [&](int j, int i) { return Bt[i * N + j]; });
    } else {
      sweep<NT, EPB, OUTER, Q, N, INNER, false, T>([&](int b) { return S(b, src); },
                                                  [&](int b) { return S(b, dst); },
                                                  [&](int j, int i) { return Bt[i * N + j]; });
    }
    __syncthreads();
  });
  for (int i = threadIdx.x; i < EPB * ND; i += NT) {
    const int el = i / ND;
    if (e0 + el < p.E) y[e0 * ND + i] = Y(el)[i - el * ND];
  }
}

template <class C>
constexpr size_t diff_lds_bytes()
{
  return (size_t)C::EPB * C::PER_ELEM * sizeof(typename C::T);
}

}  // namespace dev

// Registry of compiled (dim, nodes, points) instances of the diffusion kernel.
struct DiffSet {
  int D, N, Q;
  int (*f64)(const Params<double>&, const FieldDev<double>*, const double*, double*, hipStream_t);
  int (*f32)(const Params<float>&, const FieldDev<float>*, const float*, float*, hipStream_t);
};

struct DiffList {
  const DiffSet* sets;
  int count;
};

DiffList diff_kernels_d12();
DiffList diff_kernels_d3();
DiffList diff_kernels_d6();

template <class C>
int launch_diff(const Params<typename C::T>& prm_in, const FieldDev<typename C::T>* mu, const typename C::T* x,
                typename C::T* y, hipStream_t s)
{
  using T = typename C::T;
  auto kern = dev::diff_kernel<C>;
  const size_t lds = dev::diff_lds_bytes<C>();
  static bool attr_done = false;
  if (lds + sizeof(dev::ElemInfo<C::D>) * C::EPB > 160 * 1024)
    return fail(GD_ERR_UNSUPPORTED, "element tensors of this order/precision exceed the 160 KB LDS");
  if (lds > 64 * 1024 && !attr_done) {
    GD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    attr_done = true;
  }
  Params<T> prm = prm_in;
  const long long nb = (prm.E + C::EPB - 1) / C::EPB;
  if (nb > 0x7fffffffLL) return fail(GD_ERR_INVALID, "too many elements for one launch");
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(C::NT), lds, s, prm, *mu, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

#define GD_DSET(D, N, Q) \
  DiffSet { D, N, Q, &launch_diff<dev::DCfg<double, D, N, Q>>, &launch_diff<dev::DCfg<float, D, N, Q>> }

}  // namespace gd
