// Kronecker-factored 3D kernel for GLL elements with Q == N and "one-axis"
// separable coefficients (each component beta_d a constant or a function of
// one coordinate; sigma constant), each component of one sign at every face
// point (the formulas below are for beta_d >= 0; for beta_d <= 0 the own and
// neighbour roles of the two faces swap).
//
// With such coefficients every term of the element operator is a Kronecker
// product of three N x N matrices:
//   cell term d:  -K_d (F_x (x) F_y (x) F_z),  F_d = Gb^T W[beta] B (derivative
//                 axis), F_a = B^T W[beta] B for a != d (the beta factor sits on
//                 the one axis beta_d depends on), K_d = adv_coef detJ / h_d;
//   + face, d:    +K_d (E_{N-1,N-1} (x) the same tangential factors): the
//                 upwind flux takes the own trace since beta.n >= 0 there;
//   - face, d:    -K_d (e_0 e_{N-1}^T (x) the same factors) applied to the
//                 neighbour's node slab N-1 (beta.n <= 0: the neighbour's trace);
//   mass:         Km (M (x) M (x) M),  M = B^T W B.
// So the own-face terms fold into the cell term's derivative factor, the
// neighbour terms need only the neighbour's face slab pushed through the two
// tangential factors, and no quadrature-point data is formed at all.  This is
// the same operator as GenDiL's element-centric apply (massoperator.hpp:39-89,
// advectionoperator.hpp:106-441: B, the q-function, the collocated gradient
// Gq with Gq B = Gb, B^T, and the upwind flux on GLL face traces), regrouped.
//
// Mapping: LANES lanes per element split along z; lane h owns output z-planes
// [h*PL, h*PL+PL).  The z factors are applied first, streaming the element's
// input planes (every lane of an element loads the same planes, which the
// memory pipeline broadcasts), so no data crosses lanes; the x and y factors
// are then in-register contractions of each owned plane.  The 1D matrices of
// constant coefficients are compile-time constants; beta-dependent factors
// come from small per-position tables built once per plan.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "gd_consttab.hpp"
#include "gd_internal.hpp"

namespace gd {
namespace k3k {

template <typename T>
struct KParams {
  int E, M0, M1, M2;
  int periodic;    // bit d = periodic axis d
  int bc_zero;     // boundary faces: 0 skipped, 1 inflow with g = 0
  int negmask;     // bit d: beta_d <= 0 at every face point (else >= 0)
  int num_blocks;
  T Kt[3];         // adv_coef detJ / h_d, times beta_d's value when beta_d is constant
  T Km;            // mass_coef detJ sigma
  const T* tab[3];    // [M_axis][N][N]: B^T W[beta_d] B, or Gb^T W[beta_d] B when beta_d varies along d
  const T* fsamp[3];  // [M_d][2]: beta_d at xi = 0 and xi = 1, when beta_d varies along d
};

// axis signature (as for the register kernel): digits (axis + 1) base 4
constexpr int ksig(int b0, int b1, int b2) { return (b0 + 1) + 4 * (b1 + 1) + 16 * (b2 + 1); }
constexpr int ksig_axis(int S, int f) { return ((S >> (2 * f)) & 3) - 1; }

template <int N>
struct Mat {
  double a[N][N];
};

template <int N>
struct KTab {
  static constexpr ct::Tables<N, N> t = ct::make_gll_gauss<N, N>();
  static constexpr Mat<N> make(bool deriv)
  {
    Mat<N> m{};
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (int q = 0; q < N; ++q) s += (deriv ? t.Gb[q][i] : t.B[q][i]) * t.w[q] * t.B[q][j];
        m.a[i][j] = s;
      }
    return m;
  }
  static constexpr Mat<N> M = make(false);
  static constexpr Mat<N> D = make(true);
};

template <int I, int NN, typename F>
__device__ __forceinline__ void kfor(F&& f)
{
  if constexpr (I < NN) {
    f(std::integral_constant<int, I>{});
    kfor<I + 1, NN>(f);
  }
}

// row (h*PL + kk) of a constant matrix, h lane dependent
template <int N, int LANES, int PL, typename T>
__device__ __forceinline__ T lrow(const Mat<N>& A, int h, int kk, int col)
{
  T c = (T)A.a[kk][col];
#pragma unroll
  for (int g = 1; g < LANES; ++g) c = (h == g) ? (T)A.a[g * PL + kk][col] : c;
  return c;
}

template <typename T, int NV>
struct alignas(16) Chunk {
  T v[NV];
};

template <typename T, int COUNT>
__device__ __forceinline__ void load_block(const T* __restrict__ src, T* dst)
{
  constexpr int VW = 16 / sizeof(T);
  if constexpr ((COUNT * sizeof(T)) % 16 == 0) {
    const Chunk<T, VW>* s = reinterpret_cast<const Chunk<T, VW>*>(src);
#pragma unroll
    for (int v = 0; v < COUNT / VW; ++v) {
      const Chunk<T, VW> c = s[v];
#pragma unroll
      for (int l = 0; l < VW; ++l) dst[v * VW + l] = c.v[l];
    }
  } else {
#pragma unroll
    for (int v = 0; v < COUNT; ++v) dst[v] = src[v];
  }
}

__device__ __forceinline__ int remap(int pb, int nb)
{
  // blocks are dispatched round-robin over the 8 XCDs: give each XCD one
  // contiguous range of elements so face neighbours hit its own L2
  const int per = nb / 8, rem = nb % 8;
  const int xcd = pb % 8, idx = pb / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

template <typename T, int N, int LANES, bool ADV, bool MASS, int SIG, int NT>
__global__ void __launch_bounds__(NT) k3k_apply(KParams<T> p, const T* __restrict__ x, T* __restrict__ y)
{
  constexpr int PL = N / LANES;
  constexpr int N2 = N * N, N3 = N * N * N;
  constexpr int EPB = NT / LANES;
  constexpr int AX0 = ksig_axis(SIG, 0), AX1 = ksig_axis(SIG, 1), AX2 = ksig_axis(SIG, 2);
  // the mass term reuses term x's z factor when that factor is the plain M
  constexpr bool MSH = ADV && MASS && AX0 != 2;
  using KT = KTab<N>;

  const int h = LANES == 1 ? 0 : (int)(threadIdx.x % LANES);
  const int blk = remap(blockIdx.x, p.num_blocks);
  const int e_raw = blk * EPB + (int)(threadIdx.x / LANES);
  const bool valid = e_raw < p.E;
  const int e = valid ? e_raw : p.E - 1;

  int sub[3];
  {
    const unsigned r = (unsigned)e / (unsigned)p.M0;
    sub[0] = e - (int)r * p.M0;
    sub[2] = (int)(r / (unsigned)p.M1);
    sub[1] = (int)r - sub[2] * p.M1;
  }

  // ---- face coefficients and upwind-side neighbours ----
  // beta_d >= 0: own trace on the + face (node N-1), the - neighbour's node
  // N-1 slab on the - face (node 0).  beta_d <= 0: own trace on the - face,
  // the + neighbour's node 0 slab on the + face.  cP*/cM* act on output node
  // N-1 / 0 along d; nsl = node of the neighbour slab.
  T cPown[3] = {T(0), T(0), T(0)}, cMown[3] = {T(0), T(0), T(0)};
  T cPnb[3] = {T(0), T(0), T(0)}, cMnb[3] = {T(0), T(0), T(0)};
  int nbe[3] = {e, e, e}, nsl[3] = {N - 1, N - 1, N - 1};
  if constexpr (ADV) {
    kfor<0, 3>([&](auto d_) {
      constexpr int d = decltype(d_)::value;
      constexpr int AXd = d == 0 ? AX0 : (d == 1 ? AX1 : AX2);
      const int Ma = d == 0 ? p.M0 : (d == 1 ? p.M1 : p.M2);
      const int st = d == 0 ? 1 : (d == 1 ? p.M0 : p.M0 * p.M1);
      const int sa = sub[d];
      const bool per = (p.periodic >> d) & 1;
      const bool neg = (p.negmask >> d) & 1;
      int nlo, nhi;
      if (Ma == 1) {
        nlo = nhi = per ? e : -1;
      } else {
        nlo = sa > 0 ? e - st : (per ? e + (Ma - 1) * st : -1);
        nhi = sa < Ma - 1 ? e + st : (per ? e - (Ma - 1) * st : -1);
      }
      T f0 = T(1), f1 = T(1);
      if constexpr (AXd == d) {
        f0 = p.fsamp[d][2 * sa];
        f1 = p.fsamp[d][2 * sa + 1];
      }
      if (!neg) {
        const bool own_ok = nhi >= 0 || p.bc_zero;
        cPown[d] = own_ok ? p.Kt[d] * f1 : T(0);
        cMnb[d] = nlo >= 0 ? -p.Kt[d] * f0 : T(0);
        nbe[d] = nlo >= 0 ? nlo : e;
        nsl[d] = N - 1;
      } else {
        const bool own_ok = nlo >= 0 || p.bc_zero;
        cMown[d] = own_ok ? -p.Kt[d] * f0 : T(0);
        cPnb[d] = nhi >= 0 ? p.Kt[d] * f1 : T(0);
        nbe[d] = nhi >= 0 ? nhi : e;
        nsl[d] = 0;
      }
    });
  }

  // ---- z stage: every term's z factor, streaming the input planes ----
  // z-factor rows of the lane's output planes
  T zX[PL][N], zY[PL][N], zZ[PL][N];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int row = h * PL + kk;
      if constexpr (AX0 == 2) zX[kk][j] = p.tab[0][(sub[2] * N + row) * N + j];
      else zX[kk][j] = lrow<N, LANES, PL, T>(KT::M, h, kk, j);
      if constexpr (ADV) {
        if constexpr (AX1 == 2) zY[kk][j] = p.tab[1][(sub[2] * N + row) * N + j];
        else zY[kk][j] = lrow<N, LANES, PL, T>(KT::M, h, kk, j);
        T dz;
        if constexpr (AX2 == 2) dz = p.tab[2][(sub[2] * N + row) * N + j];
        else dz = lrow<N, LANES, PL, T>(KT::D, h, kk, j);
        zZ[kk][j] = -p.Kt[2] * dz + ((row == N - 1 && j == N - 1) ? cPown[2] : T(0)) +
                    ((row == 0 && j == 0) ? cMown[2] : T(0));
      }
    }
  T accX[PL][N2], accY[PL][N2], accZ[PL][N2], accM[(MASS && !MSH) ? PL : 1][N2];
#pragma unroll
  for (int kk = 0; kk < PL; ++kk)
#pragma unroll
    for (int m = 0; m < N2; ++m) {
      accX[kk][m] = T(0);
      accY[kk][m] = T(0);
      accZ[kk][m] = T(0);
      if constexpr (MASS && !MSH) accM[kk][m] = T(0);
    }
  const T* xe = x + (long long)e * N3;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T u[N2];
    load_block<T, N2>(xe + j * N2, u);
#pragma unroll
    for (int kk = 0; kk < PL; ++kk) {
      T mz = T(0);
      if constexpr (MASS && !MSH) mz = lrow<N, LANES, PL, T>(KT::M, h, kk, j);
#pragma unroll
      for (int m = 0; m < N2; ++m) {
        if constexpr (ADV || MSH) accX[kk][m] += zX[kk][j] * u[m];
        if constexpr (ADV) {
          accY[kk][m] += zY[kk][j] * u[m];
          accZ[kk][m] += zZ[kk][j] * u[m];
        }
        if constexpr (MASS && !MSH) accM[kk][m] += mz * u[m];
      }
    }
  }
  T sx[PL][N], sy[PL][N];
  if constexpr (ADV) {
    // upwind neighbour across a z face: its slab plane enters output plane 0
    // (beta_z >= 0) or N-1 (beta_z <= 0)
    {
      T un[N2];
      load_block<T, N2>(x + (long long)nbe[2] * N3 + nsl[2] * N2, un);
#pragma unroll
      for (int kk = 0; kk < PL; ++kk) {
        const int row = h * PL + kk;
        const T c = (row == 0 ? cMnb[2] : T(0)) + (row == N - 1 ? cPnb[2] : T(0));
#pragma unroll
        for (int m = 0; m < N2; ++m) accZ[kk][m] += c * un[m];
      }
    }
    // neighbour slabs across the x- and y- faces, through the z factors
#pragma unroll
    for (int kk = 0; kk < PL; ++kk)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        sx[kk][j] = T(0);
        sy[kk][j] = T(0);
      }
    const T* xn = x + (long long)nbe[0] * N3 + nsl[0];
    const T* yn = x + (long long)nbe[1] * N3 + nsl[1] * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      T vx[N], vy[N];
#pragma unroll
      for (int j = 0; j < N; ++j) vx[j] = xn[k * N2 + j * N];
      load_block<T, N>(yn + k * N2, vy);
#pragma unroll
      for (int kk = 0; kk < PL; ++kk)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          sx[kk][j] += zX[kk][k] * vx[j];
          sy[kk][j] += zY[kk][k] * vy[j];
        }
    }
  }

  // ---- x and y factors, one owned plane at a time ----
  // beta-dependent in-plane factors (position tables)
  auto tabm = [&](int comp, int pos, int r, int c) -> T { return p.tab[comp][(pos * N + r) * N + c]; };
#pragma unroll
  for (int kk = 0; kk < PL; ++kk) {
    T out[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m) out[m] = T(0);
    if constexpr (ADV) {
      // term x: y factor, then the derivative factor in x (+ face terms)
      T F1[N][N], F0[N][N];
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          F1[r][c] = AX0 == 1 ? tabm(0, sub[1], r, c) : (T)KT::M.a[r][c];
          F0[r][c] = -p.Kt[0] * (AX0 == 0 ? tabm(0, sub[0], r, c) : (T)KT::D.a[r][c]);
        }
      T t1[N2], s1[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        T sacc = T(0);
#pragma unroll
        for (int jj = 0; jj < N; ++jj) sacc += F1[j][jj] * sx[kk][jj];
        s1[j] = sacc;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = T(0);
#pragma unroll
          for (int jj = 0; jj < N; ++jj) a += F1[j][jj] * accX[kk][jj * N + i];
          t1[j * N + i] = a;
        }
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = out[j * N + i];
#pragma unroll
          for (int ii = 0; ii < N; ++ii) a += F0[i][ii] * t1[j * N + ii];
          out[j * N + i] = a;
        }
        out[j * N + N - 1] += cPown[0] * t1[j * N + N - 1] + cPnb[0] * s1[j];
        out[j * N] += cMown[0] * t1[j * N] + cMnb[0] * s1[j];
      }
    }
    if constexpr (MASS) {
      // Km (M (x) M) on the z-contracted planes
      T tm[N2];
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = T(0);
#pragma unroll
          for (int jj = 0; jj < N; ++jj) {
            T src;
            if constexpr (MSH) src = accX[kk][jj * N + i];
            else src = accM[kk][jj * N + i];
            a += (T)KT::M.a[j][jj] * src;
          }
          tm[j * N + i] = a;
        }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = T(0);
#pragma unroll
          for (int ii = 0; ii < N; ++ii) a += (T)KT::M.a[i][ii] * tm[j * N + ii];
          out[j * N + i] += p.Km * a;
        }
    }
    if constexpr (ADV) {
      // term y: x factor, then the derivative factor in y (+ face terms)
      T F0[N][N], F1[N][N];
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          F0[r][c] = AX1 == 0 ? tabm(1, sub[0], r, c) : (T)KT::M.a[r][c];
          F1[r][c] = -p.Kt[1] * (AX1 == 1 ? tabm(1, sub[1], r, c) : (T)KT::D.a[r][c]);
        }
      T t1[N2], s1[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T sacc = T(0);
#pragma unroll
        for (int ii = 0; ii < N; ++ii) sacc += F0[i][ii] * sy[kk][ii];
        s1[i] = sacc;
      }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = T(0);
#pragma unroll
          for (int ii = 0; ii < N; ++ii) a += F0[i][ii] * accY[kk][j * N + ii];
          t1[j * N + i] = a;
        }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = out[j * N + i];
#pragma unroll
          for (int jj = 0; jj < N; ++jj) a += F1[j][jj] * t1[jj * N + i];
          out[j * N + i] = a;
        }
#pragma unroll
      for (int i = 0; i < N; ++i) {
        out[(N - 1) * N + i] += cPown[1] * t1[(N - 1) * N + i] + cPnb[1] * s1[i];
        out[i] += cMown[1] * t1[i] + cMnb[1] * s1[i];
      }
    }
    if constexpr (ADV) {
      // term z: y factor, then x factor
      T F1[N][N], F0[N][N];
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          F1[r][c] = AX2 == 1 ? tabm(2, sub[1], r, c) : (T)KT::M.a[r][c];
          F0[r][c] = AX2 == 0 ? tabm(2, sub[0], r, c) : (T)KT::M.a[r][c];
        }
      T t1[N2];
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = T(0);
#pragma unroll
          for (int jj = 0; jj < N; ++jj) a += F1[j][jj] * accZ[kk][jj * N + i];
          t1[j * N + i] = a;
        }
#pragma unroll
      for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T a = out[j * N + i];
#pragma unroll
          for (int ii = 0; ii < N; ++ii) a += F0[i][ii] * t1[j * N + ii];
          out[j * N + i] = a;
        }
    }
    if (valid) {
      T* dst = y + (long long)e * N3 + (h * PL + kk) * N2;
      constexpr int VW = 16 / sizeof(T);
      if constexpr ((N2 * sizeof(T)) % 16 == 0) {
        Chunk<T, VW>* d = reinterpret_cast<Chunk<T, VW>*>(dst);
#pragma unroll
        for (int v = 0; v < N2 / VW; ++v) {
          Chunk<T, VW> c;
#pragma unroll
          for (int l = 0; l < VW; ++l) c.v[l] = out[v * VW + l];
          d[v] = c;
        }
      } else {
#pragma unroll
        for (int m = 0; m < N2; ++m) dst[m] = out[m];
      }
    }
  }
}

}  // namespace k3k
}  // namespace gd
