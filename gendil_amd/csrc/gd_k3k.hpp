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
  T Kc[3];         // the same for the cell terms (0: the face terms alone)
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

// Compiler-only pins: the values must be computed (in VGPRs) at this point,
// so arithmetic that produces them is not sunk past later loads.
template <int NN, typename T>
__device__ __forceinline__ void pin(T* a)
{
#pragma unroll
  for (int m = 0; m < NN; ++m) __asm__ volatile("" : "+v"(a[m]));
}

template <typename T, int NV>
struct alignas(16) Chunk {
  T v[NV];
};

// NV consecutive values aligned to their own size (NV * sizeof(T) <= 16)
template <typename T, int NV>
struct alignas(NV * sizeof(T)) Piece {
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

// Groups of G consecutive blocks per XCD turn (blocks are dispatched
// round-robin over the 8 XCDs): neighbouring rows of elements land on the
// same XCD, so face-neighbour reads hit its L2, while the 8 XCDs still stream
// through nearby memory at any moment (giving each XCD one contiguous eighth
// of the mesh instead concentrates its traffic and measured slower).  The
// tail that does not fill whole groups keeps the identity map.
template <int G>
__device__ __forceinline__ int group_remap(int pb, int nb)
{
  const int full = nb - nb % (8 * G);
  if (pb >= full) return pb;
  const int r = pb / 8, xcd = pb % 8;
  return (r / G) * 8 * G + xcd * G + r % G;
}

// Upwind neighbour of element e across an axis-d face (the element itself
// where there is none; its coefficient is then zero) and the node of the
// neighbour's slab: beta_d >= 0 takes the - neighbour's node N-1 slab, beta_d
// <= 0 the + neighbour's node 0 slab (the same rule as face_setup).
template <typename T, int N>
__device__ __forceinline__ int upwind_nb(const KParams<T>& p, int e, int d, int& node)
{
  const unsigned r = (unsigned)e / (unsigned)p.M0;
  const int s0 = e - (int)r * p.M0, s2 = (int)(r / (unsigned)p.M1), s1 = (int)r - s2 * p.M1;
  const int sa = d == 0 ? s0 : (d == 1 ? s1 : s2);
  const int Ma = d == 0 ? p.M0 : (d == 1 ? p.M1 : p.M2);
  const int st = d == 0 ? 1 : (d == 1 ? p.M0 : p.M0 * p.M1);
  const bool per = (p.periodic >> d) & 1;
  const bool neg = (p.negmask >> d) & 1;
  int nb;
  if (Ma == 1) nb = per ? e : -1;
  else if (!neg) nb = sa > 0 ? e - st : (per ? e + (Ma - 1) * st : -1);
  else nb = sa < Ma - 1 ? e + st : (per ? e - (Ma - 1) * st : -1);
  node = neg ? 0 : N - 1;
  return nb < 0 ? e : nb;
}

template <typename C>
__device__ __forceinline__ void store_nt(C* dst, const C& v)
{
  // streaming output: not kept in L2 (it is not read again by this apply)
#pragma unroll
  for (int l = 0; l < (int)(sizeof(v.v) / sizeof(v.v[0])); ++l) __builtin_nontemporal_store(v.v[l], &dst->v[l]);
}

// Face coefficients and upwind-side neighbours of one element.
// beta_d >= 0: own trace on the + face (node N-1), the - neighbour's node N-1
// slab on the - face (node 0).  beta_d <= 0: own trace on the - face, the +
// neighbour's node 0 slab on the + face.  cP*/cM* act on output node N-1 / 0
// along d; nbe = upwind neighbour (the element itself where there is none, its
// coefficient then being zero), nsl = node of the neighbour slab.
template <typename T>
struct FaceCoefs {
  T cown[3], cnb[3];  // own-trace and neighbour-slab coefficients per axis
  int nbe[3], nsl[3];
  int negmask;        // bit d: beta_d <= 0 (own trace on the - face)
  __device__ __forceinline__ T cPown(int d) const { return ((negmask >> d) & 1) ? T(0) : cown[d]; }
  __device__ __forceinline__ T cMown(int d) const { return ((negmask >> d) & 1) ? cown[d] : T(0); }
  __device__ __forceinline__ T cPnb(int d) const { return ((negmask >> d) & 1) ? cnb[d] : T(0); }
  __device__ __forceinline__ T cMnb(int d) const { return ((negmask >> d) & 1) ? T(0) : cnb[d]; }
};

template <typename T, int N, bool ADV, int SIG>
__device__ __forceinline__ void face_setup(const KParams<T>& p, int e, int sub[3], FaceCoefs<T>& fc)
{
  constexpr int AX0 = ksig_axis(SIG, 0), AX1 = ksig_axis(SIG, 1), AX2 = ksig_axis(SIG, 2);
  {
    const unsigned r = (unsigned)e / (unsigned)p.M0;
    sub[0] = e - (int)r * p.M0;
    sub[2] = (int)(r / (unsigned)p.M1);
    sub[1] = (int)r - sub[2] * p.M1;
  }
  fc.negmask = p.negmask;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    fc.cown[d] = fc.cnb[d] = T(0);
    fc.nbe[d] = e;
    fc.nsl[d] = N - 1;
  }
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
        fc.cown[d] = own_ok ? p.Kt[d] * f1 : T(0);
        fc.cnb[d] = nlo >= 0 ? -p.Kt[d] * f0 : T(0);
        fc.nbe[d] = nlo >= 0 ? nlo : e;
        fc.nsl[d] = N - 1;
      } else {
        const bool own_ok = nlo >= 0 || p.bc_zero;
        fc.cown[d] = own_ok ? -p.Kt[d] * f0 : T(0);
        fc.cnb[d] = nhi >= 0 ? p.Kt[d] * f1 : T(0);
        fc.nbe[d] = nhi >= 0 ? nhi : e;
        fc.nsl[d] = 0;
      }
    });
  }
}

// In-plane part of one owned output plane: the y and x factors of every term
// applied to its z-contracted planes (accX/accY/accZ/accM), the neighbour slab
// terms (sx, sy: the x- and y-neighbour slabs through the z factors), and the
// own-face terms folded in.  out is overwritten.
template <typename T, int N, bool ADV, bool MASS, int SIG>
__device__ __forceinline__ void inplane(const KParams<T>& p, const int sub[3], const FaceCoefs<T>& fc, const T* accX,
                                        const T* accY, const T* accZ, const T* accM, const T* sx, const T* sy,
                                        T* out)
{
  constexpr int N2 = N * N;
  constexpr int AX0 = ksig_axis(SIG, 0), AX1 = ksig_axis(SIG, 1), AX2 = ksig_axis(SIG, 2);
  constexpr bool MSH = ADV && MASS && AX0 != 2;
  using KT = KTab<N>;
  (void)accM;
  (void)sx;
  (void)sy;
  (void)AX2;
  auto tabm = [&](int comp, int pos, int r, int c) -> T { return p.tab[comp][(pos * N + r) * N + c]; };
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
        F0[r][c] = -p.Kc[0] * (AX0 == 0 ? tabm(0, sub[0], r, c) : (T)KT::D.a[r][c]);
      }
    T t1[N2], s1[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T sacc = T(0);
#pragma unroll
      for (int jj = 0; jj < N; ++jj) sacc += F1[j][jj] * sx[jj];
      s1[j] = sacc;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T a = T(0);
#pragma unroll
        for (int jj = 0; jj < N; ++jj) a += F1[j][jj] * accX[jj * N + i];
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
      out[j * N + N - 1] += fc.cPown(0) * t1[j * N + N - 1] + fc.cPnb(0) * s1[j];
      out[j * N] += fc.cMown(0) * t1[j * N] + fc.cMnb(0) * s1[j];
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
          if constexpr (MSH) src = accX[jj * N + i];
          else src = accM[jj * N + i];
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
        F1[r][c] = -p.Kc[1] * (AX1 == 1 ? tabm(1, sub[1], r, c) : (T)KT::D.a[r][c]);
      }
    T t1[N2], s1[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T sacc = T(0);
#pragma unroll
      for (int ii = 0; ii < N; ++ii) sacc += F0[i][ii] * sy[ii];
      s1[i] = sacc;
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T a = T(0);
#pragma unroll
        for (int ii = 0; ii < N; ++ii) a += F0[i][ii] * accY[j * N + ii];
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
      out[(N - 1) * N + i] += fc.cPown(1) * t1[(N - 1) * N + i] + fc.cPnb(1) * s1[i];
      out[i] += fc.cMown(1) * t1[i] + fc.cMnb(1) * s1[i];
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
        for (int jj = 0; jj < N; ++jj) a += F1[j][jj] * accZ[jj * N + i];
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
  FaceCoefs<T> fc;
  face_setup<T, N, ADV, SIG>(p, e, sub, fc);
  const T cPown[3] = {fc.cPown(0), fc.cPown(1), fc.cPown(2)};
  const T cMown[3] = {fc.cMown(0), fc.cMown(1), fc.cMown(2)};
  const T cPnb[3] = {fc.cPnb(0), fc.cPnb(1), fc.cPnb(2)};
  const T cMnb[3] = {fc.cMnb(0), fc.cMnb(1), fc.cMnb(2)};
  int(&nbe)[3] = fc.nbe;
  int(&nsl)[3] = fc.nsl;

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
        zZ[kk][j] = -p.Kc[2] * dz + ((row == N - 1 && j == N - 1) ? cPown[2] : T(0)) +
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
#pragma unroll
  for (int kk = 0; kk < PL; ++kk) {
    T out[N2];
    inplane<T, N, ADV, MASS, SIG>(p, sub, fc, accX[kk], accY[kk], accZ[kk], accM[(MASS && !MSH) ? kk : 0], sx[kk],
                                  sy[kk], out);
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

// Per-term in-plane pieces of the LDS-staged kernel, streamed a row (or
// column) at a time so only one z-contracted plane and the output plane are
// live.  Same arithmetic as inplane(); the one beta-dependent factor of a
// term (component d's table at the element's position along the axis beta_d
// depends on) is read from ft, a row-major N x N copy staged in LDS.
template <typename T, int N>
__device__ __forceinline__ T kmat(bool tabbed, const T* ft, const Mat<N>& K, int r, int c)
{
  return tabbed ? ft[r * N + c] : (T)K.a[r][c];
}

// term x: y factor, then the derivative factor in x, own/neighbour faces in x
template <typename T, int N, int SIG>
__device__ __forceinline__ void term_x(const KParams<T>& p, const FaceCoefs<T>& fc, const T* ft, const T* acc,
                                       const T* sx, T* out)
{
  constexpr int AX0 = ksig_axis(SIG, 0);
  using KT = KTab<N>;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T t1[N], s1 = T(0);
#pragma unroll
    for (int i = 0; i < N; ++i) t1[i] = T(0);
#pragma unroll
    for (int jj = 0; jj < N; ++jj) {
      const T f = kmat<T, N>(AX0 == 1, ft, KT::M, j, jj);
      s1 += f * sx[jj];
#pragma unroll
      for (int i = 0; i < N; ++i) t1[i] += f * acc[jj * N + i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T a = T(0);
#pragma unroll
      for (int ii = 0; ii < N; ++ii) a += kmat<T, N>(AX0 == 0, ft, KT::D, i, ii) * t1[ii];
      out[j * N + i] -= p.Kc[0] * a;
    }
    out[j * N + N - 1] += fc.cPown(0) * t1[N - 1] + fc.cPnb(0) * s1;
    out[j * N] += fc.cMown(0) * t1[0] + fc.cMnb(0) * s1;
  }
}

// term y: x factor, then the derivative factor in y, own/neighbour faces in y
template <typename T, int N, int SIG>
__device__ __forceinline__ void term_y(const KParams<T>& p, const FaceCoefs<T>& fc, const T* ft, const T* acc,
                                       const T* sy, T* out)
{
  constexpr int AX1 = ksig_axis(SIG, 1);
  using KT = KTab<N>;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T t1[N], s1 = T(0);
#pragma unroll
    for (int j = 0; j < N; ++j) t1[j] = T(0);
#pragma unroll
    for (int ii = 0; ii < N; ++ii) {
      const T f = kmat<T, N>(AX1 == 0, ft, KT::M, i, ii);
      s1 += f * sy[ii];
#pragma unroll
      for (int j = 0; j < N; ++j) t1[j] += f * acc[j * N + ii];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T a = T(0);
#pragma unroll
      for (int jj = 0; jj < N; ++jj) a += kmat<T, N>(AX1 == 1, ft, KT::D, j, jj) * t1[jj];
      out[j * N + i] -= p.Kc[1] * a;
    }
    out[(N - 1) * N + i] += fc.cPown(1) * t1[N - 1] + fc.cPnb(1) * s1;
    out[i] += fc.cMown(1) * t1[0] + fc.cMnb(1) * s1;
  }
}

// term z (z factor already applied): y factor, then x factor; with
// COMP < 0 the plain mass factors scaled by Km
template <typename T, int N, int SIG, int COMP>
__device__ __forceinline__ void term_plane(const KParams<T>& p, const T* ft, const T* acc, T* out)
{
  constexpr int AX = COMP < 0 ? -1 : ksig_axis(SIG, COMP);
  using KT = KTab<N>;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T t1[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t1[i] = T(0);
#pragma unroll
    for (int jj = 0; jj < N; ++jj) {
      const T f = kmat<T, N>(AX == 1, ft, KT::M, j, jj);
#pragma unroll
      for (int i = 0; i < N; ++i) t1[i] += f * acc[jj * N + i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T a = T(0);
#pragma unroll
      for (int ii = 0; ii < N; ++ii) a += kmat<T, N>(AX == 0, ft, KT::M, i, ii) * t1[ii];
      out[j * N + i] += COMP < 0 ? p.Km * a : a;
    }
  }
}

// LDS-staged form of the same operator, for planes of whole 16-byte chunks
// (N = 2, 4).  A block owns EPB consecutive elements, N lanes per element,
// lane h producing output z-plane h.
//  * Prologue (every global read of the block, issued before one barrier):
//    the block's elements with fully coalesced 16-byte loads (consecutive
//    lanes, consecutive chunks) into LDS, padded by one chunk per element so
//    the lanes of a wave, which read the same offset of 64/N elements, spread
//    over the banks; row h of the element's x-, y- and z-neighbour slabs and
//    of each beta-dependent in-plane factor (component d's table at the
//    element's position), staged in LDS for the element's other lanes; the
//    lane's row of a beta-dependent z factor, kept in registers.
//  * Compute from LDS: the z factors of all terms in one stream over the
//    element's planes (or, SEQ, one term at a time, three streams and a third
//    of the accumulator registers), then each term's in-plane factors.
//  * OUT_LDS stages the output planes through LDS so the stores are
//    coalesced too.
// The compiler-only fences and pins bound how far loads are hoisted and
// arithmetic sunk, which otherwise keeps every LDS value live at once.
template <typename T, int N, bool ADV, bool MASS, int SIG, int EPB, bool OUT_LDS, bool SEQ>
__global__ void __launch_bounds__(EPB * N, (SEQ || sizeof(T) == 4) ? 3 : 2)
    k3k_staged(KParams<T> p, const T* __restrict__ x, T* __restrict__ y)
{
  constexpr int NT = EPB * N;
  constexpr int N2 = N * N, N3 = N2 * N;
  constexpr int VW = 16 / sizeof(T);
  static_assert(N3 % VW == 0 && N2 % VW == 0, "element planes must be whole 16-byte chunks");
  constexpr int US = N3 + VW;  // element stride in LDS (values)
  constexpr int SS = N2 + VW;  // slab / table stride
  constexpr int CPE = N3 / VW;  // chunks per element
  constexpr int NCH = EPB * CPE;
  constexpr int NIT = (NCH + NT - 1) / NT;
  constexpr int AX[3] = {ksig_axis(SIG, 0), ksig_axis(SIG, 1), ksig_axis(SIG, 2)};
  constexpr bool MSH = ADV && MASS && AX[0] != 2;
  // component d's beta-dependent factor: in-plane (staged table) or z (row)
  constexpr bool TIN[3] = {ADV && (AX[0] == 0 || AX[0] == 1), ADV && (AX[1] == 0 || AX[1] == 1),
                           ADV && (AX[2] == 0 || AX[2] == 1)};
  constexpr bool TZ[3] = {ADV && AX[0] == 2, ADV && AX[1] == 2, ADV && AX[2] == 2};
  constexpr int NTB = (TIN[0] ? 1 : 0) + (TIN[1] ? 1 : 0) + (TIN[2] ? 1 : 0);
  constexpr int TBI[3] = {0, TIN[0] ? 1 : 0, (TIN[0] ? 1 : 0) + (TIN[1] ? 1 : 0)};  // slot of component d
  using KT = KTab<N>;
  using C = Chunk<T, VW>;
  __shared__ C Uc[EPB * US / VW];
  __shared__ T SL[ADV ? 3 * EPB * SS : 1];          // x-, y-, z-neighbour slabs
  __shared__ T TB[NTB > 0 ? NTB * EPB * SS : 1];    // staged beta-dependent in-plane factors

  const int tid = (int)threadIdx.x;
  const int h = tid % N;
  const int el = tid / N;

  const int blk = group_remap<16>((int)blockIdx.x, p.num_blocks);
  const int e0 = blk * EPB;
  const int nvalid = min(EPB, p.E - e0);
  const bool valid = el < nvalid;
  const int e = valid ? e0 + el : p.E - 1;
  const int last = nvalid * CPE - 1;

  int sub[3];
  FaceCoefs<T> fc;
  face_setup<T, N, ADV, SIG>(p, e, sub, fc);
  // x neighbour inside the block: its slab is read from LDS
  const int xl = fc.nbe[0] - e0;
  const bool xin = xl >= 0 && xl < nvalid;

  // ---- prologue: every global read of the block, then one barrier ----
  // own elements: coalesced 16-byte chunks (consecutive lanes, consecutive
  // chunks); the chunk index is clamped to the block's last valid chunk, so
  // every address is in bounds and only the LDS stores are guarded
  C buf[NIT];
  {
    const C* src = reinterpret_cast<const C*>(x) + (long long)e0 * CPE;
#pragma unroll
    for (int i = 0; i < NIT; ++i) buf[i] = src[min(tid + i * NT, last)];
  }
  // y- and z-neighbour slabs, cooperatively: SVW-value pieces of the slabs,
  // consecutive lanes on consecutive pieces of one element's slab
  constexpr int SVW = (N % VW == 0) ? VW : 1;
  constexpr int PPS = N2 / SVW;  // pieces per slab
  constexpr int NSI = (EPB * PPS + NT - 1) / NT;
  using SC = Piece<T, SVW>;
  SC by[ADV ? NSI : 1], bz[ADV ? NSI : 1];
  if constexpr (ADV) {
#pragma unroll
    for (int i = 0; i < NSI; ++i) {
      const int q = min(tid + i * NT, EPB * PPS - 1);
      const int ee = min(e0 + q / PPS, p.E - 1), w = q % PPS;
      int ny, nz;
      const int by_e = upwind_nb<T, N>(p, ee, 1, ny), bz_e = upwind_nb<T, N>(p, ee, 2, nz);
      const int m = w * SVW;  // slab value index: y slab (iz, ix), z slab (iy, ix)
      by[i] = *reinterpret_cast<const SC*>(x + (long long)by_e * N3 + (m / N) * N2 + ny * N + m % N);
      bz[i] = *reinterpret_cast<const SC*>(x + (long long)bz_e * N3 + nz * N2 + m);
    }
  }
  // x-neighbour slab row h where the neighbour is outside the block, and the
  // rows of the beta-dependent factors
  T vx[N], vt[3][N];
  T zt[N];  // this lane's row of a beta-dependent z factor
#pragma unroll
  for (int j = 0; j < N; ++j) zt[j] = T(0);
  if constexpr (ADV) {
    if (!xin) {
      const T* xn = x + (long long)fc.nbe[0] * N3 + h * N2 + fc.nsl[0];
#pragma unroll
      for (int j = 0; j < N; ++j) vx[j] = xn[j * N];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (TIN[d] || TZ[d]) load_block<T, N>(p.tab[d] + ((long long)sub[TZ[d] ? 2 : AX[d]] * N + h) * N, vt[d]);
  }
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int c = tid + i * NT;
    if ((NCH % NT == 0 || c < NCH) && c <= last) {
      const int ec = c / CPE;
      Uc[(ec * US) / VW + (c - ec * CPE)] = buf[i];
    }
  }
  if constexpr (ADV) {
#pragma unroll
    for (int i = 0; i < NSI; ++i) {
      const int q = tid + i * NT;
      if (EPB * PPS % NT == 0 || q < EPB * PPS) {
        *reinterpret_cast<SC*>(&SL[(1 * EPB + q / PPS) * SS + (q % PPS) * SVW]) = by[i];
        *reinterpret_cast<SC*>(&SL[(2 * EPB + q / PPS) * SS + (q % PPS) * SVW]) = bz[i];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (TIN[d]) {
#pragma unroll
        for (int j = 0; j < N; ++j) TB[(TBI[d] * EPB + el) * SS + h * N + j] = vt[d][j];
      }
      if (TZ[d]) {
#pragma unroll
        for (int j = 0; j < N; ++j) zt[j] = vt[d][j];
      }
    }
  }
  __syncthreads();
  if constexpr (ADV) {
    // x-neighbour slab row h: from the block's LDS copy where it is inside
    if (xin) {
      const T* un = reinterpret_cast<const T*>(Uc) + xl * US + h * N2 + fc.nsl[0];
#pragma unroll
      for (int j = 0; j < N; ++j) vx[j] = un[j * N];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) SL[(0 * EPB + el) * SS + h * N + j] = vx[j];
    __syncthreads();
  }
  {
  // this lane's rows of the z factors
  auto coefX = [&](int j) -> T {
    if constexpr (TZ[0]) return zt[j];
    else return lrow<N, N, 1, T>(KT::M, h, 0, j);
  };
  auto coefY = [&](int j) -> T {
    if constexpr (TZ[1]) return zt[j];
    else return lrow<N, N, 1, T>(KT::M, h, 0, j);
  };
  auto coefZ = [&](int j) -> T {
    T dz;
    if constexpr (TZ[2]) dz = zt[j];
    else dz = lrow<N, N, 1, T>(KT::D, h, 0, j);
    return -p.Kc[2] * dz + ((h == N - 1 && j == N - 1) ? fc.cPown(2) : T(0)) +
           ((h == 0 && j == 0) ? fc.cMown(2) : T(0));
  };
  auto coefM = [&](int j) -> T { return lrow<N, N, 1, T>(KT::M, h, 0, j); };
  const C* ue = Uc + (el * US) / VW;
  const T* sxe = SL + (0 * EPB + el) * SS;
  const T* sye = SL + (1 * EPB + el) * SS;
  const T* sze = SL + (2 * EPB + el) * SS;
  auto ftab = [&](int d) -> const T* { return TB + (TBI[d] * EPB + el) * SS; };
  // the neighbour slabs of x and y through this lane's z-factor row
  auto slab_row = [&](auto coef, const T* sl, T* s) {
#pragma unroll
    for (int j = 0; j < N; ++j) s[j] = T(0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const T c = coef(k);
#pragma unroll
      for (int j = 0; j < N; ++j) s[j] += c * sl[k * N + j];
    }
  };
  // upwind neighbour across a z face: its slab enters output plane 0
  // (beta_z >= 0) or N-1 (beta_z <= 0); the other lanes add it times 0
  auto zslab = [&](T* acc) {
    const T czn = (h == 0 ? fc.cMnb(2) : T(0)) + (h == N - 1 ? fc.cPnb(2) : T(0));
#pragma unroll
    for (int m = 0; m < N2; ++m) acc[m] += czn * sze[m];
  };

  T out[N2];
#pragma unroll
  for (int m = 0; m < N2; ++m) out[m] = T(0);
  if constexpr (SEQ) {
    auto zpass = [&](auto coef, T* acc) {
      // compiler-only fence (the LDS pointer is an operand, so the fence may
      // touch the block's LDS): each pass re-reads the element from LDS
      __asm__ volatile("" : : "v"(ue) : "memory");
#pragma unroll
      for (int m = 0; m < N2; ++m) acc[m] = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const T cz = coef(j);
        if (j > 0) {  // one plane in flight at a time
          pin<N2>(acc);
          __asm__ volatile("" : : "v"(ue) : "memory");
        }
#pragma unroll
        for (int v = 0; v < N2 / VW; ++v) {
          const C c = ue[j * (N2 / VW) + v];
#pragma unroll
          for (int l = 0; l < VW; ++l) acc[v * VW + l] += cz * c.v[l];
        }
      }
    };
    if constexpr (ADV || MSH) {
      T acc[N2];
      zpass(coefX, acc);
      if constexpr (ADV) {
        T sx[N];
        slab_row(coefX, sxe, sx);
        term_x<T, N, SIG>(p, fc, ftab(0), acc, sx, out);
      }
      if constexpr (MSH) term_plane<T, N, SIG, -1>(p, nullptr, acc, out);
      pin<N2>(out);
    }
    if constexpr (ADV) {
      T acc[N2];
      zpass(coefY, acc);
      T sy[N];
      slab_row(coefY, sye, sy);
      term_y<T, N, SIG>(p, fc, ftab(1), acc, sy, out);
      pin<N2>(out);
      zpass(coefZ, acc);
      zslab(acc);
      term_plane<T, N, SIG, 2>(p, ftab(2), acc, out);
      pin<N2>(out);
    }
    if constexpr (MASS && !MSH) {
      T acc[N2];
      zpass(coefM, acc);
      term_plane<T, N, SIG, -1>(p, nullptr, acc, out);
    }
  } else {
    // all terms' z factors in one stream over the element's planes
    T accX[N2], accY[N2], accZ[N2], accM[(MASS && !MSH) ? N2 : 1];
#pragma unroll
    for (int m = 0; m < N2; ++m) {
      accX[m] = accY[m] = accZ[m] = T(0);
      if constexpr (MASS && !MSH) accM[m] = T(0);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T cx = (ADV || MSH) ? coefX(j) : T(0);
      const T cy = ADV ? coefY(j) : T(0);
      const T cz = ADV ? coefZ(j) : T(0);
      const T cm = (MASS && !MSH) ? coefM(j) : T(0);
      if (j > 0) {  // one plane in flight at a time
        if constexpr (ADV || MSH) pin<N2>(accX);
        if constexpr (ADV) {
          pin<N2>(accY);
          pin<N2>(accZ);
        }
        if constexpr (MASS && !MSH) pin<N2>(accM);
        __asm__ volatile("" : : "v"(ue) : "memory");
      }
#pragma unroll
      for (int v = 0; v < N2 / VW; ++v) {
        const C c = ue[j * (N2 / VW) + v];
#pragma unroll
        for (int l = 0; l < VW; ++l) {
          const int m = v * VW + l;
          if constexpr (ADV || MSH) accX[m] += cx * c.v[l];
          if constexpr (ADV) {
            accY[m] += cy * c.v[l];
            accZ[m] += cz * c.v[l];
          }
          if constexpr (MASS && !MSH) accM[m] += cm * c.v[l];
        }
      }
    }
    if constexpr (ADV) {
      T sx[N];
      slab_row(coefX, sxe, sx);
      term_x<T, N, SIG>(p, fc, ftab(0), accX, sx, out);
    }
    if constexpr (MSH) term_plane<T, N, SIG, -1>(p, nullptr, accX, out);
    if constexpr (MASS && !MSH) term_plane<T, N, SIG, -1>(p, nullptr, accM, out);
    pin<N2>(out);
    if constexpr (ADV) {
      T sy[N];
      slab_row(coefY, sye, sy);
      term_y<T, N, SIG>(p, fc, ftab(1), accY, sy, out);
      pin<N2>(out);
      zslab(accZ);
      term_plane<T, N, SIG, 2>(p, ftab(2), accZ, out);
    }
  }

  if constexpr (OUT_LDS) {
    __syncthreads();  // every lane has read its element from U
#pragma unroll
    for (int v = 0; v < N2 / VW; ++v) {
      C c;
#pragma unroll
      for (int l = 0; l < VW; ++l) c.v[l] = out[v * VW + l];
      Uc[(el * US) / VW + h * (N2 / VW) + v] = c;
    }
    __syncthreads();
    C* dst = reinterpret_cast<C*>(y) + (long long)e0 * CPE;
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int c = tid + i * NT;
      if ((NCH % NT == 0 || c < NCH) && c <= last) {
        const int ec = c / CPE;
        store_nt(&dst[c], Uc[(ec * US) / VW + (c - ec * CPE)]);
      }
    }
  } else if (valid) {
    C* d = reinterpret_cast<C*>(y + (long long)e * N3 + h * N2);
#pragma unroll
    for (int v = 0; v < N2 / VW; ++v) {
      C c;
#pragma unroll
      for (int l = 0; l < VW; ++l) c.v[l] = out[v * VW + l];
      store_nt(&d[v], c);
    }
  }
  }
}

}  // namespace k3k
}  // namespace gd
