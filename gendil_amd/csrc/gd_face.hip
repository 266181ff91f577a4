// Face-centric upwind advection: the flux of every face computed once and
// added to the minus element, subtracted from the plus element.
//
// Reference: AdvectionLocalFaceOperator / AdvectionNonconformingLocalFaceOperator
// (advectionoperator.hpp:478-548, 596-676), driven by AdvectionFaceOperator /
// AdvectionNonconformingFaceOperator (:1105-1180, :1210-1396) over
// CartesianInteriorLocalFaceConnectivity (cartesianglobalfaceconnectivity.hpp:15-95)
// or CartesianIntermeshFaceConnectivity (cartesianintermeshfaceconnectivity.hpp:36-180).
//
// Per face: both sides' traces (the element contracted along the normal with
// the side's basis at its face) are interpolated to the q^(D-1) Gauss points
// of the shared rule, the upwind flux is formed with the minus cell's geometry
// and field, and B^T of the flux is spread back along the normal on each side.
// GenDiL accumulates with atomics on the device; here the faces of one launch
// touch disjoint elements (interior faces of one axis are coloured by the
// parity of the minus cell's index along that axis), so every update is a
// plain read-add-write and the result does not depend on scheduling.
//
// Layout: a block takes FPB faces; per face four LDS buffers of S doubles
// (S = max(n_minus, n_plus, q)^(D-1)): ping-pong for the sweeps and the two
// interpolated traces (the minus one then holds the flux).

#include "gendil_amd.h"
#include "gd_internal.hpp"
#include "gd_k3k_api.hpp"

#include <algorithm>
#include <string>

namespace gd {
namespace face {

constexpr int kThreads = 256;
constexpr int kMaxS = 1024;             // largest face tensor, entries
constexpr int kLdsBytes = 40 * 1024;    // dynamic LDS per block
constexpr int kOwnerLdsBytes = 16 * 1024;  // one-launch kernel: more, smaller blocks per CU
constexpr int kMaxFpb = 64;             // faces per block
constexpr int kMaxQN = 1024;            // basis table staged in LDS by the one-launch kernel
constexpr int kMaxN = 32;

struct Side {
  const double* x;
  double* y;
  const double* B;   // [q][n] basis at the rule points
  const double* bf;  // [n] basis at this side's face (xi = 0 or 1)
  int64_t M[kMaxDim];
  int n;
  int64_t base, stride;  // cell index along the axis = base + stride * face row
  // tangential cell index = face grid index * mul + add (1, 0: the face grid
  // is this side's own grid)
  int mul[kMaxDim], add[kMaxDim];
  // subface matrices of a coarse side: [.][q][n] blocks, the one for axis d
  // at block bsel[d]; nullptr: B
  const double* Bmap;
  int bsel[kMaxDim];
};

struct Args {
  int D, axis, q, sign, fpb, S;
  int64_t G[kMaxDim];  // face grid, first index fastest
  int64_t nfaces;
  Side s[2];           // 0 minus, 1 plus
  Params<double> geo;  // minus plan: origin, h, detJ, field, rule tables
  int xi_minus;        // 0: the minus face is at xi = 0, 1: at xi = 1
  double wscale;       // subface weight factor prod 1 / r (1: conforming)
};

__device__ __forceinline__ int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// element index of side `sd` for face f; the tangential cell indices in t[]
__device__ __forceinline__ int64_t side_elem(const Args& a, const Side& sd, int64_t f, int* sub)
{
  int64_t r = f, lin = 0, str = 1;
  for (int d = 0; d < a.D; ++d) {
    const int64_t c = r % a.G[d];
    r /= a.G[d];
    const int64_t s = d == a.axis ? sd.base + sd.stride * c : c * sd.mul[d] + sd.add[d];
    if (sub) sub[d] = (int)s;
    lin += s * str;
    str *= sd.M[d];
  }
  return lin;
}

// element-local dof index of tangential multi-index t (over the D-1
// tangential axes, first fastest, extent n) with node i along the axis
__device__ __forceinline__ int dof_index(int D, int axis, int n, int t, int i)
{
  int idx = 0, str = 1, r = t;
  for (int d = 0; d < D; ++d) {
    int c;
    if (d == axis) {
      c = i;
    } else {
      c = r % n;
      r /= n;
    }
    idx += c * str;
    str *= n;
  }
  return idx;
}

// one 1D sweep over tangential axis k of every face of the block:
// in [outer][a][inner] -> out [outer][b][inner], out = M in, M(j, i) =
// TRANS ? B[i][j] : B[j][i] with B row-major [q][n]
template <bool TRANS>
__device__ __forceinline__ void sweep(const double* B0, int n, int a, int b, int inner, int outer, int nf, int S,
                                      const double* in,
                      double* out, const int* foff = nullptr)
{
  const int per = outer * b * inner;
  for (int t = threadIdx.x; t < nf * per; t += kThreads) {
    const int f = t / per;
    int r = t - f * per;
    const int o = r / (b * inner);
    r -= o * b * inner;
    const int j = r / inner, c = r - j * inner;
    const double* src = in + f * S + o * a * inner + c;
    const double* B = foff ? B0 + foff[f * kMaxDim] : B0;  // per-face matrix (subfaces)
    double s = 0.0;
    for (int i = 0; i < a; ++i) s += (TRANS ? B[i * n + j] : B[j * n + i]) * src[i * inner];
    out[f * S + o * b * inner + j * inner + c] = s;
  }
}

// Per-block index tables in LDS: each face's element on both sides, and per
// side the element-local offset of every tangential node (node 0 along the
// axis; node i adds i * n^axis)
struct Idx {
  int64_t elem[2][kMaxFpb];
  int off[2][kMaxS];
};

// the side's trace interpolated to the q^(D-1) points, into `dst`
__device__ __forceinline__ void interp_n(int D, int n, int q, int S, const double* B, const double* bf, const double* x,
                                         int axis, const int64_t* elem, const int* off, int nf, double* A, double* Bb,
                                         double* dst, const int* foff = nullptr)
{
  const int T = D - 1;
  const int nt = ipow(n, T), nd = nt * n, ax = ipow(n, axis);
  double* first = T == 0 ? dst : A;
  for (int t = threadIdx.x; t < nf * nt; t += kThreads) {
    const int f = t / nt, tt = t - f * nt;
    const double* u = x + elem[f] * nd + off[tt];
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const double w = bf[i];
      if (w != 0.0) s += w * u[i * ax];
    }
    first[f * S + tt] = s;
  }
  __syncthreads();
  const double* src = A;
  for (int k = 0; k < T; ++k) {
    double* out = k == T - 1 ? dst : (src == A ? Bb : A);
    sweep<false>(B, n, n, q, ipow(q, k), ipow(n, T - 1 - k), nf, S, src, out, foff ? foff + k : nullptr);
    __syncthreads();
    src = out;
  }
}

// Both sides of nf faces at once (one trace pass, one pass per sweep): faces
// [0, nf) are the minus side (bf0, em), [nf, 2 nf) the plus side (bf1, ep).
// The sweeps alternate between A and U so that the last lands in U.
__device__ __forceinline__ void interp_pair(int D, int n, int q, int S, const double* B, const double* bf0,
                                            const double* bf1, const double* x, int axis, const int64_t* em,
                                            const int64_t* ep, const int* off, int nf, double* A, double* U)
{
  const int T = D - 1;
  const int nt = ipow(n, T), nd = nt * n, ax = ipow(n, axis);
  double* first = (T % 2 == 1) ? A : U;
  for (int t = threadIdx.x; t < 2 * nf * nt; t += kThreads) {
    const int f = t / nt, tt = t - f * nt;
    const bool plus = f >= nf;
    const double* bf = plus ? bf1 : bf0;
    const double* u = x + (plus ? ep[f - nf] : em[f]) * nd + off[tt];
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const double w = bf[i];
      if (w != 0.0) s += w * u[i * ax];
    }
    first[f * S + tt] = s;
  }
  __syncthreads();
  double* src = first;
  for (int k = 0; k < T; ++k) {
    double* out = src == A ? U : A;
    sweep<false>(B, n, n, q, ipow(q, k), ipow(n, T - 1 - k), 2 * nf, S, src, out);
    __syncthreads();
    src = out;
  }
}

__device__ void interp_side(const Args& a, const Side& sd, int axis, const int64_t* elem, const int* off, int nf,
                            double* A, double* Bb, double* dst, const int* foff = nullptr)
{
  interp_n(a.D, sd.n, a.q, a.S, foff ? sd.Bmap : sd.B, sd.bf, sd.x, axis, elem, off, nf, A, Bb, dst, foff);
}

// B^T of the flux (in F) spread back along the normal: y += sgn * bf (x) B^T F
__device__ void spread_side(const Args& a, const Side& sd, const int64_t* elem, const int* off, int nf, const double* F,
                            double* A, double* Bb, double sgn, const int* foff = nullptr)
{
  const int T = a.D - 1, n = sd.n, q = a.q, S = a.S;
  const int nt = ipow(n, T), nd = nt * n, ax = ipow(n, a.axis);
  const double* src = F;
  for (int k = 0; k < T; ++k) {
    double* out = (k % 2 == 0) ? A : Bb;
    sweep<true>(foff ? sd.Bmap : sd.B, n, q, n, ipow(n, k), ipow(q, T - 1 - k), nf, S, src, out,
                foff ? foff + k : nullptr);
    __syncthreads();
    src = out;
  }
  for (int t = threadIdx.x; t < nf * nd; t += kThreads) {
    const int f = t / nd, r = t - f * nd;
    const int i = r / nt, tt = r - i * nt;
    const double w = sd.bf[i];
    if (w == 0.0) continue;
    double* yy = sd.y + elem[f] * nd + off[tt] + i * ax;
    *yy += sgn * (w * src[f * S + tt]);
  }
  __syncthreads();
}

__device__ __forceinline__ double beta_axis(const Params<double>& p, int c, const int* sub, const int* k, int Q, int D)
{
  if (p.beta_kind == GD_BETA_AFFINE) {
    double v = p.affc[c];
    for (int d = 0; d < D; ++d) {
      const double xi = k[d] < Q ? p.tab.xq[k[d]] : (k[d] == Q ? 0.0 : 1.0);
      const double X = (p.origin[d] + p.h[d] * (double)sub[d]) + p.h[d] * xi;
      v += p.affA[c][d] * X;
    }
    return v;
  }
  const FieldDev<double>& fd = p.beta[c];
  double sum = 0.0;
  for (int r = 0; r < fd.rank; ++r) {
    double t = fd.scale[r];
    for (int d = 0; d < D; ++d) {
      const double* tb = fd.tab[r][d];
      if (tb) t *= tb[sub[d] * (Q + 2) + k[d]];
    }
    sum += t;
  }
  return sum;
}

__global__ void __launch_bounds__(kThreads) k_faces(Args a)
{
  extern __shared__ double lds[];
  __shared__ Idx ix;
  __shared__ int sub0[kMaxFpb][kMaxDim];  // minus cell indices per axis
  const int64_t f0 = (int64_t)blockIdx.x * a.fpb;
  const int nf = (int)min((int64_t)a.fpb, a.nfaces - f0);
  const int S = a.S, T = a.D - 1, q = a.q;
  __shared__ int boff[kMaxFpb][kMaxDim];   // the minus side's subface matrix per tangential axis
  const bool mapped = a.s[0].Bmap != nullptr;
  for (int t = threadIdx.x; t < nf; t += kThreads) {
    ix.elem[0][t] = side_elem(a, a.s[0], f0 + t, sub0[t]);
    ix.elem[1][t] = side_elem(a, a.s[1], f0 + t, nullptr);
    if (mapped) {
      int k = 0;
      for (int d = 0; d < a.D; ++d)
        if (d != a.axis) boff[t][k++] = a.s[0].bsel[d] * a.q * a.s[0].n;
    }
  }
  for (int s = 0; s < 2; ++s) {
    const int n = a.s[s].n, nt = ipow(n, T);
    for (int t = threadIdx.x; t < nt; t += kThreads) ix.off[s][t] = dof_index(a.D, a.axis, n, t, 0);
  }
  __syncthreads();
  double* A = lds;
  double* Bb = lds + a.fpb * S;
  double* Um = lds + 2 * a.fpb * S;
  double* Up = lds + 3 * a.fpb * S;
  interp_side(a, a.s[0], a.axis, ix.elem[0], ix.off[0], nf, A, Bb, Um, mapped ? &boff[0][0] : nullptr);
  interp_side(a, a.s[1], a.axis, ix.elem[1], ix.off[1], nf, A, Bb, Up);
  // upwind flux at the face points of the minus cell (:512-540)
  const int nq = ipow(q, T);
  const Params<double>& p = a.geo;
  const double scale = (double)a.sign / (double)p.h[a.axis];
  for (int t = threadIdx.x; t < nf * nq; t += kThreads) {
    const int f = t / nq;
    int r = t - f * nq;
    int k[kMaxDim];
    double w = 1.0;
    for (int d = 0; d < a.D; ++d) {
      if (d == a.axis) {
        k[d] = q + a.xi_minus;
      } else {
        k[d] = r % q;
        r /= q;
        w *= p.tab.w[k[d]];
      }
    }
    const double dot = beta_axis(p, a.axis, sub0[f], k, q, a.D) * scale;
    const int at = f * S + (t - f * nq);
    const double up = dot > 0.0 ? dot * Um[at] : dot * Up[at];
    Um[at] = (w * a.wscale) * p.detJ * up;
  }
  __syncthreads();
  spread_side(a, a.s[0], ix.elem[0], ix.off[0], nf, Um, A, Bb, 1.0, mapped ? &boff[0][0] : nullptr);
  spread_side(a, a.s[1], ix.elem[1], ix.off[1], nf, Um, A, Bb, -1.0);
}

// Interior faces with one block per run of FPB consecutive elements: the
// block visits each of its elements' 2D faces (axis by axis, lower face then
// upper face), forms that face's flux exactly as above (minus = the upper
// cell, its geometry and field) and adds only its own share, + for the lower
// face (the element is the minus side) and - for the upper face (the element
// is the plus side), into an LDS accumulator; the block ends with one
// coalesced y += acc.  Each flux is formed twice, once per side, with the
// same arithmetic, so both sides see the same value; y is read and written
// once per apply instead of once per axis and colour.
// DD, NN, QQ > 0 fix the dimension, nodes and points at compile time (the
// loops unroll and the index arithmetic folds); 0 reads them from `a`.
constexpr int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }

template <int DD, int NN, int QQ>
__global__ void __launch_bounds__(kThreads) k_faces_owner(Args a)
{
  // static LDS sized for the instance (fixed sizes: exactly what they need)
  constexpr int kNT = DD > 0 ? cpow(NN, DD - 1) : kMaxS;
  constexpr int kQN = DD > 0 ? NN * QQ : kMaxQN;
  constexpr int kNB = DD > 0 ? NN : kMaxN;
  extern __shared__ double lds[];
  __shared__ int64_t elem[2][kMaxFpb];    // the current face's minus and plus cells
  __shared__ int offs[kNT];               // element-local offset of each tangential node
  __shared__ int sub[kMaxFpb][kMaxDim];   // owner cell indices
  __shared__ int subm[kMaxFpb][kMaxDim];  // the current face's minus cell
  __shared__ int ok[kMaxFpb];             // the face is interior
  __shared__ double sB[kQN];              // basis at the rule points, [q][n]
  __shared__ double sbf[2][kNB];          // basis at xi = 0 and xi = 1
  const Side& sd = a.s[0];
  const int64_t E = a.nfaces;  // elements
  // blocks go to the 8 XCDs round robin; give each XCD one contiguous run of
  // element blocks so that a block's y and z neighbours were read into the
  // same L2 (a bijection when the grid is a multiple of 8)
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int blk = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int64_t e0 = (int64_t)blk * a.fpb;
  const int nf = (int)min((int64_t)a.fpb, E - e0);
  const int D = DD > 0 ? DD : a.D;
  const int n = NN > 0 ? NN : sd.n;
  const int q = QQ > 0 ? QQ : a.q;
  const int S = DD > 0 ? cpow(NN > QQ ? NN : QQ, DD - 1) : a.S;
  const int T = D - 1;
  const int nt = ipow(n, T), nd = nt * n, nq = ipow(q, T);
  double* A = lds;                 // 2 fpb faces of scratch
  double* U = A + 2 * a.fpb * S;   // minus faces, then plus faces
  double* acc = U + 2 * a.fpb * S;
  for (int t = threadIdx.x; t < nf; t += kThreads) {
    int64_t r = e0 + t;
    for (int d = 0; d < D; ++d) {
      sub[t][d] = (int)(r % sd.M[d]);
      r /= sd.M[d];
    }
  }
  for (int t = threadIdx.x; t < nf * nd; t += kThreads) acc[t] = 0.0;
  for (int t = threadIdx.x; t < q * n; t += kThreads) sB[t] = sd.B[t];
  for (int t = threadIdx.x; t < n; t += kThreads) {
    sbf[0][t] = a.s[0].bf[t];
    sbf[1][t] = a.s[1].bf[t];
  }
  __shared__ int nzi[2][kNB];  // nodes with a nonzero basis value at xi = 0 / xi = 1
  __shared__ int nzc[2];
  if (threadIdx.x < 2) {
    const double* bf = a.s[threadIdx.x].bf;
    int c = 0;
    for (int i = 0; i < n; ++i)
      if (bf[i] != 0.0) nzi[threadIdx.x][c++] = i;
    nzc[threadIdx.x] = c;
  }
  const Params<double>& p = a.geo;
  int64_t str = 1;
  for (int axis = 0; axis < D; ++axis) {
    const double scale = -1.0 / (double)p.h[axis];  // minus normal -e_axis
    for (int slot = 0; slot < 2; ++slot) {
      __syncthreads();  // the previous face's tables are no longer read
      for (int t = threadIdx.x; t < nt; t += kThreads) offs[t] = dof_index(D, axis, n, t, 0);
      for (int t = threadIdx.x; t < nf; t += kThreads) {
        const int64_t e = e0 + t;
        const bool in = slot == 0 ? sub[t][axis] > 0 : sub[t][axis] < sd.M[axis] - 1;
        ok[t] = in;
        elem[0][t] = in && slot == 1 ? e + str : e;  // minus: the upper cell
        elem[1][t] = in && slot == 0 ? e - str : e;  // plus: the lower cell
        for (int d = 0; d < D; ++d) subm[t][d] = sub[t][d] + ((in && slot == 1 && d == axis) ? 1 : 0);
      }
      __syncthreads();
      interp_pair(D, n, q, S, sB, sbf[0], sbf[1], sd.x, axis, elem[0], elem[1], offs, nf, A, U);
      for (int t = threadIdx.x; t < nf * nq; t += kThreads) {
        const int f = t / nq;
        const int at = f * S + (t - f * nq);
        if (!ok[f]) {
          U[at] = 0.0;
          continue;
        }
        int r = t - f * nq;
        int k[kMaxDim];
        double w = 1.0;
        for (int d = 0; d < D; ++d) {
          if (d == axis) {
            k[d] = q;  // xi = 0 of the minus cell
          } else {
            k[d] = r % q;
            r /= q;
            w *= p.tab.w[k[d]];
          }
        }
        const double dot = beta_axis(p, axis, subm[f], k, q, D) * scale;
        const double up = dot > 0.0 ? dot * U[at] : dot * U[nf * S + at];
        U[at] = w * p.detJ * up;
      }
      __syncthreads();
      // the owner's share: + b0 (x) B^T F on its lower face, - b1 (x) B^T F on its upper face
      const double* ownbf = sbf[slot];
      const double sgn = slot == 0 ? 1.0 : -1.0;
      const int ax = ipow(n, axis);
      const double* src = U;
      for (int kk = 0; kk < T; ++kk) {
        double* out = src == A ? U : A;
        sweep<true>(sB, n, q, n, ipow(n, kk), ipow(q, T - 1 - kk), nf, S, src, out);
        __syncthreads();
        src = out;
      }
      // only the nodes with a nonzero basis value at the face (one for GLL)
      const int nz = nzc[slot];
      for (int t = threadIdx.x; t < nf * nz * nt; t += kThreads) {
        const int f = t / (nz * nt), r = t - f * (nz * nt);
        const int j = r / nt, tt = r - j * nt;
        if (!ok[f]) continue;
        const int i = nzi[slot][j];
        acc[f * nd + offs[tt] + i * ax] += sgn * (ownbf[i] * src[f * S + tt]);
      }
    }
    str *= sd.M[axis];
  }
  __syncthreads();
  double* y = sd.y + e0 * nd;
  for (int t = threadIdx.x; t < nf * nd; t += kThreads) y[t] += acc[t];
}

using OwnerKernel = void (*)(Args);

// faces per block of the one-launch kernel for (D, n, q): its LDS holds four
// face tensors of q^(D-1) values and one element of n^D per face
constexpr long long owner_fpb(int D, int n, int q)
{
  long long S = 1, nd = 1;
  for (int d = 0; d < D; ++d) {
    nd *= n;
    if (d > 0) S *= q;
  }
  const long long f = kOwnerLdsBytes / ((4 * S + nd) * (long long)sizeof(double));
  return f < kMaxFpb ? f : kMaxFpb;
}

OwnerKernel owner_kernel(int D, int n, int q)
{
  // an instance is compiled only where the one-launch path can select it
  // (at least 8 faces per block, as gd_adv_face_apply_f64 requires)
#define GD_FACE_OWNER(d, nn, qq)                                                      \
  static_assert(owner_fpb(d, nn, qq) >= 8, "face instance the one launch never selects"); \
  if (D == d && n == nn && q == qq) return k_faces_owner<d, nn, qq>;
  GD_FACE_OWNER(3, 2, 2) GD_FACE_OWNER(3, 2, 3) GD_FACE_OWNER(3, 3, 3) GD_FACE_OWNER(3, 3, 4)
  GD_FACE_OWNER(3, 4, 4) GD_FACE_OWNER(3, 4, 5) GD_FACE_OWNER(2, 3, 3)
  GD_FACE_OWNER(2, 4, 4) GD_FACE_OWNER(6, 2, 2)
#undef GD_FACE_OWNER
  return k_faces_owner<0, 0, 0>;
}

int side_setup(const gd_plan* pl, int q, int xi, const double* x, double* y, Side& s)
{
  const int n = pl->desc.order + 1;
  const TableSet* ts = find_tables(pl, n, q);
  if (!ts) return fail(GD_ERR_UNSUPPORTED, "face operator: plan has no tables for the shared rule");
  s.x = x;
  s.y = y;
  s.B = ts->t64.B;
  s.bf = xi ? ts->t64.b1 : ts->t64.b0;
  for (int d = 0; d < kMaxDim; ++d) {
    s.M[d] = d < pl->dim ? pl->desc.sizes[d] : 1;
    s.mul[d] = 1;
    s.add[d] = 0;
    s.bsel[d] = 0;
  }
  s.Bmap = nullptr;
  s.n = n;
  return GD_OK;
}

int check_plan(const gd_plan* pl, const char* what)
{
  if (!pl) return fail(GD_ERR_INVALID, std::string(what) + ": null plan");
  if (pl->padapt) return fail(GD_ERR_UNSUPPORTED, std::string(what) + ": p-adaptive plans are not supported");
  return GD_OK;
}

int launch(const Args& args, hipStream_t st)
{
  if (args.nfaces <= 0) return GD_OK;
  // a small face set is spread over the device (about four blocks per CU of
  // 256 CUs) instead of filling each block's LDS: 4,096 interface faces at
  // 51 faces per block would run on 81 blocks
  constexpr long long kSpreadBlocks = 1024;
  Args a = args;
  const long long spread = (a.nfaces + kSpreadBlocks - 1) / kSpreadBlocks;
  if (spread < a.fpb) a.fpb = (int)std::max<long long>(1, spread);
  const int blocks = (int)((a.nfaces + a.fpb - 1) / a.fpb);
  const size_t lds = (size_t)4 * a.fpb * a.S * sizeof(double);
  hipLaunchKernelGGL(k_faces, dim3((unsigned)blocks), dim3(kThreads), lds, st, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? GD_OK : hip_fail(e, "face kernel launch");
}

int common_args(const gd_plan* m, const gd_plan* p, int axis, int sign, Args& a, bool refined = false)
{
  const int D = m->dim;
  if (p->dim != D) return fail(GD_ERR_INVALID, "face operator: plans differ in dimension");
  if (!m->beta_set) return fail(GD_ERR_INVALID, "face operator: advection field not set on the minus plan");
  if (m->beta_kind == GD_BETA_POINTWISE)
    return fail(GD_ERR_UNSUPPORTED, "face operator: needs a separable or affine advection field");
  const int q = m->desc.num_quad;
  if (p->desc.num_quad != q) return fail(GD_ERR_INVALID, "face operator: both plans must use the same rule");
  for (int d = 0; d < D && !refined; ++d)
    if (d != axis && m->desc.sizes[d] != p->desc.sizes[d])
      return fail(GD_ERR_INVALID, "face operator: tangential mesh sizes differ");
  const int nmax = std::max({m->desc.order + 1, p->desc.order + 1, q});
  long long S = 1;
  for (int d = 0; d < D - 1; ++d) S *= nmax;
  if (S > kMaxS) return fail(GD_ERR_UNSUPPORTED, "face operator: face tensor larger than " + std::to_string(kMaxS));
  a = Args{};
  a.D = D;
  a.axis = axis;
  a.q = q;
  a.sign = sign;
  a.S = (int)S;
  a.fpb = std::max(1, std::min(kMaxFpb, (int)(kLdsBytes / (4 * S * (long long)sizeof(double)))));
  a.xi_minus = sign > 0 ? 1 : 0;
  a.wscale = 1.0;
  const TableSet* ts = find_tables(m, m->desc.order + 1, q);
  if (!ts) return fail(GD_ERR_UNSUPPORTED, "face operator: plan has no tables for its rule");
  a.geo = make_params<double>(m, ts, 0.0, 1.0);
  return GD_OK;
}

// true when [a, a + na) and [b, b + nb) share storage
bool overlap(const double* a, int64_t na, const double* b, int64_t nb)
{
  const uintptr_t a0 = reinterpret_cast<uintptr_t>(a), b0 = reinterpret_cast<uintptr_t>(b);
  return a0 < b0 + (uintptr_t)nb * sizeof(double) && b0 < a0 + (uintptr_t)na * sizeof(double);
}

// x is read while y is added into (in several launches for the coloured and
// h-refined paths), so no input side may share storage with an output side
bool sides_overlap(const gd_plan* m, const gd_plan* p, const double* xm, const double* xp, const double* ym,
                   const double* yp)
{
  return overlap(xm, m->ndofs, ym, m->ndofs) || overlap(xm, m->ndofs, yp, p->ndofs) ||
         overlap(xp, p->ndofs, ym, m->ndofs) || overlap(xp, p->ndofs, yp, p->ndofs);
}

}  // namespace face
}  // namespace gd

extern "C" int gd_adv_face_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream)
{
  using namespace gd::face;
  if (int rc = check_plan(plan, "gd_adv_face_apply_f64")) return rc;
  if (plan->ndofs > 0 && (!x || !y)) return gd::fail(GD_ERR_INVALID, "gd_adv_face_apply_f64: null vector");
  if (overlap(x, plan->ndofs, y, plan->ndofs))
    return gd::fail(GD_ERR_INVALID, "gd_adv_face_apply_f64: x and y must not overlap");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int D = plan->dim;
  // 3D GLL, q = n, separable one-axis beta, rows of whole tiles: the
  // streaming Kronecker kernel with its cell terms off, accumulating
  // (variant 0 pinned for FP64 keeps the face kernels below)
  if (D == 3 && plan->tuned_k3_lanes[0] != 0) {
    const int r = gd::k3k::apply_faces<double>(plan, x, y, st);
    if (r != GD_ERR_UNSUPPORTED) return r;
  }
  {
    Args a;
    if (int rc = common_args(plan, plan, 0, -1, a)) return rc;
    const int n = plan->desc.order + 1;
    long long nd = 1;
    for (int d = 0; d < D; ++d) nd *= n;
    const long long per = (4LL * a.S + nd) * (long long)sizeof(double);
    const int fpb = (int)std::min<long long>(kMaxFpb, kOwnerLdsBytes / per);
    if (fpb >= 8 && n <= kMaxN && n * a.q <= kMaxQN) {
      // one launch: every element adds its own share of each of its faces;
      // a small mesh is spread over about 1,024 blocks as in launch()
      const long long spread = (plan->E + 1023) / 1024;
      const int use = (int)std::max<long long>(1, std::min<long long>(fpb, spread));
      a.fpb = use;
      a.nfaces = plan->E;
      if (int rc = side_setup(plan, a.q, 0, x, y, a.s[0])) return rc;
      if (int rc = side_setup(plan, a.q, 1, x, y, a.s[1])) return rc;
      if (a.nfaces <= 0) return GD_OK;
      const int blocks = (int)((a.nfaces + use - 1) / use);
      hipLaunchKernelGGL(owner_kernel(D, n, a.q), dim3((unsigned)blocks), dim3(kThreads), (size_t)use * per, st, a);
      const hipError_t e = hipGetLastError();
      return e == hipSuccess ? GD_OK : gd::hip_fail(e, "face kernel launch");
    }
  }
  for (int axis = 0; axis < D; ++axis) {
    // interior faces of this axis, minus cell = the upper one (reference
    // normal -e_axis), in two colours by the parity of its index
    for (int c = 0; c < 2; ++c) {
      Args a;
      if (int rc = common_args(plan, plan, axis, -1, a)) return rc;
      const int64_t Ma = plan->desc.sizes[axis];
      int64_t nf = 1;
      for (int d = 0; d < D; ++d) {
        a.G[d] = d == axis ? (Ma - c) / 2 : plan->desc.sizes[d];
        nf *= a.G[d];
      }
      for (int d = D; d < gd::kMaxDim; ++d) a.G[d] = 1;
      a.nfaces = nf;
      if (int rc = side_setup(plan, a.q, 0, x, y, a.s[0])) return rc;
      if (int rc = side_setup(plan, a.q, 1, x, y, a.s[1])) return rc;
      a.s[0].base = 1 + c;
      a.s[1].base = c;
      a.s[0].stride = a.s[1].stride = 2;
      if (int rc = launch(a, st)) return rc;
    }
  }
  return GD_OK;
}

extern "C" int gd_adv_intermesh_apply_f64(const gd_plan* minus, const gd_plan* plus, int32_t face,
                                          const double* x_minus, const double* x_plus, double* y_minus,
                                          double* y_plus, gd_stream stream)
{
  using namespace gd::face;
  if (int rc = check_plan(minus, "gd_adv_intermesh_apply_f64")) return rc;
  if (int rc = check_plan(plus, "gd_adv_intermesh_apply_f64")) return rc;
  const int D = minus->dim;
  if (face < 0 || face >= 2 * D) return gd::fail(GD_ERR_INVALID, "gd_adv_intermesh_apply_f64: face out of range");
  if (!x_minus || !x_plus || !y_minus || !y_plus)
    return gd::fail(GD_ERR_INVALID, "gd_adv_intermesh_apply_f64: null vector");
  if (sides_overlap(minus, plus, x_minus, x_plus, y_minus, y_plus))
    return gd::fail(GD_ERR_INVALID, "gd_adv_intermesh_apply_f64: x and y must not overlap");
  const int axis = face % D, sign = face >= D ? 1 : -1;
  Args a;
  if (int rc = common_args(minus, plus, axis, sign, a)) return rc;
  int64_t nf = 1;
  for (int d = 0; d < gd::kMaxDim; ++d) {
    a.G[d] = (d < D && d != axis) ? minus->desc.sizes[d] : 1;
    nf *= a.G[d];
  }
  a.nfaces = nf;
  if (int rc = side_setup(minus, a.q, sign > 0 ? 1 : 0, x_minus, y_minus, a.s[0])) return rc;
  if (int rc = side_setup(plus, a.q, sign > 0 ? 0 : 1, x_plus, y_plus, a.s[1])) return rc;
  // a +axis face: minus cells on the last slice, plus cells on the first
  a.s[0].base = sign > 0 ? minus->desc.sizes[axis] - 1 : 0;
  a.s[1].base = sign > 0 ? 0 : plus->desc.sizes[axis] - 1;
  a.s[0].stride = a.s[1].stride = 0;
  return launch(a, static_cast<hipStream_t>(stream));
}

namespace gd {
namespace face {

// l_i(x) on the plan's nodes
double lagrange(const std::vector<double>& xn, int i, double x)
{
  double v = 1.0;
  for (size_t j = 0; j < xn.size(); ++j)
    if ((int)j != i) v *= (x - xn[j]) / (xn[i] - xn[j]);
  return v;
}

// the coarse side's basis at the mapped points (o + xi) / r of every subface
// offset o < r of every tangential axis, cached on the plan
int subface_tables(const gd_plan* c, int axis, const int* r, int q, Side& sd, int* bbase)
{
  const int D = c->dim, n = c->desc.order + 1;
  std::vector<int64_t> key = {n, q, axis};
  for (int d = 0; d < D; ++d) key.push_back(r[d]);
  const TableSet* ts = find_tables(c, n, q);
  if (!ts) return fail(GD_ERR_UNSUPPORTED, "face operator: plan has no tables for the shared rule");
  int nb = 0;
  for (int d = 0; d < D; ++d) {
    bbase[d] = nb;
    if (d != axis) nb += r[d];
  }
  if (c->hmap_key != key) {
    std::vector<double> xn(n), wn(n);
    if (int rc = gd_rule_points(c->desc.basis, n, xn.data(), wn.data())) return rc;
    std::vector<double> tab((size_t)nb * q * n);
    for (int d = 0; d < D; ++d) {
      if (d == axis) continue;
      for (int o = 0; o < r[d]; ++o)
        for (int k = 0; k < q; ++k)
          for (int i = 0; i < n; ++i)
            tab[((size_t)(bbase[d] + o) * q + k) * n + i] = lagrange(xn, i, (o + ts->xq[k]) / r[d]);
    }
    if (c->hmap_dev) (void)hipFree(c->hmap_dev);
    c->hmap_dev = nullptr;
    c->hmap_key.clear();
    if (hipMalloc(&c->hmap_dev, tab.size() * sizeof(double)) != hipSuccess)
      return fail(GD_ERR_ALLOC, "face operator: subface tables");
    if (hipMemcpy(c->hmap_dev, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(GD_ERR_HIP, "face operator: upload of subface tables");
    c->hmap_key = key;
  }
  sd.Bmap = static_cast<const double*>(c->hmap_dev);
  return GD_OK;
}

}  // namespace face
}  // namespace gd

extern "C" int gd_adv_hmesh_apply_f64(const gd_plan* coarse, const gd_plan* fine, int32_t face,
                                      const double* x_coarse, const double* x_fine, double* y_coarse,
                                      double* y_fine, gd_stream stream)
{
  using namespace gd::face;
  if (int rc = check_plan(coarse, "gd_adv_hmesh_apply_f64")) return rc;
  if (int rc = check_plan(fine, "gd_adv_hmesh_apply_f64")) return rc;
  const int D = coarse->dim;
  if (fine->dim != D) return gd::fail(GD_ERR_INVALID, "gd_adv_hmesh_apply_f64: plans differ in dimension");
  if (face < 0 || face >= 2 * D) return gd::fail(GD_ERR_INVALID, "gd_adv_hmesh_apply_f64: face out of range");
  if (!x_coarse || !x_fine || !y_coarse || !y_fine)
    return gd::fail(GD_ERR_INVALID, "gd_adv_hmesh_apply_f64: null vector");
  if (sides_overlap(coarse, fine, x_coarse, x_fine, y_coarse, y_fine))
    return gd::fail(GD_ERR_INVALID, "gd_adv_hmesh_apply_f64: x and y must not overlap");
  const int axis = face % D, sign = face >= D ? 1 : -1;
  int r[gd::kMaxDim];
  double wscale = 1.0;
  for (int d = 0; d < gd::kMaxDim; ++d) r[d] = 1;
  for (int d = 0; d < D; ++d) {
    if (d == axis) continue;
    const int64_t mc = coarse->desc.sizes[d], mf = fine->desc.sizes[d];
    if (mc <= 0 || mf % mc != 0)
      return gd::fail(GD_ERR_INVALID, "gd_adv_hmesh_apply_f64: fine cells per tangential axis must be a multiple "
                                      "of the coarse ones");
    r[d] = (int)(mf / mc);
    wscale /= (double)r[d];
  }
  Args a;
  if (int rc = common_args(coarse, fine, axis, sign, a, true)) return rc;
  int64_t nf = 1;
  for (int d = 0; d < gd::kMaxDim; ++d) {
    a.G[d] = (d < D && d != axis) ? coarse->desc.sizes[d] : 1;  // one face per coarse cell per launch
    nf *= a.G[d];
  }
  a.nfaces = nf;
  a.wscale = wscale;
  if (int rc = side_setup(coarse, a.q, sign > 0 ? 1 : 0, x_coarse, y_coarse, a.s[0])) return rc;
  if (int rc = side_setup(fine, a.q, sign > 0 ? 0 : 1, x_fine, y_fine, a.s[1])) return rc;
  int bbase[gd::kMaxDim] = {};
  if (int rc = subface_tables(coarse, axis, r, a.q, a.s[0], bbase)) return rc;
  a.s[0].base = sign > 0 ? coarse->desc.sizes[axis] - 1 : 0;
  a.s[1].base = sign > 0 ? 0 : fine->desc.sizes[axis] - 1;
  a.s[0].stride = a.s[1].stride = 0;
  // the r_1 x ... subfaces of a coarse face share its cell: one launch per
  // subface position, so the faces of a launch touch disjoint cells
  int o[gd::kMaxDim] = {};
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (;;) {
    for (int d = 0; d < D; ++d) {
      if (d == axis) continue;
      a.s[1].mul[d] = r[d];
      a.s[1].add[d] = o[d];
      a.s[0].bsel[d] = bbase[d] + o[d];
    }
    if (int rc = launch(a, st)) return rc;
    int d = 0;
    for (; d < D; ++d) {
      if (d == axis) continue;
      if (++o[d] < r[d]) break;
      o[d] = 0;
    }
    if (d == D) break;
  }
  return GD_OK;
}
