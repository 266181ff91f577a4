// CPU restatement of GenDiL's element-centric matrix-free DG operator apply.
// TEST INFRASTRUCTURE ONLY (checker + "port" CPU baseline); the product
// library never links, loads or calls this file.  Reference paths are
// relative to the GenDiL tree (include/gendil/...).
//
// Parity status of each entry point:
//   gdo_points           pinned: closed forms (GL/GLL moments, symmetry) and
//                        the reference's tabulated values (tests).
//   gdo_apply            pinned: golden vectors produced by GenDiL's own
//                        SerialKernelConfiguration operators (tests/golden).
//   gdo_apply_transpose  pinned: dense transpose of gdo_apply on tiny meshes
//                        and <A^T r, v> = <r, A v> (no reference counterpart).
//   gdo_grad_beta        pinned: central finite differences of gdo_apply and
//                        linearity in beta (no reference counterpart).
//   p-adaptive apply     pinned where all elements share one rule against
//                        the uniform golden vectors, and by the conservation
//                        invariant of test-advection-p-adaptivity.cpp:106.

#include "gd_oracle.h"

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kMaxDim = GDO_MAX_DIM;

// ---------------------------------------------------------------------------
// 1D rules on [0,1] from the Legendre recurrence + Newton iteration.
// GenDiL tabulates these (NumericalIntegration/QuadraturePoints/GLpoints.hpp,
// GLLpoints.hpp) to ~15 digits; here they are computed to full precision.
// ---------------------------------------------------------------------------
void legendre(int n, double x, double& p, double& dp)
{
  // P_n(x) and P_n'(x) on [-1,1].
  double p0 = 1.0, p1 = x;
  if (n == 0) { p = 1.0; dp = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  p = p1;
  dp = (std::fabs(x) < 1.0) ? n * (x * p1 - p0) / (x * x - 1.0) : 0.5 * n * (n + 1) * std::pow(x, n + 1);
}

void gauss_legendre(int n, double* x, double* w)
{
  for (int i = 0; i < n; ++i) {
    double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, t, p, dp);
      const double dt = p / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-17) break;
    }
    double p, dp;
    legendre(n, t, p, dp);
    x[i] = 0.5 * (t + 1.0);
    w[i] = 1.0 / ((1.0 - t * t) * dp * dp);  // 2/((1-t^2)P'^2) scaled by 1/2
  }
  // symmetrize
  for (int i = 0; i < n / 2; ++i) {
    const double a = 0.5 * (x[i] + (1.0 - x[n - 1 - i]));
    x[i] = a;
    x[n - 1 - i] = 1.0 - a;
    const double b = 0.5 * (w[i] + w[n - 1 - i]);
    w[i] = b;
    w[n - 1 - i] = b;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

void gauss_lobatto(int n, double* x, double* w)
{
  if (n == 1) { x[0] = 0.5; w[0] = 1.0; return; }  // GLLpoints.hpp: one point at 0.5
  const int N = n - 1;
  for (int i = 0; i < n; ++i) {
    double t = -std::cos(M_PI * i / N);
    if (i > 0 && i < N) {
      for (int it = 0; it < 100; ++it) {
        // roots of P_N'(t): Newton on q(t) = P_N'(t), q' = (2t P_N' - N(N+1) P_N)/(1-t^2)
        double p, dp;
        legendre(N, t, p, dp);
        const double d2p = (2.0 * t * dp - N * (N + 1.0) * p) / (1.0 - t * t);
        const double dt = dp / d2p;
        t -= dt;
        if (std::fabs(dt) < 1e-17) break;
      }
    }
    double p, dp;
    legendre(N, t, p, dp);
    x[i] = (i == 0) ? 0.0 : (i == N ? 1.0 : 0.5 * (t + 1.0));
    w[i] = 1.0 / (N * (N + 1.0) * p * p);  // 2/(N(N+1)P_N^2) scaled by 1/2
  }
  for (int i = 0; i < n / 2; ++i) {
    const double a = 0.5 * (x[i] + (1.0 - x[n - 1 - i]));
    x[i] = a;
    x[n - 1 - i] = 1.0 - a;
    const double b = 0.5 * (w[i] + w[n - 1 - i]);
    w[i] = b;
    w[n - 1 - i] = b;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

// Lagrange basis on nodes xs, product form of
// ShapeFunctions/lagrangeshapefunctions.hpp:37-52 (value) and :60-82 (derivative).
double lagrange_value(const std::vector<double>& xs, int k, double x)
{
  double num = 1.0, den = 1.0;
  for (size_t i = 0; i < xs.size(); ++i)
    if ((int)i != k) {
      num *= (x - xs[i]);
      den *= (xs[k] - xs[i]);
    }
  return num / den;
}

double lagrange_deriv(const std::vector<double>& xs, int k, double x)
{
  double sum = 0.0;
  const int n = (int)xs.size();
  for (int i = 0; i < n; ++i) {
    if (i == k) continue;
    double num = 1.0, den = xs[k] - xs[i];
    for (int m = 0; m < n; ++m)
      if (m != i && m != k) {
        num *= (x - xs[m]);
        den *= (xs[k] - xs[m]);
      }
    sum += num / den;
  }
  return sum;
}

// 1D tables of KernelOperators/doftoquad.hpp:116-151 (CachedDofToQuad):
// values B[q][i] = l_i(x_q), collocated derivative Gq[a][b] = l^GL_b'(x_a),
// weights, plus the face traces l_i(0), l_i(1) (faceintegrationrules.hpp:29-56
// replaces the normal dimension by ZeroPoint / OnePoint with weight 1).
struct Tables {
  int n = 0, q = 0;
  std::vector<double> xq, wq, B, Gq, b0, b1, d0, d1;  // d0, d1: l_i'(0), l_i'(1)
};

Tables make_tables(int basis, int n, int q)
{
  Tables t;
  t.n = n;
  t.q = q;
  std::vector<double> xn(n), wn(n);
  if (basis == GDO_BASIS_GLL) gauss_lobatto(n, xn.data(), wn.data());
  else gauss_legendre(n, xn.data(), wn.data());
  t.xq.resize(q);
  t.wq.resize(q);
  gauss_legendre(q, t.xq.data(), t.wq.data());
  t.B.resize(q * n);
  t.Gq.resize(q * q);
  t.b0.resize(n);
  t.b1.resize(n);
  for (int a = 0; a < q; ++a)
    for (int i = 0; i < n; ++i) t.B[a * n + i] = lagrange_value(xn, i, t.xq[a]);
  for (int a = 0; a < q; ++a)
    for (int b = 0; b < q; ++b) t.Gq[a * q + b] = lagrange_deriv(t.xq, b, t.xq[a]);
  t.d0.resize(n);
  t.d1.resize(n);
  for (int i = 0; i < n; ++i) {
    t.b0[i] = lagrange_value(xn, i, 0.0);
    t.b1[i] = lagrange_value(xn, i, 1.0);
    t.d0[i] = lagrange_deriv(xn, i, 0.0);
    t.d1[i] = lagrange_deriv(xn, i, 1.0);
  }
  return t;
}

struct TableCache {
  std::mutex mu;
  std::map<std::tuple<int, int, int>, Tables> map;
  const Tables& get(int basis, int n, int q)
  {
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(basis, n, q);
    auto it = map.find(key);
    if (it == map.end()) it = map.emplace(key, make_tables(basis, n, q)).first;
    return it->second;
  }
};

TableCache& cache()
{
  static TableCache c;
  return c;
}

// ---------------------------------------------------------------------------
// Dense local tensors (FIFO) and 1D contractions along one dimension.
// Mirrors InterpContraction (interpolatevaluesserial.hpp:25-62): value = 0,
// then value += M * in over the contracted index in increasing order.
// ---------------------------------------------------------------------------
struct Shape {
  int D;
  int ext[kMaxDim];
  int64_t size() const
  {
    int64_t s = 1;
    for (int d = 0; d < D; ++d) s *= ext[d];
    return s;
  }
};

// out[.., j, ..] = sum_i M[j*ldm + i] in[.., i, ..], in extent ext[d] -> nout.
// transpose=true uses M[i*ldm + j].
void contract(const double* in, const Shape& s, int d, const double* M, int ldm, int nout, bool transpose,
              double* out, Shape& so)
{
  so = s;
  so.ext[d] = nout;
  int64_t inner = 1, outer = 1;
  for (int k = 0; k < d; ++k) inner *= s.ext[k];
  for (int k = d + 1; k < s.D; ++k) outer *= s.ext[k];
  const int nin = s.ext[d];
  for (int64_t o = 0; o < outer; ++o)
    for (int j = 0; j < nout; ++j)
      for (int64_t a = 0; a < inner; ++a) {
        double v = 0.0;
        for (int i = 0; i < nin; ++i) {
          const double m = transpose ? M[i * ldm + j] : M[j * ldm + i];
          v += m * in[(o * nin + i) * inner + a];
        }
        out[(o * nout + j) * inner + a] = v;
      }
}

void unravel(int64_t idx, const int64_t* ext, int D, int64_t* sub)
{
  for (int d = 0; d < D; ++d) {
    sub[d] = idx % ext[d];
    idx /= ext[d];
  }
}

int64_t ravel(const int64_t* sub, const int64_t* ext, int D)
{
  int64_t idx = 0;
  for (int d = D - 1; d >= 0; --d) idx = idx * ext[d] + sub[d];
  return idx;
}

int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// Evaluate a separable field at per-axis sample indices k[d] (0..nq+1) of
// element sub-index e[d]; table stride per axis = nq+2 (uniform rule only).
double eval_sep(const gdo_separable& f, const int64_t* e, const int* k, int D, int nq)
{
  double sum = 0.0;
  for (int r = 0; r < f.rank; ++r) {
    double t = f.scale[r];
    for (int d = 0; d < D; ++d) {
      const int64_t off = f.offset[r][d];
      if (off >= 0) t *= f.data[off + e[d] * (nq + 2) + k[d]];
    }
    sum += t;
  }
  return sum;
}

struct Problem {
  const gdo_problem* p;
  int D;
  int64_t M[kMaxDim];
  int64_t E;
  bool padapt;
  double detJ;
  double invh[kMaxDim];

  explicit Problem(const gdo_problem* pp) : p(pp)
  {
    D = p->dim;
    E = 1;
    detJ = 1.0;
    for (int d = 0; d < D; ++d) {
      M[d] = p->sizes[d];
      E *= M[d];
    }
    // Determinant of diag(h) and its inverse (computeinverseanddeterminant.hpp:35).
    for (int d = 0; d < D; ++d) {
      detJ *= p->h[d];
      invh[d] = 1.0 / p->h[d];
    }
    padapt = p->elem_order != nullptr;
  }
  int order(int64_t e) const { return padapt ? p->elem_order[e] : p->order; }
  int nq(int64_t e) const { return padapt ? p->elem_nq[e] : p->nq; }
  int64_t dof0(int64_t e) const
  {
    return padapt ? p->dof_offset[e] : e * (int64_t)ipow(p->order + 1, D);
  }
  // Neighbour across local face f (CartesianMesh::GetLocalFaceInfo,
  // cartesianndmesh.hpp:109-198): axis f mod D, sign -1 for f < D.
  // Returns -1 on a (non-periodic) boundary.
  int64_t neighbor(const int64_t* sub, int f) const
  {
    const int a = f % D;
    const int sign = f < D ? -1 : 1;
    int64_t s[kMaxDim];
    for (int d = 0; d < D; ++d) s[d] = sub[d];
    if (M[a] == 1) {
      if (!p->periodic[a]) return -1;
      s[a] = 0;
    } else if (sign < 0) {
      if (s[a] == 0) {
        if (!p->periodic[a]) return -1;
        s[a] = M[a] - 1;
      } else {
        s[a] -= 1;
      }
    } else {
      if (s[a] == M[a] - 1) {
        if (!p->periodic[a]) return -1;
        s[a] = 0;
      } else {
        s[a] += 1;
      }
    }
    return ravel(s, M, D);
  }
  // Global face index for pointwise face arrays of axis a: face f_a lies
  // between element f_a-1 (its + face) and f_a (its - face).
  int64_t face_index(const int64_t* sub, int a, int sign) const
  {
    int64_t ext[kMaxDim], s[kMaxDim];
    for (int d = 0; d < D; ++d) {
      ext[d] = M[d];
      s[d] = sub[d];
    }
    ext[a] = M[a] + (p->periodic[a] ? 0 : 1);
    if (sign > 0) {
      s[a] = sub[a] + 1;
      if (p->periodic[a] && s[a] == M[a]) s[a] = 0;
    }
    return ravel(s, ext, D);
  }
  double coord(int d, int64_t ed, double xi) const
  {
    // HyperCubeCell::GetValuesAndJacobian (hypercubecell.hpp:41-52) with the
    // cell origin of CartesianMesh::GetCell (cartesianndmesh.hpp:95-105).
    const double o = p->origin[d] + p->h[d] * (double)ed;
    return o + p->h[d] * xi;
  }
  // Advection field component c at a point: k[d] sample index per axis
  // (0..nq-1 rule point, nq: xi=0, nq+1: xi=1).
  double beta_sep(int c, const int64_t* sub, const int* k, int nq) const
  {
    if (p->beta_kind == GDO_BETA_AFFINE) {
      // beta_c = c + sum_d A[c][d] X_d at X = (o + h e) + h xi
      const Tables& R = cache().get(GDO_BASIS_GL, 1, nq);
      double v = p->beta_affine_c[c];
      for (int d = 0; d < D; ++d) {
        const double xi = k[d] < nq ? R.xq[k[d]] : (k[d] == nq ? 0.0 : 1.0);
        v += p->beta_affine_A[c][d] * coord(d, sub[d], xi);
      }
      return v;
    }
    return eval_sep(p->beta[c], sub, k, D, nq);
  }
};

// Per-element scratch.
struct Work {
  std::vector<double> u, t0, t1, t2, acc, un, face_own, face_nb, face_flux, dbu[kMaxDim];
  void ensure(int64_t sz)
  {
    for (auto* v : {&u, &t0, &t1, &t2, &acc, &un, &face_own, &face_nb, &face_flux})
      if ((int64_t)v->size() < sz) v->resize(sz);
    for (int d = 0; d < kMaxDim; ++d)
      if ((int64_t)dbu[d].size() < sz) dbu[d].resize(sz);
  }
};

// Interpolate element DoFs (n^D) to cell quadrature points (q^D), last dim
// first (InterpolateValuesImpl, interpolatevaluesserial.hpp:79-103).
void interp_cell(const Tables& T, int D, const double* u, double* out, double* tmp)
{
  Shape s{D, {}};
  for (int d = 0; d < D; ++d) s.ext[d] = T.n;
  std::vector<double> a(u, u + s.size());
  for (int d = D - 1; d >= 0; --d) {
    Shape so;
    contract(a.data(), s, d, T.B.data(), T.n, T.q, false, tmp, so);
    a.assign(tmp, tmp + so.size());
    s = so;
  }
  std::memcpy(out, a.data(), a.size() * sizeof(double));
}

// Apply B^T in every dimension: q^D -> n^D (applytestfunctionsserial.hpp:68-111).
void test_cell(const Tables& T, int D, const double* in, double* out, double* tmp)
{
  Shape s{D, {}};
  for (int d = 0; d < D; ++d) s.ext[d] = T.q;
  std::vector<double> a(in, in + s.size());
  for (int d = 0; d < D; ++d) {
    Shape so;
    contract(a.data(), s, d, T.B.data(), T.n, T.n, true, tmp, so);
    a.assign(tmp, tmp + so.size());
    s = so;
  }
  std::memcpy(out, a.data(), a.size() * sizeof(double));
}

// Collocated gradient transpose at quadrature points
// (applygradienttestfunctionsthreaded.hpp:164-254, register branch :234-250):
// out(q) = sum_d sum_p Gq[p][q_d] in_d(..p..).
void grad_transpose_cell(const Tables& T, int D, double* const* in, double* out, double* tmp)
{
  Shape s{D, {}};
  for (int d = 0; d < D; ++d) s.ext[d] = T.q;
  const int64_t sz = s.size();
  for (int64_t i = 0; i < sz; ++i) out[i] = 0.0;
  for (int d = 0; d < D; ++d) {
    Shape so;
    contract(in[d], s, d, T.Gq.data(), T.q, T.q, true, tmp, so);
    for (int64_t i = 0; i < sz; ++i) out[i] += tmp[i];
  }
}

// Collocated gradient along dim d at quadrature points: out = Gq_d in.
void grad_cell(const Tables& T, int D, int d, const double* in, double* out)
{
  Shape s{D, {}};
  for (int k = 0; k < D; ++k) s.ext[k] = T.q;
  Shape so;
  contract(in, s, d, T.Gq.data(), T.q, T.q, false, out, so);
}

// Trace of element DoFs on local face (axis a, xi_a = 0 or 1), evaluated at
// the face rule of nqf points in the tangential dims (face tables of
// doftoquad.hpp with the normal dimension replaced by Zero/OnePoint).
// Result extent: nqf in d != a, 1 in d == a.
// `row` is the normal dimension's 1-point table: the basis values (b0 / b1)
// for the trace, their derivatives (d0 / d1) for the normal derivative.
void trace_row(const Tables& Tt, int D, int a, const double* row, const double* u, double* out, double* tmp)
{
  Shape s{D, {}};
  for (int d = 0; d < D; ++d) s.ext[d] = Tt.n;
  std::vector<double> cur(u, u + s.size());
  for (int d = D - 1; d >= 0; --d) {
    Shape so;
    if (d == a) contract(cur.data(), s, d, row, Tt.n, 1, false, tmp, so);
    else contract(cur.data(), s, d, Tt.B.data(), Tt.n, Tt.q, false, tmp, so);
    cur.assign(tmp, tmp + so.size());
    s = so;
  }
  std::memcpy(out, cur.data(), cur.size() * sizeof(double));
}

void trace(const Tables& Tt, int D, int a, bool at_one, const double* u, double* out, double* tmp)
{
  trace_row(Tt, D, a, at_one ? Tt.b1.data() : Tt.b0.data(), u, out, tmp);
}

// Transpose of trace_row: add face values (extent 1 in dim a) into n^D DoFs
// (ApplyAddTestFunctions for faces, applytestfunctions.hpp:251).
void trace_row_transpose_add(const Tables& Tt, int D, int a, const double* row, const double* in, double* out,
                             double* tmp)
{
  Shape s{D, {}};
  for (int d = 0; d < D; ++d) s.ext[d] = (d == a) ? 1 : Tt.q;
  std::vector<double> cur(in, in + s.size());
  for (int d = 0; d < D; ++d) {
    Shape so;
    if (d == a) contract(cur.data(), s, d, row, 1, Tt.n, true, tmp, so);
    else contract(cur.data(), s, d, Tt.B.data(), Tt.n, Tt.n, true, tmp, so);
    cur.assign(tmp, tmp + so.size());
    s = so;
  }
  for (size_t i = 0; i < cur.size(); ++i) out[i] += cur[i];
}

void trace_transpose_add(const Tables& Tt, int D, int a, bool at_one, const double* in, double* out, double* tmp)
{
  trace_row_transpose_add(Tt, D, a, at_one ? Tt.b1.data() : Tt.b0.data(), in, out, tmp);
}

// Face quadrature point loop helper: iterates the tangential multi-index
// (FIFO, dim a fixed to 0).
struct FaceIter {
  int D, a, q;
  int64_t count;
  FaceIter(int D_, int a_, int q_) : D(D_), a(a_), q(q_)
  {
    count = 1;
    for (int d = 0; d < D; ++d)
      if (d != a) count *= q;
  }
  void sub(int64_t i, int* k) const
  {
    for (int d = 0; d < D; ++d) {
      if (d == a) {
        k[d] = 0;
        continue;
      }
      k[d] = (int)(i % q);
      i /= q;
    }
  }
};

double face_weight(const Tables& Tf, const int* k, int D, int a)
{
  // getweight.hpp:98-111: Measure() (1 for conforming faces) times the product
  // of 1D weights, the normal dimension contributing weight 1.
  double w = 1.0;
  for (int d = 0; d < D; ++d) w *= (d == a) ? 1.0 : Tf.wq[k[d]];
  return 1.0 * w;
}

double cell_weight(const Tables& T, const int* k, int D)
{
  double w = 1.0;
  for (int d = 0; d < D; ++d) w *= T.wq[k[d]];
  return w;
}

// Beta normal component at a face point as seen from element `sub` (own side).
double beta_face_normal(const Problem& P, const int64_t* sub, int a, int sign, const int* kf, int nqf,
                        int64_t qf_lin)
{
  const gdo_problem* p = P.p;
  if (p->beta_kind == GDO_BETA_POINTWISE) {
    const int64_t fi = P.face_index(sub, a, sign);
    return p->beta_face[a][fi * (int64_t)ipow(nqf, P.D - 1) + qf_lin];
  }
  int k[kMaxDim];
  for (int d = 0; d < P.D; ++d) k[d] = kf[d];
  k[a] = sign < 0 ? nqf : nqf + 1;
  return P.beta_sep(a, sub, k, nqf);
}

double bc_value(const Problem& P, const int64_t* sub, int a, int sign, const int* kf, int nqf)
{
  int k[kMaxDim];
  for (int d = 0; d < P.D; ++d) k[d] = kf[d];
  k[a] = sign < 0 ? nqf : nqf + 1;
  return eval_sep(P.p->bc, sub, k, P.D, nqf);
}

enum class Mode { Forward, Transpose };

// One element of y = mass_coef*M u + adv_coef*A u (Forward) or of the
// transpose of the linear part (Transpose).
// Forward follows MassElementOperator (massoperator.hpp:39-89) and
// AdvectionFusedOperatorWith[out]BC (advectionoperator.hpp:106-441).
void element_apply(const Problem& P, Mode mode, int64_t e, const double* x, double* y, Work& W)
{
  const gdo_problem* p = P.p;
  const int D = P.D;
  const int ord = P.order(e);
  const int n = ord + 1;
  const int q = P.nq(e);
  const Tables& T = cache().get(p->basis, n, q);
  const int64_t nd = ipow(n, D);
  const int64_t nqd = ipow(q, D);
  int64_t sub[kMaxDim];
  unravel(e, P.M, D, sub);
  const double* ue = x + P.dof0(e);
  std::vector<double> out(nd, 0.0);
  double* tmp = W.t2.data();

  // ---- cell terms ----
  double* Bu = W.t0.data();
  interp_cell(T, D, ue, Bu, tmp);
  std::vector<double> qacc(nqd, 0.0);
  int k[kMaxDim];
  if (p->mass_coef != 0.0) {
    // Du = w * detJ * sigma(X) * Bu (massoperator.hpp:65-83); symmetric, so the
    // transpose applies the same diagonal.
    std::vector<double> Du(nqd);
    for (int64_t i = 0; i < nqd; ++i) {
      int64_t t = i;
      for (int d = 0; d < D; ++d) {
        k[d] = (int)(t % q);
        t /= q;
      }
      const double w = cell_weight(T, k, D);
      const double sigma = eval_sep(p->sigma, sub, k, D, q);
      Du[i] = w * P.detJ * sigma * Bu[i];
    }
    std::vector<double> m(nd);
    test_cell(T, D, Du.data(), m.data(), tmp);
    for (int64_t i = 0; i < nd; ++i) out[i] += p->mass_coef * m[i];
  }
  if (p->adv_coef == 0.0) {
    double* ye = y + P.dof0(e);
    for (int64_t i = 0; i < nd; ++i) ye[i] = out[i];
    return;
  }
  std::vector<double> adv(nd, 0.0);
  // coefficient c_d(q) = -w * detJ * beta_d(X) / h_d (advectionoperator.hpp:301-331,
  // ApplyMappingTranspose applymappingtranspose.hpp:96).
  auto cell_coef = [&](int64_t i, double* c) {
    int64_t t = i;
    for (int d = 0; d < D; ++d) {
      k[d] = (int)(t % q);
      t /= q;
    }
    const double w = cell_weight(T, k, D);
    for (int d = 0; d < D; ++d) {
      double b;
      if (p->beta_kind == GDO_BETA_POINTWISE) b = p->beta_cell[(e * D + d) * nqd + i];
      else b = P.beta_sep(d, sub, k, q);
      c[d] = -w * P.detJ * (b * P.invh[d]);
    }
  };
  if (mode == Mode::Forward) {
    double* dbu[kMaxDim];
    for (int d = 0; d < D; ++d) dbu[d] = W.dbu[d].data();
    double c[kMaxDim];
    for (int64_t i = 0; i < nqd; ++i) {
      cell_coef(i, c);
      for (int d = 0; d < D; ++d) dbu[d][i] = c[d] * Bu[i];
    }
    std::vector<double> gt(nqd);
    grad_transpose_cell(T, D, dbu, gt.data(), tmp);
    test_cell(T, D, gt.data(), adv.data(), tmp);
  } else {
    // A_cell^T = sum_d B^T C_d Gq_d B
    std::vector<double> s(nqd, 0.0), g(nqd);
    double c[kMaxDim];
    for (int d = 0; d < D; ++d) {
      grad_cell(T, D, d, Bu, g.data());
      for (int64_t i = 0; i < nqd; ++i) {
        cell_coef(i, c);
        s[i] += c[d] * g[i];
      }
    }
    test_cell(T, D, s.data(), adv.data(), tmp);
  }

  // ---- face terms (FaceLoop, faceloop.hpp:30-52: faces 0..2D-1) ----
  for (int f = 0; f < 2 * D; ++f) {
    const int a = f % D;
    const int sign = f < D ? -1 : 1;
    const int64_t nb = P.neighbor(sub, f);
    const bool boundary = nb < 0;
    if (boundary && p->bc_mode == GDO_BC_NONE) continue;  // single-lambda FaceLoop skips them
    int64_t nsub[kMaxDim];
    int qf = q;
    const Tables* Tn = nullptr;
    if (!boundary) {
      unravel(nb, P.M, D, nsub);
      const int qn = P.nq(nb);
      qf = qn > q ? qn : q;  // p-adaptive face rule: max of the two cell rules
      if (p->plant == 1) qf = qn < q ? qn : q;
      Tn = &cache().get(p->basis, P.order(nb) + 1, qf);
    }
    const Tables& Tf = cache().get(p->basis, n, qf);
    const FaceIter it(D, a, qf);
    std::vector<double> own(it.count), nbv(it.count), flux(it.count);
    trace(Tf, D, a, sign > 0, ue, own.data(), tmp);
    if (!boundary) {
      const double* unb = x + P.dof0(nb);
      std::vector<double> swapped;
      if (p->plant == 2 && P.order(nb) != ord && D >= 3) {
        int t[2] = {-1, -1};
        for (int d = 0, k = 0; d < D && k < 2; ++d)
          if (d != a) t[k++] = d;
        const int nn = P.order(nb) + 1;
        const int64_t ndn = ipow(nn, D);
        swapped.resize(ndn);
        for (int64_t i = 0; i < ndn; ++i) {
          int idx[kMaxDim];
          int64_t r = i;
          for (int d = 0; d < D; ++d) {
            idx[d] = (int)(r % nn);
            r /= nn;
          }
          std::swap(idx[t[0]], idx[t[1]]);
          int64_t j = 0;
          for (int d = D - 1; d >= 0; --d) j = j * nn + idx[d];
          swapped[j] = unb[i];
        }
        unb = swapped.data();
      }
      trace(*Tn, D, a, sign < 0, unb, nbv.data(), tmp);
    }
    int kf[kMaxDim];
    for (int64_t i = 0; i < it.count; ++i) {
      it.sub(i, kf);
      const double w = face_weight(Tf, kf, D, a);
      // physical normal = inv_J * (sign e_a) (computephysicalnormal.hpp:27-35)
      const double nrm = (double)sign * P.invh[a];
      const double dot = beta_face_normal(P, sub, a, sign, kf, qf, i) * nrm;
      if (mode == Mode::Forward) {
        double ext;
        if (boundary) ext = bc_value(P, sub, a, sign, kf, qf);  // advectionoperator.hpp:422
        else ext = nbv[i];
        const double upwind = dot > 0.0 ? dot * own[i] : dot * ext;  // :384
        flux[i] = w * P.detJ * upwind;
      } else {
        // Transpose of the linear part: own block uses this side's upwind
        // selection; the neighbour's block uses the neighbour's own normal
        // and coefficients at the same point.
        double v = dot > 0.0 ? dot * own[i] : 0.0;
        double tv = w * P.detJ * v;
        if (!boundary) {
          int64_t dummy[kMaxDim];
          (void)dummy;
          const int nsign = -sign;
          const double ndot = beta_face_normal(P, nsub, a, nsign, kf, qf, i) * ((double)nsign * P.invh[a]);
          const double wn = face_weight(*Tn, kf, D, a);
          const double vn = ndot > 0.0 ? 0.0 : ndot * nbv[i];
          tv += wn * P.detJ * vn;
        }
        flux[i] = tv;
      }
    }
    trace_transpose_add(Tf, D, a, sign > 0, flux.data(), adv.data(), tmp);
  }
  double* ye = y + P.dof0(e);
  for (int64_t i = 0; i < nd; ++i) ye[i] = out[i] + p->adv_coef * adv[i];
}

// One element of the SIPDG diffusion apply, y overwritten: the volume term of
// DiffusionElementOperator (diffusionoperator.hpp:44-101) plus the face terms
// of DiffusionFaceOperator (:130-331), interior faces (:170-272) and boundary
// faces (:273-330).  Gradients at the cell rule points are the collocated
// derivative of the interpolated values (InterpolateGradientSerial,
// interpolategradientserial.hpp:107-123, branch num_quads >= num_dofs); on a
// face the normal derivative contracts the normal dimension with l_i'(0) or
// l_i'(1) (InterpolateGradientSumFactSerial, :39-48, num_quads = 1).
void diffusion_element(const Problem& P, const gdo_separable& mu, double sig, double kap, int64_t e, const double* x,
                       double* y, Work& W)
{
  const gdo_problem* p = P.p;
  const int D = P.D;
  const int n = p->order + 1;
  const int q = p->nq;
  const Tables& T = cache().get(p->basis, n, q);
  const int64_t nd = ipow(n, D);
  const int64_t nqd = ipow(q, D);
  int64_t sub[kMaxDim];
  unravel(e, P.M, D, sub);
  const double* ue = x + P.dof0(e);
  double* tmp = W.t2.data();
  std::vector<double> out(nd, 0.0);

  // ---- volume: B^T sum_d Gq_d^T (w detJ mu / h_d^2) Gq_d B u ----
  double* Bu = W.t0.data();
  interp_cell(T, D, ue, Bu, tmp);
  double* gu[kMaxDim];
  for (int d = 0; d < D; ++d) {
    gu[d] = W.dbu[d].data();
    grad_cell(T, D, d, Bu, gu[d]);
  }
  int k[kMaxDim];
  for (int64_t i = 0; i < nqd; ++i) {
    int64_t t = i;
    for (int d = 0; d < D; ++d) {
      k[d] = (int)(t % q);
      t /= q;
    }
    const double w = cell_weight(T, k, D) * P.detJ * eval_sep(mu, sub, k, D, q);
    for (int d = 0; d < D; ++d) {
      double g = gu[d][i] * P.invh[d];  // ApplyMapping (physical gradient)
      g *= P.invh[d];                   // ApplyMappingTranspose
      gu[d][i] = w * g;
    }
  }
  std::vector<double> gt(nqd);
  grad_transpose_cell(T, D, gu, gt.data(), tmp);
  test_cell(T, D, gt.data(), out.data(), tmp);

  // ---- faces (FaceLoop with an interior and a boundary lambda) ----
  for (int f = 0; f < 2 * D; ++f) {
    const int a = f % D;
    const int sign = f < D ? -1 : 1;
    const int64_t nb = P.neighbor(sub, f);
    const bool boundary = nb < 0;
    const FaceIter it(D, a, q);
    const double* bo = sign > 0 ? T.b1.data() : T.b0.data();
    const double* go = sign > 0 ? T.d1.data() : T.d0.data();
    std::vector<double> tv(it.count), tg(it.count), nv(it.count, 0.0), ng(it.count, 0.0);
    std::vector<double> du(it.count), dg(it.count);
    trace_row(T, D, a, bo, ue, tv.data(), tmp);
    trace_row(T, D, a, go, ue, tg.data(), tmp);
    if (!boundary) {
      // the neighbour's face is its opposite one (xi_a = 0 when this face is at 1)
      const double* unb = x + P.dof0(nb);
      const double* bn = sign > 0 ? T.b0.data() : T.b1.data();
      const double* gn = (sign > 0) != (p->plant == 4) ? T.d0.data() : T.d1.data();
      trace_row(T, D, a, bn, unb, nv.data(), tmp);
      trace_row(T, D, a, gn, unb, ng.data(), tmp);
    }
    int kf[kMaxDim];
    for (int64_t i = 0; i < it.count; ++i) {
      it.sub(i, kf);
      const double w = face_weight(T, kf, D, a);
      for (int d = 0; d < D; ++d) k[d] = kf[d];
      k[a] = sign < 0 ? q : q + 1;  // face point seen from this element
      const double m = eval_sep(mu, sub, k, D, q);
      const double g_own = tg[i] * P.invh[a];  // normal component of inv_J^T grad
      double avg, jump, half;
      if (!boundary) {
        const double g_nb = ng[i] * P.invh[a];
        avg = 0.5 * m * (g_own + g_nb);
        jump = tv[i] - nv[i];
        half = p->plant == 3 ? 1.0 : 0.5;
      } else {
        avg = m * g_own;
        jump = tv[i];
        half = 1.0;
      }
      const double nrm = (double)sign * P.invh[a];  // computephysicalnormal.hpp:27-35
      const double dot = avg * nrm;
      const double norm = std::sqrt(nrm * nrm);
      const double h_inv = norm;
      du[i] = -w * P.detJ * dot + w * P.detJ * norm * kap * h_inv * m * jump;
      const double ws = sig * w * P.detJ * jump * m * half;
      dg[i] = (ws * nrm) * P.invh[a];  // ApplyMappingTranspose of the normal
    }
    // ApplyValuesAndGradientTestFunctions<true> on this side
    trace_row_transpose_add(T, D, a, bo, du.data(), out.data(), tmp);
    trace_row_transpose_add(T, D, a, go, dg.data(), out.data(), tmp);
  }
  double* ye = y + P.dof0(e);
  for (int64_t i = 0; i < nd; ++i) ye[i] = out[i];
}

int64_t num_dofs(const Problem& P)
{
  if (P.padapt) {
    const int64_t last = P.E - 1;
    return P.p->dof_offset[last] + ipow(P.order(last) + 1, P.D);
  }
  return P.E * ipow(P.p->order + 1, P.D);
}

int64_t max_local(const Problem& P)
{
  int maxn = P.p->order + 1, maxq = P.p->nq;
  if (P.padapt) {
    for (int64_t e = 0; e < P.E; ++e) {
      maxn = std::max(maxn, P.order(e) + 1);
      maxq = std::max(maxq, (int)P.nq(e));
    }
  }
  const int m = std::max(maxn, maxq);
  return 4 * (int64_t)ipow(m, P.D) + 16;
}

template <typename F>
void parallel_elements(const Problem& P, F&& f)
{
  const int64_t sz = max_local(P);
#pragma omp parallel num_threads(P.p->num_threads > 0 ? P.p->num_threads : omp_get_max_threads())
  {
    Work W;
    W.ensure(sz);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < P.E; ++e) f(e, W);
  }
}

}  // namespace

extern "C" {

int gdo_version(void) { return 1; }

int gdo_points(int32_t basis, int32_t npts, double* pts, double* wts)
{
  if (npts < 1 || npts > 64) return 1;
  if (basis == GDO_BASIS_GLL) gauss_lobatto(npts, pts, wts);
  else gauss_legendre(npts, pts, wts);
  return 0;
}

int gdo_axis_points(const gdo_problem* p, int32_t axis, double* out)
{
  Problem P(p);
  const int q = p->nq;
  std::vector<double> xq(q), wq(q);
  gauss_legendre(q, xq.data(), wq.data());
  for (int64_t e = 0; e < P.M[axis]; ++e) {
    for (int k = 0; k < q; ++k) out[e * (q + 2) + k] = P.coord(axis, e, xq[k]);
    out[e * (q + 2) + q] = P.coord(axis, e, 0.0);
    out[e * (q + 2) + q + 1] = P.coord(axis, e, 1.0);
  }
  return 0;
}

int64_t gdo_num_dofs(const gdo_problem* p)
{
  Problem P(p);
  return num_dofs(P);
}

int gdo_apply(const gdo_problem* p, const double* x, double* y)
{
  Problem P(p);
  parallel_elements(P, [&](int64_t e, Work& W) { element_apply(P, Mode::Forward, e, x, y, W); });
  return 0;
}

int gdo_apply_elements(const gdo_problem* p, const double* x, double* y, const int64_t* elems, int64_t count)
{
  Problem P(p);
  for (int64_t i = 0; i < count; ++i)
    if (elems[i] < 0 || elems[i] >= P.E) return 1;
  const int64_t sz = max_local(P);
#pragma omp parallel num_threads(P.p->num_threads > 0 ? P.p->num_threads : omp_get_max_threads())
  {
    Work W;
    W.ensure(sz);
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < count; ++i) element_apply(P, Mode::Forward, elems[i], x, y, W);
  }
  return 0;
}

int gdo_apply_transpose(const gdo_problem* p, const double* r, double* z)
{
  Problem P(p);
  parallel_elements(P, [&](int64_t e, Work& W) { element_apply(P, Mode::Transpose, e, r, z, W); });
  return 0;
}

int gdo_grad_beta(const gdo_problem* p, const double* u, const double* r, double* dbeta_cell,
                  double* const* dbeta_face)
{
  if (p->beta_kind != GDO_BETA_POINTWISE || p->elem_order) return 1;
  Problem P(p);
  const int D = P.D;
  const int n = p->order + 1;
  const int q = p->nq;
  const Tables& T = cache().get(p->basis, n, q);
  const int64_t nd = ipow(n, D), nqd = ipow(q, D);
  parallel_elements(P, [&](int64_t e, Work& W) {
    int64_t sub[kMaxDim];
    unravel(e, P.M, D, sub);
    double* tmp = W.t2.data();
    std::vector<double> Bu(nqd), Br(nqd), g(nqd);
    interp_cell(T, D, u + e * nd, Bu.data(), tmp);
    interp_cell(T, D, r + e * nd, Br.data(), tmp);
    int k[kMaxDim];
    // cell: d/dbeta_d(q) = (Gq_d B r)(q) * (-w detJ / h_d) * Bu(q)
    for (int d = 0; d < D; ++d) {
      grad_cell(T, D, d, Br.data(), g.data());
      for (int64_t i = 0; i < nqd; ++i) {
        int64_t t = i;
        for (int dd = 0; dd < D; ++dd) {
          k[dd] = (int)(t % q);
          t /= q;
        }
        const double w = cell_weight(T, k, D);
        const double wb = -w * P.detJ * Bu[i];
        dbeta_cell[(e * D + d) * nqd + i] = p->adv_coef * g[i] * (wb * P.invh[d]);
      }
    }
    // faces: each global face is written by the element on its + side (its
    // local "-" face); the last element of a non-periodic axis also writes
    // its "+" boundary face.
    for (int f = 0; f < 2 * D; ++f) {
      const int a = f % D;
      const int sign = f < D ? -1 : 1;
      const int64_t nb = P.neighbor(sub, f);
      const bool boundary = nb < 0;
      if (sign > 0 && !boundary) continue;
      const FaceIter it(D, a, q);
      std::vector<double> uo(it.count), ro(it.count), un(it.count), rn(it.count);
      trace(T, D, a, sign > 0, u + e * nd, uo.data(), tmp);
      trace(T, D, a, sign > 0, r + e * nd, ro.data(), tmp);
      if (!boundary) {
        trace(T, D, a, sign < 0, u + nb * nd, un.data(), tmp);
        trace(T, D, a, sign < 0, r + nb * nd, rn.data(), tmp);
      }
      const int64_t fi = P.face_index(sub, a, sign);
      const int64_t nqf = it.count;
      int kf[kMaxDim];
      for (int64_t i = 0; i < nqf; ++i) {
        it.sub(i, kf);
        const double w = face_weight(T, kf, D, a);
        const double b = p->beta_face[a][fi * nqf + i];
        double grad = 0.0;
        if (boundary) {
          if (p->bc_mode == GDO_BC_NONE) {
            grad = 0.0;
          } else {
            const double nrm = (double)sign * P.invh[a];
            const double dot = b * nrm;
            const double ext = bc_value(P, sub, a, sign, kf, q);
            const double up = dot > 0.0 ? uo[i] : ext;
            grad = ro[i] * w * P.detJ * nrm * up;
          }
        } else {
          // own side (normal sign) and neighbour side (normal -sign)
          const double nrm = (double)sign * P.invh[a];
          const double dot = b * nrm;
          const double up_own = dot > 0.0 ? uo[i] : un[i];
          const double ndot = b * (-nrm);
          const double up_nb = ndot > 0.0 ? un[i] : uo[i];
          grad = ro[i] * w * P.detJ * nrm * up_own + rn[i] * w * P.detJ * (-nrm) * up_nb;
        }
        dbeta_face[a][fi * nqf + i] = p->adv_coef * grad;
      }
    }
  });
  return 0;
}

int gdo_diffusion_apply(const gdo_problem* p, const gdo_separable* mu, double sigma, double kappa, const double* x,
                        double* y)
{
  if (p->elem_order) return 1;  // uniform order only
  Problem P(p);
  parallel_elements(P, [&](int64_t e, Work& W) { diffusion_element(P, *mu, sigma, kappa, e, x, y, W); });
  return 0;
}

}  // extern "C"
