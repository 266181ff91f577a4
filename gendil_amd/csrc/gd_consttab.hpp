// Compile-time 1D tables for GLL nodes (N) and Gauss-Legendre quadrature (Q)
// on [0,1]: B[q][i] = l_i(x_q), Gb[q][i] = l_i'(x_q), weights w[q].
//
// Nodes are computed by bracketing the sign changes of the Legendre
// polynomial P_Q (Gauss) or of P_{N-1}' (interior Lobatto nodes) on a fine
// grid and polishing with Newton's method, all in constexpr double
// arithmetic, so the kernels see the tables as literal constants (no loads).
#pragma once

namespace gd {
namespace ct {

struct LegVal {
  double p, dp, d2p;
};

// P_n, P_n', P_n'' at t in [-1, 1]
constexpr LegVal legendre(int n, double t)
{
  if (n == 0) return {1.0, 0.0, 0.0};
  double p0 = 1.0, p1 = t;
  double d0 = 0.0, d1 = 1.0;
  double s0 = 0.0, s1 = 0.0;
  for (int k = 1; k < n; ++k) {
    const double p2 = ((2 * k + 1) * t * p1 - k * p0) / (k + 1);
    const double d2 = ((2 * k + 1) * (p1 + t * d1) - k * d0) / (k + 1);
    const double e2 = ((2 * k + 1) * (2 * d1 + t * s1) - k * s0) / (k + 1);
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d2;
    s0 = s1;
    s1 = e2;
  }
  return {p1, d1, s1};
}

constexpr double cabs(double v) { return v < 0 ? -v : v; }

// roots of f on (-1, 1): f = P_n (deriv = 0) or P_n' (deriv = 1), ascending
template <int COUNT>
struct Roots {
  double r[COUNT > 0 ? COUNT : 1] = {};
};

template <int COUNT>
constexpr Roots<COUNT> roots(int n, int deriv)
{
  Roots<COUNT> out;
  if constexpr (COUNT == 0) return out;
  constexpr int GRID = 4096;
  int found = 0;
  auto f = [&](double t) { const LegVal v = legendre(n, t); return deriv ? v.dp : v.p; };
  auto df = [&](double t) { const LegVal v = legendre(n, t); return deriv ? v.d2p : v.dp; };
  double a = -1.0, fa = f(-1.0 + 1e-14);
  for (int g = 1; g <= GRID && found < COUNT; ++g) {
    const double b = -1.0 + 2.0 * g / GRID;
    const double fb = f(g == GRID ? 1.0 - 1e-14 : b);
    if ((fa < 0) != (fb < 0)) {
      // bisection then Newton
      double lo = a, hi = b, flo = fa;
      for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = f(mid);
        if ((fm < 0) == (flo < 0)) {
          lo = mid;
          flo = fm;
        } else {
          hi = mid;
        }
      }
      double t = 0.5 * (lo + hi);
      for (int it = 0; it < 5; ++it) {
        const double d = df(t);
        if (d != 0.0) {
          const double step = f(t) / d;
          if (cabs(step) > 1e-300) t -= step;
        }
      }
      out.r[found++] = t;
    }
    a = b;
    fa = fb;
  }
  return out;
}

template <int N, int Q>
struct Tables {
  double xn[N] = {};  // GLL nodes on [0,1]
  double xq[Q] = {};  // Gauss points on [0,1]
  double w[Q] = {};
  double B[Q][N] = {};
  double Gb[Q][N] = {};
};

template <int N, int Q>
constexpr Tables<N, Q> make_gll_gauss()
{
  Tables<N, Q> T;
  // Gauss-Legendre
  const Roots<Q> rq = roots<Q>(Q, 0);
  for (int i = 0; i < Q; ++i) {
    double t = rq.r[i];
    // symmetrise
    const double tm = rq.r[Q - 1 - i];
    t = 0.5 * (t - tm);
    T.xq[i] = 0.5 * (t + 1.0);
    const LegVal v = legendre(Q, t);
    T.w[i] = 1.0 / ((1.0 - t * t) * v.dp * v.dp);
  }
  if (Q % 2 == 1) T.xq[Q / 2] = 0.5;
  // Gauss-Lobatto nodes: 0, roots of P_{N-1}', 1
  if constexpr (N == 1) {
    T.xn[0] = 0.5;
  } else {
    T.xn[0] = 0.0;
    T.xn[N - 1] = 1.0;
    const Roots<N - 2> rn = roots<N - 2>(N - 1, 1);
    for (int i = 0; i < N - 2; ++i) {
      const double t = 0.5 * (rn.r[i] - rn.r[N - 3 - i]);
      T.xn[i + 1] = 0.5 * (t + 1.0);
    }
    if (N % 2 == 1) T.xn[N / 2] = 0.5;
  }
  for (int q = 0; q < Q; ++q)
    for (int k = 0; k < N; ++k) {
      double num = 1.0, den = 1.0;
      for (int i = 0; i < N; ++i)
        if (i != k) {
          num *= T.xq[q] - T.xn[i];
          den *= T.xn[k] - T.xn[i];
        }
      T.B[q][k] = num / den;
      double s = 0.0;
      for (int i = 0; i < N; ++i) {
        if (i == k) continue;
        double nu = 1.0, de = T.xn[k] - T.xn[i];
        for (int j = 0; j < N; ++j)
          if (j != i && j != k) {
            nu *= T.xq[q] - T.xn[j];
            de *= T.xn[k] - T.xn[j];
          }
        s += nu / de;
      }
      T.Gb[q][k] = s;
    }
  // exact mirror symmetry about the element centre (nodes and points are
  // symmetric): B[Q-1-q][N-1-i] = B[q][i], Gb[Q-1-q][N-1-i] = -Gb[q][i], so
  // the kernels see half as many distinct constants
  for (int q = 0; q < Q; ++q)
    for (int k = 0; k < N; ++k) {
      const int qm = Q - 1 - q, km = N - 1 - k;
      if (q * N + k < qm * N + km) {
        T.B[qm][km] = T.B[q][k];
        T.Gb[qm][km] = -T.Gb[q][k];
      }
    }
  for (int q = 0; q < Q / 2; ++q) T.w[Q - 1 - q] = T.w[q];
  return T;
}

// Collocated derivative on the Q Gauss points: Gq[a][b] = l_b'(x_a), l_b the
// Lagrange polynomials on those points (antisymmetric under the mirror map).
template <int Q>
struct Colloc {
  double Gq[Q][Q] = {};
};

template <int Q>
constexpr Colloc<Q> make_gauss_colloc()
{
  const Tables<1, Q> g = make_gll_gauss<1, Q>();
  Colloc<Q> C;
  for (int a = 0; a < Q; ++a)
    for (int k = 0; k < Q; ++k) {
      double s = 0.0;
      for (int i = 0; i < Q; ++i) {
        if (i == k) continue;
        double nu = 1.0, de = g.xq[k] - g.xq[i];
        for (int j = 0; j < Q; ++j)
          if (j != i && j != k) {
            nu *= g.xq[a] - g.xq[j];
            de *= g.xq[k] - g.xq[j];
          }
        s += nu / de;
      }
      C.Gq[a][k] = s;
    }
  for (int a = 0; a < Q; ++a)
    for (int k = 0; k < Q; ++k) {
      const int am = Q - 1 - a, km = Q - 1 - k;
      if (a * Q + k < am * Q + km) C.Gq[am][km] = -C.Gq[a][k];
    }
  return C;
}

}  // namespace ct
}  // namespace gd
