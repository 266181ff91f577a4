// Plan construction: validation of the mesh / element / rule description,
// 1D basis tables (uploaded once per plan, as GenDiL builds them once per
// operator in bilinearoperator.hpp:20-106), coefficient descriptors and the
// C-ABI housekeeping entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <string>
#include <utility>

#include "gd_internal.hpp"
#include "gd_k3k_api.hpp"
#include "gd_padapt_api.hpp"

namespace gd {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what)
{
  set_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  return GD_ERR_HIP;
}

// ---------------------------------------------------------------------------
// 1D rules on [0,1].  Gauss-Legendre nodes are the roots of P_n, Gauss-Lobatto
// nodes are 0, 1 and the roots of P_{n-1}'; both refined by Newton's method
// from Chebyshev-like initial guesses.  (GenDiL hard-codes ~15-digit tables:
// NumericalIntegration/QuadraturePoints/GLpoints.hpp, GLLpoints.hpp.)
// ---------------------------------------------------------------------------
namespace {

// returns P_n(t), P_{n-1}(t)
std::pair<double, double> legendre_pair(int n, double t)
{
  double pm1 = 1.0, p = t;
  if (n == 0) return {1.0, 0.0};
  for (int k = 1; k < n; ++k) {
    const double pn = ((2 * k + 1) * t * p - k * pm1) / (k + 1);
    pm1 = p;
    p = pn;
  }
  return {p, pm1};
}

void rule_gl(int n, double* x, double* w)
{
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));  // descending roots
    for (int it = 0; it < 64; ++it) {
      auto [p, pm1] = legendre_pair(n, t);
      const double dp = n * (t * p - pm1) / (t * t - 1.0);
      const double step = p / dp;
      t -= step;
      if (std::fabs(step) <= 1e-16 * std::fabs(t) + 1e-300) break;
    }
    auto [p, pm1] = legendre_pair(n, t);
    const double dp = n * (t * p - pm1) / (t * t - 1.0);
    const double wt = 1.0 / ((1.0 - t * t) * dp * dp);  // half of 2/((1-t^2)P'^2)
    x[n - 1 - i] = 0.5 + 0.5 * t;
    x[i] = 0.5 - 0.5 * t;
    w[i] = w[n - 1 - i] = wt;
  }
  if (n % 2) x[n / 2] = 0.5;
}

void rule_gll(int n, double* x, double* w)
{
  if (n == 1) {  // single node at the centre (GLLpoints.hpp one-point rule)
    x[0] = 0.5;
    w[0] = 1.0;
    return;
  }
  const int m = n - 1;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(M_PI * i / m);
    if (i > 0) {
      for (int it = 0; it < 64; ++it) {
        auto [p, pm1] = legendre_pair(m, t);
        const double dp = m * (t * p - pm1) / (t * t - 1.0);
        const double d2p = (2.0 * t * dp - m * (m + 1.0) * p) / (1.0 - t * t);
        const double step = dp / d2p;
        t -= step;
        if (std::fabs(step) <= 1e-16 * std::fabs(t) + 1e-300) break;
      }
    }
    auto [p, pm1] = legendre_pair(m, t);
    (void)pm1;
    const double wt = 1.0 / (m * (m + 1.0) * p * p);
    x[n - 1 - i] = (i == 0) ? 1.0 : 0.5 + 0.5 * t;
    x[i] = (i == 0) ? 0.0 : 0.5 - 0.5 * t;
    w[i] = w[n - 1 - i] = wt;
  }
  if (n % 2) x[n / 2] = 0.5;
}

// Lagrange basis through nodes xs (lagrangeshapefunctions.hpp:37-86 math).
double lag(const std::vector<double>& xs, int k, double x)
{
  double num = 1.0, den = 1.0;
  for (int i = 0; i < (int)xs.size(); ++i)
    if (i != k) {
      num *= x - xs[i];
      den *= xs[k] - xs[i];
    }
  return num / den;
}

double dlag(const std::vector<double>& xs, int k, double x)
{
  double s = 0.0;
  for (int i = 0; i < (int)xs.size(); ++i) {
    if (i == k) continue;
    double num = 1.0, den = xs[k] - xs[i];
    for (int j = 0; j < (int)xs.size(); ++j)
      if (j != i && j != k) {
        num *= x - xs[j];
        den *= xs[k] - xs[j];
      }
    s += num / den;
  }
  return s;
}

int build_tables(int basis, int n, int q, TableSet& ts)
{
  ts.n = n;
  ts.q = q;
  std::vector<double> xn(n), wn(n);
  if (basis == GD_BASIS_GLL) rule_gll(n, xn.data(), wn.data());
  else rule_gl(n, xn.data(), wn.data());
  ts.xq.resize(q);
  ts.w.resize(q);
  rule_gl(q, ts.xq.data(), ts.w.data());
  ts.B.resize((size_t)q * n);
  ts.Gq.resize((size_t)q * q);
  ts.b0.resize(n);
  ts.b1.resize(n);
  for (int a = 0; a < q; ++a)
    for (int i = 0; i < n; ++i) ts.B[a * n + i] = lag(xn, i, ts.xq[a]);
  for (int a = 0; a < q; ++a)
    for (int b = 0; b < q; ++b) ts.Gq[a * q + b] = dlag(ts.xq, b, ts.xq[a]);
  for (int i = 0; i < n; ++i) {
    ts.b0[i] = lag(xn, i, 0.0);
    ts.b1[i] = lag(xn, i, 1.0);
  }
  ts.Gb.resize((size_t)q * n);
  for (int a = 0; a < q; ++a)
    for (int i = 0; i < n; ++i) ts.Gb[a * n + i] = dlag(xn, i, ts.xq[a]);
  ts.d0.resize(n);
  ts.d1.resize(n);
  for (int i = 0; i < n; ++i) {
    ts.d0[i] = dlag(xn, i, 0.0);
    ts.d1[i] = dlag(xn, i, 1.0);
  }
  // device upload: B | Gq | b0 | b1 | w | xq | Gb | d0 | d1
  std::vector<double> blob;
  blob.insert(blob.end(), ts.B.begin(), ts.B.end());
  blob.insert(blob.end(), ts.Gq.begin(), ts.Gq.end());
  blob.insert(blob.end(), ts.b0.begin(), ts.b0.end());
  blob.insert(blob.end(), ts.b1.begin(), ts.b1.end());
  blob.insert(blob.end(), ts.w.begin(), ts.w.end());
  blob.insert(blob.end(), ts.xq.begin(), ts.xq.end());
  blob.insert(blob.end(), ts.Gb.begin(), ts.Gb.end());
  blob.insert(blob.end(), ts.d0.begin(), ts.d0.end());
  blob.insert(blob.end(), ts.d1.begin(), ts.d1.end());
  std::vector<float> blob32(blob.begin(), blob.end());
  GD_HIP(hipMalloc(&ts.dev64, blob.size() * sizeof(double)));
  GD_HIP(hipMalloc(&ts.dev32, blob32.size() * sizeof(float)));
  GD_HIP(hipMemcpy(ts.dev64, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice));
  GD_HIP(hipMemcpy(ts.dev32, blob32.data(), blob32.size() * sizeof(float), hipMemcpyHostToDevice));
  auto fill = [&](auto* base, auto& t) {
    size_t o = 0;
    t.B = base + o;
    o += ts.B.size();
    t.Gq = base + o;
    o += ts.Gq.size();
    t.b0 = base + o;
    o += n;
    t.b1 = base + o;
    o += n;
    t.w = base + o;
    o += q;
    t.xq = base + o;
    o += q;
    t.Gb = base + o;
    o += ts.Gb.size();
    t.d0 = base + o;
    o += n;
    t.d1 = base + o;
  };
  fill(static_cast<double*>(ts.dev64), ts.t64);
  fill(static_cast<float*>(ts.dev32), ts.t32);
  return GD_OK;
}

int check_sep(const gd_plan* p, const gd_separable* f, const char* what, bool padapt_ok)
{
  if (!f) return GD_OK;
  if (f->rank < 0 || f->rank > GD_MAX_TERMS) return fail(GD_ERR_INVALID, std::string(what) + ": rank out of range");
  for (int r = 0; r < f->rank; ++r)
    for (int d = 0; d < p->dim; ++d) {
      const int64_t off = f->offset[r][d];
      if (off < 0) continue;
      if (p->padapt && !padapt_ok)
        return fail(GD_ERR_UNSUPPORTED, std::string(what) + ": tabulated factors need a uniform rule");
      const int64_t len = p->desc.sizes[d] * (p->desc.num_quad + 2);
      if (!f->data || off + len > f->data_len)
        return fail(GD_ERR_INVALID, std::string(what) + ": table offset beyond data_len");
    }
  return GD_OK;
}

void copy_sep(const gd_plan* p, const gd_separable* f, HostField& h, double default_scale)
{
  h = HostField{};
  if (!f) {
    h.rank = default_scale != 0.0 ? 1 : 0;
    h.scale.assign(1, default_scale);
    h.offset.assign(1, std::vector<int64_t>(GD_MAX_DIM, -1));
    return;
  }
  h.rank = f->rank;
  h.scale.assign(f->scale, f->scale + GD_MAX_TERMS);
  h.offset.assign(GD_MAX_TERMS, std::vector<int64_t>(GD_MAX_DIM, -1));
  for (int r = 0; r < f->rank; ++r)
    for (int d = 0; d < p->dim; ++d) h.offset[r][d] = f->offset[r][d];
  if (f->data && f->data_len > 0) h.data.assign(f->data, f->data + f->data_len);
}

void free_fields(gd_plan* p)
{
  for (void* t : p->field_tables) (void)hipFree(t);
  p->field_tables.clear();
  if (p->fields64) (void)hipFree(p->fields64);
  if (p->fields32) (void)hipFree(p->fields32);
  p->fields64 = p->fields32 = nullptr;
}

// Upload beta[0..dim), sigma, bc, diffusivity as FieldDev arrays (double and
// float).
int upload_fields(gd_plan* p)
{
  free_fields(p);
  const int nf = p->dim + 3;
  std::vector<FieldDev<double>> f64(nf);
  std::vector<FieldDev<float>> f32(nf);
  for (int i = 0; i < nf; ++i) {
    const HostField& h = i < p->dim ? p->beta_host[i]
                                     : (i == p->dim ? p->sigma_host : (i == p->dim + 1 ? p->bc_host : p->mu_host));
    FieldDev<double>& a = f64[i];
    FieldDev<float>& b = f32[i];
    std::memset(&a, 0, sizeof(a));
    std::memset(&b, 0, sizeof(b));
    a.rank = b.rank = h.rank;
    double* d64 = nullptr;
    float* d32 = nullptr;
    if (!h.data.empty()) {
      GD_HIP(hipMalloc(&d64, h.data.size() * sizeof(double)));
      p->field_tables.push_back(d64);
      GD_HIP(hipMalloc(&d32, h.data.size() * sizeof(float)));
      p->field_tables.push_back(d32);
      std::vector<float> h32(h.data.begin(), h.data.end());
      GD_HIP(hipMemcpy(d64, h.data.data(), h.data.size() * sizeof(double), hipMemcpyHostToDevice));
      GD_HIP(hipMemcpy(d32, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    for (int r = 0; r < h.rank; ++r) {
      a.scale[r] = h.scale[r];
      b.scale[r] = (float)h.scale[r];
      for (int d = 0; d < p->dim; ++d) {
        const int64_t off = h.offset[r][d];
        a.tab[r][d] = off >= 0 ? d64 + off : nullptr;
        b.tab[r][d] = off >= 0 ? d32 + off : nullptr;
      }
    }
  }
  p->fields64_host = f64;
  p->fields32_host = f32;
  GD_HIP(hipMalloc(&p->fields64, nf * sizeof(FieldDev<double>)));
  GD_HIP(hipMalloc(&p->fields32, nf * sizeof(FieldDev<float>)));
  GD_HIP(hipMemcpy(p->fields64, f64.data(), nf * sizeof(FieldDev<double>), hipMemcpyHostToDevice));
  GD_HIP(hipMemcpy(p->fields32, f32.data(), nf * sizeof(FieldDev<float>), hipMemcpyHostToDevice));
  return GD_OK;
}

// Every coefficient change ends here.  Applies queued earlier on any stream
// read the plan's device tables, so wait for the device to drain before they
// are replaced; then rebuild what the kernels read (separable field tables,
// the Kronecker kernel's factor tables, the p-adaptive parameter blocks), so
// the apply entry points never allocate or copy.
int commit_coefficients(gd_plan* p, bool fields)
{
  ++p->coef_version;
  GD_HIP(hipDeviceSynchronize());
  if (fields) {
    if (int rc = upload_fields(p)) return rc;
  }
  if (int rc = k3k::refresh(p)) return rc;
  return pa::refresh(p);
}

}  // namespace

const TableSet* find_tables(const gd_plan* p, int n, int q)
{
  for (const TableSet* t : p->tables)
    if (t->n == n && t->q == q) return t;
  return nullptr;
}

template <typename T>
Params<T> make_params(const gd_plan* p, const TableSet* ts, T mass_coef, T adv_coef)
{
  Params<T> k;
  std::memset(&k, 0, sizeof(k));
  const int D = p->dim;
  k.E = p->E;
  double detJ = 1.0;
  for (int d = 0; d < D; ++d) {
    k.M[d] = p->desc.sizes[d];
    k.h[d] = (T)p->desc.h[d];
    k.invh[d] = (T)(1.0 / p->desc.h[d]);
    k.origin[d] = (T)p->desc.origin[d];
    if (p->desc.periodic[d]) k.periodic_mask |= 1 << d;
    detJ *= p->desc.h[d];
  }
  k.detJ = (T)detJ;
  k.mass_coef = mass_coef;
  k.adv_coef = adv_coef;
  if constexpr (std::is_same_v<T, double>) {
    k.tab = ts->t64;
    k.beta = static_cast<const FieldDev<double>*>(p->fields64);
  } else {
    k.tab = ts->t32;
    k.beta = static_cast<const FieldDev<float>*>(p->fields32);
  }
  k.sigma = k.beta + D;
  k.bc = (p->bc_mode == GD_BC_INFLOW && p->bc_host.rank > 0) ? k.beta + D + 1 : nullptr;
  k.bc_mode = p->bc_mode;
  k.mu = k.beta + D + 2;
  k.ip_sigma = (T)p->ip_sigma;
  k.ip_kappa = (T)p->ip_kappa;
  {
    const int n = ts->n;
    bool unit = n > 0;
    for (int i = 0; i < n; ++i)
      unit = unit && ts->b0[i] == (i == 0 ? 1.0 : 0.0) && ts->b1[i] == (i == n - 1 ? 1.0 : 0.0);
    k.node_traces = unit ? 1 : 0;
  }
  k.beta_kind = p->beta_kind;
  for (int c = 0; c < D; ++c) {
    k.affc[c] = (T)p->affc[c];
    for (int d = 0; d < D; ++d) k.affA[c][d] = (T)p->affA[c][d];
  }
  if constexpr (std::is_same_v<T, double>) {
    k.beta_cell = p->beta_cell;
    for (int d = 0; d < D; ++d) k.beta_face[d] = p->beta_face[d];
  }
  return k;
}

template Params<double> make_params<double>(const gd_plan*, const TableSet*, double, double);
template Params<float> make_params<float>(const gd_plan*, const TableSet*, float, float);

}  // namespace gd

gd_plan::~gd_plan()
{
  for (gd::TableSet* t : tables) {
    if (t->dev64) (void)hipFree(t->dev64);
    if (t->dev32) (void)hipFree(t->dev32);
    delete t;
  }
  if (dof_offset_dev) (void)hipFree(dof_offset_dev);
  if (elem_order_dev) (void)hipFree(elem_order_dev);
  if (elem_quad_dev) (void)hipFree(elem_quad_dev);
  gd::free_fields(this);
  gd::pa::destroy_state(pa_state);
  gd::k3k::destroy_state(k3k_state);
  if (hmap_dev) (void)hipFree(hmap_dev);
}

using gd::fail;

extern "C" {

const char* gd_last_error(void) { return gd::g_last_error.c_str(); }

int gd_version(void) { return 1; }

int gd_rule_points(int32_t basis, int32_t npts, double* pts, double* wts)
{
  if (npts < 1 || npts > 64 || !pts || !wts) return fail(GD_ERR_INVALID, "gd_rule_points: bad arguments");
  if (basis == GD_BASIS_GLL) gd::rule_gll(npts, pts, wts);
  else if (basis == GD_BASIS_GL) gd::rule_gl(npts, pts, wts);
  else return fail(GD_ERR_INVALID, "gd_rule_points: unknown basis");
  return GD_OK;
}

int gd_plan_create(const gd_desc* desc, gd_plan** out)
{
  if (!desc || !out) return fail(GD_ERR_INVALID, "gd_plan_create: null argument");
  *out = nullptr;
  const int D = desc->dim;
  if (D < 1 || D > GD_MAX_DIM) return fail(GD_ERR_INVALID, "gd_plan_create: dim must be 1..6");
  if (desc->basis != GD_BASIS_GLL && desc->basis != GD_BASIS_GL)
    return fail(GD_ERR_INVALID, "gd_plan_create: unknown basis");
  auto p = new gd_plan();
  p->desc = *desc;
  p->dim = D;
  p->E = 1;
  for (int d = 0; d < D; ++d) {
    if (desc->sizes[d] < 1 || !(desc->h[d] > 0.0)) {
      delete p;
      return fail(GD_ERR_INVALID, "gd_plan_create: sizes must be >= 1 and h > 0");
    }
    p->E *= desc->sizes[d];
  }
  if (p->E > (int64_t)1 << 40) {
    delete p;
    return fail(GD_ERR_INVALID, "gd_plan_create: too many elements");
  }
  std::set<int> ns, qs;
  p->padapt = desc->elem_order != nullptr;
  if (p->padapt) {
    if (!desc->elem_quad) {
      delete p;
      return fail(GD_ERR_INVALID, "gd_plan_create: elem_order needs elem_quad");
    }
    p->elem_order.assign(desc->elem_order, desc->elem_order + p->E);
    p->elem_quad.assign(desc->elem_quad, desc->elem_quad + p->E);
    p->dof_offset_host.resize(p->E + 1);
    int64_t off = 0;
    for (int64_t e = 0; e < p->E; ++e) {
      const int o = p->elem_order[e], q = p->elem_quad[e];
      if (o < 0 || o + 1 > gd::kMaxOrderNodes || q < 1 || q > gd::kMaxQuad) {
        delete p;
        return fail(GD_ERR_INVALID, "gd_plan_create: per-element order/rule out of range");
      }
      ns.insert(o + 1);
      qs.insert(q);
      p->dof_offset_host[e] = off;
      int64_t nd = 1;
      for (int d = 0; d < D; ++d) nd *= (o + 1);
      off += nd;
    }
    p->dof_offset_host[p->E] = off;
    p->ndofs = off;
  } else {
    if (desc->order < 0 || desc->order + 1 > gd::kMaxOrderNodes || desc->num_quad < 1 ||
        desc->num_quad > gd::kMaxQuad) {
      delete p;
      return fail(GD_ERR_INVALID, "gd_plan_create: order must be 0..7 and num_quad 1..8");
    }
    ns.insert(desc->order + 1);
    qs.insert(desc->num_quad);
    int64_t nd = 1;
    for (int d = 0; d < D; ++d) nd *= (desc->order + 1);
    p->ndofs = p->E * nd;
  }
  // tables for every (n, q) pair that can occur (p-adaptive faces use the
  // larger of the two cell rules)
  for (int n : ns)
    for (int q : qs) {
      auto* ts = new gd::TableSet();
      p->tables.push_back(ts);
      const int rc = gd::build_tables(desc->basis, n, q, *ts);
      if (rc != GD_OK) {
        delete p;
        return rc;
      }
    }
  if (p->padapt) {
    hipError_t e1 = hipMalloc(&p->dof_offset_dev, (p->E + 1) * sizeof(int64_t));
    hipError_t e2 = hipMalloc(&p->elem_order_dev, p->E);
    hipError_t e3 = hipMalloc(&p->elem_quad_dev, p->E);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      delete p;
      return fail(GD_ERR_ALLOC, "gd_plan_create: device allocation failed");
    }
    (void)hipMemcpy(p->dof_offset_dev, p->dof_offset_host.data(), (p->E + 1) * sizeof(int64_t),
                    hipMemcpyHostToDevice);
    (void)hipMemcpy(p->elem_order_dev, p->elem_order.data(), p->E, hipMemcpyHostToDevice);
    (void)hipMemcpy(p->elem_quad_dev, p->elem_quad.data(), p->E, hipMemcpyHostToDevice);
  }
  // defaults: sigma = 1, beta = 0, no boundary terms
  gd::copy_sep(p, nullptr, p->sigma_host, 1.0);
  for (int d = 0; d < D; ++d) gd::copy_sep(p, nullptr, p->beta_host[d], 0.0);
  gd::copy_sep(p, nullptr, p->bc_host, 0.0);
  int rc = gd::upload_fields(p);
  if (rc == GD_OK) rc = gd::k3k::refresh(p);
  if (rc == GD_OK) rc = gd::pa::refresh(p);
  if (rc != GD_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return GD_OK;
}

int gd_plan_destroy(gd_plan* plan)
{
  delete plan;
  return GD_OK;
}

int64_t gd_plan_num_dofs(const gd_plan* plan) { return plan ? plan->ndofs : -1; }
int64_t gd_plan_num_elements(const gd_plan* plan) { return plan ? plan->E : -1; }
const int64_t* gd_plan_dof_offsets(const gd_plan* plan) { return plan ? plan->dof_offset_dev : nullptr; }

int gd_plan_axis_points(const gd_plan* plan, int32_t axis, double* out)
{
  if (!plan || !out || axis < 0 || axis >= plan->dim) return fail(GD_ERR_INVALID, "gd_plan_axis_points: bad arguments");
  if (plan->padapt) return fail(GD_ERR_UNSUPPORTED, "gd_plan_axis_points: p-adaptive plans have no uniform rule");
  const gd::TableSet* ts = plan->tables.front();
  const int q = ts->q;
  const double o = plan->desc.origin[axis], h = plan->desc.h[axis];
  for (int64_t e = 0; e < plan->desc.sizes[axis]; ++e) {
    const double c = o + h * (double)e;
    for (int k = 0; k < q; ++k) out[e * (q + 2) + k] = c + h * ts->xq[k];
    out[e * (q + 2) + q] = c + h * 0.0;
    out[e * (q + 2) + q + 1] = c + h * 1.0;
  }
  return GD_OK;
}

int gd_plan_set_sigma(gd_plan* plan, const gd_separable* sigma)
{
  if (!plan) return fail(GD_ERR_INVALID, "gd_plan_set_sigma: null plan");
  int rc = gd::check_sep(plan, sigma, "sigma", false);
  if (rc) return rc;
  gd::copy_sep(plan, sigma, plan->sigma_host, 1.0);
  return gd::commit_coefficients(plan, true);
}

int gd_plan_set_diffusion(gd_plan* plan, const gd_separable* mu, double sigma, double kappa)
{
  if (!plan) return fail(GD_ERR_INVALID, "gd_plan_set_diffusion: null plan");
  if (plan->padapt) return fail(GD_ERR_UNSUPPORTED, "gd_plan_set_diffusion: p-adaptive plans have no diffusion apply");
  if (!std::isfinite(sigma) || !std::isfinite(kappa))
    return fail(GD_ERR_INVALID, "gd_plan_set_diffusion: sigma and kappa must be finite");
  int rc = gd::check_sep(plan, mu, "mu", false);
  if (rc) return rc;
  gd::copy_sep(plan, mu, plan->mu_host, 1.0);
  plan->ip_sigma = sigma;
  plan->ip_kappa = kappa;
  plan->diff_set = true;
  return gd::commit_coefficients(plan, true);
}

int gd_plan_set_beta_separable(gd_plan* plan, const gd_separable* beta)
{
  if (!plan || !beta) return fail(GD_ERR_INVALID, "gd_plan_set_beta_separable: null argument");
  for (int d = 0; d < plan->dim; ++d) {
    int rc = gd::check_sep(plan, &beta[d], "beta", false);
    if (rc) return rc;
  }
  for (int d = 0; d < plan->dim; ++d) gd::copy_sep(plan, &beta[d], plan->beta_host[d], 0.0);
  plan->beta_kind = GD_BETA_SEPARABLE;
  plan->beta_set = true;
  return gd::commit_coefficients(plan, true);
}

int gd_plan_set_beta_affine(gd_plan* plan, const double* A, const double* c)
{
  if (!plan || !A || !c) return fail(GD_ERR_INVALID, "gd_plan_set_beta_affine: null argument");
  const int D = plan->dim;
  for (int i = 0; i < D; ++i) {
    plan->affc[i] = c[i];
    for (int j = 0; j < D; ++j) plan->affA[i][j] = A[i * D + j];
  }
  plan->beta_kind = GD_BETA_AFFINE;
  plan->beta_set = true;
  return gd::commit_coefficients(plan, false);
}

int gd_plan_set_beta_pointwise_f64(gd_plan* plan, const double* beta_cell, const double* const* beta_face)
{
  if (!plan || !beta_cell || !beta_face) return fail(GD_ERR_INVALID, "gd_plan_set_beta_pointwise_f64: null argument");
  if (plan->padapt) return fail(GD_ERR_UNSUPPORTED, "pointwise beta needs a uniform rule");
  plan->beta_cell = beta_cell;
  for (int d = 0; d < plan->dim; ++d) {
    if (!beta_face[d]) return fail(GD_ERR_INVALID, "gd_plan_set_beta_pointwise_f64: null face array");
    plan->beta_face[d] = beta_face[d];
  }
  plan->beta_kind = GD_BETA_POINTWISE;
  plan->beta_set = true;
  return gd::commit_coefficients(plan, false);
}

int gd_plan_set_bc(gd_plan* plan, int32_t mode, const gd_separable* g)
{
  if (!plan) return fail(GD_ERR_INVALID, "gd_plan_set_bc: null plan");
  if (mode != GD_BC_NONE && mode != GD_BC_INFLOW) return fail(GD_ERR_INVALID, "gd_plan_set_bc: unknown mode");
  int rc = gd::check_sep(plan, g, "bc", false);
  if (rc) return rc;
  plan->bc_mode = mode;
  gd::copy_sep(plan, g, plan->bc_host, 0.0);
  return gd::commit_coefficients(plan, true);
}

int gd_malloc(void** ptr, size_t bytes)
{
  if (!ptr) return fail(GD_ERR_INVALID, "gd_malloc: null");
  if (hipMalloc(ptr, bytes) != hipSuccess) return fail(GD_ERR_ALLOC, "gd_malloc: hipMalloc failed");
  return GD_OK;
}

int gd_free(void* ptr)
{
  GD_HIP(hipFree(ptr));
  return GD_OK;
}

int gd_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
  GD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GD_OK;
}

int gd_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
  GD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GD_OK;
}

int gd_memset(void* dst, int value, size_t bytes, gd_stream stream)
{
  GD_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  return GD_OK;
}

int gd_stream_sync(gd_stream stream)
{
  GD_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GD_OK;
}

int gd_stream_query(gd_stream stream, int32_t* done)
{
  if (!done) return fail(GD_ERR_INVALID, "gd_stream_query: null");
  const hipError_t e = hipStreamQuery((hipStream_t)stream);
  *done = e == hipSuccess ? 1 : 0;
  if (e != hipErrorNotReady) GD_HIP(e);
  return GD_OK;
}

int gd_device_sync(void)
{
  GD_HIP(hipDeviceSynchronize());
  return GD_OK;
}

int gd_host_alloc(void** ptr, size_t bytes)
{
  if (!ptr) return fail(GD_ERR_INVALID, "gd_host_alloc: null");
  // coherent: device stores become visible to host reads without a flush
  if (hipHostMalloc(ptr, bytes, hipHostMallocCoherent) != hipSuccess)
    return fail(GD_ERR_ALLOC, "gd_host_alloc: hipHostMalloc failed");
  return GD_OK;
}

int gd_host_free(void* ptr)
{
  GD_HIP(hipHostFree(ptr));
  return GD_OK;
}

int gd_event_create(gd_event* ev)
{
  if (!ev) return fail(GD_ERR_INVALID, "gd_event_create: null");
  hipEvent_t e = nullptr;
  GD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *ev = (gd_event)e;
  return GD_OK;
}

int gd_event_destroy(gd_event ev)
{
  GD_HIP(hipEventDestroy((hipEvent_t)ev));
  return GD_OK;
}

int gd_event_record(gd_event ev, gd_stream stream)
{
  GD_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return GD_OK;
}

int gd_event_sync(gd_event ev)
{
  GD_HIP(hipEventSynchronize((hipEvent_t)ev));
  return GD_OK;
}

}  // extern "C"
