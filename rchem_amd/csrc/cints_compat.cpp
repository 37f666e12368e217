// The libpyquante2 functions rchem binds through bindgen (cints.h:14-30,
// 40-54, 69), with their exact C signatures, served by this library's GPU
// kernels so that rchem's tho66.rs wrappers and tests link unchanged:
//   coulomb_repulsion / contr_coulomb: the library's HGP shell-quartet kernels
//     (rchem_eri_shell_quartets) with the reference's Boys function, on
//     single- or multi-primitive shells whose coefficients carry the caller's
//     norms in place of the library's own normalisation;
//   overlap, kinetic, nuclear_attraction, Fgamma: the primitive kernel of
//     cints_compat.hip.
// All calls share one lazily created device context (device 0) and are not
// thread-safe, like the single-threaded caller they serve.  On failure a
// function returns NaN and rchem_last_error(NULL) holds the reason.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/rchem_cints_compat.h"
#include "../../include/rchem_gpu.h"

namespace rchem {
struct CompatRec {
  int kind;
  int l1[3], l2[3];
  double a1, a2;
  double A[3], B[3], C[3];
  double m, x;
  double out;
};
hipError_t compat_launch(CompatRec* recs, int n, hipStream_t s);
int compat_max_power();
int compat_max_nuclear_l();
void set_global_error(const std::string& msg);
}  // namespace rchem

namespace {

rchem_ctx* g_ctx = nullptr;
rchem::CompatRec* g_rec = nullptr;

const double kNaN = std::numeric_limits<double>::quiet_NaN();

bool ensure_ctx() {
  if (g_ctx) return true;
  return rchem_ctx_create(0, &g_ctx) == RCHEM_OK;
}

double fail(const std::string& msg) {
  rchem::set_global_error(msg);
  return kNaN;
}

// radial part of the primitive norm of basis.rs:140-149 for total power L
double nrad(double a, int L) {
  return std::sqrt(std::pow(2.0, 2.0 * L + 1.5) * std::pow(a, L + 1.5) / std::pow(M_PI, 1.5));
}

double comp_norm(int l, int m, int n) {
  auto f2 = [](int k) {
    double r = 1.0;
    for (int i = k; i > 1; i -= 2) r *= i;
    return r;
  };
  return 1.0 / std::sqrt(f2(2 * l - 1) * f2(2 * m - 1) * f2(2 * n - 1));
}

// index of component (l, m, n) in rchem's order for its shell (shell.rs:1-12)
int comp_index(int l, int m, int n) {
  const int L = l + m + n;
  int idx = 0;
  for (int a = 1; a <= L + 1; ++a)
    for (int b = 1; b <= a; ++b) {
      if (L + 1 - a == l && a - b == m && b - 1 == n) return idx;
      ++idx;
    }
  return -1;
}

int ncart(int L) { return (L + 1) * (L + 2) / 2; }

struct Fn {
  int len;
  const double* exps;
  const double* coefs;  // may be null: coefficient 1
  const double* norms;
  double xyz[3];
  int p[3];
};

double contracted(const Fn f[4]) {
  if (!ensure_ctx()) return fail("no gfx950 device for the cints compatibility functions");
  double xyz[12];
  int ls[4], np[4];
  std::vector<double> ex, cf;
  for (int k = 0; k < 4; ++k) {
    const int L = f[k].p[0] + f[k].p[1] + f[k].p[2];
    if (f[k].p[0] < 0 || f[k].p[1] < 0 || f[k].p[2] < 0 || L > 2)
      return fail("cints compatibility ERIs serve powers with l + m + n <= 2");
    if (f[k].len < 1 || f[k].len > 8) return fail("1 to 8 primitives per function");
    ls[k] = L;
    np[k] = f[k].len;
    for (int x = 0; x < 3; ++x) xyz[3 * k + x] = f[k].xyz[x];
    const double cn = comp_norm(f[k].p[0], f[k].p[1], f[k].p[2]);
    for (int i = 0; i < f[k].len; ++i) {
      const double a = f[k].exps[i];
      const double c = (f[k].coefs ? f[k].coefs[i] : 1.0) * f[k].norms[i];
      ex.push_back(a);
      cf.push_back(c / (nrad(a, L) * cn));
    }
  }
  if (rchem_basis_upload(g_ctx, 4, xyz, ls, np, ex.data(), cf.data()) != RCHEM_OK)
    return fail(rchem_last_error(g_ctx));
  rchem_opts o;
  rchem_opts_default(&o);
  o.boys_mode = RCHEM_BOYS_COMPAT;
  o.schwarz_thresh = 0.0;
  const int q[4] = {0, 1, 2, 3};
  std::vector<double> blk(static_cast<size_t>(ncart(ls[0])) * ncart(ls[1]) * ncart(ls[2]) *
                          ncart(ls[3]));
  if (rchem_eri_shell_quartets(g_ctx, 1, q, blk.data(), &o) != RCHEM_OK)
    return fail(rchem_last_error(g_ctx));
  int ci[4];
  for (int k = 0; k < 4; ++k) ci[k] = comp_index(f[k].p[0], f[k].p[1], f[k].p[2]);
  return blk[((static_cast<size_t>(ci[0]) * ncart(ls[1]) + ci[1]) * ncart(ls[2]) + ci[2]) *
                 ncart(ls[3]) + ci[3]];
}

double one_body(rchem::CompatRec r) {
  if (!ensure_ctx()) return fail("no gfx950 device for the cints compatibility functions");
  const int pmax = rchem::compat_max_power();
  int total = 0;
  for (int x = 0; x < 3; ++x) {
    if (r.l1[x] < 0 || r.l2[x] < 0 || r.l1[x] > pmax || r.l2[x] > pmax)
      return fail("cints compatibility one-electron integrals serve powers 0..6 per axis");
    total += r.l1[x] + r.l2[x];
  }
  if (r.kind == 2 && total > rchem::compat_max_nuclear_l())
    return fail("nuclear attraction serves l1 + l2 <= 8 (summed over axes)");
  if (!g_rec && hipMalloc(&g_rec, sizeof(rchem::CompatRec)) != hipSuccess)
    return fail("device allocation failed");
  if (hipMemcpy(g_rec, &r, sizeof(r), hipMemcpyHostToDevice) != hipSuccess ||
      rchem::compat_launch(g_rec, 1, nullptr) != hipSuccess ||
      hipMemcpy(&r, g_rec, sizeof(r), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("compatibility kernel failed");
  return r.out;
}

rchem::CompatRec rec(int kind, double a1, int l1, int m1, int n1, double xa, double ya, double za,
                     double a2, int l2, int m2, int n2, double xb, double yb, double zb) {
  rchem::CompatRec r{};
  r.kind = kind;
  r.a1 = a1;
  r.a2 = a2;
  r.l1[0] = l1; r.l1[1] = m1; r.l1[2] = n1;
  r.l2[0] = l2; r.l2[1] = m2; r.l2[2] = n2;
  r.A[0] = xa; r.A[1] = ya; r.A[2] = za;
  r.B[0] = xb; r.B[1] = yb; r.B[2] = zb;
  return r;
}

}  // namespace

extern "C" {

double coulomb_repulsion(double xa, double ya, double za, double norma, int la, int ma, int na,
                         double alphaa, double xb, double yb, double zb, double normb, int lb,
                         int mb, int nb, double alphab, double xc, double yc, double zc,
                         double normc, int lc, int mc, int nc, double alphac, double xd,
                         double yd, double zd, double normd, int ld, int md, int nd,
                         double alphad) {
  const Fn f[4] = {{1, &alphaa, nullptr, &norma, {xa, ya, za}, {la, ma, na}},
                   {1, &alphab, nullptr, &normb, {xb, yb, zb}, {lb, mb, nb}},
                   {1, &alphac, nullptr, &normc, {xc, yc, zc}, {lc, mc, nc}},
                   {1, &alphad, nullptr, &normd, {xd, yd, zd}, {ld, md, nd}}};
  return contracted(f);
}

double contr_coulomb(int ia, double* aexps, double* acoefs, double* anorms, double xa, double ya,
                     double za, int la, int ma, int na, int ib, double* bexps, double* bcoefs,
                     double* bnorms, double xb, double yb, double zb, int lb, int mb, int nb,
                     int ic, double* cexps, double* ccoefs, double* cnorms, double xc, double yc,
                     double zc, int lc, int mc, int nc, int id, double* dexps, double* dcoefs,
                     double* dnorms, double xd, double yd, double zd, int ld, int md, int nd) {
  const Fn f[4] = {{ia, aexps, acoefs, anorms, {xa, ya, za}, {la, ma, na}},
                   {ib, bexps, bcoefs, bnorms, {xb, yb, zb}, {lb, mb, nb}},
                   {ic, cexps, ccoefs, cnorms, {xc, yc, zc}, {lc, mc, nc}},
                   {id, dexps, dcoefs, dnorms, {xd, yd, zd}, {ld, md, nd}}};
  return contracted(f);
}

double overlap(double alpha1, int l1, int m1, int n1, double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2, double xb, double yb, double zb) {
  return one_body(rec(0, alpha1, l1, m1, n1, xa, ya, za, alpha2, l2, m2, n2, xb, yb, zb));
}

double kinetic(double alpha1, int l1, int m1, int n1, double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2, double xb, double yb, double zb) {
  return one_body(rec(1, alpha1, l1, m1, n1, xa, ya, za, alpha2, l2, m2, n2, xb, yb, zb));
}

double nuclear_attraction(double x1, double y1, double z1, double norm1, int l1, int m1, int n1,
                          double alpha1, double x2, double y2, double z2, double norm2, int l2,
                          int m2, int n2, double alpha2, double x3, double y3, double z3) {
  rchem::CompatRec r = rec(2, alpha1, l1, m1, n1, x1, y1, z1, alpha2, l2, m2, n2, x2, y2, z2);
  r.C[0] = x3;
  r.C[1] = y3;
  r.C[2] = z3;
  return norm1 * norm2 * one_body(r);
}

double Fgamma(double m, double x) {
  rchem::CompatRec r{};
  r.kind = 3;
  r.m = m;
  r.x = x;
  return one_body(r);
}

}  // extern "C"
