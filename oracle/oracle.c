/*
 * oracle.c -- CPU restatement of berquist/rchem's restricted-Hartree-Fock hot
 * path, used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE.  The shipped library (rchem_amd/, librchem_gpu.so)
 * never links, loads or calls this code.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py use it, and only to produce the values
 * the GPU results are compared against (or, for the baseline, to time the
 * reference algorithm on the host cores).
 *
 * Every function names the reference file:line it restates.  The reference
 * itself is Rust (rchem) on top of C (libpyquante2); the C parts are restated
 * here from their published formulas, the Rust parts (loop order, normalisation,
 * energy expression) from their source.
 *
 * Pinning (how each function is checked before it is trusted):
 *   orc_fgamma_compat     against oracle/_ref (the reference cints.c, compiled
 *                         from /root/reference) on a grid of (m, x), and
 *                         against the exact Boys function to within the NR
 *                         series tolerance (EPS = 3e-7).
 *   orc_boys_exact        against closed forms: F_m(0) = 1/(2m+1),
 *                         F_0(x) = sqrt(pi/x) erf(sqrt x)/2, the downward
 *                         recursion identity, and mpmath (hyp1f1) values.
 *   orc_coulomb_repulsion against the reference golden (ss|ss) =
 *                         0.08608517834596989 (tho66.rs:302) and the os86
 *                         goldens (os86.rs:896,910,924), against oracle/_ref
 *                         on random primitives with l <= 2, and against the
 *                         independent HGP restatement orc_hgp_primitive.
 *   orc_hgp_primitive     against orc_coulomb_repulsion and oracle/_ref's
 *                         vrr/hrr (chgp.c).
 *   orc_overlap/kinetic/nuclear_attraction  against the os86/tho66 goldens
 *                         (os86.rs:949-1001, tho66.rs:276-285) and oracle/_ref.
 *   matrix / JK / tensor builders: by construction from the above, plus
 *                         invariants checked in tests (symmetry, Tr(DS), the
 *                         build_I / JK_inmem identity, E_nuc closed form).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Numerical-Recipes constants as fixed by the reference (cints.c:28-31). */
#define ORC_ITMAX 100
#define ORC_EPS 3.0e-7
#define ORC_FPMIN 1.0e-30
#define ORC_SMALL 1.0e-8

enum { ORC_BOYS_EXACT = 0, ORC_BOYS_COMPAT = 1 };

/* ------------------------------------------------------------------------ */
/* Integer helpers with the reference's semantics.                           */
/* ------------------------------------------------------------------------ */

/* n! for n >= 0, 1 for n <= 1 (cints.c:272-275). */
static double ifact(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}

/* (n)!! with 1 for n <= 1 (cints.c:277-280; rchem's basis.rs:108-114 returns
 * 1 for n <= 0, which agrees for every argument used, 2l-1 >= -1). */
static double ifact2(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

/* C(a, b) (cints.c:300). */
static double binom(int a, int b) {
  if (b < 0 || b > a) return 0.0;
  return ifact(a) / (ifact(b) * ifact(a - b));
}

/* a! / (b! (a-2b)!) (cints.c:389). */
static double fact_ratio2(int a, int b) {
  return ifact(a) / ifact(b) / ifact(a - 2 * b);
}

/* Integer power with the exponent's sign handled, x^0 = 1. */
static double ipow(double x, int n) {
  if (n == 0) return 1.0;
  if (n < 0) return 1.0 / ipow(x, -n);
  double r = 1.0;
  for (int k = 0; k < n; ++k) r *= x;
  return r;
}

/* Coefficient of x^s in (x + xpa)^ia (x + xpb)^ib (cints.c:291-298). */
static double binomial_prefactor(int s, int ia, int ib, double xpa, double xpb) {
  double sum = 0.0;
  for (int t = 0; t <= s; ++t)
    if (s - ia <= t && t <= ib)
      sum += binom(ia, s - t) * binom(ib, t) * ipow(xpa, ia - s + t) * ipow(xpb, ib - t);
  return sum;
}

/* ------------------------------------------------------------------------ */
/* Boys function.                                                            */
/* ------------------------------------------------------------------------ */

/* Series branch of the regularised lower incomplete gamma P(a, x), Numerical
 * Recipes `gser`, as the reference uses it (cints.c:329-348).  Returns
 * exp(gln) * gamser with the same operation order as gamm_inc (cints.c:316). */
static double compat_gser(double a, double x) {
  double gln = lgamma(a);
  if (x <= 0.0) return 0.0;
  double ap = a, del = 1.0 / a, sum = del;
  for (int n = 1; n <= ORC_ITMAX; ++n) {
    ap += 1.0;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * ORC_EPS) {
      double gamser = sum * exp(-x + a * log(x) - gln);
      return exp(gln) * gamser;
    }
  }
  /* The reference prints and returns an uninitialised value here
   * (cints.c:345-346); unreachable for a <= 16.5 and x < a + 1. */
  return NAN;
}

/* Continued-fraction branch, NR `gcf` with the modified Lentz method
 * (cints.c:351-373), returning exp(gln) * (1 - gammcf) (cints.c:319-320). */
static double compat_gcf(double a, double x) {
  double gln = lgamma(a);
  double b = x + 1.0 - a, c = 1.0 / ORC_FPMIN, d = 1.0 / b, h = d;
  int i;
  for (i = 1; i <= ORC_ITMAX; ++i) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < ORC_FPMIN) d = ORC_FPMIN;
    c = b + an / c;
    if (fabs(c) < ORC_FPMIN) c = ORC_FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < ORC_EPS) break;
  }
  double gammcf = exp(-x + a * log(x) - gln) * h;
  return exp(gln) * (1.0 - gammcf);
}

/* The reference's Boys function Fgamma(m, x) = x^(-m-1/2) gamma(m+1/2, x) / 2
 * (cints.c:302-308): |x| < 1e-8 is clamped to 1e-8, the series is used for
 * x < m + 3/2, the continued fraction otherwise (cints.c:315-321).  Its
 * truncation error (~EPS relative on the last series term) is reproduced. */
double orc_fgamma_compat(double m, double x) {
  if (fabs(x) < ORC_SMALL) x = ORC_SMALL;
  double a = m + 0.5;
  double val = (x < a + 1.0) ? compat_gser(a, x) : compat_gcf(a, x);
  return 0.5 * pow(x, -m - 0.5) * val;
}

/* Exact Boys function F_m(x) = int_0^1 t^(2m) exp(-x t^2) dt, to ~1e-16
 * relative.  For x < 120 the all-positive series
 *   F_m(x) = exp(-x) sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1))
 * is summed in long double; beyond, exp(-x) < 1e-52 and
 *   F_m(x) = (2m-1)!! / (2x)^m * sqrt(pi/x) / 2
 * is exact to double precision.  This is the accurate function rchem's V uses
 * (boys::micb25 at os86.rs:770), not the NR one of cints.c. */
double orc_boys_exact(int m, double x) {
  if (x < 0.0) return NAN;
  if (x < 120.0) {
    long double lx = x, term = 1.0L / (2 * m + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
      term *= 2.0L * lx / (2 * m + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-21L) break;
    }
    return (double)(expl(-lx) * sum);
  }
  long double f = 0.5L * sqrtl(3.14159265358979323846264338327950288L / x);
  for (int k = 0; k < m; ++k) f *= (2 * k + 1) / (2.0L * x);
  return (double)f;
}

static double boys(int mode, int m, double x) {
  return mode == ORC_BOYS_COMPAT ? orc_fgamma_compat((double)m, x) : orc_boys_exact(m, x);
}

/* ------------------------------------------------------------------------ */
/* THO primitive electron-repulsion integral (cints.c:72-149).                */
/* ------------------------------------------------------------------------ */

static double product_center(double a, double xa, double b, double xb) {
  return (a * xa + b * xb) / (a + b); /* cints.c:391-394 */
}

static double dist2(const double *a, const double *b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz; /* cints.c:282-285 */
}

/* f_B of THO eq. 2.22 (cints.c:33-40). */
static double tho_fB(int i, int l1, int l2, double p, double a, double b, int r, double g) {
  return binomial_prefactor(i, l1, l2, p - a, p - b) * fact_ratio2(i, r) * ipow(4.0 * g, r - i);
}

/* One Cartesian axis of the THO B coefficients (cints.c:117-149), written into
 * B[0 .. l1+l2+l3+l4]. */
static void tho_B_axis(double *B, int l1, int l2, int l3, int l4, double p, double a, double b,
                       double q, double c, double d, double g1, double g2, double delta) {
  int imax = l1 + l2 + l3 + l4 + 1;
  for (int i = 0; i < imax; ++i) B[i] = 0.0;
  for (int i1 = 0; i1 < l1 + l2 + 1; ++i1)
    for (int i2 = 0; i2 < l3 + l4 + 1; ++i2)
      for (int r1 = 0; r1 < i1 / 2 + 1; ++r1)
        for (int r2 = 0; r2 < i2 / 2 + 1; ++r2)
          for (int u = 0; u < (i1 + i2) / 2 - r1 - r2 + 1; ++u) {
            int e = i1 + i2 - 2 * (r1 + r2);
            int I = e - u;
            double t = tho_fB(i1, l1, l2, p, a, b, r1, g1) * ((i2 & 1) ? -1.0 : 1.0) *
                       tho_fB(i2, l3, l4, q, c, d, r2, g2) * ((u & 1) ? -1.0 : 1.0) *
                       fact_ratio2(e, u) * ipow(q - p, e - 2 * u) / ipow(delta, e - u);
            B[I] += t;
          }
}

/* Primitive ERI (ab|cd) times the four primitive norms, THO (cints.c:72-115).
 * Centres r[12] = A,B,C,D; powers pw[12] = (l,m,n) of a,b,c,d; exponents al[4];
 * norms nm[4]. */
double orc_coulomb_repulsion(const double *r, const int *pw, const double *al, const double *nm,
                             int boys_mode) {
  const double *A = r, *Bc = r + 3, *C = r + 6, *D = r + 9;
  double g1 = al[0] + al[1], g2 = al[2] + al[3];
  double P[3], Q[3];
  for (int k = 0; k < 3; ++k) {
    P[k] = product_center(al[0], A[k], al[1], Bc[k]);
    Q[k] = product_center(al[2], C[k], al[3], D[k]);
  }
  double rab2 = dist2(A, Bc), rcd2 = dist2(C, D), rpq2 = dist2(P, Q);
  double delta = (1.0 / g1 + 1.0 / g2) / 4.0;
  double Bx[32], By[32], Bz[32];
  tho_B_axis(Bx, pw[0], pw[3], pw[6], pw[9], P[0], A[0], Bc[0], Q[0], C[0], D[0], g1, g2, delta);
  tho_B_axis(By, pw[1], pw[4], pw[7], pw[10], P[1], A[1], Bc[1], Q[1], C[1], D[1], g1, g2, delta);
  tho_B_axis(Bz, pw[2], pw[5], pw[8], pw[11], P[2], A[2], Bc[2], Q[2], C[2], D[2], g1, g2, delta);
  int nx = pw[0] + pw[3] + pw[6] + pw[9] + 1;
  int ny = pw[1] + pw[4] + pw[7] + pw[10] + 1;
  int nz = pw[2] + pw[5] + pw[8] + pw[11] + 1;
  double T = 0.25 * rpq2 / delta;
  double F[64];
  for (int m = 0; m < nx + ny + nz - 2; ++m) F[m] = boys(boys_mode, m, T);
  double sum = 0.0;
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < nz; ++K) sum += Bx[I] * By[J] * Bz[K] * F[I + J + K];
  return 2.0 * pow(M_PI, 2.5) / (g1 * g2 * sqrt(g1 + g2)) * exp(-al[0] * al[1] * rab2 / g1) *
         exp(-al[2] * al[3] * rcd2 / g2) * sum * nm[0] * nm[1] * nm[2] * nm[3];
}

/* ------------------------------------------------------------------------ */
/* HGP primitive ERI via VRR + HRR (chgp.c:138-219, 412-586), an independent  */
/* second restatement that cross-checks the THO one.                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  double A[3], C[3], P[3], Q[3], W[3];
  double zeta, eta, pref; /* pref = Kab Kcd / sqrt(zeta + eta) (chgp.c:576-584) */
  double F[64];
} hgp_ctx;

/* [a 0 | c 0]^(m), lowering c first (z, y, x), then a (chgp.c:438-575). */
static double hgp_vrr(const hgp_ctx *h, const int *a, const int *c, int m) {
  int ax = -1;
  for (int k = 2; k >= 0; --k)
    if (c[k] > 0) { ax = k; break; }
  if (ax >= 0) {
    int c1[3] = {c[0], c[1], c[2]}, c2[3], a1[3];
    c1[ax] -= 1;
    double v = (h->Q[ax] - h->C[ax]) * hgp_vrr(h, a, c1, m) +
               (h->W[ax] - h->Q[ax]) * hgp_vrr(h, a, c1, m + 1);
    if (c1[ax] > 0) {
      memcpy(c2, c1, sizeof c2);
      c2[ax] -= 1;
      v += 0.5 * c1[ax] / h->eta *
           (hgp_vrr(h, a, c2, m) - h->zeta / (h->zeta + h->eta) * hgp_vrr(h, a, c2, m + 1));
    }
    if (a[ax] > 0) {
      memcpy(a1, a, sizeof a1);
      a1[ax] -= 1;
      v += 0.5 * a[ax] / (h->zeta + h->eta) * hgp_vrr(h, a1, c1, m + 1);
    }
    return v;
  }
  for (int k = 2; k >= 0; --k)
    if (a[k] > 0) { ax = k; break; }
  if (ax >= 0) {
    int a1[3] = {a[0], a[1], a[2]}, a2[3];
    a1[ax] -= 1;
    double v = (h->P[ax] - h->A[ax]) * hgp_vrr(h, a1, c, m) +
               (h->W[ax] - h->P[ax]) * hgp_vrr(h, a1, c, m + 1);
    if (a1[ax] > 0) {
      memcpy(a2, a1, sizeof a2);
      a2[ax] -= 1;
      v += 0.5 * a1[ax] / h->zeta *
           (hgp_vrr(h, a2, c, m) - h->eta / (h->zeta + h->eta) * hgp_vrr(h, a2, c, m + 1));
    }
    return v;
  }
  return h->pref * h->F[m];
}

/* (ab|cd) from [e0|f0] by the horizontal recurrence, b -> a in x,y,z order,
 * then d -> c (chgp.c:138-208). */
static double hgp_hrr(const hgp_ctx *h, const double *AB, const double *CD, const int *a,
                      const int *b, const int *c, const int *d) {
  for (int k = 0; k < 3; ++k)
    if (b[k] > 0) {
      int a1[3] = {a[0], a[1], a[2]}, b1[3] = {b[0], b[1], b[2]};
      a1[k] += 1;
      b1[k] -= 1;
      return hgp_hrr(h, AB, CD, a1, b1, c, d) + AB[k] * hgp_hrr(h, AB, CD, a, b1, c, d);
    }
  for (int k = 0; k < 3; ++k)
    if (d[k] > 0) {
      int c1[3] = {c[0], c[1], c[2]}, d1[3] = {d[0], d[1], d[2]};
      c1[k] += 1;
      d1[k] -= 1;
      return hgp_hrr(h, AB, CD, a, b, c1, d1) + CD[k] * hgp_hrr(h, AB, CD, a, b, c, d1);
    }
  return hgp_vrr(h, a, c, 0);
}

/* Primitive ERI with norms, by HGP; same arguments as orc_coulomb_repulsion. */
double orc_hgp_primitive(const double *r, const int *pw, const double *al, const double *nm,
                         int boys_mode) {
  hgp_ctx h;
  const double *A = r, *B = r + 3, *C = r + 6, *D = r + 9;
  h.zeta = al[0] + al[1];
  h.eta = al[2] + al[3];
  for (int k = 0; k < 3; ++k) {
    h.A[k] = A[k];
    h.C[k] = C[k];
    h.P[k] = product_center(al[0], A[k], al[1], B[k]);
    h.Q[k] = product_center(al[2], C[k], al[3], D[k]);
  }
  for (int k = 0; k < 3; ++k) h.W[k] = product_center(h.zeta, h.P[k], h.eta, h.Q[k]);
  double kab = sqrt(2.0) * pow(M_PI, 1.25) / h.zeta * exp(-al[0] * al[1] / h.zeta * dist2(A, B));
  double kcd = sqrt(2.0) * pow(M_PI, 1.25) / h.eta * exp(-al[2] * al[3] / h.eta * dist2(C, D));
  h.pref = nm[0] * nm[1] * nm[2] * nm[3] * kab * kcd / sqrt(h.zeta + h.eta);
  double T = h.zeta * h.eta / (h.zeta + h.eta) * dist2(h.P, h.Q);
  int L = 0;
  for (int k = 0; k < 12; ++k) L += pw[k];
  for (int m = 0; m <= L; ++m) h.F[m] = boys(boys_mode, m, T);
  double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
  double CD[3] = {C[0] - D[0], C[1] - D[1], C[2] - D[2]};
  return hgp_hrr(&h, AB, CD, pw, pw + 3, pw + 6, pw + 9);
}

/* ------------------------------------------------------------------------ */
/* One-electron primitives (THO; the same quantities as os86.rs:629-773).     */
/* ------------------------------------------------------------------------ */

/* 1-D overlap factor, THO eq. 2.12 (cints.c:199-209). */
static double overlap_1d(int l1, int l2, double pax, double pbx, double g) {
  double sum = 0.0;
  for (int i = 0; i < 1 + (l1 + l2) / 2; ++i)
    sum += binomial_prefactor(2 * i, l1, l2, pax, pbx) * ifact2(2 * i - 1) / ipow(2.0 * g, i);
  return sum;
}

/* Unnormalised primitive overlap <a|b> (cints.c:179-197; os86.rs:629-670). */
double orc_overlap(double a1, const int *p1, const double *A, double a2, const int *p2,
                   const double *B) {
  double g = a1 + a2, P[3];
  for (int k = 0; k < 3; ++k) P[k] = product_center(a1, A[k], a2, B[k]);
  double pre = pow(M_PI / g, 1.5) * exp(-a1 * a2 * dist2(A, B) / g);
  return pre * overlap_1d(p1[0], p2[0], P[0] - A[0], P[0] - B[0], g) *
         overlap_1d(p1[1], p2[1], P[1] - A[1], P[1] - B[1], g) *
         overlap_1d(p1[2], p2[2], P[2] - A[2], P[2] - B[2], g);
}

/* Unnormalised primitive kinetic-energy integral <a|-lap/2|b> (cints.c:152-177;
 * os86.rs:672-720). */
double orc_kinetic(double a1, const int *p1, const double *A, double a2, const int *p2,
                   const double *B) {
  int l2 = p2[0], m2 = p2[1], n2 = p2[2];
  int q[3];
  double t0 = a2 * (2 * (l2 + m2 + n2) + 3) * orc_overlap(a1, p1, A, a2, p2, B);
  double t1 = 0.0, t2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    memcpy(q, p2, sizeof q);
    q[k] += 2;
    t1 += orc_overlap(a1, p1, A, a2, q, B);
    if (p2[k] >= 2) {
      memcpy(q, p2, sizeof q);
      q[k] -= 2;
      t2 += p2[k] * (p2[k] - 1) * orc_overlap(a1, p1, A, a2, q, B);
    }
  }
  return t0 - 2.0 * a2 * a2 * t1 - 0.5 * t2;
}

/* THO eq. 2.18 A coefficients for one axis (cints.c:232-269). */
static void tho_A_axis(double *Ax, int l1, int l2, double pa, double pb, double cp, double g) {
  int imax = l1 + l2 + 1;
  for (int i = 0; i < imax; ++i) Ax[i] = 0.0;
  for (int i = 0; i < imax; ++i)
    for (int r = 0; r < i / 2 + 1; ++r)
      for (int u = 0; u < (i - 2 * r) / 2 + 1; ++u) {
        int I = i - 2 * r - u;
        Ax[I] += ((i & 1) ? -1.0 : 1.0) * binomial_prefactor(i, l1, l2, pa, pb) *
                 ((u & 1) ? -1.0 : 1.0) * ifact(i) * ipow(cp, i - 2 * r - 2 * u) *
                 ipow(0.25 / g, r + u) / ifact(r) / ifact(u) / ifact(i - 2 * r - 2 * u);
      }
}

/* Unnormalised primitive <a| -1/|r - C| |b> (cints.c:211-230 without the norms;
 * the sign convention of os86.rs:734).  The charge Z is applied by the caller
 * (basis.rs:332). */
double orc_nuclear_attraction(double a1, const int *p1, const double *A, double a2, const int *p2,
                              const double *B, const double *C, int boys_mode) {
  double g = a1 + a2, P[3];
  for (int k = 0; k < 3; ++k) P[k] = product_center(a1, A[k], a2, B[k]);
  double Ax[16], Ay[16], Az[16];
  tho_A_axis(Ax, p1[0], p2[0], P[0] - A[0], P[0] - B[0], P[0] - C[0], g);
  tho_A_axis(Ay, p1[1], p2[1], P[1] - A[1], P[1] - B[1], P[1] - C[1], g);
  tho_A_axis(Az, p1[2], p2[2], P[2] - A[2], P[2] - B[2], P[2] - C[2], g);
  double T = dist2(C, P) * g;
  int nx = p1[0] + p2[0] + 1, ny = p1[1] + p2[1] + 1, nz = p1[2] + p2[2] + 1;
  double F[64];
  for (int m = 0; m < nx + ny + nz - 2; ++m) F[m] = boys(boys_mode, m, T);
  double sum = 0.0;
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < nz; ++K) sum += Ax[I] * Ay[J] * Az[K] * F[I + J + K];
  return -2.0 * M_PI / g * exp(-a1 * a2 * dist2(A, B) / g) * sum;
}

/* ------------------------------------------------------------------------ */
/* Contracted-function layer with rchem's conventions.                        */
/*                                                                          */
/* A basis is n Cartesian contracted functions (rchem's CGTO, basis.rs:152-  */
/* 173): centre[3n], powers[3n], prim_off[n+1] into exps[] and coefs[].       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n;
  const double *centre;
  const int *powers;
  const int *prim_off;
  const double *exps;
  const double *coefs;
  double *norms; /* per-primitive norm, fixed at construction as PGTO::new does */
} orc_basis;

/* Primitive normalisation of rchem's PGTO (basis.rs:140-149):
 * N = [2^(2L+1.5) a^(L+1.5) / ((2l-1)!!(2m-1)!!(2n-1)!! pi^1.5)]^(1/2). */
double orc_pgto_norm(double alpha, int l, int m, int n) {
  int L = l + m + n;
  double numer = pow(2.0, 2.0 * L + 1.5) * pow(alpha, L + 1.5);
  double denom = ifact2(2 * l - 1) * ifact2(2 * m - 1) * ifact2(2 * n - 1) * pow(M_PI, 1.5);
  return pow(numer / denom, 0.5);
}

/* PGTO::new computes each primitive's norm once (basis.rs:125-134). */
static void make_basis(orc_basis *b, int n, const double *centre, const int *powers,
                       const int *prim_off, const double *exps, const double *coefs) {
  b->n = n;
  b->centre = centre;
  b->powers = powers;
  b->prim_off = prim_off;
  b->exps = exps;
  b->coefs = coefs;
  b->norms = (double *)malloc(sizeof(double) * (size_t)(prim_off[n] > 0 ? prim_off[n] : 1));
  for (int f = 0; f < n; ++f)
    for (int p = prim_off[f]; p < prim_off[f + 1]; ++p)
      b->norms[p] = orc_pgto_norm(exps[p], powers[3 * f], powers[3 * f + 1], powers[3 * f + 2]);
}

static void free_basis(orc_basis *b) { free(b->norms); }

typedef double (*orc_eri_fn)(const double *r, const int *pw, const double *al, const double *nm,
                             int boys_mode);

/* Contracted (mu nu|la si): sum of c c c c * primitive ERI with per-primitive
 * norms, as rchem's coulomb_pgto loop (basis.rs:340-381, 442-450) and cints'
 * contr_coulomb (cints.c:42-70). */
static double contracted_eri(const orc_basis *b, int i, int j, int k, int l, int boys_mode,
                             orc_eri_fn prim) {
  int f[4] = {i, j, k, l};
  double r[12];
  int pw[12];
  for (int s = 0; s < 4; ++s)
    for (int x = 0; x < 3; ++x) {
      r[3 * s + x] = b->centre[3 * f[s] + x];
      pw[3 * s + x] = b->powers[3 * f[s] + x];
    }
  double sum = 0.0;
  for (int pa = b->prim_off[i]; pa < b->prim_off[i + 1]; ++pa)
    for (int pb = b->prim_off[j]; pb < b->prim_off[j + 1]; ++pb)
      for (int pc = b->prim_off[k]; pc < b->prim_off[k + 1]; ++pc)
        for (int pd = b->prim_off[l]; pd < b->prim_off[l + 1]; ++pd) {
          double al[4] = {b->exps[pa], b->exps[pb], b->exps[pc], b->exps[pd]};
          double nm[4] = {b->norms[pa], b->norms[pb], b->norms[pc], b->norms[pd]};
          sum += b->coefs[pa] * b->coefs[pb] * b->coefs[pc] * b->coefs[pd] *
                 prim(r, pw, al, nm, boys_mode);
        }
  return sum;
}

double orc_contracted_eri(int n, const double *centre, const int *powers, const int *prim_off,
                          const double *exps, const double *coefs, int i, int j, int k, int l,
                          int boys_mode, int use_hgp) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  double v = contracted_eri(&b, i, j, k, l, boys_mode,
                            use_hgp ? orc_hgp_primitive : orc_coulomb_repulsion);
  free_basis(&b);
  return v;
}

/* Full N^4 ERI tensor [mu,nu,la,si] row-major, the layout of build_I
 * (basis.rs:430-460).  Each value is computed once for its 8-fold class and
 * stored in all eight places. */
void orc_eri_tensor(int n, const double *centre, const int *powers, const int *prim_off,
                    const double *exps, const double *coefs, int boys_mode, double *I) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  size_t N = (size_t)n;
#define IDX(a, b_, c, d) ((((size_t)(a) * N + (b_)) * N + (c)) * N + (d))
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k < n; ++k)
        for (int l = 0; l <= k; ++l) {
          if (i * (i + 1) / 2 + j < k * (k + 1) / 2 + l) continue;
          double v = contracted_eri(&b, i, j, k, l, boys_mode, orc_coulomb_repulsion);
          I[IDX(i, j, k, l)] = v;
          I[IDX(j, i, k, l)] = v;
          I[IDX(i, j, l, k)] = v;
          I[IDX(j, i, l, k)] = v;
          I[IDX(k, l, i, j)] = v;
          I[IDX(l, k, i, j)] = v;
          I[IDX(k, l, j, i)] = v;
          I[IDX(l, k, j, i)] = v;
        }
#undef IDX
  free_basis(&b);
}

/* rchem's JK_direct, literally (basis.rs:383-428): no symmetry, no screening,
 * K's integral recomputed separately, row mu of J and K written only by
 * iteration mu.  `prim` is the primitive-ERI routine (this file's THO
 * restatement, or the reference's own coulomb_repulsion compiled under
 * oracle/_ref, passed in by address).  Rows [mu_begin, mu_end) only, so a
 * bounded sample of rows can be timed. */
typedef double (*orc_cints_eri_fn)(double, double, double, double, int, int, int, double, double,
                                   double, double, double, int, int, int, double, double, double,
                                   double, double, int, int, int, double, double, double, double,
                                   double, int, int, int, double);

static orc_cints_eri_fn g_cints_eri = 0;

static double prim_via_cints(const double *r, const int *pw, const double *al, const double *nm,
                             int boys_mode) {
  (void)boys_mode;
  return g_cints_eri(r[0], r[1], r[2], nm[0], pw[0], pw[1], pw[2], al[0], r[3], r[4], r[5], nm[1],
                     pw[3], pw[4], pw[5], al[1], r[6], r[7], r[8], nm[2], pw[6], pw[7], pw[8],
                     al[2], r[9], r[10], r[11], nm[3], pw[9], pw[10], pw[11], al[3]);
}

void orc_jk_direct_rows(int n, const double *centre, const int *powers, const int *prim_off,
                        const double *exps, const double *coefs, const double *D, int boys_mode,
                        void *cints_coulomb_repulsion, int mu_begin, int mu_end, double *J,
                        double *K) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  orc_eri_fn prim = orc_coulomb_repulsion;
  if (cints_coulomb_repulsion) {
    g_cints_eri = (orc_cints_eri_fn)cints_coulomb_repulsion;
    prim = prim_via_cints;
  }
  for (int mu = mu_begin; mu < mu_end; ++mu)
    for (int nu = 0; nu < n; ++nu) {
      J[(size_t)mu * n + nu] = 0.0;
      K[(size_t)mu * n + nu] = 0.0;
    }
  for (int mu = mu_begin; mu < mu_end; ++mu)
    for (int nu = 0; nu < n; ++nu)
      for (int la = 0; la < n; ++la)
        for (int si = 0; si < n; ++si) {
          double d = D[(size_t)la * n + si];
          double j = 0.0, k = 0.0;
          /* The reference folds the coefficients and D into each primitive term
           * (basis.rs:406-417); contracted_eri sums c c c c * prim, which is
           * the same sum. */
          j = contracted_eri(&b, mu, nu, la, si, boys_mode, prim) * d;
          k = contracted_eri(&b, mu, la, nu, si, boys_mode, prim) * d;
          J[(size_t)mu * n + nu] += j;
          K[(size_t)mu * n + nu] += k;
        }
  free_basis(&b);
}

/* orc_jk_direct_rows on `nthreads` threads (OpenMP over mu): row mu of J and K
 * is written by one thread only, as each iteration of JK_direct's outer loop
 * writes only row mu (basis.rs:392-426), so the threads never share an
 * output element.  The CPU baseline's all-core leg. */
void orc_jk_direct_rows_mt(int n, const double *centre, const int *powers, const int *prim_off,
                           const double *exps, const double *coefs, const double *D,
                           int boys_mode, void *cints_coulomb_repulsion, int mu_begin, int mu_end,
                           int nthreads, double *J, double *K) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  orc_eri_fn prim = orc_coulomb_repulsion;
  if (cints_coulomb_repulsion) {
    g_cints_eri = (orc_cints_eri_fn)cints_coulomb_repulsion;
    prim = prim_via_cints;
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int mu = mu_begin; mu < mu_end; ++mu) {
    for (int nu = 0; nu < n; ++nu) {
      double j = 0.0, k = 0.0;
      for (int la = 0; la < n; ++la)
        for (int si = 0; si < n; ++si) {
          const double d = D[(size_t)la * n + si];
          j += contracted_eri(&b, mu, nu, la, si, boys_mode, prim) * d;
          k += contracted_eri(&b, mu, la, nu, si, boys_mode, prim) * d;
        }
      J[(size_t)mu * n + nu] = j;
      K[(size_t)mu * n + nu] = k;
    }
  }
  free_basis(&b);
}

/* orc_jk_direct_sample on `nthreads` threads: each thread accumulates its
 * share of the quartets into private J, K, summed at the end.  Used to time a
 * bounded sample of the reference loop on every core. */
void orc_jk_direct_sample_mt(int n, const double *centre, const int *powers, const int *prim_off,
                             const double *exps, const double *coefs, const double *D,
                             int boys_mode, void *cints_coulomb_repulsion, int nq, const int *q,
                             int nthreads, double *J, double *K) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  orc_eri_fn prim = orc_coulomb_repulsion;
  if (cints_coulomb_repulsion) {
    g_cints_eri = (orc_cints_eri_fn)cints_coulomb_repulsion;
    prim = prim_via_cints;
  }
  const size_t n2 = (size_t)n * n;
#pragma omp parallel num_threads(nthreads)
  {
    double *Jt = calloc(n2, sizeof(double)), *Kt = calloc(n2, sizeof(double));
#pragma omp for schedule(dynamic, 64)
    for (int t = 0; t < nq; ++t) {
      int mu = q[4 * t], nu = q[4 * t + 1], la = q[4 * t + 2], si = q[4 * t + 3];
      double d = D[(size_t)la * n + si];
      Jt[(size_t)mu * n + nu] += contracted_eri(&b, mu, nu, la, si, boys_mode, prim) * d;
      Kt[(size_t)mu * n + nu] += contracted_eri(&b, mu, la, nu, si, boys_mode, prim) * d;
    }
#pragma omp critical
    for (size_t i = 0; i < n2; ++i) {
      J[i] += Jt[i];
      K[i] += Kt[i];
    }
    free(Jt);
    free(Kt);
  }
  free_basis(&b);
}

/* One-electron matrices S, T, V (basis.rs:214-338): lower triangle computed and
 * mirrored; V = sum_C Z_C <mu| -1/r_C |nu>.  V uses the exact Boys function,
 * as rchem's os86 path does (os86.rs:770). */
void orc_one_electron(int n, const double *centre, const int *powers, const int *prim_off,
                      const double *exps, const double *coefs, int natom, const double *atom_xyz,
                      const double *Z, double *S, double *T, double *V) {
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu <= mu; ++nu) {
      const int *p1 = powers + 3 * mu, *p2 = powers + 3 * nu;
      const double *A = centre + 3 * mu, *B = centre + 3 * nu;
      double s = 0.0, t = 0.0, v = 0.0;
      for (int pb = prim_off[nu]; pb < prim_off[nu + 1]; ++pb)
        for (int pa = prim_off[mu]; pa < prim_off[mu + 1]; ++pa) {
          double na = orc_pgto_norm(exps[pa], p1[0], p1[1], p1[2]);
          double nb = orc_pgto_norm(exps[pb], p2[0], p2[1], p2[2]);
          double cc = coefs[pa] * coefs[pb] * na * nb;
          s += cc * orc_overlap(exps[pa], p1, A, exps[pb], p2, B);
          t += cc * orc_kinetic(exps[pa], p1, A, exps[pb], p2, B);
          for (int c = 0; c < natom; ++c)
            v += Z[c] * cc *
                 orc_nuclear_attraction(exps[pa], p1, A, exps[pb], p2, B, atom_xyz + 3 * c,
                                        ORC_BOYS_EXACT);
        }
      S[(size_t)mu * n + nu] = S[(size_t)nu * n + mu] = s;
      T[(size_t)mu * n + nu] = T[(size_t)nu * n + mu] = t;
      V[(size_t)mu * n + nu] = V[(size_t)nu * n + mu] = v;
    }
}

/* Nuclear repulsion sum_{A<B} Z_A Z_B / R_AB (absent from rchem, which prints
 * E_elec only, rchem.rs:81; closed form). */
double orc_nuclear_repulsion(int natom, const double *xyz, const double *Z) {
  double e = 0.0;
  for (int a = 0; a < natom; ++a)
    for (int b = 0; b < a; ++b) e += Z[a] * Z[b] / sqrt(dist2(xyz + 3 * a, xyz + 3 * b));
  return e;
}

/* JK_direct's inner body (basis.rs:398-423) for an explicit list of function
 * quartets q[4k..4k+3] = (mu, nu, la, si): J[mu,nu] += (mu nu|la si) D[la,si]
 * and K[mu,nu] += (mu la|nu si) D[la,si], with the primitive-ERI routine of
 * orc_jk_direct_rows.  Used to time a bounded sample of the reference loop. */
void orc_jk_direct_sample(int n, const double *centre, const int *powers, const int *prim_off,
                          const double *exps, const double *coefs, const double *D, int boys_mode,
                          void *cints_coulomb_repulsion, int nq, const int *q, double *J,
                          double *K) {
  orc_basis b;
  make_basis(&b, n, centre, powers, prim_off, exps, coefs);
  orc_eri_fn prim = orc_coulomb_repulsion;
  if (cints_coulomb_repulsion) {
    g_cints_eri = (orc_cints_eri_fn)cints_coulomb_repulsion;
    prim = prim_via_cints;
  }
  for (int t = 0; t < nq; ++t) {
    int mu = q[4 * t], nu = q[4 * t + 1], la = q[4 * t + 2], si = q[4 * t + 3];
    double d = D[(size_t)la * n + si];
    J[(size_t)mu * n + nu] += contracted_eri(&b, mu, nu, la, si, boys_mode, prim) * d;
    K[(size_t)mu * n + nu] += contracted_eri(&b, mu, la, nu, si, boys_mode, prim) * d;
  }
  free_basis(&b);
}
