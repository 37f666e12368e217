/* Drop-in replacements, with identical C signatures, for the libpyquante2
 * functions rchem binds through bindgen (reference libpyquante2/cints.h),
 * served by librchem_gpu.so's GPU kernels so that rchem's src/integrals/tho66.rs
 * wrappers and their tests link unchanged (link this library in place of
 * libpyquante2.a, build.rs:9-19).
 *
 * Semantics are the reference's: primitive Cartesian Gaussians
 * x^l y^m z^n exp(-alpha r^2) about the given centres; coulomb_repulsion and
 * nuclear_attraction multiply by the given norms; overlap and kinetic take none;
 * the Boys function is the reference's incomplete-gamma Fgamma (series /
 * continued fraction, EPS 3e-7).  ERIs serve l + m + n <= 2 per function,
 * overlap and kinetic powers up to 6 per axis, nuclear attraction a total
 * l1 + m1 + n1 + l2 + m2 + n2 <= 8.  The functions share one
 * device context (device 0) and are not thread-safe; on failure they return
 * NaN and rchem_last_error(NULL) gives the reason.
 */
#ifndef RCHEM_CINTS_COMPAT_H
#define RCHEM_CINTS_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

/* cints.h:14-21: contracted ERI; arrays of length ia, ib, ic, id */
double contr_coulomb(int ia, double *aexps, double *acoefs, double *anorms,
                     double xa, double ya, double za, int la, int ma, int na,
                     int ib, double *bexps, double *bcoefs, double *bnorms,
                     double xb, double yb, double zb, int lb, int mb, int nb,
                     int ic, double *cexps, double *ccoefs, double *cnorms,
                     double xc, double yc, double zc, int lc, int mc, int nc,
                     int id, double *dexps, double *dcoefs, double *dnorms,
                     double xd, double yd, double zd, int ld, int md, int nd);

/* cints.h:23-30: primitive ERI times norma normb normc normd */
double coulomb_repulsion(double xa, double ya, double za, double norma,
                         int la, int ma, int na, double alphaa,
                         double xb, double yb, double zb, double normb,
                         int lb, int mb, int nb, double alphab,
                         double xc, double yc, double zc, double normc,
                         int lc, int mc, int nc, double alphac,
                         double xd, double yd, double zd, double normd,
                         int ld, int md, int nd, double alphad);

/* cints.h:40-43: primitive kinetic-energy integral */
double kinetic(double alpha1, int l1, int m1, int n1,
               double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2,
               double xb, double yb, double zb);

/* cints.h:44-47: primitive overlap */
double overlap(double alpha1, int l1, int m1, int n1,
               double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2,
               double xb, double yb, double zb);

/* cints.h:50-54: <a| -1/|r - R3| |b> times norm1 norm2 (unit charge) */
double nuclear_attraction(double x1, double y1, double z1, double norm1,
                          int l1, int m1, int n1, double alpha1,
                          double x2, double y2, double z2, double norm2,
                          int l2, int m2, int n2, double alpha2,
                          double x3, double y3, double z3);

/* cints.h:69: F_m(x) = 1/2 x^(-m-1/2) gamma(m + 1/2, x) */
double Fgamma(double m, double x);

#ifdef __cplusplus
}
#endif

#endif /* RCHEM_CINTS_COMPAT_H */
