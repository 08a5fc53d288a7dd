/* CPU restatement of the reference's extended-source light-curve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product never does.  Built with -O2 -ffp-contract=off (no FMA
 * contraction, no reassociation) so that every operation below is one IEEE
 * double operation in the order written.
 *
 * What each part follows (reference files, relative to the reference root):
 *   complex arithmetic   libstdc++ std::complex<double> as compiled by g++ 11
 *                        at -O2: products (ac-bd, ad+bc) inline, quotients by
 *                        libgcc __divdc3 (Smith's method), abs by glibc cabs
 *                        (= hypot, glibc 2.35 algorithm), sqrt by glibc csqrt.
 *                        Restated below from their published algorithms;
 *                        tests compare them with the system libm on random
 *                        inputs.
 *   Laguerre             src/Laguerre/laguerre.cc:18-198
 *   lens frame           src/LensCore/lens.cc:4-40
 *   critical curves      src/CCC/ccc.cc:42-98 (coefficients), 160-203
 *   image polynomial     src/Images/imgpoint.cc:160-176 with
 *                        src/Images/imgpointcoeff.cc:223-447, evaluated in the
 *                        reference's operation order (tl_imgcoeff.c), so the
 *                        coefficients are the reference's to the bit; an
 *                        independent derivation rounded once (lens_poly_exact)
 *                        serves as a mathematical cross-check.
 *   images               src/Images/imgpoint.cc:72-90,160-274
 *   grid, boxes          src/LightCurves/lcirs.cc:39-92, src/Utils/utils.h:16-28
 *   light-curve loop     src/LightCurves/lcirs.cc:103-191
 *   caustic seeding      src/LightCurves/lcirs.cc:131-165,195-255
 *   ray map              src/LightCurves/lcirs.cc:290-341 (irsSIMD lane order)
 *   pixel mapping        src/LightCurves/lcirs.cc:355-378
 *   flood fill           src/LightCurves/lcirs.cc:381-466 and
 *                        src/Utils/amoeba.cc:15-82, with an explicit stack that
 *                        visits pixels in the same order as the recursion
 *   limb darkening       src/LimbDarkening/limbdarkeningmodel.cc:8-49
 *   point magnification  src/LightCurves/lcbase.cc:17-38
 */
#define _GNU_SOURCE
#include "tl_oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* complex arithmetic                                                         */
/* ------------------------------------------------------------------------ */

typedef struct { double re, im; } cx;

static inline cx mk(double re, double im) { cx z = {re, im}; return z; }
static inline cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static inline cx cneg(cx a) { return mk(-a.re, -a.im); }
static inline cx cconj(cx a) { return mk(a.re, -a.im); }
/* complex * complex: real = ar*br - ai*bi, imag = ar*bi + ai*br */
static inline cx cmul(cx a, cx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* complex * double and double * complex scale both parts */
static inline cx cscale(cx a, double t) { return mk(a.re * t, a.im * t); }
/* std::norm: x*x + y*y */
static inline double cnorm(cx a) { return a.re * a.re + a.im * a.im; }

/* (a + ib) / (c + id) by Smith's method, as libgcc's __divdc3 computes it for
 * operands in the normal range. */
static cx cdiv(cx n, cx d) {
  double a = n.re, b = n.im, c = d.re, e = d.im;
  double ratio, denom, x, y;
  if (fabs(c) < fabs(e)) {
    ratio = c / e;
    denom = (c * ratio) + e;
    x = ((a * ratio) + b) / denom;
    y = ((b * ratio) - a) / denom;
  } else {
    ratio = e / c;
    denom = (e * ratio) + c;
    x = ((b * ratio) + a) / denom;
    y = (b - (a * ratio)) / denom;
  }
  return mk(x, y);
}
/* double / complex: the scalar is promoted to (t, 0) and divided */
static inline cx cdiv_rc(double t, cx d) { return cdiv(mk(t, 0.0), d); }

/* hypot as glibc 2.35 computes it (Borges' corrected algorithm without FMA,
 * with the same scaling thresholds). */
static double hyp_kernel(double ax, double ay) {
  double t1, t2;
  double h = sqrt(ax * ax + ay * ay);
  if (h <= 2.0 * ay) {
    double delta = h - ay;
    t1 = ax * (2.0 * delta - ax);
    t2 = (delta - 2.0 * (ax - ay)) * delta;
  } else {
    double delta = h - ax;
    t1 = 2.0 * delta * (ax - 2.0 * ay);
    t2 = (4.0 * delta - ay) * ay + delta * delta;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

static double hyp(double x, double y) {
  const double SCALE = 0x1p-600, LARGE = 0x1p+511, TINY = 0x1p-511, EPSH = 0x1p-54;
  if (!isfinite(x) || !isfinite(y)) {
    if (isinf(x) || isinf(y)) return INFINITY;
    return x + y;
  }
  x = fabs(x);
  y = fabs(y);
  double ax = x < y ? y : x;
  double ay = x < y ? x : y;
  if (ax > LARGE) {
    if (ay <= ax * EPSH) return ax + ay;
    return hyp_kernel(ax * SCALE, ay * SCALE) / SCALE;
  }
  if (ay < TINY) {
    if (ax >= ay / EPSH) return ax + ay;
    return hyp_kernel(ax / SCALE, ay / SCALE) * SCALE;
  }
  if (ay <= ax * EPSH) return ax + ay;
  return hyp_kernel(ax, ay);
}

static inline double cabs_(cx a) { return hyp(a.re, a.im); }

/* complex square root as glibc csqrt computes it for finite operands. */
static cx csqrt_(cx z) {
  double re = z.re, im = z.im;
  if (im == 0.0) {
    if (re < 0.0) return mk(0.0, copysign(sqrt(-re), im));
    return mk(fabs(sqrt(re)), copysign(0.0, im));
  }
  if (re == 0.0) {
    double r;
    if (fabs(im) >= 2.0 * DBL_MIN) r = sqrt(0.5 * fabs(im));
    else r = 0.5 * sqrt(2.0 * fabs(im));
    return mk(r, copysign(r, im));
  }
  int scale = 0;
  if (fabs(re) > DBL_MAX / 4.0) {
    scale = 1;
    re = scalbn(re, -2);
    im = scalbn(im, -2);
  }
  if (fabs(im) > DBL_MAX / 4.0) {
    scale = 1;
    if (fabs(re) >= 4.0 * DBL_MIN) re = scalbn(re, -2);
    else re = 0.0;
    im = scalbn(im, -2);
  } else if (fabs(re) < 2.0 * DBL_MIN && fabs(im) < 2.0 * DBL_MIN) {
    scale = -((DBL_MANT_DIG + 1) / 2);
    re = scalbn(re, -2 * scale);
    im = scalbn(im, -2 * scale);
  }
  double d = hyp(re, im), r, s;
  if (re > 0.0) {
    r = sqrt(0.5 * (d + re));
    if (scale == 1 && fabs(im) < 1.0) {
      s = im / r;
      r = scalbn(r, scale);
      scale = 0;
    } else {
      s = 0.5 * (im / r);
    }
  } else {
    s = sqrt(0.5 * (d - re));
    if (scale == 1 && fabs(im) < 1.0) {
      r = fabs(im / s);
      s = scalbn(s, scale);
      scale = 0;
    } else {
      r = fabs(0.5 * (im / s));
    }
  }
  if (scale) {
    r = scalbn(r, scale);
    s = scalbn(s, scale);
  }
  return mk(r, copysign(s, im));
}

double tlo_hypot(double x, double y) { return hyp(x, y); }

/* The system library's own routines, for tests of the restatements above:
 * libm hypot and csqrt, and C complex division (libgcc __divdc3). */
double tlo_libm_hypot(double x, double y) { return hypot(x, y); }
void tlo_libm_csqrt(double re, double im, double* out) {
  double complex r = csqrt(CMPLX(re, im));
  out[0] = creal(r); out[1] = cimag(r);
}
void tlo_libm_cdiv(double a, double b, double c, double d, double* out) {
  volatile double complex n = CMPLX(a, b), q = CMPLX(c, d);
  double complex r = n / q;
  out[0] = creal(r); out[1] = cimag(r);
}
void tlo_csqrt(double re, double im, double* out) {
  cx r = csqrt_(mk(re, im));
  out[0] = r.re; out[1] = r.im;
}
void tlo_cdiv(double a, double b, double c, double d, double* out) {
  cx r = cdiv(mk(a, b), mk(c, d));
  out[0] = r.re; out[1] = r.im;
}

/* ------------------------------------------------------------------------ */
/* Laguerre (laguerre.cc)                                                     */
/* ------------------------------------------------------------------------ */

#define LAG_MAX_IT 300
#define LAG_MT 8
#define LAG_EPS DBL_EPSILON
#define MAXDEG 10

/* Fractional kicks applied every MT-th step (laguerre.cc:22,93).  The
 * reference indexes frac[i/MT] for i up to MAX_IT, reading past the 8-element
 * table once i >= 64; here the index wraps instead. */
static const double lag_frac[8] = {0.5, 0.23, 0.73, 0.13, 0.37, 0.61, 0.89, 1.0};

/* Work counters of the seed search (test instrumentation): [0] image solves
 * (getImages calls), [1] Laguerre calls, [2] Laguerre iterations. */
static __thread long g_work[3];

/* Kick steps whose table index i/MT reached 8 (test instrumentation): the
 * reference reads frac[i/MT] from an 8-entry vector (laguerre.cc:22,92), so
 * from iteration 64 on it reads past the vector, into heap memory.  The
 * oracle wraps the index modulo 8 there; a point whose solves took such a
 * step has a result the reference does not define. */
static __thread long g_frac_oob;

/* Laguerre::laguerre (laguerre.cc:18-99).  poly has m+1 coefficients. */
static int lag_iter(const cx* poly, int m, cx* xp) {
  cx x = *xp;
  const double md = (double)m;
  g_work[1]++;
  if (m == 1) {
    if (cabs_(poly[1]) != 0.0) {
      *xp = cdiv(cneg(poly[0]), poly[1]);
      return 1;
    }
    *xp = mk(999.0, 999.0);
    return 0;
  }
  for (int i = 1; i <= LAG_MAX_IT; i++) {
    g_work[2]++;
    cx b = poly[m];
    double err = cabs_(b);
    cx s = mk(0.0, 0.0), f = mk(0.0, 0.0);
    for (int j = m - 1; j >= 0; j--) {
      s = cadd(cmul(x, s), f);
      f = cadd(cmul(x, f), b);
      b = cadd(cmul(x, b), poly[j]);
      err = cabs_(b) + cabs_(x) * err;
    }
    err *= LAG_EPS;
    if (cabs_(b) <= err) { *xp = x; return 1; }
    cx g = cdiv(f, b);
    cx h = csub(cmul(g, g), cdiv(cscale(s, 2.0), b));
    cx rad = cscale(csub(cscale(h, md), cmul(g, g)), md - 1.0);
    cx gp = cadd(g, csqrt_(rad));
    cx gm = csub(g, csqrt_(rad));
    if (cabs_(gp) < cabs_(gm)) gp = gm;
    cx dx;
    if (cabs_(gp) > 0.0) dx = cdiv_rc(md, gp);
    else {
      double sn, cs;
      sincos(1.0 * i, &sn, &cs);
      dx = cscale(mk(cs, sn), 1.0 + cabs_(x));
    }
    cx x1 = csub(x, dx);
    if (x.re == x1.re && x.im == x1.im) { *xp = x; return 1; }
    if (i % LAG_MT) x = x1;
    else {
      if (i / LAG_MT >= 8) g_frac_oob++;
      x = csub(x, cmul(mk(lag_frac[(i / LAG_MT) & 7], 0.0), dx));
    }
    *xp = x;
  }
  *xp = x;
  return 0;
}

/* Laguerre::solveRoots (laguerre.cc:103-167).  coeffs: deg+1 values.
 * init: NULL, or deg+1 values where init[j] seeds the search for the root
 * found at step j = deg..1 (the reference reads initRoots[deg], one past the
 * end of its vector; callers supply that slot explicitly). */
static void lag_solve(const cx* coeffs, int deg, const cx* init, cx* roots) {
  cx tmp[MAXDEG + 1];
  for (int i = 0; i <= deg; i++) tmp[i] = coeffs[i];
  int nr = 0;
  for (int j = deg; j >= 1; j--) {
    cx x = init ? init[j] : mk(LAG_EPS, LAG_EPS);
    if (lag_iter(tmp, j, &x)) {
      roots[nr++] = x;
    } else {
      if (init) x = mk(LAG_EPS, LAG_EPS);
      roots[nr++] = x;
    }
    /* deflation by the root just found */
    cx b = tmp[j];
    for (int jj = j - 1; jj >= 0; jj--) {
      cx c = tmp[jj];
      tmp[jj] = b;
      b = cadd(cmul(x, b), c);
    }
  }
  /* polish on the full polynomial */
  for (int j = 0; j < deg; j++) lag_iter(coeffs, deg, &roots[j]);
}

/* Laguerre::polishRoots (laguerre.cc:170-183) */
static void lag_polish(const cx* coeffs, int deg, cx* roots) {
  for (int j = 0; j < deg; j++) lag_iter(coeffs, deg, &roots[j]);
}

/* Laguerre::checkRoots (laguerre.cc:188-198) */
static int lag_check(const cx* coeffs, int deg, const cx* roots) {
  cx c0 = coeffs[deg];
  for (int j = 0; j < deg; j++) c0 = cmul(c0, cneg(roots[j]));
  if (cabs_(coeffs[0]) > 1.0e-8)
    return (1.0 - cabs_(cdiv(c0, coeffs[0]))) < 1.0e-5;
  return cabs_(csub(c0, coeffs[0])) < 1.0e-10;
}

int tlo_laguerre_solve(int deg, const double* coeffs, const double* init, double* roots) {
  if (deg < 1 || deg > MAXDEG || !coeffs || !roots) return -1;
  cx c[MAXDEG + 1], in[MAXDEG + 1], r[MAXDEG];
  for (int i = 0; i <= deg; i++) c[i] = mk(coeffs[2 * i], coeffs[2 * i + 1]);
  if (init)
    for (int i = 0; i <= deg; i++) in[i] = mk(init[2 * i], init[2 * i + 1]);
  lag_solve(c, deg, init ? in : NULL, r);
  for (int i = 0; i < deg; i++) { roots[2 * i] = r[i].re; roots[2 * i + 1] = r[i].im; }
  return 0;
}

int tlo_laguerre_one(int deg, const double* coeffs, double* x) {
  if (deg < 1 || deg > MAXDEG) return -1;
  cx c[MAXDEG + 1];
  for (int i = 0; i <= deg; i++) c[i] = mk(coeffs[2 * i], coeffs[2 * i + 1]);
  cx z = mk(x[0], x[1]);
  int ok = lag_iter(c, deg, &z);
  x[0] = z.re; x[1] = z.im;
  return ok;
}

/* ------------------------------------------------------------------------ */
/* lens                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  double a, b, th, m1, m2, m3;
  cx z1, z2, z3;
} lens_t;

/* Lens::Lens and Lens::_setPos (lens.cc:4-18,35-40) */
static void lens_init(lens_t* L, double a, double b, double th, double m2, double m3) {
  L->a = a; L->b = b; L->th = th;
  L->m1 = 1.0 - m2 - m3;
  L->m2 = m2; L->m3 = m3;
  L->z1 = mk(0.0, 0.0);
  L->z2 = mk(a, 0.0);
  double sn, cs;
  sincos(th, &sn, &cs); /* g++ merges the reference's cos/sin pair into sincos */
  L->z3 = mk(b * cs, b * sn);
}

/* ---- degree-10 polynomial ---------------------------------------------- */
/* The coefficients are accumulated in double-double arithmetic (error-free
 * sums and fma-based products) and rounded once at the end, so that each is
 * the exact value of the polynomial defined by the double inputs (m1, m2,
 * m3, z2, z3, zeta), rounded to nearest: a result that does not depend on
 * the order of evaluation.  The reference's generated expressions land
 * within a few ulps of these values. */
typedef struct { double hi, lo; } dd;
typedef struct { dd re, im; } dcx;

static inline dd two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  dd r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static inline dd quick_two_sum(double a, double b) {
  double s = a + b;
  dd r = {s, b - (s - a)};
  return r;
}
static inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
static inline dd dd_neg(dd a) { dd r = {-a.hi, -a.lo}; return r; }
static inline dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p, e);
}
static inline dd dd_of(double x) { dd r = {x, 0.0}; return r; }
static inline dcx dc_of(cx z) { dcx r = {dd_of(z.re), dd_of(z.im)}; return r; }
static inline dcx dc_add(dcx a, dcx b) { dcx r = {dd_add(a.re, b.re), dd_add(a.im, b.im)}; return r; }
static inline dcx dc_sub(dcx a, dcx b) {
  dcx r = {dd_add(a.re, dd_neg(b.re)), dd_add(a.im, dd_neg(b.im))};
  return r;
}
static inline dcx dc_mul(dcx a, dcx b) {
  dcx r = {dd_add(dd_mul(a.re, b.re), dd_neg(dd_mul(a.im, b.im))),
           dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
  return r;
}
static inline dcx dc_scale(dcx a, dd t) { dcx r = {dd_mul(a.re, t), dd_mul(a.im, t)}; return r; }
static inline cx dc_round(dcx a) { return mk(a.re.hi + a.re.lo, a.im.hi + a.im.lo); }

static void dpmul(const dcx* p, int np, const dcx* q, int nq, dcx* out) {
  const dcx z = {{0.0, 0.0}, {0.0, 0.0}};
  for (int i = 0; i < np + nq - 1; i++) out[i] = z;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < nq; j++) out[i + j] = dc_add(out[i + j], dc_mul(p[i], q[j]));
}

/* Lens-equation polynomial.  Conjugating zeta = z - sum m_k/conj(z - z_k)
 * gives conj(z) = conj(zeta) + sum m_k/(z - z_k) = Nn(z)/D(z) with
 * D = prod (z - z_k).  Substituting back, with Q_k = Nn - conj(z_k) D =
 * (conj(zeta) - conj(z_k)) D + sum_j m_j D/(z - z_j):
 *   (z - zeta) Q1 Q2 Q3 - D (m1 Q2 Q3 + m2 Q1 Q3 + m3 Q1 Q2) = 0,
 * degree 10 with leading coefficient prod (conj(zeta) - conj(z_k)).  The
 * reference's polynomial (imgpointcoeff.cc:223-447) is this one times -1,
 * which is the sign kept here. */
static void lens_poly_exact(const lens_t* L, cx zeta, cx* c) {
  const dcx one = {{1.0, 0.0}, {0.0, 0.0}};
  const dcx zero = {{0.0, 0.0}, {0.0, 0.0}};
  const dd m1 = dd_of(L->m1), m2 = dd_of(L->m2), m3 = dd_of(L->m3);
  dcx f1[2] = {dc_of(cneg(L->z1)), one}, f2[2] = {dc_of(cneg(L->z2)), one},
      f3[2] = {dc_of(cneg(L->z3)), one};
  dcx d12[3], d13[3], d23[3], D[4];
  dpmul(f1, 2, f2, 2, d12);
  dpmul(f1, 2, f3, 2, d13);
  dpmul(f2, 2, f3, 2, d23);
  dpmul(d12, 3, f3, 2, D);
  dcx S[3];
  for (int i = 0; i < 3; i++)
    S[i] = dc_add(dc_add(dc_scale(d23[i], m1), dc_scale(d13[i], m2)), dc_scale(d12[i], m3));
  dcx zc = dc_of(cconj(zeta));
  dcx w1 = dc_sub(zc, dc_of(cconj(L->z1))), w2 = dc_sub(zc, dc_of(cconj(L->z2))),
      w3 = dc_sub(zc, dc_of(cconj(L->z3)));
  dcx Q1[4], Q2[4], Q3[4];
  for (int i = 0; i < 4; i++) {
    dcx s = i < 3 ? S[i] : zero;
    Q1[i] = dc_add(dc_mul(w1, D[i]), s);
    Q2[i] = dc_add(dc_mul(w2, D[i]), s);
    Q3[i] = dc_add(dc_mul(w3, D[i]), s);
  }
  dcx q12[7], q13[7], q23[7], q123[10];
  dpmul(Q1, 4, Q2, 4, q12);
  dpmul(Q1, 4, Q3, 4, q13);
  dpmul(Q2, 4, Q3, 4, q23);
  dpmul(q12, 7, Q3, 4, q123);
  dcx lin[2] = {dc_of(cneg(zeta)), one};
  dcx left[11];
  dpmul(lin, 2, q123, 10, left);
  dcx mix[7];
  for (int i = 0; i < 7; i++)
    mix[i] = dc_add(dc_add(dc_scale(q23[i], m1), dc_scale(q13[i], m2)), dc_scale(q12[i], m3));
  dcx right[10];
  dpmul(D, 4, mix, 7, right);
  for (int i = 0; i < 11; i++) {
    dcx r = i < 10 ? right[i] : zero;
    c[i] = dc_round(dc_sub(r, left[i]));
  }
}

/* The polynomial as ImgPoint::getRootsPrecalculated forms it
 * (imgpoint.cc:163-176): the position-free part (getCoeffsOptNoZ) plus the
 * position part (getCoeffsOptJustZ), added coefficient by coefficient.  The
 * reference caches the position-free part; recomputing it gives the same
 * bits. */
static void lens_poly(const lens_t* L, cx zeta, cx* c) {
  double complex fz[11], pz[11];
  const double complex z2 = CMPLX(L->z2.re, L->z2.im), z3 = CMPLX(L->z3.re, L->z3.im);
  tlo_poly_nozeta(z2, z3, L->m2, L->m3, fz);
  tlo_poly_zeta(z2, z3, L->m2, L->m3, CMPLX(zeta.re, zeta.im), pz);
  for (int i = 0; i < 11; i++) c[i] = mk(creal(fz[i]) + creal(pz[i]), cimag(fz[i]) + cimag(pz[i]));
}

int tlo_coeffs(double a, double b, double th, double m2, double m3,
               double zr, double zi, double* out) {
  lens_t L;
  lens_init(&L, a, b, th, m2, m3);
  cx c[11];
  lens_poly(&L, mk(zr, zi), c);
  for (int i = 0; i < 11; i++) { out[2 * i] = c[i].re; out[2 * i + 1] = c[i].im; }
  return 0;
}

int tlo_coeffs_exact(double a, double b, double th, double m2, double m3,
                     double zr, double zi, double* out) {
  lens_t L;
  lens_init(&L, a, b, th, m2, m3);
  cx c[11];
  lens_poly_exact(&L, mk(zr, zi), c);
  for (int i = 0; i < 11; i++) { out[2 * i] = c[i].re; out[2 * i + 1] = c[i].im; }
  return 0;
}

/* ---- point images (imgpoint.cc) ----------------------------------------- */

typedef struct {
  lens_t L;
  int binary;               /* the binary image-count rule (a, b or m3 zero) */
  cx tempRoots[MAXDEG + 1]; /* [10] is the slot read past the end */
  int haveRoots;
  cx roots[MAXDEG];
  int isImg[MAXDEG];
  cx imgs[MAXDEG];
  int nImg;
  /* Optional values of the slot read past the end (test instrumentation):
   * getImages call k of the chain starts its chained solve with
   * tempRoots[10] = slotTrace[k] instead of 0.  The reference reads
   * whatever its allocator left in that element of the roots buffer. */
  const cx* slotTrace;
  long slotCount;
  long calls;
} imgpoint_t;

/* ImgPoint::ImgPoint (imgpoint.cc:57-67): a binary lens when a, b or m3 is 0 */
static void imgpoint_init(imgpoint_t* ip, double a, double b, double th, double m2, double m3) {
  memset(ip, 0, sizeof(*ip));
  lens_init(&ip->L, a, b, th, m2, m3);
  ip->binary = a == 0.0 || b == 0.0 || m3 == 0.0;
}

/* A valid image count (imgpoint.cc:236-237): even and at least 4, or for
 * the binary lens odd and at least 3. */
static int img_count_ok(const imgpoint_t* ip) {
  return ip->binary ? (ip->nImg % 2 == 1 && ip->nImg >= 3)
                    : (ip->nImg % 2 == 0 && ip->nImg >= 4);
}

/* ImgPoint::getRootsPrecalculated (imgpoint.cc:160-205) */
static void img_roots(imgpoint_t* ip, cx zeta, int force) {
  cx c[11];
  lens_poly(&ip->L, zeta, c);
  if (!ip->haveRoots || force) {
    lag_solve(c, 10, NULL, ip->tempRoots);
    ip->haveRoots = 1;
  } else {
    cx in[MAXDEG + 1], out[MAXDEG];
    memcpy(in, ip->tempRoots, sizeof(in));
    if (ip->slotTrace && ip->calls - 1 < ip->slotCount) in[10] = ip->slotTrace[ip->calls - 1];
    lag_solve(c, 10, in, out);
    if (!lag_check(c, 10, out)) lag_solve(c, 10, NULL, out);
    memcpy(ip->tempRoots, out, sizeof(out));
  }
  /* The slot one past the end is never written by the reference's
   * push_back loop; it is kept at zero here. */
  ip->tempRoots[10] = mk(0.0, 0.0);
  memcpy(ip->roots, ip->tempRoots, sizeof(ip->roots));
}

/* ImgPoint::imgCheck (imgpoint.cc:72-90) */
static int img_check(const lens_t* L, cx img, cx zeta, double tol) {
  cx t = csub(img, cdiv_rc(L->m1, cconj(csub(img, L->z1))));
  t = csub(t, cdiv_rc(L->m2, cconj(csub(img, L->z2))));
  t = csub(t, cdiv_rc(L->m3, cconj(csub(img, L->z3))));
  return cabs_(csub(t, zeta)) < tol;
}

/* Smallest distance, in decades, between a root's lens-equation residual
 * and the 1e-6 acceptance threshold over the image tests since the last
 * reset (test instrumentation: a seed decision this close to the threshold
 * depends on rounding). */
static __thread double g_margin = 1e300;

static double img_residual(const lens_t* L, cx img, cx zeta) {
  cx t = csub(img, cdiv_rc(L->m1, cconj(csub(img, L->z1))));
  t = csub(t, cdiv_rc(L->m2, cconj(csub(img, L->z2))));
  t = csub(t, cdiv_rc(L->m3, cconj(csub(img, L->z3))));
  return cabs_(csub(t, zeta));
}

static void img_classify(imgpoint_t* ip, cx zeta) {
  ip->nImg = 0;
  for (int k = 0; k < 10; k++) {
    ip->isImg[k] = img_check(&ip->L, ip->roots[k], zeta, 1.0e-6);
    if (ip->isImg[k]) ip->imgs[ip->nImg++] = ip->roots[k];
    double r = img_residual(&ip->L, ip->roots[k], zeta);
    double m = r > 0.0 ? fabs(log10(r / 1.0e-6)) : 1e300;
    if (m < g_margin) g_margin = m;
  }
  /* Two roots converged onto one image: the solver lost a root, and which
   * one it lost follows the iteration path, i.e. the last bits of the
   * coefficients.  Reported as margin 0. */
  for (int i = 0; i < ip->nImg; i++)
    for (int j = i + 1; j < ip->nImg; j++)
      if (cabs_(csub(ip->imgs[i], ip->imgs[j])) < 1.0e-8) g_margin = 0.0;
}

/* ImgPoint::getImages(pos) for a triple lens (imgpoint.cc:208-274,281-286) */
static int img_get(imgpoint_t* ip, cx zeta) {
  g_work[0]++;
  ip->calls++;
  img_roots(ip, zeta, 0);
  img_classify(ip, zeta);
  if (!img_count_ok(ip)) {
    /* the chained solve gave an invalid image count: also a lost root */
    g_margin = 0.0;
    img_roots(ip, zeta, 1);
    img_classify(ip, zeta);
  }
  return ip->nImg;
}

int tlo_images(double a, double b, double th, double m2, double m3, int n,
               const double* pos, double* roots, int* isImg, int* nImg) {
  imgpoint_t ip;
  imgpoint_init(&ip, a, b, th, m2, m3);
  for (int i = 0; i < n; i++) {
    nImg[i] = img_get(&ip, mk(pos[2 * i], pos[2 * i + 1]));
    for (int k = 0; k < 10; k++) {
      roots[20 * i + 2 * k] = ip.roots[k].re;
      roots[20 * i + 2 * k + 1] = ip.roots[k].im;
      isImg[10 * i + k] = ip.isImg[k];
    }
  }
  return 0;
}

/* LightCurveBase::getPointAmp (lcbase.cc:17-38).  The reference squares with
 * std::pow(complex, 2.0) (log/exp form); the square here is a product, so
 * the two agree to rounding only. */
static double point_amp(imgpoint_t* ip, cx zeta) {
  img_get(ip, zeta);
  const lens_t* L = &ip->L;
  double amp = 0.0;
  for (int k = 0; k < ip->nImg; k++) {
    cx z = ip->imgs[k];
    cx d1 = csub(z, L->z1), d2 = csub(z, L->z2), d3 = csub(z, L->z3);
    cx s = cadd(cadd(cdiv_rc(L->m1, cmul(d1, d1)), cdiv_rc(L->m2, cmul(d2, d2))),
                cdiv_rc(L->m3, cmul(d3, d3)));
    double t = 1.0 / fabs(1.0 - cnorm(s));
    if (t != t) t = 0.0;
    amp += t;
  }
  return amp;
}

int tlo_point_amp(double a, double b, double th, double m2, double m3, int n,
                  const double* pos, double* out) {
  imgpoint_t ip;
  imgpoint_init(&ip, a, b, th, m2, m3);
  for (int i = 0; i < n; i++) out[i] = point_amp(&ip, mk(pos[2 * i], pos[2 * i + 1]));
  return 0;
}

/* Point magnification at the 12 disk samples around each position (the
 * samples of Gould's 2008 hexadecapole approximation: rho/2 on the axes,
 * rho on the axes, rho on the diagonals), each solved from the given roots
 * of the position's own solve: the root-reuse solve of getRootsPrecalculated
 * (imgpoint.cc:160-205) and the sum of getPointAmp (lcbase.cc:17-38);
 * cross[i] (may be NULL) counts the samples whose count of significant
 * images (each at least 1e-4 of the sum of the terms) differs from the
 * centre roots' count.  No counterpart in the reference; restates the
 * engine's ring_amp_kernel. */
/* Images carrying at least 1e-4 of the total of their terms. */
static int significant_images(const lens_t* L, const cx* imgs, int n) {
  double t[MAXDEG], amp = 0.0;
  for (int k = 0; k < n; k++) {
    cx z = imgs[k];
    cx d1 = csub(z, L->z1), d2 = csub(z, L->z2), d3 = csub(z, L->z3);
    cx s = cadd(cadd(cdiv_rc(L->m1, cmul(d1, d1)), cdiv_rc(L->m2, cmul(d2, d2))),
                cdiv_rc(L->m3, cmul(d3, d3)));
    t[k] = 1.0 / fabs(1.0 - cnorm(s));
    if (t[k] != t[k]) t[k] = 0.0;
    amp += t[k];
  }
  int c = 0;
  for (int k = 0; k < n; k++) c += t[k] >= 1.0e-4 * amp;
  return c;
}

int tlo_ring_amp(double a, double b, double th, double m2, double m3, double rho, int n,
                 const double* pos, const double* roots, double* out, int* cross) {
  static const double c = 0.70710678118654752;
  static const double off[12][2] = {{0.5, 0.0}, {0.0, 0.5}, {-0.5, 0.0}, {0.0, -0.5},
                                    {1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0},
                                    {c, c}, {-c, c}, {-c, -c}, {c, -c}};
  imgpoint_t ip;
  imgpoint_init(&ip, a, b, th, m2, m3);
  for (int i = 0; i < n; i++) {
    cx img0[MAXDEG];
    int ni0 = 0;
    for (int k = 0; k < 10; k++) {
      const cx z = mk(roots[20 * i + 2 * k], roots[20 * i + 2 * k + 1]);
      if (img_check(&ip.L, z, mk(pos[2 * i], pos[2 * i + 1]), 1.0e-6)) img0[ni0++] = z;
    }
    const int n0 = significant_images(&ip.L, img0, ni0);
    int nc = 0;
    for (int s = 0; s < 12; s++) {
      for (int k = 0; k < 10; k++) ip.tempRoots[k] = mk(roots[20 * i + 2 * k], roots[20 * i + 2 * k + 1]);
      ip.tempRoots[10] = mk(0.0, 0.0);
      ip.haveRoots = 1;
      out[12 * i + s] = point_amp(&ip, mk(pos[2 * i] + rho * off[s][0], pos[2 * i + 1] + rho * off[s][1]));
      nc += significant_images(&ip.L, ip.imgs, ip.nImg) != n0;
    }
    if (cross) cross[i] = nc;
  }
  return 0;
}

/* ---- critical curves and caustics (ccc.cc) ------------------------------ */

#define CCC_MAXLEN 100000

/* CriticalCurveCaustic::getCC + getCa for a triple lens (ccc.cc:42-98,
 * 160-203).  The coefficient expressions keep the reference's operation
 * order so that the critical curve, and with it the image-plane grid, is the
 * same to the last bit.  The light curve constructs its critical curves
 * from a LensPar (lcirs.cc:16), which leaves the binary flag unset
 * (ccc.cc:28-39): this form also serves m3 = 0. */
static void ccc_compute(const lens_t* L, int len, cx* cc, cx* ca) {
  const cx z2 = L->z2, z3 = L->z3;
  const double m2 = L->m2, m3 = L->m3;
  cx p[7];
  cx z2pz3 = cadd(z2, z3);
  p[1] = cscale(cmul(cmul(cscale(z2, 2.0), z3), z2pz3), m2 + m3 - 1.0);
  p[0] = cscale(cmul(cmul(cmul(cneg(z2), z2), z3), z3), m2 + m3 - 1.0);
  cx tmp[6];
  int have = 0;
  for (int j = 0; j < len; j++) {
    double phi = 2.0 * 3.14159265359 * j / len;
    double sn, cs;
    sincos(phi, &sn, &cs);
    cx e = mk(cs, sn);
    p[6] = cneg(e);
    p[5] = cmul(cscale(e, 2.0), z2pz3);
    /* 1.0 - complex: real 1.0 - re, imag -im (libstdc++ operator-(T, complex)) */
    cx q4 = cmul(e, cadd(cadd(cmul(z2, z2), cmul(z3, z3)), cmul(cscale(z2, 4.0), z3)));
    p[4] = mk(1.0 - q4.re, -q4.im);
    p[3] = cadd(cadd(cmul(cmul(cmul(cscale(e, 2.0), z2), z3), z2pz3),
                     cscale(cscale(z2, 2.0), m2 - 1.0)),
                cscale(cscale(z3, 2.0), m3 - 1.0));
    cx t1 = cmul(cmul(cmul(cmul(cneg(e), z2), z2), z3), z3);
    cx t2 = cmul(cscale(z2, 4.0 * (1.0 - m3 - m2)), z3);
    cx t3 = cscale(cmul(z2, z2), m2 - 1.0);
    cx t4 = cscale(cmul(z3, z3), m3 - 1.0);
    p[2] = csub(csub(cadd(t1, t2), t3), t4);
    if (!have) {
      lag_solve(p, 6, NULL, tmp);
      have = 1;
    } else {
      lag_polish(p, 6, tmp);
      if (!lag_check(p, 6, tmp)) lag_solve(p, 6, NULL, tmp);
    }
    for (int k = 0; k < 6; k++) cc[k * len + j] = tmp[k];
  }
  for (int k = 0; k < 6; k++)
    for (int j = 0; j < len; j++) {
      cx s = cc[k * len + j];
      cx v = csub(s, cdiv_rc(L->m1, cconj(csub(s, L->z1))));
      v = csub(v, cdiv_rc(L->m2, cconj(csub(s, L->z2))));
      v = csub(v, cdiv_rc(L->m3, cconj(csub(s, L->z3))));
      ca[k * len + j] = v;
    }
}

int tlo_ccc(double a, double b, double th, double m2, double m3, int len,
            double* cc, double* ca) {
  if (len < 1 || len > CCC_MAXLEN) return -1;
  lens_t L;
  lens_init(&L, a, b, th, m2, m3);
  cx* c1 = malloc(sizeof(cx) * 6 * len);
  cx* c2 = malloc(sizeof(cx) * 6 * len);
  if (!c1 || !c2) { free(c1); free(c2); return -2; }
  ccc_compute(&L, len, c1, c2);
  memcpy(cc, c1, sizeof(cx) * 6 * len);
  memcpy(ca, c2, sizeof(cx) * 6 * len);
  free(c1); free(c2);
  return 6;
}

/* ------------------------------------------------------------------------ */
/* light curve object                                                         */
/* ------------------------------------------------------------------------ */

#define CCC_LEN 500

struct tlo_lc {
  imgpoint_t ip;
  double rho, rho2, rho2pow;
  unsigned lcLength;
  long ppr;
  long N;
  double L, step, blx, bly, trx, try_, ampScale;
  double ldA, ldB, intensity;
  cx boxMin[6], boxMax[6];
  cx* ca; /* 6 x CCC_LEN caustic samples */
  /* flood-fill state */
  int32_t* head;   /* per row: first range index or -1 (N entries) */
  long* touched; long nTouched, capTouched;
  long* rxl; long* rxr; int32_t* rnext; long nRanges, capRanges;
  long* stk; long nStk, capStk; /* frames of 6 longs */
};

/* utils.h:16-28 */
static void minmax_(const cx* v, int n, cx* mn, cx* mx) {
  for (int i = 0; i < n; i++) {
    if (v[i].re < mn->re) mn->re = v[i].re;
    if (v[i].im < mn->im) mn->im = v[i].im;
    if (v[i].re > mx->re) mx->re = v[i].re;
    if (v[i].im > mx->im) mx->im = v[i].im;
  }
}

static void ld_default(tlo_lc* lc) {
  lc->ldA = 0.6; lc->ldB = 0.4;
  lc->intensity = 1.0 / (lc->ldA + 2.0 / 3.0 * lc->ldB);
}

tlo_lc* tlo_lcirs_new(double a, double b, double th, double m2, double m3,
                      double rho, unsigned lcLength, long ppr) {
  if (!(rho > 0.0) || ppr <= 0 || lcLength == 0) return NULL;
  tlo_lc* lc = calloc(1, sizeof(tlo_lc));
  if (!lc) return NULL;
  imgpoint_init(&lc->ip, a, b, th, m2, m3);
  lc->rho = rho;
  lc->rho2 = rho * rho;
  lc->rho2pow = pow(rho, 2.0);
  lc->lcLength = lcLength;
  lc->ppr = ppr;
  ld_default(lc);
  cx* cc = malloc(sizeof(cx) * 6 * CCC_LEN);
  lc->ca = malloc(sizeof(cx) * 6 * CCC_LEN);
  if (!cc || !lc->ca) { free(cc); free(lc->ca); free(lc); return NULL; }
  ccc_compute(&lc->ip.L, CCC_LEN, cc, lc->ca);
  /* _getCaBoxes (lcirs.cc:39-48): cumulative boxes including the origin */
  cx mn = mk(0.0, 0.0), mx = mk(0.0, 0.0);
  for (int k = 0; k < 6; k++) {
    minmax_(lc->ca + k * CCC_LEN, CCC_LEN, &mn, &mx);
    lc->boxMin[k] = mn;
    lc->boxMax[k] = mx;
  }
  /* _getImgPlanePars (lcirs.cc:51-92) */
  mn = mk(0.0, 0.0); mx = mk(0.0, 0.0);
  for (int k = 0; k < 6; k++) minmax_(cc + k * CCC_LEN, CCC_LEN, &mn, &mx);
  free(cc);
  double halfEdge = mx.re - mn.re;
  if (mx.im - mn.im > halfEdge) halfEdge = mx.im - mn.im;
  halfEdge *= 3.0 / 2.0;
  double cxr = (mx.re + mn.re) / 2.0, cxi = (mx.im + mn.im) / 2.0;
  lc->trx = cxr + halfEdge; lc->try_ = cxi + halfEdge;
  lc->blx = cxr - halfEdge; lc->bly = cxi - halfEdge;
  lc->N = (long)(2 * halfEdge / rho * (double)ppr);
  lc->L = 2 * halfEdge;
  lc->ampScale = lc->intensity / M_PI / pow((double)ppr, 2.0);
  lc->step = lc->L / (double)((double)lc->N - 1.0);
  return lc;
}

int tlo_set_slot_trace(tlo_lc* lc, const double* xy, long n) {
  if (!lc || n < 0) return -1;
  lc->ip.slotTrace = (const cx*)xy;
  lc->ip.slotCount = xy ? n : 0;
  lc->ip.calls = 0;
  return 0;
}

void tlo_lcirs_delete(tlo_lc* lc) {
  if (!lc) return;
  free(lc->ca); free(lc->head); free(lc->touched);
  free(lc->rxl); free(lc->rxr); free(lc->rnext); free(lc->stk);
  free(lc);
}

/* LimbDarkeningModel(v) and setLimbDarkeningModel (limbdarkeningmodel.cc:
 * 18-29, lcirs.cc:94-101) */
int tlo_set_limb_darkening(tlo_lc* lc, int model, double v) {
  if (!lc) return -1;
  if (model == 1) {
    double i0 = 1.0 / ((1.0 - v / 3.0));
    lc->ldA = i0 * (1.0 - v);
    lc->ldB = i0 * v;
    lc->intensity = 1.0;
  } else {
    ld_default(lc);
  }
  lc->ampScale = lc->intensity / M_PI / pow((double)lc->ppr, 2.0);
  return 0;
}

int tlo_grid(const tlo_lc* lc, double* out) {
  if (!lc) return -1;
  out[0] = (double)lc->N; out[1] = lc->L; out[2] = lc->step;
  out[3] = lc->blx; out[4] = lc->bly; out[5] = lc->trx; out[6] = lc->try_;
  out[7] = lc->ampScale;
  for (int k = 0; k < 6; k++) {
    out[8 + 4 * k] = lc->boxMin[k].re; out[9 + 4 * k] = lc->boxMin[k].im;
    out[10 + 4 * k] = lc->boxMax[k].re; out[11 + 4 * k] = lc->boxMax[k].im;
  }
  out[32] = lc->ldA; out[33] = lc->ldB; out[34] = lc->rho2; out[35] = lc->rho2pow;
  return 0;
}

/* ---- per-pixel ray map (lcirs.cc:290-341, 355-378) ----------------------- */

static inline double nx_to_x(const tlo_lc* lc, long nx) { return lc->blx + lc->step * (double)nx; }
static inline double ny_to_y(const tlo_lc* lc, long ny) { return lc->bly + lc->step * (double)ny; }
static inline long x_to_nx(const tlo_lc* lc, double x) {
  return (long)((x - lc->blx) * (double)lc->N / lc->L + 0.5);
}
static inline long y_to_ny(const tlo_lc* lc, double y) {
  return (long)((y - lc->bly) * (double)lc->N / lc->L + 0.5);
}

/* irsSIMD: lanes (z3, z2, z1, zeta) with weights (m3, m2, m1, 1) and the
 * zeta lane's squared distance replaced by -1; the two 128-bit halves are
 * added lane-wise (z3+z1, z2+zeta) and then the pair. */
static inline double irs_ray(const lens_t* L, double x, double y, double sx, double sy) {
  double dr3 = L->z3.re - x, dr2 = L->z2.re - x, dr1 = L->z1.re - x, drs = sx - x;
  double di3 = L->z3.im - y, di2 = L->z2.im - y, di1 = L->z1.im - y, dis = sy - y;
  double q3 = dr3 * dr3 + di3 * di3;
  double q2 = dr2 * dr2 + di2 * di2;
  double q1 = dr1 * dr1 + di1 * di1;
  double c3 = L->m3 / q3, c2 = L->m2 / q2, c1 = L->m1 / q1, cs = 1.0 / -1.0;
  double tr3 = dr3 * c3, tr2 = dr2 * c2, tr1 = dr1 * c1, trs = drs * cs;
  double ti3 = di3 * c3, ti2 = di2 * c2, ti1 = di1 * c1, tis = dis * cs;
  double r0 = tr3 + tr1, r1 = tr2 + trs;
  double i0 = ti3 + ti1, i1 = ti2 + tis;
  return (r0 + r1) * (r0 + r1) + (i0 + i1) * (i0 + i1);
}

/* irsBinary (lcirs.cc:345-353): img - m1/conj(img-z1) - m2/conj(img-z2)
 * with libgcc's complex division, then std::norm (x*x + y*y) of the
 * difference from the source. */
static inline double irs_binary(const lens_t* L, double x, double y, double sx, double sy) {
  cx img = mk(x, y);
  cx t = csub(img, cdiv_rc(L->m1, cconj(csub(img, L->z1))));
  t = csub(t, cdiv_rc(L->m2, cconj(csub(img, L->z2))));
  cx d = csub(t, mk(sx, sy));
  return d.re * d.re + d.im * d.im;
}

/* The ray test of the light-curve object: irsSIMD, or irsBinary for m3 = 0
 * (the dispatch of lcirs.cc:29-36). */
static inline double lc_irs(const tlo_lc* lc, double x, double y, double sx, double sy) {
  return lc->ip.L.m3 == 0.0 ? irs_binary(&lc->ip.L, x, y, sx, sy)
                            : irs_ray(&lc->ip.L, x, y, sx, sy);
}

double tlo_irs(const tlo_lc* lc, double x, double y, double sx, double sy) {
  return lc_irs(lc, x, y, sx, sy);
}

int tlo_hit_mask(const tlo_lc* lc, long x0, long x1, long y0, long y1, double sx, double sy,
                 uint8_t* out) {
  if (!lc || !out || x1 < x0 || y1 < y0) return -1;
  const long w = x1 - x0 + 1;
  for (long j = y0; j <= y1; j++) {
    const double y = ny_to_y(lc, j);
    for (long i = x0; i <= x1; i++)
      out[(j - y0) * w + (i - x0)] = lc_irs(lc, nx_to_x(lc, i), y, sx, sy) < lc->rho2;
  }
  return 0;
}

/* ---- amoeba (amoeba.cc) ------------------------------------------------- */

static int fill_reserve(tlo_lc* lc) {
  if (!lc->head) {
    lc->head = malloc(sizeof(int32_t) * (size_t)lc->N);
    if (!lc->head) return -1;
    for (long i = 0; i < lc->N; i++) lc->head[i] = -1;
  }
  return 0;
}

static void amoeba_reset(tlo_lc* lc) {
  for (long i = 0; i < lc->nTouched; i++) lc->head[lc->touched[i]] = -1;
  lc->nTouched = 0;
  lc->nRanges = 0;
}

static int grow(void** p, long* cap, long need, size_t elem) {
  if (need <= *cap) return 0;
  long nc = *cap ? *cap : 1024;
  while (nc < need) nc *= 2;
  void* q = realloc(*p, (size_t)nc * elem);
  if (!q) return -1;
  *p = q;
  *cap = nc;
  return 0;
}

/* Amoeba::addNode (amoeba.cc:15-30) */
static int amoeba_add(tlo_lc* lc, long nL, long nR, long ny) {
  for (int32_t r = lc->head[ny]; r >= 0; r = lc->rnext[r])
    if (lc->rxr[r] == nR && lc->rxl[r] == nL) return 0;
  long cap = lc->capRanges;
  if (grow((void**)&lc->rxl, &cap, lc->nRanges + 1, sizeof(long))) return -1;
  cap = lc->capRanges;
  if (grow((void**)&lc->rxr, &cap, lc->nRanges + 1, sizeof(long))) return -1;
  cap = lc->capRanges;
  if (grow((void**)&lc->rnext, &cap, lc->nRanges + 1, sizeof(int32_t))) return -1;
  lc->capRanges = cap;
  if (lc->head[ny] < 0) {
    if (grow((void**)&lc->touched, &lc->capTouched, lc->nTouched + 1, sizeof(long))) return -1;
    lc->touched[lc->nTouched++] = ny;
  }
  /* append at the end of the row's list (push_back) */
  long idx = lc->nRanges++;
  lc->rxl[idx] = nL; lc->rxr[idx] = nR; lc->rnext[idx] = -1;
  if (lc->head[ny] < 0) {
    lc->head[ny] = (int32_t)idx;
  } else {
    int32_t r = lc->head[ny];
    while (lc->rnext[r] >= 0) r = lc->rnext[r];
    lc->rnext[r] = (int32_t)idx;
  }
  return 0;
}

/* Amoeba::checkLine (amoeba.cc:34-52): 1 when (ny, nx) is not covered */
static int amoeba_check(const tlo_lc* lc, long ny, long nx) {
  if (ny < 0 || ny >= lc->N) return 1; /* no entry can exist for such a row */
  if (lc->head[ny] < 0) return 1;
  for (int32_t r = lc->head[ny]; r >= 0; r = lc->rnext[r])
    if (nx <= lc->rxr[r] && lc->rxl[r] <= nx) return 0;
  return 1;
}

/* Amoeba::checkLineShift (amoeba.cc:56-82) */
static int amoeba_check_shift(const tlo_lc* lc, long ny, long* nx) {
  if (ny < 0 || ny >= lc->N) { *nx += lc->N; return 0; }
  if (lc->head[ny] < 0) return 1;
  for (int32_t r = lc->head[ny]; r >= 0; r = lc->rnext[r])
    if (*nx <= lc->rxr[r] && lc->rxl[r] <= *nx) { *nx = lc->rxr[r] + 2; return 0; }
  return 1;
}

/* ---- flood fill (lcirs.cc:381-466) -------------------------------------- */

typedef struct {
  const tlo_lc* lc;
  double sx, sy;
  double sum;
  uint64_t count;
  int edge;
} fillacc_t;

static inline int ray_hit(fillacc_t* acc, long nx, double y) {
  const tlo_lc* lc = acc->lc;
  double rsq = lc_irs(lc, nx_to_x(lc, nx), y, acc->sx, acc->sy);
  if (!(rsq < lc->rho2)) return 0;
  rsq = rsq / lc->rho2;
  acc->sum += lc->ldA + lc->ldB * sqrt(1.0 - rsq);
  acc->count++;
  if (nx <= 1 || nx >= lc->N - 2) acc->edge = 1;
  return 1;
}

enum { F_NX, F_NY, F_NL, F_NR, F_NN, F_PH, F_SZ };

/* lineFloodFill(nx, ny, sPos, checked=false) and its recursion.  Frames keep
 * the recursion's locals; the order of pixel visits and of additions to the
 * brightness sum is the recursion's. */
static int flood(tlo_lc* lc, fillacc_t* acc, long nx0, long ny0) {
  const long N = lc->N;
  if (ny0 <= 0 || ny0 >= N) return 0;
  if (!amoeba_check(lc, ny0, nx0)) return 0;
  lc->nStk = 0;
  if (grow((void**)&lc->stk, &lc->capStk, F_SZ, sizeof(long))) return -1;
  long* f = lc->stk;
  f[F_NX] = nx0; f[F_NY] = ny0; f[F_PH] = 0;
  lc->nStk = 1;
  while (lc->nStk > 0) {
    f = lc->stk + (lc->nStk - 1) * F_SZ;
    if (f[F_PH] == 0) {
      long nx = f[F_NX], ny = f[F_NY];
      double y = ny_to_y(lc, ny);
      if (!ray_hit(acc, nx, y)) { lc->nStk--; continue; }
      long nR, nL;
      for (nR = nx + 1; nR < N; nR++)
        if (!ray_hit(acc, nR, y)) { nR--; break; }
      for (nL = nx - 1; nL > 0; nL--)
        if (!ray_hit(acc, nL, y)) { nL++; break; }
      if (amoeba_add(lc, nL, nR, ny)) return -1;
      f[F_NL] = nL; f[F_NR] = nR; f[F_NN] = nL; f[F_PH] = 1;
      continue;
    }
    /* phase 1: row above, phase 2: row below */
    long dir = f[F_PH] == 1 ? 1 : -1;
    if (f[F_NN] <= f[F_NR]) {
      long nn = f[F_NN];
      if (amoeba_check_shift(lc, f[F_NY] + dir, &nn)) {
        f[F_NN] = nn + 1;
        long cy = f[F_NY] + dir;
        if (grow((void**)&lc->stk, &lc->capStk, (lc->nStk + 1) * F_SZ, sizeof(long))) return -1;
        long* c = lc->stk + lc->nStk * F_SZ;
        c[F_NX] = nn; c[F_NY] = cy; c[F_PH] = 0;
        lc->nStk++;
      } else {
        f[F_NN] = nn;
      }
      continue;
    }
    if (f[F_PH] == 1) { f[F_NN] = f[F_NL]; f[F_PH] = 2; continue; }
    lc->nStk--;
  }
  return 0;
}

static void amoeba_stats(const tlo_lc* lc);
static int fill_seeds(tlo_lc* lc, double px, double py, int n, const cx* seeds,
                      uint64_t* rays, double* sum, int* edge) {
  if (fill_reserve(lc)) return -1;
  amoeba_reset(lc);
  fillacc_t acc = {lc, px, py, 0.0, 0, 0};
  for (int k = 0; k < n; k++)
    if (flood(lc, &acc, x_to_nx(lc, seeds[k].re), y_to_ny(lc, seeds[k].im))) return -1;
  amoeba_stats(lc);  /* the amoeba stays until the next fill (tlo_last_segments) */
  *rays = acc.count;
  *sum = acc.sum;
  if (edge) *edge = acc.edge;
  return 0;
}

/* Amoeba statistics of the last fill_seeds call (test instrumentation):
 * out[0] = ranges, out[1] = rows with ranges, out[2] = max ranges in a row,
 * out[3] = min row, out[4] = max row. */
static long g_stats[5];
static void amoeba_stats(const tlo_lc* lc) {
  long mx = 0, mn = -1, mxr = -1;
  for (long i = 0; i < lc->nTouched; i++) {
    long row = lc->touched[i], c = 0;
    for (int32_t r = lc->head[row]; r >= 0; r = lc->rnext[r]) c++;
    if (c > mx) mx = c;
    if (mn < 0 || row < mn) mn = row;
    if (row > mxr) mxr = row;
  }
  g_stats[0] = lc->nRanges; g_stats[1] = lc->nTouched; g_stats[2] = mx;
  g_stats[3] = mn; g_stats[4] = mxr;
}
int tlo_fill_stats(long* out) { for (int i = 0; i < 5; i++) out[i] = g_stats[i]; return 0; }

int tlo_work_counts(long* out, int reset) {
  for (int i = 0; i < 3; i++) out[i] = g_work[i];
  if (reset) for (int i = 0; i < 3; i++) g_work[i] = 0;
  return 0;
}

long tlo_last_segments(const tlo_lc* lc, long* out, long cap) {
  if (!lc) return -1;
  long n = 0;
  if (!lc->head) return 0;
  for (long t = 0; t < lc->nTouched; t++) {
    const long ny = lc->touched[t];
    for (int32_t r = lc->head[ny]; r >= 0; r = lc->rnext[r]) {
      if (out && n < cap) {
        out[3 * n] = ny;
        out[3 * n + 1] = lc->rxl[r];
        out[3 * n + 2] = lc->rxr[r];
      }
      n++;
    }
  }
  return n;
}

int tlo_fill_point(const tlo_lc* clc, double px, double py, int nSeeds,
                   const double* seedXY, uint64_t* rays, double* brightSum, int* edge) {
  tlo_lc* lc = (tlo_lc*)clc;
  if (!lc || nSeeds < 0) return -1;
  cx* s = malloc(sizeof(cx) * (nSeeds > 0 ? nSeeds : 1));
  if (!s) return -2;
  for (int k = 0; k < nSeeds; k++) s[k] = mk(seedXY[2 * k], seedXY[2 * k + 1]);
  int rc = fill_seeds(lc, px, py, nSeeds, s, rays, brightSum, edge);
  free(s);
  return rc;
}

/* ---- caustic seeding (lcirs.cc:131-165,195-255) ------------------------- */

static int intersection_check(const tlo_lc* lc, cx s, cx A, cx B, cx* trial) {
  double rsq = lc->rho2pow;
  cx bs = csub(B, s);
  double bsD = cnorm(bs);
  cx as = csub(A, s);
  double asD = cnorm(as);
  int aOut = asD > rsq, bOut = bsD > rsq;
  if (aOut != bOut) {
    *trial = bOut ? A : B;
    return 1;
  }
  if (aOut && bOut) {
    cx ab = csub(A, B);
    double abD = cnorm(ab);
    double proj = as.re * ab.re + as.im * ab.im;
    if (proj < 0) return 0;
    if (proj > abD) return 0;
    if (proj < lc->rho) {
      *trial = csub(A, cscale(B, fabs(proj) / abD));
      cx st = csub(*trial, s);
      return cnorm(st) < rsq;
    }
    return 0;
  }
  return 0;
}

/* Seeds of one position: its images, then the images of the first trial
 * point of every run of consecutive caustic segments that meet the source
 * (per caustic branch whose cumulative box meets the source square). */
static int point_seeds(tlo_lc* lc, cx pos, cx** seeds, int* cap, int* nImgOut) {
  int n = 0;
  img_get(&lc->ip, pos);
  if (nImgOut) *nImgOut = lc->ip.nImg;
  #define PUSH_IMGS() do { \
      if (n + lc->ip.nImg > *cap) { \
        int nc = *cap ? *cap * 2 : 64; \
        while (nc < n + lc->ip.nImg) nc *= 2; \
        cx* q = realloc(*seeds, sizeof(cx) * nc); \
        if (!q) return -1; \
        *seeds = q; *cap = nc; \
      } \
      for (int k_ = 0; k_ < lc->ip.nImg; k_++) (*seeds)[n++] = lc->ip.imgs[k_]; \
    } while (0)
  PUSH_IMGS();
  for (int r = 0; r < 6; r++) {
    if (pos.re + lc->rho >= lc->boxMin[r].re && pos.re - lc->rho <= lc->boxMax[r].re &&
        pos.im + lc->rho >= lc->boxMin[r].im && pos.im - lc->rho <= lc->boxMax[r].im) {
      int taken = 0;
      const cx* c = lc->ca + r * CCC_LEN;
      for (int k = 0; k < CCC_LEN - 1; k++) {
        cx trial;
        if (intersection_check(lc, pos, c[k], c[k + 1], &trial)) {
          if (!taken) {
            img_get(&lc->ip, trial);
            PUSH_IMGS();
            taken = 1;
          }
        } else {
          taken = 0;
        }
      }
    }
  }
  #undef PUSH_IMGS
  return n;
}

static inline cx traj_pos(cx s, cx e, unsigned i, unsigned len) {
  return cadd(s, cscale(csub(e, s), (double)i / (double)len));
}

int tlo_chain_states(tlo_lc* lc, double sx, double sy, double ex, double ey, unsigned i0,
                     unsigned i1, int fresh, double* states) {
  if (!lc || i1 < i0 || i1 > lc->lcLength) return -1;
  if (fresh) {
    for (int k = 0; k <= MAXDEG; k++) lc->ip.tempRoots[k] = mk(0.0, 0.0);
    lc->ip.haveRoots = 0;
  }
  cx s = mk(sx, sy), e = mk(ex, ey);
  cx* seeds = NULL;
  int cap = 0;
  for (unsigned i = i0; i <= i1; i++) {
    if (point_seeds(lc, traj_pos(s, e, i, lc->lcLength), &seeds, &cap, NULL) < 0) {
      free(seeds);
      return -2;
    }
    double* o = states + (size_t)(i - i0) * (2 * (MAXDEG + 1) + 1);
    for (int k = 0; k <= MAXDEG; k++) {
      o[2 * k] = lc->ip.tempRoots[k].re;
      o[2 * k + 1] = lc->ip.tempRoots[k].im;
    }
    o[2 * (MAXDEG + 1)] = lc->ip.haveRoots;
  }
  free(seeds);
  return 0;
}

int tlo_get_lc_irs(tlo_lc* lc, double sx, double sy, double ex, double ey,
                   double* amp, uint64_t* rays, uint8_t* edge) {
  if (!lc || !amp) return -1;
  cx s = mk(sx, sy), e = mk(ex, ey);
  cx* seeds = NULL;
  int cap = 0;
  for (unsigned i = 0; i <= lc->lcLength; i++) {
    cx pos = traj_pos(s, e, i, lc->lcLength);
    int n = point_seeds(lc, pos, &seeds, &cap, NULL);
    if (n < 0) { free(seeds); return -2; }
    uint64_t cnt; double sum; int ed;
    if (fill_seeds(lc, pos.re, pos.im, n, seeds, &cnt, &sum, &ed)) { free(seeds); return -2; }
    amp[i] = sum * lc->ampScale;
    if (rays) rays[i] = cnt;
    if (edge) edge[i] = (uint8_t)ed;
  }
  free(seeds);
  return 0;
}

int tlo_seeds_margin(tlo_lc* lc, double sx, double sy, double ex, double ey,
                     int maxSeeds, double* seedXY, int* nSeeds, int* nImg, double* margin);
int tlo_seeds(tlo_lc* lc, double sx, double sy, double ex, double ey,
              int maxSeeds, double* seedXY, int* nSeeds, int* nImg) {
  return tlo_seeds_margin(lc, sx, sy, ex, ey, maxSeeds, seedXY, nSeeds, nImg, NULL);
}

/* As tlo_seeds, and margin[i] = the smallest distance in decades between any
 * root residual of point i's solves and the 1e-6 image threshold. */
int tlo_seeds_margin(tlo_lc* lc, double sx, double sy, double ex, double ey,
                     int maxSeeds, double* seedXY, int* nSeeds, int* nImg, double* margin) {
  if (!lc || maxSeeds < 0) return -1;
  cx s = mk(sx, sy), e = mk(ex, ey);
  cx* seeds = NULL;
  int cap = 0;
  for (unsigned i = 0; i <= lc->lcLength; i++) {
    cx pos = traj_pos(s, e, i, lc->lcLength);
    g_margin = 1e300;
    int n = point_seeds(lc, pos, &seeds, &cap, nImg ? &nImg[i] : NULL);
    if (n < 0) { free(seeds); return -2; }
    if (margin) margin[i] = g_margin;
    nSeeds[i] = n;
    for (int k = 0; k < n && k < maxSeeds; k++) {
      seedXY[2 * ((size_t)i * maxSeeds + k)] = seeds[k].re;
      seedXY[2 * ((size_t)i * maxSeeds + k) + 1] = seeds[k].im;
    }
  }
  free(seeds);
  return 0;
}

/* Per point of the seed sequence (tlo_seeds' solves): the number of
 * Laguerre kick steps that read the reference's fractional-step table past
 * its 8 entries (g_frac_oob). */
int tlo_frac_oob(tlo_lc* lc, double sx, double sy, double ex, double ey, int* out) {
  if (!lc || !out) return -1;
  cx s = mk(sx, sy), e = mk(ex, ey);
  cx* seeds = NULL;
  int cap = 0;
  for (unsigned i = 0; i <= lc->lcLength; i++) {
    cx pos = traj_pos(s, e, i, lc->lcLength);
    g_frac_oob = 0;
    if (point_seeds(lc, pos, &seeds, &cap, NULL) < 0) { free(seeds); return -2; }
    out[i] = (int)g_frac_oob;
  }
  free(seeds);
  return 0;
}

/* ---- batch of independent curves over host threads ---------------------- */

typedef struct {
  int ncurves;
  const double* params; const double* traj; const unsigned* lens; const uint64_t* off;
  double* amp; uint64_t* rays;
  int next;
  int err;
  pthread_mutex_t mu;
} batch_t;

static void* batch_worker(void* p) {
  batch_t* B = p;
  for (;;) {
    pthread_mutex_lock(&B->mu);
    int k = B->next++;
    pthread_mutex_unlock(&B->mu);
    if (k >= B->ncurves) break;
    const double* q = B->params + 8 * k;
    tlo_lc* lc = tlo_lcirs_new(q[0], q[1], q[2], q[3], q[4], q[5], B->lens[k], (long)q[6]);
    int rc = lc ? 0 : -1;
    if (lc) {
      tlo_set_limb_darkening(lc, 1, q[7]);
      const double* t = B->traj + 4 * k;
      rc = tlo_get_lc_irs(lc, t[0], t[1], t[2], t[3], B->amp + B->off[k], B->rays + B->off[k], NULL);
      tlo_lcirs_delete(lc);
    }
    if (rc) { pthread_mutex_lock(&B->mu); B->err = rc; pthread_mutex_unlock(&B->mu); }
  }
  return NULL;
}

int tlo_batch(int ncurves, const double* params, const double* traj,
              const unsigned* lens, const uint64_t* off, int nthreads,
              double* amp, uint64_t* rays) {
  if (ncurves < 0 || nthreads < 1) return -1;
  batch_t B = {ncurves, params, traj, lens, off, amp, rays, 0, 0, PTHREAD_MUTEX_INITIALIZER};
  pthread_t* th = malloc(sizeof(pthread_t) * nthreads);
  if (!th) return -2;
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)16 << 20);
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], &attr, batch_worker, &B);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  pthread_attr_destroy(&attr);
  free(th);
  return B.err;
}
