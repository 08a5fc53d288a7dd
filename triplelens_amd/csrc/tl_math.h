// Lens arithmetic shared by the host set-up code and the HIP kernels.
//
// Everything here is compiled with -ffp-contract=off: each expression is the
// sequence of IEEE double operations written, in that order, on the host
// (x86-64) and on gfx950 alike, so the two produce the same bits.  The
// operation order follows the reference where the reference's result has to
// be reproduced exactly:
//   * complex products (ac-bd, ad+bc) and Smith's quotient as g++ -O2 emits
//     them for std::complex<double> (libgcc __divdc3), modulus by the glibc
//     2.35 hypot algorithm, square root by the glibc csqrt algorithm;
//   * Laguerre iteration, root polishing and checks: src/Laguerre/laguerre.cc
//     :18-198;
//   * critical-curve polynomial: src/CCC/ccc.cc:52-65 (same operation order,
//     so that the grid derived from it is the reference's to the last bit);
//   * image test: src/Images/imgpoint.cc:72-90;
//   * caustic intersection test: src/LightCurves/lcirs.cc:195-255;
//   * ray map: src/LightCurves/lcirs.cc:290-341 (irsSIMD lane order).
// The degree-10 image polynomial (src/Images/imgpointcoeff.cc:223-447, in the
// reference's operation order) is in tl_coeffs.h.
#pragma once

#include <math.h>
#include <stdint.h>
#include <float.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TL_HD __host__ __device__ __forceinline__
#else
#define TL_HD inline
#endif

namespace tl {

struct cx { double re, im; };

TL_HD cx mk(double re, double im) { cx z; z.re = re; z.im = im; return z; }
TL_HD cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
TL_HD cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
TL_HD cx cneg(cx a) { return mk(-a.re, -a.im); }
TL_HD cx cconj(cx a) { return mk(a.re, -a.im); }
TL_HD cx cmul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
TL_HD cx cscale(cx a, double t) { return mk(a.re * t, a.im * t); }
TL_HD double cnorm(cx a) { return a.re * a.re + a.im * a.im; }

// Smith's method, as libgcc's __divdc3 for operands in the normal range.
TL_HD cx cdiv(cx n, cx d) {
  double a = n.re, b = n.im, c = d.re, e = d.im, ratio, denom, x, y;
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
TL_HD cx cdiv_rc(double t, cx d) { return cdiv(mk(t, 0.0), d); }

// hypot: Borges' corrected algorithm without fma, glibc 2.35 thresholds.
TL_HD double hyp_kernel(double ax, double ay) {
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

TL_HD double hyp(double x, double y) {
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
TL_HD double cabs_(cx a) { return hyp(a.re, a.im); }

// Branch-free forms of cdiv, hyp and csqrt_ for operands in the normal range,
// with the general routines for everything else.  Each computes the values
// of both sides of the general routine's test and selects one, so the result
// is the general routine's to the bit; a wavefront then runs no divergent
// branches and independent chains can overlap.
TL_HD cx cdiv_sel(cx n, cx d) {
  const double a = n.re, b = n.im, c = d.re, e = d.im;
  const bool ce = fabs(c) < fabs(e);
  const double num = ce ? c : e, den = ce ? e : c;
  const double ratio = num / den;
  const double denom = (num * ratio) + den;
  const double ar = a * ratio, br = b * ratio;
  const double xn = ce ? ar + b : br + a;
  const double yn = ce ? br - a : b - ar;
  return mk(xn / denom, yn / denom);
}

TL_HD double hyp_sel(double x, double y) {
  const double LARGE = 0x1p+511, TINY = 0x1p-511, EPSH = 0x1p-54;
  const double fx = fabs(x), fy = fabs(y);
  const double ax = fx < fy ? fy : fx;
  const double ay = fx < fy ? fx : fy;
  if (!(isfinite(x) && isfinite(y) && !(ax > LARGE) && !(ay < TINY) && !(ay <= ax * EPSH)))
    return hyp(x, y);
  double h = sqrt(ax * ax + ay * ay);
  const double d1 = h - ay;
  const double t1a = ax * (2.0 * d1 - ax);
  const double t2a = (d1 - 2.0 * (ax - ay)) * d1;
  const double d2 = h - ax;
  const double t1b = 2.0 * d2 * (ax - 2.0 * ay);
  const double t2b = (4.0 * d2 - ay) * ay + d2 * d2;
  const bool lo = h <= 2.0 * ay;
  const double t1 = lo ? t1a : t1b, t2 = lo ? t2a : t2b;
  h -= (t1 + t2) / (2.0 * h);
  return h;
}
TL_HD double cabs_sel(cx a) { return hyp_sel(a.re, a.im); }

// Complex square root, glibc csqrt algorithm for finite operands.
TL_HD cx csqrt_(cx z) {
  double re = z.re, im = z.im;
  if (im == 0.0) {
    if (re < 0.0) return mk(0.0, copysign(sqrt(-re), im));
    return mk(fabs(sqrt(re)), copysign(0.0, im));
  }
  if (re == 0.0) {
    double r = fabs(im) >= 2.0 * DBL_MIN ? sqrt(0.5 * fabs(im)) : 0.5 * sqrt(2.0 * fabs(im));
    return mk(r, copysign(r, im));
  }
  int scale = 0;
  if (fabs(re) > DBL_MAX / 4.0) {
    scale = 1;
    re = ldexp(re, -2);
    im = ldexp(im, -2);
  }
  if (fabs(im) > DBL_MAX / 4.0) {
    scale = 1;
    re = fabs(re) >= 4.0 * DBL_MIN ? ldexp(re, -2) : 0.0;
    im = ldexp(im, -2);
  } else if (fabs(re) < 2.0 * DBL_MIN && fabs(im) < 2.0 * DBL_MIN) {
    scale = -((DBL_MANT_DIG + 1) / 2);
    re = ldexp(re, -2 * scale);
    im = ldexp(im, -2 * scale);
  }
  double d = hyp(re, im), r, s;
  if (re > 0.0) {
    r = sqrt(0.5 * (d + re));
    if (scale == 1 && fabs(im) < 1.0) {
      s = im / r;
      r = ldexp(r, scale);
      scale = 0;
    } else {
      s = 0.5 * (im / r);
    }
  } else {
    s = sqrt(0.5 * (d - re));
    if (scale == 1 && fabs(im) < 1.0) {
      r = fabs(im / s);
      s = ldexp(s, scale);
      scale = 0;
    } else {
      r = fabs(0.5 * (im / s));
    }
  }
  if (scale) {
    r = ldexp(r, scale);
    s = ldexp(s, scale);
  }
  return mk(r, copysign(s, im));
}

TL_HD cx csqrt_sel(cx z) {
  const double re = z.re, im = z.im;
  const double QMAX = DBL_MAX / 4.0, TWOMIN = 2.0 * DBL_MIN;
  if (!(im != 0.0 && re != 0.0 && fabs(re) <= QMAX && fabs(im) <= QMAX &&
        !(fabs(re) < TWOMIN && fabs(im) < TWOMIN)))
    return csqrt_(z);
  const double d = hyp_sel(re, im);
  // d + re for re > 0, d - re = d + |re| otherwise: the root of the larger one
  const double q = sqrt(0.5 * (d + fabs(re)));
  const double w = 0.5 * (im / q);
  const bool pos = re > 0.0;
  const double r = pos ? q : fabs(w);
  const double s = pos ? w : q;
  return mk(r, copysign(s, im));
}
TL_HD cx cdiv_rc_sel(double t, cx d) { return cdiv_sel(mk(t, 0.0), d); }

// ---------------------------------------------------------------------------
// Laguerre (laguerre.cc)
// ---------------------------------------------------------------------------
constexpr int kLagMaxIt = 300;
constexpr int kLagMT = 8;
constexpr double kLagEps = DBL_EPSILON;

// cos(i), sin(i) for i = 0..kLagMaxIt, filled on the host with the system libm
// (the reference's values) and copied to the device; used by the rare
// zero-denominator step of the iteration.
struct TrigTable { double c[kLagMaxIt + 1], s[kLagMaxIt + 1]; };

// Laguerre::laguerre (laguerre.cc:18-99).  poly has m+1 coefficients.  The
// fractional-kick table is indexed modulo 8 (the reference reads past its
// 8-element table once i >= 64).
TL_HD bool lag_iter(const cx* poly, int m, cx* xp, const TrigTable* trig) {
  const double frac[8] = {0.5, 0.23, 0.73, 0.13, 0.37, 0.61, 0.89, 1.0};
  cx x = *xp;
  const double md = (double)m;
  if (m == 1) {
    if (cabs_(poly[1]) != 0.0) {
      *xp = cdiv(cneg(poly[0]), poly[1]);
      return true;
    }
    *xp = mk(999.0, 999.0);
    return false;
  }
  for (int i = 1; i <= kLagMaxIt; i++) {
    cx b = poly[m];
    double err = cabs_(b);
    const double ax = cabs_(x);
    cx s = mk(0.0, 0.0), f = mk(0.0, 0.0);
    for (int j = m - 1; j >= 0; j--) {
      s = cadd(cmul(x, s), f);
      f = cadd(cmul(x, f), b);
      b = cadd(cmul(x, b), poly[j]);
      err = cabs_(b) + ax * err;
    }
    err *= kLagEps;
    if (cabs_(b) <= err) { *xp = x; return true; }
    cx g = cdiv(f, b);
    cx h = csub(cmul(g, g), cdiv(cscale(s, 2.0), b));
    cx rad = cscale(csub(cscale(h, md), cmul(g, g)), md - 1.0);
    cx sq = csqrt_(rad);
    cx gp = cadd(g, sq);
    cx gm = csub(g, sq);
    if (cabs_(gp) < cabs_(gm)) gp = gm;
    cx dx;
    if (cabs_(gp) > 0.0) dx = cdiv_rc(md, gp);
    else dx = cscale(mk(trig->c[i], trig->s[i]), 1.0 + cabs_(x));
    cx x1 = csub(x, dx);
    if (x.re == x1.re && x.im == x1.im) { *xp = x; return true; }
    if (i % kLagMT) x = x1;
    else x = csub(x, cmul(mk(frac[(i / kLagMT) & 7], 0.0), dx));
  }
  *xp = x;
  return false;
}

// Laguerre::solveRoots (laguerre.cc:103-167).  init: nullptr (fresh solve) or
// deg+1 values, init[j] seeding the root found at step j = deg..1.
template <int DEG>
TL_HD void lag_solve(const cx* coeffs, const cx* init, cx* roots, const TrigTable* trig) {
  cx tmp[DEG + 1];
  for (int i = 0; i <= DEG; i++) tmp[i] = coeffs[i];
  int nr = 0;
  for (int j = DEG; j >= 1; j--) {
    cx x = init ? init[j] : mk(kLagEps, kLagEps);
    if (!lag_iter(tmp, j, &x, trig) && init) x = mk(kLagEps, kLagEps);
    roots[nr++] = x;
    cx b = tmp[j];
    for (int jj = j - 1; jj >= 0; jj--) {
      cx c = tmp[jj];
      tmp[jj] = b;
      b = cadd(cmul(x, b), c);
    }
  }
  for (int j = 0; j < DEG; j++) lag_iter(coeffs, DEG, &roots[j], trig);
}

template <int DEG>
TL_HD void lag_polish(const cx* coeffs, cx* roots, const TrigTable* trig) {
  for (int j = 0; j < DEG; j++) lag_iter(coeffs, DEG, &roots[j], trig);
}

// Laguerre::checkRoots (laguerre.cc:188-198)
template <int DEG>
TL_HD bool lag_check(const cx* coeffs, const cx* roots) {
  cx c0 = coeffs[DEG];
  for (int j = 0; j < DEG; j++) c0 = cmul(c0, cneg(roots[j]));
  if (cabs_(coeffs[0]) > 1.0e-8) return (1.0 - cabs_(cdiv(c0, coeffs[0]))) < 1.0e-5;
  return cabs_(csub(c0, coeffs[0])) < 1.0e-10;
}

// ---------------------------------------------------------------------------
// Lens
// ---------------------------------------------------------------------------
struct Lens {
  double m1, m2, m3;
  cx z1, z2, z3;
};

// cos and sin of one argument as the reference gets them: g++ merges the two
// calls into one glibc sincos() call, whose results differ from separate
// cos() and sin() calls in the last bit for some arguments.  Host only.
inline void host_sincos(double x, double* s, double* c) { ::sincos(x, s, c); }

// Lens::Lens, Lens::_setPos (lens.cc:4-18,35-40); host only.
inline Lens make_lens(double a, double b, double th, double m2, double m3) {
  Lens L;
  L.m1 = 1.0 - m2 - m3;
  L.m2 = m2;
  L.m3 = m3;
  L.z1 = mk(0.0, 0.0);
  L.z2 = mk(a, 0.0);
  double s, c;
  host_sincos(th, &s, &c);
  L.z3 = mk(b * c, b * s);
  return L;
}

// ---- critical curves and caustics -----------------------------------------
constexpr int kCccSteps = 500;  // critical-curve samples per branch (lcirs.cc:16)

// exp(i phi_j) for the critical-curve steps, filled on the host with the
// system sincos (the reference's values).
struct CccTable { cx e[kCccSteps]; };

// CriticalCurveCaustic::getCC + getCa (ccc.cc:42-98,160-203): the degree-6
// critical-curve polynomial in the reference's operation order, solved at
// phi_j = 2*3.14159265359*j/kCccSteps with root polishing continued from step
// to step, and the caustic as the lens map of every critical point.  cc and
// ca receive 6 x kCccSteps values, branch-major.  The light curve builds
// its CriticalCurveCaustic from a LensPar (lcirs.cc:16), a constructor that
// leaves the binary flag unset (ccc.cc:28-39), so this triple-lens form
// also serves m3 = 0: two of its branches then sit at the massless z3.
TL_HD void critical_curves_into(const Lens& L, const CccTable& tab, const TrigTable* trig,
                                cx* cc, cx* ca) {
  const int len = kCccSteps;
  const cx z2 = L.z2, z3 = L.z3;
  const double m2 = L.m2, m3 = L.m3;
  cx p[7];
  const cx z2pz3 = cadd(z2, z3);
  p[1] = cscale(cmul(cmul(cscale(z2, 2.0), z3), z2pz3), m2 + m3 - 1.0);
  p[0] = cscale(cmul(cmul(cmul(cneg(z2), z2), z3), z3), m2 + m3 - 1.0);
  cx roots[6];
  bool have = false;
  for (int j = 0; j < len; j++) {
    const cx e = tab.e[j];
    p[6] = cneg(e);
    p[5] = cmul(cscale(e, 2.0), z2pz3);
    const cx q4 = cmul(e, cadd(cadd(cmul(z2, z2), cmul(z3, z3)), cmul(cscale(z2, 4.0), z3)));
    p[4] = mk(1.0 - q4.re, -q4.im);
    p[3] = cadd(cadd(cmul(cmul(cmul(cscale(e, 2.0), z2), z3), z2pz3),
                     cscale(cscale(z2, 2.0), m2 - 1.0)),
                cscale(cscale(z3, 2.0), m3 - 1.0));
    const cx t1 = cmul(cmul(cmul(cmul(cneg(e), z2), z2), z3), z3);
    const cx t2 = cmul(cscale(z2, 4.0 * (1.0 - m3 - m2)), z3);
    const cx t3 = cscale(cmul(z2, z2), m2 - 1.0);
    const cx t4 = cscale(cmul(z3, z3), m3 - 1.0);
    p[2] = csub(csub(cadd(t1, t2), t3), t4);
    if (!have) {
      lag_solve<6>(p, nullptr, roots, trig);
      have = true;
    } else {
      lag_polish<6>(p, roots, trig);
      if (!lag_check<6>(p, roots)) lag_solve<6>(p, nullptr, roots, trig);
    }
    for (int k = 0; k < 6; k++) cc[k * len + j] = roots[k];
  }
  for (int k = 0; k < 6; k++)
    for (int j = 0; j < len; j++) {
      const cx s = cc[k * len + j];
      cx v = csub(s, cdiv_rc(L.m1, cconj(csub(s, L.z1))));
      v = csub(v, cdiv_rc(L.m2, cconj(csub(s, L.z2))));
      v = csub(v, cdiv_rc(L.m3, cconj(csub(s, L.z3))));
      ca[k * len + j] = v;
    }
}

// ImgPoint::imgCheck (imgpoint.cc:72-90)
TL_HD bool img_check(const Lens& L, cx img, cx zeta, double tol) {
  cx t = csub(img, cdiv_rc(L.m1, cconj(csub(img, L.z1))));
  t = csub(t, cdiv_rc(L.m2, cconj(csub(img, L.z2))));
  t = csub(t, cdiv_rc(L.m3, cconj(csub(img, L.z3))));
  return cabs_(csub(t, zeta)) < tol;
}

// Root-reuse state of one ImgPoint (imgpoint.cc:160-205): tempRoots[10] is
// the slot the reference reads one past the end of its vector (0 here unless
// a recorded value is supplied); calls counts the getImages calls made.
struct RootChain {
  cx tempRoots[11];
  int have;
  int calls;
};

// LightCurveIRS::intersectionCheck (lcirs.cc:195-255)
TL_HD bool intersection_check(cx s, cx A, cx B, double rho, double rho2pow, cx* trial) {
  cx bs = csub(B, s);
  double bsD = cnorm(bs);
  cx as = csub(A, s);
  double asD = cnorm(as);
  bool aOut = asD > rho2pow, bOut = bsD > rho2pow;
  if (aOut != bOut) {
    *trial = bOut ? A : B;
    return true;
  }
  if (aOut && bOut) {
    cx ab = csub(A, B);
    double abD = cnorm(ab);
    double proj = as.re * ab.re + as.im * ab.im;
    if (proj < 0) return false;
    if (proj > abD) return false;
    if (proj < rho) {
      *trial = csub(A, cscale(B, fabs(proj) / abD));
      cx st = csub(*trial, s);
      return cnorm(st) < rho2pow;
    }
    return false;
  }
  return false;
}

// Position of point i of a trajectory (lcirs.cc:120).
TL_HD cx traj_pos(cx s, cx e, unsigned i, unsigned len) {
  return cadd(s, cscale(csub(e, s), (double)i / (double)len));
}

// Squared source-plane distance of the ray through (x, y) in the irsSIMD
// lane order (lcirs.cc:290-341): lanes (z3, z2, z1, zeta), weights
// (m3, m2, m1, 1), the zeta lane's squared distance replaced by -1, halves
// added lane-wise (z3 + z1, z2 + zeta) and then together.
// LightCurveIRS::irsBinary (lcirs.cc:345-353): img - m1/conj(img-z1) -
// m2/conj(img-z2) in std::complex arithmetic (each real/complex quotient is
// libgcc's complex division), and std::norm of its distance to the source.
TL_HD double irs_binary(const Lens& L, double x, double y, double sx, double sy) {
  const cx img = mk(x, y);
  cx t = csub(img, cdiv_rc_sel(L.m1, cconj(csub(img, L.z1))));
  t = csub(t, cdiv_rc_sel(L.m2, cconj(csub(img, L.z2))));
  const double dr = t.re - sx, di = t.im - sy;
  return dr * dr + di * di;
}

TL_HD double irs_ray(const Lens& L, double x, double y, double sx, double sy) {
  double dr3 = L.z3.re - x, dr2 = L.z2.re - x, dr1 = L.z1.re - x, drs = sx - x;
  double di3 = L.z3.im - y, di2 = L.z2.im - y, di1 = L.z1.im - y, dis = sy - y;
  double q3 = dr3 * dr3 + di3 * di3;
  double q2 = dr2 * dr2 + di2 * di2;
  double q1 = dr1 * dr1 + di1 * di1;
  double c3 = L.m3 / q3, c2 = L.m2 / q2, c1 = L.m1 / q1;
  double tr3 = dr3 * c3, tr2 = dr2 * c2, tr1 = dr1 * c1, trs = -drs;
  double ti3 = di3 * c3, ti2 = di2 * c2, ti1 = di1 * c1, tis = -dis;
  double r0 = tr3 + tr1, r1 = tr2 + trs;
  double i0 = ti3 + ti1, i1 = ti2 + tis;
  return (r0 + r1) * (r0 + r1) + (i0 + i1) * (i0 + i1);
}

}  // namespace tl
