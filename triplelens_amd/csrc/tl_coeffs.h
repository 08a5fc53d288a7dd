// The degree-10 image polynomial, rounded as the reference rounds it.
//
// ImgPoint::getRootsPrecalculated (src/Images/imgpoint.cc:160-176) adds two
// generated expressions per coefficient: the part free of the source position,
// ImgPoint::getCoeffsOptNoZ (src/Images/imgpointcoeff.cc:223-278), cached per
// lens, and the position part, ImgPoint::getCoeffsOptJustZ
// (imgpointcoeff.cc:280-447).  Which roots the chained Laguerre solve finds and
// whether a root passes the 1e-6 image test depend on the last bits of these
// coefficients, so both are evaluated here in the reference's operation
// order, with the reference's generated temporaries.  Each operator of `pc`
// is what g++ emits for the same std::complex<double> operation (libstdc++):
// complex products as (ac - bd, ad + bc), real scalings component-wise, a
// real added to or subtracted from the real part only.  With
// -ffp-contract=off every one is a single IEEE double operation, on the host
// and on gfx950 alike, so both give the reference's bits.
#pragma once

#include "tl_math.h"

namespace tl {

// std::complex<double> as the generated expressions use it, with implicit
// conversion from double (imaginary part 0).
struct pc {
  double re, im;
  TL_HD pc(double r = 0.0, double i = 0.0) : re(r), im(i) {}
  TL_HD explicit pc(cx z) : re(z.re), im(z.im) {}
  TL_HD cx c() const { return mk(re, im); }
};
TL_HD pc operator-(pc a) { return pc(-a.re, -a.im); }
TL_HD pc operator+(pc a, pc b) { return pc(a.re + b.re, a.im + b.im); }
TL_HD pc operator-(pc a, pc b) { return pc(a.re - b.re, a.im - b.im); }
TL_HD pc operator*(pc a, pc b) {
  return pc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
TL_HD pc operator*(pc a, double t) { return pc(a.re * t, a.im * t); }
TL_HD pc operator*(double t, pc a) { return pc(a.re * t, a.im * t); }
TL_HD pc operator+(pc a, double t) { return pc(a.re + t, a.im); }
TL_HD pc operator+(double t, pc a) { return pc(a.re + t, a.im); }
TL_HD pc operator-(pc a, double t) { return pc(a.re - t, a.im); }
TL_HD pc operator-(double t, pc a) { return pc(-a.re + t, -a.im); }
TL_HD pc cconj(pc a) { return pc(a.re, -a.im); }

// ImgPoint::getCoeffsOptNoZ (imgpointcoeff.cc:223-278): out[k] multiplies z^k.
TL_HD void image_poly_static(const Lens& L, cx* out) {
  const pc z2(L.z2), z3(L.z3);
  const double m2 = L.m2, m3 = L.m3;

  const pc z3c = cconj(z3);
  const pc a10zl = 0.0;
  const double t1 = -m2 - m3;
  const pc a9zl = t1 * z3c * z2;
  const double t3 = 0.3e1 * m3;
  const double t4 = 0.2e1 * m2;
  const pc t7 = z2 * z2;
  const double t9 = 0.2e1 * m3;
  const double t10 = 0.3e1 * m2;
  const pc t16 = z3c * m3;
  const pc a8zl = (t3 + t4) * z3c * t7 + ((t9 + t10) * z3c * z3 + m2) * z2 + t16;
  const pc t19 = t7 * z2;
  const pc t30 = z3 * z3;
  const double t32 = m2 * m3;
  const pc a7zl = (-t3 - m2) * z3c * t19 +
                     (0.6e1 * t1 * z3c * z3 + m2 * (-0.2e1 + m2)) * t7 +
                     ((-t10 - m3) * z3c * t30 + (m3 - t10 + t32) * z3 + (m2 - t3 + t32) * z3c) * z2 +
                     m3 * (-0.2e1 + m3) * z3c * z3;
  const pc t43 = t7 * t7;
  const pc t44 = t43 * z3c;
  const double t46 = 0.6e1 * m3;
  const double t50 = m2 - 0.1e1;
  const double t51 = m2 * t50;
  const double t54 = 0.6e1 * m2;
  const double t58 = m2 * m2;
  const double t59 = 0.3e1 * t58;
  const pc t62 = 0.2e1 * t32;
  const pc t68 = t30 * z3;
  const pc t69 = z3c * m2 * t68;
  const double t70 = m3 * m3;
  const pc t73 = 0.3e1 * t70;
  const pc t79 = -0.1e1 + m3;
  const pc t80 = m3 * t79;
  const pc t83 = m3 * z3;
  const pc a6zl =
      t44 * m3 + ((t46 + t10) * z3c * z3 - t51) * t19 +
      ((t54 + t3) * z3c * t30 + (-t59 + t54 - t3 - t32) * z3 + (-t4 + t58 - t62 + t3) * z3c) * t7 +
      (t69 + (t10 - t9 - t62 + t70) * t30 + (-t32 + t46 - t10 - t73) * z3c * z3 - m2) * z2 -
      t80 * z3c * t30 - t83;
  const double t91 = m2 - m3;
  const pc a5zl =
      -0.2e1 * t44 * t83 +
      (0.3e1 * t1 * z3c * t30 + (t3 - t10 + t59 - t32) * z3 - t50 * t91 * z3c) * t19 +
      (-0.2e1 * t69 + (t46 + t59 + t62 - t54 - t73) * t30 +
       (-t46 + t62 + t54 + t73 - t59) * z3c * z3 - 0.2e1 * t51) * t7 +
      (t79 * t91 * t68 + (-t3 - t32 + t73 + t10) * z3c * t30 + (t10 + t3 - 0.4e1 * t32) * z3) * z2 -
      0.2e1 * t80 * t30;
  const pc t127 = t50 * t50;
  const pc t136 = 0.8e1 * t32;
  const pc t137 = t58 * m3;
  const pc t139 = 0.6e1 * t58;
  const pc t145 = z3c * t68;
  const pc t147 = 0.6e1 * t70;
  const double t148 = m2 * t70;
  const pc t154 = t79 * t79;
  const pc a4zl =
      (t16 * t30 + m3 * t50 * z3) * t43 +
      (t69 + (t73 + t62 - t46 + t10 - t59) * t30 + (-t70 + t9 + t59 - t32 - t10) * z3c * z3 -
       m2 * t127) * t19 +
      ((t4 - t3 - t32 - t58 + t73) * t68 + (t62 + t3 - t73 - t54 + t59) * z3c * t30 +
       (t136 - t3 - t54 - 0.3e1 * t137 + t139) * z3) * t7 +
      (m2 * t79 * t145 + (-t10 + t147 - 0.3e1 * t148 + t136 - t46) * t30) * z2 -
      m3 * t154 * t68;
  const pc t166 = 0.3e1 * t32;
  const pc a3zl =
      -m3 * (m3 - 0.2e1 + t4) * t30 * t43 +
      ((-t32 + t3 - m2 - t73 + t58) * t68 + (-m3 - t59 + t70 - t32 + t10) * z3c * t30 +
       t50 * (t59 - t10 + t166 - m3) * z3) * t19 +
      (-m2 * (m2 + t9 - 0.2e1) * t145 +
       (0.6e1 * t148 - 0.16e2 * t32 - t147 - t139 + t54 + 0.6e1 * t137 + t46) * t30) * t7 +
      t79 * (t73 + t166 - t3 - m2) * t68 * z2;
  const double t186 = -0.1e1 + m2 + m3;
  const pc a2zl = t68 * t186 * m3 * t43 +
                     (t186 * z3c * m2 * t68 - t186 * (t166 - t10 + t59 - t9) * t30) * t19 -
                     t186 * (t73 + t166 - t3 - t4) * t68 * t7;
  const pc t202 = t186 * t186;
  const pc a1zl = -t202 * t1 * t68 * t19;
  const pc a0zl = 0.0;
  out[0] = a0zl.c(); out[1] = a1zl.c(); out[2] = a2zl.c(); out[3] = a3zl.c(); out[4] = a4zl.c(); out[5] = a5zl.c();
  out[6] = a6zl.c(); out[7] = a7zl.c(); out[8] = a8zl.c(); out[9] = a9zl.c(); out[10] = a10zl.c();
}

// ImgPoint::getCoeffsOptJustZ (imgpointcoeff.cc:280-447): out[k] multiplies
// z^k.
TL_HD void image_poly_zeta(const Lens& L, cx zeta_, cx* out) {
  const pc z2(L.z2), z3(L.z3), zeta(zeta_);
  const double m2 = L.m2, m3 = L.m3;

  const pc z3c = cconj(z3);
  const pc zetac = cconj(zeta);

  const pc t1 = zetac * z3c;
  const pc t2 = zetac * zetac;
  const pc t3 = -t1 + t2;
  const pc t5 = t2 * zetac;
  const pc t6 = t2 * z3c;
  const pc a10 = t3 * z2 - t5 + t6;
  const pc t7 = -t3;
  const pc t8 = 0.3e1 * t7;
  const pc t9 = z2 * z2;
  const pc t12 = 0.3e1 * zetac;
  const pc t15 = zetac * (-t12 + zeta) * z3c;
  const pc t16 = zeta * zetac;
  const pc t17 = 0.3e1 * t2;
  const pc t22 = -t6 + t5;
  const pc t23 = 0.3e1 * t22;
  const pc a9 = t8 * t9 + (t8 * z3 + t15 + zetac * (-t16 + t17 + m2 + 0.1e1)) * z2 + t23 * z3 +
                   zetac * (0.1e1 - t16 + m3) * z3c + t2 * (t16 - 0.2e1);
  const pc t30 = -t8;
  const pc t31 = t9 * z2;
  const pc t33 = 0.9e1 * t3;
  const pc t37 = zetac * (-zetac + zeta) * z3c;
  const pc t38 = 0.3e1 * t37;
  const pc t39 = 0.3e1 * t16;
  const pc t44 = z3 * z3;
  const pc t46 = 0.3e1 * t15;
  const pc t47 = 0.9e1 * t2;
  const pc t48 = 0.2e1 * m3;
  const pc t49 = 0.3e1 * m2;
  const pc t54 = zetac * m2;
  const pc t55 = 0.2e1 * t54;
  const pc t56 = zeta * t2;
  const pc t57 = 0.3e1 * t56;
  const pc t58 = zetac * m3;
  const pc t59 = 0.3e1 * t58;
  const pc t62 = 0.6e1 * zetac;
  const pc t63 = 0.3e1 * t54;
  const pc t64 = 0.2e1 * zeta;
  const pc t79 = t16 * z3c;
  const pc a8 =
      t30 * t31 + (t33 * z3 - t38 - zetac * (t17 - t39 + 0.3e1 + m2)) * t9 +
      (t30 * t44 + (-t46 - zetac * (t47 - t39 - t48 + t49 + 0.3e1)) * z3 +
       (zeta + t55 + t57 - t59 - t12) * z3c - zetac * (t57 - t62 + t63 + t64)) * z2 -
      t23 * t44 +
      (-zetac * (-t39 + m3 + 0.3e1) * z3c - 0.3e1 * t2 * (t16 + m3 - 0.2e1)) * z3 -
      0.2e1 * t79 + zetac * (t39 - 0.1e1);
  const pc t83 = t9 * t9;
  const pc t85 = -t33;
  const pc t87 = 0.3e1 * zeta;
  const pc t90 = zetac * (-zetac + t87) * z3c;
  const pc t96 = 0.9e1 * t37;
  const pc t102 = m2 * zeta;
  const pc t103 = 0.4e1 * t54;
  const pc t106 = 0.6e1 * t54;
  const pc t107 = 0.2e1 * t102;
  const pc t108 = 0.6e1 * zeta;
  const pc t113 = t44 * z3;
  const pc t115 = 0.4e1 * m3;
  const pc t120 = 0.9e1 * zetac;
  const pc t121 = zeta * m3;
  const pc t122 = 0.9e1 * t56;
  const pc t125 = 0.18e2 * zetac;
  const pc t126 = 0.9e1 * t58;
  const pc t127 = 0.9e1 * t54;
  const pc t128 = 0.2e1 * t121;
  const pc t133 = -0.3e1 + m2;
  const pc t134 = t133 * z3c;
  const pc t137 = t56 * m2;
  const pc t138 = 0.3e1 * t137;
  const pc t150 = m3 - 0.3e1;
  const pc t151 = t150 * z3c;
  const pc t154 = t16 * m3;
  const pc t156 = 0.9e1 * t16;
  const pc a7 =
      t7 * t83 + (t85 * z3 + t90 - zetac * (-t2 + t39 - 0.3e1 + m2)) * t31 +
      (t85 * t44 + (t96 + 0.3e1 * zetac * (t17 - t39 - t48 + 0.3e1 + m2)) * z3 +
       (t102 - t103 - t57 + t12 - t87 + t59) * z3c - zetac * (-t57 - t106 + t62 + t107 - t108)) * t9 +
      (t7 * t113 + (t46 + zetac * (t47 - t39 + t49 + 0.3e1 - t115)) * t44 +
       ((t120 + t121 - t87 - t106 - t122 + t59) * z3c +
        zetac * (t122 - t125 + t126 + t127 - t128 + t108)) * z3 -
       0.2e1 * t16 * t134 - t122 - zeta + t138 - t103 + t12) * z2 +
      t22 * t113 +
      (-zetac * (t39 + m3 - 0.3e1) * z3c + 0.3e1 * t2 * (t16 + t48 - 0.2e1)) * t44 +
      (-0.2e1 * t16 * t151 + zetac * (0.3e1 * t154 - t156 + 0.3e1 - t115)) * z3 - zeta * z3c + t39;
  const pc t175 = 0.2e1 * zetac;
  const pc t176 = 0.4e1 * t102;
  const pc t187 = 0.12e2 * t54;
  const pc t188 = 0.9e1 * zeta;
  const pc t189 = 0.3e1 * t102;
  const pc t190 = 0.3e1 * t121;
  const pc t198 = 0.2e1 * m2;
  const pc t199 = -0.3e1 + t198;
  const pc t203 = m2 * m2;
  const pc t204 = zetac * t203;
  const pc t205 = 0.3e1 * t204;
  const pc t206 = 0.8e1 * t54;
  const pc t207 = 0.6e1 * t137;
  const pc t208 =
      t30 * t113 + (-t96 - 0.3e1 * zetac * (t17 - t39 + m2 - t115 + 0.3e1)) * t44 +
      ((t187 + t188 - t120 + t122 - t189 - t190 - t59) * z3c +
       0.3e1 * zetac * (-t57 + t62 - t59 - t106 - t108 + t128 + t107)) * z3 +
      0.2e1 * t16 * t199 * z3c - t12 - t107 + t87 - t205 + t122 + t206 - t207;
  const pc t217 = 0.4e1 * t121;
  const pc t226 = t54 * m3;
  const pc t227 = 0.6e1 * t226;
  const pc t228 = 0.9e1 * t137;
  const pc t229 = 0.27e2 * t56;
  const pc t230 = t56 * m3;
  const pc t231 = 0.9e1 * t230;
  const pc t232 = 0.12e2 * t58;
  const pc t244 = 0.3e1 * m3;
  const pc t249 = t48 - 0.3e1;
  const pc t254 = m3 * m3;
  const pc t255 = 0.3e1 * t254;
  const pc a6 =
      (t30 * z3 - t79 + zetac * (t16 + m2 - 0.1e1)) * t83 +
      (t33 * t44 + (-0.3e1 * t90 + 0.3e1 * zetac * (-t2 + t39 + m2 - 0.3e1 + t48)) * z3 +
       (-zetac - t58 + t87 + t55 + t56 - t107) * z3c +
       zetac * (-t56 + t175 - t63 - t108 + t176)) * t31 +
      t208 * t9 +
      ((-t15 - zetac * (t17 - t16 + 0.1e1 - t48 + m2)) * t113 +
       ((-t128 + t122 + t87 + t106 - t120 + t59) * z3c -
        zetac * (t122 - t125 + 0.18e2 * t58 + t127 - t217 + t108)) * t44 +
       (0.6e1 * t16 * (m3 + m2 - 0.3e1) * z3c - t120 - t227 + t187 - t228 + t229 - t231 + t232 -
        t128 + t87) * z3 -
       zeta * t199 * z3c + 0.3e1 * t16 * t199) * z2 +
      (zetac * (t16 - 0.1e1 + m3) * z3c - t2 * (t16 + t244 - 0.2e1)) * t113 +
      (0.2e1 * t16 * t249 * z3c - zetac * (-t156 + 0.6e1 * t154 + t255 + 0.3e1 - 0.8e1 * m3)) * t44 +
      (-zeta * t249 * z3c + 0.3e1 * t16 * t249) * z3 + zeta;
  const pc t268 = 0.3e1 * t79;
  const pc t273 = m2 - 0.1e1;
  const pc t274 = zeta * t273;
  const pc t286 = 0.6e1 * t102;
  const pc t301 = 0.6e1 * m3;
  const pc t306 = 0.6e1 * t121;
  const pc t319 = 0.12e2 * t226;
  const pc t320 = t102 * m3;
  const pc t321 = 0.2e1 * t320;
  const pc t322 = 0.24e2 * t54;
  const pc t323 = 0.18e2 * t137;
  const pc t324 = 0.9e1 * t204;
  const pc t325 = -0.6e1 * t16 * (m3 + t198 - 0.3e1) * z3c + t120 - t188 + t286 + t319 - t229 -
                     t321 + t231 - t232 - t322 + t306 + t323 + t324;
  const pc t343 = zeta * t254;
  const pc t344 = zetac * t254;
  const pc t345 = 0.9e1 * t344;
  const pc t346 = 0.24e2 * t58;
  const pc t347 = 0.18e2 * t230;
  const pc t348 = -0.6e1 * t16 * (m2 - 0.3e1 + t48) * z3c + t217 - t343 - t87 + t120 - t187 +
                     t345 - t346 + t228 + t319 + t347 - t229;
  const pc t350 = m2 * m3;
  const pc t351 = 0.2e1 * t350;
  const pc t352 = 0.6e1 * m2;
  const pc t353 = 0.9e1 + t351 - t352 - t301;
  const pc t363 = -0.1e1 + m3;
  const pc t372 = zeta * t363;
  const pc a5 =
      (t8 * t44 + (t268 - zetac * (t39 + t49 + t48 - 0.3e1)) * z3 + t274 * z3c -
       0.2e1 * t16 * t273) * t83 +
      (t8 * t113 + (0.3e1 * t90 - 0.3e1 * zetac * (-t2 + t39 + m2 - 0.3e1 + t115)) * t44 +
       ((-t188 + t58 - t106 - t57 + t286 + t190 + t12) * z3c -
        0.3e1 * zetac * (-t56 - t58 - t63 + t175 + t128 - t108 + t176)) * z3 -
       0.2e1 * t16 * t273 * z3c - t273 * (-t63 + t102 - t57 - t87 + zetac)) * t31 +
      ((t38 + zetac * (t17 - t39 - t301 + m2 + 0.3e1)) * t113 +
       ((t306 - t59 - t122 + t120 - t188 + t189 - t187) * z3c -
        0.3e1 * zetac * (-t57 - t106 + t62 - 0.6e1 * t58 - t108 + t217 + t107)) * t44 +
       t325 * z3 - t274 * t134 + 0.3e1 * t16 * t273 * t133) * t9 +
      (((t12 - zeta + t121 - t59 - t57 - t55) * z3c +
        zetac * (t57 - t62 + t63 + t126 - t128 + t64)) * t113 +
       t348 * t44 + (-zeta * t353 * z3c + 0.3e1 * t16 * t353) * z3 + 0.3e1 * t274) * z2 +
      (-0.2e1 * t16 * t363 * z3c + t363 * zetac * (t244 + t39 - 0.1e1)) * t113 +
      (-t372 * t151 + 0.3e1 * t16 * t363 * t150) * t44 + 0.3e1 * t372 * z3;
  const pc t386 = t49 + m3 - 0.3e1;
  const pc t393 = t273 * t273;
  const pc t394 = zeta * t393;
  const pc t409 = t386 * z3c;
  const pc t412 = 0.12e2 * t102;
  const pc t413 = 0.4e1 * t320;
  const pc t417 = 0.3e1 * t203 * zeta;
  const pc t418 = 0.2e1 * t16 * t409 + t188 - t324 - t412 - t227 + t122 + t413 - 0.3e1 * t230 +
                     0.4e1 * t58 - t12 - t306 + t417 + t187 - t228;
  const pc t435 = 0.12e2 * t121;
  const pc t437 = 0.3e1 * t343;
  const pc t438 = 0.6e1 * t16 * (t198 + t48 - 0.3e1) * z3c + t346 - t286 - t345 - t435 - t323 +
                     t413 + t229 - 0.24e2 * t226 - t120 - t324 + t437 - t347 + t188 + t322;
  const pc t440 = 0.3e1 * t203;
  const pc t441 = 0.4e1 * t350;
  const pc t442 = 0.12e2 * m2;
  const pc t443 = 0.9e1 - t301 + t440 + t441 - t442;
  const pc t453 = -0.3e1 + t244 + m2;
  const pc t454 = t453 * z3c;
  const pc t457 = 0.2e1 * t16 * t454 - t138 - t231 - t227 + t122 - t12 + t232 - t345 + t103 +
                     zeta - t128 + t343;
  const pc t459 = 0.12e2 * m3;
  const pc t460 = 0.9e1 - t459 + t441 - t352 + t255;
  const pc t467 = t351 - t49 + 0.3e1 - t244;
  const pc t473 = t363 * t363;
  const pc t474 = zeta * t473;
  const pc a4 =
      (t3 * t113 + (-t268 + zetac * (t39 - 0.3e1 + t115 + t49)) * t44 +
       (-zeta * t386 * z3c + 0.2e1 * t16 * t386) * z3 + t394) * t83 +
      ((-t90 + zetac * (-t2 + t39 + m2 - 0.3e1 + t301)) * t113 +
       ((t57 - t306 + t58 - t12 + t188 + t106 - t286) * z3c +
        0.3e1 * zetac * (-t56 - t63 - 0.2e1 * t58 + t175 - t108 + t217 + t176)) * t44 +
       t418 * z3 + t394 * z3c - 0.3e1 * t16 * t393) * t31 +
      (((t57 + t103 - t102 + t59 + t87 - t190 - t12) * z3c +
        zetac * (-t57 + t62 - t106 - t126 - t108 + t306 + t107)) * t113 +
       t438 * t44 + (zeta * t443 * z3c - 0.3e1 * t16 * t443) * z3 + 0.3e1 * t394) * t9 +
      (t457 * t113 + (zeta * t460 * z3c - 0.3e1 * t16 * t460) * t44 + 0.3e1 * zeta * t467 * z3) * z2 +
      (t474 * z3c - 0.3e1 * t16 * t473) * t113 + 0.3e1 * t474 * t44;
  const pc t486 = t48 + t49 - 0.3e1;
  const pc t503 = t486 * z3c;
  const pc t510 = -0.2e1 * t16 * t503 - 0.8e1 * t58 + t412 + 0.3e1 * t344 + t319 + t228 -
                     0.8e1 * t320 - t122 + 0.6e1 * t230 + t12 + t324 - t417 - t187 + t435 - t437 -
                     t188;
  const pc t522 = t198 - 0.3e1 + t244;
  const pc t523 = t522 * z3c;
  const pc t526 = -0.2e1 * t16 * t523 + t319 - t437 + t107 - t122 + t231 + t207 + t12 + t205 +
                     t345 - t206 - t232 - t87 + t306 - t321;
  const pc t529 = t440 + 0.9e1 + 0.8e1 * t350 + t255 - t442 - t459;
  const pc t536 = t350 - t49 - t244 + 0.3e1;
  const pc a3 =
      ((t79 - zetac * (t16 + m2 - 0.1e1 + t48)) * t113 +
       (zeta * t486 * z3c - 0.2e1 * t16 * t486) * t44 - t274 * t486 * z3) * t83 +
      (((-t87 - t55 - t58 + t190 + t107 + zetac - t56) * z3c -
        zetac * (-t56 + t175 - t63 - t59 - t108 + t176 + t306)) * t113 +
       t510 * t44 + (-t274 * t503 + 0.3e1 * t16 * t273 * t486) * z3 + zeta * t393 * t273) * t31 +
      (t526 * t113 + (-zeta * t529 * z3c + 0.3e1 * t16 * t529) * t44 + 0.3e1 * t274 * t536 * z3) * t9 +
      ((-t372 * t523 + 0.3e1 * t16 * t363 * t522) * t113 + 0.3e1 * t372 * t536 * t44) * z2 +
      zeta * t473 * t363 * t113;
  const pc t556 = -0.1e1 + m2 + m3;
  const pc t557 = t556 * zeta;
  const pc a2 =
      ((-t557 * z3c + 0.2e1 * t557 * zetac) * t113 + t557 * t386 * t44) * t83 +
      ((0.2e1 * t557 * t1 + t556 * (-t63 + t190 - t59 - t57 - t87 + zetac + t102)) * t113 +
       (t557 * t409 - 0.3e1 * t557 * zetac * t386) * t44 - 0.3e1 * t557 * t393 * z3) * t31 +
      ((t557 * t454 - 0.3e1 * t557 * zetac * t453) * t113 - 0.3e1 * t557 * t467 * t44) * t9 -
      0.3e1 * t557 * t473 * t113 * z2;
  const pc t600 = t556 * t556;
  const pc t603 = zeta * t600;
  const pc a1 = -zeta * t113 * t600 * t83 +
                   ((-t603 * z3c + 0.3e1 * t603 * zetac) * t113 + 0.3e1 * t603 * t273 * t44) * t31 +
                   0.3e1 * t603 * t363 * t113 * t9;
  const pc a0 = -zeta * t31 * t113 * t600 * t556;
  out[0] = a0.c(); out[1] = a1.c(); out[2] = a2.c(); out[3] = a3.c(); out[4] = a4.c(); out[5] = a5.c();
  out[6] = a6.c(); out[7] = a7.c(); out[8] = a8.c(); out[9] = a9.c(); out[10] = a10.c();
}

// The polynomial the reference solves at source position zeta
// (imgpoint.cc:172-176): static part plus position part, coefficient by
// coefficient.  `fixed` is image_poly_static's result for the lens.
TL_HD void image_poly(const Lens& L, const cx* fixed, cx zeta, cx* c) {
  cx z[11];
  image_poly_zeta(L, zeta, z);
  for (int k = 0; k < 11; k++) c[k] = cadd(fixed[k], z[k]);
}

}  // namespace tl
