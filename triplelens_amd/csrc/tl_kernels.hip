// HIP kernels of the IRS engine (gfx950).
//
//  trial_kernel   per light-curve point, all at once: source position, caustic
//                 trial points (src/LightCurves/lcirs.cc:131-165) and the
//                 image polynomial of every solve (tl_coeffs.h).
//  seed_kernel    per light curve, the reference's seed sequence: point images
//                 with the root-reuse chain, then the images of caustic trial
//                 points (src/LightCurves/lcirs.cc:116-165), as pixel seeds.
//  fill_kernel    one persistent, work-stealing launch that grows the image
//                 "amoebae" of a batch of source positions: wavefronts pop
//                 row-check tasks, evaluate 64 pixels per step with the FP64
//                 ray map, record maximal hit runs in a shared run table and
//                 sweep on to the next row.  Counted set and brightness sum are
//                 those of lineFloodFill (lcirs.cc:381-466): every 4-connected
//                 component of hit pixels that holds a hit seed.
//  resolve_kernel folds the per-label partial sums into per-point results.
//
// Compiled with -ffp-contract=off: the hit test rsq < rho^2 is decided on the
// same IEEE values as the reference's irsSIMD.
#include <hip/hip_runtime.h>
#include "tl_engine.h"

namespace tl {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

// Shared state of the fill (run table, labels, queue, counters) is accessed
// only through relaxed agent-scope atomics, which are performed at the
// device's coherence point.  Where one such access must be visible before
// another, mem_order() waits until all of the wave's outstanding memory
// operations have completed; acquire/release orderings are not used because
// on this GPU they write back or invalidate the whole L2 cache.
__device__ __forceinline__ unsigned long long ald(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ void ast(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ bool acas(unsigned long long* p, unsigned long long* expected,
                                     unsigned long long desired) {
  return __hip_atomic_compare_exchange_strong(p, expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ int aldi(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT); }
__device__ __forceinline__ bool acasi(int* p, int* expected, int desired) {
  return __hip_atomic_compare_exchange_strong(p, expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ long long afaddl(long long* p, long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ void mem_order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ long long wave_sum(long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T bcast0(T v) { return __shfl(v, 0, 64); }

// Wave-uniform values in scalar registers (the first active lane's value;
// used where every lane holds the same value or lane 0 holds the result).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni64(long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(unsigned long long)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double unid(double v) {
  return __longlong_as_double(uni64(__double_as_longlong(v)));
}

// ---------------------------------------------------------------------------
// Pixel <-> plane mapping (lcirs.cc:355-378)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nx_to_x(const CurveDev& C, long long nx) {
  return C.blx + C.step * (double)nx;
}
__device__ __forceinline__ double ny_to_y(const CurveDev& C, long long ny) {
  return C.bly + C.step * (double)ny;
}
// (long)((x - BL) * N / L + 0.5), with the operand clamped so that the
// conversion is defined for any root.
__device__ __forceinline__ int to_pixel(double v, double bl, double Nd, double L) {
  double t = (v - bl) * Nd / L + 0.5;
  if (!(t > -1.0e9)) t = -1.0e9;
  if (t > 1.0e9) t = 1.0e9;
  return (int)(long long)t;
}

// ---------------------------------------------------------------------------
// Seeds
// ---------------------------------------------------------------------------
__device__ double point_amp_term(const Lens& L, cx z) {
  cx d1 = csub(z, L.z1), d2 = csub(z, L.z2), d3 = csub(z, L.z3);
  cx s = cadd(cadd(cdiv_rc(L.m1, cmul(d1, d1)), cdiv_rc(L.m2, cmul(d2, d2))),
              cdiv_rc(L.m3, cmul(d3, d3)));
  double t = 1.0 / fabs(1.0 - cnorm(s));
  return t == t ? t : 0.0;
}

// ---- wave-cooperative seed search -----------------------------------------
// One wavefront per light curve.  The root-reuse chain is serial; its vectors
// are spread over the lanes (lane k holds coefficient k, or root k) and the
// serial steps read single elements with readlane, so nothing is indexed at
// run time and everything stays in registers.  The independent parts (the
// coefficients, the 10 root polishes, the 10 image tests and the
// caustic-segment tests) run on separate lanes.  Values and their order are
// those of the serial code in tl_math.h, operation for operation.

// Laguerre's fractional kicks (laguerre.cc:22), read modulo 8 as lag_iter does.
__constant__ double kLagFrac[8] = {0.5, 0.23, 0.73, 0.13, 0.37, 0.61, 0.89, 1.0};

__device__ __forceinline__ double rl(double v, int k) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, k);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ cx rlc(const cx& v, int k) { return mk(rl(v.re, k), rl(v.im, k)); }
// One part of the quotient n / d as cdiv_sel forms it (part 0: real, 1:
// imaginary), for this lane's numerator and part: lanes that need the
// quotients of several numerators over one denominator share one division
// sequence.
__device__ __forceinline__ double smith_part(cx n, cx d, int part) {
  const double a = n.re, b = n.im, c = d.re, e = d.im;
  const bool ce = fabs(c) < fabs(e);
  const double num = ce ? c : e, den = ce ? e : c;
  const double ratio = num / den;
  const double denom = (num * ratio) + den;
  const double ar = a * ratio, br = b * ratio;
  const double xn = ce ? ar + b : br + a;
  const double yn = ce ? br - a : b - ar;
  return (part ? yn : xn) / denom;
}

// Laguerre::laguerre (laguerre.cc:18-99) for an x held by every lane; the
// polynomial's coefficient j is lane j of pv, and M is its degree.  The
// coefficients are read once per call (the degree is a template argument,
// so every lane index is a constant and the loops unroll), not once per
// iteration.  The Horner recurrence runs on all lanes and lane j keeps b_j;
// independent scalar steps of the reference run on separate lanes in one
// instruction sequence: the moduli |b_j|, |p_m| and |x|, the four parts of
// f/b and 2s/b, the moduli of the two candidate denominators and the two
// parts of m/g.  Every value is the one the serial code computes.
template <int M>
__device__ __forceinline__ bool wl_lag(const cx& pv, cx* xp, const TrigTable* trig, bool& oob) {
  const int ln = __lane_id();
  cx x = *xp;
  const double md = (double)M;
  if constexpr (M == 1) {
    const cx p0 = rlc(pv, 0), p1 = rlc(pv, 1);
    if (cabs_sel(p1) != 0.0) {
      *xp = cdiv_sel(cneg(p0), p1);
      return true;
    }
    *xp = mk(999.0, 999.0);
    return false;
  } else {
    cx p[M + 1];
#pragma unroll
    for (int j = 0; j <= M; j++) p[j] = rlc(pv, j);
    for (int i = 1; i <= kLagMaxIt; i++) {
      const cx pm = p[M];
      cx b = pm;
      cx s = mk(0.0, 0.0), f = mk(0.0, 0.0);
      // lanes past M + 1 carry 1 + i, whose modulus (unused) takes the
      // branch-free path of cabs_sel: with 0 there they would send the wave
      // through the general routine as well
      cx mine = ln == M ? pm : (ln == M + 1 ? x : mk(1.0, 1.0));
#pragma unroll
      for (int j = M - 1; j >= 0; j--) {
        s = cadd(cmul(x, s), f);
        f = cadd(cmul(x, f), b);
        b = cadd(cmul(x, b), p[j]);
        if (ln == j) mine = b;
      }
      // lane j < M: |b_j|; lane M: |p_M| (the error bound's start); lane M+1: |x|
      const double mabs = cabs_sel(mine);
      const double ax = rl(mabs, M + 1);
      double err = rl(mabs, M);
#pragma unroll
      for (int j = M - 1; j >= 0; j--) err = rl(mabs, j) + ax * err;
      err *= kLagEps;
      if (rl(mabs, 0) <= err) {
        *xp = x;
        return true;
      }
      // g = f / b (lanes 0, 1) and 2s / b (lanes 2, 3)
      const double qv = smith_part(ln < 2 ? f : cscale(s, 2.0), b, ln & 1);
      const cx g = mk(rl(qv, 0), rl(qv, 1));
      const cx h = csub(cmul(g, g), mk(rl(qv, 2), rl(qv, 3)));
      const cx rad = cscale(csub(cscale(h, md), cmul(g, g)), md - 1.0);
      const cx sq = csqrt_sel(rad);
      cx gp = cadd(g, sq);
      const cx gm = csub(g, sq);
      // |gp| (lane 0) and |gm| (lane 1)
      const double av = cabs_sel(ln == 0 ? gp : gm);
      double agp = rl(av, 0);
      const double agm = rl(av, 1);
      if (agp < agm) {
        gp = gm;
        agp = agm;
      }
      cx dx;
      if (agp > 0.0) {
        const double dv = smith_part(mk(md, 0.0), gp, ln & 1);
        dx = mk(rl(dv, 0), rl(dv, 1));
      } else {
        dx = cscale(mk(trig->c[i], trig->s[i]), 1.0 + ax);
      }
      cx x1 = csub(x, dx);
      if (x.re == x1.re && x.im == x1.im) {
        *xp = x;
        return true;
      }
      if (i % kLagMT) {
        x = x1;
      } else {
        oob |= i / kLagMT >= 8;
        x = csub(x, cmul(mk(kLagFrac[(i / kLagMT) & 7], 0.0), dx));
      }
    }
    *xp = x;
    return false;
  }
}

// Laguerre::laguerre with a different x on every lane (the polishing step):
// the coefficients are the lanes of pv (degree M), read once per call and
// used in the same order by all lanes.
template <int M>
__device__ __forceinline__ bool wl_lag_lanes(const cx& pv, cx* xp, const TrigTable* trig,
                                             bool& oob) {
  cx x = *xp;
  const double md = (double)M;
  cx p[M + 1];
#pragma unroll
  for (int j = 0; j <= M; j++) p[j] = rlc(pv, j);
  for (int i = 1; i <= kLagMaxIt; i++) {
    cx b = p[M];
    double err = cabs_sel(b);
    const double ax = cabs_sel(x);
    cx s = mk(0.0, 0.0), f = mk(0.0, 0.0);
#pragma unroll
    for (int j = M - 1; j >= 0; j--) {
      s = cadd(cmul(x, s), f);
      f = cadd(cmul(x, f), b);
      b = cadd(cmul(x, b), p[j]);
      err = cabs_sel(b) + ax * err;
    }
    err *= kLagEps;
    if (cabs_sel(b) <= err) {
      *xp = x;
      return true;
    }
    cx g = cdiv_sel(f, b);
    cx h = csub(cmul(g, g), cdiv_sel(cscale(s, 2.0), b));
    cx rad = cscale(csub(cscale(h, md), cmul(g, g)), md - 1.0);
    cx sq = csqrt_sel(rad);
    cx gp = cadd(g, sq);
    cx gm = csub(g, sq);
    // the moduli are taken once: |gp| after the choice is the chosen one's,
    // and |x| is ax
    double agp = cabs_sel(gp);
    const double agm = cabs_sel(gm);
    if (agp < agm) {
      gp = gm;
      agp = agm;
    }
    cx dx;
    if (agp > 0.0) dx = cdiv_rc_sel(md, gp);
    else dx = cscale(mk(trig->c[i], trig->s[i]), 1.0 + ax);
    cx x1 = csub(x, dx);
    if (x.re == x1.re && x.im == x1.im) {
      *xp = x;
      return true;
    }
    if (i % kLagMT) {
      x = x1;
    } else {
      oob |= i / kLagMT >= 8;
      x = csub(x, cmul(mk(kLagFrac[(i / kLagMT) & 7], 0.0), dx));
    }
  }
  *xp = x;
  return false;
}

// The deflation steps J..1 of Laguerre::solveRoots (laguerre.cc:103-167):
// every lane holds the step's root, the deflated coefficients stay on the
// lanes of tmp, and the root of step J goes to lane 10 - J of rootv.
template <int J>
__device__ __forceinline__ void wl_deflate(cx& tmp, cx& rootv, bool init, const cx& initv,
                                           const TrigTable* trig, bool& oob) {
  if constexpr (J >= 1) {
    const int ln = __lane_id();
    cx x = init ? rlc(initv, J) : mk(kLagEps, kLagEps);
    if (!wl_lag<J>(tmp, &x, trig, oob) && init) x = mk(kLagEps, kLagEps);
    if (ln == 10 - J) rootv = x;
    cx b = rlc(tmp, J);
#pragma unroll
    for (int jj = J - 1; jj >= 0; jj--) {
      const cx c = rlc(tmp, jj);
      if (ln == jj) tmp = b;
      b = cadd(cmul(x, b), c);
    }
    wl_deflate<J - 1>(tmp, rootv, init, initv, trig, oob);
  }
}

// Laguerre::solveRoots for the degree-10 image polynomial (coefficients on
// lanes 0..10 of cv): the deflation, then lane r polishes root r on the
// full polynomial.  Returns root r on lane r (r < 10; 0 elsewhere).  With
// init, initv's lane j seeds the search of step j (j = 10..1).
__device__ __forceinline__ cx wl_solve10(const cx& cv, bool init, const cx& initv,
                                         const TrigTable* trig, bool& oob) {
  const int ln = __lane_id();
  cx tmp = cv;
  cx rootv = mk(0.0, 0.0);
  wl_deflate<10>(tmp, rootv, init, initv, trig, oob);
  bool laneOob = false;
  if (ln < 10) wl_lag_lanes<10>(cv, &rootv, trig, laneOob);
  oob |= __ballot(laneOob) != 0;
  return rootv;
}

// Laguerre::checkRoots (laguerre.cc:188-198) on roots held one per lane.
__device__ __forceinline__ bool wl_check(const cx& cv, const cx& rootv) {
  cx c0 = rlc(cv, 10);
  for (int j = 0; j < 10; j++) c0 = cmul(c0, cneg(rlc(rootv, j)));
  const cx a0 = rlc(cv, 0);
  if (cabs_sel(a0) > 1.0e-8) return (1.0 - cabs_sel(cdiv_sel(c0, a0))) < 1.0e-5;
  return cabs_sel(csub(c0, a0)) < 1.0e-10;
}

// getRootsPrecalculated + getImages (imgpoint.cc:160-274) on a wavefront,
// for the polynomial whose coefficient k is lane k of cv.  chainv holds
// tempRoots[k] on lane k (lane 10 is the slot the reference reads past the
// end of its vector: 0 here, or a recorded value the caller put there).
// Returns the mask of the roots accepted as images (bit k: root k on lane k
// of chainv).
__device__ __forceinline__ unsigned long long wl_get_images(const Lens& L, const cx& cv,
                                                            cx& chainv, int& have, cx zeta,
                                                            const TrigTable* trig, bool binImg,
                                                            bool& oob) {
  const int ln = __lane_id();
  // The reference's sequence (imgpoint.cc:181-206,262-270): a solve from the
  // chain's roots (when it has some), replaced by a fresh solve when the
  // roots fail checkRoots; then, if the image count is invalid, one more
  // fresh solve.  A fresh solve of the same coefficients gives the same
  // roots, so the second one is skipped when the first solve was fresh.
  bool withInit = have != 0;
  bool second = false;
  while (true) {
    const cx r = wl_solve10(cv, withInit, chainv, trig, oob);
    if (withInit && !wl_check(cv, r)) {
      withInit = false;
      continue;
    }
    chainv = ln < 10 ? r : mk(0.0, 0.0);
    have = 1;
    const bool ok = ln < 10 && img_check(L, chainv, zeta, 1.0e-6);
    const unsigned long long m = __ballot(ok);
    const int n = __builtin_popcountll(m);
    // a valid count: even and at least 4, or for the binary lens odd and at
    // least 3 (imgpoint.cc:236-237)
    const bool bad = binImg ? (n % 2 != 1 || n < 3) : (n % 2 != 0 || n < 4);
    if (!bad || second || !withInit) return m;
    second = true;
    withInit = false;
  }
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Seed pixels of the images in `m` (root order), appended at out[cnt..].
__device__ __forceinline__ void emit_seeds(const CurveDev& C, const cx& chainv,
                                           unsigned long long m, int2* out, int& cnt,
                                           uint32_t& fl) {
  const int ln = __lane_id();
  const int n = __builtin_popcountll(m);
  if ((m >> ln) & 1ull) {
    const int slot = cnt + __builtin_popcountll(m & ((1ull << ln) - 1ull));
    if (slot < kMaxSeeds)
      out[slot] = make_int2(to_pixel(chainv.re, C.blx, C.Nd, C.Ldbl),
                            to_pixel(chainv.im, C.bly, C.Nd, C.Ldbl));
  }
  if (cnt + n > kMaxSeeds) fl |= kFlagSeedOverflow;
  cnt = cnt + n > kMaxSeeds ? kMaxSeeds : cnt + n;
}

// The caustic trial points of one source position in the reference's order
// (lcirs.cc:131-165): for each caustic branch r whose cumulative box meets
// the source square, the segments k = 0..kCccLen-2 are tested 64 at a time;
// the first hit of every run of consecutive hits is a trial point (the
// reference solves only those; `pointTaken` restarts at each branch).  For
// each step that has trial points, onStep(f, trial) receives their mask f
// (bit j: the segment on lane j, in segment order) and every lane's trial
// position.
template <typename OnStep>
__device__ __forceinline__ void scan_trials(const CurveDev& C, cx pos, const cx* caus,
                                            OnStep&& onStep) {
  const int ln = __lane_id();
  for (int r = 0; r < 6; r++) {
    if (!(pos.re + C.rho >= C.boxMin[r].re && pos.re - C.rho <= C.boxMax[r].re &&
          pos.im + C.rho >= C.boxMin[r].im && pos.im - C.rho <= C.boxMax[r].im))
      continue;
    const cx* cr = caus + r * kCccLen;
    bool prevHit = false;
    for (int k0 = 0; k0 < kCccLen - 1; k0 += 64) {
      const int k = k0 + ln;
      cx trial = mk(0.0, 0.0);
      const bool hit = k < kCccLen - 1 &&
                       intersection_check(pos, cr[k], cr[k + 1], C.rho, C.rho2pow, &trial);
      const unsigned long long hm = __ballot(hit);
      const unsigned long long f = hm & ~((hm << 1) | (prevHit ? 1ull : 0ull));
      prevHit = (hm >> 63) & 1ull;
      if (f) onStep(f, trial);
    }
  }
}

// Everything of a light-curve point that does not depend on the root chain,
// one wavefront per point, all points at once: its source position, the
// distance to the nearest caustic sample, and the list of positions the
// chain will solve for (the centre, then the caustic trial points).  The
// list goes to the pool at a place taken with one atomic add; when the pool
// is too small, the cursor still advances and nothing is written, and the
// host runs the kernel again with a bigger pool.  The solves' polynomials
// are computed afterwards by trial_poly_kernel, one lane per solve (that
// evaluation needs many registers; kept apart, this kernel stays small and
// many of its wavefronts share each SIMD).
__global__ void __launch_bounds__(64) trial_kernel(TrialArgs a) {
  const int p = blockIdx.x;
  if (p >= a.nPts) return;
  const int ln = __lane_id();
  const CurveDev& C = a.curves[a.ptCurve[p]];
  const unsigned i = (unsigned)(p - C.pointBase);
  const cx pos = traj_pos(mk(C.sx, C.sy), mk(C.ex, C.ey), i, C.lcLength);
  const cx* caus = a.caustics + C.causticBase;
  // distance to the nearest of the curve's caustic samples
  double d2 = 1e300;
  for (int q = ln; q < 6 * kCccLen; q += 64) {
    const double v = cnorm(csub(caus[q], pos));
    d2 = v < d2 ? v : d2;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const double w2 = __shfl_xor(d2, o, 64);
    d2 = w2 < d2 ? w2 : d2;
  }
  if (ln == 0) {
    a.ptPos[p] = make_double2(pos.re, pos.im);
    a.ptCausD2[p] = d2;
  }
  int n = 1;
  scan_trials(C, pos, caus, [&](unsigned long long f, const cx&) { n += __builtin_popcountll(f); });
  int base = 0;
  if (ln == 0) base = (int)atomicAdd(a.pool.cursor, (unsigned long long)n);
  base = __shfl(base, 0, 64);
  if (ln == 0) {
    a.pool.base[p] = base;
    a.pool.count[p] = n;
  }
  if ((long long)base + n > a.pool.cap) return;
  if (ln == 0) {
    a.pool.zeta[base] = pos;
    a.pool.owner[base] = p;
  }
  int next = base + 1;
  scan_trials(C, pos, caus, [&](unsigned long long f, const cx& trial) {
    if ((f >> ln) & 1ull) {
      const int sid = next + __builtin_popcountll(f & ((1ull << ln) - 1ull));
      a.pool.zeta[sid] = trial;
      a.pool.owner[sid] = p;
    }
    next += __builtin_popcountll(f);
  });
}

// The image polynomial of every solve in the pool (tl_coeffs.h: the
// reference's rounding), one lane per solve.
__global__ void __launch_bounds__(64) trial_poly_kernel(TrialArgs a, int nSolves) {
  const int sid = blockIdx.x * blockDim.x + threadIdx.x;
  if (sid >= nSolves) return;
  const CurveDev& C = a.curves[a.ptCurve[a.pool.owner[sid]]];
  cx c[11];
  image_poly(C.L, C.fixedPoly, a.pool.zeta[sid], c);
  for (int k = 0; k < 11; k++) a.pool.coef[(size_t)sid * 11 + k] = c[k];
}

int launch_trials(const TrialArgs& a, hipStream_t s) {
  if (a.nPts <= 0) return 0;
  hipLaunchKernelGGL(trial_kernel, dim3(a.nPts), dim3(64), 0, s, a);
  return (int)hipGetLastError();
}

int launch_trial_polys(const TrialArgs& a, int nSolves, hipStream_t s) {
  if (nSolves <= 0) return 0;
  hipLaunchKernelGGL(trial_poly_kernel, dim3((nSolves + 63) / 64), dim3(64), 0, s, a, nSolves);
  return (int)hipGetLastError();
}

// image_poly for n positions of one lens, one lane each (test access to the
// device evaluation of the polynomial).
__global__ void __launch_bounds__(64) image_poly_kernel(Lens L, const cx* fixedPoly,
                                                        const cx* zeta, int n, cx* coef) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  cx c[11];
  image_poly(L, fixedPoly, zeta[g], c);
  for (int k = 0; k < 11; k++) coef[(size_t)g * 11 + k] = c[k];
}

int launch_image_poly(const Lens& L, const cx* fixedPoly, const cx* zeta, int n, cx* coef,
                      hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(image_poly_kernel, dim3((n + 63) / 64), dim3(64), 0, s, L, fixedPoly, zeta,
                     n, coef);
  return (int)hipGetLastError();
}

// Workgroup (curve c, piece q): a curve with split = 1 walks the chunk's
// positions as one chain, continuing the curve's root-reuse state.  With
// split = S > 1 the chunk's positions are cut into S contiguous pieces that
// run in parallel; piece 0 continues the chain, the others start from a
// fresh solve.  Each position's solves (centre, then caustic trial points)
// and their polynomials come from trial_kernel's pool, so the chain itself
// only runs the root solves, the image tests and the seed output.
__global__ void __launch_bounds__(64) seed_kernel(SeedArgs a) {
  const int c = blockIdx.x / a.maxSplit, q = blockIdx.x % a.maxSplit;
  if (c >= a.nCurves) return;
  const int ln = __lane_id();
  const CurveDev& C = a.curves[c];
  if (q >= C.split) return;
  const Lens L = C.L;
  cx chainv = mk(0.0, 0.0);
  int have = 0, calls = 0;
  if (q > 0) {
    // a known start state (the verified speculative chain), else a fresh
    // solve at the piece's first position
    if (a.pieceStart) {
      if (ln < 10) chainv = a.pieceStart[((size_t)a.pieceBase[c] + q) * 10 + ln];
      have = 1;
    }
  } else if (a.chainIn) {
    if (ln <= 10) chainv = a.chainIn[c].tempRoots[ln <= 10 ? ln : 0];
    have = a.chainIn[c].have;
    calls = a.chainIn[c].calls;
  } else {
    if (ln <= 10) chainv = C.chain[ln <= 10 ? ln : 0];
    have = C.chainHave;
    calls = C.chainCalls;
  }
  const bool traced = q == 0 && a.slotTrace && C.slotCount > 0;
  const long long n = (long long)C.lcLength + 1;
  const long long cBegin = n * a.chunk / a.nChunks, cEnd = n * (a.chunk + 1) / a.nChunks;
  const unsigned iBegin = (unsigned)(cBegin + (cEnd - cBegin) * q / C.split);
  const unsigned iEnd = (unsigned)(cBegin + (cEnd - cBegin) * (q + 1) / C.split);
  for (unsigned i = iBegin; i < iEnd; i++) {
    const int p = C.pointBase + (int)i;
    const int base = a.pool.base[p], nSolve = a.pool.count[p];
    int cnt = 0;
    uint32_t fl = 0;
    int2* out = a.seeds + (size_t)p * kMaxSeeds;
    double amp = 0.0;
    for (int sIdx = 0; sIdx < nSolve; sIdx++) {
      const size_t sid = (size_t)base + sIdx;
      const cx cv = ln <= 10 ? a.pool.coef[sid * 11 + (ln <= 10 ? ln : 0)] : mk(0.0, 0.0);
      const cx zeta = a.pool.zeta[sid];
      if (traced && have && calls < C.slotCount && ln == 10)
        chainv = a.slotTrace[C.slotBase + calls];
      calls++;
      bool oob = false;
      const unsigned long long m = wl_get_images(L, cv, chainv, have, zeta, a.trig, C.binImg, oob);
      if (oob) fl |= kFlagKickTable;
      emit_seeds(C, chainv, m, out, cnt, fl);
      if (sIdx == 0) {
        // point-source magnification (getPointAmp, lcbase.cc:17-38) and the
        // magnification-weighted image centroid, summed in root order
        const double term = ((m >> ln) & 1ull) ? point_amp_term(L, chainv) : 0.0;
        const double wx = term * chainv.re, wy = term * chainv.im;
        double sx = 0.0, sy = 0.0;
        unsigned long long mm = m;
        while (mm) {
          const int k = __builtin_ctzll(mm);
          mm &= mm - 1;
          amp += rl(term, k);
          sx += rl(wx, k);
          sy += rl(wy, k);
        }
        // the two largest image magnifications
        double t1 = term;
        for (int o = 32; o >= 1; o >>= 1) {
          const double w2 = __shfl_xor(t1, o, 64);
          t1 = w2 > t1 ? w2 : t1;
        }
        const unsigned long long top = __ballot(term == t1 && ((m >> ln) & 1ull));
        const int topLane = top ? __builtin_ctzll(top) : -1;
        double t2 = (ln == topLane) ? 0.0 : term;
        for (int o = 32; o >= 1; o >>= 1) {
          const double w2 = __shfl_xor(t2, o, 64);
          t2 = w2 > t2 ? w2 : t2;
        }
        if (ln < 10) a.ptRoots[(size_t)p * 10 + ln] = chainv;
        if (ln == 0) {
          a.ptAmp[p] = amp;
          a.ptCen[p] = make_double2(sx / amp, sy / amp);
          const double dist = sqrt(a.ptCausD2[p]);
          // inside a caustic: more images than the 4 (binary: 3) outside
          a.ptCaus[p] = __builtin_popcountll(m) > (C.binImg ? 3 : 4) ? -dist : dist;
          a.ptPair[p] = amp > 0.0 ? (t1 + t2) / amp : 0.0;
        }
      }
    }
    if (ln == 0) {
      a.nSeeds[p] = cnt;
      a.flags[p] = fl;
      a.cost[p] = (float)(amp * 3.141592653589793 * C.ppr * C.ppr);
    }
  }
  if (q == C.split - 1) {
    if (ln <= 10) a.chainOut[c].tempRoots[ln] = chainv;
    if (ln == 0) {
      a.chainOut[c].have = have;
      a.chainOut[c].calls = calls;
    }
  }
}

// ---- verified speculative chain (SpecArgs) ---------------------------------
// The chain's solves of position p (centre, then caustic trial points), as
// seed_kernel runs them, without output.
__device__ __forceinline__ void chain_step(const SeedArgs& a, const CurveDev& C, const Lens& L,
                                           int p, cx& chainv, int& have) {
  const int ln = __lane_id();
  const int base = a.pool.base[p], nSolve = a.pool.count[p];
  for (int sIdx = 0; sIdx < nSolve; sIdx++) {
    const size_t sid = (size_t)base + sIdx;
    const cx cv = ln <= 10 ? a.pool.coef[sid * 11 + (ln <= 10 ? ln : 0)] : mk(0.0, 0.0);
    bool oob = false;
    wl_get_images(L, cv, chainv, have, a.pool.zeta[sid], a.trig, C.binImg, oob);
  }
}

// The chain state (roots 0..9 on lanes 0..9) equals the stored one, bit for bit.
__device__ __forceinline__ bool state_equal(const cx& chainv, const cx* st) {
  const int ln = __lane_id();
  bool eq = true;
  if (ln < 10) {
    const cx v = st[ln];
    eq = __double_as_longlong(v.re) == __double_as_longlong(chainv.re) &&
         __double_as_longlong(v.im) == __double_as_longlong(chainv.im);
  }
  return __ballot(!eq) == 0;
}

__device__ __forceinline__ void piece_range(const CurveDev& C, int q, int& b, int& e) {
  const long long n = (long long)C.lcLength + 1;
  b = (int)(n * q / C.split);
  e = (int)(n * (q + 1) / C.split);
}

// Round 1: piece q from a fresh solve (piece 0 from the curve's chain state),
// on into the next piece for sp.overlap positions.
__global__ void __launch_bounds__(64) spec_round1_kernel(SeedArgs a, SpecArgs sp) {
  const int c = blockIdx.x / a.maxSplit, q = blockIdx.x % a.maxSplit;
  if (c >= a.nCurves) return;
  const int ln = __lane_id();
  const CurveDev& C = a.curves[c];
  if (q >= C.split) return;
  const Lens L = C.L;
  cx chainv = mk(0.0, 0.0);
  int have = 0;
  if (q == 0) {
    if (ln <= 10) chainv = C.chain[ln <= 10 ? ln : 0];
    have = C.chainHave;
  }
  int b, e;
  piece_range(C, q, b, e);
  int stop = e;
  if (q + 1 < C.split) {
    int b2, e2;
    piece_range(C, q + 1, b2, e2);
    stop = e + sp.overlap < e2 ? e + sp.overlap : e2;
  }
  for (int i = b; i < stop; i++) {
    const int p = C.pointBase + i;
    chain_step(a, C, L, p, chainv, have);
    cx* dst = (i < e ? sp.own : sp.ext) + (size_t)p * 10;
    if (ln < 10) dst[ln] = chainv;
  }
}

// Round 2: for a piece q > 0 that does not agree with its predecessor's
// continuation, the repair chain from the predecessor's end state, through
// the following pieces until its state equals a piece's own at one of that
// piece's first sp.overlap positions.
__global__ void __launch_bounds__(64) spec_repair_kernel(SeedArgs a, SpecArgs sp) {
  const int c = blockIdx.x / a.maxSplit, q = blockIdx.x % a.maxSplit;
  if (c >= a.nCurves) return;
  const int ln = __lane_id();
  const CurveDev& C = a.curves[c];
  if (q == 0 || q >= C.split) return;
  const int gq = sp.pieceBase[c] + q;
  int b, e;
  piece_range(C, q, b, e);
  const int lim = e < b + sp.overlap ? e : b + sp.overlap;
  bool agree = false;
  for (int i = b; i < lim && !agree; i++) {
    const size_t p = (size_t)C.pointBase + i;
    cx v = mk(0.0, 0.0);
    if (ln < 10) v = sp.ext[p * 10 + ln];
    agree = state_equal(v, sp.own + p * 10);
  }
  if (ln == 0) sp.merged[gq] = agree ? 1 : 0;
  if (agree) return;
  const Lens L = C.L;
  cx chainv = mk(0.0, 0.0);
  if (ln < 10) chainv = sp.own[((size_t)C.pointBase + b - 1) * 10 + ln];
  int have = 1;
  int qq = q, qb = b, qe = e;  // the piece the chain is in
  const long long n = (long long)C.lcLength + 1;
  for (int i = b; i < n; i++) {
    if (i == qe) {
      // entering piece qq + 1: the chain's state before its first position
      qq++;
      qb = qe;
      piece_range(C, qq, qb, qe);
      const int j = qq - q - 1;
      if (j >= sp.depth) {
        if (ln == 0) sp.repTarget[gq] = -2;
        return;
      }
      if (ln < 10) sp.repState[((size_t)gq * sp.depth + j) * 10 + ln] = chainv;
    }
    const int p = C.pointBase + i;
    chain_step(a, C, L, p, chainv, have);
    if (qq > q && i < qb + sp.overlap && state_equal(chainv, sp.own + (size_t)p * 10)) {
      if (ln == 0) sp.repTarget[gq] = qq;
      return;
    }
  }
  if (ln == 0) sp.repTarget[gq] = -1;
}

// The true chain's state before each piece, walking every curve's pieces in
// order (one lane per curve): the chain known to be the true one at each
// piece start is either the previous piece's own (it agreed with the chain
// before it, or it is piece 0) or a repair chain that has not yet agreed.
__global__ void spec_resolve_kernel(SeedArgs a, SpecArgs sp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nCurves) return;
  const CurveDev& C = a.curves[c];
  const int g0 = sp.pieceBase[c];
  bool viaPiece = true;  // the true chain at the coming start is piece q - 1's own
  int rq = 0;            // else the repair chain of piece rq
  for (int q = 1; q < C.split; q++) {
    int b, e;
    piece_range(C, q, b, e);
    const cx* src;
    if (viaPiece) {
      src = sp.own + ((size_t)C.pointBase + b - 1) * 10;
      if (!sp.merged[g0 + q]) {
        viaPiece = false;
        rq = q;
      }
    } else {
      const int j = q - rq - 1;
      const int tgt = sp.repTarget[g0 + rq];
      if (j >= sp.depth || (tgt >= 0 && tgt < q)) {  // cannot happen: the chain stopped earlier
        atomicExch(sp.failed, 1);
        return;
      }
      src = sp.repState + ((size_t)(g0 + rq) * sp.depth + j) * 10;
      if (tgt == q) viaPiece = true;
    }
    for (int k = 0; k < 10; k++) sp.start[((size_t)g0 + q) * 10 + k] = src[k];
  }
}

int launch_spec(const SeedArgs& a, const SpecArgs& sp, hipStream_t s) {
  if (a.nCurves <= 0) return 0;
  if (a.maxSplit < 1 || (long long)a.nCurves * a.maxSplit > (1ll << 30))
    return (int)hipErrorInvalidValue;
  if (!sp.pieceBase || !sp.own || !sp.ext || !sp.merged || !sp.repState || !sp.repTarget ||
      !sp.start || !sp.failed || sp.overlap < 1 || sp.depth < 1 || !a.curves || !a.trig ||
      !a.pool.base || !a.pool.count || !a.pool.coef || !a.pool.zeta)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(spec_round1_kernel, dim3(a.nCurves * a.maxSplit), dim3(64), 0, s, a, sp);
  hipLaunchKernelGGL(spec_repair_kernel, dim3(a.nCurves * a.maxSplit), dim3(64), 0, s, a, sp);
  hipLaunchKernelGGL(spec_resolve_kernel, dim3((a.nCurves + 63) / 64), dim3(64), 0, s, a, sp);
  return (int)hipGetLastError();
}

// LightCurveBase::getLC / getPointAmp (lcbase.cc:17-52): positions at
// t = i/(lcLength-1), magnification sum over images (in root order) of
// 1/|1 - |sum m/(z-z_k)^2|^2| (the squares as products; the reference uses
// std::pow(z, 2.0)).  One wavefront per curve, as the seed search.
// The image polynomial at every position of the point-source curves (one
// lane per position), for point_lc_kernel: coef[(pointBase + i) * 11 + k].
__global__ void __launch_bounds__(64) point_poly_kernel(const CurveDev* curves, int nCurves,
                                                        int nTotal, cx* coef) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nTotal) return;
  int c = 0;
  while (c + 1 < nCurves && curves[c + 1].pointBase <= g) c++;
  const CurveDev& C = curves[c];
  const int i = g - C.pointBase;
  if (i < 0 || i >= (int)C.lcLength) return;
  const cx s = mk(C.sx, C.sy), e = mk(C.ex, C.ey);
  const cx pos = cadd(s, cscale(csub(e, s), i / (C.lcLength - 1.0)));
  cx cf[11];
  image_poly(C.L, C.fixedPoly, pos, cf);
  for (int k = 0; k < 11; k++) coef[(size_t)g * 11 + k] = cf[k];
}

__global__ void __launch_bounds__(64) point_lc_kernel(const CurveDev* curves, int nCurves,
                                                      const TrigTable* trig, double* out,
                                                      RootChain* chainOut, const cx* coef) {
  const int c = blockIdx.x;
  if (c >= nCurves) return;
  const int ln = __lane_id();
  const CurveDev& C = curves[c];
  const Lens L = C.L;
  cx chainv = ln <= 10 ? C.chain[ln <= 10 ? ln : 0] : mk(0.0, 0.0);
  int have = C.chainHave;
  const cx s = mk(C.sx, C.sy), e = mk(C.ex, C.ey);
  for (unsigned i = 0; i < C.lcLength; i++) {
    const cx pos = cadd(s, cscale(csub(e, s), i / (C.lcLength - 1.0)));
    const size_t g = (size_t)C.pointBase + i;
    const cx cv = ln <= 10 ? coef[g * 11 + (ln <= 10 ? ln : 0)] : mk(0.0, 0.0);
    bool oob = false;
    unsigned long long m = wl_get_images(L, cv, chainv, have, pos, trig, C.binImg, oob);
    const double term = ((m >> ln) & 1ull) ? point_amp_term(L, chainv) : 0.0;
    double amp = 0.0;
    while (m) {
      const int k = __builtin_ctzll(m);
      m &= m - 1;
      amp += rl(term, k);
    }
    if (ln == 0) out[C.pointBase + (int)i] = amp;
  }
  if (ln <= 10) chainOut[c].tempRoots[ln] = chainv;
  if (ln == 0) {
    chainOut[c].have = have;
    chainOut[c].calls = C.chainCalls + (int)C.lcLength;
  }
}

int launch_point_lc(const CurveDev* curves, int nCurves, int nTotal, const TrigTable* trig,
                    double* out, RootChain* chainOut, cx* coef, hipStream_t s) {
  if (nCurves <= 0 || nTotal <= 0) return 0;
  hipLaunchKernelGGL(point_poly_kernel, dim3((nTotal + 63) / 64), dim3(64), 0, s, curves, nCurves,
                     nTotal, coef);
  hipLaunchKernelGGL(point_lc_kernel, dim3(nCurves), dim3(64), 0, s, curves, nCurves, trig, out,
                     chainOut, coef);
  return (int)hipGetLastError();
}

// Offsets, in source radii, of the disk samples of Gould's (2008, ApJ 681,
// 1593) hexadecapole approximation: radius 1/2 on the axes, radius 1 on the
// axes, radius 1 on the diagonals.
constexpr double kSignificant = 1.0e-4;

__constant__ double kRingOff[kRingSamples][2] = {
    {0.5, 0.0}, {0.0, 0.5}, {-0.5, 0.0}, {0.0, -0.5},
    {1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0},
    {0.70710678118654752, 0.70710678118654752}, {-0.70710678118654752, 0.70710678118654752},
    {-0.70710678118654752, -0.70710678118654752}, {0.70710678118654752, -0.70710678118654752}};

// Point magnification (getPointAmp's sum, lcbase.cc:17-38) at the disk
// samples around each position, one wavefront per position.  Every sample
// is solved from the position's own centre roots (the reference's
// root-reuse solve, imgpoint.cc:160-274), so the samples are independent of
// each other and of the order they are visited in.  A sample whose count of
// significant images (each carrying at least kSignificant of the point's
// magnification) differs from the centre's lies across a caustic from it;
// the threshold keeps out the faint images next to a lens whose lens-equation
// residual sits at the acceptance threshold and that come and go between
// nearby points, while an image pair born on a fold is always bright.
// The image polynomial at every disk sample (one lane per sample) for
// ring_lane_kernel: coef[((p * kRingSamples) + s) * 11 + k].
__global__ void __launch_bounds__(64) ring_poly_kernel(const CurveDev* curves, const int* ptCurve,
                                                       const double2* ptPos, int nPts, cx* coef,
                                                       int* cross) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= (long long)nPts * kRingSamples) return;
  const int p = (int)(g / kRingSamples), smp = (int)(g % kRingSamples);
  const CurveDev& C = curves[ptCurve[p]];
  if (!C.ring) return;
  const double2 pp = ptPos[p];
  const cx zeta = mk(pp.x + C.rho * kRingOff[smp][0], pp.y + C.rho * kRingOff[smp][1]);
  cx cf[11];
  image_poly(C.L, C.fixedPoly, zeta, cf);
  for (int k = 0; k < 11; k++) coef[(size_t)g * 11 + k] = cf[k];
  if (smp == 0) cross[p] = 0;
}

// The serial Laguerre routines of tl_math.h for one lane with every array in
// registers: the degree is a template argument, so no array is indexed at
// run time (tl_math.h's run-time degree puts them in scratch memory, which
// made the disk samples several times slower).  The same operations in the
// same order, so the same values.
// lag_iter (laguerre.cc:18-99) for a polynomial of degree M.
template <int M>
__device__ __forceinline__ bool lag_iter_reg(const cx* poly, cx& xp, const TrigTable* trig) {
  if constexpr (M == 1) {
    if (cabs_(poly[1]) != 0.0) {
      xp = cdiv(cneg(poly[0]), poly[1]);
      return true;
    }
    xp = mk(999.0, 999.0);
    return false;
  } else {
    const double frac[8] = {0.5, 0.23, 0.73, 0.13, 0.37, 0.61, 0.89, 1.0};
    cx x = xp;
    const double md = (double)M;
    for (int i = 1; i <= kLagMaxIt; i++) {
      cx b = poly[M];
      double err = cabs_(b);
      const double ax = cabs_(x);
      cx s = mk(0.0, 0.0), f = mk(0.0, 0.0);
#pragma unroll
      for (int j = M - 1; j >= 0; j--) {
        s = cadd(cmul(x, s), f);
        f = cadd(cmul(x, f), b);
        b = cadd(cmul(x, b), poly[j]);
        err = cabs_(b) + ax * err;
      }
      err *= kLagEps;
      if (cabs_(b) <= err) {
        xp = x;
        return true;
      }
      cx g = cdiv(f, b);
      cx h = csub(cmul(g, g), cdiv(cscale(s, 2.0), b));
      cx rad = cscale(csub(cscale(h, md), cmul(g, g)), md - 1.0);
      cx sq = csqrt_(rad);
      cx gp = cadd(g, sq);
      cx gm = csub(g, sq);
      // each modulus taken once (|x| is ax)
      double agp = cabs_(gp);
      const double agm = cabs_(gm);
      if (agp < agm) {
        gp = gm;
        agp = agm;
      }
      cx dx;
      if (agp > 0.0) dx = cdiv_rc(md, gp);
      else dx = cscale(mk(trig->c[i], trig->s[i]), 1.0 + ax);
      cx x1 = csub(x, dx);
      if (x.re == x1.re && x.im == x1.im) {
        xp = x;
        return true;
      }
      if (i % kLagMT) x = x1;
      else x = csub(x, cmul(mk(frac[(i / kLagMT) & 7], 0.0), dx));
    }
    xp = x;
    return false;
  }
}

// lag_solve's deflation steps J..1 (laguerre.cc:103-167): the root found at
// step J goes to roots[DEG - J].
template <int DEG, int J>
__device__ __forceinline__ void lag_deflate_reg(cx* tmp, const cx* init, bool withInit, cx* roots,
                                                const TrigTable* trig) {
  if constexpr (J >= 1) {
    cx x = withInit ? init[J] : mk(kLagEps, kLagEps);
    if (!lag_iter_reg<J>(tmp, x, trig) && withInit) x = mk(kLagEps, kLagEps);
    roots[DEG - J] = x;
    cx b = tmp[J];
#pragma unroll
    for (int jj = J - 1; jj >= 0; jj--) {
      cx c = tmp[jj];
      tmp[jj] = b;
      b = cadd(cmul(x, b), c);
    }
    lag_deflate_reg<DEG, J - 1>(tmp, init, withInit, roots, trig);
  }
}

// lag_solve: the deflation, then each root polished on the full polynomial
// in root order (the roots rotate through roots[0], so one copy of the
// polishing code serves all of them).
template <int DEG>
__device__ __forceinline__ void lag_solve_reg(const cx* coeffs, const cx* init, bool withInit,
                                              cx* roots, const TrigTable* trig) {
  cx tmp[DEG + 1];
#pragma unroll
  for (int i = 0; i <= DEG; i++) tmp[i] = coeffs[i];
  lag_deflate_reg<DEG, DEG>(tmp, init, withInit, roots, trig);
#pragma unroll 1
  for (int j = 0; j < DEG; j++) {
    cx x = roots[0];
    lag_iter_reg<DEG>(coeffs, x, trig);
#pragma unroll
    for (int k = 0; k < DEG - 1; k++) roots[k] = roots[k + 1];
    roots[DEG - 1] = x;
  }
}

// lag_check (laguerre.cc:188-198)
template <int DEG>
__device__ __forceinline__ bool lag_check_reg(const cx* coeffs, const cx* roots) {
  cx c0 = coeffs[DEG];
#pragma unroll
  for (int j = 0; j < DEG; j++) c0 = cmul(c0, cneg(roots[j]));
  if (cabs_(coeffs[0]) > 1.0e-8) return (1.0 - cabs_(cdiv(c0, coeffs[0]))) < 1.0e-5;
  return cabs_(csub(c0, coeffs[0])) < 1.0e-10;
}

// Point magnification and significant-image count of the roots marked in
// `mask`, in root order (tlo_ring_amp's significant_images and point_amp):
// sig counts the images carrying at least kSignificant of the sum.
__device__ __forceinline__ double lane_amp(const Lens& L, const cx* roots, unsigned mask,
                                           int& sig) {
  double t[10], amp = 0.0;
#pragma unroll
  for (int k = 0; k < 10; k++) {
    t[k] = 0.0;
    if ((mask >> k) & 1u) {
      t[k] = point_amp_term(L, roots[k]);
      amp += t[k];
    }
  }
  sig = 0;
#pragma unroll
  for (int k = 0; k < 10; k++) sig += ((mask >> k) & 1u) && t[k] >= kSignificant * amp ? 1 : 0;
  return amp;
}

// getRootsPrecalculated + getImages (imgpoint.cc:160-274) on one lane, from
// the root-reuse state chain[0..10] (chain[10] is the slot read past the end:
// 0): the sequence wl_get_images runs on a wavefront, with the Laguerre
// routines above (the same values).  Leaves the roots in chain[0..9] and
// returns the mask of the accepted images among them.
__device__ __forceinline__ unsigned lane_images(const Lens& L, const cx* cf, cx* chain, cx zeta,
                                                const TrigTable* trig, bool binImg) {
  bool withInit = true, second = false;
  while (true) {
    cx roots[10];
    lag_solve_reg<10>(cf, chain, withInit, roots, trig);
    if (withInit && !lag_check_reg<10>(cf, roots)) {
      withInit = false;
      continue;
    }
#pragma unroll
    for (int k = 0; k < 10; k++) chain[k] = roots[k];
    chain[10] = mk(0.0, 0.0);
    unsigned mask = 0;
    int n = 0;
#pragma unroll
    for (int k = 0; k < 10; k++)
      if (img_check(L, roots[k], zeta, 1.0e-6)) {
        mask |= 1u << k;
        n++;
      }
    // a valid count: even and at least 4, or for the binary lens odd and at
    // least 3 (imgpoint.cc:236-237)
    const bool bad = binImg ? (n % 2 != 1 || n < 3) : (n % 2 != 0 || n < 4);
    if (!bad || second || !withInit) return mask;
    second = true;
    withInit = false;
  }
}

// One lane per disk sample (position g / kRingSamples, sample g %
// kRingSamples): the sample's images by a root-reuse solve from the
// position's own centre roots, its point magnification (root order), and
// whether its count of significant images differs from the centre's
// (cross[p], cleared by ring_poly_kernel, counts those samples).
__global__ void __launch_bounds__(64) ring_lane_kernel(const CurveDev* curves, const int* ptCurve,
                                                       const double2* ptPos, const cx* ptRoots,
                                                       int nPts, const TrigTable* trig,
                                                       double* ring, int* cross, const cx* coef) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= (long long)nPts * kRingSamples) return;
  const int p = (int)(g / kRingSamples), smp = (int)(g % kRingSamples);
  const CurveDev& C = curves[ptCurve[p]];
  if (!C.ring) return;
  const Lens L = C.L;
  const double2 pp = ptPos[p];
  cx chain[11];
  unsigned mask0 = 0;
#pragma unroll
  for (int k = 0; k < 10; k++) {
    chain[k] = ptRoots[(size_t)p * 10 + k];
    if (img_check(L, chain[k], mk(pp.x, pp.y), 1.0e-6)) mask0 |= 1u << k;
  }
  chain[10] = mk(0.0, 0.0);
  int sig0 = 0;
  lane_amp(L, chain, mask0, sig0);
  cx cf[11];
#pragma unroll
  for (int k = 0; k < 11; k++) cf[k] = coef[(size_t)g * 11 + k];
  const cx zeta = mk(pp.x + C.rho * kRingOff[smp][0], pp.y + C.rho * kRingOff[smp][1]);
  const unsigned mask = lane_images(L, cf, chain, zeta, trig, C.binImg);
  int sig = 0;
  ring[g] = lane_amp(L, chain, mask, sig);
  if (sig != sig0) atomicAdd(&cross[p], 1);
}

int launch_ring_amp(const CurveDev* curves, const int* ptCurve, const double2* ptPos,
                    const cx* ptRoots, int nPts, const TrigTable* trig, double* ring, int* cross,
                    cx* coef, hipStream_t s) {
  if (nPts <= 0) return 0;
  const long long lanes = (long long)nPts * kRingSamples;
  hipLaunchKernelGGL(ring_poly_kernel, dim3((unsigned)((lanes + 63) / 64)), dim3(64), 0, s, curves,
                     ptCurve, ptPos, nPts, coef, cross);
  hipLaunchKernelGGL(ring_lane_kernel, dim3((unsigned)((lanes + 63) / 64)), dim3(64), 0, s, curves,
                     ptCurve, ptPos, ptRoots, nPts, trig, ring, cross, coef);
  return (int)hipGetLastError();
}

__global__ void clear_flag_bits_kernel(uint32_t* flags, const int* pts, int n, uint32_t bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[pts[i]] &= ~bits;
}

int launch_clear_flag_bits(uint32_t* flags, const int* pts, int n, uint32_t bits, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(clear_flag_bits_kernel, dim3((n + 255) / 256), dim3(256), 0, s, flags, pts, n,
                     bits);
  return (int)hipGetLastError();
}

// Critical curves and caustics of many geometries, one per lane: the host
// routine (critical_curves_into) with the host's step table, so every value
// is the host's, and the reference's, to the bit.
__global__ void __launch_bounds__(64) ccc_kernel(const Lens* lenses, int n, const CccTable* tab,
                                                 const TrigTable* trig, cx* cc, cx* ca) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const size_t off = (size_t)g * 6 * kCccSteps;
  critical_curves_into(lenses[g], *tab, trig, cc + off, ca + off);
}

int launch_ccc(const Lens* lenses, int n, const CccTable* tab, const TrigTable* trig, cx* cc,
               cx* ca, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(ccc_kernel, dim3((n + 63) / 64), dim3(64), 0, s, lenses, n, tab, trig, cc,
                     ca);
  return (int)hipGetLastError();
}

int launch_seed(const SeedArgs& a, hipStream_t s) {
  if (a.nCurves <= 0) return 0;
  if (a.maxSplit < 1 || (long long)a.nCurves * a.maxSplit > (1ll << 30))
    return (int)hipErrorInvalidValue;
  if (a.pieceStart && !a.pieceBase) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(seed_kernel, dim3(a.nCurves * a.maxSplit), dim3(64), 0, s, a);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Run table: open addressing over 16-byte entries, one slice per group of
// kGroupPoints batch points (group b >> 14).
//   word0 = (row+1) << 39 | xL << 14 | (b & 0x3FFF)   (0 = empty)
//   word1 = label << 32 | xR                        (~0 = not yet written)
// Rows and columns are those of the grid (0 <= row, xL < 2^25 - 1; scans
// never leave the columns [0, N-1]).  A maximal run is identified by (b,
// row, xL): two floods that reach the same run compute the same (xL, xR),
// and the compare-and-swap on word0 lets exactly one of them record it.
// ---------------------------------------------------------------------------
constexpr unsigned long long kEmpty1 = ~0ull;

__device__ __forceinline__ unsigned long long run_key(int b, int row, int xl) {
  return ((unsigned long long)(row + 1) << 39) | ((unsigned long long)(unsigned)xl << 14) |
         (unsigned long long)(b & (kGroupPoints - 1));
}
// Entry of probe i from hash h in batch point b's slice.
__device__ __forceinline__ unsigned long long run_slot(const FillArgs& A, int b,
                                                       unsigned long long h, int i) {
  const int g = b >> 14;
  return A.grpBase[g] + ((h + (unsigned long long)i) & A.grpMask[g]);
}
__device__ __forceinline__ unsigned long long hash_run(unsigned long long key) {
  unsigned long long h = key;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

enum { kInsNew = 1, kInsDup = 0, kInsFull = -1 };

// Records run (b, row, xl, xr) with `label` (lane 0 works, the result is
// broadcast).  Returns kInsNew, kInsFull, or kInsDup with the stored label in
// the upper half: (label << 8) | kInsDup.  The slow path probes from the
// run's home slot on.
__device__ __forceinline__ int insert_run_slow(const FillArgs& A, int b, int row, int xl, int xr,
                                            int label) {
  int res = kInsFull;
  if (__lane_id() == 0) {
    const unsigned long long key = run_key(b, row, xl);
    const unsigned long long w1 =
        ((unsigned long long)(unsigned)label << 32) | (unsigned long long)(unsigned)xr;
    const unsigned long long h = hash_run(key);
    for (int i = 0; i < A.probeLimit; i++) {
      const unsigned long long idx = run_slot(A, b, h, i);
      unsigned long long expected = 0;
      if (acas(&A.htab[2 * idx], &expected, key)) {
        ast(&A.htab[2 * idx + 1], w1);
        res = kInsNew;
        break;
      }
      if (expected == key) {
        // the winner writes word1 right after its compare-and-swap
        unsigned long long v = ald(&A.htab[2 * idx + 1]);
        for (int spin = 0; spin < (1 << 20) && v == kEmpty1; spin++) {
          __builtin_amdgcn_s_sleep(2);
          v = ald(&A.htab[2 * idx + 1]);
        }
        const int dl = v == kEmpty1 ? -1 : (int)(v >> 32);
        res = dl < 0 ? kInsFull : ((dl << 8) | kInsDup);
        break;
      }
    }
  }
  return uni(res);
}

// Fast path: the run's home slot is free (the common case at load <= 1/2).
__device__ __forceinline__ int insert_run(const FillArgs& A, int b, int row, int xl, int xr,
                                          int label) {
  int res = -2;
  if (__lane_id() == 0) {
    const unsigned long long key = run_key(b, row, xl);
    const unsigned long long idx = run_slot(A, b, hash_run(key), 0);
    unsigned long long expected = 0;
    if (acas(&A.htab[2 * idx], &expected, key)) {
      ast(&A.htab[2 * idx + 1],
          ((unsigned long long)(unsigned)label << 32) | (unsigned long long)(unsigned)xr);
      res = kInsNew;
    }
  }
  res = uni(res);
  if (res == -2) res = insert_run_slow(A, b, row, xl, xr, label);
  return res;
}


// ---------------------------------------------------------------------------
// Labels: union-find over flood labels; label b (< nBatch) is the class of
// the reference seeds of batch point b.  Roots link larger index to smaller.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_root(int* parent, int x) {
  for (int it = 0; it < (1 << 20); it++) {
    const int p = aldi(&parent[x]);
    if (p == x) return x;
    x = p;
  }
  return x;
}

__device__ __forceinline__ void unite(const FillArgs& A, int a, int b) {
  if (__lane_id() == 0 && a >= 0 && b >= 0 && a != b) {
    for (int it = 0; it < (1 << 16); it++) {
      const int ra = find_root(A.lparent, a), rb = find_root(A.lparent, b);
      if (ra == rb) break;
      const int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
      int expected = hi;
      if (acasi(&A.lparent[hi], &expected, lo)) break;
    }
  }
}

// ---------------------------------------------------------------------------
// Global task ring (tasks handed between workgroups): bounded
// multi-producer multi-consumer ring with per-slot sequence numbers.  The
// payload is complete before the sequence number that publishes it is
// written (mem_order between them), and a slot is handed back only after its
// payload has been read.
// ---------------------------------------------------------------------------
// A task as four 64-bit words (two ints each, low half first), and back:
// built from the fields, so no task is copied through memory.
__device__ __forceinline__ unsigned long long pack2(int lo, int hi) {
  return (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32);
}
__device__ __forceinline__ void task_words(const Task& t, unsigned long long (&w)[4]) {
  w[0] = pack2(t.b, t.row);
  w[1] = pack2(t.xa, t.xb);
  w[2] = pack2(t.label, t.kd);
  w[3] = pack2(t.f1, t.f2);
}
__device__ __forceinline__ Task words_task(const unsigned long long (&w)[4]) {
  Task t;
  t.b = (int)(unsigned)w[0]; t.row = (int)(unsigned)(w[0] >> 32);
  t.xa = (int)(unsigned)w[1]; t.xb = (int)(unsigned)(w[1] >> 32);
  t.label = (int)(unsigned)w[2]; t.kd = (int)(unsigned)(w[2] >> 32);
  t.f1 = (int)(unsigned)w[3]; t.f2 = (int)(unsigned)(w[3] >> 32);
  return t;
}

__device__ __forceinline__ bool q_push(const FillArgs& A, const Task& t) {
  int ok = 0;
  if (__lane_id() == 0) {
    unsigned long long pos = ald(A.qtail);
    for (int it = 0; it < (1 << 20); it++) {
      const unsigned long long idx = pos & (A.qcap - 1);
      const unsigned long long seq = ald(&A.qseq[idx]);
      const long long dif = (long long)(seq - pos);
      if (dif == 0) {
        if (acas(A.qtail, &pos, pos + 1)) {
          unsigned long long w[4];
          task_words(t, w);
          unsigned long long* dst = reinterpret_cast<unsigned long long*>(&A.qbuf[idx]);
          ast(&dst[0], w[0]);
          ast(&dst[1], w[1]);
          ast(&dst[2], w[2]);
          ast(&dst[3], w[3]);
          mem_order();
          ast(&A.qseq[idx], pos + 1);
          ok = 1;
          break;
        }
      } else if (dif < 0) {
        break;  // full
      } else {
        pos = ald(A.qtail);
      }
    }
  }
  return bcast0(ok) != 0;
}

__device__ __forceinline__ void flag_point(const FillArgs& A, int gp, uint32_t f) {
  if (__lane_id() == 0) atomicOr(&A.flags[gp], f);
}

__device__ __forceinline__ void afaddi(int* p, int v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, AGENT);
}

// ---------------------------------------------------------------------------
// Row scanning
// ---------------------------------------------------------------------------
// Everything the ray test of one row needs, held by value so that it stays
// in registers across the scan.
struct RowCtx {
  Lens L;
  double blx, bly, step, rho2, invRho2, ldA, ldB, sumScale;
  double y, px, py;
  int N;
  int map;      // ray map of the row: kMapInRange, kMapTriple or kMapBinary
  int inRange;  // the position's lens and grid allow kMapInRange rows
};
constexpr int kMapInRange = 0;  // irsSIMD with the in-range division
constexpr int kMapTriple = 1;   // irsSIMD with IEEE division
constexpr int kMapBinary = 2;   // irsBinary (m3 = 0)

// a / b rounded to nearest, for 2^-500 <= |a| <= 2^10 and
// 2^-500 <= |b| <= 2^500: the hardware reciprocal estimate, two
// Newton-Raphson refinements, the quotient estimate and one correction by
// the residual a - b q computed exactly with a fused multiply-add
// (Markstein's correctly rounded division).  This is the IEEE double
// division sequence without its operand pre-scaling and special-value
// fix-up, which leave operands and results unchanged in that range (every
// intermediate value stays normal), so the quotient is the IEEE quotient
// to the bit.
__device__ __forceinline__ double div_in_range(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double t = fma(-b, r, 1.0);
  r = fma(r, t, r);
  t = fma(-b, r, 1.0);
  r = fma(r, t, r);
  const double q = a * r;
  const double e = fma(-b, q, a);
  return fma(e, r, q);
}

// irs_ray (the irsSIMD lane order, lcirs.cc:290-341) with its three
// quotients m_k / |z - z_k|^2 taken by div_in_range: the same value to the
// bit on rows where every |z - z_k|^2 lies in [2^-500, 2^500] and every
// mass in [2^-500, 2^10] (row_in_range checks this per row).
__device__ __forceinline__ double irs_ray_in_range(const Lens& L, double x, double y, double sx,
                                                   double sy) {
  double dr3 = L.z3.re - x, dr2 = L.z2.re - x, dr1 = L.z1.re - x, drs = sx - x;
  double di3 = L.z3.im - y, di2 = L.z2.im - y, di1 = L.z1.im - y, dis = sy - y;
  double q3 = dr3 * dr3 + di3 * di3;
  double q2 = dr2 * dr2 + di2 * di2;
  double q1 = dr1 * dr1 + di1 * di1;
  double c3 = div_in_range(L.m3, q3), c2 = div_in_range(L.m2, q2), c1 = div_in_range(L.m1, q1);
  double tr3 = dr3 * c3, tr2 = dr2 * c2, tr1 = dr1 * c1, trs = -drs;
  double ti3 = di3 * c3, ti2 = di2 * c2, ti1 = di1 * c1, tis = -dis;
  double r0 = tr3 + tr1, r1 = tr2 + trs;
  double i0 = ti3 + ti1, i1 = ti2 + tis;
  return (r0 + r1) * (r0 + r1) + (i0 + i1) * (i0 + i1);
}

// Self-test of div_in_range against IEEE division on n operand pairs drawn
// from its range: random sign, exponent and mantissa (one pair in two), the
// ends of the range, and numerators built as the rounded product of the
// denominator and a quotient half an ulp above a random double (the
// quotients nearest a rounding boundary).  bad counts the pairs whose
// results differ in any bit.
__global__ void div_check_kernel(long long n, unsigned long long seed, unsigned long long* bad) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  unsigned long long cnt = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    auto next = [&z]() {
      z += 0x9E3779B97F4A7C15ull;
      unsigned long long x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      return x ^ (x >> 31);
    };
    auto make = [](unsigned long long r, int elo, int ehi) {
      const unsigned long long mant = r & ((1ull << 52) - 1);
      const int e = elo + (int)((r >> 52) % (unsigned long long)(ehi - elo + 1));
      const unsigned long long sign = (r >> 63) << 63;
      return __longlong_as_double((long long)(sign | ((unsigned long long)(e + 1023) << 52) | mant));
    };
    const unsigned long long r0 = next(), r1 = next(), kind = next() & 7;
    double a, b;
    if (kind < 4) {
      a = make(r0, -500, 9);
      b = make(r1, -500, 499);
    } else if (kind < 6) {
      b = make(r1, -480, 480);
      const double q = make(r0, -15, 15);
      const double h = ldexp(1.0, ilogb(q) - 53);  // half an ulp of q
      a = fma(b, h, b * q);
      if (!(fabs(a) >= 0x1p-500 && fabs(a) < 0x1p10)) a = q;
    } else {
      const double ends[6] = {0x1p-500, 0x1.fffffffffffffp9, 1.0, 0x1.0000000000001p-500,
                              0x1.fffffffffffffp499, 0x1p500};
      a = ends[r0 % 3] * ((r0 >> 8) & 1 ? -1.0 : 1.0);
      b = ends[(r1 % 3) * 2 + ((r1 >> 4) & 1)] * ((r1 >> 8) & 1 ? -1.0 : 1.0);
      if (kind == 7) b = make(r1, -500, 499);
    }
    const double x = div_in_range(a, b), y = a / b;
    cnt += __double_as_longlong(x) != __double_as_longlong(y) ? 1 : 0;
  }
  if (cnt) atomicAdd(bad, cnt);
}

int launch_div_check(long long n, unsigned long long seed, unsigned long long* bad,
                     hipStream_t s) {
  if (n <= 0) return 0;
  if (!bad) return (int)hipErrorInvalidValue;
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(div_check_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256),
                     0, s, n, seed, bad);
  return (int)hipGetLastError();
}

// Whether a position's lens and pixel grid keep every row's squared lens
// distances below 2^500 and its masses in [2^-500, 2^10]: lens and pixel
// coordinates at most 2^200 in magnitude.
__device__ __forceinline__ bool lens_in_range(const Lens& L, double blx, double bly, double step,
                                              int N) {
  const double lim = 0x1p200, mlo = 0x1p-500, mhi = 0x1p10;
  const double ext = step * (double)N;
  const bool grid = fabs(blx) + ext <= lim && fabs(bly) + ext <= lim;
  const bool pos = fabs(L.z1.re) <= lim && fabs(L.z1.im) <= lim && fabs(L.z2.re) <= lim &&
                   fabs(L.z2.im) <= lim && fabs(L.z3.re) <= lim && fabs(L.z3.im) <= lim;
  const bool mass = L.m1 >= mlo && L.m1 <= mhi && L.m2 >= mlo && L.m2 <= mhi && L.m3 >= mlo &&
                    L.m3 <= mhi;
  return grid && pos && mass;
}

// Whether row y keeps every |z - z_k|^2 at least 2^-500: |y - y_k| >= 2^-250
// for the three lenses (|z - z_k|^2 >= (y - y_k)^2 for every pixel z of the
// row, and the rounded squares and sums keep that order).
__device__ __forceinline__ bool row_in_range(const Lens& L, double y) {
  const double lo = 0x1p-250;
  return fabs(L.z1.im - y) >= lo && fabs(L.z2.im - y) >= lo && fabs(L.z3.im - y) >= lo;
}

// sqrt(a) for 0 <= a <= 1 within a few units in the last place: the
// hardware reciprocal square root and two Goldschmidt refinements (the
// brightness term only feeds the fixed-point magnification sum, whose
// rounding is 2^-34 of a pixel).  a = 0 gives 0.
__device__ __forceinline__ double sqrt_unit(double a) {
  const double y = __builtin_amdgcn_rsq(a);
  double s = a * y, h = 0.5 * y;
  double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  r = fma(-s, s, a);
  s = fma(r, h, s);
  return a > 0.0 ? s : 0.0;
}

// Ray test of pixel x of the row (lcirs.cc:400-402) and its brightness
// A + B sqrt(1 - r^2) (lcirs.cc:406-407, limbdarkeningmodel.cc:45-49) as a
// multiple of 1/sumScale.  The test is the reference's to the bit (irsSIMD,
// or irsBinary for m3 = 0).  The brightness is rounded to the nearest
// integer multiple by adding 1.5 * 2^52 (exact for |value| below 2^51,
// either sign); its r^2 / rho^2 and square root are within a few units in
// the last place of the IEEE operations, far below that rounding.
// MAP is the row's ray map (R.map), fixed at compile time for the scan's
// inner loop; eval_px dispatches on R.map.
template <int MAP>
__device__ __forceinline__ bool eval_px_map(const RowCtx& R, int x, long long& val) {
  const double xx = R.blx + R.step * (double)x;
  const double rsq = MAP == kMapInRange ? irs_ray_in_range(R.L, xx, R.y, R.px, R.py)
                     : MAP == kMapTriple ? irs_ray(R.L, xx, R.y, R.px, R.py)
                                         : irs_binary(R.L, xx, R.y, R.px, R.py);
  if (!(rsq < R.rho2)) return false;
  const double v = R.ldA + R.ldB * sqrt_unit(1.0 - rsq * R.invRho2);
  const double t = fma(v, R.sumScale, 6755399441055744.0);  // 1.5 * 2^52
  val = __double_as_longlong(t) - 0x4338000000000000ll;
  return true;
}

__device__ __forceinline__ bool eval_px(const RowCtx& R, int x, long long& val) {
  return R.map == kMapInRange ? eval_px_map<kMapInRange>(R, x, val)
         : R.map == kMapTriple ? eval_px_map<kMapTriple>(R, x, val)
                               : eval_px_map<kMapBinary>(R, x, val);
}

// Follows a run leftwards from column `from` until its first miss or column
// 0; returns the run's first column and adds the pixels taken in to acc
// (lane-local) and cnt.
__device__ __forceinline__ int extend_left(const RowCtx& R, int from, long long& acc,
                                           unsigned long long& cnt, unsigned long long& evals) {
  const int ln = __lane_id();
  int xcur = from;
  while (true) {
    if (xcur < 0) return 0;
    // lane l looks at xcur - l
    const int x = xcur - ln;
    const bool valid = x >= 0;
    long long val = 0;
    const bool hit = valid && eval_px(R, x, val);
    const int nb = xcur + 1 < 64 ? xcur + 1 : 64;
    evals += (unsigned long long)nb;
    const unsigned long long m = __ballot(hit);
    const int z = ~m == 0 ? 64 : __builtin_ctzll(~m);  // hits counted from lane 0
    if (ln < z) acc += val;
    cnt += (unsigned long long)z;
    if (z < 64) return xcur - (z - 1);
    xcur -= 64;
  }
}

// A look-back recorded while a row is scanned, processed after the scan:
// row `row` over [xa, xb] (the new run's extent) without the known pixels
// [k1, k2] (the run it was reached from and its two bounding misses).
struct LookBack {
  int row, xa, xb, k1, k2, dir, label;
};
constexpr int kMaxLookBacks = 2;  // kept in two named slots (no indexed array)

// ---------------------------------------------------------------------------
// Workgroup task pool (LDS)
// ---------------------------------------------------------------------------
// A fill workgroup has one manager wave and kWorkers worker waves.  The
// workers share a pool of waiting tasks in LDS, guarded by a lock word that
// lane 0 of a wave takes.  `pending` counts the pool's tasks plus the
// workers that hold work (a task in hand or in their own ring); it changes
// only under the lock, so when it reads 0 no worker of the group holds work
// or can make any, and only the manager brings tasks in (seed tasks, or
// tasks that other groups queued in the global ring).  LDS operations of one
// wave are performed in program order; the unlock waits for the wave's LDS
// operations before it releases the word.
constexpr int kWorkers = kFillWaves - 1;
constexpr int kFeedMax = 16;     // tasks the manager brings in at a time
constexpr int kDonateMax = 8;    // tasks the manager queues for other groups at a time
constexpr int kPoolShare = 16;   // while other groups are idle, workers keep the pool at
                                 // this many tasks

struct GroupPool {
  int lock;
  int pending;     // pool tasks + workers holding work
  int head, count; // the pool ring: count tasks from head on
  int done;        // set by the manager when the fill is over
  int hungry;      // set by the manager while other groups are idle: workers then
                   // move waiting tasks to the pool, and the manager on to the
                   // global ring
  int pad[2];
  Task slot[kPoolCap];
};

#define WGS __HIP_MEMORY_SCOPE_WORKGROUP
__device__ __forceinline__ int lds_ld(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, WGS); }
__device__ __forceinline__ void lds_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, WGS);
}
// Waits until the wave's LDS (and scalar) memory operations are done.
__device__ __forceinline__ void lds_wait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt and expcnt not waited for
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Lane 0 takes and releases the pool lock.
__device__ __forceinline__ void pool_lock(GroupPool* G) {
  while (true) {
    int expected = 0;
    if (__hip_atomic_compare_exchange_strong(&G->lock, &expected, 1, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, WGS))
      break;
    __builtin_amdgcn_s_sleep(1);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void pool_unlock(GroupPool* G) {
  lds_wait();
  lds_st(&G->lock, 0);
}

__device__ __forceinline__ Task uni_task(const Task& t) {
  Task u;
  u.b = uni(t.b); u.row = uni(t.row); u.xa = uni(t.xa); u.xb = uni(t.xb);
  u.label = uni(t.label); u.kd = uni(t.kd); u.f1 = uni(t.f1); u.f2 = uni(t.f2);
  return u;
}

// A worker that holds work adds the (wave-uniform) task t; false when the
// pool is full.
__device__ __forceinline__ bool pool_push(GroupPool* G, const Task& t) {
  int ok = 0;
  if (__lane_id() == 0) {
    pool_lock(G);
    const int c = lds_ld(&G->count);
    if (c < kPoolCap) {
      G->slot[(lds_ld(&G->head) + c) & (kPoolCap - 1)] = t;
      lds_st(&G->count, c + 1);
      lds_st(&G->pending, lds_ld(&G->pending) + 1);
      ok = 1;
    }
    pool_unlock(G);
  }
  return uni(ok) != 0;
}

// A worker takes a task of the pool: the newest (newest = true) or the
// oldest.  busy: the worker holds work now (its last task is done and its
// ring is empty): it takes a task (pending - 1) or, with the pool empty,
// gives up its count (pending - 1) and is idle from then on.  !busy: an idle
// worker takes a task and is counted as holding work (pending unchanged),
// or stays idle.  The newest first keeps a group on the sweeps its scouts
// just opened (on the images still growing) before older ones: the
// surrogate's fallback fill of the C2 curve takes ~3% less time with it,
// the heaviest 41 positions ~5% less, C3's fill the same.
__device__ __forceinline__ bool pool_take(GroupPool* G, Task& t, bool busy, bool newest) {
  int ok = 0;
  Task got = t;
  if (__lane_id() == 0) {
    pool_lock(G);
    const int c = lds_ld(&G->count);
    if (c > 0) {
      const int h = lds_ld(&G->head);
      if (newest) {
        got = G->slot[(h + c - 1) & (kPoolCap - 1)];
      } else {
        got = G->slot[h];
        lds_st(&G->head, (h + 1) & (kPoolCap - 1));
      }
      lds_st(&G->count, c - 1);
      ok = 1;
    }
    if (busy) lds_st(&G->pending, lds_ld(&G->pending) - 1);
    pool_unlock(G);
  }
  ok = uni(ok);
  if (ok) t = uni_task(got);
  return ok != 0;
}

// Manager: adds the tasks held by lanes [lo, hi) (pending + their number).
// Returns how many fitted; the caller makes room first.
__device__ __forceinline__ int pool_add_lanes(GroupPool* G, const Task& t, int lo, int hi) {
  const int ln = __lane_id();
  int base = 0, fit = 0;
  if (ln == 0) {
    pool_lock(G);
    const int c = lds_ld(&G->count);
    fit = hi - lo < kPoolCap - c ? hi - lo : kPoolCap - c;
    base = lds_ld(&G->head) + c;
  }
  fit = uni(fit);
  base = uni(base);
  if (ln >= lo && ln < lo + fit) G->slot[(base + ln - lo) & (kPoolCap - 1)] = t;
  if (ln == 0) {
    lds_st(&G->count, lds_ld(&G->count) + fit);
    lds_st(&G->pending, lds_ld(&G->pending) + fit);
    pool_unlock(G);
  }
  return fit;
}

// Manager: takes up to n of the pool's oldest tasks into lanes [0, k)
// (pending - k) and returns k.
__device__ __forceinline__ int pool_remove_lanes(GroupPool* G, Task& t, int n) {
  const int ln = __lane_id();
  int k = 0, h = 0;
  if (ln == 0) {
    pool_lock(G);
    const int c = lds_ld(&G->count);
    k = n < c ? n : c;
    h = lds_ld(&G->head);
  }
  k = uni(k);
  h = uni(h);
  if (ln < k) t = G->slot[(h + ln) & (kPoolCap - 1)];
  if (ln == 0) {
    lds_st(&G->head, (h + k) & (kPoolCap - 1));
    lds_st(&G->count, lds_ld(&G->count) - k);
    lds_st(&G->pending, lds_ld(&G->pending) - k);
    pool_unlock(G);
  }
  return k;
}

// Global ring, many tasks per call (manager): lanes [0, n) offer one task
// each; the leading k slots that are free at the tail are claimed with one
// compare-and-swap and written in parallel.  Returns k (the tasks of lanes
// [0, k) are queued).
__device__ __forceinline__ int q_push_lanes(const FillArgs& A, const Task& t, int n) {
  const int ln = __lane_id();
  unsigned long long pos = 0;
  int k = 0;
  for (int it = 0; it < 64; it++) {
    if (ln == 0) pos = ald(A.qtail);
    pos = uni64(pos);
    const bool fr = ln < n && ald(&A.qseq[(pos + ln) & (A.qcap - 1)]) == pos + ln;
    const unsigned long long m = __ballot(fr);
    k = ~m == 0 ? 64 : __builtin_ctzll(~m);
    if (k == 0) return 0;  // full, or another producer moved the tail: try again later
    int ok = 0;
    if (ln == 0) {
      unsigned long long exp = pos;
      ok = acas(A.qtail, &exp, pos + k);
    }
    if (uni(ok)) break;
    k = 0;
  }
  if (ln < k) {
    const unsigned long long idx = (pos + ln) & (A.qcap - 1);
    unsigned long long w[4];
    task_words(t, w);
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(&A.qbuf[idx]);
    ast(&dst[0], w[0]);
    ast(&dst[1], w[1]);
    ast(&dst[2], w[2]);
    ast(&dst[3], w[3]);
    mem_order();
    ast(&A.qseq[idx], pos + ln + 1);
  }
  return k;
}

// Global ring, many tasks per call (manager): takes the leading k <= n
// ready slots at the head into lanes [0, k); returns k.
__device__ __forceinline__ int q_pop_lanes(const FillArgs& A, Task& t, int n) {
  const int ln = __lane_id();
  unsigned long long pos = 0;
  int k = 0;
  for (int it = 0; it < 64; it++) {
    if (ln == 0) pos = ald(A.qhead);
    pos = uni64(pos);
    const bool rd = ln < n && ald(&A.qseq[(pos + ln) & (A.qcap - 1)]) == pos + ln + 1;
    const unsigned long long m = __ballot(rd);
    k = ~m == 0 ? 64 : __builtin_ctzll(~m);
    if (k == 0) return 0;
    int ok = 0;
    if (ln == 0) {
      unsigned long long exp = pos;
      ok = acas(A.qhead, &exp, pos + k);
    }
    if (uni(ok)) break;
    k = 0;
  }
  if (ln < k) {
    const unsigned long long idx = (pos + ln) & (A.qcap - 1);
    unsigned long long w[4];
    const unsigned long long* src = reinterpret_cast<unsigned long long*>(&A.qbuf[idx]);
    w[0] = ald(const_cast<unsigned long long*>(&src[0]));
    w[1] = ald(const_cast<unsigned long long*>(&src[1]));
    w[2] = ald(const_cast<unsigned long long*>(&src[2]));
    w[3] = ald(const_cast<unsigned long long*>(&src[3]));
    mem_order();
    ast(&A.qseq[idx], pos + ln + A.qcap);
    t = words_task(w);
  }
  return k;
}

// Manager: puts the tasks of lanes [lo, hi) back into the pool, or, where it
// is full, into the global ring; waits while both are full (workers keep
// draining the pool), up to the deadline (s_memrealtime ticks).
__device__ __forceinline__ void pool_add_all(const FillArgs& A, GroupPool* G, const Task& t,
                                             int lo, int hi, long long deadline) {
  for (int it = 0; lo < hi; it++) {
    lo += pool_add_lanes(G, t, lo, hi);
    if (lo >= hi) break;
    // lanes [lo, hi) still hold tasks: offer them to the ring from lane 0 on
    Task u = t;
    const int src = __lane_id() + lo;
    u.b = __shfl(t.b, src, 64); u.row = __shfl(t.row, src, 64);
    u.xa = __shfl(t.xa, src, 64); u.xb = __shfl(t.xb, src, 64);
    u.label = __shfl(t.label, src, 64); u.kd = __shfl(t.kd, src, 64);
    u.f1 = __shfl(t.f1, src, 64); u.f2 = __shfl(t.f2, src, 64);
    lo += q_push_lanes(A, u, hi - lo);
    if (lo >= hi) break;
    if ((long long)__builtin_amdgcn_s_memrealtime() > deadline) {
      if (__lane_id() == 0) atomicExch(A.abortFlag, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(16);
  }
}

// Manager: seed tasks [st, st + k) of the batch into lanes [0, k); returns k.
__device__ __forceinline__ int claim_seeds(const FillArgs& A, Task& t, int n) {
  const int ln = __lane_id();
  int st = A.nSeedTasks;
  if (ln == 0 && aldi(A.seedNext) < A.nSeedTasks) st = atomicAdd(A.seedNext, n);
  st = uni(st);
  int k = A.nSeedTasks - st;
  k = k < 0 ? 0 : (k > n ? n : k);
  if (ln < k) {
    const int2 sk = A.seedTasks[st + ln];
    const int gp = A.batchPts[sk.x];
    const int2 px = A.seeds[(size_t)gp * kMaxSeeds + sk.y];
    const long long N = A.curves[A.ptCurve[gp]].N;
    // the reference skips seeds outside rows 1..N-1 (lcirs.cc:386-389)
    const bool rowOk = px.y > 0 && (long long)px.y < N;
    t.b = sk.x; t.row = rowOk ? px.y : -1; t.xa = px.x; t.xb = px.x; t.label = sk.x;
    t.kd = 1; t.f1 = 0; t.f2 = -1;
  }
  return k;
}

// State of the task being processed (uniform across the wave).
struct TaskState {
  const FillArgs* A;
  Task t;
  int curB;                 // batch point whose constants R holds (-1: none)
  bool heavy;               // its predicted rays reach shareRays
  RowCtx R;
  int kind, dir, gp, N;
  int xa, xb;               // the known run next to the scanned row (kind 0:
                            // the parent, kind 3: the child)
  Task cont;                // follow-up kept by this wave (the next sweep step)
  bool haveCont;
  GroupPool* G;             // the workgroup's shared pool
  int nBack;
  LookBack back0, back1;
  int accLabel;             // totals of new runs, credited to accLabel when
  long long accSum;         // the label changes or the sweep ends
  unsigned long long accCnt;
  unsigned long long evals, nIns, nDup, insTicks;
  unsigned long long nScout, nScoutRuns;  // scout tasks taken, runs they found
  bool landed;              // the scout task found a run (new or known)
  // this wave's ring of waiting tasks in LDS: lcount tasks from lhead on
  Task* ring;
  int lhead, lcount;
};

// Task hand-off.  Follow-ups a worker does not continue with go to its own
// ring in LDS (no atomics: one owner); the newest is taken next (depth
// first).  While other workers of the group are idle, a worker moves its
// oldest waiting tasks to the group's pool; a full ring spills its oldest
// task there too, and a full pool into the global ring.
__device__ __forceinline__ bool ring_donate(TaskState& S) {
  const Task t = S.ring[S.lhead];
  if (!pool_push(S.G, t) && !q_push(*S.A, t)) return false;
  S.lhead = (S.lhead + 1) & (kLocalCap - 1);
  S.lcount--;
  return true;
}

__device__ __forceinline__ void ring_push(TaskState& S, const Task& f) {
  if (S.lcount == kLocalCap && !ring_donate(S)) {
    flag_point(*S.A, S.gp, kFlagQueueFull);  // everything full: the task is lost
    return;
  }
  const int slot = (S.lhead + S.lcount) & (kLocalCap - 1);
  if (__lane_id() == 0) S.ring[slot] = f;
  S.lcount++;
}

__device__ __forceinline__ Task ring_pop(TaskState& S) {
  S.lcount--;
  return S.ring[(S.lhead + S.lcount) & (kLocalCap - 1)];
}

// Adds the signed value v to a split total (p[0], p[1]) that stands for
// p[1] * 2^32 + p[0]: p[0] sums the low 32 bits of every value added (as an
// unsigned number) and p[1] the rest (v >> 32, arithmetic).  Both adds need
// no result, so they are issued and not waited for, and any number of
// concurrent adds sums exactly as long as fewer than 2^32 values are added
// to one total.
__device__ __forceinline__ void add_split(unsigned long long* p, long long v) {
  atomicAdd(p, (unsigned long long)(v & 0xFFFFFFFFll));
  const long long hi = v >> 32;
  if (hi) atomicAdd(p + 1, (unsigned long long)hi);
}

__device__ __forceinline__ void flush_acc(TaskState& S) {
  if (S.accCnt && __lane_id() == 0) {
    atomicAdd(&S.A->lcount[S.accLabel], S.accCnt);
    add_split(&S.A->lsum[2 * (size_t)S.accLabel], S.accSum);
  }
  S.accCnt = 0;
  S.accSum = 0;
}

__device__ __forceinline__ Task make_task(int b, int row, int xa, int xb, int label, int kind,
                                          int dir, int f1, int f2) {
  Task f;
  f.b = b; f.row = row; f.xa = xa; f.xb = xb; f.label = label;
  f.kd = kind | (dir < 0 ? 4 : 0);
  f.f1 = f1; f.f2 = f2;
  return f;
}

constexpr int kQueueOnly = 8;  // kd flag: never kept by the wave
constexpr int kShare = 16;     // kd flag: to the group's pool at once (the wave's
                               // ring when the pool is full)

// A follow-up of a run, dispatched as soon as it is made: the first one that
// may stay with the wave becomes its next step, tasks marked kShare go to the
// group's pool, the others to the wave's ring.  (Collecting them first made
// the compiler keep them in a scratch array.)
__device__ __forceinline__ void follow(TaskState& S, Task f) {
  const bool queueOnly = (f.kd & kQueueOnly) != 0, share = (f.kd & kShare) != 0;
  f.kd &= ~(kQueueOnly | kShare);
  if (!queueOnly && !share && !S.haveCont) {
    S.cont = f;
    S.haveCont = true;
  } else if (!(share && pool_push(S.G, f))) {
    ring_push(S, f);
  }
}

// Rows a scout jumps from a run of w pixels: scout_jump, shortened in
// proportion where runs are wider than scout_width pixels (0: never), so
// that the rows between two scouts hold about the same number of pixels on
// wide stretches of an image as on narrow ones.  A scout task carries the
// jump it was sent with in the high half of f2 (the low half counts the
// scouts of its chain since the last leap).
constexpr int kMinJump = 16;
__device__ __forceinline__ int scout_jump_for(const FillArgs& A, int w) {
  if (A.scoutWidth <= 0 || w <= A.scoutWidth) return A.scoutJump;
  const long long j = (long long)A.scoutJump * A.scoutWidth / w;
  const int lo = kMinJump < A.scoutJump ? kMinJump : A.scoutJump;
  return j < lo ? lo : (int)j;
}

// A maximal hit run [xl, xr] of the row being scanned was found: record it,
// credit its pixels to its label and schedule the neighbouring rows.
__device__ __forceinline__ void handle_run(TaskState& S, int xl, int xr, long long sum,
                                           unsigned long long cnt) {
  const FillArgs& A = *S.A;
  const Task& t = S.t;
  int label = t.label;
  if (S.kind == 2) {
    S.landed = true;
    // a scout run starts a class of its own
    int nl = 0;
    if (__lane_id() == 0) nl = atomicAdd(A.labelNext, 1);
    nl = uni(nl);
    if (nl >= A.labelCap) {
      flag_point(A, S.gp, kFlagLabelFull);
      return;
    }
    if (__lane_id() == 0) A.lpoint[nl] = t.b;
    label = nl;
  }
  const unsigned long long ti0 = __builtin_amdgcn_s_memrealtime();
  const int ins = insert_run(A, t.b, t.row, xl, xr, label);
  S.insTicks += __builtin_amdgcn_s_memrealtime() - ti0;
  if (ins == kInsFull) {
    flag_point(A, S.gp, kFlagHashFull);
    return;
  }
  if ((ins & 0xFF) == kInsDup) {
    S.nDup++;
    if (S.kind != 2) unite(A, label, ins >> 8);
    return;
  }
  S.nIns++;
  S.nScoutRuns += S.kind == 2 ? 1 : 0;
  if (xl <= 1 || xr >= S.N - 2) flag_point(A, S.gp, kFlagEdge);
  // a run adds at most 2^60 units: flushing from 2^61 on keeps the total
  // within 2^62
  if (label != S.accLabel || S.accSum > (1ll << 61) || S.accSum < -(1ll << 61)) {
    flush_acc(S);
    S.accLabel = label;
  }
  S.accCnt += cnt;
  S.accSum += sum;
  const int row = t.row, N = S.N, center = (int)(((long long)xl + xr) / 2);
  const int d = S.dir;  // direction of travel
  if (S.kind == 1) {
    if (row + 1 < N) follow(S, make_task(t.b, row + 1, xl, xr, label, 0, 1, center, 0));
    if (row - 1 >= 0) follow(S, make_task(t.b, row - 1, xl, xr, label, 0, -1, center, 0));
  } else if (S.kind == 2) {
    // A scout landed (t.f1: the column of the run it was sent from, one jump
    // back; t.f2: that jump's length and the scouts in its chain since the
    // last leap).  The wave sends the next scout straight on at once, with
    // the shift measured over the last jump, and keeps it as its next step, so that the scouts run ahead
    // one task per jump; every leapEvery-th scout also sends a leap scout
    // leapJumps jumps ahead to the shared queue, which starts a chain of its
    // own, so that long images get many chains.  The two sweeps of the new
    // class go to the shared queue for idle waves; they start as if they had
    // just sent a scout.
    const int count = t.f2 & 0xFFFF, prevJump = (int)((unsigned)t.f2 >> 16);
    const int J = scout_jump_for(A, xr - xl + 1);
    // the shift over the last jump, rescaled to this jump
    const long long shift = ((long long)center - t.f1) * J / (prevJump > 0 ? prevJump : J);
    const long long ash = shift < 0 ? -shift : shift;
    for (int leap = 0; leap < 2 && A.scoutRows > 0; leap++) {
      if (leap == 1 && (A.leapEvery <= 0 || (count + 1) % A.leapEvery != 0)) break;
      const long long k = leap ? A.leapJumps : 1;
      const long long rt = (long long)row + (long long)d * J * k;
      if (rt < 0 || rt >= N) continue;
      const long long pc = center + shift * k;
      const long long hw = ((long long)xr - xl) / 2 + ash * k / 4 + 32 * k;
      long long sa = pc - hw, sb = pc + hw;
      if (sa < 0) sa = 0;
      if (sb > N - 1) sb = N - 1;
      if (sa > sb) continue;
      Task f = make_task(t.b, (int)rt, (int)sa, (int)sb, -1, 2, d, center,
                         (J << 16) | (leap ? 0 : (count + 1) & 0xFFFF));
      if (leap) {
        // the leap's chain measures its next shift from the column expected
        // one jump before its landing row
        f.f1 = (int)(center + shift * (A.leapJumps - 1));
        f.kd |= kQueueOnly | kShare;
      }
      follow(S, f);
    }
    Task up = make_task(t.b, row + 1, xl, xr, label, 0, 1, center, -1);
    Task dn = make_task(t.b, row - 1, xl, xr, label, 0, -1, center, -1);
    up.kd |= kShare;
    dn.kd |= kShare;
    if (row + 1 < N) follow(S, up);
    if (row - 1 >= 0) follow(S, dn);
  } else if (S.kind == 3) {
    // a run found looking back: sweep on away from the run it was reached
    // from, and look back towards that run's row where this one reaches past it
    if (row + d >= 0 && row + d < N)
      follow(S, make_task(t.b, row + d, xl, xr, label, 0, d, center, 0));
    if (row - d >= 0 && row - d < N && (xl < S.xa - 1 || xr > S.xb + 1)) {
      Task f = make_task(t.b, row - d, xl, xr, label, 3, -d, S.xa - 1, S.xb + 1);
      f.kd |= kQueueOnly;
      follow(S, f);
    }
  } else {
    // sweep step: 1. the next row in the direction of travel
    if (row + d >= 0 && row + d < N) {
      // f2 >= 0: rows since the sweep's anchor, no scout sent yet;
      // f2 < 0: -(rows since the last scout) - 1
      int a1 = t.f1, a2 = t.f2 >= 0 ? t.f2 + 1 : t.f2 - 1;
      const int rows = a2 >= 0 ? a2 : -a2 - 1;
      const int J = scout_jump_for(A, xr - xl + 1);
      const bool due = A.scoutRows > 0 && ((a2 >= 0 && rows >= A.scoutRows) ||
                                           (a2 < 0 && rows >= J));
      if (due) {
        // send a scout ahead along the run's drift, then re-anchor
        const long long rt = (long long)row + (long long)d * J;
        if (rt >= 0 && rt < N) {
          const long long drift = rows > 0 ? ((long long)center - a1) / rows : 0;
          const long long pc = center + drift * J;
          const long long adr = drift < 0 ? -drift : drift;
          const long long hw = ((long long)xr - xl) / 2 + adr * J / 4 + 32;
          long long sa = pc - hw, sb = pc + hw;
          if (sa < 0) sa = 0;
          if (sb > N - 1) sb = N - 1;
          if (sa <= sb) {
            // to the wave's ring (the shared queue while some wave is idle),
            // or to the shared queue at once on the images of positions
            // predicted heavy, whose scouts mostly land: sweeps of many
            // small images send many scouts that miss, and queuing each of
            // them at once makes the shared queue's head and tail the
            // bottleneck
            Task f = make_task(t.b, (int)rt, (int)sa, (int)sb, -1, 2, d, center, J << 16);
            f.kd |= kQueueOnly | ((A.scoutShare || S.heavy) ? kShare : 0);
            follow(S, f);
          }
        }
        a1 = center;
        a2 = -1;
      }
      follow(S, make_task(t.b, row + d, xl, xr, label, 0, d, a1, a2));
    }
    // 2. the row the sweep came from, where the new run reaches past the
    //    parent run [xa, xb] and its two bounding misses: looked at after
    //    this scan (or queued when two look-backs are already waiting)
    if (row - d >= 0 && row - d < N && (xl < S.xa - 1 || xr > S.xb + 1)) {
      LookBack lb;
      lb.row = row - d; lb.xa = xl; lb.xb = xr;
      lb.k1 = S.xa - 1; lb.k2 = S.xb + 1; lb.dir = -d; lb.label = label;
      if (S.nBack == 0) {
        S.back0 = lb;
        S.nBack = 1;
      } else if (S.nBack == 1) {
        S.back1 = lb;
        S.nBack = 2;
      } else {
        Task f = make_task(t.b, lb.row, xl, xr, label, 3, -d, lb.k1, lb.k2);
        f.kd |= kQueueOnly;
        follow(S, f);
      }
    }
  }
}

// The scan's next 64-pixel steps from column cb + 64 on: each step's hit
// mask m and pixel count nb; while a step lies wholly inside the open run its
// pixels are added to acc and cnt and the next step follows (up to column
// hi).  Returns whether the last step taken lies inside the open run; cb is
// left at that step's first column and val holds its lane's brightness.
template <int MAP>
__device__ __forceinline__ bool scan_steps(const RowCtx& R, bool open, int hi, int& cb,
                                           unsigned long long& m, int& nb, long long& val,
                                           long long& acc, unsigned long long& cnt,
                                           unsigned& evals) {
  const int ln = __lane_id();
  bool inside;
  while (true) {
    if (open && cb + 192 <= hi + 1) {
      // two whole steps inside the scan: both evaluated before either
      // ballot (twice the independent work in flight); when the first does
      // not lie wholly inside the run the second is dropped, and evaluated
      // again by the caller's next step
      const int x1 = cb + 64 + ln;
      long long v1 = 0, v2 = 0;
      const bool h1 = eval_px_map<MAP>(R, x1, v1);
      const bool h2 = eval_px_map<MAP>(R, x1 + 64, v2);
      const unsigned long long m1 = __ballot(h1), m2 = __ballot(h2);
      evals += 128u;
      cb += 64;
      nb = 64;
      if (m1 != ~0ull) {
        m = m1;
        val = v1;
        return false;
      }
      acc += v1;
      cnt += 64ull;
      cb += 64;
      m = m2;
      val = v2;
      if (m2 != ~0ull) return false;
      acc += v2;
      cnt += 64ull;
      if (cb + 64 > hi) return true;
      continue;
    }
    cb += 64;
    const int x = cb + ln;
    val = 0;
    const bool hit = x <= hi && eval_px_map<MAP>(R, x, val);
    m = __ballot(hit);
    nb = hi - cb + 1 < 64 ? hi - cb + 1 : 64;
    evals += (unsigned)nb;
    const unsigned long long full = nb == 64 ? ~0ull : ((1ull << nb) - 1ull);
    inside = open && m == full;
    if (!inside) break;
    acc += val;
    cnt += (unsigned long long)nb;
    if (cb + 64 > hi) break;
  }
  return inside;
}

// Maximal hit runs of the row that meet [s, e].  The row is evaluated over a
// window that pads [s, e] to whole 64-pixel steps on the sides where runs
// may continue (extL, extR); elsewhere the pixel next to [s, e] is a known
// miss.  Runs inside the padding that do not reach [s, e] are ignored; a run
// that reaches the window's right edge is followed past it in the same loop,
// one that reaches its left edge by extend_left.  Each run is handed to
// handle_run once both its ends are known.  A 64-pixel step that lies wholly
// inside the open run, or wholly outside any run, costs only its ray tests.
__device__ __forceinline__ void scan_window(const RowCtx& R, int s, int e, bool extL, bool extR,
                                            TaskState& S) {
  const int ln = __lane_id();
  const int len = e - s + 1;
  const int extra = ((len + 63) / 64) * 64 - len;
  const int padL = extL ? (extR ? extra / 2 : extra) : 0;
  const int padR = extR ? extra - padL : 0;
  const int ws = s - padL < 0 ? 0 : s - padL;
  const int we = e + padR > R.N - 1 ? R.N - 1 : e + padR;
  int hi = we;             // last column of the scan (grows past we while a run is followed)
  bool open = false;
  int runL = 0, cb = ws - 64, nb = 0, pos = 0;
  unsigned long long m = 0, cnt = 0;
  long long acc = 0, val = 0;  // lane-local brightness of the open run, of this lane's pixel
  for (int guard = 0;; guard++) {
    if (guard > (1 << 27)) {
      // more iterations than any row of 2^25 pixels needs: never expected;
      // the point is flagged and the launch ends
      flag_point(*S.A, S.gp, kFlagTimeout);
      if (ln == 0) atomicExch(S.A->abortFlag, 1);
      break;
    }
    bool emit = false, last = false;
    int eR = 0;
    if (pos >= nb) {
      // the next 64 pixels, or the end of the scan
      if (cb + 64 > hi) {
        if (!open) break;
        if (hi == we && extR && we < R.N - 1) {
          hi = R.N - 1;  // follow the run past the window, from we + 1 on
          cb = we + 1 - 64;
        } else {
          last = true;  // the run reaches the scan's last column
          eR = hi;
          emit = hi >= s && runL <= e;
          if (!emit) break;
        }
      }
      if (!last) {
        if (!open && cb + 64 > e) break;  // no later run can meet [s, e]
        // the next steps; while each lies wholly inside the open run it is
        // taken in without the boundary bookkeeping
        unsigned evals = 0;
        const bool inside =
            R.map == kMapInRange
                ? scan_steps<kMapInRange>(R, open, hi, cb, m, nb, val, acc, cnt, evals)
            : R.map == kMapTriple
                ? scan_steps<kMapTriple>(R, open, hi, cb, m, nb, val, acc, cnt, evals)
                : scan_steps<kMapBinary>(R, open, hi, cb, m, nb, val, acc, cnt, evals);
        S.evals += evals;
        pos = (inside || (!open && m == 0)) ? nb : 0;
        continue;
      }
    } else if (open) {
      const unsigned long long rest = ~(m >> pos);
      int z = rest == 0 ? 64 - pos : __builtin_ctzll(rest);
      if (z > nb - pos) z = nb - pos;
      if (ln >= pos && ln < pos + z) acc += val;
      cnt += (unsigned long long)z;
      pos += z;
      if (pos < nb) {
        // the run ends at cb + pos - 1
        eR = cb + pos - 1;
        emit = eR >= s && runL <= e;
        if (!emit) {
          open = false;
          acc = 0;
          cnt = 0;
          if (hi > we) break;  // the followed run has ended
        }
      }
    } else {
      const unsigned long long rest = m >> pos;
      if (rest == 0) {
        pos = nb;
        continue;
      }
      const int tz = __builtin_ctzll(rest);
      if (cb + pos + tz > e) break;  // no later run can meet [s, e]
      runL = cb + pos + tz;
      open = true;
      pos += tz;
    }
    if (emit) {
      if (runL == ws && extL && ws > 0) runL = extend_left(R, ws - 1, acc, cnt, S.evals);
      handle_run(S, runL, eR, wave_sum(acc), cnt);
      acc = 0;
      cnt = 0;
      open = false;
      if (last || hi > we) break;
    }
  }
}

// ---------------------------------------------------------------------------
// Fill
// ---------------------------------------------------------------------------
// Processes one task and the look-backs it records, as up to six scans of one
// row each: the task's row (one interval, or two around a known run), then
// each look-back (the same).  The follow-up kept for this wave (if any) is
// left in S.cont / S.haveCont.
// The task's fields as wave-uniform values, and the constants of its source
// position (kept while the wave stays on that position).
__device__ __forceinline__ void load_task(TaskState& S) {
  const FillArgs& A = *S.A;
  S.t.b = uni(S.t.b);
  S.t.row = uni(S.t.row);
  S.t.xa = uni(S.t.xa);
  S.t.xb = uni(S.t.xb);
  S.t.label = uni(S.t.label);
  S.t.kd = uni(S.t.kd);
  S.t.f1 = uni(S.t.f1);
  S.t.f2 = uni(S.t.f2);
  S.kind = S.t.kd & 3;
  S.dir = (S.t.kd & 4) ? -1 : 1;
  if (S.t.b != S.curB) {
    // constants of this source position, kept while the wave stays on it
    S.curB = S.t.b;
    S.gp = uni(A.batchPts[S.t.b]);
    const CurveDev& C = A.curves[uni(A.ptCurve[S.gp])];
    S.N = uni((int)C.N);
    S.R.L.m1 = unid(C.L.m1); S.R.L.m2 = unid(C.L.m2); S.R.L.m3 = unid(C.L.m3);
    S.R.L.z1 = mk(unid(C.L.z1.re), unid(C.L.z1.im));
    S.R.L.z2 = mk(unid(C.L.z2.re), unid(C.L.z2.im));
    S.R.L.z3 = mk(unid(C.L.z3.re), unid(C.L.z3.im));
    S.R.blx = unid(C.blx);
    S.R.bly = unid(C.bly);
    S.R.step = unid(C.step);
    S.R.rho2 = unid(C.rho2);
    S.R.invRho2 = 1.0 / S.R.rho2;
    S.R.ldA = unid(C.ldA);
    S.R.ldB = unid(C.ldB);
    S.R.sumScale = unid(C.sumScale);
    const int bin = uni(C.binIrs);
    S.R.inRange = bin ? 0 : uni((int)lens_in_range(S.R.L, S.R.blx, S.R.bly, S.R.step, S.N));
    S.R.map = bin ? kMapBinary : kMapTriple;
    S.heavy = A.shareRays > 0.0f && A.ptCost[S.gp] >= A.shareRays;
    S.R.N = S.N;
    const double2 pos = A.ptPos[S.gp];
    S.R.px = unid(pos.x);
    S.R.py = unid(pos.y);
  }
}

__device__ __forceinline__ void process_task(TaskState& S) {
  const FillArgs& A = *S.A;
  S.haveCont = false;
  S.nBack = 0;
  load_task(S);
  S.nScout += S.kind == 2 ? 1 : 0;
  if (S.t.row < 0 || S.t.row >= S.N) return;
  if (S.kind == 1 && (S.t.xa < 0 || S.t.xa > S.N - 1)) {  // seed column outside the grid
    flag_point(A, S.gp, kFlagEdge);
    return;
  }
  int k1 = S.kind == 3 ? S.t.f1 : 0, k2 = S.kind == 3 ? S.t.f2 : -1;
  S.xa = S.t.xa;
  S.xb = S.t.xb;
  const bool scout = S.kind == 2;
  const int trow = S.t.row, txa = S.t.xa, txb = S.t.xb, tf1 = S.t.f1, tf2 = S.t.f2, tdir = S.dir;
  S.landed = false;
  for (int job = 0; job < 6; job++) {
    if (job >= 2) {
      const int li = (job - 2) >> 1;
      if (li >= S.nBack) break;
      if ((job & 1) == 0) {
        const LookBack lb = li == 0 ? S.back0 : S.back1;
        S.t.row = lb.row;
        S.t.label = lb.label;
        S.kind = 3;
        S.dir = lb.dir;
        S.xa = lb.xa;
        S.xb = lb.xb;
        k1 = lb.k1;
        k2 = lb.k2;
      }
    } else if (job == 1 && k1 > k2) {
      continue;  // a task without a known run inside its interval scans once
    }
    if ((job & 1) == 0) {
      S.R.y = S.R.bly + S.R.step * (double)S.t.row;
      if (S.R.map != kMapBinary)
        S.R.map = S.R.inRange && uni((int)row_in_range(S.R.L, S.R.y)) ? kMapInRange : kMapTriple;
    }
    const int lo = S.xa < 0 ? 0 : S.xa, hi = S.xb > S.N - 1 ? S.N - 1 : S.xb;
    int s = lo, e = hi;
    bool extL = true, extR = true;
    if (k1 <= k2) {
      // two sides of the known pixels [k1, k2]
      if ((job & 1) == 0) {
        e = hi < k1 - 1 ? hi : k1 - 1;
        extR = hi < k1 - 1;
      } else {
        s = lo > k2 + 1 ? lo : k2 + 1;
        extL = lo > k2 + 1;
      }
    }
    if (s > e) continue;
    scan_window(S.R, s, e, extL, extR, S);
  }
  S.nBack = 0;
  if (scout && !S.landed && A.scoutRetry) {
    // A scout that found nothing (its chain's prediction failed, or the
    // image ends before its row): a retry from the same origin, half the
    // jump ahead, with a window of the same width about the midpoint of the
    // origin column and the failed window's centre.
    const int J = (int)((unsigned)tf2 >> 16), count = tf2 & 0xFFFF;
    const int J2 = J / 2;
    if (J2 >= kMinJump) {
      const long long rt = (long long)trow - (long long)tdir * (J - J2);
      const long long pc = ((long long)txa + txb) / 2;
      const long long c2 = tf1 + (pc - tf1) / 2;
      const long long hw = ((long long)txb - txa) / 2;
      long long sa = c2 - hw, sb = c2 + hw;
      if (sa < 0) sa = 0;
      if (sb > S.N - 1) sb = S.N - 1;
      if (rt >= 0 && rt < S.N && sa <= sb) {
        follow(S, make_task(S.t.b, (int)rt, (int)sa, (int)sb, -1, 2, tdir, tf1,
                            (J2 << 16) | count));
      }
    }
  }
}

// The manager wave of a fill workgroup.  It keeps the group's pool fed while
// workers are idle (seed tasks first, then tasks that other groups queued in
// the global ring), queues some of a large pool for other groups while any
// group is idle, and runs the end test.  A group joins the global active
// count when it starts, before it takes any task; a group whose `pending`
// count reads 0 with the seed list used up leaves the count and polls the
// global ring, and rejoins the count before it takes tasks from the ring, so
// a task is never queued while the count reads 0.  The fill is over when the
// count is 0 and the ring is empty: no group then holds work, and a group
// that starts later (its compute unit was busy with other kernels) finds
// none and leaves.  So the groups need not all be resident at once.
__device__ void fill_manager(const FillArgs& A, GroupPool* G, long long t0,
                             unsigned long long& polls) {
  const int ln = __lane_id();
  if (ln == 0) {
    afaddi(A.active, 1);
    mem_order();
  }
  bool counted = true;  // the group is in A.active
  int backoff = 0;
  unsigned long long loops = 0;
  long long nextShareCheck = 0, nextQueueCheck = 0;
  bool othersIdle = false;
  bool seedsLeft = true;  // the batch's seed list is handed out front to back, once
  Task t = make_task(0, 0, 0, 0, 0, 0, 1, 0, 0);
  while (true) {
    loops++;
    const long long now = (long long)__builtin_amdgcn_s_memrealtime();
    if (now - t0 > A.timeLimitTicks) {
      if (ln == 0) atomicExch(A.abortFlag, 1);
      break;
    }
    int pend = 0, cnt = 0;
    if (ln == 0) {
      pend = lds_ld(&G->pending);
      cnt = lds_ld(&G->count);
    }
    pend = uni(pend);
    cnt = uni(cnt);
    const int idle = kWorkers - (pend - cnt);  // workers without work
    if (pend == 0) {
      // the group is idle: seed tasks, else the global ring
      int k = seedsLeft ? claim_seeds(A, t, kFeedMax) : 0;
      seedsLeft = k > 0;
      if (k > 0) {
        pool_add_all(A, G, t, 0, k, t0 + A.timeLimitTicks);
        continue;
      }
      if (counted) {
        if (ln == 0) {
          mem_order();
          afaddi(A.active, -1);
        }
        counted = false;
      }
      polls++;
      if (backoff < A.pollBackoff) backoff++;
      for (int s = 0; s < (1 << backoff); s++) __builtin_amdgcn_s_sleep(16);
      int abortNow = 0, act = 1;
      unsigned long long qh = 0, qt = 0;
      if (ln == 0) {
        if ((polls & 3) == 0) {
          act = aldi(A.active);
          abortNow = aldi(A.abortFlag);
          mem_order();
        }
        qh = ald(A.qhead);
        qt = ald(A.qtail);
      }
      if (uni(abortNow)) break;
      act = uni(act);
      qh = uni64(qh);
      qt = uni64(qt);
      if (qt > qh) {
        // rejoin before taking tasks
        if (ln == 0) {
          afaddi(A.active, 1);
          mem_order();
        }
        counted = true;
        k = q_pop_lanes(A, t, kFeedMax);
        if (k > 0) {
          pool_add_all(A, G, t, 0, k, t0 + A.timeLimitTicks);
          backoff = 0;
        } else {
          if (ln == 0) afaddi(A.active, -1);
          counted = false;
        }
      } else if (act <= 0) {
        break;  // no group holds work and nothing is queued
      }
      continue;
    }
    backoff = 0;
    if (idle > cnt) {
      // workers are waiting: seed tasks, else (at most every ~2 us) the ring
      const int want = idle - cnt + 2 < kFeedMax ? idle - cnt + 2 : kFeedMax;
      int k = seedsLeft ? claim_seeds(A, t, want) : 0;
      seedsLeft = k > 0;
      if (k == 0 && now >= nextQueueCheck) {
        nextQueueCheck = now + 200;
        k = q_pop_lanes(A, t, want);
      }
      if (k > 0) {
        pool_add_all(A, G, t, 0, k, t0 + A.timeLimitTicks);
        continue;
      }
    } else {
      // while some other group is idle (the global count is read at most
      // every ~5 us), the pool's tasks go on to the global ring, and the
      // workers keep the pool topped up
      if (now >= nextShareCheck) {
        nextShareCheck = now + 500;
        int act = 0;
        if (ln == 0) act = aldi(A.active);
        othersIdle = uni(act) < A.nGroups;
        if (ln == 0) lds_st(&G->hungry, othersIdle ? 1 : 0);
      }
      if (othersIdle && cnt >= 2) {
        const int n = cnt / 2 < kDonateMax ? cnt / 2 : kDonateMax;
        const int k = pool_remove_lanes(G, t, n);
        const int q = q_push_lanes(A, t, k);
        if (q < k) pool_add_all(A, G, t, q, k, t0 + A.timeLimitTicks);  // ring full: keep the rest
        if (q > 0) continue;
      }
    }
    __builtin_amdgcn_s_sleep(4);
  }
  if (ln == 0) lds_st(&G->done, 1);
}

// Worker waves take work from their own ring, then from the group's pool;
// a worker that finds neither is idle and waits on the pool until the
// manager ends the fill.
__global__ void __launch_bounds__(kFillThreads) fill_kernel(FillArgs A) {
  __shared__ Task rings[kWorkers][kLocalCap];
  __shared__ GroupPool pool;
  const int ln = __lane_id();
  const int wave = threadIdx.x / 64;
  const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
  GroupPool* G = &pool;
  if (threadIdx.x == 0) {
    G->lock = 0;
    G->pending = 0;
    G->head = 0;
    G->count = 0;
    G->done = 0;
    G->hungry = 0;
  }
  __syncthreads();
  unsigned long long nTasks = 0, busyTicks = 0, idlePolls = 0;
  if (wave == 0) {
    fill_manager(A, G, t0, idlePolls);
    if (ln == 0) atomicAdd(&A.stats[6], idlePolls);
    return;
  }
  TaskState S;
  S.A = &A;
  S.G = G;
  S.curB = -1;
  S.evals = 0;
  S.nIns = 0;
  S.nDup = 0;
  S.insTicks = 0;
  S.nScout = 0;
  S.nScoutRuns = 0;
  S.accLabel = -1;
  S.accSum = 0;
  S.accCnt = 0;
  S.ring = rings[wave - 1];
  S.lhead = 0;
  S.lcount = 0;
  bool have = false, busy = false;
  while (true) {
    if (!have && S.lcount > 0) {
      S.t = ring_pop(S);
      have = true;
    }
    if (!have && busy) {
      // own work done: the pool's oldest task, or idle
      flush_acc(S);
      have = pool_take(G, S.t, true, A.poolNewest != 0);
      busy = have;
    }
    if (!have) {
      bool over = false;
      while (true) {
        idlePolls++;
        __builtin_amdgcn_s_sleep(24);
        int d = 0, c = 0;
        if (ln == 0) {
          d = lds_ld(&G->done);
          c = lds_ld(&G->count);
        }
        if (uni(d)) { over = true; break; }
        if (uni(c) > 0 && pool_take(G, S.t, false, A.poolNewest != 0)) break;
        if ((idlePolls & 255) == 0 &&
            (long long)__builtin_amdgcn_s_memrealtime() - t0 > A.timeLimitTicks + 100000000ll) {
          // the manager ends the fill at the time limit; this bound only
          // guards against a manager that never does
          over = true;
          break;
        }
      }
      if (over) break;
      have = true;
      busy = true;
    }
    nTasks++;
    const unsigned long long tb0 = __builtin_amdgcn_s_memrealtime();
    process_task(S);
    busyTicks += __builtin_amdgcn_s_memrealtime() - tb0;
    if (S.haveCont) {
      S.t = S.cont;
    } else {
      have = false;
    }
    if ((nTasks & 3) == 0 && S.lcount > 0) {
      // hand waiting tasks to idle workers of the group, and while other
      // groups are idle keep the pool topped up for them
      int pend = 0, cnt = 0, hungry = 0;
      if (ln == 0) {
        pend = lds_ld(&G->pending);
        cnt = lds_ld(&G->count);
        hungry = lds_ld(&G->hungry);
      }
      cnt = uni(cnt);
      const int idle = kWorkers - (uni(pend) - cnt);
      if (idle > cnt || (uni(hungry) && cnt < kPoolShare)) {
        ring_donate(S);
        if (S.lcount > 2) ring_donate(S);
      }
    }
    if ((nTasks & 255) == 0) {
      const long long now = (long long)__builtin_amdgcn_s_memrealtime();
      if (now - t0 > A.timeLimitTicks) {
        if (ln == 0) atomicExch(A.abortFlag, 1);
        break;
      }
      int ab = 0;
      if (ln == 0) ab = aldi(A.abortFlag);
      if (uni(ab)) break;
    }
  }
  flush_acc(S);
  if (ln == 0) {
    atomicAdd(&A.stats[0], nTasks);
    atomicAdd(&A.stats[1], S.evals);
    atomicAdd(&A.stats[2], S.nIns);
    atomicAdd(&A.stats[3], S.nDup);
    atomicAdd(&A.stats[4], busyTicks);
    atomicAdd(&A.stats[5], S.insTicks);
    atomicAdd(&A.stats[6], idlePolls);
    atomicAdd(&A.stats[7], 1ull);
    atomicAdd(&A.stats[8], S.nScout);
    atomicAdd(&A.stats[9], S.nScoutRuns);
  }
}

// Workgroups of fill_kernel that can be resident on the device at once (a
// larger launch is correct, its surplus groups run after the others).
int fill_resident_workgroups() {
  int perCu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, fill_kernel, kFillThreads, 0) !=
      hipSuccess)
    return 0;
  return perCu * cus;
}

int launch_fill(const FillArgs& a, int nWorkgroups, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(nWorkgroups), dim3(kFillThreads), 0, s, a);
  return (int)hipGetLastError();
}

__global__ void clear_u64_kernel(unsigned long long* p, unsigned long long n, unsigned long long v) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = v;
}

int launch_clear_u64(unsigned long long* p, unsigned long long n, unsigned long long v,
                     hipStream_t s) {
  if (n == 0) return 0;
  unsigned long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(clear_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return (int)hipGetLastError();
}

// Resets the run table (word0 = 0, word1 = ~0), labels, queue slots and
// counters before a fill launch.  A plain kernel (no memset nodes), so that
// the sequence can be captured in a graph.
__global__ void fill_reset_kernel(FillArgs A) {
  const unsigned long long tid = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long i = tid; i < A.htabEntries; i += stride) {
    A.htab[2 * i] = 0;
    A.htab[2 * i + 1] = kEmpty1;
  }
  for (unsigned long long i = tid; i < (unsigned long long)A.labelCap; i += stride) {
    A.lparent[i] = (int)i;
    A.lpoint[i] = i < (unsigned long long)A.nBatch ? (int)i : -1;
    A.lcount[i] = 0;
    A.lsum[2 * i] = 0;
    A.lsum[2 * i + 1] = 0;
  }
  for (unsigned long long i = tid; i < A.qcap; i += stride) A.qseq[i] = i;
  if (tid == 0) {
    *A.qhead = 0;
    *A.qtail = 0;
    *A.active = 0;
    *A.seedNext = 0;
    *A.labelNext = A.nBatch;
    *A.abortFlag = 0;
    for (int k = 0; k < 10; k++) A.stats[k] = 0;
  }
}

int launch_fill_reset(const FillArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(fill_reset_kernel, dim3(2048), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

// Adds every label whose class holds the reference-seed label of its point
// to that point's totals.
__global__ void resolve_kernel(FillArgs A, int nLabels, unsigned long long* ptCount,
                               unsigned long long* ptSum) {
  const int L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= nLabels) return;
  const int b = A.lpoint[L];
  if (b < 0 || b >= A.nBatch) return;
  if (A.lcount[L] == 0) return;
  const int r = find_root(A.lparent, L);
  const int rb = find_root(A.lparent, b);
  if (r != rb) return;
  atomicAdd(&ptCount[b], A.lcount[L]);
  // the label's split total into the point's, word by word (the point's
  // total is hi * 2^32 + lo like the labels'; its low word may carry past
  // 2^64 only after 2^32 labels)
  const unsigned long long lo = A.lsum[2 * (size_t)L], hi = A.lsum[2 * (size_t)L + 1];
  atomicAdd(&ptSum[2 * (size_t)b], lo);
  if (hi) atomicAdd(&ptSum[2 * (size_t)b + 1], hi);
}

int launch_resolve(const FillArgs& a, int nLabels, unsigned long long* ptCount,
                   unsigned long long* ptSum,
                   hipStream_t s) {
  if (nLabels <= 0) return 0;
  hipLaunchKernelGGL(resolve_kernel, dim3((nLabels + 255) / 256), dim3(256), 0, s, a, nLabels,
                     ptCount, ptSum);
  return (int)hipGetLastError();
}

}  // namespace tl
