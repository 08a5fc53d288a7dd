"""Metropolis fitting of an extended-source light curve: independent chains
of lens geometries, each proposal evaluated as a whole light curve.

The acceptance rule is the reference's point-source fitting example
(Examples/run_point_metropolis.py:117-149): a proposal with a lower chi^2 is
taken, any other with probability exp(chi^2 - chi^2_trial).  Here the
proposal steps are fixed Gaussian steps of `step` (1%) of each lens
parameter's value (a, b, m2, m3) and `step` radians of th, the trajectory,
source radius and limb darkening stay fixed, and the light curve is the
extended-source curve:

- mode "surrogate": Surrogate.light_curves (the network, with ray shooting
  of the positions its gate rejects and of a 1% audit), the chains'
  proposals of one step sharing each engine run;
- mode "irs": plain ray shooting of every position, the chains' proposals of
  one step in one batched engine run (lcirs_batch_irs).

The observed curve is the exact curve of a reference geometry with 1%
Gaussian noise; its errors are 1% of the exact magnification.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace

import numpy as np

from . import engine as E
from .workloads import Curve


@dataclass
class Chain:
    curve: Curve                   # current state (lens parameters)
    amp: np.ndarray                # its light curve
    chi2: float
    rng_prop: np.random.Generator  # proposal draws
    rng_acc: np.random.Generator   # acceptance draws
    accepted: int = 0
    proposals: int = 0
    chi2_seq: list = field(default_factory=list)


def propose(c: Curve, rng: np.random.Generator, step: float = 0.01) -> Curve | None:
    """A Gaussian step in the lens parameters; None when the masses or
    separations leave their domain (the proposal is then rejected)."""
    z = rng.standard_normal(5)
    a = c.a * (1.0 + step * z[0])
    b = c.b * (1.0 + step * z[1])
    th = c.th + step * z[2]
    m2 = c.m2 * (1.0 + step * z[3])
    m3 = c.m3 * (1.0 + step * z[4])
    if a <= 0 or b <= 0 or m2 <= 0 or m3 <= 0 or m2 + m3 >= 1:
        return None
    return replace(c, a=a, b=b, th=th, m2=m2, m3=m3)


def chi2(amp: np.ndarray, obs: np.ndarray, err: np.ndarray) -> float:
    return float(np.sum(((obs - amp) / err) ** 2))


def _new(c: Curve):
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening("linear", c.u)
    return lc


def _new_many(cs: list[Curve]):
    """Objects of many geometries, their critical curves computed together
    on the GPU (tl_ccc_batch) instead of one by one on the host."""
    if not cs:
        return []
    cc, ca = E.ccc_batch(np.array([c.lens() for c in cs]))
    out = []
    for c, k1, k2 in zip(cs, cc, ca):
        lc = E.LightCurve.with_critical_curves(*c.lens(), c.rho, c.lc_len, c.ppr, k1, k2)
        lc.set_limb_darkening("linear", c.u)
        out.append(lc)
    return out


def exact_curves(curves: list[Curve]) -> list[np.ndarray]:
    """Plain ray-shot light curves of many geometries in one engine run."""
    lcs = _new_many(curves)
    res = E.batch_irs(lcs, np.array([c.traj() for c in curves]))
    out = []
    for amp, _rays, flags in res:
        bad = int(np.bitwise_or.reduce(flags)) & ~E.FLAGS_INFO
        if bad:
            raise RuntimeError(f"ray shooting reported status flags {bad:#x}")
        out.append(amp)
    return out


def observed(truth: Curve, noise: float = 0.01, seed: int = 0):
    """Synthetic observed curve of geometry `truth` and its errors."""
    amp = exact_curves([truth])[0]
    rng = np.random.default_rng(seed)
    return amp * (1.0 + noise * rng.standard_normal(len(amp))), noise * amp


def starts_near(truth: Curve, n: int, spread: float = 0.02, seed: int = 0) -> list[Curve]:
    """n chain starts: Gaussian steps of `spread` from the true geometry."""
    rng = np.random.default_rng([seed, 99])
    out = []
    while len(out) < n:
        c = propose(truth, rng, spread)
        if c is not None:
            out.append(c)
    return out


def run(starts: list[Curve], obs, err, n_props: int, mode: str = "surrogate", sur=None,
        step: float = 0.01, seed: int = 0, progress=None) -> dict:
    """Runs len(starts) chains for n_props proposals each.  Returns the
    chains, the wall time of the proposals, the light-curve points
    evaluated, (surrogate mode) the fallback and audit counts, and the light
    curves of the first proposals by chain (both modes draw the same first
    proposals from the same starts).  progress(step, points, seconds), if
    given, is called every 16 steps."""
    if mode not in ("surrogate", "irs"):
        raise ValueError(f"unknown mode {mode}")
    if mode == "surrogate" and sur is None:
        raise ValueError("surrogate mode needs a Surrogate")
    audit_rng = np.random.default_rng(seed + 7)

    def curves_of(cs):
        if mode == "irs":
            return exact_curves(cs)
        return [a for a, _ in sur.light_curves(_new_many(cs), cs, audit_rng)]

    init = curves_of(starts)
    chains = [Chain(c, a, chi2(a, obs, err), np.random.default_rng([seed, k, 1]),
                    np.random.default_rng([seed, k, 2]))
              for k, (c, a) in enumerate(zip(starts, init))]
    n_fb = n_aud = points = 0
    first = {}
    t0 = time.perf_counter()
    for step_i in range(n_props):
        props = [propose(ch.curve, ch.rng_prop, step) for ch in chains]
        live = [k for k, p in enumerate(props) if p is not None]
        if mode == "irs":
            amps = dict(zip(live, exact_curves([props[k] for k in live]))) if live else {}
        else:
            # the step's proposals share each engine run (seed searches,
            # network, fills of the shot positions)
            amps = {}
            res = sur.light_curves(_new_many([props[k] for k in live]), [props[k] for k in live],
                                   audit_rng) if live else []
            for k, (amp, info) in zip(live, res):
                amps[k] = amp
                n_fb += int(info["fallback"].sum())
                n_aud += int(info["audited"].sum())
        if step_i == 0:
            first = dict(amps)
        for k, ch in enumerate(chains):
            ch.proposals += 1
            u = ch.rng_acc.random()
            if k not in amps:
                continue
            points += len(amps[k])
            x = chi2(amps[k], obs, err)
            # Examples/run_point_metropolis.py:117-149
            if x < ch.chi2 or u < math.exp(min(0.0, ch.chi2 - x)):
                ch.curve, ch.amp, ch.chi2 = props[k], amps[k], x
                ch.accepted += 1
            ch.chi2_seq.append(ch.chi2)
        if progress is not None and (step_i + 1) % 16 == 0:
            progress(step_i + 1, points, time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    return {"chains": chains, "wall_s": wall, "points": points, "fallback": n_fb,
            "audited": n_aud, "first": first}
