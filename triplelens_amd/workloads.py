"""Seeded synthetic workloads: the light-curve configurations the engine is
measured and tested on.

Every geometry is in the reference's lens frame (src/LensCore/lens.cc:35-40):
z1 = 0, z2 = a on the real axis, z3 = b*exp(i*th), masses m1 = 1-m2-m3, m2,
m3.  A light curve of ``n`` points is a ``lcLength = n-1`` LightCurveIRS run
(the reference evaluates lcLength+1 positions, src/LightCurves/lcirs.cc:117).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, asdict

import numpy as np


@dataclass
class Curve:
    a: float
    b: float
    th: float
    m2: float
    m3: float
    rho: float
    ppr: int
    u: float          # linear limb-darkening parameter ("linear", u)
    sx: float
    sy: float
    ex: float
    ey: float
    n_points: int     # lcLength + 1

    @property
    def lc_len(self) -> int:
        return self.n_points - 1

    def lens(self):
        return (self.a, self.b, self.th, self.m2, self.m3)

    def traj(self):
        return (self.sx, self.sy, self.ex, self.ey)

    def asdict(self):
        return asdict(self)


def c1() -> Curve:
    """Triple lens z1=(-0.5,0), z2=(0.5,0), z3=(0,0.8), masses (0.8,0.1,0.1),
    uniform disk rho=1e-3, 200 points from (-1,-0.1) to (1,0.1), pixel rho/50.
    In the lens frame (z1 at the origin) the trajectory is shifted by +0.5."""
    a = 1.0
    b = math.hypot(0.5, 0.8)
    th = math.atan2(0.8, 0.5)
    return Curve(a, b, th, 0.1, 0.1, 1e-3, 50, 0.0, -0.5, -0.1, 1.5, 0.1, 200)


def binary_c1() -> Curve:
    """C1 as a binary lens: the lenses z1=(-0.5,0), z2=(0.5,0) with masses
    (0.8, 0.2) and m3 = 0 (the reference's binary lens); C1's source, pixel
    and trajectory."""
    c = c1()
    return Curve(c.a, c.b, c.th, 0.2, 0.0, c.rho, c.ppr, 0.0, c.sx, c.sy, c.ex, c.ey, c.n_points)


def binary_planet(ca_fn) -> Curve:
    """A planetary binary lens: mass ratio 1e-3 at separation 1.1 (m3 = 0),
    linear limb darkening u = 0.5, rho = 5e-4, pixel rho/50, 200 positions
    on a chord of length 0.008 through the central caustic (the caustic
    sample nearest the primary).  ``ca_fn(a, b, th, m2, m3)`` returns the
    caustic samples."""
    a, b, th, m2, m3 = 1.1, 1.0, 1.0, 1e-3 / (1.0 + 1e-3), 0.0
    p = central_caustic_point(a, b, th, m2, m3, ca_fn(a, b, th, m2, m3))
    d = complex(0.004, 0.0015) / abs(complex(0.004, 0.0015)) * 0.004
    s, e = p - d, p + d
    return Curve(a, b, th, m2, m3, 5e-4, 50, 0.5, s.real, s.imag, e.real, e.imag, 200)


def _logu(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def central_caustic_point(a, b, th, m2, m3, ca):
    """Caustic sample nearest the primary z1 = 0 (ca: 6 x len complex)."""
    flat = ca.reshape(-1)
    return complex(flat[np.argmin(np.abs(flat))])


def c2(seed: int = 2, n_points: int = 4096, ppr: int = 200, ca_fn=None) -> Curve:
    """Separations log-uniform [0.5, 2], m2, m3 log-uniform [1e-3, 0.3],
    linear limb darkening u = 0.5, rho = 5e-4, pixel rho/200, a straight
    trajectory of length 2 through the central caustic (the caustic sample
    nearest z1) in a uniformly random direction.  ``ca_fn(a, b, th, m2, m3)``
    returns the 6 x 500 caustic samples (from the engine or the oracle)."""
    rng = np.random.default_rng(seed)
    a = _logu(rng, 0.5, 2.0)
    b = _logu(rng, 0.5, 2.0)
    th = float(rng.uniform(0.0, 2 * math.pi))
    m2 = _logu(rng, 1e-3, 0.3)
    m3 = _logu(rng, 1e-3, 0.3)
    phi = float(rng.uniform(0.0, 2 * math.pi))
    if ca_fn is None:
        raise ValueError("c2 needs ca_fn to locate the central caustic")
    p = central_caustic_point(a, b, th, m2, m3, ca_fn(a, b, th, m2, m3))
    d = complex(math.cos(phi), math.sin(phi))
    s, e = p - d, p + d
    return Curve(a, b, th, m2, m3, 5e-4, ppr, 0.5, s.real, s.imag, e.real, e.imag, n_points)


def c3(seed: int = 3, n_curves: int = 64, ppr: int = 100, n_range=(256, 8192)) -> list[Curve]:
    """64 curves: separations log-uniform [0.3, 3], mass ratios q2, q3
    log-uniform [1e-4, 1] (m1 = 1/(1+q2+q3), m_k = q_k m1), rho log-uniform
    [1e-4, 1e-2], u uniform [0, 0.8], direction uniform [0, 2pi), impact
    parameter uniform [-0.5, 0.5] about the primary, length 2, lengths
    log-uniform in n_range points, pixel rho/100."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_curves):
        a = _logu(rng, 0.3, 3.0)
        b = _logu(rng, 0.3, 3.0)
        th = float(rng.uniform(0.0, 2 * math.pi))
        q2 = _logu(rng, 1e-4, 1.0)
        q3 = _logu(rng, 1e-4, 1.0)
        m1 = 1.0 / (1.0 + q2 + q3)
        m2, m3 = q2 * m1, q3 * m1
        rho = _logu(rng, 1e-4, 1e-2)
        u = float(rng.uniform(0.0, 0.8))
        phi = float(rng.uniform(0.0, 2 * math.pi))
        imp = float(rng.uniform(-0.5, 0.5))
        n = int(round(_logu(rng, n_range[0], n_range[1])))
        d = complex(math.cos(phi), math.sin(phi))
        c = imp * d * 1j
        s, e = c - d, c + d
        out.append(Curve(a, b, th, m2, m3, rho, ppr, u, s.real, s.imag, e.real, e.imag, n))
    return out
