"""Shared fixtures.  GPU tests carry the ``gpu`` marker; everything else runs
on the CPU (the oracle, the golden fixtures, the C ABI's host side)."""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference
    if not Reference.available():
        pytest.skip("reference build oracle/_ref is absent (needs /root/reference at build time)")
    return Reference()


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def curve_from_params(p):
    """Curve from a fixture row a, b, th, m2, m3, rho, ppr, u, sx, sy, ex, ey, n."""
    from triplelens_amd.workloads import Curve
    return Curve(*[float(v) for v in p[:6]], int(p[6]), float(p[7]), *[float(v) for v in p[8:12]],
                 int(p[12]))


def pixel_seeds(grid, xy):
    """Image-plane seeds to pixels as xToNx/yToNy do (lcirs.cc:367-378)."""
    N, L, blx, bly = grid[0], grid[1], grid[3], grid[4]
    px = np.trunc((xy[..., 0] - blx) * N / L + 0.5).astype(np.int64)
    py = np.trunc((xy[..., 1] - bly) * N / L + 0.5).astype(np.int64)
    return px, py


def plain_run(g):
    """The plain reference run's (rays, amp) of a fixture: the zero-filled
    run's with the points where the plain run differs replaced."""
    rays = g["rays"].astype(np.uint64)
    amp = g["amp"].copy()
    idx = g["plain_idx"]
    rays[idx] = g["rays_plain_d"]
    amp[idx] = g["amp_plain_d"]
    return rays, amp


def slot_trace(g, k=0):
    """The plain reference run's slot values (one complex per getImages call)
    of curve k of a fixture, decoded from their run-length encoding; None for
    a curve whose plain run equals its zero-filled run (no trace stored)."""
    curves = list(g["slot_curves"])
    if k not in curves:
        return None
    j = curves.index(k)
    so = g["slot_off"]
    start, val = g["slot_start"][so[j]:so[j + 1]], g["slot_val"][so[j]:so[j + 1]]
    n = int(g["slot_calls"][j])
    return np.repeat(val, np.diff(np.append(start, n)))


# The reference's allocator settings under which its slot read gives 0
# (zero-filled allocations): the fixtures' parity runs used them.
ZEROFILL_ENV = {"GLIBC_TUNABLES": "glibc.malloc.tcache_count=0", "MALLOC_PERTURB_": "255"}


def reference_lc_zero_filled(c):
    """The reference build's light curve (rays, amp) of workloads.Curve c,
    run in a child process with zero-filled allocations."""
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from oracle.oracle import Reference\n"
        "p = [float(v) for v in sys.argv[2:]]\n"
        "amp, rays = Reference().lc_irs(*p[:6], int(p[6]), int(p[7]), p[8], *p[9:13])\n"
        "np.savez(sys.argv[1], amp=amp, rays=rays)\n")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "r.npz")
        args = [*c.lens(), c.rho, c.lc_len, c.ppr, c.u, *c.traj()]
        subprocess.run([sys.executable, "-c", code, out, *map(repr, args)],
                       env=dict(os.environ, **ZEROFILL_ENV), check=True)
        z = np.load(out)
        return z["rays"], z["amp"]
