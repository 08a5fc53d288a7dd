"""Shared fixtures.  GPU tests carry the ``gpu`` marker; everything else runs
on the CPU (the oracle, the golden fixtures, the C ABI's host side)."""
from __future__ import annotations

import os
import sys

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
