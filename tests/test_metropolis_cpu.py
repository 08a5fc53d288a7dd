"""Host logic of the Metropolis fitting driver (no GPU work)."""
from __future__ import annotations

import numpy as np

from triplelens_amd import metropolis as M
from triplelens_amd import workloads as W


def test_proposals_step_one_percent_and_stay_in_the_domain():
    c = W.c1()
    rng = np.random.default_rng(3)
    steps = []
    for _ in range(2000):
        p = M.propose(c, rng, 0.01)
        assert p is not None and p.traj() == c.traj() and p.rho == c.rho
        steps.append([p.a / c.a - 1, p.b / c.b - 1, p.th - c.th, p.m2 / c.m2 - 1,
                      p.m3 / c.m3 - 1])
    sd = np.std(np.array(steps), axis=0)
    assert np.all(np.abs(sd / 0.01 - 1) < 0.1)
    # masses that would leave the domain are rejected
    edge = W.Curve(1.0, 1.0, 1.0, 0.5, 0.4999, 1e-3, 50, 0.0, 0, 0, 1, 0, 10)
    outs = [M.propose(edge, rng, 0.05) for _ in range(200)]
    assert any(o is None for o in outs)
    assert all(o.m2 + o.m3 < 1 for o in outs if o is not None)


def test_chi2_and_starts():
    obs = np.array([1.0, 2.0, 4.0])
    err = np.array([0.1, 0.2, 0.4])
    assert abs(M.chi2(np.array([1.1, 2.0, 3.6]), obs, err) - 2.0) < 1e-12
    starts = M.starts_near(W.c1(), 5)
    assert len(starts) == 5 and len({s.a for s in starts}) == 5
