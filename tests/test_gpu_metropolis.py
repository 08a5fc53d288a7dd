"""The Metropolis fitting driver on the GPU: chains of surrogate + IRS
light curves against chains of plain ray-shot curves from the same starts."""
from __future__ import annotations

import dataclasses
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

MODEL = os.path.join(ROOT, "triplelens_amd", "data", "surrogate_c2.npz")


def test_chains_with_surrogate_and_with_ray_shooting():
    from triplelens_amd import engine as E, metropolis as M, surrogate as S, workloads as W

    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, 200).cc_ca()[1]

    truth = dataclasses.replace(W.c2(seed=2, ca_fn=ca_fn), n_points=256)
    obs, err = M.observed(truth)
    starts = M.starts_near(truth, 2)
    sur = S.Surrogate.load(MODEL)
    rs = M.run(starts, obs, err, 3, "surrogate", sur)
    ri = M.run(starts, obs, err, 3, "irs")
    for r in (rs, ri):
        assert r["points"] == 2 * 3 * 256
        for ch in r["chains"]:
            assert ch.proposals == 3 and 0 <= ch.accepted <= 3
            assert len(ch.chi2_seq) == 3
            assert np.all(np.isfinite(ch.chi2_seq)) and np.all(np.isfinite(ch.amp))
            assert ch.chi2 == pytest.approx(M.chi2(ch.amp, obs, err), rel=1e-12)
    # both modes draw the same first proposals; their curves agree within
    # the surrogate's bound
    assert sorted(rs["first"]) == sorted(ri["first"]) == [0, 1]
    for k in rs["first"]:
        assert np.max(np.abs(rs["first"][k] / ri["first"][k] - 1.0)) <= 1e-3
    assert rs["fallback"] + rs["audited"] > 0


def test_seed_union_recovers_an_image_one_split_search_misses():
    """The first proposal of the first Metropolis chain from the C2 geometry
    has a position (515 of 2048, its source disk across the caustic) where
    one split seed search misses the seed of a whole image: filled from that
    search's seeds it shows 2.58 against the single chain's 19.84.  The
    surrogate path fills its shot positions from the union of four split
    searches and matches the chain there."""
    from triplelens_amd import engine as E, metropolis as M, surrogate as S, workloads as W

    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, 200).cc_ca()[1]

    truth = dataclasses.replace(W.c2(seed=2, ca_fn=ca_fn), n_points=2048)
    p = M.propose(M.starts_near(truth, 8)[0], np.random.default_rng([0, 0, 1]), 0.01)
    sur = S.Surrogate.load(MODEL)
    ref = M._new(p)
    A_ref, rays_ref, _ = ref.lc_irs(*p.traj())
    one = M._new(p)
    one.set_seed_split(sur.seed_parts(p.n_points))
    one.find_seeds(*p.traj())
    A_one, rays_one, _ = one.fill_positions(np.array([515]))
    assert rays_one[0] < rays_ref[515] and A_one[0] < 0.5 * A_ref[515]
    A, info = sur.light_curve(M._new(p), p, np.random.default_rng(7))
    assert info["fallback"][515]
    assert abs(A[515] / A_ref[515] - 1.0) < 1e-12
    assert np.max(np.abs(A / A_ref - 1.0)) <= 1e-3
