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


def test_seed_union_never_misses_an_image_of_the_single_chain():
    """The surrogate path fills its shot positions from the union of four
    split seed searches (different piece boundaries).  A split search starts
    some positions' roots afresh, and near a caustic a fresh start can
    converge without a whole image's seed; the union guards against that.
    On the first proposals of four Metropolis chains from the C2 geometry
    (2048 positions each), no shot position has fewer rays than the
    reference's single chain gives it, and every curve is within the 1e-3
    bound of the single chain's."""
    from triplelens_amd import engine as E, metropolis as M, surrogate as S, workloads as W

    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, 200).cc_ca()[1]

    truth = dataclasses.replace(W.c2(seed=2, ca_fn=ca_fn), n_points=2048)
    sur = S.Surrogate.load(MODEL)
    n_shot = 0
    for k, start in enumerate(M.starts_near(truth, 4)):
        p = M.propose(start, np.random.default_rng([0, k, 1]), 0.01)
        A_ref, rays_ref, _ = M._new(p).lc_irs(*p.traj())
        lc = M._new(p)
        A, info = sur.light_curve(lc, p, np.random.default_rng(7))
        shot = np.nonzero(info["fallback"] | info["audited"])[0]
        n_shot += len(shot)
        A_shot, rays_shot, _ = lc.fill_positions(shot, _union_seeds(sur, lc, p, shot))
        assert np.all(rays_shot >= rays_ref[shot]), (k, shot[rays_shot < rays_ref[shot]])
        assert np.max(np.abs(A / A_ref - 1.0)) <= 1e-3, k
    assert n_shot > 0


def _union_seeds(sur, lc, p, idx):
    """The merged seeds the surrogate path filled positions idx with."""
    from triplelens_amd import surrogate as S
    ns, pix = lc.seeds()
    sd = (ns[idx], pix[idx])
    for cl in sur._extra_clones(lc, p):
        ns2, pix2 = cl.seeds()
        sd = S.union_seeds(sd[0], sd[1], ns2[idx], pix2[idx])
    return sd
