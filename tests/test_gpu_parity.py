"""Parity of the HIP engine with the reference's fixtures and the oracle.

Every test here runs the engine on the GPU through the C ABI (liblcirs.so).
Ray counts must equal the reference's at every point: the fixtures hold the
reference run with zero-filled allocations, where the slot its chained solve
reads past the previous roots (initRoots[10], laguerre.cc:111-128) is 0 as in
the engine, and the plain run with the slot values it read, which the engine
reproduces when given them (lcirs_set_slot_trace).  Magnifications must be
within 1e-6 relative of the reference's (the scope's bound; the engine sums
brightness in fixed point, the reference in double in fill order) and within
1e-10 of the oracle's.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import curve_from_params, golden, pixel_seeds, plain_run, slot_trace

pytestmark = pytest.mark.gpu

AMP_RTOL = 1e-10


@pytest.fixture(scope="module")
def E():
    from triplelens_amd import engine
    engine.load()
    engine.set_option("time_limit_s", 60)
    return engine


def _new(E, c, lc_len=None, ppr=None):
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len if lc_len is None else lc_len,
                      c.ppr if ppr is None else ppr)
    lc.set_limb_darkening("linear", c.u)
    return lc


def _assert_no_status(E, flags):
    bad = int(np.bitwise_or.reduce(flags)) & ~E.FLAGS_INFO
    assert bad == 0, f"engine status flags {bad:#x}"


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


REF_RTOL = 1e-6


def _fixture_curves(g):
    return [curve_from_params(p) for p in np.atleast_2d(g["params"])]


def _check_against_fixture(E, g, results, plain=False):
    """Every curve's ray counts equal to the fixture's and magnifications
    within REF_RTOL, except at points flagged TL_FLAG_KICK_TABLE (a Laguerre
    iteration reached step 64, where the reference reads its fractional-step
    table past the end: its result there is undefined); no other status
    flags.  Returns the flagged points per curve."""
    all_r, all_a = plain_run(g) if plain else (g["rays"], g["amp"])
    off = g["off"] if "off" in g.files else np.array([0, len(all_r)])
    undefined = []
    for k, (amp, rays, flags) in enumerate(results):
        want_r, want_a = all_r[off[k]:off[k + 1]], all_a[off[k]:off[k + 1]]
        und = (flags & E.FLAG_KICK_TABLE) != 0
        d = np.nonzero((rays != want_r) & ~und)[0]
        assert len(d) == 0, (k, d[:10])
        ok = ~und
        assert _rel(amp[ok], want_a[ok]) < REF_RTOL, k
        _assert_no_status(E, flags)
        undefined.append(np.nonzero(und)[0])
    return undefined


# --------------------------------------------------------------------- C1
def test_c1_matches_reference_golden(E):
    """The 200-point workload through the reference-style class LC_irs."""
    from triplelens_amd.ctypes_classes import LC_irs
    g = golden("c1.npz")
    c = curve_from_params(g["params"])
    lc = LC_irs(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening(b"linear", c.u)
    lc.get_lc_irs(*c.traj())
    out = np.zeros(c.lc_len)
    lc.copy_lc(out)
    rays = lc.ray_counts()
    full = lc.lc_full()
    assert np.array_equal(rays, g["rays"]), np.where(rays != g["rays"])[0][:10]
    assert _rel(full, g["amp"]) < REF_RTOL
    assert np.array_equal(out, full[:c.lc_len])


def test_c1_with_the_plain_reference_run_slots(E):
    g = golden("c1.npz")
    c = curve_from_params(g["params"])
    lc = _new(E, c)
    lc.set_slot_trace(slot_trace(g))  # None when the two runs agree
    _check_against_fixture(E, g, [lc.lc_irs(*c.traj())], plain=True)


def test_c1_matches_oracle(E, oracle):
    from triplelens_amd import workloads as W
    c = W.c1()
    amp, rays, flags = _new(E, c).lc_irs(*c.traj())
    amp_o, rays_o, edge_o = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u).lc_irs(*c.traj())
    assert np.array_equal(rays, rays_o)
    assert _rel(amp, amp_o) < AMP_RTOL
    assert np.array_equal((flags & E.FLAG_EDGE) != 0, edge_o)
    _assert_no_status(E, flags)


def test_c1_seeds_match_golden(E):
    """Seed pixels of every point: the reference's seeds mapped to pixels."""
    g = golden("c1.npz")
    c = curve_from_params(g["params"])
    lc = _new(E, c)
    lc.lc_irs(*c.traj())
    ns, pix = lc.seeds()
    off, xy = g["seed_off"], g["seed_xy"]
    px, py = pixel_seeds(g["grid"], xy)
    for i in range(c.n_points):
        ref = set(zip(px[off[i]:off[i + 1]].tolist(), py[off[i]:off[i + 1]].tolist()))
        got = set(map(tuple, pix[i, :ns[i]].tolist()))
        assert ns[i] == off[i + 1] - off[i] and got == ref, i


def test_c1_default_limb_darkening(E, oracle):
    """Without set_limb_darkening the profile is A=0.6, B=0.4."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    amp, rays, _ = lc.lc_irs(*c.traj())
    amp_o, rays_o, _ = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr).lc_irs(*c.traj())
    assert np.array_equal(rays, rays_o)
    assert _rel(amp, amp_o) < AMP_RTOL


def test_c1_point_source_curve(E, oracle):
    """get_lc: point-source magnification at t = i/(lcLength-1)."""
    from triplelens_amd import workloads as W
    c = W.c1()
    got = _new(E, c).point_lc(*c.traj())
    t = np.arange(c.lc_len) / (c.lc_len - 1)
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    ref = oracle.point_amp(*c.lens(), pos)
    assert _rel(got, ref) < 1e-12


# --------------------------------------------------------------------- C2
@pytest.mark.parametrize("plain", [False, True], ids=["zero_slot", "plain_slots"])
def test_c2_full_resolution_matches_reference(E, plain):
    """The bench curve (4096 points, pixel rho/200): every ray count equal to
    the reference's, every magnification within 1e-6."""
    g = golden("c2_full.npz")
    c = _fixture_curves(g)[0]
    lc = _new(E, c)
    if plain:
        lc.set_slot_trace(slot_trace(g))
    _check_against_fixture(E, g, [lc.lc_irs(*c.traj())], plain)


def test_speculative_chain_equals_serial_chain(E):
    """The verified speculative chain (pieces in parallel, repair chains
    where a piece disagrees with its predecessor) gives the single chain's
    seeds and the reference's ray counts on the bench curve; so does the
    serial chain, and a repair depth too small to resolve the curve falls
    back to the serial chain with the same results."""
    g = golden("c2_full.npz")
    c = _fixture_curves(g)[0]
    runs = {}
    try:
        for name, opts in [("spec", {}), ("serial", {"spec_piece": 0}),
                           ("short_pieces", {"spec_piece": 16, "spec_overlap": 8,
                                             "spec_depth": 256}),
                           ("fallback", {"spec_depth": 1})]:
            for k, v in opts.items():
                E.set_option(k, v)
            lc = _new(E, c)
            res = lc.lc_irs(*c.traj())
            st = E.last_stats()
            runs[name] = (res, lc.seeds(), st)
            for k in opts:
                E.set_option(k, {"spec_piece": 32, "spec_overlap": 16, "spec_depth": 4096}[k])
    finally:
        for k, v in {"spec_piece": 32, "spec_overlap": 16, "spec_depth": 4096}.items():
            E.set_option(k, v)
    assert runs["spec"][2]["spec_pieces"] == 4096 // 32 and runs["spec"][2]["spec_fallback"] == 0
    assert runs["serial"][2]["spec_pieces"] == 0
    assert runs["short_pieces"][2]["spec_pieces"] == 4096 // 16
    assert runs["fallback"][2]["spec_fallback"] == 1
    ns0, pix0 = runs["serial"][1]
    for name, (res, (ns, pix), st) in runs.items():
        _check_against_fixture(E, g, [res])
        assert np.array_equal(ns, ns0), name
        assert np.array_equal(pix, pix0), name


def test_c2_window_full_resolution_golden(E):
    """A 65-point curve around the C2 magnification peak at pixel rho/200."""
    g = golden("c2_window.npz")
    c = curve_from_params(g["params"])
    _check_against_fixture(E, g, [_new(E, c).lc_irs(*c.traj())])


def test_c2_ppr20_matches_oracle_and_golden(E, oracle):
    """All 4096 C2 points at pixel rho/20: identical to the reference and
    the oracle."""
    g = golden("c2_ppr20.npz")
    c = curve_from_params(g["params"])
    res = _new(E, c).lc_irs(*c.traj())
    _check_against_fixture(E, g, [res])
    amp_o, rays_o, _ = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u).lc_irs(*c.traj())
    assert np.array_equal(res[1], rays_o)
    assert _rel(res[0], amp_o) < AMP_RTOL


def test_c3_batch_matches_oracle(E, oracle):
    """16 curves of C3 at pixel rho/20 in one batched engine run."""
    g = golden("c3_ppr20.npz")
    curves = _fixture_curves(g)
    lcs = [_new(E, c) for c in curves]
    res = E.batch_irs(lcs, np.array([c.traj() for c in curves]))
    _check_against_fixture(E, g, res)
    for k, (c, (amp, rays, flags)) in enumerate(zip(curves[:4], res)):
        amp_o, rays_o, _ = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u).lc_irs(*c.traj())
        assert np.array_equal(rays, rays_o), (k, np.where(rays != rays_o)[0][:10])
        assert _rel(amp, amp_o) < AMP_RTOL, k


@pytest.mark.parametrize("plain", [False, True], ids=["zero_slot", "plain_slots"])
def test_c3_full_batch_matches_reference(E, plain):
    """All 64 C3 curves at their pixel rho/100 (amoebae up to ~10^6 pixels,
    curves of 256-8192 points) in one batched engine run: every ray count
    equal to the reference's, magnifications within 1e-6."""
    g = golden("c3_full.npz")
    curves = _fixture_curves(g)
    lcs = [_new(E, c) for c in curves]
    if plain:
        for k, lc in enumerate(lcs):
            lc.set_slot_trace(slot_trace(g, k))
    res = E.batch_irs(lcs, np.array([c.traj() for c in curves]))
    undefined = _check_against_fixture(E, g, res, plain)
    assert sum(len(u) for u in undefined) <= 8


def test_engine_flags_the_oracles_undefined_points(E, oracle):
    """TL_FLAG_KICK_TABLE marks exactly the points where the oracle's
    Laguerre iterations reach step 64 (all 64 C3 curves at rho/100)."""
    from concurrent.futures import ThreadPoolExecutor
    g = golden("c3_full.npz")
    curves = _fixture_curves(g)
    lcs = [_new(E, c) for c in curves]
    E.batch_find_seeds(lcs, np.array([c.traj() for c in curves]))

    def oracle_points(c):
        lo = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u)
        return np.nonzero(lo.frac_oob(*c.traj()))[0]

    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(oracle_points, curves))
    for k, lc in enumerate(lcs):
        got = np.nonzero(lc.flags() & E.FLAG_KICK_TABLE)[0]
        assert np.array_equal(got, want[k]), (k, got, want[k])
    assert sum(len(w) for w in want) > 0


def test_bright_point_brightness_sum_does_not_wrap(E, oracle):
    """A nearly isolated primary (m2 = m3 = 1e-6) with a source 0.2 rho from
    it at C2's rho and limb darkening (u = 0.5), pixel rho/250: an Einstein
    ring of about 7.8e8 rays (A of about 4200) whose fixed-point brightness
    total (about 1.5 * 2^63 units of 2^-34) is beyond a 64-bit integer.  Ray
    counts equal the oracle's and magnifications are within 1e-10 (the
    reference sums in double, lcirs.cc:407)."""
    lens, rho, u, ppr = (1.0, 1.0, 1.0, 1e-6, 1e-6), 5e-4, 0.5, 250
    traj = (1e-4, 0.0, 1e-4, 1e-5)
    lc = E.LightCurve(*lens, rho, 1, ppr)
    lc.set_limb_darkening("linear", u)
    amp, rays, flags = lc.lc_irs(*traj)
    _assert_no_status(E, flags)
    lo = oracle.new(*lens, rho, 1, ppr, u)
    amp_o, rays_o, _ = lo.lc_irs(*traj)
    assert np.array_equal(rays, rays_o)
    assert _rel(amp, amp_o) < AMP_RTOL
    total = amp / lo.grid()[7]  # the brightness sum, in pixel units
    assert np.all(total * 2.0 ** 34 > 2.0 ** 63) and np.all(amp > 3000)


def test_image_polynomial_on_device_equals_host_and_reference(E):
    """tl_image_poly: the GPU kernel's coefficients equal the host's and the
    reference's fixture values (imgpointcoeff.cc:223-447) to the bit."""
    rows = golden("coeffs.npz")["rows"]
    for r in rows:
        a, b, th, m2, m3, zr, zi = r[:7]
        ref = r[7:].view(np.complex128)
        z = np.array([complex(zr, zi)])
        host = E.image_poly(a, b, th, m2, m3, z)[0]
        dev = E.image_poly(a, b, th, m2, m3, z, on_device=True)[0]
        assert np.array_equal(host.view(np.float64), ref.view(np.float64))
        assert np.array_equal(dev.view(np.float64), ref.view(np.float64))
    rng = np.random.default_rng(3)
    z = rng.uniform(-2, 2, 4096) + 1j * rng.uniform(-2, 2, 4096)
    for lens in [(0.7, 1.3, 2.0, 0.05, 0.002), (1.1, 1.0, 1.0, 1e-3 / 1.001, 0.0)]:
        host = E.image_poly(*lens, z)
        dev = E.image_poly(*lens, z, on_device=True)
        assert np.array_equal(host.view(np.float64), dev.view(np.float64))


def test_in_range_division_equals_ieee_division(E):
    """tl_div_selftest: the fill's in-range quotient (the ray test's
    m_k / |z - z_k|^2) equals IEEE division to the bit on 2^26 operand pairs
    from its range, among them quotients half an ulp from a rounding
    boundary and the ends of the range."""
    lib = E.load()
    lib.tl_div_selftest.argtypes = [ctypes.c_longlong, ctypes.c_ulonglong]
    lib.tl_div_selftest.restype = ctypes.c_int
    for seed in (1, 0x5DEECE66D):
        assert lib.tl_div_selftest(1 << 26, seed) == 0


def test_full_c2_curve_properties(E, oracle):
    """The full benchmark curve (4096 points, pixel rho/200): no status
    flags, bitwise-identical repeat runs, every point's seed set equal to the
    oracle's, and sampled points filled by the oracle from those seeds."""
    from triplelens_amd import workloads as W
    c = W.c2(seed=2, ca_fn=lambda *p: oracle.ccc(*p)[1])
    lc = _new(E, c)
    amp, rays, flags = lc.lc_irs(*c.traj())
    _assert_no_status(E, flags)
    amp2, rays2, _ = lc.lc_irs(*c.traj())
    assert np.array_equal(rays, rays2) and np.array_equal(amp, amp2)
    ns, pix = lc.seeds()
    lo = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u)
    xy, ns_o, _ = lo.seeds(*c.traj())
    g = lo.grid()
    px, py = pixel_seeds(g, xy)
    for i in range(c.n_points):
        ref = set(zip(px[i, :ns_o[i]].tolist(), py[i, :ns_o[i]].tolist()))
        assert ns[i] == ns_o[i] and set(map(tuple, pix[i, :ns[i]].tolist())) == ref, i
    picks = sorted(set(np.linspace(0, c.n_points - 1, 12).astype(int).tolist()
                       + np.argsort(amp)[-4:].tolist()))
    for i in picks:
        t = i / c.lc_len
        r, s, _ = lo.fill_point(c.sx + (c.ex - c.sx) * t, c.sy + (c.ey - c.sy) * t,
                                xy[i, :ns_o[i]])
        assert rays[i] == r, i
        assert abs(amp[i] - s * g[7]) <= AMP_RTOL * abs(s * g[7]), i


def test_scout_options_keep_results(E):
    """Scouts only add labels that merge: without them, with the retry of
    missed scouts off, with the jump shortened on wide runs and with the
    pool handing out its oldest task first, the window around the C2 peak
    gives the reference's ray counts and the same magnifications."""
    g = golden("c2_window.npz")
    c = curve_from_params(g["params"])
    runs = []
    defaults = {"scout_rows": 8, "scout_retry": 1, "scout_width": 768, "pool_newest": 1}
    try:
        for opts in [{"scout_rows": 0}, {"scout_retry": 0}, {"scout_width": 0},
                     {"pool_newest": 0}, {}]:
            for n, v in opts.items():
                E.set_option(n, v)
            amp, rays, flags = _new(E, c).lc_irs(*c.traj())
            for n in opts:
                E.set_option(n, defaults[n])
            assert np.array_equal(rays, g["rays"]), opts
            _assert_no_status(E, flags)
            runs.append(amp)
    finally:
        for n, v in defaults.items():
            E.set_option(n, v)
    for amp in runs[1:]:
        assert np.array_equal(amp, runs[0])


def test_graph_launches_keep_results(E):
    """With use_graph each fill batch (reset, clears, fill, resolve) runs as
    one captured hipGraph.  Alternating graph and direct runs of C1 and of
    the window around the C2 peak give the direct runs' ray counts and
    magnifications every time."""
    from triplelens_amd import workloads as W
    g = golden("c2_window.npz")
    cases = [W.c1(), curve_from_params(g["params"])]
    ref = [_new(E, c).lc_irs(*c.traj()) for c in cases]
    try:
        for rep in range(3):
            for use in (1, 0):
                E.set_option("use_graph", use)
                for c, (A0, r0, _) in zip(cases, ref):
                    A, r, flags = _new(E, c).lc_irs(*c.traj())
                    assert np.array_equal(r, r0) and np.array_equal(A, A0), (rep, use)
                    _assert_no_status(E, flags)
    finally:
        E.set_option("use_graph", 1)
    assert np.array_equal(ref[1][1], g["rays"])
    # the same launch again replays the graph captured by the run before
    # (no new capture), with the engine's own work between the replays
    c, (A0, r0, _) = cases[1], ref[1]
    for rep in range(4):
        A, r, _ = _new(E, c).lc_irs(*c.traj())
        assert np.array_equal(r, r0) and np.array_equal(A, A0), rep
        if rep > 0:
            assert E.last_stats()["graph_builds"] == 0, rep


# ------------------------------------------------------------ seeded fills
def test_fill_from_given_seeds(E, oracle):
    """lcirs_fill_seeds on hand-made seed lists: none, duplicates, misses,
    seeds of one image twice over, and the seed search's own seeds."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    lc.lc_irs(*c.traj())
    ns, pix = lc.seeds()
    lo = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u)
    g = lo.grid()
    N, L, blx, bly = int(g[0]), g[1], g[3], g[4]
    rng = np.random.default_rng(7)
    idx = [0, 25, 99, 150, 199]
    pos, nsl, pixl = [], [], []
    for i in idx:
        t = i / c.lc_len
        p = (c.sx + (c.ex - c.sx) * t, c.sy + (c.ey - c.sy) * t)
        own = pix[i, :ns[i]]
        cases = [own[:0], own, np.concatenate([own, own]), own[:1],
                 rng.integers(0, N, size=(5, 2)), np.concatenate([own, rng.integers(0, N, (3, 2))])]
        for s in cases:
            s = np.asarray(s, np.int32)[:E.MAX_SEEDS]
            buf = np.zeros((E.MAX_SEEDS, 2), np.int32)
            buf[:len(s)] = s
            pos.append(p)
            nsl.append(len(s))
            pixl.append(buf)
    amp, rays, flags = lc.fill_seeds(np.array(pos), np.array(nsl), np.array(pixl))
    _assert_no_status(E, flags)
    for k in range(len(pos)):
        s = pixl[k][:nsl[k]]
        # pixel centres in image coordinates map back to the same pixels
        sxy = np.stack([blx + s[:, 0] * (L / N), bly + s[:, 1] * (L / N)], axis=1)
        px, py = pixel_seeds(g, sxy)
        assert np.array_equal(px, s[:, 0]) and np.array_equal(py, s[:, 1])
        r, b, _ = lo.fill_point(pos[k][0], pos[k][1], sxy)
        assert rays[k] == r, k
        assert abs(amp[k] - b * g[7]) <= AMP_RTOL * max(abs(b * g[7]), 1e-300), k


def test_seed_rows_of_chosen_positions(E):
    """lcirs_copy_seeds_at: the seed rows of chosen positions equal those
    rows of the whole table (lcirs_copy_seeds); an index outside the
    positions is rejected.  Slots past a position's seed count are zero."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    lc.find_seeds(*c.traj())
    ns, pix = lc.seeds()
    past = np.arange(pix.shape[1])[None, :] >= ns[:, None]
    assert ns.min() < pix.shape[1] and not pix[past].any()
    idx = np.array([0, 7, 7, 150, c.lc_len], np.int32)
    ns2, pix2 = lc.seeds(idx)
    assert np.array_equal(ns2, ns[idx]) and np.array_equal(pix2, pix[idx])
    assert lc.seeds(np.zeros(0, np.int32))[0].shape == (0,)
    with pytest.raises(E.EngineError, match="out of range"):
        lc.seeds(np.array([c.lc_len + 1], np.int32))


def test_seed_union_equals_the_python_union(E):
    """lcirs_copy_seeds_union over an object and two clones after split seed
    searches of different piece counts equals surrogate.union_seeds applied
    in turn to their whole seed tables, at every position."""
    from triplelens_amd import workloads as W
    from triplelens_amd.surrogate import union_seeds
    c = W.c1()
    lc = _new(E, c)
    clones = [lc.clone(), lc.clone()]
    for o, parts in zip([lc] + clones, (5, 4, 7)):
        o.set_seed_split(parts)
    E.batch_find_seeds([lc] + clones, np.array([c.traj()] * 3))
    idx = np.arange(c.lc_len + 1, dtype=np.int32)
    ns, pix = lc.seeds()
    for o in clones:
        ns2, pix2 = o.seeds()
        ns, pix = union_seeds(ns, pix, ns2, pix2)
    ns_u, pix_u = lc.seeds_union(clones, idx)
    assert np.array_equal(ns_u, ns)
    for i in idx:
        assert np.array_equal(pix_u[i, :ns[i]], pix[i, :ns[i]])
    sel = np.array([3, 3, 150], np.int32)
    ns_s, pix_s = lc.seeds_union(clones, sel)
    assert np.array_equal(ns_s, ns[sel]) and np.array_equal(pix_s, pix_u[sel])
    with pytest.raises(E.EngineError, match="out of range"):
        lc.seeds_union(clones, np.array([-1], np.int32))


def test_fill_seeds_rejects_bad_arguments(E):
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    with pytest.raises(E.EngineError):
        lc.fill_seeds(np.zeros((1, 2)), np.array([E.MAX_SEEDS + 1]), np.zeros((1, E.MAX_SEEDS, 2)))


def test_amoeba_print_out_matches_reference(E, tmp_path):
    """set_amoeba_printout: one file per point with the runs of its fill, and
    the parameter file, as the reference writes them (five C1 points)."""
    from triplelens_amd.ctypes_classes import LC_irs
    g = golden("amoeba_c1.npz")
    c = curve_from_params(g["params"])
    lc = LC_irs(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening("linear", c.u)
    lc.set_amoeba_printout(str(tmp_path / "Amoeba_"), str(tmp_path / "Par"))
    lc.get_lc_irs(*c.traj())
    off = g["off"]
    for i in range(c.n_points):
        a = np.loadtxt(tmp_path / f"Amoeba_{i}", delimiter=",", dtype=np.int64).reshape(-1, 3)
        a = a[np.lexsort((a[:, 1], a[:, 0]))]
        assert np.array_equal(a, g["segs"][off[i]:off[i + 1]]), i
    assert (tmp_path / "Par").read_text() == str(g["par"])


def test_lc_cuda_entry_points_give_the_irs_curve(E):
    """LC_cuda (lccuda_new / get_lc_cuda / set_limb_darkening_cuda) computes
    the same curve as LC_irs."""
    from triplelens_amd.ctypes_classes import LC_cuda, LC_irs
    g = golden("c1.npz")
    c = curve_from_params(g["params"])
    out = []
    for cls, run in ((LC_irs, "get_lc_irs"), (LC_cuda, "get_lc_cuda")):
        lc = cls(*c.lens(), c.rho, c.lc_len, c.ppr)
        lc.set_limb_darkening(b"linear", c.u)
        getattr(lc, run)(*c.traj())
        v = np.zeros(c.lc_len)
        lc.copy_lc(v)
        out.append(v)
    assert np.array_equal(out[0], out[1])
    assert _rel(out[1], g["amp"][:c.lc_len]) < 1e-6


def test_critical_curves_batch_equal_host(E):
    """tl_ccc_batch: the critical curves and caustics of 96 geometries on the
    GPU equal the host set-up's (which equals the reference's) to the bit."""
    from triplelens_amd import workloads as W
    rng = np.random.default_rng(11)
    rows = [list(c.lens()) for c in W.c3()]
    for _ in range(32):
        rows.append([float(np.exp(rng.uniform(np.log(0.3), np.log(3)))),
                     float(np.exp(rng.uniform(np.log(0.3), np.log(3)))),
                     float(rng.uniform(0, 2 * np.pi)),
                     float(np.exp(rng.uniform(np.log(1e-4), np.log(0.3)))),
                     float(np.exp(rng.uniform(np.log(1e-4), np.log(0.3))))])
    cc, ca = E.ccc_batch(np.array(rows))
    for g, r in enumerate(rows):
        hc, ha = E.LightCurve(*r, 1e-3, 10, 10).cc_ca()
        assert np.array_equal(cc[g], hc) and np.array_equal(ca[g], ha), g


def test_point_magnification_and_centroid_at_irs_positions(E, oracle):
    """lcirs_copy_point_amp: A_point at the IRS positions i/lcLength equals the
    oracle's getPointAmp chain (to 1e-10; the two chains differ in the trial
    solves in between) and the centroid the weighted mean of the oracle's
    images."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    lc.lc_irs(*c.traj())
    amp, cen, caus = lc.point_amp()
    t = np.arange(c.n_points) / c.lc_len
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    ref = oracle.point_amp(*c.lens(), pos)
    assert _rel(amp, ref) < 1e-10
    roots, is_img, _ = oracle.images(*c.lens(), pos)
    L = c.lens()
    z1, z2 = 0.0, L[0]
    z3 = L[1] * np.exp(1j * L[2])
    m1, m2, m3 = 1 - L[3] - L[4], L[3], L[4]
    for i in range(0, c.n_points, 7):
        z = roots[i][is_img[i]]
        s = m1 / (z - z1) ** 2 + m2 / (z - z2) ** 2 + m3 / (z - z3) ** 2
        w = 1 / np.abs(1 - np.abs(s) ** 2)
        want = np.sum(w * z) / np.sum(w)
        assert abs(complex(*cen[i]) - want) < 1e-9 * max(1.0, abs(want)), i
    # caustic distance: the nearest of the curve's caustic samples, signed by
    # the image count
    _, ca = lc.cc_ca()
    for i in range(0, c.n_points, 13):
        d = np.min(np.abs(ca.ravel() - pos[i]))
        want = -d if is_img[i].sum() > 4 else d
        assert abs(caus[i] - want) <= 1e-12 * max(1.0, d), i


def test_disk_samples_match_oracle(E, oracle):
    """lcirs_set_ring: the 12 disk-sample magnifications around each IRS
    position equal the oracle's (tlo_ring_amp) from the same centre roots,
    and those roots give the position's point magnification."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    lc.set_ring(True)
    lc.find_seeds(*c.traj())
    amp, _, _ = lc.point_amp()
    ring, roots, cross = lc.ring_amp()
    t = np.arange(c.n_points) / c.lc_len
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    want, want_cross = oracle.ring_amp(*c.lens(), c.rho, pos, roots)
    assert _rel(ring, want) < 1e-13
    assert np.array_equal(cross, want_cross)
    L = c.lens()
    zs = np.array([0.0, L[0], L[1] * np.exp(1j * L[2])])
    ms = np.array([1 - L[3] - L[4], L[3], L[4]])
    for i in range(0, c.n_points, 11):
        z = roots[i][:, None]
        res = np.abs(z[:, 0] - np.sum(ms / np.conj(z - zs), axis=1) - pos[i])
        s = np.sum(ms / (z - zs) ** 2, axis=1)
        a = np.sum(1 / np.abs(1 - np.abs(s[res < 1e-6]) ** 2))
        assert abs(a - amp[i]) <= 1e-12 * amp[i], i
    # without set_ring the copy is refused
    lc2 = _new(E, c)
    lc2.find_seeds(*c.traj())
    with pytest.raises(E.EngineError):
        lc2.ring_amp()


def test_split_seed_search(E, oracle):
    """lcirs_set_seed_split: the first piece continues the chain and gives
    the single chain's seeds and ray counts to the bit; every other piece
    starts from a fresh solve (the oracle's from-scratch roots, to the bit);
    and almost every position's ray count is the single chain's.  Where a
    solve is ill-conditioned the two can differ: at C1's caustic entry
    (position 26) the reference's chain loses an image pair that the piece
    started afresh at position 25 finds."""
    from triplelens_amd import workloads as W
    c = W.c1()
    ref = _new(E, c)
    ref.set_ring(True)
    A0, rays0, _ = ref.lc_irs(*c.traj())
    ns0, pix0 = ref.seeds()
    parts = 8
    lc = _new(E, c)
    lc.set_ring(True)
    lc.set_seed_split(parts)
    A1, rays1, _ = lc.lc_irs(*c.traj())
    ns1, pix1 = lc.seeds()
    _, roots, _ = lc.ring_amp()
    n = c.n_points
    first = n // parts  # piece q holds positions [n q / P, n (q+1) / P)
    assert np.array_equal(rays1[:first], rays0[:first])
    assert np.array_equal(ns1[:first], ns0[:first]) and np.array_equal(pix1[:first], pix0[:first])
    t = np.arange(n) / c.lc_len
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    for q in range(1, parts):
        i0 = n * q // parts
        fresh, _, _ = oracle.images(*c.lens(), pos[i0:i0 + 1])
        assert np.array_equal(roots[i0], fresh[0]), q
    assert np.mean(rays1 == rays0) > 0.95
    with pytest.raises(E.EngineError):
        lc.set_seed_split(0)


def test_fill_positions_of_the_last_seed_search(E):
    """lcirs_fill_positions: ray shooting chosen positions of the last seed
    search gives get_lc_irs's ray counts and magnifications at them."""
    from triplelens_amd import workloads as W
    c = W.c1()
    ref = _new(E, c)
    A0, rays0, _ = ref.lc_irs(*c.traj())
    lc = _new(E, c)
    lc.find_seeds(*c.traj())
    idx = np.array([199, 0, 26, 27, 150, 3, 100])
    A, rays, flags = lc.fill_positions(idx)
    assert np.array_equal(rays, rays0[idx]) and np.array_equal(A, A0[idx])
    A, rays, _ = lc.fill_positions(np.arange(c.n_points))
    assert np.array_equal(rays, rays0) and np.array_equal(A, A0)
    with pytest.raises(E.EngineError):
        lc.fill_positions(np.array([c.n_points]))
    # lcirs_fill_positions_seeds: the search's own seeds give the same
    # results; a seed set with duplicates and an extra seed of an image
    # already seeded counts every ray once; no seeds count nothing
    ns, pix = lc.seeds()
    A, rays, _ = lc.fill_positions(idx, (ns[idx], pix[idx]))
    assert np.array_equal(rays, rays0[idx]) and np.array_equal(A, A0[idx])
    ns2, pix2 = ns[idx].copy(), pix[idx].copy()
    for i in range(len(idx)):
        k = ns2[i]
        pix2[i, k:2 * k] = pix2[i, :k]
        ns2[i] = 2 * k
    A, rays, _ = lc.fill_positions(idx, (ns2, pix2))
    assert np.array_equal(rays, rays0[idx]) and np.array_equal(A, A0[idx])
    _, rays, _ = lc.fill_positions(idx, (np.zeros(len(idx), np.int32), pix[idx]))
    assert not rays.any()


def test_clones_and_batched_seed_searches(E):
    """lcirs_clone copies an object's set-up (its light curve is the
    original's to the bit), and lcirs_batch_find_seeds gives each object the
    seeds its own lcirs_find_seeds gives."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lc = _new(E, c)
    A0, rays0, _ = lc.lc_irs(*c.traj())
    cl = _new(E, c).clone()
    A1, rays1, _ = cl.lc_irs(*c.traj())
    assert np.array_equal(rays0, rays1) and np.array_equal(A0, A1)
    parts = [7, 13, 50]
    seq = []
    for p in parts:
        o = _new(E, c)
        o.set_seed_split(p)
        o.find_seeds(*c.traj())
        seq.append(o.seeds())
    base = _new(E, c)
    clones = [base.clone() for _ in parts]
    for o, p in zip(clones, parts):
        o.set_seed_split(p)
    E.batch_find_seeds(clones, np.array([c.traj()] * len(parts)))
    for o, (ns, pix) in zip(clones, seq):
        ns2, pix2 = o.seeds()
        assert np.array_equal(ns, ns2)
        for i in range(len(ns)):
            assert np.array_equal(pix[i, :ns[i]], pix2[i, :ns[i]]), i


def test_binary_lens_matches_oracle_and_golden(E, oracle):
    """The binary lens (m3 = 0, irsBinary and the binary image-count rule):
    the C1 binary's and the planetary binary's ray counts equal the
    reference's and the oracle's at every point, and the C1 binary's seeds
    the reference's."""
    g = golden("binary.npz")
    for k, p in enumerate(g["params"]):
        c = curve_from_params(p)
        lc = _new(E, c)
        amp, rays, flags = lc.lc_irs(*c.traj())
        _assert_no_status(E, flags)
        amp_o, rays_o, _ = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u).lc_irs(*c.traj())
        assert np.array_equal(rays, rays_o), (k, np.where(rays != rays_o)[0][:10])
        assert _rel(amp, amp_o) < AMP_RTOL, k
        und = (flags & E.FLAG_KICK_TABLE) != 0
        assert und.sum() <= 1, k
        assert np.array_equal(rays[~und], g["rays"][k][~und]), k
        assert _rel(amp[~und], g["amp"][k][~und]) < REF_RTOL, k
        if k == 0:
            ns, pix = lc.seeds()
            off, xy = g["seed_off"][0], g["seed_xy0"]
            px, py = pixel_seeds(g["grid"][0], xy)
            for i in range(c.n_points):
                ref = set(zip(px[off[i]:off[i + 1]].tolist(), py[off[i]:off[i + 1]].tolist()))
                assert set(map(tuple, pix[i, :ns[i]].tolist())) == ref, i


def test_autotuner_keeps_results(E, tmp_path):
    """The autotuner times option sets on C1, finds every set's ray counts
    identical, and its saved choice applies."""
    from triplelens_amd import autotune, workloads as W
    space = {"fill_workgroups": [0, 512], "scout_jump": [64, 256], "run_pixels": [128]}
    best, table = autotune.tune(W.c1(), space, repeats=1, log=lambda s: None)
    assert len(table) == 4 and best in [row[0] for row in table]
    p = tmp_path / "tuning.json"
    p.write_text(__import__("json").dumps({"options": best}))
    try:
        assert autotune.apply(str(p)) == best
    finally:
        for n, v in autotune.DEFAULTS.items():
            E.set_option(n, v)
