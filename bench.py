#!/usr/bin/env python3
"""Light-curve benchmark of the MI355X IRS engine: surrogate + IRS against
plain IRS on the C2 workload.

The C2 workload: a seeded triple lens (separations log-uniform [0.5, 2],
m2, m3 log-uniform [1e-3, 0.3]), linear limb darkening u = 0.5,
rho = 5e-4, pixel rho/200, 4096 positions on a straight line of length 2
through the central caustic (geometry seed 2, which the surrogate's
training and calibration curves exclude).

Two timed loops over the same curve, each of --steps light curves:
  * plain IRS: get_lc_irs through the C ABI (seed search with the
    reference's single root chain, flood fill and ray map on the GPU);
  * surrogate + IRS (the default mode): the split seed search with the
    disk samples, the network on the GPU, and ray shooting of the positions
    the gate sends back plus a 1% audit (Surrogate.light_curve).
The printed value is the surrogate loop's light-curve points per second.
Accuracy is measured against the reference itself: tests/golden/c2_full.npz
holds the reference LightCurveIRS's ray counts and magnifications of this
curve at all 4096 positions (written by tests/golden/make_golden.py from the
reference build), and every timed curve of both loops is compared with it:
the plain loop's ray counts must be equal at every position, and the
largest relative magnification error of the surrogate loop over all its
timed curves is the reported error (the bound is 1e-3).  With --mode plain
only the plain loop runs and its rate is the value.

Roofline figures are those of the fill kernel, where the path's arithmetic
is (43 FP64 operations per counted ray): its launch durations come from HIP
events around it on the engine's stream, over the plain-IRS timed loop, and
its HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
summary under profiles/.

The CPU baseline is the reference LightCurveIRS compiled from its own
sources (oracle/_ref: the parity flags -O2 -ffp-contract=off -mavx, and, as
a second figure, the reference's own -Ofast -mavx) run over the same curve
on the host cores, contiguous parts of the trajectory spread over one worker
process per core (the box's CPU share for one GPU, at most 16); it is timed
before the GPU is first used.  Where the reference build is absent the C
restatement in oracle/ is timed instead.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FLOP_PER_RAY = 43          # irsSIMD 35 + pixel position 2 + profile and sum 6
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector peak (the fill runs on the vector ALUs)
ROOF_BOUND = "fp64-valu"   # FP64 vector arithmetic, not memory and not the matrix cores
BOUND = 1e-3               # largest relative magnification error allowed
REF_C2 = os.path.join(HERE, "tests", "golden", "c2_full.npz")
METRIC = "IRS light-curve points/s at ≤1e-3 rel err: surrogate+IRS vs plain IRS, 1 GPU"
MODEL = os.path.join(HERE, "triplelens_amd", "data", "surrogate_c2.npz")


def c2_curve():
    """C2 geometry; the central caustic is located with the engine's own
    critical curves (host set-up, no GPU work)."""
    from triplelens_amd import engine as E
    from triplelens_amd import workloads as W

    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, 200).cc_ca()[1]

    return W.c2(seed=2, ca_fn=ca_fn)


def reference_c2(c):
    """The reference's ray counts and magnifications of curve c (the
    committed C2 fixture), or None when the fixture is absent or holds
    another curve."""
    try:
        z = np.load(REF_C2, allow_pickle=False)
    except OSError:
        return None
    p = z["params"][0]
    want = [*c.lens(), c.rho, c.ppr, c.u, *c.traj(), c.n_points]
    if not np.array_equal(p, np.array(want, np.float64)):
        return None
    return z["rays"].astype(np.uint64), z["amp"]


def host_cpu():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


# ---------------------------------------------------------------- CPU baseline
def _ref_worker(args):
    import ctypes
    lib_path, lens, rho, ppr, u, traj, n = args
    D = ctypes.c_double
    lib = ctypes.CDLL(lib_path)
    lib.tlref_lc_irs.argtypes = [D] * 6 + [ctypes.c_uint, ctypes.c_long, ctypes.c_int, D, D, D,
                                           D, D, ctypes.c_void_p, ctypes.c_void_p]
    amp = np.zeros(n)
    rays = np.zeros(n, np.uint64)
    t = time.perf_counter()
    lib.tlref_lc_irs(*lens, rho, n - 1, ppr, 1, u, *traj, amp.ctypes.data, rays.ctypes.data)
    return time.perf_counter() - t, int(rays.sum()), n


def _split(c, parts):
    """Contiguous parts of the trajectory: (start, end, points) per part."""
    n = c.n_points
    bounds = np.linspace(0, n, parts + 1).astype(int)
    out = []
    for k in range(parts):
        i0, i1 = bounds[k], bounds[k + 1]
        if i1 - i0 < 2:
            continue
        t0, t1 = i0 / c.lc_len, (i1 - 1) / c.lc_len
        out.append(((c.sx + (c.ex - c.sx) * t0, c.sy + (c.ey - c.sy) * t0,
                     c.sx + (c.ex - c.sx) * t1, c.sy + (c.ey - c.sy) * t1), i1 - i0))
    return out


def measured_traffic():
    """HBM bytes per fill launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by a plain copy kernel's readings), or None."""
    path = os.path.join(HERE, "profiles", "fill_kernel_c2_traffic.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def _time_ref(lib, c, parts, cores):
    jobs = [(lib, c.lens(), c.rho, c.ppr, c.u, tr, n) for tr, n in parts]
    # largest parts first so that the pool stays balanced
    jobs.sort(key=lambda j: -j[-1])
    t = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = list(pool.imap_unordered(_ref_worker, jobs, chunksize=1))
    wall = time.perf_counter() - t
    return wall, sum(r[2] for r in res) / sum(r[0] for r in res)


def cpu_baseline(c, cores):
    from oracle import oracle as orc
    parts = _split(c, 8 * cores)
    extra = {}
    if os.path.exists(orc.REF_SO):
        wall, one_core = _time_ref(orc.REF_SO, c, parts, cores)
        kind = "reference"
        if os.path.exists(orc.REF_FAST_SO):
            wall_f, one_f = _time_ref(orc.REF_FAST_SO, c, parts, cores)
            pts = sum(n for _, n in parts)
            extra = {"value_ofast": float(pts / wall_f), "value_ofast_1core": float(one_f),
                     "ofast_flags": "-Ofast -mavx (the reference makefile's, without "
                                    "-march=native)"}
    else:
        O = orc.Oracle()
        params = np.array([[*c.lens(), c.rho, c.ppr, c.u] for _ in parts])
        trajs = np.array([tr for tr, _ in parts])
        lens = np.array([n - 1 for _, n in parts])
        t = time.perf_counter()
        O.batch(params, trajs, lens, cores)
        wall = time.perf_counter() - t
        kind = "port"
        one_core = None
    pts = sum(n for _, n in parts)
    model, nproc = host_cpu()
    out = {"value": float(pts / wall), "unit": "points/s", "cores": cores, "kind": kind,
           "sample": f"the full {c.n_points}-point C2 curve split into {len(parts)} contiguous "
                     f"parts over {cores} worker processes, wall {wall:.1f} s",
           "flags": "-O2 -ffp-contract=off -mavx (parity flags)" if kind == "reference"
                    else "-O2 -ffp-contract=off",
           "host_cpu": model, "host_nproc": nproc}
    if one_core is not None:
        out["value_1core"] = float(one_core)
    out.update(extra)
    return out


# ---------------------------------------------------------------- copy kernel
def calibrate_copy(mib):
    """Three device-to-device copies of `mib` MiB by torch's copy kernel: each
    reads and writes exactly mib MiB of HBM, so the FETCH_SIZE and WRITE_SIZE
    that rocprofv3 reports for it give the counters' scale."""
    import torch
    n = mib * (1 << 20)
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
    return 0


# ---------------------------------------------------------------- C3 batch
def c3_bench(args):
    """64 seeded light curves (separations 0.3-3, mass ratios 1e-4-1, rho
    1e-4-1e-2, lengths 256-8192 points, pixel rho/100) by plain IRS in one
    batched engine run per step (lcirs_batch_irs)."""
    from triplelens_amd import engine as E, workloads as W
    if args.tuning:
        from triplelens_amd import autotune
        autotune.apply(args.tuning)
    name, cus = E.device_info()
    curves = W.c3()
    trajs = np.array([c.traj() for c in curves])

    def objects():
        out = []
        for c in curves:
            lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
            lc.set_limb_darkening("linear", c.u)
            out.append(lc)
        return out

    for _ in range(args.warmup):
        E.batch_irs(objects(), trajs)
    fill_ms = launches = rays = seed_ms = builds = 0.0
    flags_all = 0
    wall = 0.0
    for _ in range(args.steps):
        lcs = objects()  # a fresh root chain per curve, as a new run has
        t = time.perf_counter()
        res = E.batch_irs(lcs, trajs)
        wall += time.perf_counter() - t
        st = E.last_stats()
        fill_ms += st["fill_ms"]
        seed_ms += st["seed_ms"]
        launches += st["fill_launches"]
        builds += st["graph_builds"]
        for _a, r, f in res:
            rays += float(r.sum())
            flags_all |= int(np.bitwise_or.reduce(f))
    pts = sum(c.n_points for c in curves)
    fill_launch_ms = fill_ms / launches
    achieved = FLOP_PER_RAY * (rays / launches) / (fill_launch_ms * 1e-3) / 1e12
    line = {
        "metric": "C3 batched IRS light-curve points/s, 1 GPU",
        "value": pts * args.steps / wall,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "C3: 64 seeded triple-lens curves (separations 0.3-3, mass ratios "
                        "1e-4-1, rho 1e-4-1e-2, u 0-0.8, 256-8192 points), pixel rho/100, one "
                        "batched engine run per step",
            "device": name,
            "cus": cus,
            "points_per_step": pts,
            "rays_per_step": rays / args.steps,
            "rays_per_s": rays / wall,
            "seed_kernel_ms_per_step": seed_ms / args.steps,
            "fill_kernel_ms_per_step": fill_ms / args.steps,
            "fill_launches_per_step": launches / args.steps,
            "fill_graph_captures_per_step": builds / args.steps,
            "status_flags": flags_all,
        },
        "roofline": {
            "bound": ROOF_BOUND,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "fill_kernel",
            "kernel_ms": fill_launch_ms,
            "flop_per_ray": FLOP_PER_RAY,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


# ---------------------------------------------------------------- Metropolis
def metropolis_bench(args, c2):
    """The Metropolis fitting workload at reduced size: args.chains chains of
    args.proposals proposals, each a light curve of args.points positions on
    the C2 trajectory, 1% Gaussian steps of the lens parameters.  Times the
    surrogate + IRS fallback chains and all-IRS chains (batched across
    chains) from the same starts."""
    import dataclasses
    from triplelens_amd import engine as E, metropolis as M, surrogate as S
    if args.tuning:
        from triplelens_amd import autotune
        autotune.apply(args.tuning)
    name, cus = E.device_info()
    truth = dataclasses.replace(c2, n_points=args.points)
    obs, err = M.observed(truth)
    starts = M.starts_near(truth, args.chains)
    sur = S.Surrogate.load(args.model)
    for _ in range(args.warmup):
        M.run(starts, obs, err, 1, "surrogate", sur)
        M.run(starts, obs, err, 1, "irs")
    def progress(mode):
        def f(step, points, secs):
            print(f"metropolis {mode}: {step} steps, {points} points, {secs:.1f} s", file=sys.stderr,
                  flush=True)
        return f

    rs = M.run(starts, obs, err, args.proposals, "surrogate", sur, progress=progress("surrogate"))
    n_irs = args.irs_proposals if args.irs_proposals > 0 else args.proposals
    ri = M.run(starts, obs, err, min(n_irs, args.proposals), "irs", progress=progress("irs"))
    sur_pps = rs["points"] / rs["wall_s"]
    irs_pps = ri["points"] / ri["wall_s"]
    errs = [float(np.max(np.abs(rs["first"][k] / ri["first"][k] - 1.0)))
            for k in rs["first"] if k in ri["first"]]
    n_prop = args.chains * args.proposals
    line = {
        "metric": "Metropolis fitting: light-curve points/s, surrogate+IRS fallback, 1 GPU",
        "value": sur_pps,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.proposals,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * rs["wall_s"] / args.proposals,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"Metropolis fitting: {args.chains} chains x {args.proposals} proposals x "
                        f"{args.points}-point light curves on the C2 trajectory, 1% Gaussian steps "
                        "of a, b, th, m2, m3, observed curve with 1% noise",
            "device": name,
            "cus": cus,
            "all_irs_points_per_s": irs_pps,
            "all_irs_proposals_per_chain": min(n_irs, args.proposals),
            "speedup_vs_all_irs": sur_pps / irs_pps,
            "fallback_fraction": rs["fallback"] / max(1, rs["points"]),
            "audit_fraction": rs["audited"] / max(1, rs["points"]),
            "acceptance": sum(ch.accepted for ch in rs["chains"]) / n_prop,
            "acceptance_all_irs": sum(ch.accepted for ch in ri["chains"]) /
                                  (args.chains * min(n_irs, args.proposals)),
            "points": rs["points"],
            "wall_s": rs["wall_s"],
            "max_rel_err_first_proposals": max(errs) if errs else None,
            "bound_met": bool(errs) and max(errs) <= 1e-3,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


# ---------------------------------------------------------------- main
def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["surrogate", "plain", "metropolis", "c3"],
                    default="surrogate",
                    help="surrogate: the C2 headline (surrogate + IRS vs plain IRS); plain: "
                         "plain IRS alone; metropolis: the Metropolis fitting workload; c3: "
                         "64 seeded curves of uneven lengths in one batched engine run")
    ap.add_argument("--chains", type=int, default=8, help="metropolis: chains")
    ap.add_argument("--proposals", type=int, default=8, help="metropolis: proposals per chain")
    ap.add_argument("--points", type=int, default=2048, help="metropolis: points per curve")
    ap.add_argument("--irs-proposals", type=int, default=0,
                    help="metropolis: proposals per chain of the all-IRS comparison (default: "
                         "all; the all-IRS rate is measured on this many)")
    ap.add_argument("--model", default=MODEL, help="trained surrogate (.npz)")
    ap.add_argument("--tuning", default="", help="engine options saved by triplelens_amd.autotune")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1),
                    help="worker processes of the CPU baseline (default: the host's CPUs, at "
                         "most 16, the box's CPU share for one GPU)")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args(argv)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the "
              "engine has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        return 2

    c = c2_curve()
    if args.mode in ("surrogate", "metropolis") and not os.path.exists(args.model):
        print(f"{args.model} is missing: train it with python -m triplelens_amd.train_surrogate "
              f"--out {args.model}", file=sys.stderr)
        return 2
    if args.mode == "metropolis":
        return metropolis_bench(args, c)
    if args.mode == "c3":
        return c3_bench(args)
    cpu = None if args.no_cpu else cpu_baseline(c, args.cpu_cores)

    from triplelens_amd import engine as E
    name, cus = E.device_info()
    tuning = None
    if args.tuning:
        from triplelens_amd import autotune
        tuning = autotune.apply(args.tuning)
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening("linear", c.u)
    for _ in range(args.warmup):
        lc.lc_irs(*c.traj())
    fill_ms = launches = rays = seed_ms = 0.0
    flags_all = 0
    curves_plain = []

    def fresh_objects():
        # one object per timed curve, made before the clock starts: each
        # timed curve starts the root chain from scratch, as the reference's
        # first getLCIRS on a new LightCurveIRS does
        out = []
        for _ in range(args.steps):
            o = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
            o.set_limb_darkening("linear", c.u)
            out.append(o)
        return out

    objs = fresh_objects()
    t = time.perf_counter()
    for lc in objs:
        A_plain, r, f = lc.lc_irs(*c.traj())
        st = E.last_stats()
        fill_ms += st["fill_ms"]
        seed_ms += st["seed_ms"]
        launches += st["fill_launches"]
        rays += float(r.sum())
        flags_all |= int(np.bitwise_or.reduce(f))
        curves_plain.append((A_plain, r))
    wall_plain = time.perf_counter() - t
    ref = reference_c2(c)
    plain_pps = c.n_points * args.steps / wall_plain
    fill_launch_ms = fill_ms / launches
    rays_per_launch = rays / launches
    achieved = FLOP_PER_RAY * rays_per_launch / (fill_launch_ms * 1e-3) / 1e12
    config = {
        "workload": "C2: seeded triple lens, rho=5e-4, linear limb darkening u=0.5, "
                    "pixel rho/200, 4096-point trajectory through the central caustic",
        "n_points": c.n_points,
        "geometry_seed": 2,
        "lens": {"a": c.a, "b": c.b, "th": c.th, "m2": c.m2, "m3": c.m3},
        "device": name,
        "cus": cus,
        "engine_options": tuning or "defaults",
        "plain_irs_points_per_s": plain_pps,
        "plain_irs_ms_per_curve": 1e3 * wall_plain / args.steps,
        "plain_rays_per_curve": rays / args.steps,
        "plain_status_flags": flags_all,
        "plain_seed_kernel_ms_per_curve": seed_ms / args.steps,
        "plain_fill_kernel_ms_per_curve": fill_ms / args.steps,
        "plain_fill_launches_per_curve": launches / args.steps,
    }
    if ref is not None:
        config["plain_ray_mismatches_vs_reference"] = int(max(
            np.count_nonzero(r != ref[0]) for _, r in curves_plain))
        config["plain_max_rel_err_vs_reference"] = float(max(
            np.max(np.abs(A / ref[1] - 1.0)) for A, _ in curves_plain))
    if args.mode == "plain":
        config["mode"] = "plain IRS (no surrogate)"
        value, ms_per_step = plain_pps, 1e3 * wall_plain / args.steps
    else:
        from triplelens_amd import surrogate as S
        sur = S.Surrogate.load(args.model)
        lcs = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        lcs.set_limb_darkening("linear", c.u)
        rng = np.random.default_rng(1)
        for _ in range(args.warmup):
            sur.light_curve(lcs, c, rng)
        stats = []
        shot_rays = n_fb = n_aud = n_fail = 0
        audit_err = 0.0
        curves_sur = []
        objs = fresh_objects()
        for o in objs:
            sur._extra_clones(o, c)  # the seed union's clones, made with the object
        t = time.perf_counter()
        for lcs in objs:
            A_sur, info = sur.light_curve(lcs, c, rng, stats)
            shot_rays += info["rays"]
            n_fb += int(info["fallback"].sum())
            n_aud += int(info["audited"].sum())
            n_fail += int(info["audit_failed"])
            audit_err = max(audit_err, float(info["audit_rel_err"].max(initial=0.0)))
            curves_sur.append((A_sur, info["fallback"] | info["audited"]))
        wall_sur = time.perf_counter() - t
        value = c.n_points * args.steps / wall_sur
        ms_per_step = 1e3 * wall_sur / args.steps
        # every timed curve against the plain curve and the reference
        rel = np.max([np.abs(A / A_plain - 1.0) for A, _ in curves_sur], axis=0)
        acc = ~curves_sur[-1][1]
        if ref is not None:
            rel_ref = np.max([np.abs(A / ref[1] - 1.0) for A, _ in curves_sur], axis=0)
            config.update({
                "max_rel_err_vs_reference": float(rel_ref.max()),
                "max_rel_err_vs_reference_accepted": float(np.max(
                    [np.max(np.abs(A / ref[1] - 1.0)[~shot], initial=0.0)
                     for A, shot in curves_sur])),
                "reference": "reference LightCurveIRS ray counts and magnifications of this "
                             "curve (tests/golden/c2_full.npz, all 4096 positions)",
            })
        config.update({
            "mode": "surrogate + IRS fallback (split seed search, disk samples, BF16 network, "
                    "gate, 1% audit)",
            "model": os.path.relpath(args.model, HERE),
            "gate_radii": sur.gate_radii,
            "hex_max": sur.hex_max,
            "band_edges": sur.band_edges.tolist(),
            "sigma_max": sur.sigma_max.tolist(),
            "speedup_vs_plain_irs": value / plain_pps,
            "max_rel_err_vs_plain_irs": float(rel.max()),
            "max_rel_err_accepted": float(rel[acc].max(initial=0.0)),
            "bound_met": bool((rel_ref.max() if ref is not None else rel.max()) <= BOUND),
            "bound_measured_against": "reference" if ref is not None else "plain engine IRS",
            "fallback_fraction": n_fb / (args.steps * c.n_points),
            "audit_fraction": n_aud / (args.steps * c.n_points),
            "max_audit_rel_err": audit_err,
            "failed_audits": n_fail,
            "ray_share_shot": shot_rays / rays,
            "surrogate_seed_kernel_ms_per_curve": sum(s["seed_ms"] for s in stats) / args.steps,
            "surrogate_fill_kernel_ms_per_curve": sum(s["fill_ms"] for s in stats) / args.steps,
        })
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {
            "bound": ROOF_BOUND,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(),
            "kernel": "fill_kernel",
            "kernel_ms": fill_launch_ms,
            "flop_per_ray": FLOP_PER_RAY,
            "timed_loop": "plain IRS",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
