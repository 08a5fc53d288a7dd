"""Autotuner of the engine's launch options.

    python -m triplelens_amd.autotune --workload c2 --out tuning.json

Times plain ray shooting (get_lc_irs) of a workload curve under every
combination of the fill's options (`tl_set_option`):

    fill_workgroups  workgroups of 4 wavefronts (0 = one per CU)
    scout_jump       rows a scout jumps ahead
    scout_width      runs wider than this shorten the jump (0 = never)
    run_pixels       predicted rays per run (sizes the run table and batches)
    queue_log2       capacity of the shared task queue (2^k tasks)

checks that every combination gives the same ray counts (the options only
schedule the work), and writes the fastest as JSON.  `apply(path)` sets
them in the engine.
"""
from __future__ import annotations

import argparse
import itertools
import json
import time

import numpy as np

from . import engine as E
from . import workloads as W

DEFAULT_SPACE = {
    "fill_workgroups": [0, 512],
    "scout_jump": [256, 512, 1024],
    "scout_width": [0, 512, 768, 2048],
    "run_pixels": [64, 128],
}
DEFAULTS = {"fill_workgroups": 0, "scout_jump": 512, "scout_width": 768, "run_pixels": 128,
            "queue_log2": 21}


def workload(name: str):
    if name == "c1":
        return W.c1()
    if name == "c2":
        def ca_fn(a, b, th, m2, m3):
            return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, 200).cc_ca()[1]
        return W.c2(seed=2, ca_fn=ca_fn)
    raise ValueError(f"unknown workload {name}")


def _run(c):
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening("linear", c.u)
    t = time.perf_counter()
    _, rays, flags = lc.lc_irs(*c.traj())
    return time.perf_counter() - t, rays, E.last_stats()["fill_ms"]


def tune(c, space=None, repeats: int = 2, log=print):
    """Best option set for curve c: the one with the lowest median fill
    time.  Returns (best, table) with table rows (options, fill_ms, wall_s)."""
    space = DEFAULT_SPACE if space is None else space
    names = list(space)
    rays0 = None
    table = []
    try:
        for combo in itertools.product(*(space[n] for n in names)):
            opts = dict(zip(names, combo))
            for n, v in opts.items():
                E.set_option(n, v)
            fills, walls = [], []
            for _ in range(repeats):
                wall, rays, fill_ms = _run(c)
                if rays0 is None:
                    rays0 = rays
                elif not np.array_equal(rays, rays0):
                    raise RuntimeError(f"ray counts changed under {opts}")
                fills.append(fill_ms)
                walls.append(wall)
            row = (opts, float(np.median(fills)), float(np.median(walls)))
            table.append(row)
            log(f"{opts}: fill {row[1]:.1f} ms, wall {1e3 * row[2]:.1f} ms")
    finally:
        for n, v in DEFAULTS.items():
            E.set_option(n, v)
    best = min(table, key=lambda r: r[1])[0]
    return best, table


def apply(path: str) -> dict:
    """Sets the options saved by tune (a JSON object name -> value)."""
    with open(path) as f:
        opts = json.load(f)["options"]
    for n, v in opts.items():
        E.set_option(n, v)
    return opts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--repeats", type=int, default=2)
    args = ap.parse_args(argv)
    c = workload(args.workload)
    _run(c)  # warm-up: first-call allocations
    best, table = tune(c, repeats=args.repeats, log=lambda s: print(s, flush=True))
    with open(args.out, "w") as f:
        json.dump({"workload": args.workload, "options": best,
                   "table": [{"options": o, "fill_ms": fm, "wall_s": w} for o, fm, w in table]},
                  f, indent=1)
    print(f"best {best} -> {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
