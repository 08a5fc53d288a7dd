"""Train the surrogate network for the C2 workload family.

    python -m triplelens_amd.train_surrogate --out triplelens_amd/data/surrogate_c2.npz

Training data are ray-shot by the engine itself (batched get_lc_irs with the
disk samples on): C2-distribution geometries (separations log-uniform
[0.5, 2], m2, m3 log-uniform [1e-3, 0.3], rho = 5e-4, u = 0.5) with
  * near-caustic curves: 512 positions over 100 source radii about the
    central-caustic crossing, where the finite-source corrections are, and
  * full C2 curves: --full-points positions (default 1024; 4096 is the C2
    trajectory itself) over the whole length-2 trajectory.
Geometry seeds 10000 and up are used for training, 20000 and up for the
calibration of the gate and 30000 and up for the held-out test set, so the
benchmark geometries (small seeds) stay out.  The network is trained on the
positions the gate can accept, then sigma_max is calibrated on the separate
calibration curves (Surrogate.calibrate).  With --test-zoom/--test-full the
held-out curves are ray-shot too and --report receives their errors against
ray shooting (the engine's FP64 IRS, whose ray counts are the reference's):
the network's over all positions and over the positions its gate accepts,
beside the hexadecapole estimate and the point-source magnification.
Progress lines go to stdout.

The 2M-position configuration (about 2.1M training, 0.1M calibration and
0.2M held-out positions):
    python -m triplelens_amd.train_surrogate --out MODEL --zoom 1024 --full 384 \
        --full-points 4096 --calib-zoom 64 --calib-full 16 --test-zoom 128 \
        --test-full 33 --steps 60000 --report REPORT.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import engine as E
from . import surrogate as S
from . import workloads as W


def c2_curves(seeds, n_points, ppr, zoom):
    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, ppr).cc_ca()[1]

    out = [W.c2(seed=s, n_points=n_points, ppr=ppr, ca_fn=ca_fn) for s in seeds]
    return [S.caustic_zoom(c, 0.025) for c in out] if zoom else out


def ray_shot(seeds, n_points, ppr, zoom, label, t0, chunk=16):
    Xs, ys, cs = [], [], []
    seeds = list(seeds)
    for k in range(0, len(seeds), chunk):
        X, y, raw = S.training_set(c2_curves(seeds[k:k + chunk], n_points, ppr, zoom))
        Xs.append(X)
        ys.append(y)
        cs.append(raw["cross"])
        print(f"{label}: {min(k + chunk, len(seeds))}/{len(seeds)} curves, "
              f"{sum(len(v) for v in ys)} positions, {time.time() - t0:.0f} s", flush=True)
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(cs)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--ppr", type=int, default=200)
    ap.add_argument("--zoom", type=int, default=192, help="near-caustic training curves")
    ap.add_argument("--full", type=int, default=64, help="full-length training curves")
    ap.add_argument("--full-points", type=int, default=1024,
                    help="positions of the full-length curves (training and calibration)")
    ap.add_argument("--calib-zoom", type=int, default=48)
    ap.add_argument("--calib-full", type=int, default=16)
    ap.add_argument("--test-zoom", type=int, default=0, help="held-out near-caustic curves")
    ap.add_argument("--test-full", type=int, default=0, help="held-out full-length curves")
    ap.add_argument("--report", default="", help="held-out errors (.json)")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--gate", type=float, default=2.0, help="gate radius in source radii")
    ap.add_argument("--bound", type=float, default=1e-3)
    ap.add_argument("--safety", type=float, default=2.0)
    ap.add_argument("--hex-max", type=float, default=1e-3,
                    help="largest hexadecapole term / A_point the gate accepts")
    ap.add_argument("--save-data", default="", help="also write the ray-shot data (.npz)")
    ap.add_argument("--load-data", default="",
                    help="train on ray-shot data written by --save-data instead of shooting")
    args = ap.parse_args(argv)
    t0 = time.time()
    if args.load_data:
        z = np.load(args.load_data, allow_pickle=False)
        X, y, cr, Xc, yc, crc = z["X"], z["y"], z["cross"], z["Xc"], z["yc"], z["cross_c"]
    else:
        fp = args.full_points
        parts = [ray_shot(range(10000, 10000 + args.zoom), 512, args.ppr, True, "train near", t0),
                 ray_shot(range(11000, 11000 + args.full), fp, args.ppr, False, "train full",
                          t0)]
        calib = [ray_shot(range(20000, 20000 + args.calib_zoom), 512, args.ppr, True,
                          "calib near", t0),
                 ray_shot(range(21000, 21000 + args.calib_full), fp, args.ppr, False,
                          "calib full", t0)]
        X, y, cr = (np.concatenate(v) for v in zip(*parts))
        Xc, yc, crc = (np.concatenate(v) for v in zip(*calib))
    if args.save_data:
        np.savez(args.save_data, X=X, y=y, cross=cr, Xc=Xc, yc=yc, cross_c=crc)
    sur = S.Surrogate(S.Net(seed=1), gate_radii=args.gate, hex_max=args.hex_max)
    losses = np.zeros(0)
    done = 0
    while done < args.steps:
        n = min(5000, args.steps - done)
        # one Adam run per slice: the optimiser state persists in the network
        if done == 0:
            losses = sur.fit(X, y, cr, steps=n, batch=args.batch, lr=args.lr, seed=2)
        else:
            keep = sur.trainable(X, y, cr)
            losses = sur.net.train(X[keep], y[keep] - X[keep, S.F_HEX], steps=n,
                                   batch=args.batch, lr=args.lr, seed=2 + done)
        done += n
        print(f"train: {done}/{args.steps} steps, loss {losses[-200:].mean():.3f}, "
              f"{time.time() - t0:.0f} s", flush=True)
    smax = sur.calibrate(Xc, yc, crc, bound=args.bound, safety=args.safety)
    mu, sigma = sur.predict(Xc)
    acc = sur.accepts(Xc, crc, sigma)
    err = np.abs(np.expm1(mu - yc))
    print(f"calibration: sigma_max {np.array2string(smax, formatter={'float': '{:.3e}'.format})} per band "
          f"(edges {sur.band_edges} rho), accepts {acc.mean():.3f} of {len(yc)} positions, "
          f"max rel err {err[acc].max(initial=0):.2e}", flush=True)
    sur.save(args.out)
    print(f"wrote {args.out}", flush=True)
    if args.test_zoom or args.test_full:
        test = [ray_shot(range(30000, 30000 + args.test_zoom), 512, args.ppr, True,
                         "test near", t0),
                ray_shot(range(31000, 31000 + args.test_full), args.full_points, args.ppr,
                         False, "test full", t0)]
        Xt, yt, crt = (np.concatenate(v) for v in zip(*test))
        rep = held_out_report(sur, Xt, yt, crt)
        rep.update({"train_positions": int(len(y)), "calibration_positions": int(len(yc)),
                    "steps": args.steps, "batch": args.batch, "ppr": args.ppr,
                    "seconds": round(time.time() - t0, 1)})
        print("held-out:", json.dumps(rep), flush=True)
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rep, f, indent=1)
    return 0


def held_out_report(sur, X, y, cross):
    """Errors on held-out positions against ray shooting: the network (all
    positions; the gate's accepted ones), the hexadecapole estimate and the
    point-source magnification."""
    mu, sigma = sur.predict(X)
    err = np.abs(np.expm1(mu - y))
    acc = sur.accepts(X, cross, sigma)
    err_hex = np.abs(np.expm1(X[:, S.F_HEX] - y))
    err_pt = np.abs(np.expm1(-y))
    cand = sur.candidates(X, cross)

    def q(e, sel=None):
        e = e if sel is None else e[sel]
        if len(e) == 0:
            return None
        return {k: float(v) for k, v in zip(("p50", "p90", "p99", "max"),
                                              np.percentile(e, [50, 90, 99, 100]))}

    return {"held_out_positions": int(len(y)),
            "network_all": q(err), "network_gate_candidates": q(err, cand),
            "gate_accepts": float(acc.mean()), "network_accepted": q(err, acc),
            "accepted_within_1e-3": float(np.mean(err[acc] <= 1e-3)) if acc.any() else None,
            "hexadecapole_gate_candidates": q(err_hex, cand),
            "point_source_all": q(err_pt), "sigma_max": sur.sigma_max.tolist()}


if __name__ == "__main__":
    raise SystemExit(main())
