"""Train the surrogate network for the C2 workload family.

    python -m triplelens_amd.train_surrogate --out triplelens_amd/data/surrogate_c2.npz

Training data are ray-shot by the engine itself (batched get_lc_irs with the
disk samples on): C2-distribution geometries (separations log-uniform
[0.5, 2], m2, m3 log-uniform [1e-3, 0.3], rho = 5e-4, u = 0.5) with
  * near-caustic curves: 512 positions over 100 source radii about the
    central-caustic crossing, where the finite-source corrections are, and
  * full C2 curves: 1024 positions over the whole length-2 trajectory.
Geometry seeds 10000 and up are used for training and 20000 and up for the
calibration of the gate, so the benchmark geometries (small seeds) stay out.
The network is trained on the positions the gate can accept, then sigma_max
is calibrated on the separate calibration curves (Surrogate.calibrate).
Progress lines go to stdout.
"""
from __future__ import annotations

import argparse
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
    ap.add_argument("--calib-zoom", type=int, default=48)
    ap.add_argument("--calib-full", type=int, default=16)
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
        parts = [ray_shot(range(10000, 10000 + args.zoom), 512, args.ppr, True, "train near", t0),
                 ray_shot(range(11000, 11000 + args.full), 1024, args.ppr, False, "train full",
                          t0)]
        calib = [ray_shot(range(20000, 20000 + args.calib_zoom), 512, args.ppr, True,
                          "calib near", t0),
                 ray_shot(range(21000, 21000 + args.calib_full), 1024, args.ppr, False,
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
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
