"""Surrogate magnification with ray-shooting fallback.

A small network (include/tl_surrogate.h, kernels in csrc/tl_mlp.hip) predicts,
for a source position, mu = ln(A_finite / A_point) and s = ln(sigma), from
fourteen features:

    a, b, th                      lens separations and angle
    log10(m2/m1), log10(m3/m1)    mass ratios
    log10(rho), u                 source radius and linear limb darkening
    dx, dy                        offset of the magnification-weighted image
                                  centroid from the source position
    d / rho                       signed distance to the nearest caustic
                                  sample (negative inside a caustic), in
                                  source radii, clipped to [-1000, 1000]
    ln A_point                    point-source magnification
    pair share                    share of A_point in the two brightest
                                  images (near a fold: the merging pair)
    h = ln(A_hex / A_point)       Gould's (2008, ApJ 681, 1593) hexadecapole
                                  estimate of the finite-source
                                  magnification, from 12 point
                                  magnifications on the source disk
                                  (lcirs_set_ring) and the exact disk
                                  moments of the limb-darkening profile
    t4                            the hexadecapole term relative to A_point,
                                  log-compressed: large where the expansion
                                  in rho breaks down

A_point, the centroid, d, the pair share and the disk samples come from
the engine's seed search (lcirs_find_seeds), which every position needs
anyway.  The network's mean output is a correction to h (a skip
connection): mu = h + the network's first output, so it learns where and by
how much the expansion in rho fails rather than the expansion itself.  A position is
computed by inverse ray shooting instead when it lies within `gate_radii`
source radii of a caustic, when sigma exceeds `sigma_max`, or when the
prediction is not finite; a random `audit` fraction of the accepted
positions is ray-shot as well and compared.  The network is trained on
ray-shooting output the engine generates itself (training_set).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import engine as _E

N_FEATURES = 14
F_DIST, F_HEX = 9, 12
_P = ctypes.c_void_p


class SurrogateError(RuntimeError):
    pass


def _check_flags(flags):
    bad = int(np.bitwise_or.reduce(flags)) & ~_E.FLAGS_INFO
    if bad:
        raise SurrogateError(f"ray shooting reported status flags {bad:#x}")


def _lib():
    lib = _E.load()
    if getattr(lib, "_tl_mlp_declared", False):
        return lib
    lib.tl_mlp_new.argtypes = [ctypes.c_ulonglong]
    lib.tl_mlp_new.restype = _P
    lib.tl_mlp_delete.argtypes = [_P]
    lib.tl_mlp_param_count.restype = ctypes.c_long
    lib.tl_mlp_get_params.argtypes = [_P, _P]
    lib.tl_mlp_set_params.argtypes = [_P, _P]
    lib.tl_mlp_set_scaling.argtypes = [_P, _P, _P]
    lib.tl_mlp_predict.argtypes = [_P, ctypes.c_int, _P, _P]
    lib.tl_mlp_grads.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P]
    lib.tl_mlp_train.argtypes = [_P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_ulonglong, _P]
    lib.tl_mlp_last_error.restype = ctypes.c_char_p
    lib._tl_mlp_declared = True
    return lib


def _ok(rc):
    if rc != 0:
        raise SurrogateError(_lib().tl_mlp_last_error().decode() or f"surrogate call failed ({rc})")


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


class Net:
    """The network on the GPU."""

    def __init__(self, seed: int = 1):
        lib = _lib()
        self._lib = lib
        self.h = lib.tl_mlp_new(seed)
        if not self.h:
            raise SurrogateError(lib.tl_mlp_last_error().decode())
        self.mean = np.zeros(N_FEATURES, np.float32)
        self.scale = np.ones(N_FEATURES, np.float32)
        # the network works on standardised targets (y - y_mean) / y_scale, so
        # that its BF16 activations carry O(1) values; predict() undoes it
        self.y_mean, self.y_scale = 0.0, 1.0

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self._lib.tl_mlp_delete(h)
            self.h = None

    @staticmethod
    def param_count() -> int:
        return int(_lib().tl_mlp_param_count())

    def params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        _ok(self._lib.tl_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, theta):
        t = _f32(theta)
        if t.size != self.param_count():
            raise ValueError("wrong parameter count")
        _ok(self._lib.tl_mlp_set_params(self.h, t.ctypes.data))

    def set_scaling(self, mean, scale):
        self.mean, self.scale = _f32(mean).copy(), _f32(scale).copy()
        _ok(self._lib.tl_mlp_set_scaling(self.h, self.mean.ctypes.data, self.scale.ctypes.data))

    def fit_scaling(self, X, y=None):
        """Feature standardisation (on the GPU) and, given targets, a robust
        target standardisation (median and scaled median absolute
        deviation), so that the rare large targets do not set the scale."""
        X = np.asarray(X, np.float64)
        mean, sd = X.mean(axis=0), X.std(axis=0)
        # a column that is constant up to rounding (rho of a fixed-rho
        # workload) keeps scale 1: dividing its rounding noise by a ~1e-16
        # deviation would feed the network large meaningless inputs
        const = sd <= 1e-9 * np.maximum(1.0, np.abs(mean))
        self.set_scaling(mean, np.where(const, 1.0, sd))
        if y is not None:
            y = np.asarray(y, np.float64)
            self.y_mean = float(np.median(y))
            mad = float(np.median(np.abs(y - self.y_mean)))
            self.y_scale = max(1.4826 * mad, 1e-7)

    def predict_raw(self, X) -> np.ndarray:
        """The network's outputs for standardised targets."""
        X = _f32(X, (-1, N_FEATURES))
        out = np.zeros((len(X), 2), np.float32)
        _ok(self._lib.tl_mlp_predict(self.h, len(X), X.ctypes.data, out.ctypes.data))
        return out

    def predict(self, X) -> np.ndarray:
        """(mu, ln sigma) per row, in the units of the targets."""
        o = self.predict_raw(X).astype(np.float64)
        return np.stack([self.y_mean + self.y_scale * o[:, 0],
                         o[:, 1] + math.log(self.y_scale)], axis=1)

    def grads(self, X, y):
        """Loss and gradient of the network for raw outputs vs targets y."""
        X = _f32(X, (-1, N_FEATURES))
        y = _f32(y)
        g = np.zeros(self.param_count(), np.float32)
        loss = ctypes.c_float(0.0)
        _ok(self._lib.tl_mlp_grads(self.h, len(X), X.ctypes.data, y.ctypes.data, g.ctypes.data,
                                   ctypes.byref(loss)))
        return float(loss.value), g

    def train(self, X, y, steps: int, batch: int = 512, lr: float = 1e-3, seed: int = 1):
        """Adam on the standardised targets; returns the per-step losses."""
        X = _f32(X, (-1, N_FEATURES))
        y = _f32((np.asarray(y, np.float64) - self.y_mean) / self.y_scale)
        losses = np.zeros(steps, np.float32)
        _ok(self._lib.tl_mlp_train(self.h, len(X), X.ctypes.data, y.ctypes.data, steps, batch,
                                   lr, seed, losses.ctypes.data))
        return losses


# ---------------------------------------------------------------- features
def disk_moments(u: float):
    """<r^2> and <r^4> (in units of rho) over the source disk weighted by the
    linear limb-darkened profile 1 - u (1 - sqrt(1 - r^2))."""
    norm = (1.0 - u) + 2.0 * u / 3.0
    return (((1.0 - u) / 2.0 + u * 4.0 / 15.0) / norm,
            ((1.0 - u) / 3.0 + u * 16.0 / 105.0) / norm)


def hexadecapole(amp, ring, u):
    """Gould's (2008) expansion of the finite-source magnification from the
    centre value and the disk samples of lcirs_set_ring: the four-point
    means at rho/2 (axes), rho (axes) and rho (diagonals) give the rho^2 and
    rho^4 coefficients a2, a4 of the circle mean of A, and the disk mean is
    A0 + a2 <r^2> + a4 <r^4>.  Returns (A_hex, a4 <r^4>)."""
    A0 = np.asarray(amp, np.float64)
    ring = np.asarray(ring, np.float64)

    def mean4(lo):
        # numpy's mean over four columns, summed in the same order (the same
        # values), without the general reduction's per-call cost
        return (((ring[:, lo] + ring[:, lo + 1]) + ring[:, lo + 2]) + ring[:, lo + 3]) / 4.0

    half = mean4(0) - A0
    full = mean4(4) - A0
    diag = mean4(8) - A0
    a2 = (16.0 * half - full) / 3.0
    a4 = 0.5 * (full + diag) - a2
    x2, x4 = disk_moments(u)
    return A0 + a2 * x2 + a4 * x4, a4 * x4


def features(lens, rho, u, pos, centroid, caus, amp, pair, ring) -> np.ndarray:
    """Feature rows for positions pos (n, 2) of one lens (a, b, th, m2, m3)."""
    a, b, th, m2, m3 = lens
    m1 = 1.0 - m2 - m3
    n = len(pos)
    X = np.empty((n, N_FEATURES))
    X[:, 0] = a
    X[:, 1] = b
    X[:, 2] = th
    X[:, 3] = math.log10(m2 / m1)
    X[:, 4] = math.log10(m3 / m1)
    X[:, 5] = math.log10(rho)
    X[:, 6] = u
    X[:, 7] = centroid[:, 0] - pos[:, 0]
    X[:, 8] = centroid[:, 1] - pos[:, 1]
    X[:, F_DIST] = np.clip(caus / rho, -1000.0, 1000.0)
    amp = np.asarray(amp, np.float64)
    safe = np.maximum(amp, 1e-300)
    X[:, 10] = np.log(safe)
    X[:, 11] = pair
    A_hex, t4 = hexadecapole(amp, ring, u)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.log(A_hex / safe)
    X[:, F_HEX] = np.clip(np.nan_to_num(h, nan=5.0, posinf=5.0, neginf=-5.0), -5.0, 5.0)
    r4 = np.nan_to_num(t4 / safe, nan=1.0, posinf=1.0, neginf=-1.0)
    X[:, 13] = np.sign(r4) * np.log10(1.0 + np.abs(r4) * 1e6) / 6.0
    return X


def hex_term(X) -> np.ndarray:
    """|hexadecapole term| / A_point per row, from the compressed feature."""
    return (10.0 ** (6.0 * np.abs(np.asarray(X)[:, 13])) - 1.0) * 1e-6


def _point_features(lc, c):
    """A_point, the feature rows and the disk samples' caustic crossings of
    the last seed search of lc over curve c."""
    ap, cen, caus = lc.point_amp()
    ring, _, cross = lc.ring_amp(with_roots=False)
    X = features(c.lens(), c.rho, c.u, positions(c), cen, caus, ap, lc.pair_share(), ring)
    return ap, X, cross


def caustic_zoom(c, half_length: float):
    """Curve c with its trajectory shortened to 2 half_length about its
    midpoint (for the C2 workload: the central-caustic crossing), so that
    most positions lie within a few tens of source radii of the caustic."""
    from .workloads import Curve
    mx, my = 0.5 * (c.sx + c.ex), 0.5 * (c.sy + c.ey)
    dx, dy = c.ex - c.sx, c.ey - c.sy
    nrm = math.hypot(dx, dy)
    dx, dy = dx / nrm * half_length, dy / nrm * half_length
    return Curve(c.a, c.b, c.th, c.m2, c.m3, c.rho, c.ppr, c.u, mx - dx, my - dy, mx + dx, my + dy,
                 c.n_points)


def positions(c) -> np.ndarray:
    t = np.arange(c.n_points) / c.lc_len
    return np.stack([c.sx + (c.ex - c.sx) * t, c.sy + (c.ey - c.sy) * t], axis=1)


def _new_curve(c):
    lc = _E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_limb_darkening("linear", c.u)
    lc.set_ring(True)
    return lc


def training_set(curves):
    """Ray-shot curves (one batched engine run) -> X (n, 14), y = ln(A/A_point),
    and the raw arrays (A_irs, A_point, caustic distance / rho, disk samples
    across a caustic)."""
    lcs = [_new_curve(c) for c in curves]
    res = _E.batch_irs(lcs, np.array([c.traj() for c in curves]))
    Xs, ys, A, Ap, d, cr = [], [], [], [], [], []
    for c, lc, (amp, rays, flags) in zip(curves, lcs, res):
        ap, X, cross = _point_features(lc, c)
        ok = np.isfinite(ap) & (ap > 0) & (amp > 0)
        cr.append(cross[ok])
        Xs.append(X[ok])
        ys.append(np.log(amp[ok] / ap[ok]))
        A.append(amp[ok])
        Ap.append(ap[ok])
        d.append(X[ok, F_DIST])
    return (np.concatenate(Xs), np.concatenate(ys),
            {"A": np.concatenate(A), "A_point": np.concatenate(Ap), "d": np.concatenate(d),
             "cross": np.concatenate(cr)})


def union_seeds(ns1, pix1, ns2, pix2):
    """Per position, the seed pixels of either search, each once (first
    search's order, then the second's new ones), at most 64."""
    n = len(ns1)
    m = _E.MAX_SEEDS
    ns1 = np.asarray(ns1, np.int64)
    ns2 = np.asarray(ns2, np.int64)
    # every seed as (position, rank, x, y); rank orders the first search's
    # seeds before the second's
    slot = np.arange(m)
    take1 = slot[None, :] < ns1[:, None]
    take2 = slot[None, :] < ns2[:, None]
    pos = np.concatenate([np.nonzero(take1)[0], np.nonzero(take2)[0]])
    rank = np.concatenate([np.nonzero(take1)[1], m + np.nonzero(take2)[1]])
    xy = np.concatenate([np.asarray(pix1)[take1], np.asarray(pix2)[take2]]).astype(np.int64)
    ns = np.zeros(n, np.int32)
    pix = np.zeros((n, m, 2), np.int32)
    if len(pos) == 0:
        return ns, pix
    order = np.lexsort((rank, xy[:, 1], xy[:, 0], pos))
    pos, rank, xy = pos[order], rank[order], xy[order]
    first = np.ones(len(pos), bool)
    first[1:] = (pos[1:] != pos[:-1]) | np.any(xy[1:] != xy[:-1], axis=1)
    pos, rank, xy = pos[first], rank[first], xy[first]
    order = np.lexsort((rank, pos))
    pos, xy = pos[order], xy[order]
    start = np.searchsorted(pos, np.arange(n))
    k = np.arange(len(pos)) - start[pos]
    keep = k < m
    pix[pos[keep], k[keep]] = xy[keep]
    ns[:] = np.minimum(np.bincount(pos, minlength=n), m)
    return ns, pix


class Surrogate:
    """Network plus gate: surrogate where it is trusted, ray shooting
    elsewhere and on a random audit sample.  The gate accepts a position
    when it lies at least gate_radii source radii from the caustic samples,
    no disk sample lies across a caustic from it, the hexadecapole term is
    at most hex_max of A_point (the expansion in rho has converged), and the
    network's sigma is at most the sigma_max of the position's distance band
    (bands split at band_edges source radii).  Near the caustics sigma has
    to separate the positions where the expansion fails; far from them it
    mostly reflects the ray-shooting noise of low magnifications, so each
    band gets its own calibrated threshold."""

    def __init__(self, net: Net, gate_radii: float = 2.0, sigma_max=3e-5,
                 audit: float = 0.01, seed_piece: int | None = 16, hex_max: float = 1e-3,
                 band_edges=(5.0,)):
        self.net = net
        self.gate_radii = gate_radii
        self.hex_max = hex_max
        self.band_edges = np.asarray(band_edges, np.float64)
        self.sigma_max = np.broadcast_to(np.asarray(sigma_max, np.float64),
                                         (len(self.band_edges) + 1,)).copy()
        self.audit = audit
        # an audited position further than this (relative) from its ray-shot
        # value fails the audit: the whole curve is then ray-shot
        self.audit_bound = 1e-3
        # shot positions are filled from the union of four split searches;
        # positions per piece of the three extra ones (None: seed_piece)
        self.seed_union = True
        self.extra_piece = None
        # positions per independent piece of the seed search
        # (lcirs_set_seed_split); None keeps the reference's single chain.
        # On the C2 curve, pieces of 64/32/16/8 positions take 13/6.7/4.8/2.8
        # ms and give ray counts other than the single chain's at 4/12/16/19
        # of 4096 positions.
        self.seed_piece = seed_piece

    def seed_parts_extra(self, n_positions: int) -> list:
        """Pieces of the extra seed searches of the shot positions: P+1, P-1
        and P+3 for the first search's P, so that the searches share no
        piece boundary but the curve's start and every position is reached
        by different chain histories."""
        p = self.seed_parts(n_positions)
        if self.extra_piece:
            p = max(1, min(65536, n_positions // self.extra_piece))
        first = self.seed_parts(n_positions)
        return [q for q in (p + 1, p - 1, p + 3) if q >= 1 and q != first]

    def seed_parts(self, n_positions: int) -> int:
        if not self.seed_piece:
            return 1
        return max(1, min(65536, n_positions // self.seed_piece))

    def trainable(self, X, y, cross, margin: float = 0.5):
        """Rows near what the gate can accept (|d| / rho at least
        gate_radii - margin, no disk sample across a caustic, hexadecapole
        term at most 3 hex_max): the rows to train on."""
        return (np.abs(X[:, F_DIST]) >= self.gate_radii - margin) & (np.asarray(cross) == 0) & \
            (hex_term(X) <= 3.0 * self.hex_max) & np.isfinite(y) & np.isfinite(X).all(axis=1)

    def fit(self, X, y, cross, steps: int, batch: int = 512, lr: float = 1e-3, seed: int = 1):
        """Train the network on the trainable rows, for the residual y - h of
        the hexadecapole estimate.  Returns the per-step losses."""
        X = np.asarray(X, np.float64)
        y = np.asarray(y, np.float64)
        keep = self.trainable(X, y, cross)
        Xk, rk = X[keep], y[keep] - X[keep, F_HEX]
        if len(rk) < batch:
            raise ValueError("fewer trainable rows than one batch")
        self.net.fit_scaling(Xk, rk)
        return self.net.train(Xk, rk, steps=steps, batch=batch, lr=lr, seed=seed)

    def save(self, path: str):
        """Weights, scalings and gate in one .npz file."""
        n = self.net
        np.savez(path, params=n.params(), mean=n.mean, scale=n.scale,
                 y=np.array([n.y_mean, n.y_scale]),
                 gate=np.array([self.gate_radii, self.audit, float(self.seed_piece or 0),
                                self.hex_max]),
                 band_edges=self.band_edges, sigma_max=self.sigma_max)

    @classmethod
    def load(cls, path: str) -> "Surrogate":
        z = np.load(path, allow_pickle=False)
        net = Net()
        net.set_params(z["params"])
        net.set_scaling(z["mean"], z["scale"])
        net.y_mean, net.y_scale = float(z["y"][0]), float(z["y"][1])
        g = z["gate"]
        return cls(net, gate_radii=float(g[0]), sigma_max=z["sigma_max"], audit=float(g[1]),
                   seed_piece=int(g[2]) or None, hex_max=float(g[3]),
                   band_edges=z["band_edges"])

    def band(self, X) -> np.ndarray:
        return np.searchsorted(self.band_edges, np.abs(np.asarray(X)[:, F_DIST]), side="right")

    def candidates(self, X, cross) -> np.ndarray:
        """Rows the gate considers (everything but sigma)."""
        return (np.abs(X[:, F_DIST]) >= self.gate_radii) & (np.asarray(cross) == 0) & \
            (hex_term(X) <= self.hex_max)

    def accepts(self, X, cross, sigma) -> np.ndarray:
        """The gate's rows (before the audit)."""
        return self.candidates(X, cross) & (sigma <= self.sigma_max[self.band(X)])

    def calibrate(self, X, y, cross, bound: float = 1e-3, safety: float = 2.0) -> np.ndarray:
        """Set sigma_max, per distance band, from calibration rows the
        network was not trained on: the largest threshold at which every
        calibration row of the band that the gate accepts is within
        bound / safety relative error.  Returns sigma_max."""
        X = np.asarray(X, np.float64)
        y = np.asarray(y, np.float64)
        mu, sigma = self.predict(X)
        err = np.abs(np.expm1(mu - y))
        cand = self.candidates(X, cross) & np.isfinite(err) & np.isfinite(sigma)
        bands = self.band(X)
        for b in range(len(self.sigma_max)):
            sel = cand & (bands == b)
            sig, e = sigma[sel], err[sel]
            order = np.argsort(sig, kind="stable")
            # the running maximum is non-decreasing, so the rows within the
            # bound are a prefix of the sigma order
            worst = np.maximum.accumulate(e[order])
            k = int(np.count_nonzero(worst <= bound / safety)) - 1
            self.sigma_max[b] = float(sig[order[k]]) if k >= 0 else 0.0
        return self.sigma_max.copy()

    def predict(self, X):
        """(mu, sigma) per row: mu = h + the network's correction."""
        X = np.asarray(X, np.float64)
        p = self.net.predict(X)
        return X[:, F_HEX] + p[:, 0], np.exp(np.clip(p[:, 1], -30.0, 30.0))

    def _extra_clones(self, lc, c):
        """Clones of engine object lc for the extra seed searches, made once
        and kept with it, set to their piece counts without disk samples."""
        extra = self.seed_parts_extra(c.n_points)
        clones = getattr(lc, "_seed_clones", None)
        if clones is None or len(clones) != len(extra):
            clones = [lc.clone() for _ in extra]
            lc._seed_clones = clones
        for cl, parts in zip(clones, extra):
            cl.set_ring(False)
            cl.set_seed_split(parts)
        return clones

    def light_curve(self, lc, c, rng=None, stats=None):
        """Magnification at the lcLength+1 positions of curve c on engine object
        lc.  Returns (A, info) with info = fallback, audited masks, the
        audit's relative errors and audit_failed.  Audited positions carry
        their ray-shot value.  When an audit misses the bound (audit_bound),
        every remaining position of the curve is ray-shot as well (and
        marked fallback), so the curve returned is plain ray shooting.  With
        a list `stats`, the engine statistics (engine.last_stats) of each
        engine run are appended to it."""
        return self.light_curves([lc], [c], rng, stats)[0]

    def light_curves(self, lcs, curves, rng=None, stats=None):
        """light_curve of several curves (engine objects lcs[k] over curves[k])
        with each engine step shared: one seed launch for all of them, one
        network call, one fill of every shot position.  Returns a list of
        (A, info)."""
        rng = np.random.default_rng(0) if rng is None else rng
        # A split search can miss the seed of a whole image where a caustic
        # trial solve converges differently from its fresh start.  Positions
        # that are ray-shot take the union of the seeds of four split
        # searches with different piece boundaries; a seed only adds a
        # component that is really there (its pixel's ray lands in the
        # source).  The three extra searches run on clones of each object
        # (kept with it) in the same seed launch as the first.
        union = bool(self.seed_union and self.seed_piece)
        objs, trajs, clones_of = [], [], []
        for lc, c in zip(lcs, curves):
            lc.set_ring(True)
            lc.set_seed_split(self.seed_parts(c.n_points))
            cl = self._extra_clones(lc, c) if union else []
            clones_of.append(cl)
            objs += [lc] + cl
            trajs += [c.traj()] * (1 + len(cl))
        _E.batch_find_seeds(objs, np.array(trajs))
        if stats is not None:
            stats.append(_E.last_stats())
        feats = [_point_features(lc, c) for lc, c in zip(lcs, curves)]
        mu_all, sigma_all = self.predict(np.concatenate([f[1] for f in feats]))
        parts, shots, seeds_list, shot_k = [], [], [], []
        o = 0
        for k, ((ap, X, cross), lc) in enumerate(zip(feats, lcs)):
            m = len(ap)
            mu, sigma = mu_all[o:o + m], sigma_all[o:o + m]
            o += m
            A = ap * np.exp(mu)
            fallback = ~self.accepts(X, cross, sigma) | ~np.isfinite(A) | ~(ap > 0)
            audited = ~fallback & (rng.random(len(A)) < self.audit)
            shoot = fallback | audited
            parts.append([A, fallback, audited])
            if shoot.any():
                idx = np.nonzero(shoot)[0]
                sd = lc.seeds_union(clones_of[k], idx)
                shots.append(idx)
                seeds_list.append(sd)
                shot_k.append(k)
        out = [None] * len(lcs)
        failed = []  # curves whose audit missed the bound
        if shots:
            res = _E.batch_fill_positions([lcs[k] for k in shot_k], shots,
                                          seeds_list if union else None)
            if stats is not None:
                stats.append(_E.last_stats())
            for k, idx, (A_irs, rays, flags) in zip(shot_k, shots, res):
                _check_flags(flags)
                A, fallback, audited = parts[k]
                full = np.zeros(len(A))
                full[idx] = A_irs
                errs = np.abs(A[audited] - full[audited]) / full[audited]
                A = np.where(fallback | audited, full, A)
                out[k] = (A, {"fallback": fallback, "audited": audited, "audit_rel_err": errs,
                              "audit_failed": False, "rays": int(rays.sum())})
                if errs.size and not errs.max() <= self.audit_bound:
                    failed.append(k)
        for k in range(len(lcs)):
            if out[k] is None:
                A, fallback, audited = parts[k]
                out[k] = (A, {"fallback": fallback, "audited": audited,
                              "audit_rel_err": np.zeros(0), "audit_failed": False, "rays": 0})
        if failed:
            # the network is off on these curves: ray-shoot everything it
            # answered, from the same seeds as the other shot positions
            rest, rest_seeds = [], []
            for k in failed:
                A, info = out[k]
                idx = np.nonzero(~(info["fallback"] | info["audited"]))[0]
                sd = lcs[k].seeds_union(clones_of[k], idx)
                rest.append(idx)
                rest_seeds.append(sd)
            res = _E.batch_fill_positions([lcs[k] for k in failed], rest,
                                          rest_seeds if union else None)
            if stats is not None:
                stats.append(_E.last_stats())
            for k, idx, (A_irs, rays, flags) in zip(failed, rest, res):
                _check_flags(flags)
                A, info = out[k]
                A = A.copy()
                A[idx] = A_irs
                fb = info["fallback"].copy()
                fb[idx] = True
                info.update(fallback=fb, audit_failed=True, rays=info["rays"] + int(rays.sum()))
                out[k] = (A, info)
        return out
