"""Flood fill of the oracle (lineFloodFill, lcirs.cc:381-466).

What the fill counts is fixed by its semantics: the hit pixels of every
4-connected component of hits that contains a hit seed.  Brute force: label
the hit pixels of a box around the images (scipy.ndimage) and sum the
components that hold a seed; the oracle's fill must count exactly those
pixels, with the uniform-disk magnification an integer multiple of ampScale.
Plus the closed form ampScale = intensity / (pi ppr^2) (lcirs.cc:87,98)."""
import math

import numpy as np
import pytest
from scipy import ndimage

from conftest import pixel_seeds
from triplelens_amd import workloads as W


def _brute_count(lc, grid, px, py, seeds_pix, pad):
    """Hit pixels of the 4-connected hit components holding a hit seed, found
    in a box of +-pad pixels around each seed (columns and rows 1..N-1)."""
    N = int(grid[0])
    kept = set()
    for sx, sy in seeds_pix:
        x0, x1 = max(1, sx - pad), min(N - 1, sx + pad)
        y0, y1 = max(1, sy - pad), min(N - 1, sy + pad)
        if not (x0 <= sx <= x1 and y0 <= sy <= y1):
            continue
        hit = lc.hit_mask(x0, x1, y0, y1, px, py)
        if not hit[sy - y0, sx - x0]:
            continue
        lab, _ = ndimage.label(hit)  # 4-connectivity
        m = lab == lab[sy - y0, sx - x0]
        # the component must lie inside the box (or end at the grid's edge)
        assert not ((m[0].any() and y0 > 1) or (m[-1].any() and y1 < N - 1) or
                    (m[:, 0].any() and x0 > 1) or (m[:, -1].any() and x1 < N - 1)), "box too small"
        ys, xs = np.nonzero(m)
        kept.update(zip((ys + y0).tolist(), (xs + x0).tolist()))
    return len(kept)


@pytest.mark.parametrize("idx", [0, 26, 60, 120])
def test_fill_equals_components_holding_seeds(oracle, idx):
    c = W.c1()
    ppr = 6  # coarse grid: images of a few thousand pixels
    lc = oracle.new(*c.lens(), c.rho, c.lc_len, ppr, 0.0)
    g = lc.grid()
    xy, ns, _ = lc.seeds(*c.traj())
    t = idx / c.lc_len
    px, py = c.sx + (c.ex - c.sx) * t, c.sy + (c.ey - c.sy) * t
    sxy = xy[idx, :ns[idx]]
    sx, sy = pixel_seeds(g, sxy)
    rays, bsum, edge = lc.fill_point(px, py, sxy)
    want = _brute_count(lc, g, px, py, list(zip(sx.tolist(), sy.tolist())), pad=400)
    assert rays == want
    assert not edge
    # uniform disk: brightness 1 per ray
    assert bsum == float(rays)


def test_amp_scale_closed_form(oracle):
    c = W.c1()
    for ppr, v in [(50, 0.0), (7, 0.5), (200, 0.3)]:
        lc = oracle.new(*c.lens(), c.rho, 10, ppr, v)
        assert lc.grid()[7] == 1.0 / math.pi / (ppr * ppr) or \
            abs(lc.grid()[7] - 1.0 / math.pi / ppr ** 2) <= 1e-16 / ppr ** 2
    lc = oracle.new(*c.lens(), c.rho, 10, 50)  # default profile A=0.6, B=0.4
    assert abs(lc.grid()[7] - (1.0 / (0.6 + 0.4 * 2 / 3)) / math.pi / 2500) < 1e-19


def test_missing_seed_and_out_of_grid_rows(oracle):
    c = W.c1()
    lc = oracle.new(*c.lens(), c.rho, c.lc_len, 10, 0.0)
    g = lc.grid()
    # no seeds: nothing counted
    assert lc.fill_point(0.1, 0.0, np.zeros((0, 2)))[0] == 0
    # a seed whose row falls outside 1..N-1 is skipped (lcirs.cc:386-389)
    far = np.array([[0.0, g[4] - 1.0]])
    assert lc.fill_point(0.1, 0.0, far)[0] == 0
