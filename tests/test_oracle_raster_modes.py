"""Pins of the oracle's extension raster modes: the cull modes (the reference culls back faces
only, main.c:856) and the opt-in D3D top-left fill rule (the reference's rule is inclusive,
main.c:1015-1028, with the top-left rule a TODO at main.c:1027).

- Top-left: a rectangle of pixel corners [x0, x1] x [y0, y1] tiled by triangles of both windings,
  all vertices on pixel corners (so many corners lie exactly on shared edges): under the top-left
  rule every corner with x0 <= x < x1 and y0 <= y < y1 is covered exactly once and no other corner
  is covered; under the inclusive rule the closed rectangle is covered, shared-edge corners more
  than once.
- Cull modes: back + front assembled counts equal none's plus the zero-area triangles (kept by
  every mode), and per pixel corner the coverage counts of back and front add up to none's.
"""
from __future__ import annotations

import numpy as np

from tests.test_oracle_pins import W32, orient, passthrough_vertices


def _draw(oracle, tris, cull, fill):
    verts = np.concatenate([passthrough_vertices(t) for t in tris])
    return oracle.Draw(np.arange(3 * len(tris), dtype=np.uint32), verts, oracle.VS_PASSTHROUGH,
                       oracle.PS_PASSTHROUGH, viewport=(0, 0, W32, W32, 0, 1), cull_mode=cull, fill_rule=fill)


def _counts(oracle, tris, cull, fill):
    """Per pixel corner (y, x) of the 32x32 target: how many assembled triangles cover it."""
    tr = oracle.primitive_assembly_array(_draw(oracle, tris, cull, fill))
    c = np.zeros((W32, W32), np.int64)
    for t in tr:
        for oy in range(0, W32, 8):
            for ox in range(0, W32, 8):
                m = oracle.coverage_mask(t["edge"], ox, oy, int(t["excl"]))
                for b in range(64):
                    if (m >> b) & 1:
                        c[oy + b // 8, ox + b % 8] += 1
    return c, len(tr)


def _tiling(rng, x0, y0, x1, y1, step):
    """Triangles (28.4 points) tiling [x0, x1] x [y0, y1] (pixel corners): each step x step cell
    split by a random diagonal, each triangle with a random winding."""
    tris = []
    for y in range(y0, y1, step):
        for x in range(x0, x1, step):
            a, b, c, d = (x, y), (x + step, y), (x + step, y + step), (x, y + step)
            halves = [(a, b, c), (a, c, d)] if rng.integers(2) else [(a, b, d), (b, c, d)]
            for h in halves:
                h = [(16 * p[0], 16 * p[1]) for p in h]
                tris.append(h if rng.integers(2) else [h[0], h[2], h[1]])
    return tris


def test_top_left_rule_partitions_a_tiled_rectangle(oracle):
    rng = np.random.default_rng(5)
    x0, y0, x1, y1 = 4, 3, 28, 27
    tris = _tiling(rng, x0, y0, x1, y1, 4)
    yy, xx = np.mgrid[0:W32, 0:W32]
    half_open = (xx >= x0) & (xx < x1) & (yy >= y0) & (yy < y1)
    closed = (xx >= x0) & (xx <= x1) & (yy >= y0) & (yy <= y1)
    c, n = _counts(oracle, tris, oracle.CULL_NONE, oracle.FILL_TOP_LEFT)
    assert n == len(tris)
    np.testing.assert_array_equal(c, half_open.astype(np.int64))
    ci, _ = _counts(oracle, tris, oracle.CULL_NONE, oracle.FILL_INCLUSIVE)
    np.testing.assert_array_equal(ci > 0, closed)
    assert ci.max() > 1  # shared edges are covered by both of their triangles


def test_cull_modes_split_the_triangles(oracle):
    rng = np.random.default_rng(9)
    tris = []
    for _ in range(300):
        pts = [tuple(int(v) for v in rng.integers(0, 16 * W32 + 1, 2)) for _ in range(3)]
        if rng.random() < 0.05:  # some zero-area triangles
            pts[2] = pts[1]
        tris.append(pts)
    zero = sum(1 for t in tris if orient(*t) == 0)
    assert zero > 0
    cb, nb = _counts(oracle, tris, oracle.CULL_BACK, oracle.FILL_INCLUSIVE)
    cf, nf = _counts(oracle, tris, oracle.CULL_FRONT, oracle.FILL_INCLUSIVE)
    cn, nn = _counts(oracle, tris, oracle.CULL_NONE, oracle.FILL_INCLUSIVE)
    assert nn == len(tris) and nb + nf == nn + zero
    np.testing.assert_array_equal(cb + cf, cn)
    # the reference's mode keeps exactly the triangles of signed area <= 0 (main.c:856)
    kept = [t for t in tris if orient(*t) <= 0]
    ck, nk = _counts(oracle, kept, oracle.CULL_NONE, oracle.FILL_INCLUSIVE)
    assert nk == nb
    np.testing.assert_array_equal(ck, cb)


def _depth_only_draws(ps_b, ps_ac=("passthrough", "passthrough")):
    """Three overlapping draws of the binning stress (1500 small and a few large triangles each,
    interleaved in depth); the middle one's pixel shader is ps_b, the others' ps_ac."""
    from malevich_amd import scenes as S
    from tests.render_utils import DrawSpec
    out = []
    for k, ps in enumerate((ps_ac[0], ps_b, ps_ac[1])):
        st = S.binning_stress(1500, 2, 48, 40, seed=70 + k, big_every=500)
        out.append(DrawSpec(st.indices, st.vertices, "passthrough", ps))
    return out


def test_depth_only_pass(oracle):
    """The depth-only extension (no pixel shader): depth, visibility and stats are those of the
    same draws with a shader, and the colour target is what the other draws leave: [A, B none]
    has [A]'s colours and [A, B]'s depth and visibility."""
    from malevich_amd import scenes as S
    from tests.render_utils import render_oracle
    W, H = 48, 40
    shaded = _depth_only_draws("passthrough")
    depth_only = _depth_only_draws("none")
    two_shaded = render_oracle(oracle, W, H, shaded[:2], S.CLEAR_COLOR)
    two_none = render_oracle(oracle, W, H, depth_only[:2], S.CLEAR_COLOR)
    first = render_oracle(oracle, W, H, shaded[:1], S.CLEAR_COLOR)
    np.testing.assert_array_equal(two_none.colors, first.colors)
    np.testing.assert_array_equal(two_none.depth.view(np.uint32), two_shaded.depth.view(np.uint32))
    np.testing.assert_array_equal(two_none.vis_tri, two_shaded.vis_tri)
    np.testing.assert_array_equal(two_none.vis_draw, two_shaded.vis_draw)
    assert two_none.stats == two_shaded.stats
    assert (two_none.vis_draw == 1).sum() > 100  # the depth-only draw (tag 1) wins pixels
    assert not np.array_equal(two_shaded.colors, first.colors)
