"""Footprint counting (the measurement mode behind the roofline's algorithmic bytes).

Bar (written here): with counting on, the frame is bit-identical to the frame without it; the
visible-triangle count equals the number of distinct (draw, triangle) winners in the visibility
buffer, which equals the oracle's; the pixels-written count equals the covered pixels; the
distinct texels of a screen-filling quad sampled one texel per pixel (level 0) or two (level 1)
equal the bilinear taps' set computed in closed form (floor(w u - 1/2) and the next texel,
clamped), in count and bytes.
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import scenes as S
from oracle import mips as M
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def test_counts_match_the_visibility_buffer(device, oracle):
    W, H = 320, 184
    sc = S.c2_scene(W, H, n_instances=10, seed=31, texture_size=0)
    irr, rad, albedo = S.c2_textures(n_instances=10, seed=31, albedo_size=64)
    rad = rad[::8, ::8].copy()
    rad_chain = M.chain(rad)
    draws = [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_textured", cb=sc.cb, vs_tex=irr,
                      ps_tex=rad_chain, ps_tex1=M.chain(a)) for m, a in zip(sc.instances, albedo)]
    print(f"case: footprint counts {W}x{H}, {len(draws)} draws", flush=True)
    plain = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
    device.set_footprint_counting(True)
    try:
        device.footprint()
        counted = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
        fp = device.footprint()
    finally:
        device.set_footprint_counting(False)
    for a, b in ((plain.colors, counted.colors), (plain.vis_tri, counted.vis_tri)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(plain.depth.view(np.uint32), counted.depth.view(np.uint32))
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    assert_frames_match(counted, o, color_tol=1)
    cov = o.vis_tri != 0xFFFFFFFF
    winners = {(int(d), int(t)) for d, t in zip(o.vis_draw[cov], o.vis_tri[cov])}
    print(f"  {fp}", flush=True)
    assert fp["visible_triangles"] == len(winners)
    assert fp["shaded_pixels"] == int(cov.sum())
    assert fp["texels"] > 0 and fp["texel_bytes"] >= 4 * fp["texels"]


def _taps(coord, n_tex):
    """Texel indices the bilinear taps of texel coordinates `coord` (= n u - 1/2) read, clamped."""
    f = np.floor(coord).astype(np.int64)
    return set(np.clip(f, 0, n_tex - 1).tolist()) | set(np.clip(f + 1, 0, n_tex - 1).tolist())


@pytest.mark.parametrize("level", [0, 1])
def test_texel_count_of_a_quad(device, level):
    """A quad whose uv advance 2^level texels per pixel reads level `level` only; its distinct texels
    are the product of the column and row tap sets."""
    W = H = 64
    tw = 64
    sizes = M.level_sizes(tw, tw)
    tex = [np.full((h, w), 0x11223344, np.uint32) for w, h in sizes]
    p = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32)
    span = 2 ** level
    uv = np.array([[0, 0], [span, 0], [span, span], [0, span]], np.float32)
    verts = np.concatenate([p, np.full((4, 1), 0.5, np.float32), np.tile([[0, 0, 1]], (4, 1)).astype(np.float32),
                            uv], 1)
    cb = np.zeros(48, np.float32)
    cb[0:16] = np.eye(4, dtype=np.float32).ravel()
    d = DrawSpec(np.array([0, 1, 2, 2, 3, 0, 0, 2, 1, 2, 0, 3], np.uint32), verts, "basic", "basic_lambert", cb=cb,
                 ps_tex=tex, ps_cb=np.array([0, 0, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0], np.float32))
    device.set_footprint_counting(True)
    try:
        device.footprint()
        g = render_gpu(device, W, H, [d], (0, 0, 0, 0))
        fp = device.footprint()
    finally:
        device.set_footprint_counting(False)
    assert (g.vis_tri != 0xFFFFFFFF).all()
    n = sizes[level][0]
    # corners x, y = 0..63 (w = 1: affine): u = span x / W, v = span (1 - y / H); the taps are at
    # s = n u - 1/2 and t = n (1 - v) - 1/2 (never an integer here, so the sets are exact)
    c = np.arange(W, dtype=np.float64)
    cols = _taps(n * (span * c / W) - 0.5, n)
    rows = _taps(n * (1 - span * (1 - c / H)) - 0.5, n)
    print(f"case: quad level {level}: {fp}, expected {len(cols)} x {len(rows)}", flush=True)
    assert fp["shaded_pixels"] == W * H
    assert fp["texels"] == len(cols) * len(rows)
    assert fp["texel_bytes"] == 4 * fp["texels"]
