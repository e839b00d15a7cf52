"""The extension raster modes on the GPU against the oracle: cull modes (back, the reference's;
none; front), the opt-in top-left fill rule and the depth-only pass (no pixel shader).

Bar: visibility, depth and every stat bit-exact, colour within 1/255; a rectangle tiled by
triangles of both windings with every vertex on a pixel corner is covered exactly once per
corner of its half-open extent under the top-left rule (the oracle's own pin, checked here on
the GPU's visibility buffer).
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle
from tests.test_oracle_pins import passthrough_vertices

pytestmark = pytest.mark.gpu

N = 64


def _tiled(seed, x0, y0, x1, y1, step, flat_depths=None):
    """Squares of `step` pixels split into two triangles each (either diagonal, either winding),
    corners on pixel corners; depths random per vertex, or one of `flat_depths` per triangle."""
    rng = np.random.default_rng(seed)
    verts = []
    for y in range(y0, y1, step):
        for x in range(x0, x1, step):
            a, b, c, d = (x, y), (x + step, y), (x + step, y + step), (x, y + step)
            halves = [(a, b, c), (a, c, d)] if rng.integers(2) else [(a, b, d), (b, c, d)]
            for h in halves:
                h = [(16 * p[0], 16 * p[1]) for p in h]
                h = h if rng.integers(2) else [h[0], h[2], h[1]]
                z = tuple(rng.uniform(0.2, 0.8, 3)) if flat_depths is None else (float(rng.choice(flat_depths)),) * 3
                v = passthrough_vertices(h, z=z, size=N)
                v[:, 4:7] = rng.uniform(0, 1, 3)
                verts.append(v)
    v = np.concatenate(verts)
    return v, np.arange(len(v), dtype=np.uint32)


@pytest.mark.parametrize("cull,fill", [(L.CULL_NONE, L.FILL_TOP_LEFT), (L.CULL_NONE, L.FILL_INCLUSIVE),
                                       (L.CULL_FRONT, L.FILL_TOP_LEFT), (L.CULL_BACK, L.FILL_TOP_LEFT)])
def test_tiled_rectangle(device, oracle, cull, fill):
    x0, y0, x1, y1 = 5, 3, 57, 59  # multiples of the step apart
    v, idx = _tiled(21, x0, y0, x1, y1, 4)
    d = DrawSpec(idx, v, "passthrough", "passthrough", cull=cull, fill=fill)
    print(f"case: tiled rectangle cull={cull} fill={fill}, {len(idx) // 3} triangles", flush=True)
    g = render_gpu(device, N, N, [d], (0, 0, 0, 0))
    o = render_oracle(oracle, N, N, [d], (0, 0, 0, 0))
    assert_frames_match(g, o, color_tol=0)
    if cull == L.CULL_NONE:
        yy, xx = np.mgrid[0:N, 0:N]
        want = ((xx >= x0) & (xx < x1) & (yy >= y0) & (yy < y1)) if fill == L.FILL_TOP_LEFT else \
               ((xx >= x0) & (xx <= x1) & (yy >= y0) & (yy <= y1))
        np.testing.assert_array_equal(g.vis_tri != 0xFFFFFFFF, want)


@pytest.mark.parametrize("depth_func,fill", [(0, L.FILL_INCLUSIVE), (1, L.FILL_INCLUSIVE), (0, L.FILL_TOP_LEFT),
                                             (1, L.FILL_TOP_LEFT)])
def test_tiny_triangle_layers(device, oracle, depth_func, fill):
    """Three draws tiling the target with 2-pixel triangles, each triangle at one of three flat
    depths (so equal depths meet and the later triangle must win them): every bin walks long runs
    of survivors with a handful of pixel corners each, which the tile kernel rasterises one
    triangle per lane and merges per pixel."""
    draws = []
    for layer in range(3):
        v, idx = _tiled(100 + layer, 0, 0, N, N, 2, flat_depths=(0.3, 0.5, 0.7))
        draws.append(DrawSpec(idx, v, "passthrough", "passthrough", cull=L.CULL_NONE, fill=fill))
    clear = 0.0 if depth_func == 0 else 1.0
    gpu_func = L.DEPTH_GREATER_EQUAL if depth_func == 0 else L.DEPTH_LESS_EQUAL
    print(f"case: tiny triangle layers depth_func={depth_func} fill={fill}, 3 x {len(draws[0].indices) // 3} "
          f"triangles", flush=True)
    g = render_gpu(device, N, N, draws, (0, 0, 0, 0), clear_depth=clear, depth_func=gpu_func)
    o = render_oracle(oracle, N, N, draws, (0, 0, 0, 0), clear_depth=clear, depth_func=depth_func)
    assert_frames_match(g, o, color_tol=0)
    assert (g.vis_tri != 0xFFFFFFFF).all()


@pytest.mark.parametrize("depth_func", [0, 1])
@pytest.mark.parametrize("path", ["heavy", "wide"])
def test_tiny_triangles_heavy_and_wide_tiles(device, oracle, depth_func, path):
    """One draw of 2-pixel triangles at three flat depths (every tile list holds one draw), walked
    by the heavy-tile path (four waves per bin, their keys merged) or by 32x32 tiles, with the
    triangle-parallel merge in both."""
    v, idx = _tiled(200, 0, 0, N, N, 2, flat_depths=(0.3, 0.5, 0.7))
    d = DrawSpec(idx, v, "passthrough", "passthrough", cull=L.CULL_NONE)
    clear = 0.0 if depth_func == 0 else 1.0
    gpu_func = L.DEPTH_GREATER_EQUAL if depth_func == 0 else L.DEPTH_LESS_EQUAL
    keep = device.tuning()
    t = dict(keep, heavy_tile_pairs=0, tile_edge=16) if path == "heavy" else \
        dict(keep, heavy_tile_pairs=0xFFFFFFFF, tile_edge=32)
    print(f"case: tiny triangles {path} depth_func={depth_func}, {len(idx) // 3} triangles", flush=True)
    try:
        device.set_tuning(**t)
        g = render_gpu(device, N, N, [d], (0, 0, 0, 0), clear_depth=clear, depth_func=gpu_func)
    finally:
        device.set_tuning(**keep)
    o = render_oracle(oracle, N, N, [d], (0, 0, 0, 0), clear_depth=clear, depth_func=depth_func)
    assert_frames_match(g, o, color_tol=0)


@pytest.mark.parametrize("cull,fill", [(L.CULL_FRONT, L.FILL_INCLUSIVE), (L.CULL_NONE, L.FILL_TOP_LEFT)])
def test_sphere_modes(device, oracle, cull, fill):
    W, H = 160, 96
    mesh, cb, tex = S.c1_scene(W, H)
    d = DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex, cull=cull, fill=fill)
    print(f"case: sphere cull={cull} fill={fill}", flush=True)
    g = render_gpu(device, W, H, [d], S.CLEAR_COLOR)
    o = render_oracle(oracle, W, H, [d], S.CLEAR_COLOR)
    assert (o.vis_tri != 0xFFFFFFFF).sum() > 1000
    assert_frames_match(g, o, color_tol=1)


@pytest.mark.parametrize("order", [("passthrough", "none", "passthrough"), ("none", "passthrough", "none"),
                                   ("none", "none", "none")])
def test_depth_only_pass(device, oracle, order):
    """Three overlapping draws, some without a pixel shader (a depth-only draw after a shaded one,
    and before one): colour, depth, visibility and stats bit-exact against the oracle (whose
    depth-only pass is pinned on the CPU)."""
    from tests.test_oracle_raster_modes import _depth_only_draws
    W, H = 48, 40
    draws = _depth_only_draws(order[1], (order[0], order[2]))
    print(f"case: depth-only draws {order}", flush=True)
    g = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    assert_frames_match(g, o, color_tol=0)
    assert (o.vis_draw == 1).sum() > 100  # the depth-only draw wins pixels
