"""Parity of the gfx950 draw pipeline with the CPU oracle on seeded scenes.

Bar: triangle-id visibility, draw tags and depth bit-exact; RGBA8 exact for
passthrough shading and within 1/255 per channel where libm transcendentals
stand in for SVML (pow/exp/acos); the reference's stats counters exact.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def both(device, oracle, W, H, draws, clear=S.CLEAR_COLOR, depth=0.0, depth_func=0):
    print(f"case: {W}x{H}, {len(draws)} draw(s), depth_func={depth_func}", flush=True)
    g = render_gpu(device, W, H, draws, clear, depth, depth_func)
    o = render_oracle(oracle, W, H, draws, clear, depth, depth_func)
    return g, o


def test_suprematism_1200x720(device, oracle):
    d = DrawSpec(S.SUPREMATISM_INDICES, S.SUPREMATISM_VERTICES, "passthrough", "passthrough")
    g, o = both(device, oracle, 1200, 720, [d])
    assert_frames_match(g, o, color_tol=0)
    assert g.stats["assembled_triangle_count"] == 8


def test_c1_sphere_basic_shading(device, oracle):
    W, H = 640, 360
    mesh, cb, tex = S.c1_scene(W, H)
    d = DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex)
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=1)
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 30000


def test_fullscreen_env_lighting(device, oracle):
    W, H = 320, 184
    env = S.smooth_env_map(256, 128, seed=3)
    cb = S.camera_cb((0.5, -0.3, 0.2), (-1.0, 0.4, 0.1), 75.0, W, H)
    d = DrawSpec(S.FULLSCREEN_INDICES, S.FULLSCREEN_VERTICES, "fullscreen", "env_lighting", cb=cb, ps_tex=env)
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=1)
    assert (g.vis_tri != 0xFFFFFFFF).all()


def test_vertex_lighting_and_env_pixel_lighting(device, oracle):
    W, H = 256, 256
    env = S.smooth_env_map(256, 128, seed=5)
    mesh = S.sphere_mesh(12, displace=0.2, seed=4)
    cb = S.camera_cb((2.6, 0.8, 0.9), (0, 0, 0), 60.0, W, H)
    draws = [
        DrawSpec(S.FULLSCREEN_INDICES, S.FULLSCREEN_VERTICES, "fullscreen", "env_lighting", cb=cb, ps_tex=env),
        DrawSpec(mesh.indices, mesh.vertices, "vertex_lighting", "passthrough", cb=cb, vs_tex=env),
        DrawSpec(mesh.indices, mesh.vertices, "basic", "env_lighting", cb=cb, ps_tex=env),
    ]
    g, o = both(device, oracle, W, H, draws)
    assert_frames_match(g, o, color_tol=1)


def test_fp16_env_maps(device, oracle):
    """vertex_lighting_vs and env_lighting_ps sampling RGBA16F maps (an extension format): the
    frame equals the oracle's on the maps' exact float32 values."""
    W, H = 160, 96
    irr = S.smooth_env_map(256, 128, seed=12, lo=0.05, hi=2.0).astype(np.float16)
    rad = S.smooth_env_map(512, 256, seed=13, lo=0.05, hi=2.0).astype(np.float16)
    mesh = S.sphere_mesh(12, displace=0.1, seed=4)
    cb = S.camera_cb((2.8, 0.7, 0.6), (0, 0, 0), 60.0, W, H)
    d = DrawSpec(mesh.indices, mesh.vertices, "vertex_lighting", "env_lighting", cb=cb, vs_tex=irr, ps_tex=rad)
    g, o = both(device, oracle, W, H, [d])
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 1000
    assert_frames_match(g, o, color_tol=1)


def test_clipping_near_plane_and_screen_edges(device, oracle):
    W, H = 192, 128
    tex = S.random_rgba8(64, seed=9)
    mesh = S.sphere_mesh(10, displace=0.15, seed=2)
    cb = S.camera_cb((1.04, 0.0, 0.05), (0.0, 0.1, 0.0), 95.0, W, H, near=0.05)
    d = DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex)
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=1)


def test_clipping_against_several_planes(device, oracle):
    """Large clip-space triangles, many crossing two or more planes and some with vertices
    behind the eye (w < 0): the clipped polygons' fan triangles and interpolated colours are
    bit-identical to the oracle's serial clipper."""
    W, H = 160, 96
    rng = np.random.default_rng(17)
    n = 3000
    w = rng.uniform(0.1, 2.0, (n, 3)) * np.where(rng.random((n, 3)) < 0.15, -1.0, 1.0)
    v = np.zeros((n, 3, 8), np.float32)
    v[..., 0] = rng.uniform(-3.0, 3.0, (n, 3)) * np.abs(w)
    v[..., 1] = rng.uniform(-3.0, 3.0, (n, 3)) * np.abs(w)
    v[..., 2] = rng.uniform(-0.5, 1.5, (n, 3)) * np.abs(w)
    v[..., 3] = w
    v[..., 4:7] = rng.uniform(0.0, 1.0, (n, 1, 3))
    v[..., 7] = 1.0
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    flip = rng.random(n) < 0.5  # both windings, so both cull outcomes occur
    idx[flip] = idx[flip][:, [0, 2, 1]]
    d = DrawSpec(idx.reshape(-1), v.reshape(-1, 8), "passthrough", "passthrough")
    for depth_func in (0, 1):
        g, o = both(device, oracle, W, H, [d], depth=0.0 if depth_func == 0 else 1.0, depth_func=depth_func)
        assert (g.depth != (0.0 if depth_func == 0 else 1.0)).mean() > 0.5
        assert_frames_match(g, o, color_tol=0)


def test_binning_stress_lite(device, oracle):
    W, H = 256, 192
    st = S.binning_stress(20000, 20, W, H, seed=7, big_every=1000)
    d = DrawSpec(st.indices, st.vertices, "passthrough", "passthrough")
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=0)


def test_multi_draw_depth_and_hiz(device, oracle):
    W, H = 320, 240
    tex = S.random_rgba8(128, seed=1)
    draws = []
    for k, (cx, r) in enumerate([(0.0, 1.0), (0.6, 0.7), (-0.4, 0.9), (0.2, 0.5)]):
        T = np.eye(4)
        T[:3, :3] *= r
        T[:3, 3] = (cx, 0.3 * k - 0.4, 0.1 * k)
        m = S.sphere_mesh(8, displace=0.1, seed=k, transform=T)
        cb = S.camera_cb((4.0, 0.5, 0.8), (0, 0, 0), 50.0, W, H)
        draws.append(DrawSpec(m.indices, m.vertices, "basic", "basic", cb=cb, ps_tex=tex))
    st = S.binning_stress(2000, 2, W, H, seed=11, big_every=700)
    draws.append(DrawSpec(st.indices, st.vertices, "passthrough", "passthrough"))
    g, o = both(device, oracle, W, H, draws)
    assert_frames_match(g, o, color_tol=1)
    assert len(np.unique(g.vis_draw)) >= 4


def test_less_equal_depth(device, oracle):
    W, H = 160, 120
    st = S.binning_stress(4000, 4, W, H, seed=13, big_every=900)
    d = DrawSpec(st.indices, st.vertices, "passthrough", "passthrough")
    g, o = both(device, oracle, W, H, [d, d], depth=1.0, depth_func=1)
    assert_frames_match(g, o, color_tol=0)


def test_resolution_not_multiple_of_16(device, oracle):
    W, H = 200, 120
    mesh, cb, tex = S.c1_scene(W, H)
    d = DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex)
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=1)


def test_start_index_and_base_vertex(device, oracle):
    W, H = 128, 128
    mesh, cb, tex = S.c1_scene(W, H)
    pad_v = np.zeros((5, 8), np.float32)
    verts = np.concatenate([pad_v, mesh.vertices])
    idx = np.concatenate([np.zeros(30, np.uint32), mesh.indices])
    n = 3 * 4000
    d = DrawSpec(idx, verts, "basic", "basic", cb=cb, ps_tex=tex, start_index=30 + 3 * 100, base_vertex=5,
                 index_count=n)
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=1)


def test_out_of_range_indices_are_dropped_and_counted(device, oracle):
    W, H = 64, 64
    st = S.binning_stress(300, 1, W, H, seed=2, big_every=100)
    idx = st.indices.copy()
    idx[3 * 17 + 1] = 10 ** 7
    idx[3 * 40] = st.vertices.shape[0]
    d = DrawSpec(idx, st.vertices, "passthrough", "passthrough")
    g, o = both(device, oracle, W, H, [d])
    assert_frames_match(g, o, color_tol=0)
    assert g.stats["oob_triangle_count"] == 2


def test_empty_and_degenerate_draws(device, oracle):
    W, H = 64, 48
    d0 = DrawSpec(np.zeros(3, np.uint32), S.SUPREMATISM_VERTICES, "passthrough", "passthrough", index_count=0)
    d1 = DrawSpec(np.zeros(24, np.uint32), S.SUPREMATISM_VERTICES, "passthrough", "passthrough")
    g, o = both(device, oracle, W, H, [d0, d1])
    assert_frames_match(g, o, color_tol=0)


def test_invalid_state_reports_errors(device):
    dev = device
    dev.ia_set_primitive_topology(L.TOPOLOGY_UNDEFINED)
    with pytest.raises(L.MalevichError) as e:
        dev.draw_indexed(3)
    assert e.value.code == L.ERROR_INVALID_STATE
    dev.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
    with pytest.raises(L.MalevichError) as e:
        dev.draw_indexed(4)
    assert e.value.code == L.ERROR_INVALID_ARGUMENT


_OVERFLOW_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
from malevich_amd import pipeline as P, scenes as S
from tests.render_utils import DrawSpec, render_gpu, render_oracle, assert_frames_match
from oracle import binding as ob
st = S.binning_stress(6000, 6, 192, 128, seed=21, big_every=900)
d = DrawSpec(st.indices, st.vertices, "passthrough", "passthrough")
dev = P.Device(0)
g = render_gpu(dev, 192, 128, [d], S.CLEAR_COLOR)
o = render_oracle(ob, 192, 128, [d], S.CLEAR_COLOR)
assert g.stats["overflow_count"] == 1, g.stats
assert_frames_match(g, o, color_tol=0)
dev.close()
print("OVERFLOW_PATH_OK")
"""


def test_bin_list_overflow_slow_path_is_exact():
    """A bin-list arena too small for the draw switches the tile kernel to scanning every
    triangle; the image must not change."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MLV_PAIR_CAPACITY="1000")
    print("case: bin-list overflow child process", flush=True)
    r = subprocess.run([sys.executable, "-c", _OVERFLOW_CHILD, root], env=env, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0 and "OVERFLOW_PATH_OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("depth_func", [0, 1])
def test_heavy_tiles_multi_draw(device, oracle, depth_func):
    """Tiles with thousands of entries from several draws (walked by 16 waves per bin, merged
    at draw boundaries) equal the single-threaded reference order, Hi-Z refreshes included."""
    W, H = 64, 48
    draws = []
    for k in range(8):
        st = S.binning_stress(3000, 3, W, H, seed=100 + k, big_every=900)
        draws.append(DrawSpec(st.indices, st.vertices, "passthrough", "passthrough"))
    clear_depth = 0.0 if depth_func == 0 else 1.0
    saved = device.tuning()
    device.set_tuning(**dict(saved, heavy_tile_pairs=384, tile_edge=16))  # past 384: 16 waves
    try:
        g, o = both(device, oracle, W, H, draws, depth=clear_depth, depth_func=depth_func)
    finally:
        device.set_tuning(**saved)
    assert_frames_match(g, o, color_tol=0)
    assert len(np.unique(g.vis_draw)) >= 6


@pytest.mark.parametrize("size", [(1200, 720), (37, 23)])
def test_present_is_the_dib_view_of_the_frame(device, oracle, size):
    """mlv_present gives the image paint_window shows: each u32 read through the DIB masks
    R 0x00FF0000, G 0x0000FF00, B 0x000000FF (main.c:305-307), top row first."""
    W, H = size
    print(f"case: present {W}x{H}", flush=True)
    d = device
    rt, ds = d.create_render_target(W, H), d.create_depth_stencil(W, H)
    ib = d.create_buffer(L.BUFFER_INDEX_U32, S.SUPREMATISM_INDICES)
    vb = d.create_buffer(L.BUFFER_VERTEX, S.SUPREMATISM_VERTICES)
    try:
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.ia_set_index_buffer(ib)
        d.ia_set_vertex_buffer(vb, 32)
        d.vs_set_shader(L.VS_PASSTHROUGH)
        d.ps_set_shader(L.PS_PASSTHROUGH)
        d.rs_set_viewport(0, 0, W, H, 0, 1)
        d.om_set_render_targets(rt, ds)
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        d.clear_render_target_view(rt, S.CLEAR_COLOR)
        d.clear_depth_stencil_view(ds, 0.0)
        d.draw_indexed(S.SUPREMATISM_INDICES.size)
        img = rt.present()
        fb = rt.read()
    finally:
        for h in (ib, vb, rt, ds):
            h.release()
    if W % 8 == 0 and H % 8 == 0:  # the oracle's targets are whole 8x8 bins
        o = render_oracle(oracle, W, H, [DrawSpec(S.SUPREMATISM_INDICES, S.SUPREMATISM_VERTICES, "passthrough",
                                                   "passthrough")], S.CLEAR_COLOR)
        np.testing.assert_array_equal(fb, o.colors)
    want = np.stack([(fb >> 16) & 255, (fb >> 8) & 255, fb & 255], -1).astype(np.uint8)
    np.testing.assert_array_equal(img, want)
    assert len(np.unique(img.reshape(-1, 3), axis=0)) >= 3
