"""Launch parameters change speed, never results.

Bar: frames and stats drawn under extreme tunings (every long tile on the 16-wave path, or none;
one heavy-tile block; few pair blocks; 16x16 and 32x32 tiles; occupancy caps; triangles without a
pixel corner binned or not) are bit-identical to the oracle's; the autotuner returns parameters
from its candidate set, leaves the recorded frame in the targets (bit-identical to the eager
frame) and the context keeps the chosen parameters.
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu

DEFAULT = dict(heavy_tile_pairs=0xFFFFFFFF, heavy_blocks=1024, pair_blocks=2048, tile_stage=256, tile_edge=16,
               tile_waves_per_cu=0, setup_waves_per_cu=0, skip_cornerless=1)


def _draws(W, H):
    out = []
    for k in range(6):
        st = S.binning_stress(2500, 3, W, H, seed=300 + k, big_every=800)
        out.append(DrawSpec(st.indices, st.vertices, "passthrough", "passthrough"))
    return out


@pytest.mark.parametrize("tuning", [dict(heavy_tile_pairs=0, heavy_blocks=1, pair_blocks=3, tile_stage=256),
                                    dict(heavy_tile_pairs=0xFFFFFFFF, heavy_blocks=7, pair_blocks=1, tile_stage=128),
                                    dict(heavy_tile_pairs=100, heavy_blocks=4096, pair_blocks=4096, tile_stage=128),
                                    dict(heavy_tile_pairs=0, heavy_blocks=64, pair_blocks=512, tile_stage=256,
                                         skip_cornerless=0, tile_waves_per_cu=8, setup_waves_per_cu=16),
                                    dict(heavy_tile_pairs=0, heavy_blocks=64, pair_blocks=5, tile_stage=128,
                                         tile_edge=32, tile_waves_per_cu=16, setup_waves_per_cu=24),
                                    dict(heavy_tile_pairs=0xFFFFFFFF, heavy_blocks=1024, pair_blocks=2048,
                                         tile_stage=256, tile_edge=32, skip_cornerless=0)])
def test_tunings_give_the_oracle_frame(device, oracle, tuning):
    W, H = 96, 64
    draws = _draws(W, H)
    print(f"case: tuning {tuning}", flush=True)
    try:
        device.set_tuning(**tuning)
        g = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
    finally:
        device.set_tuning(**DEFAULT)
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    assert_frames_match(g, o, color_tol=0)


@pytest.mark.parametrize("skip", [1, 0])
@pytest.mark.parametrize("edge", [16, 32])
def test_triangles_without_a_pixel_corner(device, oracle, skip, edge):
    """A dense mesh in a small view: most of its triangles hold no pixel corner. Skipped or binned,
    the frame and every stat (assembled triangles, active bins, triangle-bin pairs) equal the
    oracle's."""
    W, H = 80, 56
    mesh, _ = S.sphere_of_about(60000, 0.1, 77, np.eye(4))
    cb = S.camera_cb(eye=(0.0, 0.4, 2.6), target=(0.0, 0.0, 0.0), fov_y_deg=50.0, width=W, height=H)
    tex = S.random_rgba8(64, seed=5)
    draws = [DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex)]
    big = S.binning_stress(300, 2, W, H, seed=9, big_every=100)
    draws.append(DrawSpec(big.indices, big.vertices, "passthrough", "passthrough"))
    print(f"case: cornerless skip={skip} edge={edge}", flush=True)
    try:
        device.set_tuning(**dict(DEFAULT, skip_cornerless=skip, tile_edge=edge))
        g = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
    finally:
        device.set_tuning(**DEFAULT)
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    assert_frames_match(g, o, color_tol=1)


def test_autotune_keeps_results_and_choice(device, oracle):
    W, H = 320, 184
    sc = S.c2_scene(W, H, n_instances=8, seed=21, texture_size=0)
    env = S.smooth_env_map(256, 128, seed=6, lo=0.05, hi=2.0)
    print(f"case: autotune {W}x{H}, {len(sc.instances)} draws", flush=True)
    d = device
    vbs = [d.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.instances]
    ibs = [d.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.instances]
    cb = d.create_buffer(L.BUFFER_CONSTANT, sc.cb)
    tex = d.create_texture2d(env)
    rt, ds = d.create_render_target(W, H), d.create_depth_stencil(W, H)

    def frame():
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.vs_set_shader(L.VS_BASIC)
        d.ps_set_shader(L.PS_ENV_LIGHTING)
        d.vs_set_constant_buffer(0, cb)
        d.ps_set_shader_resource(0, tex)
        d.rs_set_viewport(0, 0, W, H, 0.0, 1.0)
        d.om_set_render_targets(rt, ds)
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        d.clear_render_target_view(rt, S.CLEAR_COLOR)
        d.clear_depth_stencil_view(ds, 0.0)
        for vb, ib, m in zip(vbs, ibs, sc.instances):
            d.ia_set_vertex_buffer(vb, 32)
            d.ia_set_index_buffer(ib)
            d.draw_indexed(m.indices.size)

    try:
        frame()
        ref = (rt.read().copy(), ds.read().copy())
        d.reset_stats()
        d.autotune_begin()
        frame()
        best = d.autotune_end(repeats=2)
        print(f"  chosen: {best}", flush=True)
        # the tuned frame counts once in the stats, however many times the tuner ran it
        st = d.stats()
        draws = [DrawSpec(m.indices, m.vertices, "basic", "env_lighting", cb=sc.cb, ps_tex=env)
                 for m in sc.instances]
        o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
        for k in ("vertex_count", "input_triangle_count", "assembled_triangle_count", "active_bin_count",
                  "total_triangle_count_in_bins", "oob_triangle_count"):
            assert st[k] == o.stats[k], (k, st[k], o.stats[k])
        assert st["draw_count"] == len(sc.instances)
        assert best["heavy_tile_pairs"] in (0, 32, 64, 192, 384, 768, 1536, 0xFFFFFFFF)
        assert best["heavy_blocks"] in (256, 512, 1024, 2048)
        assert best["pair_blocks"] in (512, 1024, 2048, 4096)
        assert best["tile_stage"] in (128, 256)
        assert best["tile_edge"] in (16, 32) and best["skip_cornerless"] in (0, 1)
        assert best["tile_waves_per_cu"] in (0, 8, 16, 24) and best["setup_waves_per_cu"] in (0, 16, 24)
        assert best["pair_capacity"] > 0
        np.testing.assert_array_equal(rt.read(), ref[0])
        np.testing.assert_array_equal(ds.read().view(np.uint32), ref[1].view(np.uint32))
        now = d.tuning()
        assert {k: now[k] for k in DEFAULT} == {k: best[k] for k in DEFAULT}
    finally:
        d.set_tuning(**DEFAULT)
        for h in vbs + ibs + [cb, tex, rt, ds]:
            h.release()
