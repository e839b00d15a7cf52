"""Many views of many meshes recorded together (one flush covers the draws of every view) give
exactly the frames of flushing after every draw, and the oracle's frames."""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, Frame, assert_frames_match, render_oracle

pytestmark = pytest.mark.gpu


def render_views(dev, sc, per_draw: bool):
    keep, targets = [], []
    try:
        vbs = [dev.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.meshes]
        ibs = [dev.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.meshes]
        texs = [dev.create_texture2d(t) for t in sc.textures]
        cbs = [dev.create_buffer(L.BUFFER_CONSTANT, v.cb) for v in sc.views]
        keep += vbs + ibs + texs + cbs
        dev.reset_stats()
        dev.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        dev.vs_set_shader(L.VS_BASIC)
        dev.ps_set_shader(L.PS_BASIC)
        dev.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        for v in sc.views:
            rt = dev.create_render_target(v.width, v.height)
            ds = dev.create_depth_stencil(v.width, v.height)
            vt = dev.create_buffer(L.BUFFER_VERTEX, nbytes=v.width * v.height * 4)
            vd = dev.create_buffer(L.BUFFER_VERTEX, nbytes=v.width * v.height * 4)
            targets.append((rt, ds, vt, vd))
            dev.om_set_render_targets(rt, ds)
            dev.om_set_visibility_buffer(vt, vd)
            dev.clear_render_target_view(rt, S.CLEAR_COLOR)
            dev.clear_depth_stencil_view(ds, 0.0)
        for v, cb, (rt, ds, vt, vd) in zip(sc.views, cbs, targets):
            dev.rs_set_viewport(0, 0, v.width, v.height, 0, 1)
            dev.om_set_render_targets(rt, ds)
            dev.om_set_visibility_buffer(vt, vd)
            dev.vs_set_constant_buffer(0, cb)
            for vb, ib, tex, m in zip(vbs, ibs, texs, sc.meshes):
                dev.ia_set_vertex_buffer(vb, 32)
                dev.ia_set_index_buffer(ib)
                dev.ps_set_shader_resource(0, tex)
                dev.draw_indexed(m.indices.size)
                if per_draw:
                    dev.flush()
        frames = []
        for v, (rt, ds, vt, vd) in zip(sc.views, targets):
            frames.append(Frame(rt.read(), ds.read(), vt.read(np.uint32).reshape(v.height, v.width),
                                vd.read(np.uint32).reshape(v.height, v.width)))
        stats = dev.stats()
        return frames, stats
    finally:
        dev.om_set_visibility_buffer(None, None)
        for t in targets:
            for h in t:
                h.release()
        for k in keep:
            k.release()


# (views, meshes, min/max triangles): 30 draws in one flush, and 320 (a flush of 256 draws or
# more looks each block's draw up in a table filled on the device instead of searching)
SCENES = {"few_draws": (6, 5, 200, 6000), "many_draws": (4, 80, 20, 400)}


@pytest.mark.parametrize("scene", list(SCENES))
def test_batched_views_equal_per_draw_and_oracle(device, oracle, scene):
    nv, nm, lo, hi = SCENES[scene]
    sc = S.batched_scene(n_views=nv, n_meshes=nm, min_tris=lo, max_tris=hi, resolutions=(64, 96, 128),
                         tex_sizes=(32, 64), seed=8)
    print(f"case: batched {len(sc.views)} views x {len(sc.meshes)} meshes", flush=True)
    fb, sb = render_views(device, sc, per_draw=False)
    fp, sp = render_views(device, sc, per_draw=True)
    assert sb["draw_count"] == sp["draw_count"] == len(sc.views) * len(sc.meshes)
    for a, b in zip(fb, fp):
        np.testing.assert_array_equal(a.colors, b.colors)
        np.testing.assert_array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32))
        np.testing.assert_array_equal(a.vis_tri, b.vis_tri)
    for k in ("assembled_triangle_count", "total_triangle_count_in_bins", "active_bin_count"):
        assert sb[k] == sp[k], k
    covered = 0
    for i, (v, g) in enumerate(zip(sc.views, fb)):
        draws = [DrawSpec(m.indices, m.vertices, "basic", "basic", cb=v.cb, ps_tex=t)
                 for m, t in zip(sc.meshes, sc.textures)]
        o = render_oracle(oracle, v.width, v.height, draws, S.CLEAR_COLOR)
        # draw tags count draws across the whole batch; the oracle numbers them per view
        g.vis_draw = np.where(g.vis_draw == 0xFFFFFFFF, g.vis_draw, g.vis_draw - i * len(sc.meshes)).astype(np.uint32)
        assert_frames_match(g, o, color_tol=1, check_stats=False)
        covered += int((g.vis_tri != 0xFFFFFFFF).sum())
    assert covered > 1000
