"""A frame recorded and captured as a HIP graph replays exactly like the same frame run eagerly.

Bar: every replay's colour, depth and visibility buffers are bit-identical to the eager frame
with the same constant-buffer contents; replays alternate with eager frames and camera updates
(so a replay that left stale state behind, or read stale scan status or tables, would differ);
the stats advance by one frame per launch; flushing while recording is refused.
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from malevich_amd._lib import MalevichError

pytestmark = pytest.mark.gpu


class Frame:
    def __init__(self, dev, sc, env):
        W, H = sc.width, sc.height
        self.dev, self.sc = dev, sc
        self.vbs = [dev.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.instances]
        self.ibs = [dev.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.instances]
        self.cb = dev.create_buffer(L.BUFFER_CONSTANT, sc.cb)
        self.env = dev.create_texture2d(env)
        self.rt = dev.create_render_target(W, H)
        self.ds = dev.create_depth_stencil(W, H)
        self.vt = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
        self.vd = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)

    def record(self):
        d, sc = self.dev, self.sc
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.vs_set_shader(L.VS_BASIC)
        d.ps_set_shader(L.PS_ENV_LIGHTING)
        d.vs_set_constant_buffer(0, self.cb)
        d.ps_set_shader_resource(0, self.env)
        d.rs_set_viewport(0, 0, sc.width, sc.height, 0.0, 1.0)
        d.om_set_render_targets(self.rt, self.ds)
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        d.om_set_visibility_buffer(self.vt, self.vd)
        d.clear_render_target_view(self.rt, S.CLEAR_COLOR)
        d.clear_depth_stencil_view(self.ds, 0.0)
        for vb, ib, m in zip(self.vbs, self.ibs, sc.instances):
            d.ia_set_vertex_buffer(vb, 32)
            d.ia_set_index_buffer(ib)
            d.draw_indexed(m.indices.size)

    def read(self):
        H, W = self.sc.height, self.sc.width
        return (self.rt.read().copy(), self.ds.read().view(np.uint32).copy(),
                self.vt.read(np.uint32).reshape(H, W).copy(), self.vd.read(np.uint32).reshape(H, W).copy())

    def release(self):
        self.dev.om_set_visibility_buffer(None, None)
        for h in self.vbs + self.ibs + [self.cb, self.env, self.rt, self.ds, self.vt, self.vd]:
            h.release()


def _same(a, b, what):
    for x, y, n in zip(a, b, ("colour", "depth", "tri id", "draw tag")):
        np.testing.assert_array_equal(x, y, err_msg=f"{what}: {n}")


def test_graph_replays_match_eager_frames(device):
    W, H = 320, 184
    sc = S.c2_scene(W, H, n_instances=10, seed=9, texture_size=0)
    env = S.smooth_env_map(256, 128, seed=4, lo=0.05, hi=2.0)
    cam_b = S.camera_cb((4.0, 26.0, 9.0), (0.0, 0.0, 0.0), 55.0, W, H)
    print(f"case: graph {W}x{H}, {len(sc.instances)} draws", flush=True)
    f = Frame(device, sc, env)
    g = None
    try:
        device.reset_stats()
        f.record()
        ref_a = f.read()
        keys = ("assembled_triangle_count", "active_bin_count", "total_triangle_count_in_bins")
        st_a = {k: device.stats()[k] for k in keys}
        f.cb.update(cam_b)
        f.record()
        ref_b = f.read()
        st_ab = {k: device.stats()[k] for k in keys}
        assert (ref_a[2] != 0xFFFFFFFF).sum() > 1000 and (ref_b[2] != 0xFFFFFFFF).sum() > 1000
        assert not np.array_equal(ref_a[0], ref_b[0])

        f.cb.update(sc.cb)
        device.graph_capture_begin()
        f.record()
        with pytest.raises(MalevichError) as e:
            device.flush()  # anything that would run work is refused while recording
        assert e.value.code == L.ERROR_INVALID_STATE
        g = device.graph_capture_end()  # runs the frame once, then captures it
        # the capture's first run counts once, like an eager frame (its sizing retries do not)
        st = device.stats()
        for k in keys:
            assert st[k] == st_ab[k] + st_a[k], (k, st[k], st_ab[k], st_a[k])
        # draw tags count draws since the last reset: the recorded draws got theirs after the
        # two eager frames, and every replay writes those
        tag0 = 2 * len(sc.instances)
        ref_a_tags = (ref_a[0], ref_a[1], ref_a[2], np.where(ref_a[3] == 0xFFFFFFFF, ref_a[3], ref_a[3] + tag0))
        _same(f.read(), ref_a_tags, "capture run")
        draws0 = device.stats()["draw_count"]
        for i in range(3):
            f.cb.update(sc.cb)
            g.launch()
            _same(f.read(), ref_a_tags, f"replay {i} camera A")
            f.cb.update(cam_b)
            g.launch()
            got = f.read()
            _same(got[:3], ref_b[:3], f"replay {i} camera B")
            # eager work between replays
            f.record()
            eager = f.read()
            _same(eager[:3], ref_b[:3], f"eager frame after replay {i}")
        assert device.stats()["draw_count"] == draws0 + 9 * len(sc.instances)
    finally:
        if g is not None:
            g.release()
        f.release()


def test_graph_of_batched_views(device):
    sc = S.batched_scene(n_views=4, n_meshes=3, min_tris=200, max_tris=4000, resolutions=(64, 128),
                         tex_sizes=(32,), seed=13)
    print(f"case: graph of {len(sc.views)} views x {len(sc.meshes)} meshes", flush=True)
    d = device
    vbs = [d.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.meshes]
    ibs = [d.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.meshes]
    texs = [d.create_texture2d(t) for t in sc.textures]
    cbs = [d.create_buffer(L.BUFFER_CONSTANT, v.cb) for v in sc.views]
    tg = [(d.create_render_target(v.width, v.height), d.create_depth_stencil(v.width, v.height)) for v in sc.views]
    g = None

    def frame():
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.vs_set_shader(L.VS_BASIC)
        d.ps_set_shader(L.PS_BASIC)
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        for v, cb, (rt, ds) in zip(sc.views, cbs, tg):
            d.rs_set_viewport(0, 0, v.width, v.height, 0, 1)
            d.om_set_render_targets(rt, ds)
            d.clear_render_target_view(rt, S.CLEAR_COLOR)
            d.clear_depth_stencil_view(ds, 0.0)
            d.vs_set_constant_buffer(0, cb)
            for vb, ib, tex, m in zip(vbs, ibs, texs, sc.meshes):
                d.ia_set_vertex_buffer(vb, 32)
                d.ia_set_index_buffer(ib)
                d.ps_set_shader_resource(0, tex)
                d.draw_indexed(m.indices.size)

    def read():
        return [(rt.read().copy(), ds.read().view(np.uint32).copy()) for rt, ds in tg]

    try:
        frame()
        ref = read()
        d.graph_capture_begin()
        frame()
        g = d.graph_capture_end()
        for i in range(3):
            got = read()
            for (c, z), (rc, rz) in zip(got, ref):
                np.testing.assert_array_equal(c, rc, err_msg=f"replay {i}")
                np.testing.assert_array_equal(z, rz, err_msg=f"replay {i}")
            for rt, ds in tg:  # scribble over the targets; the replay must redraw them
                d.clear_render_target_view(rt, (0.0, 0.0, 0.0, 0.0))
                d.clear_depth_stencil_view(ds, 1.0)
            d.synchronize()
            g.launch()
    finally:
        if g is not None:
            g.release()
        for h in vbs + ibs + texs + cbs:
            h.release()
        for rt, ds in tg:
            rt.release()
            ds.release()
