"""Render the same draw list through the gfx950 library and through the CPU oracle."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from malevich_amd import _lib as L
from malevich_amd import pipeline as P

VS = {"basic": (L.VS_BASIC, 0), "vertex_lighting": (L.VS_VERTEX_LIGHTING, 1), "passthrough": (L.VS_PASSTHROUGH, 2),
      "fullscreen": (L.VS_FULLSCREEN, 3)}
PS = {"basic": (L.PS_BASIC, 0), "env_lighting": (L.PS_ENV_LIGHTING, 1), "passthrough": (L.PS_PASSTHROUGH, 2),
      "basic_lambert": (L.PS_BASIC_LAMBERT, 3), "env_textured": (L.PS_ENV_TEXTURED, 4), "none": (L.PS_NONE, 5)}


@dataclass
class DrawSpec:
    indices: np.ndarray
    vertices: np.ndarray
    vs: str
    ps: str
    cb: np.ndarray | None = None
    vs_tex: np.ndarray | None = None
    ps_tex: np.ndarray | None = None
    start_index: int = 0
    base_vertex: int = 0
    index_count: int | None = None
    ps_tex1: np.ndarray | list | None = None  # extension shaders: texture slot 1
    ps_cb: np.ndarray | None = None           # extension shaders: 12 floats
    cull: int = 0                             # extension: 0 back (the reference's), 1 none, 2 front
    fill: int = 0                             # extension: 0 inclusive (the reference's), 1 top-left
    # a texture given as a list is a mip chain (level 0 first)


@dataclass
class Frame:
    colors: np.ndarray
    depth: np.ndarray
    vis_tri: np.ndarray
    vis_draw: np.ndarray
    stats: dict = field(default_factory=dict)


def render_gpu(dev: P.Device, W: int, H: int, draws: list[DrawSpec], clear_color, clear_depth=0.0,
               depth_func: int = L.DEPTH_GREATER_EQUAL) -> Frame:
    rt = dev.create_render_target(W, H)
    ds = dev.create_depth_stencil(W, H)
    vis_t = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    vis_d = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    keep = []
    try:
        dev.reset_stats()
        dev.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        dev.rs_set_viewport(0, 0, W, H, 0, 1)
        dev.om_set_depth_func(depth_func)
        dev.om_set_render_targets(rt, ds)
        dev.om_set_visibility_buffer(vis_t, vis_d)
        dev.clear_render_target_view(rt, clear_color)
        dev.clear_depth_stencil_view(ds, clear_depth)
        for d in draws:
            ib = dev.create_buffer(L.BUFFER_INDEX_U32, np.ascontiguousarray(d.indices, np.uint32))
            vb = dev.create_buffer(L.BUFFER_VERTEX, np.ascontiguousarray(d.vertices, np.float32))
            keep += [ib, vb]
            dev.ia_set_index_buffer(ib)
            dev.ia_set_vertex_buffer(vb, d.vertices.shape[1] * 4)
            dev.vs_set_shader(VS[d.vs][0])
            dev.ps_set_shader(PS[d.ps][0])
            if d.cb is not None:
                cb = dev.create_buffer(L.BUFFER_CONSTANT, np.ascontiguousarray(d.cb, np.float32))
                keep.append(cb)
                dev.vs_set_constant_buffer(0, cb)
            for tex, setter, slot in ((d.vs_tex, dev.vs_set_shader_resource, 0),
                                      (d.ps_tex, dev.ps_set_shader_resource, 0),
                                      (d.ps_tex1, dev.ps_set_shader_resource, 1)):
                if tex is not None:
                    t = dev.create_texture2d(tex)
                    keep.append(t)
                    setter(slot, t)
                else:
                    setter(slot, None)
            if d.ps_cb is not None:
                pcb = dev.create_buffer(L.BUFFER_CONSTANT, np.ascontiguousarray(d.ps_cb, np.float32))
                keep.append(pcb)
                dev.ps_set_constant_buffer(0, pcb)
            else:
                dev.ps_set_constant_buffer(0, None)
            dev.rs_set_cull(d.cull)
            dev.rs_set_fill_rule(d.fill)
            n = d.indices.size if d.index_count is None else d.index_count
            dev.draw_indexed(n, d.start_index, d.base_vertex)
        dev.rs_set_cull(L.CULL_BACK)
        dev.rs_set_fill_rule(L.FILL_INCLUSIVE)
        colors = rt.read()
        depth = ds.read()
        vt = vis_t.read(np.uint32).reshape(H, W)
        vd = vis_d.read(np.uint32).reshape(H, W)
        stats = dev.stats()
    finally:
        dev.om_set_visibility_buffer(None, None)
        for k in keep:
            k.release()
        vis_t.release()
        vis_d.release()
        rt.release()
        ds.release()
    return Frame(colors, depth, vt, vd, stats)


def render_oracle(oracle, W: int, H: int, draws: list[DrawSpec], clear_color, clear_depth=0.0,
                  depth_func: int = 0) -> Frame:
    rt = oracle.RenderTarget(W, H)
    rt.clear_render_target_view(clear_color)
    rt.clear_depth_stencil_view(clear_depth, depth_func)
    for d in draws:
        od = oracle.Draw(d.indices, d.vertices, VS[d.vs][1], PS[d.ps][1], cb=d.cb, vs_tex=oracle_texture(oracle, d.vs_tex),
                         ps_tex=oracle_texture(oracle, d.ps_tex), viewport=(0, 0, W, H, 0, 1),
                         start_index=d.start_index, base_vertex=d.base_vertex, index_count=d.index_count,
                         depth_func=depth_func, ps_tex1=oracle_texture(oracle, d.ps_tex1), ps_cb=d.ps_cb,
                         cull_mode=d.cull, fill_rule=d.fill)
        oracle.draw_indexed(rt, od)
    st = {k: getattr(rt.stats, k) for k, _ in oracle.Stats._fields_}
    return Frame(rt.colors.copy(), rt.depth.copy(), rt.vis_tri.copy(), rt.vis_draw.copy(), st)


def oracle_texture(oracle, t):
    """A texture for the oracle: RGBA16F texels are exact in float32, so the oracle samples their
    float32 values; a list is a mip chain."""
    if t is None:
        return None
    if isinstance(t, (list, tuple)):
        return oracle.MipTexture(list(t))
    return t.astype(np.float32) if t.dtype == np.float16 else t


def channel_diff(a: np.ndarray, b: np.ndarray) -> int:
    d = 0
    for s in (0, 8, 16, 24):
        d = max(d, int(np.abs(((a >> s) & 255).astype(np.int64) - ((b >> s) & 255).astype(np.int64)).max(initial=0)))
    return d


def assert_frames_match(g: Frame, o: Frame, color_tol: int, check_stats: bool = True) -> None:
    np.testing.assert_array_equal(g.vis_tri, o.vis_tri, err_msg="visibility (triangle id) differs")
    np.testing.assert_array_equal(g.vis_draw, o.vis_draw, err_msg="visibility (draw tag) differs")
    np.testing.assert_array_equal(g.depth.view(np.uint32), o.depth.view(np.uint32), err_msg="depth differs")
    if color_tol == 0:
        np.testing.assert_array_equal(g.colors, o.colors, err_msg="colour differs")
    else:
        assert channel_diff(g.colors, o.colors) <= color_tol
    if check_stats:
        for k in ("vertex_count", "input_triangle_count", "assembled_triangle_count", "active_bin_count",
                  "total_triangle_count_in_bins", "oob_triangle_count"):
            assert g.stats[k] == o.stats[k], (k, g.stats[k], o.stats[k])
