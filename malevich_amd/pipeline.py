"""D3D11-style host API over the gfx950 C ABI.

Mirrors the reference's pipeline interface (source/main.c): the `Pipeline`
state struct with IA / VS / RS / PS / OM parts (main.c:71-115) and the three
commands clear_render_target_view (main.c:1191), clear_depth_stencil_view
(main.c:1204) and draw_indexed (main.c:1219). Each call goes straight to the
C ABI of libmalevich_hip.so; errors come back as MalevichError with the
library's status code and message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import MalevichError


def _check(ctx, rc: int) -> None:
    if rc != L.OK:
        msg = L.load().mlv_get_last_error(ctx).decode(errors="replace") if ctx else ""
        raise MalevichError(rc, msg)


def device_count() -> int:
    n = C.c_int(0)
    L.load().mlv_device_count(C.byref(n))
    return n.value


class Buffer:
    def __init__(self, dev: "Device", handle: int, nbytes: int, kind: int):
        self.dev, self.handle, self.nbytes, self.kind = dev, handle, nbytes, kind

    @property
    def device_ptr(self) -> int:
        return L.load().mlv_buffer_device_ptr(self.handle)

    def update(self, data, offset: int = 0) -> None:
        a = np.ascontiguousarray(data)
        _check(self.dev.ctx, L.load().mlv_buffer_update(self.dev.ctx, self.handle, offset, a.nbytes, a.ctypes.data))

    def read(self, dtype=np.uint8, count: int | None = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        _check(self.dev.ctx, L.load().mlv_buffer_readback(self.dev.ctx, self.handle, offset, out.nbytes,
                                                          out.ctypes.data))
        return out

    def release(self) -> None:
        if self.handle:
            L.load().mlv_buffer_destroy(self.dev.ctx, self.handle)
            self.handle = None


def _level_sizes(width: int, height: int, levels: int) -> list:
    out = [(width, height)]
    while len(out) < levels:
        w, h = out[-1]
        out.append((max(w >> 1, 1), max(h >> 1, 1)))
    return out


def _texel_shape(fmt: int, w: int, h: int):
    return ((h, w), np.uint32) if fmt == L.FORMAT_RGBA8_UNORM else \
        ((h, w, 4), np.float32 if fmt == L.FORMAT_RGBA32F else np.float16)


class Texture2D:
    def __init__(self, dev, handle, width, height, fmt, mip_levels: int = 1):
        self.dev, self.handle, self.width, self.height, self.format = dev, handle, width, height, fmt
        self.mip_levels = mip_levels

    def generate_mips(self) -> None:
        """Levels 1.. of the chain from level 0 on the device (2x2 box filter)."""
        _check(self.dev.ctx, L.load().mlv_texture2d_generate_mips(self.dev.ctx, self.handle))

    def read_levels(self) -> list:
        """Every level of the chain, level 0 first, shaped as read() shapes level 0."""
        sizes = _level_sizes(self.width, self.height, self.mip_levels)
        shapes = [_texel_shape(self.format, w, h) for w, h in sizes]
        flat = np.empty(sum(int(np.prod(s)) for s, _ in shapes), shapes[0][1])
        _check(self.dev.ctx, L.load().mlv_texture2d_readback(self.dev.ctx, self.handle, flat.ctypes.data))
        out, o = [], 0
        for shape, _ in shapes:
            n = int(np.prod(shape))
            out.append(flat[o:o + n].reshape(shape))
            o += n
        return out

    def srgb_to_linear(self) -> None:
        """load_texture's sRGB pre-pass (main.c:546-558) on the device, in place (RGBA8 only)."""
        _check(self.dev.ctx, L.load().mlv_texture2d_srgb_to_linear(self.dev.ctx, self.handle))

    def read(self) -> np.ndarray:
        """Level-0 texels: (h, w) uint32 for RGBA8, (h, w, 4) float32 / float16 for the float formats."""
        if self.mip_levels > 1:
            return self.read_levels()[0]
        if self.format == L.FORMAT_RGBA8_UNORM:
            out = np.empty((self.height, self.width), np.uint32)
        elif self.format == L.FORMAT_RGBA32F:
            out = np.empty((self.height, self.width, 4), np.float32)
        else:
            out = np.empty((self.height, self.width, 4), np.float16)
        _check(self.dev.ctx, L.load().mlv_texture2d_readback(self.dev.ctx, self.handle, out.ctypes.data))
        return out

    def release(self) -> None:
        if self.handle:
            L.load().mlv_texture2d_destroy(self.dev.ctx, self.handle)
            self.handle = None


class RenderTarget:
    def __init__(self, dev, handle, width, height):
        self.dev, self.handle, self.width, self.height = dev, handle, width, height

    @property
    def device_ptr(self) -> int:
        return L.load().mlv_render_target_device_ptr(self.handle)

    def read(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.uint32)
        _check(self.dev.ctx, L.load().mlv_render_target_readback(self.dev.ctx, self.handle, out.ctypes.data))
        return out

    def present(self) -> np.ndarray:
        """(H, W, 3) uint8: the image the reference's window shows for this frame (mlv_present)."""
        out = np.empty((self.height, self.width, 3), np.uint8)
        _check(self.dev.ctx, L.load().mlv_present(self.dev.ctx, self.handle, out.ctypes.data))
        return out

    def release(self) -> None:
        if self.handle:
            L.load().mlv_render_target_destroy(self.dev.ctx, self.handle)
            self.handle = None


class DepthStencil:
    def __init__(self, dev, handle, width, height):
        self.dev, self.handle, self.width, self.height = dev, handle, width, height

    @property
    def device_ptr(self) -> int:
        return L.load().mlv_depth_stencil_device_ptr(self.handle)

    def read(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.float32)
        _check(self.dev.ctx, L.load().mlv_depth_stencil_readback(self.dev.ctx, self.handle, out.ctypes.data))
        return out

    def release(self) -> None:
        if self.handle:
            L.load().mlv_depth_stencil_destroy(self.dev.ctx, self.handle)
            self.handle = None


class Graph:
    """A recorded frame captured as a HIP graph (mlv_graph_*)."""

    def __init__(self, dev: "Device", handle: int):
        self.dev, self.handle = dev, handle

    def launch(self) -> None:
        _check(self.dev.ctx, L.load().mlv_graph_launch(self.dev.ctx, self.handle))

    def release(self) -> None:
        if self.handle:
            L.load().mlv_graph_destroy(self.dev.ctx, self.handle)
            self.handle = None


class Device:
    """One mlv_context: the pipeline state plus one HIP stream on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.mlv_context_create(device, C.byref(h))
        if rc != L.OK:
            raise MalevichError(rc, "mlv_context_create failed (no usable GPU?)")
        self.ctx = h.value

    def close(self) -> None:
        if self.ctx:
            self.lib.mlv_context_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- resources
    def create_buffer(self, kind: int, data=None, nbytes: int | None = None) -> Buffer:
        a = None if data is None else np.ascontiguousarray(data)
        n = a.nbytes if a is not None else int(nbytes)
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_buffer_create(self.ctx, kind, n, None if a is None else a.ctypes.data,
                                                    C.byref(h)))
        return Buffer(self, h.value, n, kind)

    def wrap_buffer(self, kind: int, device_ptr: int, nbytes: int) -> Buffer:
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_buffer_wrap(self.ctx, kind, device_ptr, nbytes, C.byref(h)))
        return Buffer(self, h.value, nbytes, kind)

    def create_texture2d(self, data, fmt: int | None = None, mip_levels: int = 1,
                         generate_mips: bool = False) -> Texture2D:
        """data: (h, w) uint32 RGBA8 (R in the low byte), (h, w, 4) float32 or float16: level 0, or a
        list of every level (level 0 first). With mip_levels > 1 and level 0 alone, generate_mips
        fills the other levels on the device."""
        levels = [np.ascontiguousarray(x) for x in data] if isinstance(data, (list, tuple)) else \
            [np.ascontiguousarray(data)]
        a = levels[0]
        if fmt is None:
            fmt = {np.dtype(np.uint32): L.FORMAT_RGBA8_UNORM, np.dtype(np.float32): L.FORMAT_RGBA32F,
                   np.dtype(np.float16): L.FORMAT_RGBA16F}[a.dtype]
        h_, w_ = a.shape[:2]
        if len(levels) > 1:
            mip_levels = len(levels)
        sizes = _level_sizes(w_, h_, mip_levels)
        if len(levels) == 1 and mip_levels > 1:
            if not generate_mips:
                raise ValueError("mip_levels > 1 needs every level or generate_mips=True")
            rest = [np.zeros(_texel_shape(fmt, w, h)[0], a.dtype) for w, h in sizes[1:]]
            levels = levels + rest
        for lv, (w, h) in zip(levels, sizes):
            assert lv.shape[:2] == (h, w), (lv.shape, (h, w))
        flat = np.concatenate([lv.reshape(-1) for lv in levels]) if len(levels) > 1 else a
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_texture2d_create(self.ctx, w_, h_, fmt, mip_levels, flat.ctypes.data,
                                                       C.byref(h)))
        t = Texture2D(self, h.value, w_, h_, fmt, mip_levels)
        if generate_mips and mip_levels > 1:
            t.generate_mips()
        return t

    def create_render_target(self, width: int, height: int) -> RenderTarget:
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_render_target_create(self.ctx, width, height, C.byref(h)))
        return RenderTarget(self, h.value, width, height)

    def create_depth_stencil(self, width: int, height: int) -> DepthStencil:
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_depth_stencil_create(self.ctx, width, height, C.byref(h)))
        return DepthStencil(self, h.value, width, height)

    # ---------------------------------------------------------------- IA
    def ia_set_index_buffer(self, buf: Buffer | None, offset: int = 0) -> None:
        _check(self.ctx, self.lib.mlv_ia_set_index_buffer(self.ctx, buf.handle if buf else None, offset))

    def ia_set_vertex_buffer(self, buf: Buffer | None, stride: int, offset: int = 0) -> None:
        _check(self.ctx, self.lib.mlv_ia_set_vertex_buffer(self.ctx, buf.handle if buf else None, stride, offset))

    def ia_set_primitive_topology(self, topology: int = L.TOPOLOGY_TRIANGLELIST) -> None:
        _check(self.ctx, self.lib.mlv_ia_set_primitive_topology(self.ctx, topology))

    # ---------------------------------------------------------------- VS
    def vs_set_shader(self, vs: int) -> None:
        _check(self.ctx, self.lib.mlv_vs_set_shader(self.ctx, vs))

    def vs_set_constant_buffer(self, slot: int, buf: Buffer | None) -> None:
        _check(self.ctx, self.lib.mlv_vs_set_constant_buffer(self.ctx, slot, buf.handle if buf else None))

    def vs_set_shader_resource(self, slot: int, tex: Texture2D | None) -> None:
        _check(self.ctx, self.lib.mlv_vs_set_shader_resource(self.ctx, slot, tex.handle if tex else None))

    # ---------------------------------------------------------------- RS
    def rs_set_viewport(self, top_left_x=0.0, top_left_y=0.0, width=0.0, height=0.0, min_depth=0.0,
                        max_depth=1.0) -> None:
        vp = L.Viewport(top_left_x, top_left_y, width, height, min_depth, max_depth)
        _check(self.ctx, self.lib.mlv_rs_set_viewport(self.ctx, C.byref(vp)))

    def rs_set_cull(self, mode: int) -> None:
        """L.CULL_BACK (the reference's), L.CULL_NONE or L.CULL_FRONT."""
        _check(self.ctx, self.lib.mlv_rs_set_cull(self.ctx, mode))

    def rs_set_fill_rule(self, rule: int) -> None:
        """L.FILL_INCLUSIVE (the reference's) or L.FILL_TOP_LEFT."""
        _check(self.ctx, self.lib.mlv_rs_set_fill_rule(self.ctx, rule))

    # ---------------------------------------------------------------- PS
    def ps_set_shader(self, ps: int) -> None:
        _check(self.ctx, self.lib.mlv_ps_set_shader(self.ctx, ps))

    def ps_set_shader_resource(self, slot: int, tex: Texture2D | None) -> None:
        _check(self.ctx, self.lib.mlv_ps_set_shader_resource(self.ctx, slot, tex.handle if tex else None))

    def ps_set_constant_buffer(self, slot: int, buf: Buffer | None) -> None:
        """Extension: the Lambert shader's light (12 floats: L.xyz _, colour.rgb _, ambient.rgb _)."""
        _check(self.ctx, self.lib.mlv_ps_set_constant_buffer(self.ctx, slot, buf.handle if buf else None))

    # ---------------------------------------------------------------- OM
    def om_set_render_targets(self, rt: RenderTarget | None, ds: DepthStencil | None) -> None:
        _check(self.ctx, self.lib.mlv_om_set_render_targets(self.ctx, rt.handle if rt else None,
                                                            ds.handle if ds else None))

    def om_set_depth_func(self, func: int) -> None:
        _check(self.ctx, self.lib.mlv_om_set_depth_func(self.ctx, func))

    def om_set_visibility_buffer(self, tri_ids: Buffer | None, draw_tags: Buffer | None) -> None:
        _check(self.ctx, self.lib.mlv_om_set_visibility_buffer(
            self.ctx, tri_ids.device_ptr if tri_ids else None, draw_tags.device_ptr if draw_tags else None))

    def om_set_primitive_id_buffer(self, device_ptr: int | None) -> None:
        """Per-pixel input-triangle index of the visible triangle (u32 device array, W*H)."""
        _check(self.ctx, self.lib.mlv_om_set_primitive_id_buffer(self.ctx, device_ptr))

    # ---------------------------------------------------------- commands
    def clear_render_target_view(self, rt: RenderTarget, rgba) -> None:
        c = (C.c_float * 4)(*[float(x) for x in rgba])
        _check(self.ctx, self.lib.mlv_clear_render_target_view(self.ctx, rt.handle, c))

    def clear_depth_stencil_view(self, ds: DepthStencil, depth: float) -> None:
        _check(self.ctx, self.lib.mlv_clear_depth_stencil_view(self.ctx, ds.handle, float(depth)))

    def draw_indexed(self, index_count: int, start_index: int = 0, base_vertex: int = 0) -> None:
        _check(self.ctx, self.lib.mlv_draw_indexed(self.ctx, index_count, start_index, base_vertex))

    def synchronize(self) -> None:
        _check(self.ctx, self.lib.mlv_synchronize(self.ctx))

    @property
    def stream(self) -> int:
        return self.lib.mlv_context_stream(self.ctx)

    def stats(self) -> dict:
        s = L.Stats()
        _check(self.ctx, self.lib.mlv_get_stats(self.ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Stats._fields_}

    def set_profiling(self, enable: bool) -> None:
        _check(self.ctx, self.lib.mlv_context_set_profiling(self.ctx, int(bool(enable))))

    def profile(self) -> dict:
        """Per-stage GPU milliseconds since the last call (HIP events on the context stream)."""
        pr = L.Profile()
        _check(self.ctx, self.lib.mlv_get_profile(self.ctx, C.byref(pr)))
        out = {name: pr.stage_ms[i] for i, name in enumerate(L.STAGES)}
        out["draws"] = pr.draws
        out["flushes"] = pr.flushes
        out["tile_kernel"] = pr.tile_kernel_ms
        return out

    def set_footprint_counting(self, enable: bool) -> None:
        """Measurement mode: mark visible triangles and sampled texels (adds atomics; not for timing)."""
        _check(self.ctx, self.lib.mlv_context_set_footprint_counting(self.ctx, int(bool(enable))))

    def footprint(self) -> dict:
        """Visible triangles, pixels written, distinct texels and their bytes since the last call."""
        f = L.Footprint()
        _check(self.ctx, self.lib.mlv_get_footprint(self.ctx, C.byref(f)))
        return {k: getattr(f, k) for k, _ in L.Footprint._fields_}

    def flush(self) -> None:
        """Enqueue the recorded draws now (they otherwise run when their results are needed)."""
        _check(self.ctx, self.lib.mlv_flush(self.ctx))

    def reset_stats(self) -> None:
        _check(self.ctx, self.lib.mlv_reset_stats(self.ctx))

    # ---------------------------------------------------------- cameras
    def camera_constant_buffers(self, cameras, device_ptr: int) -> None:
        """cameras: [(position xyz, yaw_rad, pitch_rad, fov_y_deg, near_plane, width, height)];
        writes 48 floats per camera (the reference's PerFrameCB) at device_ptr on the device."""
        arr = (L.Camera * len(cameras))()
        for k, (pos, yaw, pitch, fov, near, w, h) in enumerate(cameras):
            arr[k] = L.Camera((C.c_float * 3)(*pos), yaw, pitch, fov, near, w, h)
        _check(self.ctx, self.lib.mlv_camera_constant_buffers(self.ctx, arr, len(cameras), device_ptr))

    # ---------------------------------------------------------- tuning
    def tuning(self) -> dict:
        t = L.Tuning()
        _check(self.ctx, self.lib.mlv_get_tuning(self.ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.Tuning._fields_}

    def set_tuning(self, heavy_tile_pairs: int, heavy_blocks: int, pair_blocks: int, tile_stage: int = 256,
                   tile_edge: int = 16, tile_waves_per_cu: int = 0, setup_waves_per_cu: int = 0,
                   skip_cornerless: int = 1, pair_capacity: int = 0) -> None:
        """Launch parameters (mlv_tuning); pair_capacity is reported by tuning() and ignored here."""
        t = L.Tuning(heavy_tile_pairs, heavy_blocks, pair_blocks, tile_stage, 0, tile_edge, tile_waves_per_cu,
                     setup_waves_per_cu, skip_cornerless)
        _check(self.ctx, self.lib.mlv_set_tuning(self.ctx, C.byref(t)))

    def autotune_begin(self) -> None:
        """Record the following clears and draws as the frame to tune on."""
        _check(self.ctx, self.lib.mlv_autotune_begin(self.ctx))

    def autotune_end(self, repeats: int = 2) -> dict:
        """Time the recorded frame under candidate launch parameters and keep the fastest."""
        t = L.Tuning()
        _check(self.ctx, self.lib.mlv_autotune_end(self.ctx, repeats, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.Tuning._fields_}

    # ---------------------------------------------------------- graphs
    def graph_capture_begin(self) -> None:
        """Record the following clears, draws and batched differentiable passes instead of running
        them."""
        _check(self.ctx, self.lib.mlv_graph_capture_begin(self.ctx))

    def graph_capture_end(self) -> Graph:
        """Run the recorded frame once and capture it as a graph to replay with Graph.launch()."""
        h = C.c_void_p()
        _check(self.ctx, self.lib.mlv_graph_capture_end(self.ctx, C.byref(h)))
        return Graph(self, h.value)
