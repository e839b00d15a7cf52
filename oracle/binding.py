"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (built by oracle/Makefile). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

VS_BASIC, VS_VERTEX_LIGHTING, VS_PASSTHROUGH, VS_FULLSCREEN = 0, 1, 2, 3
PS_BASIC, PS_ENV_LIGHTING, PS_PASSTHROUGH, PS_BASIC_LAMBERT, PS_ENV_TEXTURED, PS_NONE = 0, 1, 2, 3, 4, 5
FMT_RGBA8, FMT_RGBA32F = 0, 1
DEPTH_GE, DEPTH_LE = 0, 1
CULL_BACK, CULL_NONE, CULL_FRONT = 0, 1, 2  # extension; CULL_BACK is the reference's (main.c:856)
FILL_INCLUSIVE, FILL_TOP_LEFT = 0, 1  # extension; FILL_INCLUSIVE is the reference's (main.c:1015-1028)


class Texture(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("levels", C.c_uint32)]


class Viewport(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("top_left_x", "top_left_y", "width", "height", "min_depth", "max_depth")]


class DrawDesc(C.Structure):
    _fields_ = [
        ("index_buffer", C.c_void_p),
        ("index_count", C.c_uint32),
        ("start_index", C.c_uint32),
        ("base_vertex", C.c_int32),
        ("vertex_buffer", C.c_void_p),
        ("vertex_stride", C.c_uint32),
        ("vertex_count", C.c_uint32),
        ("vs", C.c_int32),
        ("ps", C.c_int32),
        ("cb", C.c_void_p),
        ("vs_srv0", Texture),
        ("ps_srv0", Texture),
        ("viewport", Viewport),
        ("depth_func", C.c_int32),
        ("ps_srv1", Texture),
        ("ps_cb", C.c_void_p),
        ("cull_mode", C.c_int32),
        ("fill_rule", C.c_int32),
    ]


class Target(C.Structure):
    _fields_ = [
        ("colors", C.c_void_p),
        ("depth", C.c_void_p),
        ("hiz", C.c_void_p),
        ("vis_tri", C.c_void_p),
        ("vis_draw", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("draw_counter", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("vertex_count", "input_triangle_count", "assembled_triangle_count",
                                           "active_bin_count", "total_triangle_count_in_bins",
                                           "oob_triangle_count")]


class Triangle(C.Structure):
    _fields_ = [
        ("edge", (C.c_int32 * 3) * 3),
        ("rw", C.c_float * 3),
        ("one_over_area", C.c_float),
        ("max_depth", C.c_float),
        ("min_bounds", C.c_int32 * 2),
        ("max_bounds", C.c_int32 * 2),
        ("prim", C.c_uint32),
        ("attr", (C.c_float * 12) * 3),
        ("excl", C.c_uint32),
    ]


class DiffScene(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("colors", C.c_void_p), ("uvs", C.c_void_p), ("normals", C.c_void_p),
        ("positions_f", C.c_void_p), ("uvs_f", C.c_void_p), ("indices", C.c_void_p),
        ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
        ("texture", C.c_void_p), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32), ("tex_levels", C.c_uint32),
        ("sh", C.c_void_p), ("clip_from_world", C.c_double * 16), ("clip_from_world_f", C.c_float * 16),
        ("width", C.c_uint32), ("height", C.c_uint32),
    ]


class DiffGrads(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("positions", "colors", "uvs", "normals", "texture", "sh")]


_lib = None


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("yaw_rad", C.c_float), ("pitch_rad", C.c_float),
                ("fov_y_deg", C.c_float), ("near_plane", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32)]


def camera_cb(position, yaw_rad, pitch_rad, fov_y_deg, near_plane, width, height) -> np.ndarray:
    """The reference's PerFrameCB (48 float32) for a camera state (main.c:1422-1562)."""
    c = Camera((C.c_float * 3)(*position), yaw_rad, pitch_rad, fov_y_deg, near_plane, width, height)
    out = np.empty(48, np.float32)
    lib().orc_camera_cb(C.byref(c), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def inv4(m: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.empty(16, np.float32)
    lib().orc_inv4(a.ctypes.data_as(C.POINTER(C.c_float)), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out.reshape(4, 4)


def det4(m: np.ndarray) -> float:
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    return float(lib().orc_det4(a.ctypes.data_as(C.POINTER(C.c_float))))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing; run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        L.orc_clear_render_target_view.argtypes = [P(Target), P(C.c_float)]
        L.orc_clear_depth_stencil_view.argtypes = [P(Target), C.c_float, C.c_int32]
        L.orc_draw_indexed.argtypes = [P(Target), P(DrawDesc), P(Stats)]
        L.orc_vertex_shader.argtypes = [P(DrawDesc), C.c_uint32, P(C.c_float)]
        L.orc_primitive_assembly.argtypes = [P(DrawDesc), P(Triangle), C.c_uint32, P(C.c_uint32), P(C.c_uint64)]
        L.orc_coverage_mask.argtypes = [P(C.c_int32), C.c_int32, C.c_int32]
        L.orc_coverage_mask.restype = C.c_uint64
        L.orc_coverage_mask_excl.argtypes = [P(C.c_int32), C.c_uint32, C.c_int32, C.c_int32]
        L.orc_coverage_mask_excl.restype = C.c_uint64
        L.orc_bilinear.argtypes = [P(Texture), C.c_float, C.c_float, P(C.c_float)]
        L.orc_nearest_level.argtypes = [C.c_float, C.c_uint32]
        L.orc_nearest_level.restype = C.c_uint32
        L.orc_bilinear_level.argtypes = [P(Texture), C.c_uint32, C.c_float, C.c_float, P(C.c_float)]
        L.orc_latlon.argtypes = [P(Texture), P(C.c_float), P(C.c_float)]
        L.orc_encode_output.argtypes = [P(C.c_float)]
        L.orc_encode_output.restype = C.c_uint32
        L.orc_encode_clear.argtypes = [P(C.c_float)]
        L.orc_encode_clear.restype = C.c_uint32
        L.orc_srgb_from_linear_approx.argtypes = [C.c_float]
        L.orc_srgb_from_linear_approx.restype = C.c_float
        L.orc_tonemap.argtypes = [C.c_float]
        L.orc_tonemap.restype = C.c_float
        L.orc_normalize3.argtypes = [P(C.c_float), P(C.c_float)]
        L.orc_cvt_rne.argtypes = [C.c_float]
        L.orc_cvt_rne.restype = C.c_int32
        L.orc_clip_planes.argtypes = [P(C.c_float)]
        L.orc_clip_polygon.argtypes = [P(C.c_float), C.c_int, P(C.c_float)]
        L.orc_clip_polygon.restype = C.c_int
        L.orc_threads.restype = C.c_int
        L.orc_srgb_to_linear_rgba8.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.orc_camera_cb.argtypes = [P(Camera), P(C.c_float)]
        L.orc_det4.argtypes = [P(C.c_float)]
        L.orc_det4.restype = C.c_float
        L.orc_inv4.argtypes = [P(C.c_float), P(C.c_float)]
        L.orc_diff_forward.argtypes = [P(DiffScene), C.c_void_p, C.c_void_p]
        L.orc_diff_backward.argtypes = [P(DiffScene), C.c_void_p, C.c_void_p, P(DiffGrads)]
        L.orc_sh9.argtypes = [P(C.c_double), P(C.c_double)]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


class MipTexture:
    """A texture with a chain of levels for the oracle: the levels (level 0 first) concatenated,
    RGBA8 as uint32 or float32 (RGBA16F texels are exact in float32)."""

    def __init__(self, levels: list):
        l0 = levels[0]
        self.width, self.height = l0.shape[1], l0.shape[0]
        self.levels = len(levels)
        if l0.dtype == np.uint32:
            self.flat = np.ascontiguousarray(np.concatenate([x.reshape(-1) for x in levels]))
            self.format = FMT_RGBA8
        else:
            self.flat = np.ascontiguousarray(np.concatenate([x.reshape(-1, 4) for x in levels]).astype(np.float32))
            self.format = FMT_RGBA32F


def make_texture(arr) -> Texture:
    """arr: (h, w) uint32 RGBA8, (h, w, 4) float32, or a MipTexture."""
    if arr is None:
        return Texture(None, 0, 0, 0, 1)
    if isinstance(arr, MipTexture):
        return Texture(arr.flat.ctypes.data, arr.width, arr.height, arr.format, arr.levels)
    arr = np.ascontiguousarray(arr)
    if arr.dtype == np.uint32:
        h, w = arr.shape
        return Texture(arr.ctypes.data, w, h, FMT_RGBA8, 1)
    assert arr.dtype == np.float32 and arr.ndim == 3 and arr.shape[2] == 4
    h, w, _ = arr.shape
    return Texture(arr.ctypes.data, w, h, FMT_RGBA32F, 1)


class RenderTarget:
    """Colour + depth + 8x8 tile minima (+ optional visibility) in host memory."""

    def __init__(self, width: int, height: int, with_vis: bool = True):
        self.width, self.height = width, height
        self.colors = np.zeros((height, width), np.uint32)
        self.depth = np.zeros((height, width), np.float32)
        self.hiz = np.zeros(((height // 8) * (width // 8),), np.float32)
        self.vis_tri = np.full((height, width), 0xFFFFFFFF, np.uint32) if with_vis else None
        self.vis_draw = np.full((height, width), 0xFFFFFFFF, np.uint32) if with_vis else None
        self.t = Target(self.colors.ctypes.data, self.depth.ctypes.data, self.hiz.ctypes.data,
                        _ptr(self.vis_tri), _ptr(self.vis_draw), width, height, 0)
        self.stats = Stats()

    def clear_render_target_view(self, rgba) -> None:
        c = (C.c_float * 4)(*rgba)
        rc = lib().orc_clear_render_target_view(C.byref(self.t), c)
        assert rc == 0, rc

    def clear_depth_stencil_view(self, depth: float, depth_func: int = DEPTH_GE) -> None:
        rc = lib().orc_clear_depth_stencil_view(C.byref(self.t), C.c_float(depth), depth_func)
        assert rc == 0, rc


class Draw:
    """Keeps the numpy inputs alive for the lifetime of the descriptor."""

    def __init__(self, indices, vertices, vs, ps, cb=None, vs_tex=None, ps_tex=None, viewport=None,
                 start_index=0, base_vertex=0, index_count=None, depth_func=DEPTH_GE, ps_tex1=None, ps_cb=None,
                 cull_mode=0, fill_rule=0):
        self.indices = np.ascontiguousarray(indices, np.uint32).ravel()
        self.vertices = np.ascontiguousarray(vertices, np.float32)
        assert self.vertices.ndim == 2
        self.cb = np.ascontiguousarray(cb if cb is not None else np.zeros(48), np.float32).ravel()
        assert self.cb.size == 48
        keep = lambda t: t if t is None or isinstance(t, MipTexture) else np.ascontiguousarray(t)
        self.vs_tex_arr, self.ps_tex_arr, self.ps_tex1_arr = keep(vs_tex), keep(ps_tex), keep(ps_tex1)
        self.ps_cb = None if ps_cb is None else np.ascontiguousarray(ps_cb, np.float32).ravel()
        assert self.ps_cb is None or self.ps_cb.size >= 12
        n = self.indices.size if index_count is None else index_count
        self.desc = DrawDesc(
            self.indices.ctypes.data, n, start_index, base_vertex,
            self.vertices.ctypes.data, self.vertices.shape[1] * 4, self.vertices.shape[0],
            vs, ps, self.cb.ctypes.data, make_texture(self.vs_tex_arr), make_texture(self.ps_tex_arr),
            Viewport(*(viewport if viewport is not None else (0, 0, 0, 0, 0, 1))), depth_func,
            make_texture(self.ps_tex1_arr), None if self.ps_cb is None else self.ps_cb.ctypes.data, cull_mode,
            fill_rule)


def draw_indexed(rt: RenderTarget, draw: Draw) -> None:
    rc = lib().orc_draw_indexed(C.byref(rt.t), C.byref(draw.desc), C.byref(rt.stats))
    if rc != 0:
        raise RuntimeError(f"orc_draw_indexed failed: {rc}")


def primitive_assembly(draw: Draw, capacity: int | None = None) -> list:
    """Runs VS + primitive assembly; returns a structured array of Triangle records."""
    T = draw.desc.index_count // 3
    cap = capacity or (T + max(2 * T, 512))
    buf = (Triangle * cap)()
    n = C.c_uint32(0)
    oob = C.c_uint64(0)
    rc = lib().orc_primitive_assembly(C.byref(draw.desc), buf, cap, C.byref(n), C.byref(oob))
    if rc != 0:
        raise RuntimeError(f"orc_primitive_assembly failed: {rc}")
    return list(buf[: n.value])


TRIANGLE_DTYPE = np.dtype([("edge", np.int32, (3, 3)), ("rw", np.float32, 3), ("one_over_area", np.float32),
                           ("max_depth", np.float32), ("min_bounds", np.int32, 2), ("max_bounds", np.int32, 2),
                           ("prim", np.uint32), ("attr", np.float32, (3, 12)), ("excl", np.uint32)])
assert TRIANGLE_DTYPE.itemsize == C.sizeof(Triangle)


def primitive_assembly_array(draw: Draw, capacity: int | None = None) -> np.ndarray:
    """As primitive_assembly, as a structured numpy array (for millions of triangles)."""
    T = draw.desc.index_count // 3
    cap = capacity or (T + max(2 * T, 512))
    out = np.zeros(cap, TRIANGLE_DTYPE)
    n = C.c_uint32(0)
    oob = C.c_uint64(0)
    rc = lib().orc_primitive_assembly(C.byref(draw.desc), out.ctypes.data_as(C.POINTER(Triangle)), cap,
                                      C.byref(n), C.byref(oob))
    if rc != 0:
        raise RuntimeError(f"orc_primitive_assembly failed: {rc}")
    return out[: n.value].copy()


def vertex_shader(draw: Draw, vertex_index: int) -> np.ndarray:
    """The 12-float VS output record of one vertex (registers 0..2, main.c:714-727)."""
    out = np.zeros(12, np.float32)
    rc = lib().orc_vertex_shader(C.byref(draw.desc), vertex_index, out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise RuntimeError(f"orc_vertex_shader failed: {rc}")
    return out


def sh9(n) -> np.ndarray:
    """The oracle's 9 real SH basis values at the unit direction n."""
    d = (C.c_double * 3)(*[float(x) for x in n])
    out = (C.c_double * 9)()
    lib().orc_sh9(d, out)
    return np.array(out[:], np.float64)


def coverage_mask(edges, ox: int, oy: int, excl: int = 0) -> int:
    e = (C.c_int32 * 9)(*np.asarray(edges, np.int64).astype(np.int32).ravel().tolist())
    return int(lib().orc_coverage_mask_excl(e, excl, ox, oy))


def bilinear_level(tex, level: int, u: float, v: float) -> np.ndarray:
    t = make_texture(tex)
    out = (C.c_float * 4)()
    lib().orc_bilinear_level(C.byref(t), level, C.c_float(u), C.c_float(v), out)
    return np.array(out[:], np.float32)


def nearest_level(rho2: float, levels: int) -> int:
    return int(lib().orc_nearest_level(C.c_float(rho2), levels))


def bilinear(tex: np.ndarray, u: float, v: float) -> np.ndarray:
    t = make_texture(tex)
    out = (C.c_float * 4)()
    lib().orc_bilinear(C.byref(t), C.c_float(u), C.c_float(v), out)
    return np.array(out[:], np.float32)


def latlon(tex: np.ndarray, d) -> np.ndarray:
    t = make_texture(tex)
    dd = (C.c_float * 3)(*[float(x) for x in d])
    out = (C.c_float * 4)()
    lib().orc_latlon(C.byref(t), dd, out)
    return np.array(out[:], np.float32)


def encode_output(rgb) -> int:
    return int(lib().orc_encode_output((C.c_float * 3)(*[float(x) for x in rgb])))


def encode_clear(rgba) -> int:
    return int(lib().orc_encode_clear((C.c_float * 4)(*[float(x) for x in rgba])))


def clip_polygon(vertices: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    n = v.shape[0]
    out = np.zeros((16, 12), np.float32)
    k = lib().orc_clip_polygon(v.ctypes.data_as(C.POINTER(C.c_float)), n, out.ctypes.data_as(C.POINTER(C.c_float)))
    if k < 0:
        raise RuntimeError("clip polygon overflow")
    return out[:k].copy()


def clip_planes() -> np.ndarray:
    out = np.zeros((6, 4), np.float32)
    lib().orc_clip_planes(out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


class DiffInputs:
    """FP64 copies of a differentiable-shading scene (the f32 originals kept for the mip choice).

    positions/colors/normals (V,3), uvs (V,2), indices (T,3), texture_chain (N,4) RGBA mip chain,
    sh (9,3), clip_from_world (4,4), image width/height."""

    def __init__(self, positions, colors, uvs, normals, indices, texture_chain, tex_width, tex_height, tex_levels,
                 sh, clip_from_world, width, height):
        self.positions_f = np.ascontiguousarray(positions, np.float32)
        self.uvs_f = np.ascontiguousarray(uvs, np.float32)
        self.positions = self.positions_f.astype(np.float64)
        self.colors = np.ascontiguousarray(colors, np.float32).astype(np.float64)
        self.uvs = self.uvs_f.astype(np.float64)
        self.normals = np.ascontiguousarray(normals, np.float32).astype(np.float64)
        self.indices = np.ascontiguousarray(indices, np.uint32)
        self.texture = np.ascontiguousarray(texture_chain, np.float32).astype(np.float64)
        self.tex_width, self.tex_height, self.tex_levels = tex_width, tex_height, tex_levels
        self.sh = np.ascontiguousarray(sh, np.float32).astype(np.float64)
        self.M_f = np.ascontiguousarray(clip_from_world, np.float32).reshape(16)
        self.M = self.M_f.astype(np.float64)
        self.width, self.height = width, height

    def struct(self) -> DiffScene:
        for name in ("positions", "colors", "uvs", "normals", "texture", "sh"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), np.float64))
        return DiffScene(self.positions.ctypes.data, self.colors.ctypes.data, self.uvs.ctypes.data,
                         self.normals.ctypes.data, self.positions_f.ctypes.data, self.uvs_f.ctypes.data,
                         self.indices.ctypes.data, len(self.positions), len(self.indices),
                         self.texture.ctypes.data, self.tex_width, self.tex_height, self.tex_levels,
                         self.sh.ctypes.data, (C.c_double * 16)(*self.M.tolist()),
                         (C.c_float * 16)(*self.M_f.tolist()), self.width, self.height)


def diff_forward(d: DiffInputs, prim: np.ndarray) -> np.ndarray:
    prim = np.ascontiguousarray(prim, np.uint32)
    out = np.zeros((d.height, d.width, 4), np.float64)
    sc = d.struct()
    rc = lib().orc_diff_forward(C.byref(sc), prim.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_diff_forward failed: {rc}")
    return out


def diff_backward(d: DiffInputs, prim: np.ndarray, grad_rgba: np.ndarray) -> dict:
    prim = np.ascontiguousarray(prim, np.uint32)
    g_out = np.ascontiguousarray(grad_rgba, np.float64)
    grads = {"positions": np.zeros_like(d.positions), "colors": np.zeros_like(d.colors),
             "uvs": np.zeros_like(d.uvs), "normals": np.zeros_like(d.normals),
             "texture": np.zeros_like(d.texture), "sh": np.zeros_like(d.sh)}
    gs = DiffGrads(*[grads[k].ctypes.data for k in ("positions", "colors", "uvs", "normals", "texture", "sh")])
    sc = d.struct()
    rc = lib().orc_diff_backward(C.byref(sc), prim.ctypes.data, g_out.ctypes.data, C.byref(gs))
    if rc != 0:
        raise RuntimeError(f"orc_diff_backward failed: {rc}")
    return grads


def srgb_to_linear_rgba8(texels: np.ndarray) -> np.ndarray:
    """load_texture's sRGB pre-pass (main.c:546-558) over u32 RGBA8 texels."""
    src = np.ascontiguousarray(texels, np.uint32)
    out = np.empty_like(src)
    lib().orc_srgb_to_linear_rgba8(src.ctypes.data, out.ctypes.data, src.size)
    return out


def threads() -> int:
    """OpenMP threads the oracle's assembly and bin loops run on."""
    return int(lib().orc_threads())

