"""ctypes loader for libmalevich_hip.so (the C ABI declared in include/malevich.h).

The library is built in-tree by malevich_amd/build.py. There is no fallback:
if the shared library is missing, loading raises, so nothing ever runs a
host-side substitute of the GPU pipeline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmalevich_hip.so"

OK = 0
ERROR_INVALID_ARGUMENT = -1
ERROR_OUT_OF_MEMORY = -2
ERROR_DEVICE = -3
ERROR_INVALID_STATE = -4
ERROR_CAPACITY = -5
ERROR_NO_DEVICE = -6

BUFFER_VERTEX, BUFFER_INDEX_U32, BUFFER_CONSTANT = 1, 2, 3
FORMAT_RGBA8_UNORM, FORMAT_RGBA32F, FORMAT_RGBA16F = 1, 2, 3
TOPOLOGY_UNDEFINED, TOPOLOGY_TRIANGLELIST = 0, 1
VS_BASIC, VS_VERTEX_LIGHTING, VS_PASSTHROUGH, VS_FULLSCREEN = 1, 2, 3, 4
PS_BASIC, PS_ENV_LIGHTING, PS_PASSTHROUGH, PS_BASIC_LAMBERT, PS_ENV_TEXTURED, PS_NONE = 1, 2, 3, 4, 5, 6
DEPTH_GREATER_EQUAL, DEPTH_LESS_EQUAL = 0, 1
CULL_BACK, CULL_NONE, CULL_FRONT = 0, 1, 2
FILL_INCLUSIVE, FILL_TOP_LEFT = 0, 1


class Viewport(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("top_left_x", "top_left_y", "width", "height", "min_depth", "max_depth")]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "vertex_count", "input_triangle_count", "assembled_triangle_count", "active_bin_count",
        "total_triangle_count_in_bins", "oob_triangle_count", "draw_count", "overflow_count")]


class Profile(C.Structure):
    _fields_ = [("stage_ms", C.c_double * 4), ("draws", C.c_uint64), ("flushes", C.c_uint64),
                ("tile_kernel_ms", C.c_double)]


class Footprint(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("visible_triangles", "shaded_pixels", "texels", "texel_bytes")]


STAGES = ("vertex", "setup", "binning", "tile")

_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_SZ = C.c_size_t
_PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/malevich.h
class DiffScene(C.Structure):
    """mlv_diff_scene"""
    _fields_ = [
        ("positions", C.c_void_p), ("colors", C.c_void_p), ("uvs", C.c_void_p), ("normals", C.c_void_p),
        ("indices", C.c_void_p), ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
        ("texture", C.c_void_p), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32), ("tex_levels", C.c_uint32),
        ("sh", C.c_void_p), ("clip_from_world", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32),
    ]


class DiffGrads(C.Structure):
    """mlv_diff_grads"""
    _fields_ = [(n, C.c_void_p) for n in ("positions", "colors", "uvs", "normals", "texture", "sh")]


class Camera(C.Structure):
    """mlv_camera"""
    _fields_ = [("position", C.c_float * 3), ("yaw_rad", C.c_float), ("pitch_rad", C.c_float),
                ("fov_y_deg", C.c_float), ("near_plane", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32)]


class Tuning(C.Structure):
    """mlv_tuning"""
    _fields_ = [(n, C.c_uint32) for n in ("heavy_tile_pairs", "heavy_blocks", "pair_blocks", "tile_stage",
                                          "pair_capacity", "tile_edge", "tile_waves_per_cu", "setup_waves_per_cu",
                                          "skip_cornerless")]


class DiffMesh(C.Structure):
    """mlv_diff_mesh"""
    _fields_ = [
        ("positions", C.c_void_p), ("colors", C.c_void_p), ("uvs", C.c_void_p), ("normals", C.c_void_p),
        ("indices", C.c_void_p), ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
        ("texture", C.c_void_p), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32), ("tex_levels", C.c_uint32),
        ("grad_positions", C.c_void_p), ("grad_colors", C.c_void_p), ("grad_uvs", C.c_void_p),
        ("grad_normals", C.c_void_p), ("grad_texture", C.c_void_p), ("pos_accum", C.c_void_p),
    ]


class DiffView(C.Structure):
    """mlv_diff_view"""
    _fields_ = [
        ("clip_from_world", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32),
        ("prim_ids", C.c_void_p), ("draw_tags", C.c_void_p), ("first_tag", C.c_uint32), ("draw_count", C.c_uint32),
        ("draw_mesh", C.POINTER(C.c_uint32)), ("rgba", C.c_void_p), ("grad_rgba", C.c_void_p),
    ]


SIGNATURES = {
    "mlv_context_create": (_I, [_I, _PP]),
    "mlv_context_destroy": (_I, [_P]),
    "mlv_get_last_error": (C.c_char_p, [_P]),
    "mlv_synchronize": (_I, [_P]),
    "mlv_context_stream": (_P, [_P]),
    "mlv_device_count": (_I, [C.POINTER(C.c_int)]),
    "mlv_buffer_create": (_I, [_P, _I, _SZ, _P, _PP]),
    "mlv_buffer_wrap": (_I, [_P, _I, _P, _SZ, _PP]),
    "mlv_buffer_update": (_I, [_P, _P, _SZ, _SZ, _P]),
    "mlv_buffer_readback": (_I, [_P, _P, _SZ, _SZ, _P]),
    "mlv_buffer_destroy": (_I, [_P, _P]),
    "mlv_buffer_device_ptr": (_P, [_P]),
    "mlv_texture2d_create": (_I, [_P, _U32, _U32, _I, _U32, _P, _PP]),
    "mlv_texture2d_destroy": (_I, [_P, _P]),
    "mlv_texture2d_generate_mips": (_I, [_P, _P]),
    "mlv_texture2d_srgb_to_linear": (_I, [_P, _P]),
    "mlv_texture2d_readback": (_I, [_P, _P, _P]),
    "mlv_texture2d_device_ptr": (_P, [_P]),
    "mlv_render_target_create": (_I, [_P, _U32, _U32, _PP]),
    "mlv_render_target_destroy": (_I, [_P, _P]),
    "mlv_render_target_device_ptr": (_P, [_P]),
    "mlv_depth_stencil_create": (_I, [_P, _U32, _U32, _PP]),
    "mlv_depth_stencil_destroy": (_I, [_P, _P]),
    "mlv_depth_stencil_device_ptr": (_P, [_P]),
    "mlv_ia_set_index_buffer": (_I, [_P, _P, _U32]),
    "mlv_ia_set_vertex_buffer": (_I, [_P, _P, _U32, _U32]),
    "mlv_ia_set_primitive_topology": (_I, [_P, _I]),
    "mlv_vs_set_shader": (_I, [_P, _I]),
    "mlv_vs_set_constant_buffer": (_I, [_P, _U32, _P]),
    "mlv_vs_set_shader_resource": (_I, [_P, _U32, _P]),
    "mlv_rs_set_viewport": (_I, [_P, C.POINTER(Viewport)]),
    "mlv_rs_set_cull": (_I, [_P, _I]),
    "mlv_rs_set_fill_rule": (_I, [_P, _I]),
    "mlv_ps_set_shader": (_I, [_P, _I]),
    "mlv_ps_set_shader_resource": (_I, [_P, _U32, _P]),
    "mlv_ps_set_constant_buffer": (_I, [_P, _U32, _P]),
    "mlv_om_set_render_targets": (_I, [_P, _P, _P]),
    "mlv_om_set_depth_func": (_I, [_P, _I]),
    "mlv_om_set_visibility_buffer": (_I, [_P, _P, _P]),
    "mlv_om_set_primitive_id_buffer": (_I, [_P, _P]),
    "mlv_clear_render_target_view": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "mlv_clear_depth_stencil_view": (_I, [_P, _P, C.c_float]),
    "mlv_draw_indexed": (_I, [_P, _U32, _U32, C.c_int32]),
    "mlv_render_target_readback": (_I, [_P, _P, _P]),
    "mlv_present": (_I, [_P, _P, _P]),
    "mlv_depth_stencil_readback": (_I, [_P, _P, _P]),
    "mlv_get_stats": (_I, [_P, C.POINTER(Stats)]),
    "mlv_reset_stats": (_I, [_P]),
    "mlv_context_set_profiling": (_I, [_P, _I]),
    "mlv_context_set_footprint_counting": (_I, [_P, _I]),
    "mlv_get_footprint": (_I, [_P, C.POINTER(Footprint)]),
    "mlv_flush": (_I, [_P]),
    "mlv_get_profile": (_I, [_P, C.POINTER(Profile)]),
    "mlv_camera_constant_buffers": (_I, [_P, C.POINTER(Camera), _U32, _P]),
    "mlv_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "mlv_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "mlv_autotune_begin": (_I, [_P]),
    "mlv_autotune_end": (_I, [_P, _U32, C.POINTER(Tuning)]),
    "mlv_graph_capture_begin": (_I, [_P]),
    "mlv_graph_capture_end": (_I, [_P, _PP]),
    "mlv_graph_launch": (_I, [_P, _P]),
    "mlv_graph_destroy": (_I, [_P, _P]),
    "mlv_diff_forward": (_I, [_P, C.POINTER(DiffScene), _P, _P]),
    "mlv_diff_backward": (_I, [_P, C.POINTER(DiffScene), _P, _P, C.POINTER(DiffGrads), _I]),
    "mlv_diff_backward_noalloc": (_I, [_P, C.POINTER(DiffScene), _P, _P, C.POINTER(DiffGrads)]),
    "mlv_diff_batch_forward": (_I, [_P, C.POINTER(DiffMesh), _U32, C.POINTER(DiffView), _U32, _P]),
    "mlv_diff_batch_backward": (_I, [_P, C.POINTER(DiffMesh), _U32, C.POINTER(DiffView), _U32, _P, _P, _I]),
    "mlv_diff_l2_grad": (_I, [_P, _P, _P, C.c_uint64, C.c_float, _P]),
    "mlv_mip_chain_size": (_I, [_U32, _U32, C.POINTER(_U32), C.POINTER(_U32)]),
    "mlv_mip_chain": (_I, [_P, _P, _U32, _U32, _P]),
    "mlv_mip_chain_adjoint": (_I, [_P, _P, _U32, _U32, _P]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the pipeline has no CPU fallback)")
        # PyTorch-ROCm ships its own HIP runtime under the same soname (libamdhip64.so.7). A
        # process that loads this library first and torch afterwards would hold two HIP
        # runtimes; loading torch first makes the library bind to torch's runtime instead.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class MalevichError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def mip_chain_size(width: int, height: int) -> tuple[int, int]:
    """(texels, levels) of the mip chain mlv_mip_chain builds for a width x height texture."""
    n, levels = C.c_uint32(0), C.c_uint32(0)
    rc = load().mlv_mip_chain_size(width, height, C.byref(n), C.byref(levels))
    if rc != 0:
        raise MalevichError(rc, f"mip_chain_size: bad size {width}x{height}")
    return int(n.value), int(levels.value)
