"""The C ABI library loads and exports every entry point include/malevich.h declares.

CPU-only: no compute call is made; without a GPU, context creation must report
MLV_ERROR_NO_DEVICE instead of aborting.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from malevich_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    text = (ROOT / "include" / "malevich.h").read_text()
    return set(re.findall(r"MLV_API\s+[^;(]*?\b(mlv_\w+)\s*\(", text))


def test_header_declares_the_reference_command_set():
    syms = declared_symbols()
    # the reference's three commands (main.c:1191, 1204, 1219) and its state parts (main.c:71-107)
    for name in ("mlv_clear_render_target_view", "mlv_clear_depth_stencil_view", "mlv_draw_indexed",
                 "mlv_ia_set_index_buffer", "mlv_ia_set_vertex_buffer", "mlv_ia_set_primitive_topology",
                 "mlv_vs_set_shader", "mlv_vs_set_constant_buffer", "mlv_vs_set_shader_resource",
                 "mlv_rs_set_viewport", "mlv_ps_set_shader", "mlv_ps_set_shader_resource",
                 "mlv_om_set_render_targets"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert declared_symbols() == set(L.SIGNATURES)


def test_load_sets_prototypes():
    lib = L.load()
    assert lib.mlv_draw_indexed.argtypes is not None


def test_null_arguments_are_rejected_without_a_device():
    lib = L.load()
    assert lib.mlv_context_create(0, None) == L.ERROR_INVALID_ARGUMENT
    assert lib.mlv_draw_indexed(None, 3, 0, 0) == L.ERROR_INVALID_ARGUMENT
    assert lib.mlv_synchronize(None) == L.ERROR_INVALID_ARGUMENT
    n = C.c_int(-1)
    assert lib.mlv_device_count(C.byref(n)) == L.OK
    assert n.value >= 0


def test_context_create_reports_missing_gpu():
    lib = L.load()
    n = C.c_int(0)
    lib.mlv_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    h = C.c_void_p()
    assert lib.mlv_context_create(0, C.byref(h)) == L.ERROR_NO_DEVICE
    assert not h.value
