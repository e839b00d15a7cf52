"""mlv_diff_backward_noalloc (the differentiable backward pass on the library's fixed scratch
array, no allocation) against mlv_diff_backward (atomic mode) and the FP64 oracle.

Bar (written here): on two calls in a row, the second on a smaller scene and image, every
gradient group of the fixed-scratch entry point is within 1e-6 relative of the existing entry
point's (both sum with atomics, in a run-dependent order) and within 1e-4 relative of the FP64
oracle's adjoint. Calls with recorded draws pending, or needing more scratch than the array
holds, are refused before any launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from malevich_amd import _lib as L
from malevich_amd import diff as D
from malevich_amd import pipeline as P
from malevich_amd import scenes as S
from tests.diff_utils import oracle_inputs, rel_err

pytestmark = pytest.mark.gpu

GROUPS = ("positions", "colors", "uvs", "normals", "texture", "sh")


def _pass(device, oracle, sc, seed):
    m = sc.mesh
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    t = dict(positions=cuda(m.positions), colors=cuda(m.colors), uvs=cuda(m.uvs), normals=cuda(m.normals),
             indices=cuda(m.indices.astype(np.int32)), chain=cuda(sc.chain), sh=cuda(sc.sh))
    M = torch.from_numpy(sc.clip_from_world.copy())
    r = D.DiffRenderer(device, sc.width, sc.height, deterministic=False)
    try:
        verts = torch.cat([t["positions"], t["normals"], t["uvs"]], 1).contiguous()
        eye = torch.eye(4).reshape(-1)
        cb = torch.cat([M.reshape(-1), eye, eye]).cuda()
        prim = r.rasterize(verts, t["indices"], cb).clone()
        st = r.scene_struct(t["positions"], t["colors"], t["uvs"], t["normals"], t["indices"], t["chain"],
                            sc.texture.shape[1], sc.texture.shape[0], sc.levels, t["sh"], M)
        out = r.shade(st, prim)
        torch.cuda.synchronize()
        rng = np.random.default_rng(seed)
        go_np = (rng.normal(0, 1, (sc.height, sc.width, 4)) * (out.cpu().numpy()[..., 3:] > 0)).astype(np.float32)
        go = cuda(go_np)
        want = {k: (t[k if k != "texture" else "chain"].shape) for k in GROUPS}
        ref = {k: v.cpu().numpy() for k, v in r.shade_backward(st, prim, go, want).items()}
        got = {k: torch.full(shape, float("nan"), dtype=torch.float32, device="cuda") for k, shape in want.items()}
        gs = L.DiffGrads(*[got[k].data_ptr() for k in GROUPS])
        torch.cuda.synchronize()
        device.synchronize()
        rc = device.lib.mlv_diff_backward_noalloc(device.ctx, C.byref(st), prim.data_ptr(), go.data_ptr(),
                                                  C.byref(gs))
        P._check(device.ctx, rc)
        device.synchronize()
        got = {k: v.cpu().numpy() for k, v in got.items()}
        g64 = oracle.diff_backward(oracle_inputs(oracle, sc), prim.cpu().numpy().view(np.uint32),
                                   go_np.astype(np.float64))
        for k in GROUPS:
            assert np.all(np.isfinite(got[k])), k
            e_ref = rel_err(got[k], ref[k])
            e_64 = rel_err(got[k], g64[k])
            print(f"  {sc.width}x{sc.height} {k}: vs mlv_diff_backward {e_ref:.1e}, vs FP64 oracle {e_64:.1e}",
                  flush=True)
            assert e_ref <= 1e-6, (k, e_ref)
            assert e_64 <= 1e-4, (k, e_64)
        return st, prim, go
    finally:
        r.release()


def test_noalloc_backward_matches_the_allocating_entry_point(device, oracle):
    big = S.c4_scene(width=96, height=64, rings=16, segments=24, tex_size=64, seed=5)
    small = S.c4_scene(width=48, height=32, rings=10, segments=14, tex_size=32, seed=6)
    print("case: noalloc backward, 96x64 then 48x32", flush=True)
    _pass(device, oracle, big, 1)
    st, prim, go = _pass(device, oracle, small, 2)

    # refused before any launch: more scratch than the array holds (24 B per vertex)
    huge = L.DiffScene.from_buffer_copy(st)
    huge.vertex_count = 2_000_000
    gs = L.DiffGrads(None, None, None, None, None, None)
    rc = device.lib.mlv_diff_backward_noalloc(device.ctx, C.byref(huge), prim.data_ptr(), go.data_ptr(),
                                              C.byref(gs))
    assert rc == L.ERROR_CAPACITY, rc
    # refused with draws recorded and not flushed
    rt = device.create_render_target(16, 16)
    ds = device.create_depth_stencil(16, 16)
    vb = device.create_buffer(L.BUFFER_VERTEX, S.SUPREMATISM_VERTICES)
    ib = device.create_buffer(L.BUFFER_INDEX_U32, S.SUPREMATISM_INDICES)
    try:
        device.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        device.ia_set_vertex_buffer(vb, S.SUPREMATISM_VERTICES.shape[1] * 4)
        device.ia_set_index_buffer(ib)
        device.vs_set_shader(L.VS_PASSTHROUGH)
        device.ps_set_shader(L.PS_PASSTHROUGH)
        device.rs_set_viewport(0, 0, 16, 16, 0, 1)
        device.om_set_render_targets(rt, ds)
        device.draw_indexed(S.SUPREMATISM_INDICES.size)
        rc = device.lib.mlv_diff_backward_noalloc(device.ctx, C.byref(st), prim.data_ptr(), go.data_ptr(),
                                                  C.byref(gs))
        assert rc == L.ERROR_INVALID_STATE, rc
        device.flush()
        device.synchronize()
    finally:
        for h in (vb, ib, rt, ds):
            h.release()
