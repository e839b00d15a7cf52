"""Helpers shared by the differentiable-shading tests."""
from __future__ import annotations

import numpy as np

NONE = 0xFFFFFFFF


def oracle_prim_ids(ob, sc) -> np.ndarray:
    """Primitive id per pixel from the oracle's raster of the scene (basic_vs, GREATER_EQUAL,
    clear depth 0): the assembled-triangle visibility mapped to input triangles."""
    W, H = sc.width, sc.height
    verts = sc.mesh.basic_vertices()
    idx = sc.mesh.indices.reshape(-1)
    rt = ob.RenderTarget(W, H)
    rt.clear_render_target_view((0.0, 0.0, 0.0, 0.0))
    rt.clear_depth_stencil_view(0.0)
    d = ob.Draw(idx, verts, ob.VS_BASIC, ob.PS_PASSTHROUGH, cb=sc.cb, viewport=(0, 0, W, H, 0, 1))
    ob.draw_indexed(rt, d)
    tris = ob.primitive_assembly_array(d)
    vis = rt.vis_tri.copy()
    prim = np.full((H, W), NONE, np.uint32)
    m = vis != NONE
    prim[m] = tris["prim"][vis[m].astype(np.int64)]
    return prim


def oracle_inputs(ob, sc):
    m = sc.mesh
    return ob.DiffInputs(m.positions, m.colors, m.uvs, m.normals, m.indices, sc.chain, sc.texture.shape[1],
                         sc.texture.shape[0], sc.levels, sc.sh, sc.clip_from_world, sc.width, sc.height)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))
