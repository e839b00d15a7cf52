"""Pins of the FP64 differentiable-shading oracle (CPU only).

- The analytic adjoint equals central finite differences of the oracle's own forward pass
  for every parameter group (positions, colours, UVs, normals, texels, SH), at fixed
  visibility and mip choice.
- Perspective-correct barycentrics equal an independent FP64 ray / triangle-plane
  intersection at every covered pixel.
- The mip choice follows the texel footprint: a quad that maps k texels to a pixel picks the
  level nearest log2(k).
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import scenes as S
from tests.diff_utils import NONE, oracle_inputs, oracle_prim_ids


@pytest.fixture(scope="module")
def small(oracle):
    sc = S.c4_scene(width=64, height=64, rings=12, segments=16, tex_size=32, seed=9)
    prim = oracle_prim_ids(oracle, sc)
    assert (prim != NONE).sum() > 500
    return sc, prim


def _loss_and_grad(ob, d, prim, target):
    out = ob.diff_forward(d, prim)
    r = out - target
    return float((r * r).sum()), 2.0 * r


@pytest.mark.parametrize("group,h,count", [("positions", 1e-6, 12), ("colors", 1e-6, 12), ("uvs", 1e-7, 12),
                                           ("normals", 1e-6, 12), ("texture", 1e-6, 16), ("sh", 1e-6, 27)])
def test_adjoint_matches_central_differences(oracle, small, group, h, count):
    sc, prim = small
    d = oracle_inputs(oracle, sc)
    rng = np.random.default_rng(3)
    target = rng.uniform(0.0, 1.0, (sc.height, sc.width, 4))
    _, g_out = _loss_and_grad(oracle, d, prim, target)
    grads = oracle.diff_backward(d, prim, g_out)
    g = grads[group].ravel()
    arr = getattr(d, group)
    flat = arr.reshape(-1)
    # parameters that matter: the largest analytic gradients plus random ones
    nz = np.nonzero(g)[0]
    assert len(nz) > 0
    pick = list(nz[np.argsort(-np.abs(g[nz]))[: count // 2]]) + list(rng.choice(nz, count - count // 2))
    for i in pick:
        old = flat[i]
        flat[i] = old + h
        lp, _ = _loss_and_grad(oracle, d, prim, target)
        flat[i] = old - h
        lm, _ = _loss_and_grad(oracle, d, prim, target)
        flat[i] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (group, i, fd, g[i])


def test_barycentrics_equal_ray_plane_intersection(oracle, small):
    sc, prim = small
    m = sc.mesh
    d = oracle_inputs(oracle, sc)
    # colours = 1, texture = 1, SH = ambient 1: the output is srgb(1) everywhere; instead read the
    # barycentrics through the colour channels: vertex k of every triangle gets colour e_k
    # by de-indexing the mesh.
    tri = m.indices.astype(np.int64)
    T = len(tri)
    pos = m.positions[tri].reshape(-1, 3)
    col = np.tile(np.eye(3, dtype=np.float32), (T, 1))
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    ones = np.ones((len(sc.chain), 4), np.float32)
    sh = np.zeros((9, 3), np.float32)
    sh[0] = 1.0 / 0.282094792
    dd = oracle.DiffInputs(pos, col, m.uvs[tri].reshape(-1, 2), m.normals[tri].reshape(-1, 3), idx, ones,
                           sc.texture.shape[1], sc.texture.shape[0], sc.levels, sh, sc.clip_from_world,
                           sc.width, sc.height)
    out = oracle.diff_forward(dd, prim)
    ys, xs = np.nonzero(prim != NONE)
    # invert the sRGB curve: c = ((o + 0.055) / 1.055) ^ (1 / p)
    p = float(np.float32(0.416666667))
    lam = ((out[ys, xs, :3] + 0.055) / 1.055) ** (1.0 / p)
    # independent: intersect the pixel's view ray with the triangle's plane in world space (FP64)
    Mc = sc.clip_from_world.astype(np.float64)
    inv = np.linalg.inv(Mc)
    u = 2.0 * xs / sc.width - 1.0
    v = 1.0 - 2.0 * ys / sc.height
    near = (inv @ np.stack([u, v, np.full_like(u, 1.0), np.ones_like(u)]))
    far = (inv @ np.stack([u, v, np.full_like(u, 0.5), np.ones_like(u)]))
    o = (near[:3] / near[3]).T
    dvec = (far[:3] / far[3]).T - o
    P = m.positions.astype(np.float64)[tri[prim[ys, xs].astype(np.int64)]]
    e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    n = np.cross(e1, e2)
    t = np.einsum("ij,ij->i", P[:, 0] - o, n) / np.einsum("ij,ij->i", dvec, n)
    X = o + t[:, None] * dvec
    # barycentrics of X in the triangle (least squares in the plane)
    A = np.stack([e1, e2], 2)
    ATA = np.einsum("nki,nkj->nij", A, A)
    rhs = np.einsum("nki,nk->ni", A, X - P[:, 0])
    b12 = np.linalg.solve(ATA, rhs[..., None])[..., 0]
    ref = np.stack([1 - b12[:, 0] - b12[:, 1], b12[:, 0], b12[:, 1]], 1)
    inside = (ref > 1e-3).all(1)  # the sRGB inversion is exact only where every weight is positive
    assert inside.sum() > 200
    np.testing.assert_allclose(lam[inside], ref[inside], atol=2e-9, rtol=1e-7)


def test_mip_choice_follows_the_footprint(oracle):
    W = H = 32
    # a quad over the whole screen in clip space (identity matrix, w = 1), uv spanning [0, 1]
    pos = np.array([[-1.2, -1.2, 0.5], [1.2, -1.2, 0.5], [1.2, 1.2, 0.5], [-1.2, 1.2, 0.5]], np.float32)
    uv = np.array([[-0.1, -0.1], [1.1, -0.1], [1.1, 1.1], [-0.1, 1.1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    col = np.ones((4, 3), np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    sh = np.zeros((9, 3), np.float32)
    sh[0] = 1.0 / 0.282094792
    M = np.eye(4, dtype=np.float32)
    for size, expect in ((32, 0), (64, 1), (128, 2), (256, 3)):
        levels = int(np.log2(size)) + 1
        dims = [max(size >> l, 1) for l in range(levels)]
        chain = np.concatenate([np.full((dd * dd, 4), 0.05 * (l + 1), np.float32) for l, dd in enumerate(dims)])
        d = oracle.DiffInputs(pos, col, uv, nrm, idx, chain, size, size, levels, sh, M, W, H)
        prim = np.zeros((H, W), np.uint32)
        prim[:, W // 2:] = 1
        out = oracle.diff_forward(d, prim)
        p = float(np.float32(0.416666667))
        lin = ((out[..., :3] + 0.055) / 1.055) ** (1.0 / p)
        lvl = np.rint(lin[..., 0] / 0.05 - 1).astype(int)
        # uv spans 1.2 over 2.4 NDC units, i.e. 1/W per pixel: size/W texels per pixel
        assert (lvl == expect).mean() > 0.95, (size, np.bincount(lvl.ravel()))
