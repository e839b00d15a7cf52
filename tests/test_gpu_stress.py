"""Binning stress at 1920x1080: tiny random triangles with screen-covering ones interleaved
in submission order, passthrough_vs/ps, reversed-Z GREATER_EQUAL.

- 1M tiny + 100 big: every pixel's visibility, depth and colour equal the oracle's.
- 10M tiny + 1k big (the full stress size): the oracle's primitive assembly of the same draw
  gives the assembled triangles; every covered pixel's winner covers it (integer edge test) and
  its depth is that triangle's interpolated depth, bit for bit; on sampled 8x8 bins the winner is
  the brute-force maximum of (depth, submission order) over all triangles covering the pixel;
  and two renders of the frame are identical.
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
NONE = 0xFFFFFFFF


def _wrap32(x: np.ndarray) -> np.ndarray:
    return ((x.astype(np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)).astype(np.int64)


def _edges_at(tri: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """E_k = ((a*x)<<4) + ((b*y)<<4) + c in i32 arithmetic (main.c:1015-1025) -> (..., 3) int64."""
    a = tri["edge"][..., 0].astype(np.int64)
    b = tri["edge"][..., 1].astype(np.int64)
    c = tri["edge"][..., 2].astype(np.int64)
    return _wrap32(a * x[..., None] * 16 + b * y[..., None] * 16 + c)


def _depth_at(tri: np.ndarray, e: np.ndarray) -> np.ndarray:
    """Register-0 interpolation with the quantised barycentrics (main.c:1083-1154), f32."""
    ooa = tri["one_over_area"].astype(np.float32)
    b1 = (e[..., 1] >> 8).astype(np.float32) * ooa
    b2 = (e[..., 2] >> 8).astype(np.float32) * ooa
    z0 = tri["attr"][..., 0, 2]
    z1 = tri["attr"][..., 1, 2]
    z2 = tri["attr"][..., 2, 2]
    z = z0 + (z1 - z0) * b1
    return (z + (z2 - z0) * b2).astype(np.float32)


def test_c3_binning_stress_1080p_1m_exact(device, oracle):
    st = S.binning_stress(1_000_000, 100, W, H, seed=21)
    d = DrawSpec(st.indices, st.vertices, "passthrough", "passthrough")
    print("case: 1M tiny + 100 big at 1920x1080", flush=True)
    g = render_gpu(device, W, H, [d], (0.0, 0.0, 0.0, 0.0))
    o = render_oracle(oracle, W, H, [d], (0.0, 0.0, 0.0, 0.0))
    assert_frames_match(g, o, color_tol=0)


@pytest.mark.timeout(600)
def test_c3_full_size_winners(device, oracle):
    st = S.binning_stress(10_000_000, 1000, W, H, seed=3)
    d = DrawSpec(st.indices, st.vertices, "passthrough", "passthrough")
    print("case: 10M tiny + 1k big at 1920x1080, render 1", flush=True)
    g = render_gpu(device, W, H, [d], (0.0, 0.0, 0.0, 0.0))
    print("case: 10M tiny + 1k big at 1920x1080, render 2", flush=True)
    g2 = render_gpu(device, W, H, [d], (0.0, 0.0, 0.0, 0.0))
    np.testing.assert_array_equal(g.vis_tri, g2.vis_tri)
    np.testing.assert_array_equal(g.depth.view(np.uint32), g2.depth.view(np.uint32))
    np.testing.assert_array_equal(g.colors, g2.colors)

    od = oracle.Draw(st.indices, st.vertices, oracle.VS_PASSTHROUGH, oracle.PS_PASSTHROUGH,
                     viewport=(0, 0, W, H, 0, 1))
    tris = oracle.primitive_assembly_array(od)
    assert g.stats["assembled_triangle_count"] == len(tris)
    assert g.stats["input_triangle_count"] == 10_001_000

    # every covered pixel: its winner covers it and the stored depth is the winner's depth
    ys, xs = np.nonzero(g.vis_tri != NONE)
    assert len(ys) > 0.9 * W * H
    win = tris[g.vis_tri[ys, xs].astype(np.int64)]
    e = _edges_at(win, xs.astype(np.int64), ys.astype(np.int64))
    assert ((e[:, 0] | e[:, 1] | e[:, 2]) > 0).all(), "a winner does not cover its pixel"
    np.testing.assert_array_equal(_depth_at(win, e).view(np.uint32), g.depth[ys, xs].view(np.uint32))

    # sampled bins: brute-force maximum of (depth, order) over every covering triangle
    rng = np.random.default_rng(5)
    mnx, mny = tris["min_bounds"][:, 0], tris["min_bounds"][:, 1]
    mxx, mxy = tris["max_bounds"][:, 0], tris["max_bounds"][:, 1]
    for _ in range(48):
        bx, by = int(rng.integers(0, W // 8)), int(rng.integers(0, H // 8))
        x0, y0 = bx * 8, by * 8
        cand = np.nonzero((mnx <= x0 + 7) & (mxx >= x0) & (mny <= y0 + 7) & (mxy >= y0))[0]
        px, py = np.meshgrid(np.arange(x0, x0 + 8), np.arange(y0, y0 + 8))
        px, py = px.ravel().astype(np.int64), py.ravel().astype(np.int64)
        ct = tris[cand]
        ee = _edges_at(ct[None, :], px[:, None], py[:, None])  # (64, n, 3)
        cov = (ee[..., 0] | ee[..., 1] | ee[..., 2]) > 0
        z = _depth_at(ct[None, :], ee)
        ok = cov & (z >= 0.0)
        for k in range(64):
            idx = np.nonzero(ok[k])[0]
            got = int(g.vis_tri[py[k], px[k]])
            if len(idx) == 0:
                assert got == NONE, (px[k], py[k], got)
                continue
            zk = z[k, idx]
            best = idx[zk == zk.max()].max()  # equal depths: the later submission wins
            assert got == int(cand[best]), (px[k], py[k], got, int(cand[best]))
