"""Writes tests/golden/ftm_scene.npz: the reference's FTM scene geometry (assets/ftm_*_mesh.octrn,
the meshes init() loads for SceneType_FTM, main.c:1317-1361), for GPU parity tests on the box where
the reference's files are absent.

The shipped meshes repeat every vertex per triangle; identical vertex rows are merged and the
indices remapped (each triangle keeps its three vertices' exact bits), which halves the file. Run
from the repository root: python tests/golden/make_ftm_fixture.py [reference_root]
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from malevich_amd import assets as A  # noqa: E402


def merged(m: A.OctrnMesh) -> tuple[np.ndarray, np.ndarray]:
    v = np.ascontiguousarray(m.vertices, dtype=np.float32)
    rows, inverse = np.unique(v.view(np.dtype((np.void, 32))), return_inverse=True)
    verts = rows.view(np.float32).reshape(-1, 8)
    idx = inverse.reshape(-1).astype(np.uint32)[m.indices]
    assert np.array_equal(verts[idx].view(np.uint32), v[m.indices].view(np.uint32))
    return verts, idx


def main(root: str = "/root/reference") -> None:
    out = {}
    for name in A.FTM_MESHES:
        verts, idx = merged(A.read_octrn(Path(root) / "assets" / f"{name}_mesh.octrn"))
        out[f"{name}_vertices"], out[f"{name}_indices"] = verts, idx
    dst = Path(__file__).with_name("ftm_scene.npz")
    np.savez_compressed(dst, names=np.array(A.FTM_MESHES), **out)
    print(f"{dst}: {sum(v.size for k, v in out.items() if k.endswith('_indices')) // 3} triangles")


if __name__ == "__main__":
    main(*sys.argv[1:])
