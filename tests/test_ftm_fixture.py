"""The committed FTM geometry fixture (tests/golden/ftm_scene.npz) against the reference's asset
files, where they are present: every triangle's three vertices are the shipped ones, bit for bit,
in the shipped order (main.c:1317-1361 draws the meshes in this order)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from malevich_amd import assets as A

FIXTURE = Path(__file__).resolve().parent / "golden" / "ftm_scene.npz"
ASSETS = Path("/root/reference/assets")


def test_fixture_shape():
    z = np.load(FIXTURE)
    names = [str(n) for n in z["names"]]
    assert names == list(A.FTM_MESHES)
    tris = 0
    for n in names:
        v, i = z[f"{n}_vertices"], z[f"{n}_indices"]
        assert v.dtype == np.float32 and v.shape[1] == 8 and i.dtype == np.uint32
        assert i.size % 3 == 0 and int(i.max()) < v.shape[0]
        tris += i.size // 3
    assert tris == 47936


@pytest.mark.skipif(not ASSETS.is_dir(), reason="the reference's asset files are not on this machine")
def test_fixture_equals_the_shipped_meshes():
    z = np.load(FIXTURE)
    for n in A.FTM_MESHES:
        m = A.read_octrn(ASSETS / f"{n}_mesh.octrn")
        v, i = z[f"{n}_vertices"], z[f"{n}_indices"]
        np.testing.assert_array_equal(v[i].view(np.uint32), np.ascontiguousarray(m.vertices)[m.indices].view(np.uint32))
