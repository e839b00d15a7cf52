"""The .octrn asset reader against the reference's own asset files, and load_texture's sRGB
pre-pass (main.c:546-558) in the oracle against the sRGB transfer function (CPU)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from malevich_amd import assets as A

REF_ASSETS = Path("/root/reference/assets")
# vertex (= index) counts of the shipped meshes; every shipped mesh is de-indexed
MESH_COUNTS = {"ftm_piedras": 9480, "ftm_madera": 13368, "ftm_leaves": 39048, "ftm_dec": 52392, "ftm_roof": 22440,
               "ftm_ground": 1896, "ftm_sky": 5184, "toon_house": 22056, "toon_sky": 5184}


@pytest.mark.skipif(not REF_ASSETS.exists(), reason="the reference's asset folder is not on this machine")
def test_reference_meshes():
    total_ftm = 0
    for name, n in MESH_COUNTS.items():
        m = A.read_octrn(REF_ASSETS / f"{name}_mesh.octrn")
        assert isinstance(m, A.OctrnMesh)
        assert m.vertices.shape == (n, 8) and m.indices.shape == (n,), name
        assert m.indices.size % 8 == 0  # input_assembler_stage asserts this (main.c:668)
        np.testing.assert_array_equal(m.indices, np.arange(n, dtype=np.uint32))
        assert np.all(np.isfinite(m.vertices))
        np.testing.assert_allclose(np.linalg.norm(m.vertices[:, 3:6], axis=1), 1.0, atol=1e-3)
        if name.startswith("ftm"):
            total_ftm += n
    assert total_ftm == 143808
    scene = A.load_scene(REF_ASSETS, A.FTM_MESHES, tex_size=16)
    assert [m.vertices.shape[0] for m, _ in scene] == [MESH_COUNTS[n] for n in A.FTM_MESHES]


@pytest.mark.skipif(not REF_ASSETS.exists(), reason="the reference's asset folder is not on this machine")
def test_reference_irradiance_map():
    im = A.read_octrn(REF_ASSETS / "ninomaru_teien_panorama_irradiance.octrn")
    assert isinstance(im, A.OctrnImage)
    assert (im.format, im.width, im.height, im.mip_levels) == (A.FORMAT_RGBA32F, 256, 128, 1)
    rgb = im.texels[..., :3]
    assert 0.05 < rgb.min() < 0.1 and 2.0 < rgb.max() < 2.2
    np.testing.assert_array_equal(im.texels[..., 3], 1.0)


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    v = rng.normal(size=(40, 8)).astype(np.float32)
    i = rng.integers(0, 40, 48).astype(np.uint32)
    A.write_octrn_mesh(tmp_path / "m.octrn", v, i)
    m = A.read_octrn(tmp_path / "m.octrn")
    np.testing.assert_array_equal(m.vertices, v)
    np.testing.assert_array_equal(m.indices, i)
    t = rng.integers(0, 2 ** 32, (5, 7), dtype=np.uint64).astype(np.uint32)
    A.write_octrn_image(tmp_path / "i.octrn", t, A.FORMAT_RGBA8_SRGB)
    im = A.read_octrn(tmp_path / "i.octrn")
    assert (im.format, im.width, im.height) == (A.FORMAT_RGBA8_SRGB, 7, 5)
    np.testing.assert_array_equal(im.texels, t)
    with pytest.raises(ValueError):
        (tmp_path / "bad.octrn").write_bytes(b"notoctarine.....")
        A.read_octrn(tmp_path / "bad.octrn")


def _srgb_eotf(c: np.ndarray) -> np.ndarray:
    """IEC 61966-2-1 sRGB -> linear, in float64."""
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_srgb_pre_pass_against_the_transfer_function(oracle):
    b = np.arange(256, dtype=np.uint32)
    texels = b | (b << 8) | (b << 16) | (b << 24)
    out = oracle.srgb_to_linear_rgba8(texels)
    chans = [(out >> (8 * k)) & 255 for k in range(4)]
    for c in chans[1:]:
        np.testing.assert_array_equal(c, chans[0])  # every channel, alpha included, maps alike
    got = chans[0].astype(np.int64)
    want = np.floor(255.0 * _srgb_eotf(b / 255.0)).astype(np.int64)  # truncating re-encode
    assert got[0] == 0 and got[255] == 255
    assert np.all(np.diff(got) >= 0)
    assert np.max(np.abs(got - want)) <= 1
    assert (got == want).sum() >= 250  # off by one only where f32 rounding meets an integer
    assert got[128] == 55  # 0.2158605 * 255 = 55.04
