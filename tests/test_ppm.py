"""PPM files written for presented frames read back unchanged (CPU)."""
import numpy as np

from malevich_amd.ppm import read_ppm, write_ppm


def test_ppm_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    p = tmp_path / "f.ppm"
    write_ppm(p, img)
    assert p.read_bytes().startswith(b"P6\n53 37\n255\n")
    np.testing.assert_array_equal(read_ppm(p), img)


def test_ppm_header_with_comment(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3)
    p = tmp_path / "c.ppm"
    p.write_bytes(b"P6\n# a comment\n3 2\n255\n" + img.tobytes())
    np.testing.assert_array_equal(read_ppm(p), img)
