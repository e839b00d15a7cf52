"""The order-preserving depth code of the tile kernel's triangle-parallel merge
(malevich_amd/csrc/mlv_kernels.hip, depth_key / depth_of_key): per-pixel maxima of 64-bit keys
whose high word codes the depth must pick the same fragment as the reference's float compares
(main.c:1166, GREATER_EQUAL; LESS_EQUAL mirrored). Checked here on the code's definition over
every float of a dense sample: the unsigned order of the codes equals the float order, +0.0 and
-0.0 share a code (they compare equal), and the code with the -0.0 flag restores every depth's
bits. The kernel itself is checked bit-exactly against the oracle by the GPU raster tests."""
from __future__ import annotations

import numpy as np


def depth_key(z: np.ndarray, ge: bool) -> np.ndarray:
    u = z.view(np.uint32).copy()
    u[(u << np.uint32(1)) == 0] = 0
    sign = (u >> np.uint32(31)).astype(bool)
    o = np.where(sign, ~u, u ^ np.uint32(0x80000000)).astype(np.uint32)
    return o if ge else ~o


def depth_of_key(hi: np.ndarray, neg_zero: np.ndarray, ge: bool) -> np.ndarray:
    o = hi if ge else ~hi
    u = np.where((o & np.uint32(0x80000000)) != 0, o ^ np.uint32(0x80000000), ~o).astype(np.uint32)
    u[neg_zero] = np.uint32(0x80000000)
    return u.view(np.float32)


def _sample(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    special = np.array([0x00000000, 0x80000000, 0x00000001, 0x80000001, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F800000,
                        0xFF800000, 0x3F800000, 0xBF800000], np.uint32)
    z = np.concatenate([bits, special]).view(np.float32)
    return z[~np.isnan(z)]  # a NaN depth never takes part in the merge


def test_code_order_equals_float_order():
    z = _sample(2_000_000, 1)
    a, b = z[: len(z) // 2], z[len(z) // 2: 2 * (len(z) // 2)]
    for ge in (True, False):
        ka, kb = depth_key(a, ge), depth_key(b, ge)
        better = (a > b) if ge else (a < b)
        np.testing.assert_array_equal(ka > kb, better)
        np.testing.assert_array_equal(ka == kb, a == b)


def test_code_restores_every_depth():
    z = _sample(4_000_000, 2)
    neg_zero = z.view(np.uint32) == np.uint32(0x80000000)
    for ge in (True, False):
        back = depth_of_key(depth_key(z, ge), neg_zero, ge)
        np.testing.assert_array_equal(back.view(np.uint32), z.view(np.uint32))


def test_signed_zeros_share_a_code():
    z = np.array([0.0, -0.0], np.float32)
    for ge in (True, False):
        k = depth_key(z, ge)
        assert k[0] == k[1]
