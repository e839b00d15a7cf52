"""The C host program (malevich_amd/host/mlv_render.c) drives the library through the C ABI
alone; its frame equals the oracle's render of the same Suprematism scene."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from malevich_amd import build as B
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, render_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H", [(1200, 720), (328, 184)])
def test_c_host_program_matches_oracle(oracle, tmp_path, W, H):
    exe = B.HOST_PROGRAM
    if not exe.exists():
        pytest.fail(f"{exe} is not built; run __graft_entry__.build()")
    out = tmp_path / "frame.raw"
    print(f"case: C host program {W}x{H}", flush=True)
    proc = subprocess.run([str(exe), str(out), str(W), str(H)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "8 assembled triangles" in proc.stdout, proc.stdout
    got = np.fromfile(out, np.uint32).reshape(H, W)
    d = DrawSpec(S.SUPREMATISM_INDICES, S.SUPREMATISM_VERTICES, "passthrough", "passthrough")
    o = render_oracle(oracle, W, H, [d], S.CLEAR_COLOR)
    np.testing.assert_array_equal(got, o.colors)
