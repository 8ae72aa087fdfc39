"""The reference's own splat test suite (tests/test_splat.cpp of the reference), compiled
unchanged against the drop-in C++ API (host/include, libu3d_splat.so over libu3d_hip.so) by
`make -C oracle dropin`, run on the GPU: gradient check against finite differences,
closed-form projection, D-SSIM gradient, training recovery, PLY round trip, crop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "u3d_tests_dropin")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in suite not built (needs the reference sources)")]


@pytest.mark.parametrize("name", ["splat_gradients", "splat_projection", "ssim_loss",
                                  "training_recovers", "ply_round_trip", "crop"])
def test_reference_suite_on_gpu(name, tmp_path):
    r = subprocess.run([BIN, name], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
