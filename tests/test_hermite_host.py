"""The boundary-cell update rebuilds only the two populations it needs
(recon2_pop / recon3_pop); they must be bit-identical to the corresponding
entries of the all-population reconstructions the bulk tiles and the oracle
follow.  The header is compiled for the host with the library's flags and
checked over random moment sets (no GPU needed)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_single_population_reconstruction_is_bit_exact(tmp_path):
    exe = tmp_path / "hermite_pop_check"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", str(exe), os.path.join(HERE, "native", "hermite_pop_check.hip")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    r = subprocess.run([str(exe), "300000"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 2D 0 3D 0" in r.stdout
