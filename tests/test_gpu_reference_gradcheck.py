"""The reference's own gradient check (tests/gradcheck.cpp of the reference: analytic
gradients from one Adam step with a huge epsilon against central differences of a
double-precision objective, on 64 gaussians at 32x32) run on the HIP backend through the
drop-in adapter: its five runs — L1, SSIM, L1 + SSIM, view-dependent colour, and the
prior-guided run (aux rasterizer, depth / sky / fill priors, learned sky, exposure) —
with the reference's own tolerance (5% relative plus 1e-3 of the largest gradient). The
program is built from the reference's source by oracle/Makefile (target gradcheck)."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
PROGRAM = os.path.join(ROOT, "oracle", "_ref", "gradcheck_hip")


@pytest.mark.skipif(not os.path.exists(PROGRAM), reason="oracle/_ref/gradcheck_hip not built (needs the reference "
                                                        "source: make -C oracle gradcheck)")
def test_reference_gradcheck_passes_on_the_hip_backend():
    r = subprocess.run([PROGRAM], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    assert r.stdout.count("gradients match") == 5
    assert "priors, sky, exposure" in r.stdout
