"""Every matrix-core tile variant of the batched complex GEMMs and of the grouped
contraction on the reference's projections.  The autotuner picks a variant per
GEMM signature by timing, so any of them can serve a run; TG_GEMM_VARIANT=k
forces tile k for every zz / dz / zd GEMM of a process and TG_GROUPED_VARIANT=k
tile k of the contraction (both read once per process), so each variant runs in
a child process on the projection parity tests of test_gpu_parity.py (all angles,
the reference's outputs, tolerances of that file: energy 1e-8 relative, H20 and
the projected matrices 1e-10)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = [("TG_GEMM_VARIANT", v) for v in range(11)] + \
    [("TG_GROUPED_VARIANT", v) for v in range(5)]


@pytest.mark.parametrize("var,variant", VARIANTS)
def test_projection_with_forced_tile(var, variant):
    env = dict(os.environ, **{var: str(variant)})
    print(var, variant, flush=True)
    cmd = [sys.executable, "-m", "pytest", os.path.join(HERE, "test_gpu_parity.py"), "-x", "-q",
           "-p", "no:cacheprovider", "-k",
           "(test_projection_matches_reference and not True) or test_gradient_matches_reference"]
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(HERE), capture_output=True, text=True,
                       timeout=240)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-15:])
    assert r.returncode == 0, tail
    assert " passed" in r.stdout and "7 passed" in r.stdout, tail
