"""The selectable Gauss-Jordan kernels of tg_inverse against the references the
default kernels are checked against.  TG_GJ=1 forces 16-column panels at every
order (one thread per row), TG_GJ=2 16-column panels with two threads per row
(n <= 384); the variable is read once per process, so each runs in a child
process on
  * test_batched_inverse_large_orders (LAPACK inverse and determinant, A A^-1 = 1,
    singular info, orders 24 ... 512), and
  * the e_max = 6 projection against the reference's FP64 run (225 inverses of
    order 336; tolerances of test_gpu_large.py)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("gj", [1, 2])
def test_inverse_and_projection_with_forced_gj_kernel(gj):
    env = dict(os.environ, TG_GJ=str(gj))
    print("TG_GJ", gj, flush=True)
    cmd = [sys.executable, "-m", "pytest",
           os.path.join(HERE, "test_gpu_parity.py") + "::test_batched_inverse_large_orders",
           os.path.join(HERE, "test_gpu_large.py") +
           "::test_large_projection_fp64_matches_reference[emax6]",
           "-x", "-q", "-p", "no:cacheprovider"]
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(HERE), capture_output=True, text=True,
                       timeout=300)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-15:])
    assert r.returncode == 0, tail
    assert "2 passed" in r.stdout, tail
