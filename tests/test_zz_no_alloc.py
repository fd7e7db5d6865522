"""tg_zskpfa_noalloc -- the allocation-free variant of tg_zskpfa, through the
C ABI: the same Pfaffians (same kernels on the same inputs: identical values)
on two calls in a row, the second with a smaller input, with the scratch array
left dirty by the first call; a call above the scratch limit is refused with
TG_E_ARG and no device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _skew(rng, count, n):
    X = rng.normal(size=(count, n, n)) + 1j * rng.normal(size=(count, n, n))
    return X - np.transpose(X, (0, 2, 1))


def test_noalloc_matches_zskpfa_two_calls():
    from taurus_vap_amd import gpu
    rng = np.random.default_rng(11)
    for count, n in ((12, 336), (3, 40)):
        print("noalloc: %d matrices of order %d" % (count, n), flush=True)
        A = _skew(rng, count, n)
        pf_ref, info_ref = gpu.zskpfa(A)
        pf, info = gpu.zskpfa_noalloc(A)
        assert np.array_equal(info, info_ref)
        assert np.array_equal(pf, pf_ref), np.max(np.abs(pf - pf_ref) / np.abs(pf_ref))
    # odd order and order 0 are answered on the host
    assert gpu.zskpfa_noalloc(np.zeros((3, 3)))[0] == 0
    assert gpu.zskpfa_noalloc(np.zeros((0, 0)))[0] == 1


def test_noalloc_refuses_oversize():
    from taurus_vap_amd import gpu
    n = 512
    count = (64 << 20) // (2 * 16 * n * n) + 1     # just above the scratch limit
    A = np.zeros((count, n, n), np.complex128)
    with pytest.raises(gpu.TaurusGPUError, match="scratch"):
        gpu.zskpfa_noalloc(A)
