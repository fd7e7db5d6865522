"""The native TBME producer (taurus_vap_amd/csrc/hamil_producer.cpp: J-T
uncoupling into the m-scheme list, the storage order of the reference and the
time-reversal codes) against the reduced files the reference wrote itself.

Every fixture below comes from a reference run that built its Hamiltonian from
the .sho file (hamil_read = 0: set_hamiltonian_2body, uncouple_JT,
reconstruct_2body_timerev and write_hamiltonian_reduced, src/module_hamiltonian.f90:
247-401, 828-885, 970-1018, 608-627) on the same seeded synthetic interaction:
  * sd (24 states), pf (40) and sdpf (64): the whole list, element by element:
    abcd and trperm identical, H2 bit-identical in FP64 (rH2 = real64) and, for
    sd, in FP32 (the stock rH2 = real32);
  * e_max = 4 (140 states, 5.5e5 elements) and e_max = 6 (336 states, 1.3e7
    elements): the SHA-256 of the element payload (abcd, H2, trperm in file order),
    FP64 and FP32.
"""
import hashlib

import numpy as np
import pytest

from conftest import load_golden


def _produce(wname, seed=1234):
    from taurus_vap_amd import hamiltonian as H, inputs as I
    return H.workload_hamiltonian(I.WORKLOADS[wname], seed)


def _sha(abcd, H2, dt, trperm):
    d = hashlib.sha256()
    # 0-based indices (the file stores them 1-based)
    d.update(np.ascontiguousarray(np.asarray(abcd) - 1, dtype="<i2").tobytes())
    d.update(np.ascontiguousarray(H2, dtype=dt).tobytes())
    d.update(np.ascontiguousarray(trperm, dtype="<i1").tobytes())
    return d.hexdigest()


@pytest.mark.parametrize("case,wname", [("sd64", "sd"), ("sd32", "sd"), ("pf64", "pf"),
                                        ("batch_sdpf", "sdpf")])
def test_producer_list_equals_reference_reduced_file(case, wname):
    g = load_golden(case)
    h = _produce(wname)
    assert h["abcd"].shape == g["abcd"].shape
    assert np.array_equal(h["abcd"], g["abcd"])
    assert np.array_equal(h["trperm"], g["trperm"])
    if g["H2"].dtype == np.float32:
        assert np.array_equal(h["H2"].astype(np.float32), g["H2"])
    else:
        assert np.array_equal(h["H2"], g["H2"])


@pytest.mark.parametrize("case", ["emax4_11", "emax6"])
def test_producer_payload_hash_equals_reference(case):
    g = load_golden(case)
    h = _produce(str(g["workload"]), int(g["seed"]))
    assert len(h["trperm"]) == int(g["ntb"])
    for bits, dt in ((64, "<f8"), (32, "<f4")):
        key = "red_sha256_%d" % bits
        if key in g:
            H2 = h["H2"] if bits == 64 else h["H2"].astype(np.float32)
            assert _sha(h["abcd"], H2, dt, h["trperm"]) == str(g[key]), bits
