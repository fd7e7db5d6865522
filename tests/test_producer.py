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
The interaction-file reader (hamiltonian.read_hamiltonian, the reference's
set_hamiltonian with hamil_read = 0) against the same kind of fixtures:
  * the reference's own usdb.sho (ANTOINE, mass-scaled TBMEs: idens = 1,
    zmass = 0.3) for 24Mg: the list the reference wrote, H2 bit-identical in FP32;
  * seeded J-scheme interactions (uncouple_J, the six isospin channels and
    find_iso, src/module_hamiltonian.f90:747-814, 911-959): type 3 with its .01b
    one-body file, and type 4 (kinetic one-body part, set_kinetic_bare :692-733)
    with the centre-of-mass correction (.com, :496-525): abcd, H2 (FP64), trperm,
    H0 and H1 identical to the reference's.
"""
import os
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


def test_read_hamiltonian_antoine_usdb():
    from taurus_vap_amd import hamiltonian as H
    g = load_golden("a19_mg24_g1")
    here = os.path.dirname(os.path.abspath(__file__))
    h = H.read_hamiltonian(os.path.join(here, "golden", "usdb.sho"), float(g["valence_Z"]),
                           float(g["valence_N"]))
    assert h["type"] == 1 and h["nucleus_A"] == 24.0
    for k in ("sp_2mj", "sp_2mt", "sp_l", "sp_tr"):
        assert np.array_equal(getattr(h["basis"], k), g[k])
    assert np.array_equal(h["abcd"], g["abcd"])
    assert np.array_equal(h["trperm"], g["trperm"])
    assert np.array_equal(h["H2"].astype(np.float32), g["H2"])
    assert np.array_equal(h["H1"], g["H1"])


@pytest.mark.parametrize("case", ["jscheme_t3", "jscheme_t4_com"])
def test_read_hamiltonian_jscheme_equals_reference(case, tmp_path):
    from taurus_vap_amd import hamiltonian as H, inputs as I
    g = load_golden(case)
    htype, com = int(g["htype"]), int(g["hamil_com"])
    stem = str(tmp_path / "jscheme")
    I.write_jscheme_synthetic(stem, htype, com, int(g["seed"]))
    h = H.read_hamiltonian(stem + ".sho", float(g["valence_Z"]), float(g["valence_N"]),
                           hamil_com=com)
    assert h["type"] == htype
    for k in ("sp_2mj", "sp_2mt", "sp_l", "sp_tr"):
        assert np.array_equal(getattr(h["basis"], k), g[k])
    assert h["abcd"].shape == g["abcd"].shape == (int(g["ntb"]), 4)
    assert np.array_equal(h["abcd"], g["abcd"])
    assert np.array_equal(h["trperm"], g["trperm"])
    assert np.array_equal(h["H2"], g["H2"])
    assert _sha(h["abcd"], h["H2"], "<f8", h["trperm"]) == str(g["red_sha256_64"])
    assert h["H0"] == complex(g["H0"])
    assert np.array_equal(h["H1"], g["H1"])


def test_jscheme_reconstruction_images():
    """After the symmetry reconstruction every stored J-scheme element equals each
    of its seven images times the exchange phases, in the channel the permutation
    leads to: a channel map written out by hand from which particles are protons
    (0 pp pp, 1 pn pn, 2 pn np, 3 np pn, 4 np np, 5 nn nn)."""
    from taurus_vap_amd import hamiltonian as H
    rng = np.random.default_rng(3)
    sh_2j = [1, 3, 1]
    nsh, jmax = 3, 3
    V = np.zeros((6, jmax + 1, nsh, nsh, nsh, nsh), order="F")
    # representatives a<=b, c<=d, (a,b)<=(c,d) only, so no two are images of each other
    for a in range(nsh):
        for b in range(a, nsh):
            for c in range(a, nsh):
                for d in range(c, nsh):
                    if (a, b) > (c, d) or (a, b) == (c, d) or a == b or c == d:
                        continue
                    V[:, :, a, b, c, d] = rng.normal(size=(6, jmax + 1))
    H.reconstruct(V, 3, sh_2j)
    # channel after swapping the bra pair, the ket pair, or bra with ket
    swap_bra = {0: 0, 1: 3, 2: 4, 3: 1, 4: 2, 5: 5}
    swap_ket = {0: 0, 1: 2, 2: 1, 3: 4, 4: 3, 5: 5}
    bra_ket = {0: 0, 1: 1, 2: 3, 3: 2, 4: 4, 5: 5}
    t, J, a, b, c, d = np.nonzero(np.abs(V) > 1e-12)
    assert t.size > 0
    for t_, J_, a_, b_, c_, d_ in zip(t, J, a, b, c, d):
        v = V[t_, J_, a_, b_, c_, d_]
        pab = (-1.0) ** ((sh_2j[a_] + sh_2j[b_]) // 2 + J_ + 1)
        pcd = (-1.0) ** ((sh_2j[c_] + sh_2j[d_]) // 2 + J_ + 1)
        assert V[swap_ket[t_], J_, a_, b_, d_, c_] == v * pcd
        assert V[swap_bra[t_], J_, b_, a_, c_, d_] == v * pab
        assert V[bra_ket[t_], J_, c_, d_, a_, b_] == v
    # idempotent
    W = V.copy(order="F")
    H.reconstruct(W, 3, sh_2j)
    assert np.array_equal(W, V)
