"""Pins the CPU oracle (oracle/pnvap.py) before it is trusted as the checker.

* against the reference itself: the golden fixtures were written by the
  reference TAURUS_vap (oracle/_ref, built from its own sources) through
  oracle/ref_harness.f90 -- every stage of the gauge-angle loop is compared;
* against what the mathematics fixes: Pfaffian identities and brute force,
  the dense-tensor form of the contraction, overlap invariants, projection
  invariants and a finite-difference check of the projected gradient.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import load_golden, golden_ham, golden_canon
from oracle import pnvap as P

CASES = ["sd64", "sd32", "pf64", "sd64_m1", "sd64_m4", "mg24_32"]


def _basis(g):
    return dict(sp_2mj=g["sp_2mj"], sp_tr=g["sp_tr"] - 1)


def _project(g, keep=False, canon=True):
    return P.project_wavefunction(
        g["U0"], g["V0"], golden_ham(g), _basis(g), int(g["Mphip"]), int(g["Mphin"]),
        float(g["facpi"]), float(g["valence_Z"]), float(g["valence_N"]),
        canon=golden_canon(g) if canon else None, factor_delta=float(g["factor_delta"]),
        keep_angles=keep)


# ---------------------------------------------------------------------------
# Oracle vs the reference (golden fixtures)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", CASES)
def test_projection_matches_reference(case):
    g = load_golden(case)
    r = _project(g, keep=True)
    over = g["pnp_over"]
    assert abs(r["over"] - over) <= 1e-12 * max(1.0, abs(over))
    e_ref = (g["pnp_ener"] / over).real
    assert abs((r["ener"] / r["over"]).real - e_ref) <= 1e-10 * abs(e_ref)
    for key, ref in (("A", "pnp_A"), ("H20", "pnp_H20"), ("AH20", "pnp_AH20")):
        assert np.max(np.abs(r[key] - g[ref])) < 1e-12
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-12
    # every angle's scalars
    for name in ("over", "ener", "prot", "neut", "prot2", "neut2"):
        got = np.array([a[name] for a in r["angles"]])
        assert np.max(np.abs(got - g["rot_" + name])) < 1e-11, name
    # full matrices for the sampled angles
    pairs = [("A", "rot_A"), ("Ut", "Utilde"), ("Vt", "Vtilde"), ("rho", "rhoLR"),
             ("kLR", "kappaLR"), ("kRL", "kappaRL"), ("G", "gammaLR"), ("hsp", "hspLR"),
             ("DLR", "deltaLR"), ("DRL", "deltaRL"), ("H20", "rot_H20")]
    for k, ia in enumerate(g["angle_sel"]):
        a = r["angles"][int(ia)]
        for mine, ref in pairs:
            assert np.max(np.abs(a[mine] - g["ang_" + ref][k])) < 1e-12, (ia, mine)


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_own_canonical_basis(case):
    """construct_canonical_basis restated with numpy's symmetric eigensolver:
    same occupation counts; projected gradient unchanged to the conditioning of
    the v^2 ~ 1 eigenvalues (the overlap itself is basis independent)."""
    g = load_golden(case)
    cb = P.construct_canonical_basis(g["U0"], g["V0"])
    assert cb["nocc"] == int(g["nocc0"]) and cb["nemp"] == int(g["nemp0"])
    assert abs(cb["ovac"] - float(g["ovac0"])) < 1e-8
    r = _project(g, canon=False)
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-10
    assert abs(r["over"] - g["pnp_over"]) < 1e-9 * abs(g["pnp_over"])


def eigen_for_gradient(g):
    """diagonalize_hsp_and_H11(0) restated (module_gradient.f90:158-453) on the
    fixture's state; used when gradient_type > 0."""
    ncons = int(g["constraint_dim"])
    return P.diagonalize_hsp_and_H11(
        g["U0"], g["V0"], golden_ham(g), _basis(g), g["constraint_HO"] if ncons else None,
        g["lagrange_lambda1"] if ncons else [], ncons, int(g["constraint_pair"]),
        float(g["factor_delta"]), int(g["Mphip"]))


def test_gradient_step_matches_reference():
    """calculate_gradient, with the heavy-ball recipe of gradient_type 1 fed by
    the restated hsp/H11 diagonalisation for the Mg24 example."""
    for case in ("sd64", "pf64", "mg24_32"):
        g = load_golden(case)
        ncons = int(g["constraint_dim"])
        lam = g["lagrange_lambda1"] if ncons else []
        c20 = g["constraint_20"] if ncons else None
        n = int(g["ndim"])
        ehsp, eH11 = eigen_for_gradient(g)
        Zi, norm, eta, mu = P.calculate_gradient(g["field_H20v"], c20, lam, np.zeros(n * n),
                                                 int(g["gradient_type"]), float(g["gradient_eta"]),
                                                 float(g["gradient_mu"]), eigen_H11=eH11,
                                                 eigen_hsp=ehsp)
        assert np.max(np.abs(Zi - g["gradient_Zi"])) < 1e-13
        assert abs(norm - float(g["gradient_norm"])) < 1e-10 * float(g["gradient_norm"])
    assert int(load_golden("mg24_32")["gradient_type"]) == 1


# ---------------------------------------------------------------------------
# Pfaffian: reference values and identities
# ---------------------------------------------------------------------------

def test_pfaffian_matches_reference():
    g = load_golden("pfaffian")
    rng = np.random.default_rng(int(g["seed"]))
    sizes = list(g["sizes"])
    for i, n in enumerate(sizes[:-1]):
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = X - X.T
        if ("mat%d" % i) in g:
            assert np.array_equal(M, g["mat%d" % i])
        pf, info = P.zskpfa(M)
        assert abs(pf - g["pf"][i]) <= 1e-12 * abs(g["pf"][i]) * max(1, n / 16)
    pf, info = P.zskpfa(g["mat%d" % (len(sizes) - 1)])
    assert pf == 0 and g["pf"][-1] == 0 and g["info"][-1] == 0


def _pf_brute(A):
    """Pfaffian by the permutation expansion over perfect matchings."""
    n = A.shape[0]
    if n == 0:
        return 1.0
    total = 0
    for j in range(1, n):
        rest = [k for k in range(1, n) if k != j]
        sub = A[np.ix_(rest, rest)]
        total += (-1) ** (j - 1) * A[0, j] * _pf_brute(sub)
    return total


def test_pfaffian_identities():
    rng = np.random.default_rng(3)
    for n in (2, 4, 6, 8):
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = X - X.T
        pf, _ = P.zskpfa(M)
        assert abs(pf - _pf_brute(M)) < 1e-12 * max(1, abs(pf))
    for n in (10, 40, 80):
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = X - X.T
        pf, _ = P.zskpfa(M)
        det = np.linalg.det(M)
        assert abs(pf * pf - det) < 1e-10 * abs(det)
        # swapping two indices flips the sign
        p = np.arange(n); p[[1, 5]] = p[[5, 1]]
        pf2, _ = P.zskpfa(M[np.ix_(p, p)])
        assert abs(pf2 + pf) < 1e-10 * abs(pf)
    # block diagonal: product of the 2x2 Pfaffians
    x = rng.normal(size=5) + 1j * rng.normal(size=5)
    M = np.zeros((10, 10), complex)
    for i in range(5):
        M[2 * i, 2 * i + 1], M[2 * i + 1, 2 * i] = x[i], -x[i]
    assert abs(P.zskpfa(M)[0] - np.prod(x)) < 1e-13
    # odd size
    assert P.zskpfa(np.zeros((3, 3)))[0] == 0
    # a lower triangle is never read (UPLO = 'U')
    X = rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12)); M = X - X.T
    M2 = M.copy(); M2[np.tril_indices(12, -1)] = 99.0
    assert P.zskpfa(M)[0] == P.zskpfa(M2)[0]


# ---------------------------------------------------------------------------
# Overlap invariants
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", ["sd64", "pf64", "mg24_32"])
def test_overlap_invariants(case):
    g = load_golden(case)
    cb = golden_canon(g)
    n = int(g["ndim"])
    args = (cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"]) * 2
    # <Phi|Phi> = 1; the Pfaffian formula multiplies exp(ovac) (here ~e^-21)
    # by a Pfaffian of size ~e^+21, so 1 is met to ~1e-10 relative
    assert abs(P.calculate_overlap(*args, np.eye(n, dtype=complex)) - 1.0) < 1e-8
    U, V = g["U0"].astype(complex), g["V0"].astype(complex)
    for phip, phin in ((0.3, 1.1), (2.0, -0.7)):
        R = P.rotation_gauge(phip, phin, n)
        Ub, Vb = P.rotate_wavefunction(R, U, V)
        o = P.calculate_overlap(*args, R)
        assert abs(abs(o) - P.calculate_norm(U, V, Ub, Vb)) < 1e-8 * max(abs(o), 1e-3)


# ---------------------------------------------------------------------------
# Contraction: dense antisymmetrised tensor
# ---------------------------------------------------------------------------

def _dense_V(ham, sp_2mj, sp_tr, n):
    V = np.zeros((n, n, n, n))
    for a, b, c, d, h in P._tbme_passes(ham, sp_2mj, sp_tr):
        for (i, j, k, l) in ((a, b, c, d), (c, d, a, b)):
            V[i, j, k, l] = h
            V[j, i, k, l] = -h
            V[i, j, l, k] = -h
            V[j, i, l, k] = h
    return V


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_fields_dense_tensor(case):
    g = load_golden(case)
    n = int(g["ndim"])
    ham = golden_ham(g)
    V = _dense_V(ham, g["sp_2mj"], g["sp_tr"] - 1, n)
    # the stored list is hermitian and antisymmetric by construction
    assert np.allclose(V, V.transpose(2, 3, 0, 1)) and np.allclose(V, -V.transpose(1, 0, 2, 3))
    rng = np.random.default_rng(5)
    rho = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    kLR = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)); kLR = kLR - kLR.T
    kRL = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)); kRL = kRL - kRL.T
    G, hsp, DLR, DRL = P.calculate_fields(rho, kLR, kRL, ham, g["sp_2mj"], g["sp_tr"] - 1)
    assert np.max(np.abs(G - np.einsum("abcd,db->ac", V, rho))) < 1e-11
    assert np.max(np.abs(DLR - 0.5 * np.einsum("abcd,cd->ab", V, kLR))) < 1e-11
    assert np.max(np.abs(DRL - 0.5 * np.einsum("abcd,cd->ab", V, kRL))) < 1e-11
    assert np.max(np.abs(hsp - G - ham["H1"])) < 1e-14


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_fields_diag_equals_general(case):
    g = load_golden(case)
    ham = golden_ham(g)
    rho, kap = g["rhoRR"].astype(complex), g["kappaRR"].astype(complex)
    a = P.calculate_fields_diag(rho, kap, ham, g["sp_2mj"], g["sp_tr"] - 1)
    b = P.calculate_fields(rho, kap, kap, ham, g["sp_2mj"], g["sp_tr"] - 1)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) < 1e-12


# ---------------------------------------------------------------------------
# Projection invariants and the gradient
# ---------------------------------------------------------------------------

def test_projection_exact_particle_number():
    """With M = 13 > 12 (sd shell holds 0..12 of each species) the Fomenko
    sum is an exact projector: <Z>, <N> equal the targets, variances vanish."""
    g = load_golden("sd64")
    r = P.project_wavefunction(g["U0"], g["V0"], golden_ham(g), _basis(g), 13, 13,
                               float(g["facpi"]), float(g["valence_Z"]), float(g["valence_N"]),
                               canon=golden_canon(g))
    o = r["over"]
    assert abs(r["prot"] / o - 4.0) < 1e-10 and abs(r["neut"] / o - 4.0) < 1e-10
    assert abs(r["prot2"] / o - 16.0) < 1e-9 and abs(r["neut2"] / o - 16.0) < 1e-9


def test_gradient_finite_difference():
    """Central differences (h = 1e-6) of E_proj along real skew Thouless
    directions Z = t (e_ij - e_ji) (evolve_wavefunction) equal
    2 * field_H20v(i,j): field_H20v is antisymmetric and
    dE/dt = sum_kl field_H20v(k,l) dZ_kl/dt."""
    g = load_golden("sd64")
    n = int(g["ndim"])
    ham, basis = golden_ham(g), _basis(g)
    kw = dict(Mphip=7, Mphin=7, facpi=float(g["facpi"]), valZ=4.0, valN=4.0)

    def energy(U, V):
        r = P.project_wavefunction(U, V, ham, basis, **kw)
        return (r["ener"] / r["over"]).real

    r0 = P.project_wavefunction(g["U0"], g["V0"], ham, basis, **kw)
    H20 = r0["H20v"].reshape((n, n), order="F")
    assert np.max(np.abs(H20 + H20.T)) < 1e-12
    h = 1e-6
    for (i, j) in ((0, 13), (3, 20), (7, 8)):
        Z = np.zeros((n, n)); Z[i, j], Z[j, i] = 1.0, -1.0
        Up, Vp = P.evolve_wavefunction(g["U0"], g["V0"], h * Z)
        Um, Vm = P.evolve_wavefunction(g["U0"], g["V0"], -h * Z)
        fd = (energy(Up, Vp) - energy(Um, Vm)) / (2 * h)
        assert abs(fd - 2 * H20[i, j]) < 1e-6 * max(1.0, abs(fd)), (i, j, fd, H20[i, j])
