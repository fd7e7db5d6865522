"""The device symmetric eigen-solver (tg_sym_eigensystem: Householder
tridiagonalization, divide and conquer, back-transformation) that gives the
canonical basis its eigenvectors, in place of DSYEV('V','U')
(construct_canonical_basis, src/module_wavefunctions.f90:806).

Checked against what the mathematics fixes, for matrices whose eigen-structure
stresses the merges: random symmetric matrices of orders across the panel and
merge boundaries, spectra with exact and near multiplicities (identity, zero,
repeated diagonal, Wilkinson's W+, a block-diagonal glue of equal blocks),
rank-deficient V V^T, graded entries, and the density V V^T of the e_max = 6 seed
state of the reference (tests/golden/emax6.npz):
  * eigenvalues ascending and equal to LAPACK's (numpy.linalg.eigvalsh) to
    1e-13 ||A||;
  * residual ||A Z - Z diag(w)|| <= 1e-12 ||A|| and ||Z^T Z - I|| <= 1e-13 n;
  * only the upper triangle is read (the lower one holds garbage in one case).
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _check(A, w, Z, name):
    n = A.shape[0]
    S = np.triu(A) + np.triu(A, 1).T
    nrm = max(1.0, np.max(np.abs(S)))
    wr = np.linalg.eigvalsh(S)
    assert np.all(np.diff(w) >= 0), name
    dw = np.max(np.abs(w - wr)) / nrm
    res = np.max(np.abs(S @ Z - Z * w)) / nrm
    orth = np.max(np.abs(Z.T @ Z - np.eye(n)))
    print("%-14s n=%4d |dw| %.2e residual %.2e orth %.2e" % (name, n, dw, res, orth), flush=True)
    assert dw <= 1e-13, name
    assert res <= 1e-12, name
    assert orth <= 1e-13 * max(n, 10), name


def _sym(rng, n):
    X = rng.normal(size=(n, n))
    return X + X.T


@pytest.mark.parametrize("n", [1, 2, 3, 5, 31, 32, 33, 34, 64, 65, 66, 100, 129, 336, 500, 700])
def test_random_symmetric(n):
    from taurus_vap_amd import gpu
    A = _sym(np.random.default_rng(n), n)
    w, Z = gpu.sym_eigensystem(A)
    _check(A, w, Z, "random")


def test_batch_and_upper_triangle():
    from taurus_vap_amd import gpu
    rng = np.random.default_rng(7)
    n = 97
    As = np.stack([_sym(rng, n) for _ in range(5)])
    garbage = As.copy()
    for k in range(5):
        garbage[k][np.tril_indices(n, -1)] = rng.normal(size=n * (n - 1) // 2) * 1e3
    w, Z = gpu.sym_eigensystem(garbage)
    for k in range(5):
        _check(As[k], w[k], Z[k], "batch %d" % k)


def _cases():
    rng = np.random.default_rng(11)
    n = 200
    out = {
        "identity": np.eye(n),
        "zero": np.zeros((n, n)),
        "repeated": np.diag(np.repeat(np.arange(10.0), 20)),
        "wilkinson": (np.diag(np.abs(np.arange(-100, 101)).astype(float)) +
                      np.diag(np.ones(200), 1) + np.diag(np.ones(200), -1)),
    }
    # glue of 10 equal blocks by weak couplings: eigenvalues in tight clusters
    B = _sym(rng, 20)
    G = np.kron(np.eye(10), B)
    for k in range(1, 10):
        G[20 * k - 1, 20 * k] = G[20 * k, 20 * k - 1] = 1e-9
    out["glued"] = G
    # rank-deficient V V^T, and graded entries
    V = rng.normal(size=(n, 40))
    out["rank40"] = V @ V.T
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    out["graded"] = (Q * np.logspace(0, -12, n)) @ Q.T
    # a random orthogonal similarity of a spectrum with multiplicities
    out["multiple"] = (Q * np.repeat([-1.0, 0.0, 1e-8, 1.0, 2.0], 40)) @ Q.T
    return out


@pytest.mark.parametrize("case", ["identity", "zero", "repeated", "wilkinson", "glued", "rank40",
                                  "graded", "multiple"])
def test_structured_spectra(case):
    from taurus_vap_amd import gpu
    A = _cases()[case]
    w, Z = gpu.sym_eigensystem(A)
    _check(A, w, Z, case)


def test_density_of_emax6_state():
    from taurus_vap_amd import gpu
    g = load_golden("emax6")
    V = np.asarray(g["V0"])
    rho = V @ V.T
    w, Z = gpu.sym_eigensystem(rho)
    _check(rho, w, Z, "emax6 rho")
