"""diagonalize_hsp_and_H11(0) on the device (src/module_gradient.f90:158-453)
and the gradient recipes it feeds (calculate_gradient, :79-143, gradient_type 1
and 2), against the reference's own runs.

Fixtures (tests/golden/make_golden_large.py, the harness's 'proj' mode): for the
Mg24 example (USDB, Q21 constraint, 5x5 angles, recipes 1 and 2), the pf deck
(Q20, 9x9, recipes 1 and 2) and the sd deck with one gauge angle (recipe 1, the
M = 1 branch of :177-179), the matrices the reference hands to DSYEV (h and H11
with the constraint terms removed), their eigenvalues, the gradient of the
first iteration and three iterations of history.
Tolerances: matrices and eigenvalues 1e-10 x max(1, |spectrum|), energies 1e-8
relative, gradient norms 1e-8 relative.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

A19 = ["a19_mg24_g1", "a19_mg24_g2", "a19_pf_g1", "a19_pf_g2", "a19_sd_m1_g1"]


def _ctx(g):
    from taurus_vap_amd import gpu
    return gpu.Context(int(g["ndim"]), g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"],
                       complex(g["H0"]), g["H1"], g["abcd"], g["H2"], g["trperm"])


def _set(ctx, g, U, V):
    from taurus_vap_amd import host
    cb = host.construct_canonical_basis(U, V)
    ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"],
                  int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
                  float(g["valence_N"]), float(g["factor_delta"]))


def test_sym_eigenvalues_random_and_reference_matrices():
    """The device solver (Householder + bisection) against LAPACK on random
    symmetric matrices across the LDS / global-memory boundary, and against the
    reference's own DSYEV eigenvalues of its h and H11 matrices."""
    from taurus_vap_amd import gpu
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 24, 64, 111, 112, 113, 140, 336):
        X = rng.normal(size=(3, n, n))
        A = X + np.transpose(X, (0, 2, 1))
        w = gpu.sym_eigenvalues(A)
        print("sym_eigenvalues n=%d" % n, flush=True)
        for k in range(3):
            ref = np.linalg.eigvalsh(A[k])
            scale = max(1.0, np.max(np.abs(ref)))
            assert np.max(np.abs(w[k] - ref)) < 1e-12 * n * scale, (n, k)
    # degenerate spectrum (multiplicities of a Kramers-like pairing)
    Q, _ = np.linalg.qr(rng.normal(size=(40, 40)))
    A = Q @ np.diag(np.repeat(np.arange(20.0), 2)) @ Q.T
    assert np.max(np.abs(gpu.sym_eigenvalues(A) - np.repeat(np.arange(20.0), 2))) < 1e-12 * 40 * 20
    for case in A19:
        g = load_golden(case)
        for mat, ev in (("hsp_RR", "eigen_hsp"), ("H11", "eigen_H11")):
            w = gpu.sym_eigenvalues(np.asarray(g[mat]))
            ref = np.asarray(g[ev])
            assert np.max(np.abs(w - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), (case, mat)


@pytest.mark.parametrize("case", A19)
def test_diagonalize_hsp_and_H11_matches_reference(case):
    g = load_golden(case)
    ctx = _ctx(g)
    try:
        _set(ctx, g, np.asarray(g["U0"]), np.asarray(g["V0"]))
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        eh, e11, hm, h11 = ctx.diagonalize_hsp_and_H11(np.asarray(g["lagrange_lambda1"]))
        print(case, "max|dh| %.2e max|dH11| %.2e" % (np.max(np.abs(hm - g["hsp_RR"])),
                                                     np.max(np.abs(h11 - g["H11"]))), flush=True)
        for got, key in ((hm, "hsp_RR"), (h11, "H11"), (eh, "eigen_hsp"), (e11, "eigen_H11")):
            ref = np.asarray(g[key])
            assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), key
    finally:
        ctx.close()


@pytest.mark.parametrize("case", A19)
def test_recipe_iterations_match_reference(case):
    """tg_iterate with gradient_type 1 / 2: every iteration diagonalises h and
    H11 on the device and takes the recipe's eta, mu; the reference's first
    three iterations (energies, particle numbers, gradient norms: the later
    iterations start from states evolved with the recipe's eta, mu) and its
    first spectra are reproduced.  (The device's Z after tg_iterate includes the
    constraint corrections of adjust_constraints, so it is not the dumped
    gradient_Zi of calculate_gradient.)"""
    g = load_golden(case)
    gtype = int(g["gradient_type"])
    assert gtype in (1, 2)
    ctx = _ctx(g)
    try:
        ctx.set_gradient(float(g["gradient_eta"]), float(g["gradient_mu"]))
        ctx.set_gradient_type(gtype)
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
        for it in range(int(g["niter"])):
            print("%s iteration %d" % (case, it + 1), flush=True)
            _set(ctx, g, U, V)
            r = ctx.iterate(0.0)
            e = (r["ener"] / r["over"]).real
            assert abs(e - g["hist_ener"][it]) <= 1e-8 * abs(g["hist_ener"][it]), (it, e)
            assert abs((r["prot"] / r["over"]).real - g["hist_prot"][it]) <= 1e-8
            assert abs((r["neut"] / r["over"]).real - g["hist_neut"][it]) <= 1e-8
            gn = g["hist_gradnorm"][it]
            assert abs(r["gradient_norm"] - gn) <= 1e-8 * gn, (it, r["gradient_norm"], gn)
            if it == 0:
                eh, e11, eta, mu = ctx.hsp_h11_eigenvalues()
                for got, key in ((eh, "eigen_hsp"), (e11, "eigen_H11")):
                    ref = np.asarray(g[key])
                    assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), key
            U, V = ctx.get_state()
    finally:
        ctx.close()
