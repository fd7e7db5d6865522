"""Per-angle observables of the final projection (tg_project_observables, the
iopt = 1 branch of project_wavefunction) against the reference's own per-angle
quantities.

The golden fixtures hold, for selected gauge angles, the reference's rot_over,
rot_prot/neut/prot2/neut2 and its transition densities and fields (rho_LR,
kappa_LR, kappa_RL, Gamma_LR, Delta_LR).  The observables of the device are
checked against the reference's routines applied to those matrices:
calculate_decompo_energy (module_hamiltonian.f90:1134-1260),
calculate_occupation_number (module_particlenumber.f90:157-176),
calculate_expectval_obo_cplx (module_operators.f90:146-164) and the traces of
calculate_expectval_obos_cplx (:227-257), with seeded random one-body operators.
Tolerance 1e-10 relative.  The whole printing path (sum_gauge, print_results) is
checked against the reference build in test_gpu_dropin.py.
"""
import numpy as np
import pytest

from conftest import load_golden, golden_canon

pytestmark = pytest.mark.gpu


def _decompo(rho, kRL, G, D, H1):
    """calculate_decompo_energy: ecomp(1..12)."""
    n = rho.shape[0]
    h = n // 2
    P, N = slice(0, h), slice(h, n)
    blk = lambda X, a, b: X[a, b]
    ph = {}
    pp = {}
    for ka, a in (("p", P), ("n", N)):
        for kb, b in (("p", P), ("n", N)):
            ph[ka + kb] = 0.5 * np.trace(blk(G, a, b) @ blk(rho, b, a))
            pp[ka + kb] = 0.5 * np.trace(blk(D, a, b) @ blk(kRL, a, b).T)
    A1, A2, A3 = H1 @ rho, G @ rho, D @ kRL.T
    e1p, e1n = np.trace(A1[P, P]), np.trace(A1[N, N])
    ephp, ephn = 0.5 * np.trace(A2[P, P]), 0.5 * np.trace(A2[N, N])
    eppp, eppn = 0.5 * np.trace(A3[P, P]), 0.5 * np.trace(A3[N, N])
    return np.array([e1p, e1n, ph["pp"], ph["pn"], ph["np"], ph["nn"], pp["pp"], pp["pn"],
                     pp["np"], pp["nn"], e1p + ephp + eppp, e1n + ephn + eppn])


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_observables_match_reference_angles(case):
    from taurus_vap_amd import gpu
    g = load_golden(case)
    n = int(g["ndim"])
    h = n // 2
    ctx = gpu.Context(n, g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"], complex(g["H0"]), g["H1"],
                      g["abcd"], g["H2"], g["trperm"])
    rng = np.random.default_rng(3)
    obo = [rng.normal(size=(n, n)) for _ in range(2)]
    O, O2 = rng.normal(size=(n, n)), rng.normal(size=(n, n))
    nshell = 3
    shell = np.array([(i % h) % nshell for i in range(n)], dtype=np.int32)
    try:
        cb = golden_canon(g)
        ctx.set_state(g["U0"], g["V0"], cb["D"].real, cb["nocc"], cb["nemp"], cb["ovac"],
                      cb["voveru"], int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                      float(g["valence_Z"]), float(g["valence_N"]), float(g["factor_delta"]))
        rec, rho_sum = ctx.project_observables(obo, [(O, O2)], shell, nshell)
    finally:
        ctx.close()
    assert rec.shape[0] == int(g["nangles"])
    H1 = np.asarray(g["H1"])
    for i, a in enumerate(g["angle_sel"]):
        q = rec[a]
        rho, kLR, kRL = g["ang_rhoLR"][i], g["ang_kappaLR"][i], g["ang_kappaRL"][i]
        G, D = g["ang_gammaLR"][i], g["ang_deltaLR"][i]
        close = lambda x, y, what: np.testing.assert_allclose(x, y, rtol=1e-10, atol=1e-10,
                                                              err_msg="%s angle %d" % (what, a))
        close(q[0], g["rot_over"][a], "rot_over")
        close(q[2:6], [g["rot_prot"][a], g["rot_neut"][a], g["rot_prot2"][a], g["rot_neut2"][a]],
              "particle numbers")
        close(q[7:19], _decompo(rho, kRL, G, D, H1), "ecomp")
        occ = np.zeros((nshell, 2), complex)
        for s in range(n):
            occ[shell[s], int(s >= h)] += rho[s, s]
        close(q[19:19 + 2 * nshell], occ.reshape(-1), "occnum")
        b0 = 19 + 2 * nshell
        for k, Ok in enumerate(obo):
            A1 = Ok @ rho
            close(q[b0 + 2 * k:b0 + 2 * k + 2], [np.trace(A1[:h, :h]), np.trace(A1[h:, h:])], "obo")
        b1 = b0 + 2 * len(obo)
        A1 = O @ rho
        t6 = np.trace(O.T @ kRL.T @ O @ kLR)
        close(q[b1:b1 + 4], [np.trace(A1), np.trace(O2 @ rho), np.trace(A1 @ A1), t6], "obos")
    assert np.all(np.isfinite(rho_sum))
