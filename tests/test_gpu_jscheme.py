"""The device path on J-scheme interactions (hamil_type 3 and 4) against the
reference's own runs.

Fixtures (tests/golden/make_golden_large.py, jscheme_*): seeded J-scheme files on
the e_max = 2 space (40 states), built into the m-scheme list by the reference
itself (hamil_read = 0: uncouple_J, src/module_hamiltonian.f90:747-814): type 3
with its .01b one-body file, 3p/4n; type 4 (kinetic one-body part) with the
centre-of-mass correction, 4p/4n; 5x5 gauge angles, two iterations.  The list
fed to the device is the one the interaction-file reader produces
(hamiltonian.read_hamiltonian), which test_producer pins element by element to
the reference's.  Tolerances: projected energy 1e-8 relative, particle numbers
1e-10, field_H20v 1e-10 absolute; iterations as the large-basis tests.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

CASES = ["jscheme_t3", "jscheme_t4_com"]


def _context(g, tmp_path):
    from taurus_vap_amd import gpu, hamiltonian as H, inputs as I
    stem = str(tmp_path / "jscheme")
    I.write_jscheme_synthetic(stem, int(g["htype"]), int(g["hamil_com"]), int(g["seed"]))
    h = H.read_hamiltonian(stem + ".sho", float(g["valence_Z"]), float(g["valence_N"]),
                           hamil_com=int(g["hamil_com"]))
    assert np.array_equal(h["abcd"], g["abcd"]) and np.array_equal(h["H2"], g["H2"])
    b = h["basis"]
    return gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"],
                       h["H2"], h["trperm"], tbme_fp32=False)


@pytest.mark.parametrize("case", CASES)
def test_jscheme_projection_matches_reference(case, tmp_path):
    from taurus_vap_amd import host
    g = load_golden(case)
    print(case, flush=True)
    ctx = _context(g, tmp_path)
    try:
        U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
        cb = host.construct_canonical_basis(U, V)
        r = ctx.project_wavefunction(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"],
                                     cb["voveru"], int(g["Mphip"]), int(g["Mphin"]),
                                     float(g["facpi"]), float(g["valence_Z"]),
                                     float(g["valence_N"]), float(g["factor_delta"]))
    finally:
        ctx.close()
    over = g["pnp_over"]
    assert abs(r["over"] - over) < 1e-10 * max(1.0, abs(over))
    e_ref = (g["pnp_ener"] / over).real
    e = (r["ener"] / r["over"]).real
    print("E_proj %.12f reference %.12f" % (e, e_ref), flush=True)
    assert abs(e - e_ref) < 1e-8 * abs(e_ref)
    for key in ("prot", "neut", "prot2", "neut2"):
        ref = g["pnp_" + key] / over
        assert abs(r[key] / r["over"] - ref) < 1e-10 * max(1.0, abs(ref)), key
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-10


@pytest.mark.parametrize("case", CASES)
def test_jscheme_iterations_match_reference(case, tmp_path):
    """tg_iterate with the reference's <Z>, <N> constraint operators and the device
    canonical basis between the iterations: energies, particle numbers, gradient
    norms and the first multipliers of the reference's history."""
    from taurus_vap_amd import host
    g = load_golden(case)
    ctx = _context(g, tmp_path)
    try:
        ctx.set_gradient(float(g["gradient_eta"]), float(g["gradient_mu"]))
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
        cb = host.construct_canonical_basis(U, V)
        ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"],
                      int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                      float(g["valence_Z"]), float(g["valence_N"]), float(g["factor_delta"]))
        for it in range(int(g["niter"])):
            print("%s iteration %d" % (case, it + 1), flush=True)
            r = ctx.iterate(0.0)
            e = (r["ener"] / r["over"]).real
            assert abs(e - g["hist_ener"][it]) <= 1e-8 * abs(g["hist_ener"][it]), (it, e)
            assert abs((r["prot"] / r["over"]).real - g["hist_prot"][it]) <= 1e-8
            assert abs((r["neut"] / r["over"]).real - g["hist_neut"][it]) <= 1e-8
            gn = g["hist_gradnorm"][it]
            assert abs(r["gradient_norm"] - gn) <= 1e-8 * gn, (it, r["gradient_norm"], gn)
            if it == 0:
                lam = np.asarray(g["lagrange_lambda1"])
                tol = np.maximum(1e-8 * np.abs(lam), 1e-10)
                assert np.all(np.abs(r["lagrange_lambda1"] - lam) <= tol)
            ctx.canonical_basis()
    finally:
        ctx.close()
