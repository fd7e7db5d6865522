"""Parity of the gfx950 library (through its C ABI) with the reference.

The expected values are the reference TAURUS_vap's own outputs (golden fixtures
written by oracle/_ref through oracle/ref_harness.f90) and, where a fixture does
not cover a case, the CPU oracle (oracle/pnvap.py) on the same seeded inputs.
Tolerances: projected energy 1e-8 relative, fields and the
gradient 1e-10 absolute.
"""
import numpy as np
import pytest

from conftest import load_golden, golden_ham, golden_canon
from oracle import pnvap as P

pytestmark = pytest.mark.gpu

CASES = ["sd64", "sd32", "pf64", "sd64_m1", "sd64_m4", "mg24_32"]


def _ctx(g, plain=False, fp32=None, conj_pairs=False):
    from taurus_vap_amd import gpu
    return gpu.Context(int(g["ndim"]), g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"],
                       complex(g["H0"]), g["H1"], g["abcd"], g["H2"], g["trperm"], plain=plain,
                       tbme_fp32=fp32, conj_pairs=conj_pairs)


def _project(ctx, g, **kw):
    return ctx.project_wavefunction(g["U0"], g["V0"], g["zD0"], int(g["nocc0"]), int(g["nemp0"]),
                                    float(g["ovac0"]), g["voveru0"], int(g["Mphip"]),
                                    int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
                                    float(g["valence_N"]), float(g["factor_delta"]), **kw)


def test_zskpfa_matches_reference():
    from taurus_vap_amd import gpu
    g = load_golden("pfaffian")
    rng = np.random.default_rng(int(g["seed"]))
    sizes = list(g["sizes"])
    for i, n in enumerate(sizes[:-1]):
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = X - X.T
        pf, info = gpu.zskpfa(M)
        print("zskpfa n=%d" % n, flush=True)
        assert info == 0
        assert abs(pf - g["pf"][i]) <= 1e-11 * abs(g["pf"][i]) * max(1, n / 16)
    pf, info = gpu.zskpfa(g["mat%d" % (len(sizes) - 1)])
    assert pf == 0 and info == 0
    # batched call, odd size, n = 0
    Ms = np.stack([g["mat1"], g["mat1"] * 2])
    pfs, _ = gpu.zskpfa(Ms)
    assert abs(pfs[1] - pfs[0] * 2 ** 2) < 1e-12 * abs(pfs[1])
    assert gpu.zskpfa(np.zeros((3, 3)))[0] == 0
    assert gpu.zskpfa(np.zeros((0, 0)))[0] == 1


@pytest.mark.parametrize("case", ["sd64", "sd32", "pf64", "mg24_32"])
def test_fields_match_reference(case):
    g = load_golden(case)
    ctx = _ctx(g)
    for k in range(len(g["angle_sel"])):
        print("fields", case, "angle", int(g["angle_sel"][k]), flush=True)
        G, h, dLR, dRL = ctx.calculate_fields(g["ang_rhoLR"][k], g["ang_kappaLR"][k],
                                              g["ang_kappaRL"][k])
        assert np.max(np.abs(G - g["ang_gammaLR"][k])) < 1e-10
        assert np.max(np.abs(h - g["ang_hspLR"][k])) < 1e-10
        assert np.max(np.abs(dLR - g["ang_deltaLR"][k])) < 1e-10
        assert np.max(np.abs(dRL - g["ang_deltaRL"][k])) < 1e-10
    ctx.close()


def test_fields_diag_matches_oracle():
    g = load_golden("pf64")
    ctx = _ctx(g)
    rho, kap = g["rhoRR"].astype(complex), g["kappaRR"].astype(complex)
    got = ctx.calculate_fields_diag(rho, kap)
    ref = P.calculate_fields_diag(rho, kap, golden_ham(g), g["sp_2mj"], g["sp_tr"] - 1)
    for x, y in zip(got, ref):
        assert np.max(np.abs(x - y)) < 1e-10
    ctx.close()


@pytest.mark.parametrize("case", ["sd64", "pf64", "mg24_32", "sd64_m4"])
def test_overlap_matches_reference(case):
    from taurus_vap_amd import gpu
    g = load_golden(case)
    cb = golden_canon(g)
    n = int(g["ndim"])
    for a in (0, 1, int(g["nangles"]) // 2, int(g["nangles"]) - 1):
        R = P.rotation_gauge(float(g["phip"][a]), float(g["phin"][a]), n)
        o = gpu.calculate_overlap(cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"],
                                  cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"], R)
        ref = g["rot_over"][a]
        assert abs(o - ref) < 1e-9 * max(1.0, abs(ref)), (a, o, ref)


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("conj_pairs", [False, True])
def test_projection_matches_reference(case, plain, conj_pairs):
    """All gauge angles, or (conj_pairs) one angle of each (phi, -phi) pair
    counted as 2 Re(.), against the reference's full-grid sums."""
    g = load_golden(case)
    print("project", case, "plain" if plain else "batched", "pairs" if conj_pairs else "", flush=True)
    ctx = _ctx(g, plain=plain, conj_pairs=conj_pairs)
    r = _project(ctx, g)
    over = g["pnp_over"]
    assert abs(r["over"] - over) < 1e-10 * max(1.0, abs(over))
    e_ref = (g["pnp_ener"] / over).real
    assert abs((r["ener"] / r["over"]).real - e_ref) < 1e-8 * abs(e_ref)
    for key in ("prot", "neut", "prot2", "neut2"):
        assert abs(r[key] / r["over"] - g["pnp_" + key] / over) < 1e-10
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-10
    for key in ("A", "H20", "AH20"):
        assert np.max(np.abs(r[key] - g["pnp_" + key])) < 1e-10 * max(1.0, abs(over))
    ctx.close()


def test_parity_projection_matches_reference():
    """iprint > 0: parity overlaps of the projected state (rot_pari, pnp_pari,
    module_projection.f90:279-284, 439) as the reference computed them for the
    Mg24 deck (iter_print = 1); and the same against the oracle for sd/pf."""
    g = load_golden("mg24_32")
    assert int(g["iter_print"]) == 1
    for pairs in (False, True):
        ctx = _ctx(g, conj_pairs=pairs)
        r = _project(ctx, g, iprint=1)
        ctx.close()
        assert abs(r["pari"] - g["pnp_pari"]) < 1e-10, (pairs, r["pari"], g["pnp_pari"])
    for case in ("sd64", "pf64"):
        g = load_golden(case)
        cb = golden_canon(g)
        n = int(g["ndim"])
        par = np.where(g["sp_l"] % 2 == 0, 1.0, -1.0)
        ref = 0j
        for a in range(int(g["nangles"])):
            R = P.rotation_gauge(float(g["phip"][a]), float(g["phin"][a]), n) * par[:, None]
            o = P.calculate_overlap(cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"],
                                    cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"], R)
            ref += o * g["weip"][a] * g["wein"][a]
        ctx = _ctx(g)
        r = _project(ctx, g, iprint=1)
        ctx.close()
        assert abs(r["pari"] - ref) < 1e-9 * max(1.0, abs(ref)), (case, r["pari"], ref)


def test_projection_fp32_storage_matches_stock_reference():
    """The stock reference stores TBMEs in real32 (rH2 = real32) and accumulates
    in real64; the FP32 device copy of the same values reproduces it."""
    g = load_golden("sd32")
    ctx = _ctx(g, fp32=True)
    r = _project(ctx, g)
    assert abs((r["ener"] / r["over"]).real - (g["pnp_ener"] / g["pnp_over"]).real) < 1e-8 * 40
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-10
    ctx.close()


def test_gradient_matches_reference():
    from taurus_vap_amd import gpu
    from test_oracle import eigen_for_gradient
    for case in ("sd64", "pf64", "mg24_32"):
        g = load_golden(case)
        n = int(g["ndim"])
        ncons = int(g["constraint_dim"])
        ehsp, eH11 = eigen_for_gradient(g)
        Zi, norm = gpu.calculate_gradient(g["field_H20v"], g["constraint_20"] if ncons else None,
                                          g["lagrange_lambda1"] if ncons else None, np.zeros(n * n),
                                          int(g["gradient_type"]), float(g["gradient_eta"]),
                                          float(g["gradient_mu"]), eigen_H11=eH11, eigen_hsp=ehsp)
        assert np.max(np.abs(Zi - g["gradient_Zi"])) < 1e-12
        assert abs(norm - float(g["gradient_norm"])) < 1e-10 * float(g["gradient_norm"])


def test_projection_own_canonical_basis():
    """Canonical basis from the host restatement instead of LAPACK's: the
    projected gradient is basis independent (to the v^2 ~ 1 conditioning)."""
    g = load_golden("pf64")
    cb = P.construct_canonical_basis(g["U0"], g["V0"])
    ctx = _ctx(g)
    r = ctx.project_wavefunction(g["U0"], g["V0"], cb["D"].real, cb["nocc"], cb["nemp"], cb["ovac"],
                                 cb["voveru"], int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                                 float(g["valence_Z"]), float(g["valence_N"]))
    assert np.max(np.abs(r["H20v"] - g["field_H20v"])) < 1e-9
    ctx.close()


def test_projection_emax4_vs_oracle():
    """No-core e_max=4 basis (140 states, 5.5e5 stored TBMEs from the native
    producer), 3x3 angles: the GPU path against the CPU oracle on the same
    seeded inputs (blocked inverse and reduced Pfaffian at a mid size)."""
    from taurus_vap_amd import hamiltonian as H, host, inputs as I
    w = I.WORKLOADS["emax4"]
    h = H.workload_hamiltonian(w, 1234)
    b = h["basis"]
    U, V = host.seed_state(w, b, 1234)
    cb = host.construct_canonical_basis(U, V)
    from taurus_vap_amd import gpu
    ctx = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"],
                      h["H2"], h["trperm"])
    r = ctx.project_wavefunction(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"],
                                 3, 3, 2 * np.pi, w.Z, w.N)
    ctx.close()
    ham = dict(H0=h["H0"], H1=h["H1"], abcd=h["abcd"].astype(np.int64) - 1, H2=h["H2"],
               trperm=h["trperm"].astype(np.int64))
    canon = dict(D=cb["D"].astype(complex), nocc=cb["nocc"], nemp=cb["nemp"], ovac=cb["ovac"],
                 voveru=cb["voveru"])
    ref = P.project_wavefunction(U, V, ham, dict(sp_2mj=b.sp_2mj, sp_tr=b.sp_tr - 1), 3, 3,
                                 2 * np.pi, w.Z, w.N, canon=canon)
    e, e_ref = (r["ener"] / r["over"]).real, (ref["ener"] / ref["over"]).real
    print("emax4 E_gpu %.12f E_oracle %.12f" % (e, e_ref), flush=True)
    assert abs(r["over"] - ref["over"]) < 1e-9 * abs(ref["over"])
    assert abs(e - e_ref) < 1e-8 * abs(e_ref)
    assert np.max(np.abs(r["H20v"] - ref["H20v"])) < 1e-9


def test_adjoint_gradient_emax4_finite_difference():
    """e_max=4 no-core basis (140 states), 11x11 gauge angles, FP64 TBMEs: the
    GPU projected gradient field_H20v (the adjoint dE_proj/dZ of
    calculate_projected_H20) against central finite differences (h = 1e-6) of
    the GPU projected energy along real skew Thouless directions
    (evolve_wavefunction, src/module_constraints.f90:497-590): one single-pair
    direction Z = e_ij - e_ji (dE/dt = 2 H20v_ij) and two dense random skew
    directions (dE/dt = sum_kl H20v_kl Z_kl).  The canonical basis of every
    displaced state is rebuilt on the host, as the reference does per
    iteration."""
    from taurus_vap_amd import gpu, hamiltonian as H, host, inputs as I
    w = I.WORKLOADS["emax4"]
    h = H.workload_hamiltonian(w, 1234)
    b = h["basis"]
    n = b.dim
    U0, V0 = host.seed_state(w, b, 1234)
    ctx = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"],
                      h["H2"], h["trperm"])
    M = w.M

    def project(U, V):
        cb = host.construct_canonical_basis(U, V)
        return ctx.project_wavefunction(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"],
                                        cb["voveru"], M, M, 2 * np.pi, w.Z, w.N)

    def energy(U, V):
        r = project(U, V)
        return (r["ener"] / r["over"]).real

    r0 = project(U0, V0)
    G = r0["H20v"].reshape((n, n), order="F")
    assert np.max(np.abs(G + G.T)) < 1e-12
    hstep = 1e-6
    rng = np.random.default_rng(5)
    dirs = []
    Z = np.zeros((n, n)); Z[3, 90], Z[90, 3] = 1.0, -1.0
    dirs.append(Z)
    for _ in range(2):
        X = rng.normal(size=(n, n))
        dirs.append(np.triu(X, 1) - np.triu(X, 1).T)
    try:
        for Z in dirs:
            Up, Vp = host.evolve(U0, V0, hstep * Z)
            Um, Vm = host.evolve(U0, V0, -hstep * Z)
            fd = (energy(Up, Vp) - energy(Um, Vm)) / (2 * hstep)
            pred = float(np.sum(G * Z))
            print("emax4 adjoint: fd %.9e  H20v.Z %.9e" % (fd, pred), flush=True)
            # rounding of E (~1e-12 relative of ~1e2-1e3 MeV) over 2h bounds the FD error
            assert abs(fd - pred) < 1e-5 * abs(pred) + 2e-3, (fd, pred)
    finally:
        ctx.close()


def test_batched_inverse_large_orders():
    """tg_inverse (the ZGETRF + ZGETRI pair of calculate_thouless) at the orders of
    the e_max = 4 / 6 inverses and up to 512 (each workgroup size of the blocked
    Gauss-Jordan kernels; above 256 the 32-column panels, whose last panel has 1,
    16, 17 or 32 columns at n = 257, 272, 273, 288): A A^-1 = 1, the LAPACK
    inverse and determinant, and singular matrices' info."""
    from taurus_vap_amd import gpu
    rng = np.random.default_rng(21)
    for n in (24, 64, 128, 129, 140, 257, 272, 273, 288, 300, 336, 384, 512):
        # entries of variance 1/n keep the determinant within the FP64 range
        A = (rng.normal(size=(3, n, n)) + 1j * rng.normal(size=(3, n, n))) / np.sqrt(2.0 * n)
        print("inverse n=%d" % n, flush=True)
        inv, det, info = gpu.inverse(A)
        assert np.all(info == 0)
        for k in range(3):
            ref = np.linalg.inv(A[k])
            err = np.max(np.abs(inv[k] - ref)) / np.max(np.abs(ref))
            assert err < 1e-11 * n, (n, k, err)
            assert np.max(np.abs(A[k] @ inv[k] - np.eye(n))) < 1e-11 * n
            sign, logdet = np.linalg.slogdet(A[k])
            assert abs(np.log(abs(det[k])) - logdet) < 1e-10 * n, (n, k)
            assert abs(det[k] / abs(det[k]) - sign) < 1e-10 * n, (n, k)
    # a zero column: info = first zero pivot column (1-based), det = 0; at
    # n = 336 in the first and in the second half of a 32-column panel
    for n, z in ((200, 57), (336, 40), (336, 57), (336, 335)):
        A = rng.normal(size=(n, n)) + 0j
        A[:, z] = 0
        _, det, info = gpu.inverse(A)
        assert info == z + 1 and det == 0, (n, z, info, det)
