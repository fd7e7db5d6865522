"""The headline sizes against the reference's own runs.

Fixtures (tests/golden/make_golden_large.py): the reference TAURUS_vap built
from its own sources, with its own Hamiltonian built from the .sho file, run
on the seeded synthetic decks of
  * e_max = 6: 336 states, 15x15 gauge angles, 10p/12n, constraints <Z>, <N>,
    Q20 = 0.2 and Jx = 0, FP64 TBMEs (rH2 = real64) and the stock FP32 build;
  * e_max = 4: 140 states, 11x11 gauge angles, 8p/8n, FP64 TBMEs.
Each fixture holds the state of the reference's first iteration (after its
own adjust_constraints), its gauge sums, field_H20v, multipliers, gradient and
two iterations of history, and the SHA-256 of the element payload of the
reduced Hamiltonian file the reference wrote (abcd, H2, trperm): the device
runs on the native producer's list, which the hash pins to the reference's.
Tolerances: projected energy 1e-8 relative, field_H20v and the gradient 1e-10
absolute, the FP32-storage energy within 1 keV of the FP64 reference.
"""
import hashlib

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

_HAM = {}


def _hamiltonian(g):
    from taurus_vap_amd import hamiltonian as H, inputs as I
    name = str(g["workload"])
    if name not in _HAM:
        _HAM.clear()
        _HAM[name] = H.workload_hamiltonian(I.WORKLOADS[name], int(g["seed"]))
    return _HAM[name]


def payload_sha256(abcd, H2, h2_dtype, trperm):
    d = hashlib.sha256()
    # 0-based indices (the file stores them 1-based)
    d.update(np.ascontiguousarray(np.asarray(abcd) - 1, dtype="<i2").tobytes())
    d.update(np.ascontiguousarray(H2, dtype=h2_dtype).tobytes())
    d.update(np.ascontiguousarray(trperm, dtype="<i1").tobytes())
    return d.hexdigest()


def _context(g, fp32, **kw):
    from taurus_vap_amd import gpu
    h = _hamiltonian(g)
    b = h["basis"]
    assert len(h["trperm"]) == int(g["ntb"])
    key = "red_sha256_32" if fp32 else "red_sha256_64"
    got = payload_sha256(h["abcd"], h["H2"].astype(np.float32) if fp32 else h["H2"],
                         "<f4" if fp32 else "<f8", h["trperm"])
    assert got == str(g[key]), "producer list differs from the reference's reduced file"
    for k in ("sp_2mj", "sp_2mt", "sp_l", "sp_tr"):
        assert np.array_equal(getattr(b, k), g[k])
    H2 = h["H2"].astype(np.float32) if fp32 else h["H2"]
    return gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"],
                       H2, h["trperm"], tbme_fp32=fp32, **kw)


def _project(ctx, g):
    from taurus_vap_amd import host
    U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
    cb = host.construct_canonical_basis(U, V)
    assert cb["nocc"] == int(g["nocc0"]) and cb["nemp"] == int(g["nemp0"])
    assert abs(cb["ovac"] - float(g["ovac0"])) < 1e-8 * max(1.0, abs(float(g["ovac0"])))
    return ctx.project_wavefunction(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"],
                                    int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                                    float(g["valence_Z"]), float(g["valence_N"]),
                                    float(g["factor_delta"]))


def _check_projection(r, g, prefix=""):
    over = g[prefix + "pnp_over"]
    assert abs(r["over"] - over) < 1e-10 * max(1.0, abs(over))
    e_ref = (g[prefix + "pnp_ener"] / over).real
    e = (r["ener"] / r["over"]).real
    print("E_proj %.12f reference %.12f" % (e, e_ref), flush=True)
    assert abs(e - e_ref) < 1e-8 * abs(e_ref)
    for key in ("prot", "neut", "prot2", "neut2"):
        assert abs(r[key] / r["over"] - g[prefix + "pnp_" + key] / over) < 1e-10 * max(
            1.0, abs(g[prefix + "pnp_" + key] / over))
    dh = np.max(np.abs(r["H20v"] - g[prefix + "field_H20v"]))
    print("max |dH20v| %.3e" % dh, flush=True)
    assert dh < 1e-10
    return e


@pytest.mark.parametrize("case", ["emax6", "emax4_11"])
def test_large_projection_fp64_matches_reference(case):
    g = load_golden(case)
    print(case, "FP64 TBMEs", flush=True)
    ctx = _context(g, fp32=False)
    try:
        _check_projection(_project(ctx, g), g)
    finally:
        ctx.close()


def test_emax6_projection_fp32_storage():
    """FP32-stored TBMEs, FP64 arithmetic: equal to the stock (rH2 = real32)
    reference on the same deck, and within 1 keV of the FP64 reference."""
    g = load_golden("emax6")
    ctx = _context(g, fp32=True)
    try:
        e32 = _check_projection(_project(ctx, g), g, prefix="fp32_")
    finally:
        ctx.close()
    e64_ref = (g["pnp_ener"] / g["pnp_over"]).real
    print("|E(FP32 TBMEs, device) - E(FP64 reference)| = %.3e keV" % (abs(e32 - e64_ref) * 1e3),
          flush=True)
    assert abs(e32 - e64_ref) < 1e-3


@pytest.mark.parametrize("mode", ["sequential", "overlapped"])
@pytest.mark.parametrize("case", ["emax6", "emax4_11"])
def test_large_iterations_match_reference(case, mode):
    """tg_iterate (constraint_20, projection, multipliers, gradient step,
    adjust_constraints) from the reference's first state, with the reference's
    constraint operators (e_max = 6: <Z>, <N>, Q20, Jx); the canonical basis of
    the evolved state is built on the device between the two iterations:
    "sequential" by tg_canonical_basis, "overlapped" inside tg_iterate_canonical
    (the bench's step: the canonical basis beside the second projection)."""
    g = load_golden(case)
    assert int(g["gradient_type"]) == 0
    n = int(g["ndim"])
    ctx = _context(g, fp32=False)
    try:
        from taurus_vap_amd import host
        ctx.set_gradient(float(g["gradient_eta"]), float(g["gradient_mu"]))
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
        cb = host.construct_canonical_basis(U, V)
        ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"],
                      int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                      float(g["valence_Z"]), float(g["valence_N"]), float(g["factor_delta"]))
        if mode == "overlapped":
            ctx.enable_timing(True)
        for it in range(int(g["niter"])):
            print("%s %s iteration %d" % (case, mode, it + 1), flush=True)
            if mode == "overlapped" and it > 0:
                r, _ = ctx.iterate_canonical()
            else:
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
                assert np.all(np.abs(r["lagrange_lambda1"] - lam) <= tol), (r["lagrange_lambda1"], lam)
            if mode == "sequential":
                ctx.canonical_basis()
        if mode == "overlapped":
            assert ctx.timing("canonical_basis_beside")[1] == int(g["niter"]) - 1
    finally:
        ctx.close()
