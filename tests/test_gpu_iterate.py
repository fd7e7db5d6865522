"""Device-resident gradient iteration (tg_iterate: constraint_20 update,
projection, Lagrange multipliers, gradient step, adjust_constraints with the
Cholesky evolution) against the reference's own iterations of the same decks.

The expected values are the reference TAURUS_vap's outputs for its first
iterations (golden fixtures written by oracle/ref_harness.f90: hist_ener,
hist_prot, hist_neut, hist_gradnorm per iteration; lagrange_lambda1 of the first
iteration).  Every iteration after the first starts from the state that
tg_iterate left on the device, so the later rows also check the evolved state
and the adjusted constraints.  Between iterations the canonical basis of the new
state is built on the host (construct_canonical_basis), as the reference does at
the start of each projection.
Tolerances: projected energy 1e-8 relative (the projection's bar), particle
numbers 1e-8 absolute, gradient norm 1e-8 relative, multipliers 1e-8 relative
(1e-10 absolute when they vanish).
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _ctx(g, **kw):
    from taurus_vap_amd import gpu
    return gpu.Context(int(g["ndim"]), g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"],
                       complex(g["H0"]), g["H1"], g["abcd"], g["H2"], g["trperm"], **kw)


def _set(ctx, g, U, V):
    from taurus_vap_amd import host
    canon = host.construct_canonical_basis(U, V)
    ctx.set_state(U, V, canon["D"], canon["nocc"], canon["nemp"], canon["ovac"], canon["voveru"],
                  int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
                  float(g["valence_N"]), float(g["factor_delta"]))


@pytest.mark.parametrize("case,opts", [
    ("sd64", {}),
    ("pf64", {}),
    ("pf64", {"graph": True, "conj_pairs": True}),
])
def test_iterations_match_reference(case, opts):
    g = load_golden(case)
    assert int(g["gradient_type"]) == 0
    n = int(g["ndim"])
    niter = int(g["niter"])
    ctx = _ctx(g, **opts)
    try:
        ctx.set_gradient(float(g["gradient_eta"]), float(g["gradient_mu"]))
        dim = int(g["constraint_dim"])
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        U, V = np.asarray(g["U0"]), np.asarray(g["V0"])
        for it in range(niter):
            print("%s %s iteration %d" % (case, opts, it + 1), flush=True)
            _set(ctx, g, U, V)
            r = ctx.iterate(0.0)
            over = r["over"]
            e = (r["ener"] / over).real
            assert abs(e - g["hist_ener"][it]) <= 1e-8 * abs(g["hist_ener"][it]), (it, e)
            assert abs((r["prot"] / over).real - g["hist_prot"][it]) <= 1e-8, it
            assert abs((r["neut"] / over).real - g["hist_neut"][it]) <= 1e-8, it
            gn = g["hist_gradnorm"][it]
            assert abs(r["gradient_norm"] - gn) <= 1e-8 * gn, (it, r["gradient_norm"], gn)
            if it == 0 and dim > 0:
                lam = np.asarray(g["lagrange_lambda1"])
                tol = np.maximum(1e-8 * np.abs(lam), 1e-10)
                assert np.all(np.abs(r["lagrange_lambda1"] - lam) <= tol), (r["lagrange_lambda1"], lam)
            assert not r["converged"]
            assert 1 <= r["adjust_iterations"] <= 100
            U, V = ctx.get_state()
            # the evolved state is a Bogoliubov transformation: U^T U + V^T V = 1
            assert np.max(np.abs(U.T @ U + V.T @ V - np.eye(n))) <= 1e-10
    finally:
        ctx.close()


def test_iterate_requires_state_after_evolution():
    """After an evolving iteration the canonical basis is stale: tg_iterate
    refuses to run until a new state is set."""
    from taurus_vap_amd import gpu
    g = load_golden("pf64")
    ctx = _ctx(g)
    try:
        ctx.set_gradient(float(g["gradient_eta"]), float(g["gradient_mu"]))
        ctx.set_constraints(np.asarray(g["constraint_HO"]).reshape(-1), g["constraint_val"],
                            int(g["constraint_pair"]), 1e-6)
        _set(ctx, g, np.asarray(g["U0"]), np.asarray(g["V0"]))
        ctx.iterate(0.0)
        with pytest.raises(gpu.TaurusGPUError):
            ctx.iterate(0.0)
        # a converged call (gradient_eps above the norm) keeps the state
        U, V = ctx.get_state()
        _set(ctx, g, U, V)
        r = ctx.iterate(1e6)
        assert r["converged"] and r["adjust_iterations"] == 0
        U2, V2 = ctx.get_state()
        assert np.array_equal(U, U2) and np.array_equal(V, V2)
    finally:
        ctx.close()
