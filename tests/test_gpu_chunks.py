"""Memory-short mode of the batched projection: the (state, angle) items run
chunk by chunk (tg_set_batch_limit, or automatically when the work matrices of
all items do not fit), and only A_phi and H20_phi of every item are kept for
sum_gauge.  Chunked runs must give the results of the all-at-once run, which the
projection tests pin to the reference: scalars to 1e-11 relative, field_H20v and
the gradient to 1e-10.
"""
import numpy as np
import pytest

from conftest import load_golden, golden_canon

pytestmark = pytest.mark.gpu


def _ctx(g, **kw):
    from taurus_vap_amd import gpu
    return gpu.Context(int(g["ndim"]), g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"],
                       complex(g["H0"]), g["H1"], g["abcd"], g["H2"], g["trperm"], **kw)


def _same(r, ref, Z, Zref):
    for k in ("over", "ener", "prot", "neut", "prot2", "neut2", "pari"):
        assert abs(r[k] - ref[k]) <= 1e-11 * max(1.0, abs(ref[k])), k
    assert np.max(np.abs(r["H20v"] - ref["H20v"])) <= 1e-10
    assert np.max(np.abs(Z - Zref)) <= 1e-10


@pytest.mark.parametrize("case", ["sd64", "pf64"])
@pytest.mark.parametrize("graph", [False, True])
def test_chunked_projection_matches_all_at_once(case, graph):
    g = load_golden(case)
    n = int(g["ndim"])
    cb = golden_canon(g)
    rng = np.random.default_rng(9)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    lam = [0.01, 0.02]
    runs = {}
    for limit in (0, 7, 1):
        ctx = _ctx(g, graph=graph)
        try:
            ctx.set_batch_limit(limit)
            ctx.set_gradient(0.01, 0.3, c20, lam)
            ctx.set_state(g["U0"], g["V0"], cb["D"].real, cb["nocc"], cb["nemp"], cb["ovac"],
                          cb["voveru"], int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                          float(g["valence_Z"]), float(g["valence_N"]), float(g["factor_delta"]),
                          iprint=1)
            for rep in range(3):    # eager, captured, replayed when graph
                ctx.project_resident()
            runs[limit] = (ctx.get_results(), ctx.get_gradient()[0])
            if limit == 7 and graph:
                # a new limit on the same context: the captured graph is not reused
                ctx.set_batch_limit(5)
                ctx.set_state(g["U0"], g["V0"], cb["D"].real, cb["nocc"], cb["nemp"], cb["ovac"],
                              cb["voveru"], int(g["Mphip"]), int(g["Mphin"]),
                              float(g["facpi"]), float(g["valence_Z"]), float(g["valence_N"]),
                              float(g["factor_delta"]), iprint=1)
                ctx.project_resident()
                r5 = ctx.get_results()
                assert abs(r5["ener"] - runs[7][0]["ener"]) <= 1e-11 * abs(runs[7][0]["ener"])
        finally:
            ctx.close()
    ref, Zref = runs[0]
    assert abs((ref["ener"] / ref["over"]).real - float(g["pnp_ener"].real / g["pnp_over"].real)) \
        <= 1e-8 * abs(float(g["pnp_ener"].real / g["pnp_over"].real))
    for limit in (7, 1):
        r, Z = runs[limit]
        _same(r, ref, Z, Zref)


def test_chunked_states_hold_whole_states():
    """Three states in one context with a limit of two states plus a few items:
    chunks of two whole states, then one."""
    from taurus_vap_amd import host
    g = load_golden("batch_pf")
    A = int(g["Mphip"]) * int(g["Mphin"])
    states = []
    for k in range(3):
        U, V = np.asarray(g["U0"][k]), np.asarray(g["V0"][k])
        states.append((U, V, host.construct_canonical_basis(U, V)))
    setup = (int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
             float(g["valence_N"]), float(g["factor_delta"]))
    out = {}
    for limit in (0, 2 * A + 5):
        ctx = _ctx(g)
        try:
            ctx.set_batch_limit(limit)
            ctx.set_gradient(0.01, 0.0)
            ctx.set_states(states, *setup)
            ctx.project_resident()
            out[limit] = [(ctx.get_results(k), ctx.get_gradient(k)[0]) for k in range(3)]
        finally:
            ctx.close()
    for (r, Z), (ref, Zref) in zip(out[2 * A + 5], out[0]):
        for k in ("over", "ener", "prot", "neut"):
            assert abs(r[k] - ref[k]) <= 1e-11 * max(1.0, abs(ref[k])), k
        assert np.max(np.abs(r["H20v"] - ref["H20v"])) <= 1e-10
        assert np.max(np.abs(Z - Zref)) <= 1e-10


def test_observables_in_blocks_match_all_at_once():
    """tg_project_observables when the angles do not run all at once: blocks of the
    memory-short chunk size (batch limit 5) and one angle at a time (the plain
    per-angle mode) give the per-angle records and the rho sum of the all-at-once
    run; a projection after it (conjugate pairs, gradient step) still gives its
    own results: the caller's angle set-up, iprint and gradient are restored."""
    g = load_golden("sd64")
    n = int(g["ndim"])
    cb = golden_canon(g)
    rng = np.random.default_rng(4)
    ops = [rng.normal(size=(n, n)) for _ in range(2)]
    pairs = [(ops[0], ops[0] @ ops[0])]
    shell = np.arange(n) % 3
    out = {}
    for mode in ("all", "blocks", "plain"):
        ctx = _ctx(g, plain=(mode == "plain"), conj_pairs=(mode == "all"))
        try:
            if mode == "blocks":
                ctx.set_batch_limit(5)
            ctx.set_gradient(0.01, 0.2)
            ctx.set_state(g["U0"], g["V0"], cb["D"].real, cb["nocc"], cb["nemp"], cb["ovac"],
                          cb["voveru"], int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]),
                          float(g["valence_Z"]), float(g["valence_N"]), float(g["factor_delta"]))
            ctx.project_resident()
            before = ctx.get_results()
            print("observables", mode, flush=True)
            out[mode] = ctx.project_observables(ops, pairs, shell, 3)
            ctx.project_resident()
            after = ctx.get_results()
            for k in ("over", "ener", "prot", "neut"):
                assert abs(after[k] - before[k]) <= 1e-13 * max(1.0, abs(before[k])), (mode, k)
            assert np.max(np.abs(after["H20v"] - before["H20v"])) <= 1e-13, mode
        finally:
            ctx.close()
    rec, rho = out["all"]
    for mode in ("blocks", "plain"):
        r2, rho2 = out[mode]
        assert np.max(np.abs(r2 - rec)) <= 1e-11 * max(1.0, np.max(np.abs(rec))), mode
        assert np.max(np.abs(rho2 - rho)) <= 1e-11 * max(1.0, np.max(np.abs(rho))), mode
