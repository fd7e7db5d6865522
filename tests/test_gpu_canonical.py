"""Canonical basis on the device (tg_canonical_basis), construct_canonical_basis
of src/module_wavefunctions.f90:777-945.

Expected values:
* against the reference's own canonical basis: the golden fixtures hold the
  reference's nocc0, nemp0, ovac0, voveru0 and zD0 of their state.  nocc and
  nemp agree exactly, ovac to 1e-9 relative, |v/u| to 1e-7 relative (the signs
  follow the eigenvector signs of each eigensolver, so only magnitudes are
  compared).  ovac = sum log sqrt(1 - v2) and v/u amplify the eigenvalues'
  rounding by 1 / (2 (1 - v2)), up to 5e7 for v2 just below the 1 - eps cut.
* through the projection: a context whose state got its canonical basis on the
  device gives the overlap, energy and particle numbers (2e-9 relative: the
  overlap carries exp(ovac)), H20 and the gradient (1e-9) of a context given
  the reference's canonical basis: the projected quantities do not depend on
  which canonical basis is used.
* through whole iterations: the three reference runs of each batch fixture,
  with the canonical basis of every new state taken on the device, reproduce
  the reference histories to the tolerances of test_gpu_batched.py.
* a state whose rho has a four-fold degenerate eigenvalue (kappa not canonical
  in any eigenbasis the solver picks): the device brings the four-dimensional
  block to canonical form and the projection matches the host route, whose
  canonical basis takes the real Schur form of the block (scipy's, as the
  reference's dgees).
"""
import numpy as np
import pytest

from conftest import load_golden, golden_canon

pytestmark = pytest.mark.gpu


def _ctx(g, **kw):
    from taurus_vap_amd import gpu
    return gpu.Context(int(g["ndim"]), g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"],
                       complex(g["H0"]), g["H1"], g["abcd"], g["H2"], g["trperm"], **kw)


def _setup(g):
    return (int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
            float(g["valence_N"]), float(g["factor_delta"]))


def _project(g, states, c20, lam, device_canon):
    ctx = _ctx(g)
    try:
        ctx.set_gradient(0.01, 0.3, c20, lam)
        ctx.set_states(states, *_setup(g))
        cb = ctx.canonical_basis() if device_canon else None
        ctx.project_resident()
        out = [(ctx.get_results(k), ctx.get_gradient(k)[0]) for k in range(len(states))]
    finally:
        ctx.close()
    return cb, out


def _same_projection(a, b):
    for (r, Z), (q, Zq) in zip(a, b):
        for key in ("over", "ener", "prot", "neut", "prot2", "neut2"):
            assert abs(r[key] - q[key]) <= 2e-9 * max(1.0, abs(q[key])), key
        assert np.max(np.abs(r["H20v"] - q["H20v"])) <= 1e-9
        assert np.max(np.abs(Z - Zq)) <= 1e-9


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_device_canonical_basis_matches_reference(case):
    g = load_golden(case)
    n = int(g["ndim"])
    ref = golden_canon(g)
    ref["D"] = ref["D"].real
    rng = np.random.default_rng(11)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    states = [(g["U0"], g["V0"], ref)]
    cb, dev = _project(g, states, c20, [0.01, 0.02], device_canon=True)
    _, host = _project(g, states, c20, [0.01, 0.02], device_canon=False)
    d = cb[0]
    assert (d["nocc"], d["nemp"]) == (ref["nocc"], ref["nemp"])
    assert abs(d["ovac"] - ref["ovac"]) <= 1e-9 * abs(ref["ovac"])
    m = n - ref["nocc"]
    a, b = np.abs(d["voveru"][:m]), np.abs(np.asarray(ref["voveru"])[:m])
    assert np.all(np.abs(a - b) <= 1e-7 * np.maximum(b, 1e-300)), np.max(np.abs(a - b) / b)
    _same_projection(dev, host)


@pytest.mark.parametrize("case", ["batch_sd", "batch_pf", "batch_sdpf"])
def test_device_canonical_basis_batched_states(case):
    """Three states of one context: each state's device canonical basis gives its
    host-route projection."""
    from taurus_vap_amd import host
    g = load_golden(case)
    n = int(g["ndim"])
    states = []
    for k in range(len(g["betas"])):
        U, V = np.asarray(g["U0"][k]), np.asarray(g["V0"][k])
        states.append((U, V, host.construct_canonical_basis(U, V)))
    rng = np.random.default_rng(12)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    cb, dev = _project(g, states, c20, [0.02, -0.01], device_canon=True)
    _, hst = _project(g, states, c20, [0.02, -0.01], device_canon=False)
    for d, (_, _, h) in zip(cb, states):
        assert (d["nocc"], d["nemp"]) == (h["nocc"], h["nemp"])
        assert abs(d["ovac"] - h["ovac"]) <= 1e-9 * abs(h["ovac"])
    _same_projection(dev, hst)


@pytest.mark.parametrize("case", ["batch_sd", "batch_pf"])
def test_iterations_with_device_canonical_basis_match_reference(case):
    """tg_iterate followed by tg_canonical_basis, the state never leaving the
    device, against the three reference runs."""
    from taurus_vap_amd import host
    from test_gpu_batched import _check_iteration, _prepare
    g = load_golden(case)
    n = int(g["ndim"])
    ctx = _ctx(g)
    try:
        _prepare(ctx, g)
        states = []
        for k in range(len(g["betas"])):
            U, V = np.asarray(g["U0"][k]), np.asarray(g["V0"][k])
            states.append((U, V, host.construct_canonical_basis(U, V)))
        ctx.set_states(states, *_setup(g))
        for it in range(int(g["niter"])):
            print("%s iteration %d" % (case, it + 1), flush=True)
            res = ctx.iterate(0.0)
            for k, r in enumerate(res):
                _check_iteration(g, k, it, r)
            ctx.canonical_basis()
        Us, Vs = ctx.get_state()
        for U, V in zip(Us, Vs):
            assert np.max(np.abs(U.T @ U + V.T @ V - np.eye(n))) <= 1e-10
    finally:
        ctx.close()


def _degenerate_state(n, seed):
    """Canonical pairs (u_k, v_k) rotated by a random orthogonal matrix; pairs 0
    and 1 share v, so rho has a four-fold degenerate eigenvalue."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(0.2, 0.9, n // 2)
    v[1] = v[0]
    u = np.sqrt(1.0 - v ** 2)
    Uc = np.zeros((n, n))
    Vc = np.zeros((n, n))
    for k in range(n // 2):
        Uc[2 * k, 2 * k] = Uc[2 * k + 1, 2 * k + 1] = u[k]
        Vc[2 * k, 2 * k + 1] = v[k]
        Vc[2 * k + 1, 2 * k] = -v[k]
    O, _ = np.linalg.qr(rng.normal(size=(n, n)))
    return O @ Uc, O @ Vc


def test_wide_degenerate_subspace():
    from taurus_vap_amd import host
    g = load_golden("sd64")
    n = int(g["ndim"])
    U, V = _degenerate_state(n, 5)
    assert np.max(np.abs(U.T @ U + V.T @ V - np.eye(n))) <= 1e-12
    hc = host.construct_canonical_basis(U, V)
    states = [(U, V, hc)]
    rng = np.random.default_rng(13)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    _, hst = _project(g, states, c20, [0.0, 0.0], device_canon=False)
    cb, dev = _project(g, states, c20, [0.0, 0.0], device_canon=True)
    d = cb[0]
    assert (d["nocc"], d["nemp"]) == (hc["nocc"], hc["nemp"]) == (0, 0)
    assert abs(d["ovac"] - hc["ovac"]) <= 1e-9 * abs(hc["ovac"])
    assert np.allclose(np.abs(d["voveru"]), np.abs(hc["voveru"]), rtol=1e-9, atol=0)
    _same_projection(dev, hst)


@pytest.mark.parametrize("case", ["batch_sd", "batch_pf"])
def test_iterate_canonical_matches_sequential_calls(case):
    """tg_iterate_canonical (canonical basis beside the projection) against
    tg_canonical_basis + tg_iterate on a second context from the same states:
    every iteration's results and canonical data agree to rounding, both follow
    the reference histories, and the canonical basis ran on the second stream in
    every iteration after the first (the first projection uploads its set-up)."""
    from taurus_vap_amd import host
    from test_gpu_batched import _check_iteration, _prepare
    g = load_golden(case)
    n = int(g["ndim"])
    states = []
    for k in range(len(g["betas"])):
        U, V = np.asarray(g["U0"][k]), np.asarray(g["V0"][k])
        states.append((U, V, host.construct_canonical_basis(U, V)))
    seq, ovl = _ctx(g), _ctx(g)
    try:
        for ctx in (seq, ovl):
            _prepare(ctx, g)
            ctx.set_states(states, *_setup(g))
        ovl.enable_timing(True)
        niter = int(g["niter"])
        for it in range(niter):
            print("%s iteration %d" % (case, it + 1), flush=True)
            if it == 0:
                rs, rc = seq.iterate(0.0), ovl.iterate(0.0)
            else:
                cs = seq.canonical_basis()
                rs = seq.iterate(0.0)
                rc, cc = ovl.iterate_canonical()
                for a, b in zip(cs, cc):
                    assert (a["nocc"], a["nemp"]) == (b["nocc"], b["nemp"])
                    assert abs(a["ovac"] - b["ovac"]) <= 1e-12 * max(1.0, abs(a["ovac"]))
                    assert np.max(np.abs(a["voveru"] - b["voveru"]), initial=0.0) <= 1e-9
            for k, (a, b) in enumerate(zip(rs, rc)):
                _check_iteration(g, k, it, b)
                for key in ("over", "ener", "prot", "neut", "prot2", "neut2"):
                    assert abs(a[key] - b[key]) <= 1e-11 * max(1.0, abs(a[key])), (it, k, key)
                assert abs(a["gradient_norm"] - b["gradient_norm"]) <= 1e-11 * a["gradient_norm"]
                assert np.allclose(a["lagrange_lambda1"], b["lagrange_lambda1"], rtol=1e-10,
                                   atol=1e-12)
                assert a["adjust_iterations"] == b["adjust_iterations"]
        assert ovl.timing("canonical_basis_beside")[1] == niter - 1
        (Us, Vs), (Uo, Vo) = seq.get_state(), ovl.get_state()
        for a, b in zip(Us + Vs, Uo + Vo):
            assert np.max(np.abs(a - b)) <= 1e-11
    finally:
        seq.close()
        ovl.close()
