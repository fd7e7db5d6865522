"""hipGraph replay of the projection (TG_FLAG_GRAPH) against the eager launch
sequence of the same library on the same inputs.

The graph is captured on the second projection of a configuration and
replayed afterwards; every replay is compared with an eager context, with
eager work of the library (a field contraction on the same context and
stream) and host synchronisations between the replays, and a new state of the
same configuration (only its set-up re-uploaded) must replay correctly too.
The gradient step inside the graph carries its momentum term from replay to
replay, so equal gradients also show that the replays ran in order.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _contexts(g):
    from taurus_vap_amd import gpu
    n = int(g["ndim"])
    args = (n, g["sp_2mj"], g["sp_2mt"], g["sp_l"], g["sp_tr"], complex(g["H0"]), g["H1"],
            g["abcd"], g["H2"], g["trperm"])
    return gpu.Context(*args, graph=False), gpu.Context(*args, graph=True)


def _set(ctx, g, U, V, canon):
    ctx.set_state(U, V, canon["D"], canon["nocc"], canon["nemp"], canon["ovac"], canon["voveru"],
                  int(g["Mphip"]), int(g["Mphin"]), float(g["facpi"]), float(g["valence_Z"]),
                  float(g["valence_N"]))


@pytest.mark.parametrize("case", ["sd64", "pf64"])
def test_graph_replay_matches_eager(case, capfd, monkeypatch):
    monkeypatch.setenv("TG_GRAPH_VERBOSE", "1")
    from taurus_vap_amd import host
    g = load_golden(case)
    n = int(g["ndim"])
    eager, graph = _contexts(g)
    rng = np.random.default_rng(11)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    lam = [0.03, -0.01]
    U0, V0 = np.asarray(g["U0"]), np.asarray(g["V0"])
    X = rng.normal(size=(n, n)) * 0.02
    U1, V1 = host.evolve(U0, V0, np.triu(X, 1) - np.triu(X, 1).T)
    rho = np.asarray(g["ang_rhoLR"][0])
    kLR, kRL = np.asarray(g["ang_kappaLR"][0]), np.asarray(g["ang_kappaRL"][0])
    try:
        for ctx in (eager, graph):
            ctx.set_gradient(0.01, 0.3, c20, lam)
        for U, V in ((U0, V0), (U1, V1)):
            canon = host.construct_canonical_basis(U, V)
            for ctx in (eager, graph):
                _set(ctx, g, U, V, canon)
            for rep in range(4):
                eager.project_resident()
                graph.project_resident()
                re, rg = eager.get_results(), graph.get_results()
                Ze, ne = eager.get_gradient()
                Zg, ng = graph.get_gradient()
                for k in ("over", "ener", "prot", "neut", "prot2", "neut2"):
                    assert abs(rg[k] - re[k]) <= 1e-12 * max(1.0, abs(re[k])), (rep, k)
                assert np.max(np.abs(rg["H20v"] - re["H20v"])) <= 1e-12
                assert np.max(np.abs(Zg - Ze)) <= 1e-12
                assert abs(ng - ne) <= 1e-12 * max(1.0, ne)
                # eager library work on the graph's context and stream between replays
                gam, _, _, _ = graph.calculate_fields(rho, kLR, kRL)
                assert np.all(np.isfinite(gam))
    finally:
        eager.close()
        graph.close()
    err = capfd.readouterr().err
    assert "projection captured" in err, "the graph path was not taken"
