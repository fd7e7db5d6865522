"""The many-solves workload as the bench runs it: 24 sd (n = 24), 24 pf (n = 40)
and 16 sdpf (n = 64) seeded states, 9x9 gauge angles, one context per basis
holding its solves, the three contexts one group replayed as a captured
hipGraph.

Expected values:
* a sample of the solves (the first, middle and last state of each basis)
  against the CPU oracle's projection of the same state (oracle/pnvap.py,
  project_wavefunction, FP64 TBMEs): overlap 1e-9 relative, projected energy
  1e-8 relative, field_H20v 1e-9 absolute (the tolerances of the e_max = 4
  oracle test);
* every solve of the graph replay equals its eager projection (1e-12 relative:
  the replay re-runs the same kernels on the same data);
* with the TBMEs stored in FP32, every solve's projected energy is within 1 keV of
  its FP64-TBME energy (the north_star tolerance; both on the device)."""
import numpy as np
import pytest

from oracle import pnvap as P

pytestmark = pytest.mark.gpu

SEED = 1234


def _contexts(fp32):
    from taurus_vap_amd import gpu, hamiltonian as H, host, inputs as I
    ctxs, meta = [], []
    for wname, count in I.BATCHED:
        w = I.WORKLOADS[wname]
        h = H.workload_hamiltonian(w, SEED)
        b = h["basis"]
        H2 = h["H2"].astype(np.float32) if fp32 else h["H2"]
        c = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"], H2,
                        h["trperm"], tbme_fp32=fp32)
        sts = []
        for k in range(count):
            U, V = host.seed_state(w, b, SEED + 101 * k)
            sts.append((U, V, host.construct_canonical_basis(U, V)))
        c.set_gradient(0.01, 0.3)
        c.set_states(sts, 9, 9, 2 * np.pi, float(w.Z), float(w.N))
        ctxs.append(c)
        meta.append((w, h, sts))
    return ctxs, meta


def _results(ctxs):
    return [[c.get_results(k) for k in range(c.nstates)] for c in ctxs]


def test_group_of_64_solves_against_oracle():
    from taurus_vap_amd import gpu
    ctxs, meta = _contexts(False)
    try:
        grp = gpu.Group(ctxs, graph=True)
        grp.project()                       # eager
        for c in ctxs:
            c.sync()
        eager = _results(ctxs)
        for _ in range(3):                  # captured, then replayed
            grp.project()
        for c in ctxs:
            c.sync()
        replay = _results(ctxs)
        grp.close()
        for per_e, per_r in zip(eager, replay):
            for a, b in zip(per_e, per_r):
                for key in ("ener", "over"):
                    assert abs(a[key] - b[key]) <= 1e-12 * max(1.0, abs(a[key]))
                assert np.max(np.abs(a["H20v"] - b["H20v"])) <= 1e-12
        for c, (w, h, sts), res in zip(ctxs, meta, replay):
            b = h["basis"]
            ham = dict(H0=h["H0"], H1=h["H1"], abcd=h["abcd"].astype(np.int64) - 1, H2=h["H2"],
                       trperm=h["trperm"].astype(np.int64))
            for k in sorted({0, len(sts) // 2, len(sts) - 1}):
                U, V, cb = sts[k]
                canon = dict(D=cb["D"].astype(complex), nocc=cb["nocc"], nemp=cb["nemp"],
                             ovac=cb["ovac"], voveru=cb["voveru"])
                ref = P.project_wavefunction(U, V, ham, dict(sp_2mj=b.sp_2mj, sp_tr=b.sp_tr - 1), 9, 9,
                                             2 * np.pi, w.Z, w.N, canon=canon)
                r = res[k]
                e, e_ref = (r["ener"] / r["over"]).real, (ref["ener"] / ref["over"]).real
                print("%s state %d: E_gpu %.12f E_oracle %.12f" % (w.name, k, e, e_ref), flush=True)
                assert abs(r["over"] - ref["over"]) < 1e-9 * abs(ref["over"])
                assert abs(e - e_ref) < 1e-8 * abs(e_ref)
                assert np.max(np.abs(r["H20v"] - ref["H20v"])) < 1e-9
    finally:
        for c in ctxs:
            c.close()


def test_group_of_64_solves_fp32_storage_within_1_keV():
    from taurus_vap_amd import gpu
    E = {}
    for fp32 in (False, True):
        ctxs, _ = _contexts(fp32)
        try:
            grp = gpu.Group(ctxs, graph=False)
            grp.project()
            for c in ctxs:
                c.sync()
            E[fp32] = np.array([(r["ener"] / r["over"]).real for per in _results(ctxs) for r in per])
            grp.close()
        finally:
            for c in ctxs:
                c.close()
    assert E[True].size == 64
    dE = np.abs(E[True] - E[False])
    print("max |E(FP32 storage) - E(FP64)| = %.3e keV" % (1e3 * dE.max()), flush=True)
    assert dE.max() < 1e-3
