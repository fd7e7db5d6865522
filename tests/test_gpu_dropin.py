"""The reference's own driver with the GPU projection (integration build of
taurus_vap_amd/fortran/build_dropin.sh) on the reference's Mg24 example
(examples/gradient_mg24: USDB interaction, initial_wf.txt, recipe 1, Q21
constraint), against the published benchmark energies
(data/results_vap_benchmark:2, data/results_hfb_benchmark:2) and against the
first iterations of the CPU reference run (golden fixture mg24_32).

Iteration counts are rounding sensitive (the reference itself converges at 82
instead of the published 83 for HFB when rebuilt), so they are checked within a
band, not exactly.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, GOLDEN, load_golden

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "taurus_vap_amd", "fortran", "_build", "taurus_vap_gpu32.exe")
MG24 = os.path.join(GOLDEN, "mg24")


def _run(M: int, xo: int = 1, iters: int = 3000, plain: bool = False, extra_env=None):
    if not os.path.exists(EXE):
        pytest.skip("integration build not present (taurus_vap_amd/fortran/build_dropin.sh)")
    with tempfile.TemporaryDirectory() as td:
        for f in ("usdb.sho", "initial_wf.txt"):
            with open(os.path.join(MG24, f)) as src, open(os.path.join(td, f), "w") as dst:
                dst.write(src.read())
        deck = open(os.path.join(MG24, "template_input.txt")).read()
        deck = (deck.replace("ETA", "000").replace("MUU", "000").replace("XGGG", str(M))
                .replace("XO", str(xo)))
        deck = deck.replace("Maximum no. of iterations     3000",
                            "Maximum no. of iterations     %4d" % iters)
        with open(os.path.join(td, "input.txt"), "w") as f:
            f.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="4")
        if plain:
            env["TAURUS_GPU_PLAIN"] = "1"
        env.update(extra_env or {})
        with open(os.path.join(td, "input.txt")) as fin:
            r = subprocess.run([EXE], stdin=fin, capture_output=True, text=True, cwd=td, env=env,
                               timeout=600)
    out = r.stdout
    assert r.returncode == 0, out[-2000:] + r.stderr[-2000:]
    assert "This is the end, my only friend, the end." in out
    return out


def _final_energy(out: str) -> float:
    # the example's own extraction: first "Full H" line, 6th field (launch_example.sh:84-87)
    for line in out.splitlines():
        if line.startswith("Full H   "):
            return float(line.split()[5])
    raise AssertionError("no Full H line")


def _iterations(out: str):
    """(iteration, gradient norm, energy) rows of the print_iteration table."""
    rows, on = [], False
    for line in out.splitlines():
        if line.startswith("Iteration") and "Gradient" in line:
            on = True
            continue
        if not on:
            continue
        if "converged" in line or "Maximum number" in line:
            break
        f = line.split()
        if len(f) >= 8 and f[0].isdigit():
            rows.append((int(f[0]), float(f[1]), float(f[2])))
    return rows


def _published(fname: str) -> tuple[int, float]:
    line = open(os.path.join(MG24, fname)).read().splitlines()[1]   # "Recipe 1: Iter = .. Ener = .."
    f = line.split()
    return int(f[4]), float(f[7])


def test_dropin_first_iterations_match_cpu_reference():
    """Iterations 1-3 (5x5 angles, recipe 1) of the drop-in against the
    reference's own CPU run of the same deck (golden mg24_32, printed to 1e-6)."""
    out = _run(5, iters=3)
    rows = _iterations(out)
    g = load_golden("mg24_32")
    e_ref = g["hist_ener"]
    assert len(rows) >= 3
    for k in range(3):
        assert rows[k][0] == k + 1
        assert abs(rows[k][2] - e_ref[k]) < 2e-6, (k, rows[k], e_ref[k])


@pytest.mark.parametrize("plain", [False, True])
def test_dropin_mg24_vap_published(plain):
    """VAP (5x5 gauge angles), recipe 1: published 276 iterations,
    E = -82.773573 MeV."""
    it_pub, e_pub = _published("results_vap_benchmark")
    out = _run(5, plain=plain)
    e = _final_energy(out)
    rows = _iterations(out)
    print("drop-in VAP: %d iterations, E = %.6f (published %d, %.6f)" % (rows[-1][0], e, it_pub, e_pub))
    assert abs(e - e_pub) <= 1.5e-6
    assert abs(rows[-1][0] - it_pub) <= 10


def test_dropin_mg24_vap_graph_conj_pairs():
    """The same VAP run with the projection replayed as a captured hipGraph
    (every iteration re-uploads its state) and one angle of each conjugate
    pair computed: published 276 iterations, E = -82.773573 MeV."""
    it_pub, e_pub = _published("results_vap_benchmark")
    out = _run(5, extra_env={"TAURUS_GPU_GRAPH": "1", "TAURUS_GPU_CONJ_PAIRS": "1"})
    e = _final_energy(out)
    rows = _iterations(out)
    print("drop-in VAP graph+pairs: %d iterations, E = %.6f" % (rows[-1][0], e))
    assert abs(e - e_pub) <= 1.5e-6
    assert abs(rows[-1][0] - it_pub) <= 10


def test_dropin_mg24_hfb_published():
    """HFB (no projection), recipe 1: published E = -80.959670 MeV (83 iterations)."""
    it_pub, e_pub = _published("results_hfb_benchmark")
    out = _run(1)
    e = _final_energy(out)
    rows = _iterations(out)
    print("drop-in HFB: %d iterations, E = %.6f (published %d, %.6f)" % (rows[-1][0], e, it_pub, e_pub))
    assert abs(e - e_pub) <= 1.5e-6
    assert abs(rows[-1][0] - it_pub) <= 10


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "taurus_vap32.exe")


def _run_exe(exe: str, M: int, iters: int, extra_env=None):
    """stdout and occupation_numbers.dat of a run of `exe` on the Mg24 deck."""
    with tempfile.TemporaryDirectory() as td:
        for f in ("usdb.sho", "initial_wf.txt"):
            with open(os.path.join(MG24, f)) as src, open(os.path.join(td, f), "w") as dst:
                dst.write(src.read())
        deck = open(os.path.join(MG24, "template_input.txt")).read()
        deck = (deck.replace("ETA", "000").replace("MUU", "000").replace("XGGG", str(M))
                .replace("XO", "1"))
        deck = deck.replace("Maximum no. of iterations     3000",
                            "Maximum no. of iterations     %4d" % iters)
        with open(os.path.join(td, "input.txt"), "w") as f:
            f.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="4")
        env.update(extra_env or {})
        with open(os.path.join(td, "input.txt")) as fin:
            r = subprocess.run([exe], stdin=fin, capture_output=True, text=True, cwd=td, env=env,
                               timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        occ = open(os.path.join(td, "occupation_numbers.dat")).read()
    return r.stdout, occ


def _numbers(text: str):
    """(label, numbers) of every line that carries numbers."""
    rows = []
    for line in text.splitlines():
        toks = line.split()
        nums, words = [], []
        for t in toks:
            try:
                nums.append(float(t))
            except ValueError:
                words.append(t)
        if nums:
            rows.append((" ".join(words), nums))
    return rows


@pytest.mark.parametrize("M", [5, 1])
def test_dropin_final_printout_matches_reference(M):
    """print_results (projected and quasiparticle-state tables: energy
    decomposition, multipoles, radii, angular momentum, parity, particle
    numbers) and occupation_numbers.dat of the drop-in, whose final projection
    runs on the GPU (tg_project_observables), against the reference's own CPU
    build (oracle/_ref/taurus_vap32.exe) on the same deck after 3 iterations.
    Numbers agree to the printed precision (2e-6 absolute, 2e-6 relative)."""
    if not os.path.exists(EXE) or not os.path.exists(REF_EXE):
        pytest.skip("integration build or reference build not present")
    out_gpu, occ_gpu = _run_exe(EXE, M, 3)
    out_ref, occ_ref = _run_exe(REF_EXE, M, 3)
    checked = 0
    start = "PROJECTED STATE PROPERTIES" if M > 1 else "QUASIPARTICLE STATE PROPERTIES"
    tables = lambda out: out[out.index(start):out.index("COMPLEMENTARY FILES")]
    for text_g, text_r, min_rows in ((tables(out_gpu), tables(out_ref), 30), (occ_gpu, occ_ref, 4)):
        rg, rr = _numbers(text_g), _numbers(text_r)
        assert len(rg) == len(rr) and len(rg) >= min_rows, (len(rg), len(rr))
        for (lg, ng), (lr, nr) in zip(rg, rr):
            if lg.startswith("Label"):
                continue
            assert lg == lr and len(ng) == len(nr), (lg, lr)
            for a, b in zip(ng, nr):
                assert abs(a - b) <= 2e-6 + 2e-6 * abs(b), (lg, ng, nr)
                checked += 1
    print("final printout: %d numbers agree" % checked)
