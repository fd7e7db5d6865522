"""Writes the large golden fixtures under tests/golden/ from runs of the
reference TAURUS_vap (built from its own sources into oracle/_ref by
oracle/build_ref.sh) in the harness's 'proj' mode (oracle/ref_harness.f90),
which dumps only what stays small at large bases:

* emax6.npz    -- e_max = 6 (336 states), 15x15 gauge angles, 10p/12n, Q20 + Jx
                  constraints, 2 iterations, FP64 TBMEs (rH2 = real64) and the
                  stock FP32 build on the same deck;
* emax4_11.npz -- e_max = 4 (140 states), 11x11 angles, 8p/8n, 2 iterations, FP64;
* a19_<case>.npz -- small decks with gradient_type 1 or 2: the h and H11 matrices
                  diagonalize_hsp_and_H11 hands to dsyev, their eigenvalues and the
                  recipe steps (src/module_gradient.f90:79-453);
* jscheme_<case>.npz -- J-scheme interactions (hamil_type 3 with a .01b file,
                  hamil_type 4 with the centre-of-mass correction), e_max = 2: the
                  reference's reduced list, H0 and H1, and the projection of the
                  first iterations.

In every case the reference builds its own Hamiltonian from the .sho file
(hamil_read = 0) and writes its reduced file; the fixture keeps the SHA-256 of
that file's element payload (abcd, H2, trperm), which pins the native producer.

Run in the build container (needs /root/reference and oracle/_ref), e.g.
    python tests/golden/make_golden_large.py emax4_11 a19_pf_g1
The fixtures are data only: inputs and the reference's outputs.
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from taurus_vap_amd import inputs as I          # noqa: E402
from oracle import pnvap as P                    # noqa: E402

REF = os.environ.get("TAURUS_REFERENCE", "/root/reference")
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def red_payload_sha256(path: str, ndim: int, rh2_bytes: int) -> tuple[str, int]:
    """SHA-256 over the element arrays of a reduced file: abcd (0-based int16, 4
    per element), H2 (float32 or float64) and trperm (int8), in file order."""
    h = P.read_red(path, ndim, rh2_bytes)
    d = hashlib.sha256()
    d.update(np.ascontiguousarray(h["abcd"], dtype="<i2").tobytes())
    d.update(np.ascontiguousarray(h["H2"], dtype="<f4" if rh2_bytes == 4 else "<f8").tobytes())
    d.update(np.ascontiguousarray(h["trperm"], dtype="<i1").tobytes())
    return d.hexdigest(), len(h["trperm"])


def run_proj(workdir: str, prec: int, kmax: int, threads: int, deck: str = "input.txt"):
    exe = os.path.join(REFBIN, "harness%d.exe" % prec)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
               OMP_STACKSIZE="256M")
    with open(os.path.join(workdir, deck)) as fin, \
            open(os.path.join(workdir, "out%d.txt" % prec), "w") as fout:
        subprocess.run(["bash", "-c", "ulimit -s unlimited; exec \"$0\" \"$@\"", exe, "proj",
                        "dump%d.bin" % prec, str(kmax)],
                       stdin=fin, stdout=fout, stderr=subprocess.STDOUT, cwd=workdir,
                       check=True, env=env)
    return P.read_harness_dump(os.path.join(workdir, "dump%d.bin" % prec))


def one(d, k):
    return d[k][0]


SCALARS = ("ndim", "valence_Z", "valence_N", "Mphip", "Mphin", "facpi", "gradient_type",
           "gradient_eta", "gradient_mu", "constraint_dim", "constraint_pair", "factor_delta",
           "H0", "nocc0", "nemp0", "ovac0", "pnp_over", "pnp_ener", "pnp_prot", "pnp_neut",
           "pnp_prot2", "pnp_neut2", "gradient_norm", "niter")
ARRAYS = ("sp_2mj", "sp_2mt", "sp_l", "sp_2j", "sp_tr", "voveru0", "field_H20v", "gradient_Zi",
          "hist_ener", "hist_prot", "hist_neut", "hist_gradnorm", "U0", "V0")
OPTIONAL = ("constraint_HO", "constraint_val", "lagrange_lambda1", "hsp_RR", "H11", "eigen_hsp",
            "eigen_H11", "info_hsp", "info_H11")
PER_ANGLE = ("phip", "phin", "weip", "wein", "rot_over", "rot_ener", "rot_prot", "rot_neut",
             "rot_prot2", "rot_neut2")


def pack(d: dict, prefix: str = "", keep_state: bool = True) -> dict:
    out = {}
    for k in SCALARS:
        out[prefix + k] = np.asarray(one(d, k))
    for k in ARRAYS:
        if k in ("U0", "V0") and not keep_state:
            continue
        out[prefix + k] = np.asarray(one(d, k))
    for k in OPTIONAL:
        if k in d:
            out[prefix + k] = np.asarray(one(d, k))
    for k in PER_ANGLE:
        out[prefix + k] = np.asarray(d[k])
    return out


def workload_case(w: I.Workload, kmax: int, precs=(64,), seed: int = 1234, threads: int = 8,
                  **kw) -> dict:
    """Reference runs of workload w at each precision (concurrently), Hamiltonian
    built by the reference from the .sho (hamil_read = 0)."""
    dirs = {p: tempfile.mkdtemp(prefix="tg_golden_%s_%d_" % (w.name, p)) for p in precs}
    try:
        for p in precs:
            I.write_workload(w, dirs[p], seed=seed, hamil_read=0, iters=kmax, **kw)
        with ThreadPoolExecutor(len(precs)) as ex:
            futs = {p: ex.submit(run_proj, dirs[p], p, kmax, max(1, threads // len(precs)))
                    for p in precs}
            dumps = {p: f.result() for p, f in futs.items()}
        out = dict(workload=w.name, seed=seed)
        for i, p in enumerate(precs):
            d = dumps[p]
            n = int(one(d, "ndim"))
            sha, ntb = red_payload_sha256(os.path.join(dirs[p], w.name + ".red"), n,
                                          4 if p == 32 else 8)
            out["red_sha256_%d" % p] = np.asarray(sha)
            out["ntb"] = np.asarray(ntb)
            out.update(pack(d, "" if i == 0 else "fp%d_" % p, keep_state=(i == 0)))
            if i > 0:
                # every build starts from the same deck and seed state
                assert np.array_equal(one(d, "U0"), one(dumps[precs[0]], "U0")) or p == 32
        return out
    finally:
        for p in precs:
            shutil.rmtree(dirs[p], ignore_errors=True)


def mg24_a19_case(M: int, kmax: int, gtype: int) -> dict:
    with tempfile.TemporaryDirectory() as td:
        shutil.copy(os.path.join(REF, "int", "usdb.sho"), td)
        shutil.copy(os.path.join(REF, "examples", "gradient_mg24", "data", "initial_wf.txt"), td)
        with open(os.path.join(REF, "examples", "gradient_mg24", "data", "template_input.txt")) as f:
            deck = f.read()
        deck = (deck.replace("ETA", "000").replace("MUU", "000").replace("XGGG", str(M))
                .replace("XO", str(gtype)))
        deck = deck.replace("Maximum no. of iterations     3000",
                            "Maximum no. of iterations     %4d" % kmax)
        with open(os.path.join(td, "input.txt"), "w") as f:
            f.write(deck)
        d = run_proj(td, 32, kmax, 1)
        n = int(one(d, "ndim"))
        sha, ntb = red_payload_sha256(os.path.join(td, "usdb.red"), n, 4)
        h = P.read_red(os.path.join(td, "usdb.red"), n, 4)
        out = pack(d)
        out.update(red_sha256_32=np.asarray(sha), ntb=np.asarray(ntb),
                   abcd=(h["abcd"] + 1).astype(np.int16), H2=h["H2"].astype(np.float32),
                   trperm=h["trperm"].astype(np.int8), H1=h["H1"], prec=np.asarray(32))
        return out


def small_a19_case(wname: str, M: int, kmax: int, gtype: int, eta: float = 0.0,
                   mu: float = 0.0) -> dict:
    w = dataclasses.replace(I.WORKLOADS[wname], M=M, gradient_type=gtype, eta=eta, mu=mu)
    with tempfile.TemporaryDirectory() as td:
        I.write_workload(w, td, seed=1234, hamil_read=0, iters=kmax)
        d = run_proj(td, 64, kmax, 1)
        n = int(one(d, "ndim"))
        h = P.read_red(os.path.join(td, w.name + ".red"), n, 8)
        sha, ntb = red_payload_sha256(os.path.join(td, w.name + ".red"), n, 8)
        out = pack(d)
        out.update(red_sha256_64=np.asarray(sha), ntb=np.asarray(ntb),
                   abcd=(h["abcd"] + 1).astype(np.int16), H2=h["H2"],
                   trperm=h["trperm"].astype(np.int8),
                   H1=h["H1"], prec=np.asarray(64))
        return out


def jscheme_case(htype: int, com: int, Z: int, N: int, M: int, kmax: int,
                 seed: int = 11) -> dict:
    """Seeded J-scheme interaction on the e_max = 2 space (40 states), written by
    inputs.write_jscheme_synthetic; the reference builds it (hamil_read = 0) and runs kmax
    iterations at M x M gauge angles."""
    name = "jscheme%d" % htype
    with tempfile.TemporaryDirectory() as td:
        shells = I.write_jscheme_synthetic(os.path.join(td, name), htype, com, seed)
        I.write_deck(os.path.join(td, "input.txt"), name, Z, N, M, M, hamil_read=0, hamil_com=com,
                     iter_max=kmax, grad_eps=1e-12)
        basis = I.Basis(shells)
        U, V = I.seed_wavefunction(basis, Z, N, seed=seed + 2, sigma=0.1)
        I.write_wavefunction_text(os.path.join(td, "initial_wf.txt"), shells, U, V, jscheme=True)
        d = run_proj(td, 64, kmax, 1)
        n = int(one(d, "ndim"))
        h = P.read_red(os.path.join(td, name + ".red"), n, 8)
        sha, ntb = red_payload_sha256(os.path.join(td, name + ".red"), n, 8)
        out = pack(d)
        out.update(red_sha256_64=np.asarray(sha), ntb=np.asarray(ntb), htype=np.asarray(htype),
                   hamil_com=np.asarray(com), seed=np.asarray(seed),
                   abcd=(h["abcd"] + 1).astype(np.int16), H2=h["H2"],
                   trperm=h["trperm"].astype(np.int8), H0=np.asarray(h["H0"]), H1=h["H1"],
                   prec=np.asarray(64))
        return out


CASES = {
    # e_max = 6, 15x15, Q20 + Jx, FP64 and FP32 TBMEs, 2 iterations
    "emax6": lambda: workload_case(I.WORKLOADS["emax6"], 2, precs=(64, 32), threads=8),
    # e_max = 4, 11x11, 2 iterations, FP64
    "emax4_11": lambda: workload_case(I.WORKLOADS["emax4"], 2, precs=(64,), threads=8),
    # diagonalize_hsp_and_H11 + the recipe steps (3 iterations each)
    "a19_mg24_g1": lambda: mg24_a19_case(5, 3, 1),
    "a19_mg24_g2": lambda: mg24_a19_case(5, 3, 2),
    "a19_pf_g1": lambda: small_a19_case("pf", 9, 3, 1),
    "a19_pf_g2": lambda: small_a19_case("pf", 9, 3, 2),
    # one gauge angle: the h/Delta of the M = 1 projection feed H11 directly
    "a19_sd_m1_g1": lambda: small_a19_case("sd", 1, 3, 1),
    # J-scheme interactions: type 3 (.01b one-body file), 3p/4n, and type 4 (kinetic
    # one-body part) with the centre-of-mass correction (.com), 4p/4n; 5x5 angles
    "jscheme_t3": lambda: jscheme_case(3, 0, 3, 4, 5, 2),
    "jscheme_t4_com": lambda: jscheme_case(4, 1, 4, 4, 5, 2),
}


def main():
    only = sys.argv[1:] or list(CASES)
    for name in only:
        c = CASES[name]()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **c)
        print("wrote", path, os.path.getsize(path), flush=True)


if __name__ == "__main__":
    main()
