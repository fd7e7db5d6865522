"""Writes the golden fixtures under tests/golden/ by running the reference
TAURUS_vap (built from its own sources into oracle/_ref by oracle/build_ref.sh)
through oracle/ref_harness.f90 on seeded synthetic decks and on the reference's
own Mg24 example (examples/gradient_mg24, int/usdb.sho).

Run in the build container (needs /root/reference and oracle/_ref):
    python tests/golden/make_golden.py
The fixtures are data only: inputs and the reference's outputs.
"""
from __future__ import annotations

import dataclasses
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from taurus_vap_amd import inputs as I          # noqa: E402
from oracle import pnvap as P                    # noqa: E402

REF = os.environ.get("TAURUS_REFERENCE", "/root/reference")
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def run_harness(workdir: str, prec: int, deck: str = "input.txt"):
    exe = os.path.join(REFBIN, "harness%d.exe" % prec)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    with open(os.path.join(workdir, deck)) as fin, open(os.path.join(workdir, "out%d.txt" % prec), "w") as fout:
        subprocess.run([exe, "iter", "dump%d.bin" % prec], stdin=fin, stdout=fout, cwd=workdir,
                       check=True, env=env)
    return P.read_harness_dump(os.path.join(workdir, "dump%d.bin" % prec))


def pack_case(d: dict, ham: dict, angle_sel, prec: int) -> dict:
    n = int(d["ndim"][0])
    out = dict(
        ndim=n, prec=prec,
        sp_2mj=d["sp_2mj"][0], sp_2mt=d["sp_2mt"][0], sp_l=d["sp_l"][0], sp_2j=d["sp_2j"][0],
        sp_tr=d["sp_tr"][0],
        valence_Z=d["valence_Z"][0], valence_N=d["valence_N"][0],
        Mphip=d["Mphip"][0], Mphin=d["Mphin"][0], facpi=d["facpi"][0],
        gradient_type=d["gradient_type"][0], gradient_eta=d["gradient_eta"][0],
        gradient_mu=d["gradient_mu"][0], constraint_dim=d["constraint_dim"][0],
        constraint_pair=d["constraint_pair"][0], factor_delta=d["factor_delta"][0],
        H0=d["H0"][0], H1=d["H1"][0],
        abcd=(ham["abcd"] + 1).astype(np.int16), H2=ham["H2"].astype(np.float32 if prec == 32 else np.float64),
        trperm=ham["trperm"].astype(np.int8),
        U0=d["U0"][0], V0=d["V0"][0], rhoRR=d["rhoRR"][0], kappaRR=d["kappaRR"][0],
        nocc0=d["nocc0"][0], nemp0=d["nemp0"][0], ovac0=d["ovac0"][0], voveru0=d["voveru0"][0],
        zD0=d["zD0"][0].real,
        nangles=d["nangles"][0],
        phip=np.array(d["phip"]), phin=np.array(d["phin"]),
        weip=np.array(d["weip"]), wein=np.array(d["wein"]),
        rot_over=np.array(d["rot_over"]), rot_ener=np.array(d["rot_ener"]),
        rot_prot=np.array(d["rot_prot"]), rot_neut=np.array(d["rot_neut"]),
        rot_prot2=np.array(d["rot_prot2"]), rot_neut2=np.array(d["rot_neut2"]),
        pnp_over=d["pnp_over"][0], pnp_ener=d["pnp_ener"][0], pnp_prot=d["pnp_prot"][0],
        pnp_neut=d["pnp_neut"][0], pnp_prot2=d["pnp_prot2"][0], pnp_neut2=d["pnp_neut2"][0],
        pnp_pari=d["pnp_pari"][0], iter_print=d["iter_print"][0],
        pnp_A=d["pnp_A"][0], pnp_H20=d["pnp_H20"][0], pnp_AH20=d["pnp_AH20"][0],
        field_H20v=d["field_H20v"][0],
        gradient_Zi=d["gradient_Zi"][0], gradient_norm=d["gradient_norm"][0],
        niter=d["niter"][0], hist_ener=d["hist_ener"][0], hist_prot=d["hist_prot"][0],
        hist_neut=d["hist_neut"][0], hist_gradnorm=d["hist_gradnorm"][0],
        angle_sel=np.array(angle_sel, dtype=np.int32),
    )
    if int(out["constraint_dim"]) > 0:
        out["constraint_HO"] = d["constraint_HO"][0]
        out["constraint_val"] = d["constraint_val"][0]
        out["constraint_20"] = d["constraint_20"][0]
        out["lagrange_lambda1"] = d["lagrange_lambda1"][0]
    for key in ("rot_A", "Utilde", "Vtilde", "rhoLR", "kappaLR", "kappaRL", "gammaLR", "hspLR",
                "deltaLR", "deltaRL", "rot_H20"):
        out["ang_" + key] = np.stack([d[key][k] for k in angle_sel])
    return out


def synthetic_case(wname: str, prec: int, M: int, iters: int, angle_sel, seed: int = 1234,
                   **kw) -> dict:
    w = I.WORKLOADS[wname]
    with tempfile.TemporaryDirectory() as td:
        I.write_workload(w, td, seed=seed, M=M, iters=iters, **kw)
        d = run_harness(td, prec)
        n = int(d["ndim"][0])
        ham = P.read_red(os.path.join(td, w.name + ".red"), n, 4 if prec == 32 else 8)
        return pack_case(d, ham, angle_sel, prec)


def batch_case(wname: str, betas, prec: int = 64, M: int = 9, iters: int = 3,
               seed: int = 1234) -> dict:
    """Independent Q20-constrained solves of one basis (one reference run per
    beta2 value, same seed and Hamiltonian): the common inputs once, and per
    solve its iteration-1 state, constraint values and iteration history."""
    w0 = I.WORKLOADS[wname]
    runs = []
    for b in betas:
        w = dataclasses.replace(w0, M=M, constraints={"Constraint multipole Q20": float(b)})
        with tempfile.TemporaryDirectory() as td:
            I.write_workload(w, td, seed=seed, M=M, iters=iters)
            d = run_harness(td, prec)
            n = int(d["ndim"][0])
            ham = P.read_red(os.path.join(td, w.name + ".red"), n, 4 if prec == 32 else 8)
        runs.append((d, ham))
    d0, ham = runs[0]
    n = int(d0["ndim"][0])
    for d, _ in runs:
        assert int(d["niter"][0]) == iters
        for k in ("facpi", "constraint_dim", "constraint_pair", "Mphip", "Mphin"):
            assert np.array_equal(np.asarray(d[k][0]), np.asarray(d0[k][0])), k
        assert np.array_equal(np.asarray(d["constraint_HO"][0]), np.asarray(d0["constraint_HO"][0]))
    st = lambda key: np.stack([np.asarray(d[key][0]) for d, _ in runs])
    return dict(
        ndim=n, prec=prec, betas=np.asarray(betas, dtype=np.float64),
        sp_2mj=d0["sp_2mj"][0], sp_2mt=d0["sp_2mt"][0], sp_l=d0["sp_l"][0], sp_tr=d0["sp_tr"][0],
        valence_Z=d0["valence_Z"][0], valence_N=d0["valence_N"][0],
        Mphip=d0["Mphip"][0], Mphin=d0["Mphin"][0], facpi=d0["facpi"][0],
        gradient_type=d0["gradient_type"][0], gradient_eta=d0["gradient_eta"][0],
        gradient_mu=d0["gradient_mu"][0], constraint_dim=d0["constraint_dim"][0],
        constraint_pair=d0["constraint_pair"][0], factor_delta=d0["factor_delta"][0],
        H0=d0["H0"][0], H1=d0["H1"][0],
        abcd=(ham["abcd"] + 1).astype(np.int16),
        H2=ham["H2"].astype(np.float32 if prec == 32 else np.float64),
        trperm=ham["trperm"].astype(np.int8),
        constraint_HO=d0["constraint_HO"][0],
        U0=st("U0"), V0=st("V0"), constraint_val=st("constraint_val"),
        lagrange_lambda1=st("lagrange_lambda1"), niter=iters,
        hist_ener=st("hist_ener"), hist_prot=st("hist_prot"), hist_neut=st("hist_neut"),
        hist_gradnorm=st("hist_gradnorm"),
    )


def mg24_case(prec: int, M: int, iters: int, angle_sel, xo: int = 1) -> dict:
    with tempfile.TemporaryDirectory() as td:
        shutil.copy(os.path.join(REF, "int", "usdb.sho"), td)
        shutil.copy(os.path.join(REF, "examples", "gradient_mg24", "data", "initial_wf.txt"), td)
        with open(os.path.join(REF, "examples", "gradient_mg24", "data", "template_input.txt")) as f:
            deck = f.read()
        deck = (deck.replace("ETA", "000").replace("MUU", "000").replace("XGGG", str(M))
                .replace("XO", str(xo)))
        deck = deck.replace("Maximum no. of iterations     3000",
                            "Maximum no. of iterations     %4d" % iters)
        with open(os.path.join(td, "input.txt"), "w") as f:
            f.write(deck)
        d = run_harness(td, prec)
        n = int(d["ndim"][0])
        ham = P.read_red(os.path.join(td, "usdb.red"), n, 4 if prec == 32 else 8)
        return pack_case(d, ham, angle_sel, prec)


def pfaffian_cases(seed: int = 7):
    """Random complex skew matrices through ZSKPFA('u','p'); sizes cover the
    unblocked (ZSKTF2) and blocked (ZLASKTRF, N >= 256) branches."""
    rng = np.random.default_rng(seed)
    sizes = [2, 4, 6, 8, 10, 16, 30, 48, 64, 100, 128, 130, 256, 300]
    mats = []
    for n in sizes:
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        mats.append(X - X.T)
    # degenerate: zero column forces INFO > 0 -> pf = 0
    Z = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    Z = Z - Z.T
    Z[:, 5] = 0; Z[5, :] = 0
    mats.append(Z)
    sizes.append(6)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as f:
            f.write(np.int32(len(mats)).tobytes())
            for M in mats:
                f.write(np.int32(M.shape[0]).tobytes())
                f.write(np.asfortranarray(M).astype("<c16").tobytes(order="F"))
        subprocess.run([os.path.join(REFBIN, "harness64.exe"), "pfaff", fin, os.path.join(td, "out.bin")],
                       check=True)
        raw = open(os.path.join(td, "out.bin"), "rb").read()
    rec = np.dtype([("pf", "<c16"), ("info", "<i4")])
    res = np.frombuffer(raw, dtype=rec)
    small = {("mat%d" % i): M for i, M in enumerate(mats) if M.shape[0] <= 64}
    return dict(seed=seed, sizes=np.array(sizes), pf=res["pf"].copy(), info=res["info"].copy(), **small)


def main():
    cases = {
        # sd: 7x7 angles, 3 iterations, FP64 TBMEs (oracle64) and stock FP32
        "sd64": lambda: synthetic_case("sd", 64, 7, 3, [0, 1, 24]),
        "sd32": lambda: synthetic_case("sd", 32, 7, 3, [0, 10]),
        # pf: odd number parity, Q20 constraint, 9x9, 2 iterations
        "pf64": lambda: synthetic_case("pf", 64, 9, 2, [0, 1, 40]),
        # HFB limit (one angle) and an even-M grid (Fomenko half-shift)
        "sd64_m1": lambda: synthetic_case("sd", 64, 1, 2, [0]),
        "sd64_m4": lambda: synthetic_case("sd", 64, 4, 1, [0, 5]),
        # reference's own example: Mg24, USDB, Q21 constraint, 5x5 angles
        "mg24_32": lambda: mg24_case(32, 5, 3, [0, 12]),
        # independent Q20-constrained solves of three basis sizes (9x9 angles,
        # 3 iterations each), the cases of the batched many-solve path.  The
        # beta2 values are ones whose constraint adjustment converges: where
        # adjust_constraints runs out its 100 steps the trajectory depends on
        # rounding and no implementation can reproduce it.
        "batch_sd": lambda: batch_case("sd", [-0.1, 0.0, 0.2]),
        "batch_pf": lambda: batch_case("pf", [-0.1, 0.0, 0.2]),
        "batch_sdpf": lambda: batch_case("sdpf", [-0.1, 0.0, 0.2]),
    }
    only = sys.argv[1:]
    for name, fn in cases.items():
        if only and name not in only:
            continue
        c = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **c)
        print("wrote", path, os.path.getsize(path))
    if not only or "pfaffian" in only:
        c = pfaffian_cases()
        path = os.path.join(HERE, "pfaffian.npz")
        np.savez_compressed(path, **c)
        print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
