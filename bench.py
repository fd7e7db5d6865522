#!/usr/bin/env python3
"""PNVAP gradient iterations/s on one MI355X.

Workload: e_max = 6 harmonic-oscillator basis (336 single-particle states),
15 x 15 Fomenko gauge angles, 10 protons / 12 neutrons, constraints <Z>, <N>,
Q20 = 0.2 and Jx = 0; seeded synthetic JT-coupled interaction
V ~ N(0, 2 MeV) exp(-(Na+Nb+Nc+Nd)/4) (12.9M stored m-scheme elements, the list
the reference builds from the same .sho file, bit for bit), SPEs
hbar*omega (N + 3/2).  The state and the constraint operators are those of the
reference's own first iteration on this deck (tests/golden/emax6.npz, written by
the reference after its adjust_constraints).  TBMEs stored in FP32, all
arithmetic FP64.

One step = one full gradient iteration of the reference's loop
(taurus_vap.f90:77-113) with the state resident in HBM: tg_iterate
(update_constraints_qpbasis, project_wavefunction over all 225 angles,
calculate_lagrange_multipliers, calculate_gradient, adjust_constraints with its
Cholesky evolution) with, first, construct_canonical_basis of the state (which
project_wavefunction does first in the reference): tg_iterate_canonical, whose
canonical basis runs on the second stream beside the projection's Thouless /
field / H20 chain.

Prints one JSON line.  --gpus N > 1 is not supported (one GPU per run).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PNVAP gradient iters/s (e_max=6, 15x15 angles) vs plain FP64 GPU, |dE|<1 keV"
PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
# Gauge integral over [0, 2 pi): the seed states carry proton-neutron mixing
# (their Thouless matrix has proton-neutron entries), so check_symmetries keeps
# the default proj_facpi = 2 pi (src/module_projection.f90:37,
# src/module_constraints.f90:749-753), as in the reference's own runs of them.
FACPI = 2.0 * math.pi


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_flops(ctx_info, n, A, nb):
    """Algorithmic FP64 flops per step (one projection of all A angles) of the
    timed kernels (real flops; complex multiply-add = 8, real x complex = 4)."""
    return {
        "kernel:grouped_ph": 4.0 * ctx_info["wph"] * A,
        "kernel:grouped_pp": 4.0 * ctx_info["wpp"] * 2 * A,
        "kernel:gj_inverse": 8.0 * n ** 3 * A,
        "kernel:pfaffian": (4.0 / 3.0) * nb ** 3 * A,
        # complex batched GEMMs of the loop body on the main stream: A_phi (n^3),
        # the two H20 products over the inner dimension 2n and the one over n
        "kernel:zgemm": 8.0 * A * 6 * n ** 3,
        # the reduced overlap's angle-independent parts of K on the second stream:
        # three real nb x nb x nb products per state (the per-angle combination is
        # element-wise)
        "kernel:k_parts": 3 * 2.0 * nb ** 3,
    }


def measured_traffic(workload, angles_per_launch, kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs recorded under profiles/ (scaled by the copy-kernel calibration stated
    there), when they were taken on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if int(t.get("angles_per_launch", -1)) != angles_per_launch:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else float(k["hbm_bytes_per_launch"])


def contraction_sizes(h):
    """sum of s_g^2 of the ph and pp block channels (as built by tg_init)."""
    from collections import Counter
    b = h["basis"]
    m, t, l = b.sp_2mj.astype(int), b.sp_2mt.astype(int), b.sp_l.astype(int)
    n = b.dim
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ph = Counter(zip((m[x] - m[y]).ravel(), (t[x] - t[y]).ravel(), ((l[x] + l[y]) & 1).ravel()))
    sel = x > y
    pp = Counter(zip((m[x] + m[y])[sel], (t[x] + t[y])[sel], ((l[x] + l[y]) & 1)[sel]))
    return dict(wph=float(sum(v * v for v in ph.values())), wpp=float(sum(v * v for v in pp.values())))


FIXTURE = os.path.join(ROOT, "tests", "golden", "emax6.npz")


def load_state():
    """The reference's first-iteration state of the e_max = 6 deck, its constraint
    operators and its FP64 / FP32 projected energies (data written by the
    reference; no reference code runs here)."""
    return dict(np.load(FIXTURE))


def payload_sha256(ham, bits):
    """SHA-256 of the element payload (0-based abcd, H2, trperm) in the reference's
    reduced-file layout, to compare with the reference-written file's."""
    import hashlib
    d = hashlib.sha256()
    d.update(np.ascontiguousarray(np.asarray(ham["abcd"]) - 1, dtype="<i2").tobytes())
    d.update(np.ascontiguousarray(ham["H2"].astype(np.float32) if bits == 32 else ham["H2"],
                                  dtype="<f4" if bits == 32 else "<f8").tobytes())
    d.update(np.ascontiguousarray(ham["trperm"], dtype="<i1").tobytes())
    return d.hexdigest()


def make_context(ham, fix, plain, fp32, graph=False, conj_pairs=False):
    from taurus_vap_amd import gpu, host
    b = ham["basis"]
    H2 = ham["H2"].astype(np.float32) if fp32 else ham["H2"]
    ctx = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, ham["H0"], ham["H1"],
                      ham["abcd"], H2, ham["trperm"], plain=plain, tbme_fp32=fp32, graph=graph,
                      conj_pairs=conj_pairs)
    ctx.set_gradient(float(fix["gradient_eta"]), float(fix["gradient_mu"]))
    ctx.set_constraints(np.asarray(fix["constraint_HO"]).reshape(-1), fix["constraint_val"],
                        int(fix["constraint_pair"]), 1e-6)
    U, V = np.asarray(fix["U0"]), np.asarray(fix["V0"])
    cb = host.construct_canonical_basis(U, V)
    M = int(fix["Mphip"])
    ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], M, M,
                  float(fix["facpi"]), float(fix["valence_Z"]), float(fix["valence_N"]),
                  float(fix["factor_delta"]))
    return ctx, cb


def next_step(ctx, fix):
    """One step from an evolved state: tg_iterate_canonical (its canonical basis,
    run beside the projection, then the iteration).  A degenerate kappa block
    wider than the device form takes the host route (counted): the host's
    construct_canonical_basis, tg_set_state, tg_iterate."""
    from taurus_vap_amd import gpu, host
    try:
        r, _ = ctx.iterate_canonical()
        return r, 0
    except gpu.TaurusGPUError as e:
        if "(-7)" not in str(e):
            raise
    U, V = ctx.get_state()
    cb = host.construct_canonical_basis(U, V)
    M = int(fix["Mphip"])
    ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], M, M,
                  float(fix["facpi"]), float(fix["valence_Z"]), float(fix["valence_N"]),
                  float(fix["factor_delta"]))
    return ctx.iterate(0.0), 1


def iterate_gpu(args, ham, fix, plain, fp32, steps, warmup, timing=True, graph=False,
                conj_pairs=False):
    """max(1, W) untimed iterations from the reference's state (the first on the
    state as given, with its canonical basis), then K timed steps, each the
    canonical basis of the evolved state and the next iteration.  Returns the
    wall time of the K steps, the projected energy of the first iteration (the
    fixture state) and the per-kernel timings of the timed steps."""
    ctx, cb = make_context(ham, fix, plain, fp32, graph=graph, conj_pairs=conj_pairs)
    timing = timing and not graph     # per-kernel events run the eager launch sequence
    host_route = 0
    first = ctx.iterate(0.0)
    for _ in range(max(1, warmup) - 1):
        host_route += next_step(ctx, fix)[1]
    ctx.sync()
    if timing:
        ctx.enable_timing(True)
        ctx.reset_timing()
    nadj, energies = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        r, h = next_step(ctx, fix)
        host_route += h
        nadj.append(r["adjust_iterations"])
        energies.append(float((r["ener"] / r["over"]).real))
    ctx.sync()
    t1 = time.perf_counter()
    kt = {}
    if timing:
        for key in ("total", "kernel:grouped_ph", "kernel:grouped_pp", "kernel:gj_inverse",
                    "kernel:pfaffian", "kernel:zgemm", "kernel:k_parts", "contraction", "gemm",
                    "gemm_h20", "overlap", "thouless_inverse", "canonical_basis_beside"):
            kt[key] = ctx.timing(key)
        ctx.enable_timing(False)
    ctx.close()
    n = ham["basis"].dim
    M = int(fix["Mphip"])
    return dict(seconds=t1 - t0, steps=steps, timing=kt,
                E_first=float((first["ener"] / first["over"]).real), first=first, n=n,
                A=((M * M + 1) // 2 if conj_pairs else M * M),
                nb=n - cb["nocc"] - cb["nemp"], adjust_steps=nadj, energies=energies,
                host_route_states=host_route)


def projection_only(args, ham, fix, steps, warmup):
    """The projection of the fixture state alone (tg_project_resident: all angles,
    sum_gauge, projected H20, gradient step), FP32 TBMEs: the loop body without
    the constraint machinery and the canonical basis."""
    ctx, _ = make_context(ham, fix, plain=False, fp32=True)
    for _ in range(warmup):
        ctx.project_resident()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.project_resident()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.close()
    return dict(value=steps / (t1 - t0), unit="projections/s", ms_per_step=1e3 * (t1 - t0) / steps)


def batched_solves(args):
    """Independent solves of uneven basis size in one launch per step: 24 sd
    (n = 24), 24 pf (n = 40) and 16 sdpf (n = 64) seeded states, 9x9 angles,
    one context per basis holding its solves, the three contexts one group
    replayed as a captured hipGraph.  A step = the projected gradient of every
    solve (same hot path as the main line).  Also the largest |dE| of the FP32-
    TBME run against the FP64-TBME run over the 64 solves."""
    from taurus_vap_amd import gpu, hamiltonian as H, host, inputs as I

    def build(fp32):
        ctxs, info = [], []
        for wname, count in I.BATCHED:
            w = I.WORKLOADS[wname]
            h = H.workload_hamiltonian(w, args.seed)
            b = h["basis"]
            H2 = h["H2"].astype(np.float32) if fp32 else h["H2"]
            c = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"],
                            h["abcd"], H2, h["trperm"], tbme_fp32=fp32)
            sts = []
            for k in range(count):
                U, V = host.seed_state(w, b, args.seed + 101 * k)
                sts.append((U, V, host.construct_canonical_basis(U, V)))
            c.set_gradient(args.eta, args.mu)
            c.set_states(sts, 9, 9, FACPI, float(w.Z), float(w.N))
            ctxs.append(c)
            info.append((wname, b.dim, count))
        return ctxs, info

    def energies(ctxs):
        out = []
        for c in ctxs:
            for k in range(c.nstates):
                r = c.get_results(k)
                out.append((r["ener"] / r["over"]).real)
        return np.array(out)

    ctxs, info = build(True)
    grp = gpu.Group(ctxs, graph=True)
    for _ in range(3):          # eager, captured, replayed
        grp.project()
    ctxs[0].sync()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        grp.project()
    for c in ctxs:
        c.sync()
    t1 = time.perf_counter()
    E32 = energies(ctxs)
    grp.close()
    for c in ctxs:
        c.close()
    ctxs64, _ = build(False)
    grp = gpu.Group(ctxs64, graph=False)
    grp.project()
    for c in ctxs64:
        c.sync()
    E64 = energies(ctxs64)
    grp.close()
    for c in ctxs64:
        c.close()
    nsol = sum(cnt for _, _, cnt in info)
    sec = (t1 - t0) / args.batch_steps
    return {"solves": nsol,
            "bases": " + ".join("%d x %s (n=%d)" % (cnt, wn, nd) for wn, nd, cnt in info),
            "angles": "9x9", "graph": True,
            "value": nsol / sec, "unit": "solve-iterations/s", "ms_per_step": 1e3 * sec,
            "max_abs_dE_fp32_vs_fp64_keV": float(np.max(np.abs(E32 - E64))) * 1e3}


def cpu_baseline(args, ham, fix, threads):
    """The reference TAURUS_vap built from its own sources (oracle/_ref, FP64 TBMEs,
    OpenMP + threaded OpenBLAS) on the same deck, reduced Hamiltonian file and
    state: oracle/ref_harness.f90 ('timefull' mode) times one full iteration of
    the reference's loop -- project_wavefunction over the 15 x 15 angles (with
    its construct_canonical_basis), calculate_lagrange_multipliers,
    calculate_gradient and adjust_constraints.  The reference's set-up (reading
    the 220 MB reduced file) is not timed."""
    import resource
    exe = os.path.join(ROOT, "oracle", "_ref", "harness64.exe")
    if not os.path.exists(exe):
        log("reference binary not built: cpu_baseline = null")
        return None
    from taurus_vap_amd import inputs as I, hamiltonian as H
    w = I.WORKLOADS[args.workload]

    def big_stack():
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    with tempfile.TemporaryDirectory() as td:
        I.write_workload(w, td, seed=args.seed, hamil_read=1, iters=1,
                         eta=float(fix["gradient_eta"]), mu=float(fix["gradient_mu"]),
                         grad_eps=1e-14, UV=(np.asarray(fix["U0"]), np.asarray(fix["V0"])))
        H.write_red(os.path.join(td, w.name + ".red"), "synthetic %s seed %d" % (w.name, args.seed),
                    1, ham["H0"], ham["H1"], ham["abcd"], ham["H2"], ham["trperm"], rh2_bytes=8)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
                   OMP_STACKSIZE="256M")
        t0 = time.perf_counter()
        with open(os.path.join(td, "input.txt")) as fin, open(os.path.join(td, "out.txt"), "w") as fout:
            r = subprocess.run([exe, "timefull", "times.txt", str(args.cpu_iterations)], stdin=fin,
                               stdout=fout, stderr=subprocess.STDOUT, cwd=td, env=env,
                               timeout=1500, preexec_fn=big_stack)
        log("reference harness run: %.1f s (rc %d)" % (time.perf_counter() - t0, r.returncode))
        if r.returncode != 0:
            return None
        rows = np.loadtxt(os.path.join(td, "times.txt"), ndmin=2)
    per = rows[:, 2] + rows[:, 3] + rows[:, 4]
    t_iter = float(np.mean(per))
    return dict(value=1.0 / t_iter, unit="iterations/s", cores=threads, kind="reference",
                sample="%d full iteration(s) of the reference's loop on the same deck, state and "
                       "Hamiltonian, timed inside the reference build: project_wavefunction over "
                       "%d angles %.2f s, multipliers + gradient %.3f s, adjust_constraints %.3f s "
                       "(mean per iteration); set-up not timed"
                       % (len(rows), int(rows[0, 1]), float(np.mean(rows[:, 2])),
                          float(np.mean(rows[:, 3])), float(np.mean(rows[:, 4]))),
                E_first=float(rows[0, 5]))


def copy_probe(mib: int):
    """A plain copy of known size: reads and writes mib MiB per launch."""
    import torch
    nel = mib * (1 << 20) // 8
    src = torch.ones(nel, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy probe: 3 copies of %d MiB (%d bytes read + %d written per launch)"
        % (mib, nel * 8, nel * 8))


def dgemm_probe(n: int):
    """The vendor FP64 GEMM (torch.matmul -> rocBLAS / hipBLASLt) on two n x n
    matrices: the FP64 matrix-core rate a library GEMM sustains on this part."""
    import torch
    a = torch.rand(n, n, dtype=torch.float64, device="cuda")
    b = torch.rand(n, n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    for _ in range(3):
        torch.matmul(a, b, out=c)
    torch.cuda.synchronize()
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.matmul(a, b, out=c)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    log("dgemm probe: n = %d, %.3f ms per product, %.2f TFLOP/s (FP64)"
        % (n, ms, 2.0 * n ** 3 / ms / 1e9))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="emax6")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tbme", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed iterations of the plain FP64 per-angle path (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=1,
                    help="full reference iterations timed for cpu_baseline")
    ap.add_argument("--conj-pairs", action="store_true",
                    help="compute one gauge angle of each (phi, -phi) pair (exact for real U, V)")
    ap.add_argument("--pair-steps", type=int, default=5,
                    help="timed iterations of the conjugate-pair variant reported beside value (0: skip)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the projection as a captured hipGraph (no per-kernel timing)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--batch-steps", type=int, default=5,
                    help="timed steps of the 64 batched independent solves (0: skip)")
    ap.add_argument("--proj-steps", type=int, default=5,
                    help="timed projections alone, without the constraint machinery (0: skip)")
    ap.add_argument("--eta", type=float, default=0.005, help="step of the batched solves")
    ap.add_argument("--mu", type=float, default=0.3, help="momentum of the batched solves")
    ap.add_argument("--dgemm-probe", type=int, default=0, metavar="N",
                    help="time the vendor FP64 GEMM on N x N matrices and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibration "
                         "of the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the batched gauge-angle path is measured on a "
              "single MI355X (--gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.dgemm_probe:
        dgemm_probe(args.dgemm_probe)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    from taurus_vap_amd import hamiltonian as H, inputs as I
    w = I.WORKLOADS[args.workload]
    fix = load_state()
    t = time.perf_counter()
    ham = H.workload_hamiltonian(w, args.seed)
    same_list = all(payload_sha256(ham, b) == str(fix["red_sha256_%d" % b]) for b in (64, 32))
    log("hamiltonian %s: n=%d, %d stored TBMEs (%.1f s), list %s the reference's"
        % (args.workload, ham["basis"].dim, len(ham["trperm"]), time.perf_counter() - t,
           "equal to" if same_list else "DIFFERENT FROM"))
    sizes = contraction_sizes(ham)
    fp32 = args.tbme == "fp32"

    main_run = iterate_gpu(args, ham, fix, plain=False, fp32=fp32, steps=args.steps,
                           warmup=args.warmup, graph=args.graph, conj_pairs=args.conj_pairs)
    ips = main_run["steps"] / main_run["seconds"]
    E64_ref = float((fix["pnp_ener"] / fix["pnp_over"]).real)
    E32_ref = float((fix["fp32_pnp_ener"] / fix["fp32_pnp_over"]).real)
    dE_keV = abs(main_run["E_first"] - E64_ref) * 1e3
    cb_ms, cb_cnt = main_run["timing"].get("canonical_basis_beside", (0.0, 0))
    log("iterations: %.3f it/s (%.2f ms per step; canonical basis %.2f ms of it on the second "
        "stream in %d of %d steps), first E = %.9f MeV (reference FP64 %.9f, |dE| = %.3e keV), "
        "adjustment steps %s"
        % (ips, 1e3 / ips, cb_ms / max(1, cb_cnt), cb_cnt, main_run["steps"], main_run["E_first"],
           E64_ref, dE_keV, main_run["adjust_steps"]))
    for k, (ms, c) in main_run["timing"].items():
        if c:
            log("  %-22s %9.3f ms/step  (%d intervals)" % (k, ms / main_run["steps"], c))

    # per-kernel rooflines (in-run HIP events on the stream each kernel ran on);
    # the dominant kernel is the one with the largest summed time per step, on
    # either stream
    n, A, nb = main_run["n"], main_run["A"], main_run["nb"]
    M = int(fix["Mphip"])
    fl = algorithmic_flops(sizes, n, A, nb)
    kern = {k: v for k, v in main_run["timing"].items() if k.startswith("kernel:") and v[1] > 0}
    roofline = None
    by_kernel = {}
    for k, (ms, cnt) in kern.items():
        avg_s = ms / cnt / 1e3
        per_step = cnt / main_run["steps"]
        flop_launch = fl[k] / per_step
        ach = flop_launch / avg_s / 1e12
        name = k.split(":", 1)[1]
        by_kernel[name] = {"ms_per_step": ms / main_run["steps"],
                           "launches_per_step": per_step,
                           "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop_launch,
                           "achieved": ach, "frac": ach / PEAK_FP64_TFLOPS,
                           "traffic": measured_traffic(args.workload, A, name)}
    step_s = main_run["seconds"] / main_run["steps"]
    proj_flop = sum(fl.values())
    if kern:
        dom = max(kern, key=lambda k: kern[k][0])
        d = by_kernel[dom.split(":", 1)[1]]
        roofline = {"bound": "mfma", "kernel": dom.split(":", 1)[1], "achieved": d["achieved"],
                    "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": d["frac"],
                    "traffic": d["traffic"], "flop_per_launch": d["flop_per_launch"],
                    "avg_launch_ms": d["avg_launch_ms"],
                    "whole_step": {"flop_per_step": proj_flop, "ms_per_step": 1e3 * step_s,
                                   "achieved": proj_flop / step_s / 1e12,
                                   "frac": proj_flop / step_s / 1e12 / PEAK_FP64_TFLOPS,
                                   "counted": "the projection's algorithmic flops only "
                                              "(constraint machinery and canonical basis "
                                              "not counted)"}}

    # plain FP64 per-angle drop-in (the metric's denominator), same iteration
    plain = None
    if args.plain_steps > 0:
        pr = iterate_gpu(args, ham, fix, plain=True, fp32=False, steps=args.plain_steps, warmup=0,
                         timing=False)
        plain = dict(value=pr["steps"] / pr["seconds"], unit="iterations/s",
                     E_first=pr["E_first"],
                     dE_vs_oracle64_keV=abs(pr["E_first"] - E64_ref) * 1e3)
        log("plain FP64: %.4f it/s, first E = %.9f MeV" % (plain["value"], pr["E_first"]))

    pairs = None
    if not args.conj_pairs and args.pair_steps > 0:
        cp = iterate_gpu(args, ham, fix, plain=False, fp32=fp32, steps=args.pair_steps,
                         warmup=args.warmup, timing=False, conj_pairs=True)
        pairs = dict(value=cp["steps"] / cp["seconds"], unit="iterations/s",
                     ms_per_step=1e3 * cp["seconds"] / cp["steps"], angles_per_launch=cp["A"],
                     dE_vs_all_angles_keV=abs(cp["E_first"] - main_run["E_first"]) * 1e3)
        log("conjugate pairs: %.3f it/s" % pairs["value"])

    proj = None
    if args.proj_steps > 0:
        proj = projection_only(args, ham, fix, args.proj_steps, args.warmup)
        log("projection alone: %.3f projections/s" % proj["value"])

    batched = None
    if args.batch_steps > 0:
        try:
            batched = batched_solves(args)
            log("batched solves: %.1f solve-iterations/s (%.3f ms per step), |dE| fp32 vs fp64 "
                "<= %.3e keV" % (batched["value"], batched["ms_per_step"],
                                 batched["max_abs_dE_fp32_vs_fp64_keV"]))
        except Exception as e:      # a failed side measurement must not hide the main one
            log("batched solves failed: %r" % (e,))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args, ham, fix, args.cpu_threads)
        except Exception as e:      # a failed baseline must not hide the measurement
            log("cpu baseline failed: %r" % (e,))

    out = {
        "metric": METRIC,
        "value": ips,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%s: HO basis of %d sp states, %dx%d gauge angles, %dp/%dn, constraints "
                        "<Z>, <N>, Q20 = 0.2, Jx = 0, seeded synthetic JT interaction (%d stored "
                        "m-scheme TBMEs, TBMEs %s / FP64 arithmetic), state = the reference's "
                        "first iteration; step = one full gradient iteration "
                        "(tg_iterate_canonical: the device canonical basis of the state beside "
                        "the projection, then multipliers, gradient, adjust_constraints), state "
                        "resident in HBM" % (
                            args.workload, n, M, M, w.Z, w.N, len(ham["trperm"]),
                            args.tbme.upper()),
            "mode": "batched" + ("+graph" if args.graph else "") +
                    ("+conj_pairs" if args.conj_pairs else ""),
            "parallelism": "single GPU, gauge angles batched",
        },
        "roofline": roofline,
        "roofline_by_kernel": by_kernel,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain_fp64_gpu": None if plain is None else ips / plain["value"],
        "dE_vs_oracle64_keV": dE_keV,
        "E_first_MeV": main_run["E_first"],
        "E_oracle64_MeV": E64_ref,
        "E_oracle32_MeV": E32_ref,
        "hamiltonian_equals_reference": same_list,
        "iteration_split_ms": {"step": 1e3 / ips,
                               "canonical_basis_beside": cb_ms / max(1, cb_cnt),
                               "steps_with_canonical_beside": cb_cnt,
                               "adjust_steps": main_run["adjust_steps"],
                               "host_route_states": main_run["host_route_states"]},
        "projection_only": proj,
        "conj_pairs": pairs,
        "batched_solves": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
