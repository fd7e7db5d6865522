#!/usr/bin/env python3
"""PNVAP gradient iterations/s on one MI355X.

Workload (default): e_max = 6 harmonic-oscillator basis (336 single-particle
states), 15 x 15 Fomenko gauge angles, 10 protons / 12 neutrons, seeded
synthetic JT-coupled interaction V ~ N(0, 2 MeV) exp(-(Na+Nb+Nc+Nd)/4)
(12.9M stored m-scheme elements), SPEs hbar*omega (N + 3/2), seed state a real
Thouless rotation (Z ~ N(0, 0.01)) of the lowest-filled Slater determinant,
brought to <Z> = 10, <N> = 12.
TBMEs stored in FP32, all arithmetic FP64.

One step = one projected-gradient iteration of the hot path with the state
resident in HBM: the gauge-angle loop of project_wavefunction over all 225
angles (Thouless matrices, transition densities, TBME contraction, overlaps,
energies, H20), sum_gauge, calculate_projected_H20 and the gradient step.

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
# timing keys of kernels launched on the second (overlap) stream
SECOND_STREAM_KERNELS = ("kernel:pfaffian", "kernel:zgemm_q")
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
        # the reduced-overlap product Q^T (J Q) on the second stream: its upper
        # row bands (the skew result's upper triangle is what the Pfaffian reads)
        "kernel:zgemm_q": 8.0 * A * nb * sum(min(rb, nb - r0) * (nb - r0)
                                             for rb in [q_band_rows(nb)]
                                             for r0 in range(0, nb, rb)),
    }


def q_band_rows(nb):
    """Row band height of the reduced-overlap product (as the library's)."""
    return (((nb + 2) // 3 + 15) // 16) * 16 if nb >= 128 else nb


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


def run_gpu(args, ham, plain, fp32, steps, warmup, timing=True, graph=False, conj_pairs=False):
    from taurus_vap_amd import host
    ctx, h, w, U, V = host.workload_context(args.workload, seed=args.seed, tbme_fp32=fp32,
                                            plain=plain, ham=ham, graph=graph,
                                            conj_pairs=conj_pairs)
    timing = timing and not graph     # per-kernel events run the eager launch sequence
    cb = host.construct_canonical_basis(U, V)
    n = h["basis"].dim
    M = args.angles if args.angles else w.M
    ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], M, M,
                  FACPI, float(w.Z), float(w.N))
    # gradient step with two constraint arrays (the shape of Q20 + Jx)
    rng = np.random.default_rng(args.seed + 7)
    c20 = rng.normal(size=(2, n * n)) * 0.01
    ctx.set_gradient(args.eta, args.mu, c20, [0.05, -0.02])
    for _ in range(warmup):
        ctx.project_resident()
    ctx.sync()
    if timing:
        ctx.enable_timing(True)
        ctx.reset_timing()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.project_resident()
    ctx.sync()
    t1 = time.perf_counter()
    res = ctx.get_results()
    Zi, gnorm = ctx.get_gradient()
    kt = {}
    if timing:
        for key in ("total", "kernel:grouped_ph", "kernel:grouped_pp", "kernel:gj_inverse",
                    "kernel:pfaffian", "kernel:zgemm", "kernel:zgemm_q", "contraction", "gemm",
                    "gemm_h20",
                    "overlap", "thouless_inverse"):
            kt[key] = ctx.timing(key)
        ctx.enable_timing(False)
    boundary = None
    if timing:
        # the host-buffer boundary as the Fortran drop-in calls it: canonical basis
        # on the host, then tg_project_wavefunction (upload, all angles, sum_gauge,
        # projected H20, download of pnp_* and field_H20v); not part of `value`
        tc0 = time.perf_counter()
        cb2 = host.construct_canonical_basis(U, V)
        tc1 = time.perf_counter()
        for _ in range(2):
            ctx.project_wavefunction(U, V, cb2["D"], cb2["nocc"], cb2["nemp"], cb2["ovac"],
                                     cb2["voveru"], M, M, FACPI, float(w.Z), float(w.N))
        tc2 = time.perf_counter()
        boundary = dict(canonical_basis_ms=1e3 * (tc1 - tc0), host_call_ms=1e3 * (tc2 - tc1) / 2)
    ctx.close()
    E = (res["ener"] / res["over"]).real
    nbloc = n - cb["nocc"] - cb["nemp"]
    return dict(seconds=t1 - t0, steps=steps, E=E, Z=(res["prot"] / res["over"]).real,
                N=(res["neut"] / res["over"]).real, gnorm=gnorm, timing=kt, n=n,
                A=((M * M + 1) // 2 if conj_pairs else M * M),   # angles per launch
                nb=nbloc, nocc=cb["nocc"], nemp=cb["nemp"], UV=(U, V), boundary=boundary)


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


def device_iteration(args, ham):
    """Full gradient iterations of the workload on the device: tg_iterate
    (constraint_20, projection, Lagrange multipliers, gradient, adjust_constraints)
    then tg_canonical_basis of each new state, which never leaves the device,
    with <Z>, <N> constrained.  A state the device canonical basis refuses
    (degenerate kappa subspace wider than one pair) takes the host route and is
    counted.  Not part of `value`."""
    from taurus_vap_amd import host, gpu
    ctx, h, w, U, V = host.workload_context(args.workload, seed=args.seed, tbme_fp32=True, ham=ham)
    M = args.angles if args.angles else w.M
    Zop, Nop = host.number_operators(h["basis"])
    ctx.set_gradient(args.eta, 0.0)
    ctx.set_constraints(np.stack([Zop.T, Nop.T]).reshape(-1), [float(w.Z), float(w.N)], 0, 1e-6)
    cb = host.construct_canonical_basis(U, V)
    ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], M, M,
                  FACPI, float(w.Z), float(w.N))
    it_ms, cb_ms, nadj, host_route = [], [], [], 0
    for it in range(args.iterate_steps + 1):
        t0 = time.perf_counter()
        r = ctx.iterate(0.0)
        t1 = time.perf_counter()
        try:
            ctx.canonical_basis()
        except gpu.TaurusGPUError as e:
            if "(-7)" not in str(e):
                raise
            U, V = ctx.get_state()
            cb = host.construct_canonical_basis(U, V)
            ctx.set_state(U, V, cb["D"], cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], M, M,
                          FACPI, float(w.Z), float(w.N))
            host_route += 1
        ctx.sync()
        t2 = time.perf_counter()
        if it > 0:                  # the first call allocates and tunes
            it_ms.append(1e3 * (t1 - t0))
            cb_ms.append(1e3 * (t2 - t1))
            nadj.append(r["adjust_iterations"])
    ctx.close()
    per = (np.mean(it_ms) + np.mean(cb_ms)) / 1e3
    return {"value": 1.0 / per, "unit": "iterations/s", "iterations": args.iterate_steps,
            "iterate_ms": float(np.mean(it_ms)), "canonical_basis_ms": float(np.mean(cb_ms)),
            "canonical_basis": "device", "host_route_states": host_route,
            "adjust_steps": nadj, "constraints": "<Z>, <N>"}


def cpu_baseline(args, ham, threads, UV):
    """Reference TAURUS_vap built from its own sources (oracle/_ref, FP64 TBMEs,
    OpenMP + threaded OpenBLAS) on the same deck, .red file and state: inside one
    run of oracle/ref_harness.f90 ('time' mode) the reference's
    project_wavefunction(0,...) is timed at 1x2 and 3x3 gauge angles (best of 3),
    plus its gradient step; the iteration time at 15x15 angles is the linear fit
    in the angle count.  The reference's own set-up (reading the 220 MB .red file,
    ~1 min) is not timed."""
    import resource
    exe = os.path.join(ROOT, "oracle", "_ref", "harness64.exe")
    if not os.path.exists(exe):
        log("reference binary not built: cpu_baseline = null")
        return None
    from taurus_vap_amd import inputs as I, hamiltonian as H
    w = I.WORKLOADS[args.workload]
    A = (args.angles or w.M) ** 2

    def big_stack():
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    with tempfile.TemporaryDirectory() as td:
        I.write_workload(w, td, seed=args.seed, hamil_read=1, iters=1, eta=args.eta, mu=args.mu,
                         grad_eps=1e-14, UV=UV)
        H.write_red(os.path.join(td, w.name + ".red"), "synthetic %s seed %d" % (w.name, args.seed),
                    1, ham["H0"], ham["H1"], ham["abcd"], ham["H2"], ham["trperm"], rh2_bytes=8)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
                   OMP_STACKSIZE="256M")
        t0 = time.perf_counter()
        with open(os.path.join(td, "input.txt")) as fin, open(os.path.join(td, "out.txt"), "w") as fout:
            r = subprocess.run([exe, "time", "times.txt", "3"], stdin=fin, stdout=fout,
                               stderr=subprocess.STDOUT, cwd=td, env=env, timeout=1500,
                               preexec_fn=big_stack)
        log("reference harness run: %.1f s (rc %d)" % (time.perf_counter() - t0, r.returncode))
        if r.returncode != 0:
            return None
        rows = np.loadtxt(os.path.join(td, "times.txt"), ndmin=2)
    best = {}
    for k, a, tp, tg, e in rows:
        a = int(a)
        best[a] = min(best.get(a, (1e30, 1e30))[0], tp), min(best.get(a, (1e30, 1e30))[1], tg)
    (t2, g2), (t9, g9) = best[2], best[9]
    per_angle = (t9 - t2) / 7.0
    fixed = t2 - 2.0 * per_angle
    if per_angle <= 0.0 or fixed < 0.0:
        # timing noise larger than the spread: no fixed part, slope of the larger sample
        per_angle, fixed = t9 / 9.0, 0.0
    t_iter = fixed + A * per_angle + min(g2, g9)
    return dict(value=1.0 / t_iter, unit="iterations/s", cores=threads, kind="reference",
                sample="reference project_wavefunction timed at 1x2 (%.2f s) and 3x3 (%.2f s) gauge "
                       "angles + gradient step (%.3f s), best of 3, linear fit %.3f s + %.3f s/angle "
                       "extrapolated to %d angles" % (t2, t9, min(g2, g9), fixed, per_angle, A))


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
    ap.add_argument("--angles", type=int, default=0, help="gauge angles per species (default: workload)")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--eta", type=float, default=0.005)
    ap.add_argument("--mu", type=float, default=0.3)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--tbme", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed steps of the plain FP64 per-angle path (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--conj-pairs", action="store_true",
                    help="compute one gauge angle of each (phi, -phi) pair (exact for real U, V)")
    ap.add_argument("--pair-steps", type=int, default=5,
                    help="timed steps of the conjugate-pair variant reported beside value (0: skip)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the projection as a captured hipGraph (no per-kernel timing)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--batch-steps", type=int, default=5,
                    help="timed steps of the 64 batched independent solves (0: skip)")
    ap.add_argument("--iterate-steps", type=int, default=3,
                    help="timed full device iterations with the constraint machinery (0: skip)")
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
    t = time.perf_counter()
    ham = H.workload_hamiltonian(w, args.seed)
    log("hamiltonian %s: n=%d, %d stored TBMEs (%.1f s)" % (args.workload, ham["basis"].dim,
                                                           len(ham["trperm"]), time.perf_counter() - t))
    sizes = contraction_sizes(ham)

    main_run = run_gpu(args, ham, plain=(args.mode == "plain"), fp32=(args.tbme == "fp32"),
                       steps=args.steps, warmup=args.warmup, graph=args.graph,
                       conj_pairs=args.conj_pairs)
    ips = main_run["steps"] / main_run["seconds"]
    log("optimized: %.3f it/s, E = %.9f MeV, <Z> = %.6f, <N> = %.6f, |grad| = %.4e"
        % (ips, main_run["E"], main_run["Z"], main_run["N"], main_run["gnorm"]))
    for k, (ms, c) in main_run["timing"].items():
        if c:
            log("  %-22s %9.3f ms/step  (%d intervals)" % (k, ms / main_run["steps"], c))

    # dominant timed kernel and its roofline
    n, A, nb = main_run["n"], main_run["A"], main_run["nb"]
    M = args.angles if args.angles else w.M
    fl = algorithmic_flops(sizes, n, A, nb)
    # every timed kernel (in-run HIP events on the stream each launch ran on);
    # the dominant one is the kernel with the largest summed time per step on the
    # context stream: the second stream's kernels (overlap chain) share the CUs
    # with the context stream's, so their event times are not their own
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
    if kern:
        ctx_kern = [k for k in kern if k not in SECOND_STREAM_KERNELS] or list(kern)
        dom = max(ctx_kern, key=lambda k: kern[k][0])
        d = by_kernel[dom.split(":", 1)[1]]
        roofline = {"bound": "mfma", "kernel": dom.split(":", 1)[1], "achieved": d["achieved"],
                    "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": d["frac"],
                    "traffic": d["traffic"], "flop_per_launch": d["flop_per_launch"],
                    "avg_launch_ms": d["avg_launch_ms"]}

    # plain FP64 per-angle drop-in (denominator of the metric) and |dE| vs FP64
    plain = None
    if args.plain_steps > 0:
        pr = run_gpu(args, ham, plain=True, fp32=False, steps=args.plain_steps, warmup=0, timing=False)
        plain = dict(value=pr["steps"] / pr["seconds"], E=pr["E"])
        log("plain FP64: %.4f it/s, E = %.9f MeV" % (plain["value"], pr["E"]))
    else:
        pr = run_gpu(args, ham, plain=False, fp32=False, steps=1, warmup=0, timing=False)
        plain = dict(value=None, E=pr["E"])
    dE_keV = abs(main_run["E"] - plain["E"]) * 1e3

    # the same iteration computing one angle of each (phi, -phi) pair (exact for
    # real U, V: the other angle's quantities are complex conjugates)
    pairs = None
    if not args.conj_pairs and args.pair_steps > 0:
        cp = run_gpu(args, ham, plain=(args.mode == "plain"), fp32=(args.tbme == "fp32"),
                     steps=args.pair_steps, warmup=args.warmup, timing=False, conj_pairs=True)
        pairs = dict(value=cp["steps"] / cp["seconds"], unit="iterations/s",
                     ms_per_step=1e3 * cp["seconds"] / cp["steps"], angles_per_launch=cp["A"],
                     dE_vs_all_angles_keV=abs(cp["E"] - main_run["E"]) * 1e3)
        log("conjugate pairs: %.3f it/s, E = %.9f MeV" % (pairs["value"], cp["E"]))

    batched = None
    if args.batch_steps > 0:
        try:
            batched = batched_solves(args)
            log("batched solves: %.1f solve-iterations/s (%.3f ms per step), |dE| fp32 vs fp64 "
                "<= %.3e keV" % (batched["value"], batched["ms_per_step"],
                                 batched["max_abs_dE_fp32_vs_fp64_keV"]))
        except Exception as e:      # a failed side measurement must not hide the main one
            log("batched solves failed: %r" % (e,))
    devit = None
    if args.iterate_steps > 0:
        try:
            devit = device_iteration(args, ham)
            log("device iteration: %.3f it/s (tg_iterate %.2f ms, device canonical basis %.2f ms, "
                "%d states on the host route)" % (devit["value"], devit["iterate_ms"],
                                                  devit["canonical_basis_ms"],
                                                  devit["host_route_states"]))
        except Exception as e:
            log("device iteration failed: %r" % (e,))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args, ham, args.cpu_threads, main_run["UV"])
        except Exception as e:      # a failed baseline must not hide the measurement
            log("cpu baseline failed: %r" % (e,))

    out = {
        "metric": METRIC,
        "value": ips,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * main_run["seconds"] / main_run["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%s: HO basis of %d sp states, %dx%d gauge angles, %dp/%dn, seeded "
                        "synthetic JT interaction (%d stored m-scheme TBMEs), TBMEs %s / FP64 "
                        "arithmetic, state resident in HBM" % (
                            args.workload, n, M, M, w.Z, w.N, len(ham["trperm"]),
                            args.tbme.upper()),
            "mode": args.mode + ("+graph" if args.graph else "") +
                    ("+conj_pairs" if args.conj_pairs else ""),
            "parallelism": "single GPU, gauge angles batched",
        },
        "roofline": roofline,
        "roofline_by_kernel": by_kernel,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": None if plain["value"] is None else {"value": plain["value"],
                                                               "unit": "iterations/s"},
        "speedup_vs_plain_fp64_gpu": None if plain["value"] is None else ips / plain["value"],
        "host_boundary": main_run["boundary"],
        "dE_vs_fp64_keV": dE_keV,
        "E_proj_MeV": main_run["E"],
        "conj_pairs": pairs,
        "batched_solves": batched,
        "device_iteration": devit,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
