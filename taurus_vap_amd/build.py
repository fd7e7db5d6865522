"""Builds the gfx950 shared library libtaurus_gpu.so in-tree with hipcc.

    python -m taurus_vap_amd.build        (or build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtaurus_gpu.so")
SOURCES = ["tg_gemm.hip", "tg_mfma.hip", "tg_linalg.hip", "tg_fields.hip", "tg_project.hip",
           "tg_iterate.hip", "tg_canon.hip", "tg_a19.hip", "tg_noalloc.hip", "tg_eig.hip"]
ARCH = os.environ.get("TAURUS_GPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = False) -> str:
    objs = []
    flags = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=%s" % ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        out = p.communicate()[0].decode(errors="replace")
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l)
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write("---- %s\n%s\n" % (src, out))
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    # the HIP runtime is the only library linked
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=%s" % ARCH, *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
