"""Two-body Hamiltonian in the reference's reduced format.

* ``jt_table`` -- the JT-scheme array hamil_H2cpd(0:3,0:jmax,sh,sh,sh,sh) of an
  ANTOINE interaction after read_hamiltonian_2body_antoine and the symmetry
  reconstruction (src/module_hamiltonian.f90:409-456, 284-319);
* ``mscheme`` -- the stored m-scheme list (abcd, H2, trperm) through the native
  producer libtaurus_hamil.so (include/taurus_hamil.h);
* ``write_red`` / ``read_red`` -- the reduced binary file of
  write_hamiltonian_reduced / read_hamiltonian_reduced (:533-627).
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np

from . import inputs as I

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtaurus_hamil.so")
_lib = None


def build_producer() -> str:
    import subprocess
    src = os.path.join(HERE, "csrc", "hamil_producer.cpp")
    tmp = LIBPATH + ".tmp"
    subprocess.run(["g++", "-O3", "-fPIC", "-shared", "-o", tmp, src], check=True)
    os.replace(tmp, LIBPATH)
    return LIBPATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libtaurus_hamil.so not built (taurus_vap_amd.build)")
        lib = ctypes.CDLL(LIBPATH)
        P, I_, D, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        lib.th_mscheme.argtypes = [I_, P, P, P, P, P, I_, I_, P, D, L, P, P]
        lib.th_mscheme.restype = L
        lib.th_timerev.argtypes = [I_, P, P, L, P, P]
        lib.th_clebsch_gordan.argtypes = [I_] * 6
        lib.th_clebsch_gordan.restype = D
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def jt_table(shells, tbme) -> tuple[np.ndarray, int]:
    """hamil_H2cpd for a type-1 ANTOINE interaction given by its representatives
    (a,b,c,d,T,J,v) (1-based shells, T = 0/1): T=0 fills index 0, T=1 indices
    1,2,3 (all M_T), then the 7 permutations with the reference's phases."""
    nsh = len(shells)
    jmax = max(s[2] for s in shells)
    V = np.zeros((4, jmax + 1, nsh, nsh, nsh, nsh))
    for (a, b, c, d, T, J, v) in tbme:
        for mmt in ((0,) if T == 0 else (1, 2, 3)):
            V[mmt, J, a - 1, b - 1, c - 1, d - 1] = v
    # symmetry reconstruction: every nonzero entry spreads to its 7 images
    # (done on the representatives, whose orbits are disjoint)
    mmt, J, a, b, c, d = np.nonzero(np.abs(V) > 1e-12)
    v = V[mmt, J, a, b, c, d]
    two_j = np.array([s[2] for s in shells])
    xt = (mmt != 0).astype(np.int64)
    pab = np.where(((two_j[a] + two_j[b]) // 2 + J + xt) % 2 == 0, 1.0, -1.0)
    pcd = np.where(((two_j[c] + two_j[d]) // 2 + J + xt) % 2 == 0, 1.0, -1.0)
    V[mmt, J, a, b, d, c] = v * pcd
    V[mmt, J, b, a, c, d] = v * pab
    V[mmt, J, b, a, d, c] = v * pab * pcd
    V[mmt, J, c, d, a, b] = v
    V[mmt, J, c, d, b, a] = v * pab
    V[mmt, J, d, c, a, b] = v * pcd
    V[mmt, J, d, c, b, a] = v * pab * pcd
    return V, jmax


def mscheme(basis: I.Basis, V: np.ndarray, jmax: int, scaling_2b: float = 1.0):
    """Stored m-scheme elements (abcd 1-based int16 [ntb,4], H2 float64, trperm int8)."""
    lib = _load()
    n = basis.dim
    nsh = V.shape[2]
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in
            (basis.sp_2j, basis.sp_2mj, basis.sp_2mt, basis.sp_l, basis.sp_sh)]
    Vf = np.asfortranarray(V).ravel(order="F")
    ub = lib.th_mscheme(n, *[_p(x) for x in arrs], nsh, jmax, None, 1.0, 0, None, None)
    if ub < 0:
        raise RuntimeError("th_mscheme: bad arguments")
    abcd = np.zeros((max(ub, 1), 4), dtype=np.int16)
    H2 = np.zeros(max(ub, 1))
    cnt = lib.th_mscheme(n, *[_p(x) for x in arrs], nsh, jmax, _p(Vf), scaling_2b, ub, _p(abcd), _p(H2))
    if cnt < 0:
        raise RuntimeError("th_mscheme failed (%d)" % cnt)
    abcd = np.ascontiguousarray(abcd[:cnt])
    H2 = H2[:cnt].copy()
    trperm = np.zeros(max(cnt, 1), dtype=np.int8)
    sp_2j = np.ascontiguousarray(basis.sp_2j, dtype=np.int32)
    sp_tr = np.ascontiguousarray(basis.sp_tr, dtype=np.int32)
    if lib.th_timerev(n, _p(sp_2j), _p(sp_tr), cnt, _p(abcd), _p(trperm)) != 0:
        raise RuntimeError("th_timerev failed")
    return abcd, H2, trperm[:cnt].copy()


def one_body_antoine(basis: I.Basis, spe) -> np.ndarray:
    """hamil_H1 of a type-1 ANTOINE file: diagonal SPEs, same for p and n
    (read_hamiltonian_1body_antoine, src/module_hamiltonian.f90:180-189)."""
    n = basis.dim
    e = np.array([spe[s - 1] for s in basis.sp_sh], dtype=np.float64)
    return np.diag(e).astype(np.complex128)


def write_red(path: str, name: str, htype: int, H0: complex, H1, abcd, H2, trperm, rh2_bytes=4):
    n = H1.shape[0]
    with open(path, "wb") as f:
        f.write(name.encode()[:100].ljust(100, b" "))
        f.write(struct.pack("<i", htype))
        f.write(struct.pack("<dd", complex(H0).real, complex(H0).imag))
        f.write(np.asfortranarray(H1, dtype=np.complex128).tobytes(order="F"))
        f.write(struct.pack("<q", len(trperm)))
        f.write(np.ascontiguousarray(abcd, dtype="<i2").tobytes())
        f.write(np.ascontiguousarray(H2, dtype="<f4" if rh2_bytes == 4 else "<f8").tobytes())
        f.write(np.ascontiguousarray(trperm, dtype="<i1").tobytes())


def read_red(path: str, ndim: int, rh2_bytes: int = 4) -> dict:
    with open(path, "rb") as f:
        raw = f.read()
    off = 100
    htype = struct.unpack_from("<i", raw, off)[0]; off += 4
    H0 = complex(*struct.unpack_from("<dd", raw, off)); off += 16
    H1 = np.frombuffer(raw, "<c16", ndim * ndim, off).reshape((ndim, ndim), order="F").copy()
    off += 16 * ndim * ndim
    dim = struct.unpack_from("<q", raw, off)[0]; off += 8
    abcd = np.frombuffer(raw, "<i2", 4 * dim, off).reshape((dim, 4)).copy(); off += 8 * dim
    H2 = np.frombuffer(raw, "<f4" if rh2_bytes == 4 else "<f8", dim, off).copy(); off += rh2_bytes * dim
    trperm = np.frombuffer(raw, "<i1", dim, off).copy(); off += dim
    if off != len(raw):
        raise ValueError("reduced file size mismatch")
    return dict(name=raw[:100].decode("latin-1").rstrip(), type=htype, H0=H0, H1=H1, abcd=abcd,
                H2=H2, trperm=trperm)


def workload_hamiltonian(w: I.Workload, seed: int = 1234):
    """(basis, H0, H1, abcd, H2 float64, trperm, spe) of a named synthetic workload,
    identical to what the reference builds from the .sho that inputs.write_workload writes."""
    shells = I.space_shells(w.space)
    basis = I.Basis(shells)
    spe = I.workload_spe(w, shells)
    tb = I.synth_jt_tbme(shells, seed=seed, sigma=w.sigma, pairing_shift=w.pairing, decay=w.decay)
    V, jmax = jt_table(shells, tb)
    abcd, H2, trperm = mscheme(basis, V, jmax)
    H1 = one_body_antoine(basis, spe)
    return dict(basis=basis, H0=0j, H1=H1, abcd=abcd, H2=H2, trperm=trperm, spe=spe)
