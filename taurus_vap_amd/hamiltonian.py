"""Two-body Hamiltonian in the reference's reduced format.

* ``read_hamiltonian`` -- set_hamiltonian with hamil_read = 0 (src/module_hamiltonian.
  f90:79-106): the basis, H0, H1 and the stored m-scheme list (abcd, H2, trperm)
  from the reference's own interaction files: ANTOINE JT-scheme ``.sho`` (types
  1-2) or the J-scheme ``.sho`` + ``.01b`` + ``.2b`` (+ ``.com``) (types 3-4);
* ``jt_table`` / ``reconstruct`` -- hamil_H2cpd after the readers and the
  symmetry reconstruction (:284-319), the reconstruction in the native producer;
* ``mscheme`` -- the stored m-scheme list through the native producer
  libtaurus_hamil.so (include/taurus_hamil.h);
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
        for f in (lib.th_mscheme, lib.th_mscheme_j):
            f.argtypes = [I_, P, P, P, P, P, I_, I_, P, D, L, P, P]
            f.restype = L
        lib.th_reconstruct_2body.argtypes = [I_, I_, I_, P, P]
        lib.th_timerev.argtypes = [I_, P, P, L, P, P]
        lib.th_clebsch_gordan.argtypes = [I_] * 6
        lib.th_clebsch_gordan.restype = D
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def reconstruct(V: np.ndarray, htype: int, sh_2j) -> np.ndarray:
    """The symmetry reconstruction of set_hamiltonian_2body (src/module_hamiltonian.
    f90:284-319), in place on hamil_H2cpd V (channels, 0:jmax, sh, sh, sh, sh),
    Fortran-ordered, in the reference's loop order (th_reconstruct_2body)."""
    if not V.flags.f_contiguous or V.dtype != np.float64:
        raise ValueError("V must be a Fortran-ordered float64 array")
    nch, jp1, nsh = V.shape[0], V.shape[1], V.shape[2]
    if nch != (4 if htype <= 2 else 6):
        raise ValueError("hamil_type %d takes %d channels" % (htype, 4 if htype <= 2 else 6))
    two_j = np.ascontiguousarray(sh_2j, dtype=np.int32)
    if _load().th_reconstruct_2body(htype, nsh, jp1 - 1, _p(two_j), _p(V)) != 0:
        raise RuntimeError("th_reconstruct_2body: bad arguments")
    return V


def jt_table(shells, tbme) -> tuple[np.ndarray, int]:
    """hamil_H2cpd for a type-1 ANTOINE interaction given by its representatives
    (a,b,c,d,T,J,v) (1-based shells, T = 0/1): T=0 fills index 0, T=1 indices
    1,2,3 (all M_T) as read_hamiltonian_2body_antoine does for tmax = 1, then the
    symmetry reconstruction."""
    nsh = len(shells)
    jmax = max(s[2] for s in shells)
    V = np.zeros((4, jmax + 1, nsh, nsh, nsh, nsh), order="F")
    for (a, b, c, d, T, J, v) in tbme:
        for mmt in ((0,) if T == 0 else (1, 2, 3)):
            V[mmt, J, a - 1, b - 1, c - 1, d - 1] = v
    return reconstruct(V, 1, [s[2] for s in shells]), jmax


def mscheme(basis: I.Basis, V: np.ndarray, jmax: int, scaling_2b: float = 1.0, htype: int = 1):
    """Stored m-scheme elements (abcd 1-based int16 [ntb,4], H2 float64, trperm int8):
    uncouple_JT for htype 1-2 (V with 4 channels), uncouple_J for 3-4 (6 channels)."""
    lib = _load()
    produce = lib.th_mscheme if htype <= 2 else lib.th_mscheme_j
    n = basis.dim
    nsh = V.shape[2]
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in
            (basis.sp_2j, basis.sp_2mj, basis.sp_2mt, basis.sp_l, basis.sp_sh)]
    Vf = np.asfortranarray(V).ravel(order="F")
    ub = produce(n, *[_p(x) for x in arrs], nsh, jmax, None, 1.0, 0, None, None)
    if ub < 0:
        raise RuntimeError("th_mscheme: bad arguments")
    abcd = np.zeros((max(ub, 1), 4), dtype=np.int16)
    H2 = np.zeros(max(ub, 1))
    cnt = produce(n, *[_p(x) for x in arrs], nsh, jmax, _p(Vf), scaling_2b, ub, _p(abcd), _p(H2))
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


# --------------------------------------------------------------------------
# The reference's interaction files (set_hamiltonian with hamil_read = 0)
# --------------------------------------------------------------------------

HBARC = 197.3269788                       # MeV fm (src/module_constants.f90:45-52)
MASS_MP, MASS_MN = 938.27208816, 939.56542052
MASS_MA = (MASS_MP + MASS_MN) / 2
HBARMASS = HBARC * HBARC / (2 * MASS_MA)


class _Records:
    """Fortran list-directed reads over the lines of a text file: every read
    starts at a new line, takes as many values as it needs (from as many lines),
    and drops what is left of its last line."""

    def __init__(self, path: str):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.pos = 0

    def skip(self, k: int = 1):
        self.pos += k

    def eof(self) -> bool:
        return self.pos >= len(self.lines)

    def read(self, count: int) -> list[str]:
        out: list[str] = []
        while len(out) < count:
            if self.pos >= len(self.lines):
                raise EOFError
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            out.extend(toks)
        return out[:count]


def _num(t: str) -> float:
    return float(t.replace("D", "E").replace("d", "e"))


def _shell_index(names, label: int) -> int:
    idx = [i for i, na in enumerate(names) if na == label]
    if not idx:
        raise ValueError("shell %d is not in the model space" % label)
    return idx[-1]


def _read_coupled_blocks(path: str, names, V: np.ndarray, antoine_rec: _Records | None = None,
                         add: float | None = None):
    """J-scheme blocks 'tmin tmax a b c d jmin jmax' + one line of channels tmin..tmax
    per J (read_hamiltonian_2body_Jscheme / _com, src/module_hamiltonian.f90:463-525).
    With ``add`` the values are added times ``add`` (the centre-of-mass file; the
    channels a block does not list keep the previous block's values, as H2com does)."""
    rec = _Records(path)
    rec.skip()
    h2com = np.zeros(6)
    while True:
        try:
            hdr = rec.read(8)
        except EOFError:
            break
        try:
            tmin, tmax, *nlj, jmin, jmax = (int(x) for x in hdr)
        except ValueError:
            break
        a, b, c, d = (_shell_index(names, x) for x in nlj)
        for J in range(jmin, jmax + 1):
            vals = [_num(x) for x in rec.read(tmax - tmin + 1)]
            if add is None:
                V[tmin:tmax + 1, J, a, b, c, d] = vals
            else:
                h2com[tmin:tmax + 1] = vals
                V[:, J, a, b, c, d] = V[:, J, a, b, c, d] + add * h2com


def kinetic_bare(basis: I.Basis, hw: float, htype: int, hamil_com: int, nucleus_A: float,
                 H1: np.ndarray) -> np.ndarray:
    """set_kinetic_bare (src/module_hamiltonian.f90:692-733): adds
    (hw/2)(delta_{htype,4} - com/A) times the HO kinetic matrix to H1."""
    factor = (0.5 * hw) * ((1.0 if htype == 4 else 0.0) - hamil_com / nucleus_A)
    n_, l_, j_, m_, t_ = basis.sp_n, basis.sp_l, basis.sp_2j, basis.sp_2mj, basis.sp_2mt
    na, nb = n_[:, None].astype(np.float64), n_[None, :].astype(np.float64)
    la, lb = l_[:, None].astype(np.float64), l_[None, :].astype(np.float64)
    same = ((t_[:, None] == t_[None, :]) & (l_[:, None] == l_[None, :]) &
            (j_[:, None] == j_[None, :]) & (m_[:, None] == m_[None, :]))
    f0 = 2.0 * nb + lb + 1.5
    fm = np.sqrt(nb * (nb + lb + 0.5))
    fp = np.sqrt(na * (na + la + 0.5))
    d0 = (n_[:, None] == n_[None, :]).astype(np.float64)
    dm = (n_[:, None] == n_[None, :] - 1).astype(np.float64)
    dp = (n_[:, None] == n_[None, :] + 1).astype(np.float64)
    term = d0 * f0 + dm * fm + dp * fp
    H1 = H1.copy()
    H1[same] = H1[same] + factor * term[same]
    return H1


def read_hamiltonian(sho_path: str, valence_Z: float, valence_N: float, hamil_com: int = 0) -> dict:
    """The Hamiltonian the reference builds from its interaction files when
    hamil_read = 0 (set_basis, set_nucleus, set_hamiltonian_01body,
    set_hamiltonian_2body: src/module_basis.f90:73-189, src/module_nucleus.f90:46-74,
    src/module_hamiltonian.f90:115-401).  ``sho_path`` is the main file; a
    J-scheme interaction (types 3, 4) takes its companions from the same stem
    (``.01b``, ``.2b``, ``.com``).  Returns basis, type, name, HO_hw, nucleus_A,
    H0, H1, abcd (1-based), H2 (float64) and trperm."""
    stem = sho_path[:-4] if sho_path.endswith(".sho") else sho_path
    rec = _Records(sho_path)
    name = rec.lines[0] if rec.lines else ""
    rec.skip()
    htype = int(rec.read(1)[0])
    rec.pos = 1
    if htype in (1, 2):
        nsh = int(rec.read(2)[1])
        rec.pos = 1
        names = [int(x) for x in rec.read(2 + nsh)[2:]]
        facn = 1000
    elif htype in (3, 4):
        rec.pos = 2
        nsh = int(rec.read(1)[0])
        rec.pos = 2
        names = [int(x) for x in rec.read(1 + nsh)[1:]]
        facn = 10000
    else:
        raise ValueError("Format of the hamiltonian not recognized: type %d" % htype)
    shells = []
    for na in names:
        n = na // facn
        l = (na - n * facn) // 100
        shells.append((n, l, na - n * facn - l * 100))
    basis = I.Basis(shells)
    # cores (set_nucleus) and the oscillator frequency (set_basis)
    after_names = rec.pos
    if htype in (1, 2):
        line = rec.lines[after_names + htype].replace(",", " ").split()
        core_Z, core_N = _num(line[1]), _num(line[2])
        opt_hw, hw = (0, 0.0) if len(line) == 4 else (int(line[4]), _num(line[5]))
    else:
        rec.pos = after_names
        core_Z, core_N = (_num(x) for x in rec.read(2))
        opt_hw, hw = rec.read(2)
        opt_hw, hw = int(opt_hw), _num(hw)
    nucleus_A = (core_Z + core_N) + (valence_Z + valence_N)
    if opt_hw == 0:
        hw = 45.0 * nucleus_A ** (-1.0 / 3.0) - 25.0 * nucleus_A ** (-2.0 / 3.0)
    elif opt_hw == 1:
        hw = 41.0 * nucleus_A ** (-1.0 / 3.0)
    elif opt_hw != 2:
        raise ValueError("opt_hw = %d should be 0, 1 or 2" % opt_hw)
    n = basis.dim
    jm2 = max(s[2] for s in shells)
    H0 = 0j
    H1 = np.zeros((n, n), dtype=np.complex128)
    scaling_2b = 1.0
    if htype in (1, 2):
        # read_hamiltonian_1body_antoine (:150-193)
        rec.pos = after_names
        zepsi = [[_num(x) for x in rec.read(nsh)] for _ in range(htype)]
        idens, _, _, zmass = rec.read(4)
        idens, zmass = int(idens), _num(zmass)
        x1, x2 = nucleus_A, (core_Z + core_N) + 2
        scaling_1b = 1.0
        if idens == 1:
            scaling_2b = (x2 / x1) ** zmass
        elif idens == 2:
            scaling_1b = scaling_2b = (x2 / x1) ** zmass
        e = np.array([zepsi[0][s - 1] for s in basis.sp_sh[: n // 2]]) * scaling_1b
        f = np.array([zepsi[htype - 1][s - 1] for s in basis.sp_sh[n // 2:]]) * scaling_1b
        H1[np.arange(n), np.arange(n)] = np.concatenate([e, f])
        # read_hamiltonian_2body_antoine (:409-456), continuing the same file
        V = np.zeros((4, jm2 + 1, nsh, nsh, nsh, nsh), order="F")
        while True:
            try:
                hdr = rec.read(8)
                tmin, tmax, *nlj, jmin, jmax = (int(x) for x in hdr)
            except (EOFError, ValueError):
                break
            a, b, c, d = (_shell_index(names, x) for x in nlj)
            for T in range(tmin, tmax + 1):
                h = np.zeros(jm2 + 1)
                h[jmin:jmax + 1] = [_num(x) for x in rec.read(jmax - jmin + 1)]
                if T == 0:
                    V[0, :, a, b, c, d] = h
                elif T == 1:
                    for i in range(2, 5 - tmax + 1):
                        V[1 if i == 4 else i, :, a, b, c, d] = h
                elif T == 2:
                    V[T + (-1) ** (tmax + 1), :, a, b, c, d] = h
                elif T == 3:
                    V[1, :, a, b, c, d] = h
                else:
                    raise ValueError("Wrong interaction (Antoine) format: isospin!")
    else:
        if htype == 3:
            # read_hamiltonian_1body_general (:200-239)
            r1 = _Records(stem + ".01b")
            r1.skip()
            H0 = complex(_num(r1.read(1)[0]))
            m, sh = basis.sp_2mj[: n // 2], basis.sp_sh[: n // 2]
            label = np.array([names[s - 1] for s in sh])
            while True:
                try:
                    a, b, hp, hn = r1.read(4)
                    a, b, hp, hn = int(a), int(b), _num(hp), _num(hn)
                except (EOFError, ValueError):
                    break
                k, l = np.nonzero((m[:, None] == m[None, :]) & (label[:, None] == a) &
                                  (label[None, :] == b))
                H1[k, l] = hp
                H1[l, k] = hp
                H1[k + n // 2, l + n // 2] = hn
                H1[l + n // 2, k + n // 2] = hn
            if hamil_com == 1:
                H1 = kinetic_bare(basis, hw, htype, hamil_com, nucleus_A, H1)
        else:
            H1 = kinetic_bare(basis, hw, htype, hamil_com, nucleus_A, H1)
        V = np.zeros((6, jm2 + 1, nsh, nsh, nsh, nsh), order="F")
        _read_coupled_blocks(stem + ".2b", names, V)
        if hamil_com == 1:
            _read_coupled_blocks(stem + ".com", names, V, add=hw / (nucleus_A * HBARMASS))
    reconstruct(V, htype, [s[2] for s in shells])
    abcd, H2, trperm = mscheme(basis, V, jm2, scaling_2b, htype)
    return dict(basis=basis, type=htype, name=name, HO_hw=hw, nucleus_A=nucleus_A, H0=H0, H1=H1,
                abcd=abcd, H2=H2, trperm=trperm)


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
