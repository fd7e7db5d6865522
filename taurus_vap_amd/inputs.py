"""Seeded synthetic inputs in the reference's own file formats.

Writes what the reference TAURUS_vap reads, so the same files drive the GPU path
and the reference CPU build:

* ANTOINE JT-scheme Hamiltonian ``<name>.sho`` (hamil_type 1), read by
  ``read_hamiltonian_1body_antoine`` / ``read_hamiltonian_2body_antoine``
  (src/module_hamiltonian.f90:150-193, 409-456) and ``set_basis``
  (src/module_basis.f90:73-189).
* J-scheme Hamiltonians (hamil_type 3 and 4): ``<name>.sho``, ``.01b``, ``.2b`` and
  ``.com``, read by ``read_hamiltonian_1body_general`` and
  ``read_hamiltonian_2body_Jscheme`` / ``_com`` (src/module_hamiltonian.f90:
  200-239, 463-525).
* The fixed-format STDIN deck read by ``read_input``
  (src/module_initialization.f90:86-155).
* ``initial_wf.txt`` in the text layout of ``read_wavefunction``
  (src/module_wavefunctions.f90:568-655).

The single-particle ordering, time-reversal partner and quantum numbers follow
``set_basis`` (src/module_basis.f90:209-238).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# Model spaces
# --------------------------------------------------------------------------

def ho_shells(nmax: int) -> list[tuple[int, int, int]]:
    """All HO shells (n, l, 2j) with 2n+l <= nmax, ordered by major shell N,
    then by decreasing j within N (0s1/2 | 0p3/2 0p1/2 | 0d5/2 0d3/2 1s1/2 ...)."""
    out = []
    for N in range(nmax + 1):
        shell = []
        for l in range(N % 2, N + 1, 2):
            n = (N - l) // 2
            for j2 in (2 * l + 1, 2 * l - 1):
                if j2 > 0:
                    shell.append((n, l, j2))
        shell.sort(key=lambda s: (-s[2], s[0]))
        out.extend(shell)
    return out


NAMED_SPACES = {
    # sd: 0d5/2 1s1/2 0d3/2 (ANTOINE order of usdb.sho)
    "sd": [(0, 2, 5), (1, 0, 1), (0, 2, 3)],
    # pf: 0f7/2 1p3/2 0f5/2 1p1/2
    "pf": [(0, 3, 7), (1, 1, 3), (0, 3, 5), (1, 1, 1)],
    # sdpf: sd + pf (64 sp states)
    "sdpf": [(0, 2, 5), (1, 0, 1), (0, 2, 3), (0, 3, 7), (1, 1, 3), (0, 3, 5), (1, 1, 1)],
}


def space_shells(space: str) -> list[tuple[int, int, int]]:
    if space in NAMED_SPACES:
        return list(NAMED_SPACES[space])
    if space.startswith("emax"):
        return ho_shells(int(space[4:]))
    raise ValueError(f"unknown model space {space!r}")


def antoine_name(sh: tuple[int, int, int]) -> int:
    n, l, j2 = sh
    return 1000 * n + 100 * l + j2


@dataclass
class Basis:
    """m-scheme single-particle basis in the reference ordering
    (src/module_basis.f90:209-238): protons 1..n/2 then neutrons, shell by
    shell, m_j running from +j down to -j."""
    shells: list
    sp_n: np.ndarray = field(init=False)
    sp_l: np.ndarray = field(init=False)
    sp_2j: np.ndarray = field(init=False)
    sp_2mj: np.ndarray = field(init=False)
    sp_2mt: np.ndarray = field(init=False)
    sp_sh: np.ndarray = field(init=False)   # 1-based shell index
    sp_tr: np.ndarray = field(init=False)   # 1-based time-reversal partner

    def __post_init__(self):
        n_, l_, j_, m_, t_, s_, tr_ = [], [], [], [], [], [], []
        half = sum(sh[2] + 1 for sh in self.shells)
        k = 0
        rows = []
        for i, (n, l, j2) in enumerate(self.shells):
            for mj in range(j2, -j2 - 1, -2):
                k += 1
                rows.append((n, l, j2, mj, i + 1, k + mj))
        for tz in (-1, 1):
            for (n, l, j2, mj, sh, tr) in rows:
                n_.append(n); l_.append(l); j_.append(j2); m_.append(mj)
                t_.append(tz); s_.append(sh)
                tr_.append(tr + (half if tz == 1 else 0))
        self.sp_n = np.array(n_, dtype=np.int32)
        self.sp_l = np.array(l_, dtype=np.int32)
        self.sp_2j = np.array(j_, dtype=np.int32)
        self.sp_2mj = np.array(m_, dtype=np.int32)
        self.sp_2mt = np.array(t_, dtype=np.int32)
        self.sp_sh = np.array(s_, dtype=np.int32)
        self.sp_tr = np.array(tr_, dtype=np.int32)

    @property
    def dim(self) -> int:
        return int(self.sp_n.size)


# --------------------------------------------------------------------------
# Synthetic JT-coupled two-body interaction
# --------------------------------------------------------------------------

def synth_jt_tbme(shells, seed: int, sigma: float = 1.0, pairing_shift: float = 0.0,
                  decay: float | None = None) -> list[tuple]:
    """Random JT-scheme matrix elements <ab;JT|V|cd;JT> ~ N(0, sigma) for one
    representative per symmetry class: a<=b, c<=d, (a,b)<=(c,d) (1-based shell
    indices), parity-conserving, J+T odd when a==b or c==d (antisymmetry).
    ``pairing_shift`` is added on the J=0 T=1 diagonals <aa|V|aa>.
    ``decay`` multiplies by exp(-(N_a+N_b+N_c+N_d)/decay) (N = 2n+l).
    The reference reconstructs all other permutations itself
    (src/module_hamiltonian.f90:284-319).  Returns (a,b,c,d,T,J,value)."""
    rng = np.random.default_rng(seed)
    ns = len(shells)
    pairs = [(a, b) for a in range(ns) for b in range(a, ns)]
    out = []
    for i1, (a, b) in enumerate(pairs):
        for (c, d) in pairs[i1:]:
            la, lb, lc, ld = shells[a][1], shells[b][1], shells[c][1], shells[d][1]
            if (la + lb) % 2 != (lc + ld) % 2:
                continue
            ja, jb, jc, jd = shells[a][2], shells[b][2], shells[c][2], shells[d][2]
            jmin = max(abs(ja - jb), abs(jc - jd)) // 2
            jmax = min(ja + jb, jc + jd) // 2
            if jmin > jmax:
                continue
            fac = 1.0
            if decay is not None:
                Ns = sum(2 * shells[x][0] + shells[x][1] for x in (a, b, c, d))
                fac = math.exp(-Ns / decay)
            for T in (0, 1):
                for J in range(jmin, jmax + 1):
                    if (a == b or c == d) and (J + T) % 2 == 0:
                        continue
                    v = float(rng.normal(0.0, sigma)) * fac
                    if pairing_shift and T == 1 and J == 0 and a == b and (a, b) == (c, d):
                        v += pairing_shift
                    # the value is defined as the 10-decimal number the file holds
                    out.append((a + 1, b + 1, c + 1, d + 1, T, J, float("%.10f" % v)))
    return out


def write_antoine(path: str, name: str, shells, spe, tbme, core=(8, 8), opt_hw=None):
    """ANTOINE type-1 file (same SPE for protons and neutrons)."""
    ns = len(shells)
    by_quad: dict = {}
    for (a, b, c, d, T, J, v) in tbme:
        by_quad.setdefault((a, b, c, d), {}).setdefault(T, {})[J] = v
    with open(path, "w") as f:
        f.write(f" {name}\n")
        f.write(" 1 %d " % ns + " ".join(str(antoine_name(s)) for s in shells) + "\n")
        f.write(" " + " ".join("%.6f" % e for e in spe) + "\n")
        tail = "" if opt_hw is None else "  2 %.4f" % opt_hw
        f.write(" 1 %d %d 0.000000%s\n" % (core[0], core[1], tail))
        for (a, b, c, d), byT in by_quad.items():
            ja, jb, jc, jd = (shells[x - 1][2] for x in (a, b, c, d))
            jmin = max(abs(ja - jb), abs(jc - jd)) // 2
            jmax = min(ja + jb, jc + jd) // 2
            names = [antoine_name(shells[x - 1]) for x in (a, b, c, d)]
            f.write(" 0 1 %d %d %d %d %d %d\n" % (*names, jmin, jmax))
            for T in (0, 1):
                vals = [byT.get(T, {}).get(J, 0.0) for J in range(jmin, jmax + 1)]
                f.write(" " + " ".join("%.10f" % v for v in vals) + "\n")


# --------------------------------------------------------------------------
# J-scheme interactions (hamil_type 3 and 4)
# --------------------------------------------------------------------------

def jscheme_name(sh: tuple[int, int, int]) -> int:
    """Shell label of the J-scheme files: 10000 n + 100 l + 2j (set_basis,
    src/module_basis.f90:164-175)."""
    n, l, j2 = sh
    return 10000 * n + 100 * l + j2


def synth_j_tbme(shells, seed: int, sigma: float = 1.0, decay: float | None = None) -> list[tuple]:
    """Random J-scheme matrix elements for one representative per symmetry class
    (a<=b, c<=d, (a,b)<=(c,d), 1-based shells), parity conserving, independently
    drawn in the six channels of the J-scheme files (0 pp pp, 1 pn pn, 2 pn np,
    3 np pn, 4 np np, 5 nn nn); the pp and nn channels are zero for J odd when
    a==b or c==d.  Where a degenerate class (a==b, c==d or (a,b)==(c,d)) maps one
    channel onto another, the symmetry reconstruction of the reference decides
    which value stays (src/module_hamiltonian.f90:284-319).
    Returns (a, b, c, d, J, [v_0 .. v_5])."""
    rng = np.random.default_rng(seed)
    ns = len(shells)
    pairs = [(a, b) for a in range(ns) for b in range(a, ns)]
    out = []
    for i1, (a, b) in enumerate(pairs):
        for (c, d) in pairs[i1:]:
            la, lb, lc, ld = (shells[x][1] for x in (a, b, c, d))
            if (la + lb) % 2 != (lc + ld) % 2:
                continue
            ja, jb, jc, jd = (shells[x][2] for x in (a, b, c, d))
            jmin = max(abs(ja - jb), abs(jc - jd)) // 2
            jmax = min(ja + jb, jc + jd) // 2
            if jmin > jmax:
                continue
            fac = 1.0
            if decay is not None:
                Ns = sum(2 * shells[x][0] + shells[x][1] for x in (a, b, c, d))
                fac = math.exp(-Ns / decay)
            for J in range(jmin, jmax + 1):
                v = [float("%.10f" % (rng.normal(0.0, sigma) * fac)) for _ in range(6)]
                if (a == b or c == d) and J % 2 == 1:
                    v[0] = v[5] = 0.0
                out.append((a + 1, b + 1, c + 1, d + 1, J, v))
    return out


def write_jscheme(stem: str, name: str, shells, htype: int, core=(0, 0), opt_hw: int = 2,
                  hw: float = 20.0, H0: float = 0.0, h1=None, tbme=(), com_tbme=None):
    """``<stem>.sho`` and ``<stem>.2b`` of a J-scheme interaction, with ``<stem>.01b``
    for type 3 and ``<stem>.com`` when com_tbme is given (set_basis,
    src/module_basis.f90:79-124; read_hamiltonian_1body_general, _2body_Jscheme,
    _2body_com, src/module_hamiltonian.f90:200-239, 463-525).
    h1: {(a, b): (e_p, e_n)} over 1-based shells; tbme / com_tbme as synth_j_tbme."""
    with open(stem + ".sho", "w") as f:
        f.write(" %s\n %d\n" % (name, htype))
        f.write(" %d " % len(shells) + " ".join(str(jscheme_name(s)) for s in shells) + "\n")
        f.write(" %d %d\n" % core)
        f.write(" %d %.6f\n" % (opt_hw, hw))
    if htype == 3:
        with open(stem + ".01b", "w") as f:
            f.write(" %s\n %.10f\n" % (name, H0))
            for (a, b), (ep, en) in (h1 or {}).items():
                f.write(" %d %d %.10f %.10f\n" % (jscheme_name(shells[a - 1]),
                                                 jscheme_name(shells[b - 1]), ep, en))

    def blocks(path, rows_in):
        by_quad: dict = {}
        for (a, b, c, d, J, vals) in rows_in:
            by_quad.setdefault((a, b, c, d), []).append((J, vals))
        with open(path, "w") as f:
            f.write(" %s\n" % name)
            for (a, b, c, d), rows in by_quad.items():
                names = [jscheme_name(shells[x - 1]) for x in (a, b, c, d)]
                f.write(" 0 5 %d %d %d %d %d %d\n" % (*names, rows[0][0], rows[-1][0]))
                for J, vals in rows:
                    f.write(" " + " ".join("%.10f" % v for v in vals) + "\n")

    blocks(stem + ".2b", tbme)
    if com_tbme is not None:
        blocks(stem + ".com", com_tbme)


def write_jscheme_synthetic(stem: str, htype: int, com: int, seed: int = 11, emax: int = 2):
    """The seeded J-scheme interaction of the J-scheme tests on the e_max space:
    random couplings of synth_j_tbme (sigma 1, decay 4), for type 3 a .01b file with
    H0 = 1.5, hw (N + 3/2) diagonals shifted per shell and one 0s1/2 - 1s1/2
    coupling; hw = 20 (type 3) or 18 (type 4); with com a .com file of width 0.1."""
    shells = ho_shells(emax)
    tb = synth_j_tbme(shells, seed=seed, sigma=1.0, decay=4.0)
    com_tb = synth_j_tbme(shells, seed=seed + 1, sigma=0.1) if com else None
    h1 = {(i + 1, i + 1): (20.0 * (2 * sh[0] + sh[1] + 1.5) + 0.1 * i,
                           20.0 * (2 * sh[0] + sh[1] + 1.5) - 0.05 * i)
          for i, sh in enumerate(shells)}
    s1 = [i + 1 for i, sh in enumerate(shells) if sh[1] == 0 and sh[2] == 1]
    if len(s1) >= 2:
        h1[(s1[0], s1[1])] = (0.7, -0.4)
    write_jscheme(stem, "synthetic J-scheme type %d seed %d" % (htype, seed), shells, htype,
                  hw=18.0 if htype == 4 else 20.0, H0=1.5, h1=h1, tbme=tb, com_tbme=com_tb)
    return shells


# --------------------------------------------------------------------------
# STDIN deck (fixed format, src/module_initialization.f90:93-155)
# --------------------------------------------------------------------------

_CONSTR_NAMES = [
    "Constraint multipole Q10", "Constraint multipole Q11", "Constraint multipole Q20",
    "Constraint multipole Q21", "Constraint multipole Q22", "Constraint multipole Q30",
    "Constraint multipole Q31", "Constraint multipole Q32", "Constraint multipole Q33",
    "Constraint multipole Q40", "Constraint multipole Q41", "Constraint multipole Q42",
    "Constraint multipole Q43", "Constraint multipole Q44", "Constraint radius sqrt(r^2)",
    "Constraint ang. mom. Jx", "Constraint ang. mom. Jy", "Constraint ang. mom. Jz",
    "Constraint pair P_T00_J10", "Constraint pair P_T00_J1m1", "Constraint pair P_T00_J1p1",
    "Constraint pair P_T10_J00", "Constraint pair P_T1m1_J00", "Constraint pair P_T1p1_J00",
    "Constraint field Delta",
]


def _l30(s: str) -> str:
    return (s + " " * 30)[:30]


def write_deck(path: str, hamil: str, Z: float, N: float, Mphip: int, Mphin: int, *,
               hamil_read: int = 0, hamil_com: int = 0, seed_type: int = 1, seed_text: int = 1,
               seed_symm: int = 0, iter_max: int = 100, iter_print: int = 0,
               gradient_type: int = 0, eta: float = 0.01, mu: float = 0.0,
               grad_eps: float = 1e-4, enforce_NZ: int = 1, opt_betalm: int = 1,
               constraint_eps: float = 1e-6, constraints: dict | None = None):
    """Writes the 69-line STDIN deck.  ``constraints`` maps a name from
    _CONSTR_NAMES (e.g. "Constraint multipole Q20") to its value."""
    constraints = constraints or {}
    L = []
    L += ["Interaction", "-----------", _l30("Master name hamil. files") + hamil,
          _l30("Center-of-mass correction") + str(hamil_com), _l30("Read reduced hamiltonian") + str(hamil_read),
          _l30("No. of MPI proc per H team") + "    0", "",
          "Particle Number", "---------------",
          _l30("Number of active protons") + "%7.2f" % Z, _l30("Number of active neutrons") + "%7.2f" % N,
          _l30("No. of gauge angles protons") + "%5d" % Mphip,
          _l30("No. of gauge angles neutrons") + "%5d" % Mphin, "",
          "Wave Function", "-------------",
          _l30("Type of seed wave function") + str(seed_type),
          _l30("Number of QP to block") + "    0",
          _l30("No symmetry simplifications") + str(seed_symm),
          _l30("Seed random number generation") + "    0",
          _l30("Read/write wf file as text") + str(seed_text),
          _l30("Cutoff occupied s.-p. states") + " 0.000E+00",
          _l30("Include all empty sp states") + "0",
          _l30("Spatial one-body density") + "0",
          _l30("Discretization for x/r") + "  0  0.00",
          _l30("Discretization for y/theta") + "  0  0.00",
          _l30("Discretization for z/phi") + "  0  0.00", "",
          "Iterative Procedure", "-------------------",
          _l30("Maximum no. of iterations") + "%5d" % iter_max,
          _l30("Step intermediate wf writing") + "    0",
          _l30("More intermediate printing") + str(iter_print),
          _l30("Type of gradient") + str(gradient_type),
          _l30("Parameter eta for gradient") + "%10.3E" % eta,
          _l30("Parameter mu  for gradient") + "%10.3E" % mu,
          _l30("Tolerance for gradient") + "%10.3E" % grad_eps, "",
          "Constraints", "-----------",
          _l30("Force constraint N/Z") + str(enforce_NZ),
          _l30("Constraint beta_lm") + str(opt_betalm),
          _l30("Pair coupling scheme") + "1",
          _l30("Tolerence for constraints") + "%10.3E" % constraint_eps]
    for name in _CONSTR_NAMES:
        if name in constraints:
            L.append(_l30(name) + "1 %7.3f" % constraints[name])
        else:
            L.append(_l30(name) + "0   0.000")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


# --------------------------------------------------------------------------
# Seed wave function: real Thouless rotation of the lowest-filled Slater vacuum
# --------------------------------------------------------------------------

def slater_uv(basis: Basis, Z: int, N: int):
    """Bogoliubov matrices of the Slater determinant that fills the first Z
    proton and first N neutron states of the basis: U = 1 on empty, V = 1 on
    occupied single-particle states."""
    n = basis.dim
    h = n // 2
    occ = np.zeros(n, dtype=bool)
    occ[:Z] = True
    occ[h:h + N] = True
    U = np.diag((~occ).astype(np.float64))
    V = np.diag(occ.astype(np.float64))
    return U, V


def thouless_evolve(U, V, Z):
    """U1 = (U + V Z) L^-T, V1 = (V + U Z) L^-T with L L^T = 1 + Z^T Z, the
    evolution of src/module_constraints.f90:497-590 (real Z)."""
    n = U.shape[0]
    L = np.linalg.cholesky(np.eye(n) + Z.T @ Z)
    Linv = np.linalg.inv(L)
    return (U + V @ Z) @ Linv.T, (V + U @ Z) @ Linv.T


def seed_wavefunction(basis: Basis, Z: int, N: int, seed: int, sigma: float = 0.1):
    """Real skew Thouless matrix Z ~ N(0, sigma) applied to the lowest-filled
    Slater vacuum."""
    rng = np.random.default_rng(seed)
    n = basis.dim
    A = rng.normal(0.0, sigma, size=(n, n))
    Zt = np.triu(A, 1)
    Zt = Zt - Zt.T
    U0, V0 = slater_uv(basis, Z, N)
    return thouless_evolve(U0, V0, Zt)


def write_wavefunction_text(path: str, shells, U, V, label: int = 0, jscheme: bool = False):
    """Shells are labelled as the interaction labels them: the ANTOINE
    1000 n + 100 l + 2j, or 10000 n + 100 l + 2j for a J-scheme interaction."""
    name = jscheme_name if jscheme else antoine_name
    with open(path, "w") as f:
        f.write("%12d\n" % len(shells))
        for s in shells:
            f.write("%12d\n" % name(s))
        f.write("%22d\n" % label)
        for M in (U, V):
            for v in np.asarray(M, dtype=np.float64).flatten(order="F"):
                f.write("%25.17E\n" % v)


def read_wavefunction_text(path: str):
    with open(path) as f:
        tok = f.read().split()
    nsh = int(tok[0])
    names = [int(t) for t in tok[1:1 + nsh]]
    vals = np.array([float(t.replace("D", "E")) for t in tok[2 + nsh:]], dtype=np.float64)
    n = int(round(math.sqrt(vals.size / 2)))
    U = vals[:n * n].reshape((n, n), order="F")
    V = vals[n * n:].reshape((n, n), order="F")
    return names, U, V


# --------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    space: str
    Z: int
    N: int
    M: int
    sigma: float
    pairing: float
    spe: object           # list per shell, or "hw" for hbar*omega*(N+3/2)
    decay: float | None = None
    hw: float = 20.0
    iters: int = 100
    constraints: dict | None = None
    gradient_type: int = 0
    eta: float = 0.01
    mu: float = 0.0
    seed_sigma: float = 0.1   # width of the random Thouless matrix of the seed state


WORKLOADS = {
    # Plain PNVAP gradient descent, sd shell, 4p+4n, 7x7 angles, 200 iterations
    "sd": Workload("sd", "sd", 4, 4, 7, 1.0, -2.0, [0.0, 1.0, 2.0], iters=200),
    # pf shell, 5p/4n (odd number parity), Q20 at beta2=0.2, 9x9 angles
    "pf": Workload("pf", "pf", 5, 4, 9, 1.0, -1.5, [0.0, 4.0 / 3.0, 8.0 / 3.0, 4.0], iters=300,
                   constraints={"Constraint multipole Q20": 0.2}),
    # no-core e_max=4 (140 states), 8p/8n, 11x11
    "emax4": Workload("emax4", "emax4", 8, 8, 11, 2.0, 0.0, "hw", decay=4.0, iters=10,
                      seed_sigma=0.02),
    # sdpf (64 states), 6p/6n, Q20, 9x9 angles: the largest basis of the batched solves
    "sdpf": Workload("sdpf", "sdpf", 6, 6, 9, 1.0, -1.5,
                     [0.0, 1.0, 2.0, 4.0, 16.0 / 3.0, 20.0 / 3.0, 8.0], iters=50,
                     constraints={"Constraint multipole Q20": 0.2}),
    # e_max=6 (336 states), 10p/12n, Q20 + Jx, 15x15 angles
    "emax6": Workload("emax6", "emax6", 10, 12, 15, 2.0, 0.0, "hw", decay=4.0, iters=100,
                      constraints={"Constraint multipole Q20": 0.2,
                                   "Constraint ang. mom. Jx": 0.0}, seed_sigma=0.01),
}


# Independent constrained solves of uneven basis size run together: (workload,
# number of solves) with Q20 values on a beta2 grid from -0.4 to 0.6, 9x9 angles.
BATCHED = [("sd", 24), ("pf", 24), ("sdpf", 16)]
BATCHED_BETA2 = (-0.4, 0.6)


def batched_betas(count: int) -> list[float]:
    lo, hi = BATCHED_BETA2
    return [lo + (hi - lo) * k / max(1, count - 1) for k in range(count)]


def workload_spe(w: Workload, shells):
    """Single-particle energies as written in the .sho (6 decimals)."""
    if w.spe == "hw":
        e = [w.hw * (2 * n + l + 1.5) for (n, l, j2) in shells]
    else:
        e = list(w.spe)
    return [float("%.6f" % x) for x in e]


def write_workload(w: Workload, outdir: str, seed: int = 1234, M: int | None = None,
                   iters: int | None = None, hamil_read: int = 0, iter_print: int = 0,
                   gradient_type: int | None = None, eta: float | None = None,
                   mu: float | None = None, grad_eps: float = 1e-12, UV=None):
    """Writes <outdir>/{<name>.sho, input.txt, initial_wf.txt} for workload w
    (the seed state is UV = (U, V) when given).  Returns (basis, U, V)."""
    os.makedirs(outdir, exist_ok=True)
    shells = space_shells(w.space)
    basis = Basis(shells)
    spe = workload_spe(w, shells)
    tb = synth_jt_tbme(shells, seed=seed, sigma=w.sigma, pairing_shift=w.pairing, decay=w.decay)
    core = (0, 0) if w.space.startswith("emax") else ((8, 8) if w.space == "sd" else (20, 20))
    write_antoine(os.path.join(outdir, w.name + ".sho"), "synthetic %s seed %d" % (w.name, seed),
                  shells, spe, tb, core=core,
                  opt_hw=w.hw if w.space.startswith("emax") else None)
    Mv = w.M if M is None else M
    write_deck(os.path.join(outdir, "input.txt"), w.name, w.Z, w.N, Mv, Mv,
               hamil_read=hamil_read, iter_max=w.iters if iters is None else iters,
               iter_print=iter_print,
               gradient_type=w.gradient_type if gradient_type is None else gradient_type,
               eta=w.eta if eta is None else eta, mu=w.mu if mu is None else mu,
               grad_eps=grad_eps, constraints=w.constraints)
    if UV is None:
        U, V = seed_wavefunction(basis, w.Z, w.N, seed=seed + 1, sigma=w.seed_sigma)
    else:
        U, V = UV
    write_wavefunction_text(os.path.join(outdir, "initial_wf.txt"), shells, U, V)
    return basis, U, V
