"""CPU restatement (numpy) of the TAURUS_vap projected-gradient hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py may import it.  The
product (taurus_vap_amd) never imports, links or calls anything under oracle/.

Each function restates one reference routine and cites it (paths relative to the
reference repository).  Arrays are numpy, 0-based, column-major semantics kept by
using plain 2-D arrays indexed [row, col] (the Fortran A(i,j) is A[i-1, j-1]).

Pinning (see tests/test_oracle.py):
  * every stage below is compared with the reference itself (oracle/_ref
    harness dumps, committed as tests/golden/*.npz) at sd and pf size;
  * zskpfa: Pf^2 = det, brute-force permutation expansion for n <= 8,
    block-diagonal products, sign under a swap;
  * calculate_overlap: = 1 at phi = 0; |overlap| = Onishi sqrt|det| formula;
  * calculate_fields: dense antisymmetric V_abcd rebuilt from the stored list
    and contracted by einsum;
  * calculate_fields_diag: equals calculate_fields when L = R;
  * projection: <Z>_proj = Z for enough angles; M = 1 limit equals the
    unprojected energy.
"""
from __future__ import annotations

import math
import struct

import numpy as np
import scipy.linalg as sla

# ----------------------------------------------------------------------------
# File formats
# ----------------------------------------------------------------------------

def read_red(path: str, ndim: int, rh2_bytes: int = 4):
    """Reduced Hamiltonian file (stream), src/module_hamiltonian.f90:608-627:
    name[100] | int32 type | c128 H0 | c128 H1[n*n] (col-major) | int64 dim |
    int16 abcd[4*dim] | real(rH2) H2[dim] | int8 trperm[dim]."""
    with open(path, "rb") as f:
        raw = f.read()
    off = 0
    name = raw[off:off + 100].decode("latin-1").rstrip(); off += 100
    htype = struct.unpack_from("<i", raw, off)[0]; off += 4
    H0 = complex(*struct.unpack_from("<dd", raw, off)); off += 16
    H1 = np.frombuffer(raw, dtype="<c16", count=ndim * ndim, offset=off).reshape((ndim, ndim), order="F").copy()
    off += 16 * ndim * ndim
    dim = struct.unpack_from("<q", raw, off)[0]; off += 8
    abcd = np.frombuffer(raw, dtype="<i2", count=4 * dim, offset=off).reshape((dim, 4)).astype(np.int64)
    off += 8 * dim
    dt = "<f4" if rh2_bytes == 4 else "<f8"
    H2 = np.frombuffer(raw, dtype=dt, count=dim, offset=off).astype(np.float64); off += rh2_bytes * dim
    perm = np.frombuffer(raw, dtype="<i1", count=dim, offset=off).astype(np.int64); off += dim
    assert off == len(raw), (off, len(raw))
    return dict(name=name, type=htype, H0=H0, H1=H1, abcd=abcd - 1, H2=H2, trperm=perm)


_KIND = {1: ("<i4", 4), 2: ("<f8", 8), 3: ("<c16", 16), 4: ("<i8", 8), 5: ("<i1", 1), 6: ("<i2", 2)}


def read_harness_dump(path: str) -> dict:
    """Reads oracle/ref_harness.f90 dumps: name -> list of arrays (in order)."""
    with open(path, "rb") as f:
        raw = f.read()
    off = 0
    out: dict = {}
    while off < len(raw):
        name = raw[off:off + 32].decode().strip(); off += 32
        kind, rank = struct.unpack_from("<ii", raw, off); off += 8
        dims = struct.unpack_from("<%dq" % rank, raw, off) if rank else (); off += 8 * rank
        dt, sz = _KIND[kind]
        cnt = int(np.prod(dims)) if rank else 1
        arr = np.frombuffer(raw, dtype=dt, count=cnt, offset=off).copy(); off += sz * cnt
        arr = arr.reshape(dims, order="F") if rank else arr[0]
        out.setdefault(name, []).append(arr)
    return out


# ----------------------------------------------------------------------------
# a1  assign_angles  (src/module_projection.f90:112-185)
# ----------------------------------------------------------------------------

def assign_angles(Mphip: int, Mphin: int, facpi: float, valZ: float, valN: float):
    phip, phin, weip, wein = [], [], [], []
    for jp in range(1, Mphip + 1):
        for jn in range(1, Mphin + 1):
            pp = facpi * (jp - 1.0) / Mphip if Mphip % 2 == 1 else facpi * (jp - 0.5) / Mphip
            pn = facpi * (jn - 1.0) / Mphin if Mphin % 2 == 1 else facpi * (jn - 0.5) / Mphin
            phip.append(pp); phin.append(pn)
            weip.append(np.exp(-1j * pp * valZ) / Mphip)
            wein.append(np.exp(-1j * pn * valN) / Mphin)
    return np.array(phip), np.array(phin), np.array(weip), np.array(wein)


# ----------------------------------------------------------------------------
# a3  construct_canonical_basis  (src/module_wavefunctions.f90:777-945)
#     + v/u set-up of project_wavefunction (src/module_projection.f90:219-234)
# ----------------------------------------------------------------------------

def construct_canonical_basis(U, V, eps: float = 1e-8, allemp: int = 0):
    n = U.shape[0]
    rho = V @ V.T
    ev, Dc = np.linalg.eigh(rho)              # DSYEV('V','U'), ascending
    rhoc = Dc.T @ rho @ Dc
    nocc = nemp = 0
    ovac = 0.0
    for i in range(n):
        v2 = abs(rhoc[i, i])
        u = math.sqrt(abs(1.0 - v2))
        if v2 >= 1.0 - eps:
            nocc += 1
        elif v2 <= eps and allemp == 0:
            nemp += 1
        else:
            ovac += math.log(u)
    kappa = V @ U.T
    kappac = Dc.T @ kappa @ Dc
    # degenerate blocks (1-based loops of :843-875 translated to 0-based)
    submax = np.zeros(n, dtype=int)
    for i in range(nemp, n - nocc):
        submax[i] = i
        for j in range(i + 1, n - nocc):
            if abs(kappac[j, i]) > math.sqrt(eps) * 1.0e-2:
                submax[i] = j
    k = nemp
    while k < n - nocc - 1:
        j = submax[k]
        m = submax[k:j + 1].max()
        submax[k:j + 1] = m
        if m == j:
            k = j + 1
    subdim = np.zeros(n, dtype=int)
    k = nemp
    while k < n - nocc - 1:
        subdim[k] = submax[k] - k + 1
        if subdim[k] > 2 and subdim[k] % 2 != 0:
            subdim[k] = 2
        k += subdim[k]
    A1 = np.eye(n)
    k = nemp
    while k < n - nocc - 1:
        m = subdim[k]
        if m > 2:
            B1 = kappac[k:k + m, k:k + m].copy()
            T, Z = sla.schur(B1, output="real")
            A1[k:k + m, k:k + m] = Z
            kappac[k:k + m, k:k + m] = T
        k += max(m, 1)
    if subdim.max() > 2:
        Dc = Dc @ A1
    zUc = np.zeros((n, n), dtype=complex)
    zVc = np.zeros((n, n), dtype=complex)
    kk = 0
    for i in range(n):
        v2 = abs(rhoc[i, i])
        v = math.sqrt(v2)
        zUc[i, i] = math.sqrt(abs(1.0 - v2))
        if v2 > 1.0 - eps or v2 <= eps:
            zVc[i, i] = v
        else:
            if kk % 2 == 0:
                zVc[i, i + 1] = math.copysign(v, kappac[i, i + 1])
            else:
                zVc[i, i - 1] = math.copysign(v, kappac[i, i - 1])
            kk += 1
    if n - nocc - nemp == 0:
        ovac = 0.0
    vou = np.zeros(n - nocc)
    j = 1
    for i in range(nemp, n - nocc):
        if j % 2 == 1:
            vou[i] = (zVc[i, i + 1] / zUc[i, i]).real
        else:
            vou[i] = (zVc[i, i - 1] / zUc[i, i]).real
        j += 1
    return dict(D=Dc.astype(complex), Uc=zUc, Vc=zVc, ovac=ovac, nocc=nocc, nemp=nemp, voveru=vou)


# ----------------------------------------------------------------------------
# a4 generate_rotation_gauge (:569-588), a5 rotate_wavefunction (:641-653)
# ----------------------------------------------------------------------------

def rotation_gauge(phip: float, phin: float, n: int):
    r = np.empty(n, dtype=complex)
    r[:n // 2] = np.exp(1j * phip)
    r[n // 2:] = np.exp(1j * phin)
    return np.diag(r)


def rotate_wavefunction(R, U, V):
    return R @ U, np.conj(R) @ V


# ----------------------------------------------------------------------------
# a7  ZSKPFA('U','P') (src/subroutines_pfaffian.f:12-256, Parlett-Reid of
#     ZSKTF2 :1694-2015, upper branch; pivot = IZAMAX = first max of |Re|+|Im|)
# ----------------------------------------------------------------------------

def zskpfa(A_in):
    """Pfaffian of a complex skew-symmetric matrix; only the strictly upper
    triangle of A_in is read (UPLO='U').  Returns (pf, info)."""
    n = A_in.shape[0]
    if n == 0:
        return 1.0 + 0j, 0
    if n % 2 == 1:
        return 0j, 0
    U = np.triu(np.asarray(A_in, dtype=complex), 1)
    A = U - U.T
    info = 0
    pf = 1.0 + 0j
    nswap = 0
    for k in range(n - 1, 0, -2):          # 0-based k = K-1, K = N, N-2, ..., 2
        col = A[:k, k]
        mag = np.abs(col.real) + np.abs(col.imag)
        kp = int(np.argmax(mag))
        colmax = abs(col[kp])
        if colmax == 0.0:
            if info == 0:
                info = k
            kp = k - 1
        kk = k - 1
        if kp != kk:
            A[[kk, kp], :] = A[[kp, kk], :]
            A[:, [kk, kp]] = A[:, [kp, kk]]
            nswap += 1
        if colmax != 0.0:
            p = A[kk, k]
            x = A[:kk, k].copy()
            y = A[:kk, kk].copy()
            A[:kk, :kk] += (np.outer(x, y) - np.outer(y, x)) / p
        pf *= A[kk, k]
    if info > 0:
        return 0j, 0
    return pf * (-1) ** nswap, 0


# ----------------------------------------------------------------------------
# a6  calculate_overlap  (src/module_projection.f90:782-937)
# ----------------------------------------------------------------------------

def calculate_overlap(nocc0, nemp0, ovac0, vou0, D0, nocc1, nemp1, ovac1, vou1, D1, ROT):
    n = D0.shape[0]
    nempm = min(nemp1, nemp0)
    nbloc0 = n - nocc0 - nempm
    nbloc1 = n - nocc1 - nempm
    nreg = nbloc1 + nbloc0
    n0, n1, n2 = n - nocc0, n - nocc1, n - nempm
    fac1 = (-1) ** ((n2 * (n2 + 1) // 2) % 2)
    fac2 = (-1) ** ((nocc0 * (nocc0 - 1) // 2) % 2)
    fac3 = (-1) ** ((nocc0 * (nocc0 + 1) // 2 + nocc1 * (nocc1 + 1) // 2 + nocc0 * (n0 + n2) + nocc1 * n1) % 2)
    signfac = fac1 * fac2 * fac3
    normvac = 0.5 * (ovac0 + ovac1)
    isla0 = 1 if n0 - nemp0 == 0 else 0
    isla1 = 2 if n1 - nemp1 == 0 else 0
    overlap = 0j
    s = isla0 + isla1
    if s == 1 and nocc0 > n - nemp1:
        return 0j
    if s == 2 and nocc1 > n - nemp0:
        return 0j
    if s == 3:
        if nocc0 == nocc1:
            overlap = 1.0 + 0j
        else:
            return 0j
    Rn = np.conj(D0).T @ (ROT @ D1)
    Rn2 = Rn[nempm:, nempm:]
    lu, piv = sla.lu_factor(Rn2)             # ZGETRF
    detR = np.prod(np.diag(lu))
    sgn = (-1) ** int(np.sum(piv != np.arange(n2)))
    detR = sgn * detR
    Rinv = sla.lu_solve((lu, piv), np.eye(n2, dtype=complex))   # ZGETRI
    if s != 3:
        M = np.zeros((nreg, nreg), dtype=complex)
        for i in range(1, nbloc1 + 1):
            if i % 2 == 1:
                M[i - 1, i] = vou1[i - 1 + nempm]
            else:
                M[i - 1, i - 2] = vou1[i - 1 + nempm]
        for j in range(1, nbloc0 + 1):
            i = j + nbloc1
            if i % 2 == 1:
                M[i - 1, i] = -vou0[j - 1 + nempm]
            else:
                M[i - 1, i - 2] = -vou0[j - 1 + nempm]
        M[:nbloc1, nbloc1:] = -Rinv[:nbloc1, :nbloc0]
        M[nbloc1:, :nbloc1] = Rinv[:nbloc1, :nbloc0].T
        overlap, _ = zskpfa(M)
    if abs(overlap) > 0.0:
        ph = 1j * math.atan2(overlap.imag, overlap.real)
        overlap = detR * signfac * np.exp(ph + math.log(abs(overlap)) + normvac)
    return complex(overlap)


def calculate_norm(UL, VL, UR, VR):
    """Onishi modulus, src/module_projection.f90:951-978."""
    UI = np.conj(UR).T @ UL + np.conj(VR).T @ VL
    return math.sqrt(abs(np.linalg.det(UI)))


# ----------------------------------------------------------------------------
# a8  calculate_thouless (src/module_projection.f90:684-718)
# ----------------------------------------------------------------------------

def calculate_thouless(U, V, Ubar, Vbar):
    Uphi = V.T @ Vbar + U.T @ Ubar
    Vphi = V.T @ Ubar + U.T @ Vbar
    Uinv = np.linalg.inv(Uphi)
    A = np.conj(Vphi @ Uinv)
    return A, U + V @ A, V + U @ A


# ----------------------------------------------------------------------------
# a9  calculate_densities (src/module_wavefunctions.f90:1073-1087)
# ----------------------------------------------------------------------------

def calculate_densities(UL, VL, UR, VR):
    URc, VRc = np.conj(UR), np.conj(VR)
    return VRc @ VL.T, VRc @ UL.T, VL @ URc.T


# ----------------------------------------------------------------------------
# a12 find_timerev (src/module_hamiltonian.f90:1028-1089) on index arrays
# ----------------------------------------------------------------------------

_TR_ORDER = {0: (0, 1, 2, 3), 1: (1, 0, 2, 3), 2: (0, 1, 3, 2), 3: (1, 0, 3, 2),
             4: (2, 3, 0, 1), 5: (2, 3, 1, 0), 6: (3, 2, 0, 1), 7: (3, 2, 1, 0)}


def find_timerev(perm, a, b, c, d, sp_tr):
    """Vectorised find_timerev; perm in {-8..7}; indices 0-based; sp_tr 0-based."""
    t = np.stack([sp_tr[a], sp_tr[b], sp_tr[c], sp_tr[d]])
    code = np.where(perm < 0, perm + 8, perm)
    out = np.empty_like(t)
    for p, order in _TR_ORDER.items():
        m = code == p
        for slot in range(4):
            out[slot, m] = t[order[slot], m]
    return out[0], out[1], out[2], out[3]


def _tbme_passes(ham, sp_2mj, sp_tr):
    """The (up to) two passes of the contraction loop, :128-134."""
    a, b, c, d = ham["abcd"].T
    h = ham["H2"]
    perm = ham["trperm"]
    yield a, b, c, d, h
    m = (sp_2mj[a] + sp_2mj[b]) != 0
    a2, b2, c2, d2 = find_timerev(perm[m], a[m], b[m], c[m], d[m], sp_tr)
    sgn = np.where(perm[m] < 0, -1.0, 1.0)        # sign(one, perm*one)
    yield a2, b2, c2, d2, sgn * h[m]


# ----------------------------------------------------------------------------
# a10 calculate_fields (src/module_fields.f90:99-189)
# ----------------------------------------------------------------------------

def calculate_fields(rho, kLR, kRL, ham, sp_2mj, sp_tr):
    n = rho.shape[0]
    G = np.zeros((n, n), dtype=complex)
    DLR = np.zeros((n, n), dtype=complex)
    DRL = np.zeros((n, n), dtype=complex)
    for a, b, c, d, h in _tbme_passes(ham, sp_2mj, sp_tr):
        f = h * (1 - ((a == c) & (b == d)))
        np.add.at(G, (a, c), h * rho[d, b])
        np.add.at(G, (a, d), -h * rho[c, b])
        np.add.at(G, (b, c), -h * rho[d, a])
        np.add.at(G, (b, d), h * rho[c, a])
        np.add.at(G, (c, a), f * rho[b, d])
        np.add.at(G, (c, b), -f * rho[a, d])
        np.add.at(G, (d, a), -f * rho[b, c])
        np.add.at(G, (d, b), f * rho[a, c])
        np.add.at(DLR, (b, a), h * kLR[d, c])
        np.add.at(DRL, (b, a), h * kRL[d, c])
        np.add.at(DLR, (d, c), f * kLR[b, a])
        np.add.at(DRL, (d, c), f * kRL[b, a])
    hsp = G + ham["H1"]
    iu = np.triu_indices(n, 1)
    DLR[iu] = -DLR.T[iu]
    DRL[iu] = -DRL.T[iu]
    return G, hsp, DLR, DRL


# ----------------------------------------------------------------------------
# a11 calculate_fields_diag (src/module_fields.f90:204-317)
# ----------------------------------------------------------------------------

def calculate_fields_diag(rho, kLR, ham, sp_2mj, sp_tr):
    n = rho.shape[0]
    G = np.zeros((n, n), dtype=complex)
    DLR = np.zeros((n, n), dtype=complex)
    for a, b, c, d, h in _tbme_passes(ham, sp_2mj, sp_tr):
        np.add.at(G, (a, c), h * rho[d, b])
        np.add.at(G, (a, d), -h * rho[c, b])
        m = (c == a) & (b != d)
        np.add.at(G, (c[m], a[m]), h[m] * rho[b[m], d[m]])
        m = b <= d
        np.add.at(G, (b[m], d[m]), h[m] * rho[c[m], a[m]])
        m = b <= c
        np.add.at(G, (b[m], c[m]), -h[m] * rho[d[m], a[m]])
        ne = (a != c) | (b != d)
        m = ne & (d <= b)
        np.add.at(G, (d[m], b[m]), h[m] * rho[a[m], c[m]])
        m = ne & (c <= b)
        np.add.at(G, (c[m], b[m]), -h[m] * rho[a[m], d[m]])
        np.add.at(DLR, (b, a), h * kLR[d, c])
        np.add.at(DLR, (d[ne], c[ne]), h[ne] * kLR[b[ne], a[ne]])
    il = np.tril_indices(n, -1)
    G[il] = np.conj(G.T[il])
    hsp = G + ham["H1"]
    iu = np.triu_indices(n, 1)
    DLR[iu] = -DLR.T[iu]
    return G, hsp, DLR, np.conj(DLR)


# ----------------------------------------------------------------------------
# a13 calculate_particle_number (src/module_particlenumber.f90:78-143)
# ----------------------------------------------------------------------------

def calculate_particle_number(rho, kLR, kRL):
    n = rho.shape[0]
    h = n // 2
    prot = np.trace(rho[:h, :h])
    neut = np.trace(rho[h:, h:])
    rp, rn = rho[:h, :h], rho[h:, h:]
    tr1 = np.trace(rp @ rp)
    tr2 = np.trace(kRL[:h, :h].T @ kLR[:h, :h])
    tr3 = np.trace(rn @ rn)
    tr4 = np.trace(kRL[h:, h:].T @ kLR[h:, h:])
    return prot, neut, prot + prot ** 2 - tr1 + tr2, neut + neut ** 2 - tr3 + tr4


# ----------------------------------------------------------------------------
# a14 calculate_expectval_energy (src/module_hamiltonian.f90:1098-1126)
# ----------------------------------------------------------------------------

def calculate_expectval_energy(rho, kRL, G, DLR, H0, H1):
    e1 = np.trace(H1 @ rho)
    eph = np.trace(G @ rho)
    epp = np.trace(DLR @ kRL.T)
    return H0 + e1 + 0.5 * (eph + epp)


# ----------------------------------------------------------------------------
# a15 calculate_H20 (src/module_fields.f90:366-391)
# ----------------------------------------------------------------------------

def calculate_H20(U, V, hsp, DLR, DRL, factor_delta=1.0):
    A2 = U.T @ hsp @ V
    A4 = U.T @ DLR @ U
    A5 = V.T @ DRL @ V
    return A2 - A2.T + (A4 - A5) * factor_delta


# ----------------------------------------------------------------------------
# a17 calculate_projected_H20 (src/module_projection.f90:729-759)
# ----------------------------------------------------------------------------

def calculate_projected_H20(pnp):
    over = pnp["over"]
    ener = pnp["ener"] / over
    A = 0.5 * (pnp["A"] - pnp["A"].T) / over
    H20 = 0.5 * (pnp["H20"] - pnp["H20"].T) / over
    AH20 = 0.5 * (pnp["AH20"] - pnp["AH20"].T) / over
    X = AH20 - ener * A + H20
    imag = ((pnp["AH20"] - ener * pnp["A"] + pnp["H20"]) / over).imag
    if np.any(imag > 1e-8):
        raise RuntimeError("The gradient has non-zero imaginary part: %g" % imag.max())
    return X.real.flatten(order="F")


# ----------------------------------------------------------------------------
# a2 project_wavefunction(iopt=0) (src/module_projection.f90:201-373) with
# a16 sum_gauge (:425-458)
# ----------------------------------------------------------------------------

def project_wavefunction(U0, V0, ham, basis, Mphip, Mphin, facpi, valZ, valN,
                         canon=None, factor_delta=1.0, keep_angles=False):
    """Returns dict with the projected scalars, pnp_A/H20/AH20, field_H20v,
    and (if keep_angles) the per-angle stage values."""
    n = U0.shape[0]
    sp_2mj, sp_tr = basis["sp_2mj"], basis["sp_tr"]
    zU, zV = U0.astype(complex), V0.astype(complex)
    cb = canon if canon is not None else construct_canonical_basis(U0, V0)
    phip, phin, weip, wein = assign_angles(Mphip, Mphin, facpi, valZ, valN)
    pnp = dict(over=0j, ener=0j, prot=0j, neut=0j, prot2=0j, neut2=0j,
               A=np.zeros((n, n), complex), H20=np.zeros((n, n), complex),
               AH20=np.zeros((n, n), complex))
    angles = []
    for k in range(phip.size):
        R = rotation_gauge(phip[k], phin[k], n)
        Ub, Vb = rotate_wavefunction(R, zU, zV)
        if max(Mphip, Mphin) > 1:
            over = calculate_overlap(cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"],
                                     cb["nocc"], cb["nemp"], cb["ovac"], cb["voveru"], cb["D"], R)
        else:
            over = 1.0 + 0j
        A, Ut, Vt = calculate_thouless(zU, zV, Ub, Vb)
        rho, kLR, kRL = calculate_densities(zU, zV, np.conj(Ut), np.conj(Vt))
        if max(Mphip, Mphin) > 1:
            G, hsp, DLR, DRL = calculate_fields(rho, kLR, kRL, ham, sp_2mj, sp_tr)
        else:
            G, hsp, DLR, DRL = calculate_fields_diag(rho, kLR, ham, sp_2mj, sp_tr)
        prot, neut, prot2, neut2 = calculate_particle_number(rho, kLR, kRL)
        ener = calculate_expectval_energy(rho, kRL, G, DLR, ham["H0"], ham["H1"])
        H20 = calculate_H20(Ut, Vt, hsp, DLR, DRL, factor_delta)
        w = weip[k] * wein[k]
        f1 = w * over
        pnp["over"] += f1
        pnp["ener"] += ener * f1
        pnp["prot"] += prot * f1
        pnp["neut"] += neut * f1
        pnp["prot2"] += prot2 * f1
        pnp["neut2"] += neut2 * f1
        pnp["A"] += A * f1
        pnp["H20"] += H20 * f1
        pnp["AH20"] += A * (f1 * ener)
        if keep_angles:
            angles.append(dict(over=over, ener=ener, prot=prot, neut=neut, prot2=prot2,
                               neut2=neut2, A=A, Ut=Ut, Vt=Vt, rho=rho, kLR=kLR, kRL=kRL,
                               G=G, hsp=hsp, DLR=DLR, DRL=DRL, H20=H20))
    pnp["H20v"] = calculate_projected_H20(pnp)
    pnp["canon"] = cb
    if keep_angles:
        pnp["angles"] = angles
    return pnp


# ----------------------------------------------------------------------------
# Operators in the qp basis (src/module_operators.f90:49-97)
# ----------------------------------------------------------------------------

def set_operator_qpbasis(part, U, V, A):
    if part == "f11":
        fac, A2, A3 = 1.0, U.T @ A @ U, V.T @ A.T @ V
    elif part == "g11":
        fac, A2, A3 = 1.0, U.T @ A @ V, V.T @ A @ U
    elif part == "f20":
        fac, A2, A3 = -1.0, U.T @ A @ V, V.T @ A.T @ U
    elif part == "g20":
        fac, A2, A3 = -1.0, U.T @ A @ U, V.T @ A @ V
    else:
        raise ValueError(part)
    A1 = A2 - A3
    return 0.5 * (A1 + fac * A1.T)


# ----------------------------------------------------------------------------
# a18 calculate_gradient (src/module_gradient.f90:79-143)
# ----------------------------------------------------------------------------

def calculate_gradient(H20v, constraint_20, lambdas, Zim1, gradient_type, eta_in, mu_in,
                       eigen_H11=None, eigen_hsp=None, info_H11=0, info_hsp=0):
    n2 = H20v.size
    n = int(round(math.sqrt(n2)))
    opt = gradient_type
    if info_H11 != 0 and gradient_type == 1:
        opt = 0
    if info_hsp != 0 and gradient_type == 2:
        opt = 0
    if opt == 0:
        eta, mu = eta_in, mu_in
    else:
        ev = eigen_H11 if opt == 1 else eigen_hsp
        if opt == 1:
            emax, emin = 4.0 * np.max(np.abs(ev)), 2.0 * np.min(np.abs(ev))
        else:
            emax, emin = 2.0 * np.max(np.abs(ev)), 2.0 * np.min(np.abs(ev))
        cond = emax / emin
        eta = (2.0 / (math.sqrt(emax) + math.sqrt(emin))) ** 2
        mu = ((math.sqrt(cond) - 1) / (math.sqrt(cond) + 1)) ** 2
    s = np.zeros(n2)
    if constraint_20 is not None and len(lambdas):
        C = np.asarray(constraint_20).reshape((len(lambdas), n2))
        s = np.asarray(lambdas) @ C
    Zi = -eta * (H20v - s) + mu * Zim1
    Zm = Zi.reshape((n, n), order="F")
    norm = math.sqrt(np.trace(Zm.T @ Zm)) / eta
    return Zi, norm, eta, mu


# ----------------------------------------------------------------------------
# a19 diagonalize_hsp_and_H11(opt=0) (src/module_gradient.f90:158-453)
#     with calculate_H11_real (src/module_fields.f90:327-353)
# ----------------------------------------------------------------------------

def diagonalize_hsp_and_H11(U0, V0, ham, basis, constraint_HO, lambdas, ncons, npair,
                            factor_delta=1.0, MphiP=2, hspLR=None, deltaRL=None):
    n = U0.shape[0]
    if MphiP == 1 and hspLR is not None:
        hspRR = hspLR.real.copy()
        deltaRR = deltaRL.real.copy()
    else:
        rho = V0 @ V0.T
        kap = V0 @ U0.T
        _, hsp, dLR, dRL = calculate_fields_diag(rho.astype(complex), kap.astype(complex), ham,
                                                 basis["sp_2mj"], basis["sp_tr"])
        hspRR, deltaRR = hsp.real.copy(), dRL.real.copy()
    A2 = U0.T @ hspRR @ U0
    A3 = V0.T @ hspRR.T @ V0
    A4 = U0.T @ deltaRR @ V0
    H11 = A2 - A3 + (A4 + A4.T) * factor_delta
    n2 = n * n
    for m in range(ncons):
        C = np.asarray(constraint_HO[m * n2:(m + 1) * n2]).reshape((n, n), order="F")
        C11 = set_operator_qpbasis("f11", U0, V0, C) if m < ncons - npair else np.zeros((n, n))
        H11 = H11 - lambdas[m] * C11
        hspRR = hspRR - lambdas[m] * C
    return np.linalg.eigvalsh(hspRR), np.linalg.eigvalsh(H11)


# ----------------------------------------------------------------------------
# Thouless evolution used by finite differences (src/module_constraints.f90:497-590)
# ----------------------------------------------------------------------------

def evolve_wavefunction(U0, V0, Z):
    n = U0.shape[0]
    L = np.linalg.cholesky(np.eye(n) + Z.T @ Z)
    Linv = np.linalg.inv(L)
    return (U0 + V0 @ Z) @ Linv.T, (V0 + U0 @ Z) @ Linv.T
