"""Host side of one PNVAP gradient iteration around the device library.

The reference computes the canonical basis of the current state on the host
once per iteration (construct_canonical_basis, src/module_wavefunctions.f90:
777-945, called from project_wavefunction :219-234) and passes it into the
gauge-angle loop.  This module does the same for the drivers of this package
(tests, bench.py): it is host code that runs once per iteration, outside the
angle loop; all per-angle work runs in libtaurus_gpu.so.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

from . import gpu
from . import hamiltonian as H
from . import inputs as I


def construct_canonical_basis(U, V, eps: float = 1.0e-8, allemp: int = 0) -> dict:
    """Canonical (BMZ) basis of the real Bogoliubov state (U, V):
    D diagonalises rho = V V^T (ascending eigenvalues), degenerate blocks of
    kappa are brought to their real Schur form, nocc/nemp count the fully
    occupied/empty states with cut eps, ovac = sum log u over the paired ones,
    voveru the v/u of the paired states as read by the overlap (upper entries).
    Returns dict(D, nocc, nemp, ovac, voveru)."""
    n = U.shape[0]
    rho = V @ V.T
    _, D = np.linalg.eigh(rho)
    rhoc = D.T @ rho @ D
    v2 = np.abs(np.diag(rhoc))
    occ = v2 >= 1.0 - eps
    emp = (v2 <= eps) & (allemp == 0) & ~occ
    nocc, nemp = int(occ.sum()), int(emp.sum())
    paired = ~(occ | emp)
    ovac = float(np.sum(np.log(np.sqrt(np.abs(1.0 - v2[paired]))))) if paired.any() else 0.0
    kap = D.T @ (V @ U.T) @ D
    lo, hi = nemp, n - nocc
    # blocks of kappa wider than one canonical pair: extent of the coupling
    thr = math.sqrt(eps) * 1.0e-2
    reach = np.arange(n)
    for i in range(lo, hi):
        nz = np.nonzero(np.abs(kap[i + 1:hi, i]) > thr)[0]
        reach[i] = i + 1 + nz[-1] if nz.size else i
    k = lo
    T = np.eye(n)
    changed = False
    while k < hi - 1:
        j = reach[k]
        m = reach[k:j + 1].max()
        while m > j:
            j = m
            m = reach[k:j + 1].max()
        size = j - k + 1
        if size > 2 and size % 2 != 0:
            size = 2
        if size > 2:
            B = kap[k:k + size, k:k + size]
            S, Z = sla.schur(B, output="real")
            T[k:k + size, k:k + size] = Z
            kap[k:k + size, k:k + size] = S
            changed = True
        k += max(size, 1)
    if changed:
        D = D @ T
    vou = np.zeros(n - nocc)
    par = 0
    for i in range(lo, hi):
        u = math.sqrt(abs(1.0 - v2[i]))
        v = math.sqrt(v2[i])
        if par % 2 == 0:
            vou[i] = math.copysign(v, kap[i, i + 1]) / u
        else:
            vou[i] = math.copysign(v, kap[i, i - 1]) / u
        par += 1
    if n - nocc - nemp == 0:
        ovac = 0.0
    return dict(D=D, nocc=nocc, nemp=nemp, ovac=ovac, voveru=vou)


def operator_qp20(U, V, A):
    """'f20' part of a one-body operator in the qp basis (set_operator_qpbasis,
    src/module_operators.f90:73-96): B = 1/2 (A1 - A1^T), A1 = U^T A V - V^T A^T U."""
    A1 = U.T @ A @ V - V.T @ A.T @ U
    return 0.5 * (A1 - A1.T)


def evolve(U, V, Z):
    """Thouless evolution U1 = (U + V Z) L^-T, V1 = (V + U Z) L^-T, L L^T = 1 + Z^T Z
    (cholesky_factorization + evolve_wavefunction, src/module_constraints.f90:497-590)."""
    n = U.shape[0]
    L = np.linalg.cholesky(np.eye(n) + Z.T @ Z)
    Li = sla.solve_triangular(L, np.eye(n), lower=True)
    return (U + V @ Z) @ Li.T, (V + U @ Z) @ Li.T


def adjust_constraints(U, V, ops, vals, Z=None, eps: float = 1.0e-6, maxit: int = 100):
    """adjust_constraints (src/module_constraints.f90:270-373) for one-body
    constraints: evolve with Z, then add the constraint directions that solve
    A x = b, A_ij = Tr(C20_i^T C20_j), b_i = val_i - <O_i>, until |b| <= eps.
    Returns (U1, V1, Z, iterations)."""
    n = U.shape[0]
    Z = np.zeros((n, n)) if Z is None else Z.copy()
    for it in range(1, maxit + 1):
        U1, V1 = evolve(U, V, Z)
        rho = V1 @ V1.T
        b = np.array([val - np.sum(rho * O.T) for O, val in zip(ops, vals)])
        C = [operator_qp20(U1, V1, O) for O in ops]
        if np.all(np.abs(b) <= eps):
            return U1, V1, Z, it
        A = np.array([[np.trace(Ci.T @ Cj) for Cj in C] for Ci in C])
        x = np.linalg.solve(A, b)
        for xm, Cm in zip(x, C):
            Z = Z + xm * Cm
    return U1, V1, Z, maxit


def number_operators(basis: I.Basis):
    n = basis.dim
    Zop = np.diag((basis.sp_2mt < 0).astype(np.float64))
    Nop = np.diag((basis.sp_2mt > 0).astype(np.float64))
    return Zop, Nop


def seed_state(w: I.Workload, basis: I.Basis, seed: int):
    """Seed state of a workload: real Thouless rotation of the lowest-filled
    Slater determinant (inputs.seed_wavefunction, sigma per workload), brought
    to <Z> = Z, <N> = N by the particle-number constraints."""
    U, V = I.seed_wavefunction(basis, w.Z, w.N, seed=seed + 1, sigma=w.seed_sigma)
    Zop, Nop = number_operators(basis)
    U, V, _, _ = adjust_constraints(U, V, [Zop, Nop], [float(w.Z), float(w.N)])
    return U, V


def workload_context(name: str, seed: int = 1234, tbme_fp32: bool = True, plain: bool = False,
                     ham: dict | None = None, graph: bool = False, conj_pairs: bool = False):
    """Device context + seed state of a named synthetic workload."""
    w = I.WORKLOADS[name]
    h = ham if ham is not None else H.workload_hamiltonian(w, seed)
    b = h["basis"]
    H2 = h["H2"].astype(np.float32) if tbme_fp32 else h["H2"]
    ctx = gpu.Context(b.dim, b.sp_2mj, b.sp_2mt, b.sp_l, b.sp_tr, h["H0"], h["H1"], h["abcd"], H2,
                      h["trperm"], plain=plain, tbme_fp32=tbme_fp32, graph=graph,
                      conj_pairs=conj_pairs)
    U, V = seed_state(w, b, seed)
    return ctx, h, w, U, V
