"""Host-side binding of the gfx950 library (libtaurus_gpu.so) through its C ABI.

The functions mirror the reference routines they replace -- same names,
argument meaning and error behaviour (a nonzero status raises, where the
reference stops) -- so tests and drivers read like the reference's own code:

  set_hamiltonian storage        -> Context(...)          (tg_init)
  calculate_fields               -> Context.calculate_fields
  calculate_fields_diag          -> Context.calculate_fields_diag
  ZSKPFA('U','P')                -> zskpfa
  calculate_overlap              -> calculate_overlap
  project_wavefunction(iopt=0)   -> Context.project_wavefunction
  calculate_gradient             -> calculate_gradient

There is no CPU fallback: importing works anywhere, but every call needs the
built library and a HIP device and raises otherwise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("TAURUS_GPU_LIB", os.path.join(HERE, "libtaurus_gpu.so"))

TG_FLAG_PLAIN = 1
TG_FLAG_TBME_FP32 = 2
TG_FLAG_GRAPH = 4
TG_FLAG_CONJ_PAIRS = 8

EXPORTS = [
    "tg_last_error", "tg_init", "tg_finalize", "tg_set_mode", "tg_calculate_fields",
    "tg_calculate_fields_diag", "tg_zskpfa", "tg_calculate_overlap",
    "tg_project_wavefunction", "tg_set_state", "tg_project_resident", "tg_sync",
    "tg_get_results", "tg_calculate_gradient", "tg_set_gradient", "tg_get_gradient",
    "tg_enable_timing", "tg_reset_timing", "tg_get_timing", "tg_set_constraints",
    "tg_iterate", "tg_get_state", "tg_set_states", "tg_get_results_state",
    "tg_get_gradient_state", "tg_group_create", "tg_group_destroy", "tg_group_project",
    "tg_group_iterate", "tg_group_iterate_results", "tg_project_observables",
    "tg_set_batch_limit", "tg_canonical_basis", "tg_set_gradient_type",
    "tg_get_hsp_h11_eigenvalues", "tg_diagonalize_hsp_and_H11", "tg_sym_eigenvalues",
    "tg_zskpfa_noalloc", "tg_inverse", "tg_sym_eigensystem", "tg_iterate_canonical",
]

_lib = None


class TaurusGPUError(RuntimeError):
    pass


def load_library():
    """Loads libtaurus_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise TaurusGPUError("libtaurus_gpu.so not built (run taurus_vap_amd.build)")
    lib = ctypes.CDLL(LIBPATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    lib.tg_last_error.restype = ctypes.c_char_p
    lib.tg_init.argtypes = [ctypes.POINTER(P), I, P, P, P, P, D, D, P, ctypes.c_int64, P, P, I, P, I]
    lib.tg_finalize.argtypes = [P]
    lib.tg_set_mode.argtypes = [P, I]
    lib.tg_calculate_fields.argtypes = [P] * 8
    lib.tg_calculate_fields_diag.argtypes = [P] * 7
    lib.tg_zskpfa.argtypes = [I, I, P, I, P, P]
    lib.tg_zskpfa_noalloc.argtypes = [I, I, P, I, P, P]
    lib.tg_inverse.argtypes = [I, I, P, I, P, P]
    lib.tg_calculate_overlap.argtypes = [I, I, I, D, P, P, I, I, D, P, P, P, P]
    lib.tg_project_wavefunction.argtypes = [P, I, P, P, P, I, I, D, P, I, I, D, D, D, D, P, P, P, P, P, P]
    lib.tg_set_state.argtypes = [P, P, P, P, I, I, D, P, I, I, D, D, D, D, I]
    lib.tg_project_resident.argtypes = [P]
    lib.tg_sync.argtypes = [P]
    lib.tg_get_results.argtypes = [P, P, P, P, P, P, P]
    lib.tg_calculate_gradient.argtypes = [I, I, D, D, I, I, P, P, P, I, P, P, P, P, P]
    lib.tg_set_gradient.argtypes = [P, D, D, I, P, P]
    lib.tg_get_gradient.argtypes = [P, P, P]
    lib.tg_enable_timing.argtypes = [P, I]
    lib.tg_reset_timing.argtypes = [P]
    lib.tg_get_timing.argtypes = [P, ctypes.c_char_p, P, P]
    lib.tg_set_constraints.argtypes = [P, I, I, P, I, P, D]
    lib.tg_iterate.argtypes = [P, D, P, P, P, P, P]
    lib.tg_get_state.argtypes = [P, P, P]
    lib.tg_set_states.argtypes = [P, I, P, P, P, P, P, P, P, I, I, D, D, D, D, I]
    lib.tg_get_results_state.argtypes = [P, I, P, P, P, P, P, P]
    lib.tg_get_gradient_state.argtypes = [P, I, P, P]
    lib.tg_group_create.argtypes = [ctypes.POINTER(P), I, P, I]
    lib.tg_group_destroy.argtypes = [P]
    lib.tg_group_project.argtypes = [P]
    lib.tg_group_iterate.argtypes = [P, D]
    lib.tg_group_iterate_results.argtypes = [P, I, P, P, P, P, P]
    lib.tg_project_observables.argtypes = [P, I, P, I, P, I, P, P, P]
    lib.tg_set_batch_limit.argtypes = [P, I]
    lib.tg_canonical_basis.argtypes = [P, D, P, P, P, P]
    lib.tg_iterate_canonical.argtypes = [P, D, D, P, P, P, P, P, P, P, P, P]
    lib.tg_set_gradient_type.argtypes = [P, I]
    lib.tg_get_hsp_h11_eigenvalues.argtypes = [P, I, P, P, P, P]
    lib.tg_diagonalize_hsp_and_H11.argtypes = [P, I, P, P, P, P, P, P, P]
    lib.tg_sym_eigenvalues.argtypes = [I, I, P, I, P]
    lib.tg_sym_eigensystem.argtypes = [I, I, P, I, P, P]
    _lib = lib
    return lib


def _check(status: int, what: str):
    if status != 0:
        msg = load_library().tg_last_error().decode(errors="replace")
        raise TaurusGPUError("%s failed (%d): %s" % (what, status, msg))


def _c(a, dtype):
    """Contiguous column-major copy with the given dtype."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """Device context: basis + Hamiltonian uploaded once (tg_init)."""

    def __init__(self, ndim, sp_2mj, sp_2mt, sp_l, sp_tr, H0, H1, abcd, H2, trperm,
                 plain=False, tbme_fp32=None, graph=False, conj_pairs=False):
        lib = load_library()
        self.n = int(ndim)
        H2 = np.ascontiguousarray(H2)
        if H2.dtype not in (np.float32, np.float64):
            raise ValueError("H2 must be float32 or float64")
        if tbme_fp32 is None:
            tbme_fp32 = H2.dtype == np.float32
        flags = ((TG_FLAG_PLAIN if plain else 0) | (TG_FLAG_TBME_FP32 if tbme_fp32 else 0)
                 | (TG_FLAG_GRAPH if graph else 0) | (TG_FLAG_CONJ_PAIRS if conj_pairs else 0))
        self._keep = [np.ascontiguousarray(x, dtype=np.int32) for x in (sp_2mj, sp_2mt, sp_l, sp_tr)]
        H1 = _c(H1, np.complex128)
        abcd = np.ascontiguousarray(np.asarray(abcd, dtype=np.int16).reshape(-1))
        trperm = np.ascontiguousarray(trperm, dtype=np.int8)
        ntb = trperm.size
        if abcd.size != 4 * ntb or H2.size != ntb:
            raise ValueError("abcd / H2 / trperm sizes disagree")
        h = ctypes.c_void_p()
        H0 = complex(H0)
        _check(lib.tg_init(ctypes.byref(h), self.n, *[_p(x) for x in self._keep], H0.real, H0.imag,
                           _p(H1), ntb, _p(abcd), _p(H2), H2.itemsize, _p(trperm), flags), "tg_init")
        self._h = h
        self.nstates = 1
        self.constraint_dim = 0

    def close(self):
        if getattr(self, "_h", None):
            load_library().tg_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_batch_limit(self, max_items: int):
        """At most max_items (state, angle) items per launch sequence (0: all)."""
        _check(load_library().tg_set_batch_limit(self._h, int(max_items)), "tg_set_batch_limit")

    def set_mode(self, plain: bool):
        _check(load_library().tg_set_mode(self._h, 1 if plain else 0), "tg_set_mode")

    # calculate_fields (src/module_fields.f90:99-189)
    def calculate_fields(self, rhoLR, kappaLR, kappaRL):
        n = self.n
        ins = [_c(x, np.complex128) for x in (rhoLR, kappaLR, kappaRL)]
        outs = [np.zeros((n, n), np.complex128, order="F") for _ in range(4)]
        _check(load_library().tg_calculate_fields(self._h, *[_p(x) for x in ins],
                                                  *[_p(x) for x in outs]), "tg_calculate_fields")
        return tuple(outs)   # gammaLR, hspLR, deltaLR, deltaRL

    # calculate_fields_diag (src/module_fields.f90:204-317)
    def calculate_fields_diag(self, rhoLR, kappaLR):
        n = self.n
        ins = [_c(x, np.complex128) for x in (rhoLR, kappaLR)]
        outs = [np.zeros((n, n), np.complex128, order="F") for _ in range(4)]
        _check(load_library().tg_calculate_fields_diag(self._h, *[_p(x) for x in ins],
                                                       *[_p(x) for x in outs]),
               "tg_calculate_fields_diag")
        return tuple(outs)

    # project_wavefunction(iopt=0) + calculate_projected_H20 (module_projection.f90:201, 729)
    def project_wavefunction(self, U0, V0, D0, nocc0, nemp0, ovac0, voveru0, Mphip, Mphin,
                             facpi, valence_Z, valence_N, factor_delta=1.0, iprint=0):
        n = self.n
        U0, V0, D0 = (_c(x, np.float64) for x in (U0, V0, D0))
        vou = np.ascontiguousarray(np.asarray(voveru0, dtype=np.float64).reshape(-1))
        if vou.size == 0:
            vou = np.zeros(1)
        scal = np.zeros(7, np.complex128)
        pA, pH, pAH = (np.zeros((n, n), np.complex128, order="F") for _ in range(3))
        H20v = np.zeros(n * n)
        imag = np.zeros(1)
        _check(load_library().tg_project_wavefunction(
            self._h, int(iprint), _p(U0), _p(V0), _p(D0), int(nocc0), int(nemp0), float(ovac0), _p(vou),
            int(Mphip), int(Mphin), float(facpi), float(valence_Z), float(valence_N),
            float(factor_delta), _p(scal), _p(pA), _p(pH), _p(pAH), _p(H20v), _p(imag)),
            "tg_project_wavefunction")
        return dict(over=scal[0], ener=scal[1], prot=scal[2], neut=scal[3], prot2=scal[4],
                    neut2=scal[5], pari=scal[6], A=pA, H20=pH, AH20=pAH, H20v=H20v,
                    imag=float(imag[0]))

    # resident-state API used for timing
    def set_state(self, U0, V0, D0, nocc0, nemp0, ovac0, voveru0, Mphip, Mphin, facpi,
                  valence_Z, valence_N, factor_delta=1.0, iprint=0):
        U0, V0, D0 = (_c(x, np.float64) for x in (U0, V0, D0))
        vou = np.ascontiguousarray(np.asarray(voveru0, dtype=np.float64).reshape(-1))
        if vou.size == 0:
            vou = np.zeros(1)
        self._state_keep = (U0, V0, D0, vou)
        self.nstates = 1
        self._angles = int(Mphip) * int(Mphin)
        _check(load_library().tg_set_state(self._h, _p(U0), _p(V0), _p(D0), int(nocc0), int(nemp0),
                                           float(ovac0), _p(vou), int(Mphip), int(Mphin),
                                           float(facpi), float(valence_Z), float(valence_N),
                                           float(factor_delta), int(iprint)), "tg_set_state")

    def set_states(self, states, Mphip, Mphin, facpi, valence_Z, valence_N, factor_delta=1.0,
                   iprint=0):
        """Several independent states of this basis, projected together.
        states: list of (U0, V0, canon) with canon = construct_canonical_basis(U0, V0)."""
        n = self.n
        S = len(states)
        U = np.zeros((S, n, n))
        V = np.zeros((S, n, n))
        D = np.zeros((S, n, n))
        vou = np.zeros((S, n))
        nocc = np.zeros(S, np.int32)
        nemp = np.zeros(S, np.int32)
        ovac = np.zeros(S)
        for k, (Uk, Vk, cn) in enumerate(states):
            # state k column-major at + k n^2: the transposed C-order slab
            U[k] = np.asarray(Uk, dtype=np.float64).T
            V[k] = np.asarray(Vk, dtype=np.float64).T
            D[k] = np.asarray(cn["D"], dtype=np.float64).T
            v = np.asarray(cn["voveru"], dtype=np.float64).reshape(-1)
            vou[k, :v.size] = v
            nocc[k], nemp[k], ovac[k] = int(cn["nocc"]), int(cn["nemp"]), float(cn["ovac"])
        self._state_keep = (U, V, D, vou, nocc, nemp, ovac)
        _check(load_library().tg_set_states(self._h, S, _p(U), _p(V), _p(D), _p(nocc), _p(nemp),
                                            _p(ovac), _p(vou), int(Mphip), int(Mphin), float(facpi),
                                            float(valence_Z), float(valence_N), float(factor_delta),
                                            int(iprint)), "tg_set_states")
        self.nstates = S

    def project_resident(self):
        _check(load_library().tg_project_resident(self._h), "tg_project_resident")

    def sync(self):
        _check(load_library().tg_sync(self._h), "tg_sync")

    def get_results(self, state=0):
        n = self.n
        scal = np.zeros(7, np.complex128)
        H20v = np.zeros(n * n)
        imag = np.zeros(1)
        _check(load_library().tg_get_results_state(self._h, int(state), _p(scal), None, None, None,
                                                   _p(H20v), _p(imag)), "tg_get_results")
        return dict(over=scal[0], ener=scal[1], prot=scal[2], neut=scal[3], prot2=scal[4],
                    neut2=scal[5], pari=scal[6], H20v=H20v, imag=float(imag[0]))

    # calculate_gradient (gradient_type 0) on the device after each resident step
    def set_gradient(self, gradient_eta, gradient_mu, constraint_20=None, lagrange_lambda1=None):
        lam = np.ascontiguousarray(np.asarray(lagrange_lambda1 if lagrange_lambda1 is not None else [],
                                              dtype=np.float64).reshape(-1))
        ncons = lam.size
        c20 = (np.ascontiguousarray(constraint_20, dtype=np.float64).reshape(-1) if ncons
               else np.zeros(1))
        if ncons == 0:
            lam = np.zeros(1)
        _check(load_library().tg_set_gradient(self._h, float(gradient_eta), float(gradient_mu), ncons,
                                              _p(c20), _p(lam)), "tg_set_gradient")

    def get_gradient(self, state=0):
        Zi = np.zeros(self.n * self.n)
        norm = np.zeros(1)
        _check(load_library().tg_get_gradient_state(self._h, int(state), _p(Zi), _p(norm)),
               "tg_get_gradient")
        return Zi, float(norm[0])

    # device-resident gradient iteration (gradient_type 0)
    def set_constraints(self, constraint_HO, constraint_val, constraint_pair=0,
                        constraint_eps=1e-6):
        """constraint_val: one set (constraint_dim values) for every state, or
        one row per state (nstates x constraint_dim)."""
        val2 = np.atleast_2d(np.asarray(constraint_val, dtype=np.float64))
        nsets, dim = val2.shape
        val = np.ascontiguousarray(val2.reshape(-1))
        HO = (np.ascontiguousarray(constraint_HO, dtype=np.float64).reshape(-1) if dim
              else np.zeros(1))
        if dim and HO.size != dim * self.n * self.n:
            raise ValueError("constraint_HO must hold constraint_dim * ndim**2 values")
        self._cons_keep = (HO, val)
        _check(load_library().tg_set_constraints(self._h, dim, int(constraint_pair), _p(HO),
                                                 max(1, nsets), _p(val if dim else np.zeros(1)),
                                                 float(constraint_eps)), "tg_set_constraints")
        self.constraint_dim = dim

    def _iterate_buffers(self):
        S, dim = self.nstates, self.constraint_dim
        return (np.zeros(7 * S, np.complex128), np.zeros(S), np.zeros(max(1, S * dim)),
                np.zeros(S, np.int32), np.zeros(S, np.int32))

    def _iterate_dicts(self, scal, gn, lam, nadj, conv):
        S, dim = self.nstates, self.constraint_dim
        out = []
        for k in range(S):
            sc = scal[7 * k:7 * k + 7]
            out.append(dict(over=sc[0], ener=sc[1], prot=sc[2], neut=sc[3], prot2=sc[4],
                            neut2=sc[5], pari=sc[6], gradient_norm=float(gn[k]),
                            lagrange_lambda1=lam[k * dim:(k + 1) * dim].copy(),
                            adjust_iterations=int(nadj[k]), converged=bool(conv[k])))
        return out

    def iterate(self, gradient_eps=0.0):
        """One iteration of the gradient loop on the device (projection of the
        current states, Lagrange multipliers, gradient, constraint adjustment).
        Returns, per state, the projected scalars, the gradient norm, the
        multipliers and the number of adjustment steps (a dict for a single
        state, else a list of dicts); the states on the device become the
        evolved ones (read them with get_state)."""
        bufs = self._iterate_buffers()
        _check(load_library().tg_iterate(self._h, float(gradient_eps), *[_p(b) for b in bufs]),
               "tg_iterate")
        out = self._iterate_dicts(*bufs)
        return out[0] if self.nstates == 1 else out

    def iterate_canonical(self, canon_eps=1.0e-8, gradient_eps=0.0):
        """canonical_basis(canon_eps) followed by iterate(gradient_eps) in one call
        (tg_iterate_canonical); the canonical basis runs beside the projection's
        Thouless / field chain.  Returns (iterate results, canonical_basis list)."""
        n, S = self.n, self.nstates
        nocc = np.zeros(S, np.int32)
        nemp = np.zeros(S, np.int32)
        ovac = np.zeros(S)
        vou = np.zeros(S * n)
        bufs = self._iterate_buffers()
        _check(load_library().tg_iterate_canonical(self._h, float(canon_eps), float(gradient_eps),
                                                   *[_p(b) for b in bufs], _p(nocc), _p(nemp),
                                                   _p(ovac), _p(vou)), "tg_iterate_canonical")
        out = self._iterate_dicts(*bufs)
        canon = [dict(nocc=int(nocc[k]), nemp=int(nemp[k]), ovac=float(ovac[k]),
                      voveru=vou[k * n:k * n + n - int(nocc[k])].copy()) for k in range(S)]
        return (out[0] if S == 1 else out), canon

    def set_gradient_type(self, gradient_type):
        """gradient_type of the device iteration: 0 (eta, mu as given), 1 / 2 (the
        recipes from the H11 / h spectra, diagonalize_hsp_and_H11 on the device)."""
        _check(load_library().tg_set_gradient_type(self._h, int(gradient_type)),
               "tg_set_gradient_type")

    def hsp_h11_eigenvalues(self, state=0):
        """(eigen_hsp, eigen_H11, eta, mu) of the last gradient_type 1 / 2 iteration."""
        eh = np.zeros(self.n)
        e11 = np.zeros(self.n)
        em = np.zeros(2)
        _check(load_library().tg_get_hsp_h11_eigenvalues(self._h, int(state), _p(eh), _p(e11),
                                                         _p(em[:1]), _p(em[1:])),
               "tg_get_hsp_h11_eigenvalues")
        return eh, e11, float(em[0]), float(em[1])

    def diagonalize_hsp_and_H11(self, lagrange_lambda1=None):
        """diagonalize_hsp_and_H11(0) of the resident state: (eigen_hsp, eigen_H11,
        hsp_RR, field_H11) with the constraints of set_constraints."""
        dim = getattr(self, "constraint_dim", 0)
        lam = np.ascontiguousarray(np.asarray(lagrange_lambda1 if dim else [0.0], dtype=np.float64))
        n = self.n
        eh, e11 = np.zeros(n), np.zeros(n)
        hm = np.zeros((n, n), order="F")
        h11 = np.zeros((n, n), order="F")
        info = np.zeros(2, np.int32)
        _check(load_library().tg_diagonalize_hsp_and_H11(self._h, dim, _p(lam), _p(eh), _p(e11),
                                                         _p(info[:1]), _p(info[1:]), _p(hm),
                                                         _p(h11)), "tg_diagonalize_hsp_and_H11")
        return eh, e11, hm, h11

    def get_state(self):
        """(U, V) of the device state, or lists of them for several states."""
        n, S = self.n, self.nstates
        U = np.zeros(S * n * n)
        V = np.zeros(S * n * n)
        _check(load_library().tg_get_state(self._h, _p(U), _p(V)), "tg_get_state")
        Us = [U[k * n * n:(k + 1) * n * n].reshape((n, n), order="F") for k in range(S)]
        Vs = [V[k * n * n:(k + 1) * n * n].reshape((n, n), order="F") for k in range(S)]
        return (Us[0], Vs[0]) if S == 1 else (Us, Vs)

    def canonical_basis(self, eps=1.0e-8):
        """construct_canonical_basis of every resident state on the device
        (tg_canonical_basis): the states are set for the next projection or
        iteration without leaving the device.  Returns a list of dict(nocc, nemp,
        ovac, voveru), one per state.  Raises TaurusGPUError (code -7) when a
        degenerate kappa subspace needs its real Schur form (host route)."""
        n, S = self.n, self.nstates
        nocc = np.zeros(S, np.int32)
        nemp = np.zeros(S, np.int32)
        ovac = np.zeros(S)
        vou = np.zeros(S * n)
        _check(load_library().tg_canonical_basis(self._h, float(eps), _p(nocc), _p(nemp), _p(ovac),
                                                 _p(vou)), "tg_canonical_basis")
        return [dict(nocc=int(nocc[k]), nemp=int(nemp[k]), ovac=float(ovac[k]),
                     voveru=vou[k * n:k * n + n - int(nocc[k])].copy()) for k in range(S)]

    def project_observables(self, obo_ops=(), obos_pairs=(), sp_shell=None, nshell=0):
        """Per-angle observables of the final projection (project_wavefunction
        iopt = 1) of the resident state: returns (records A x R complex, pnp_rho)
        with the record layout of tg_project_observables."""
        n = self.n
        obo = np.ascontiguousarray(np.asarray([np.asarray(o, float).reshape(n, n).T for o in obo_ops]
                                              or np.zeros((1, n, n))))
        obs = np.ascontiguousarray(np.asarray([[np.asarray(o, float).reshape(n, n).T,
                                                 np.asarray(o2, float).reshape(n, n).T]
                                                for o, o2 in obos_pairs] or np.zeros((1, 2, n, n))))
        shell = np.ascontiguousarray(np.asarray(sp_shell if sp_shell is not None else np.zeros(n),
                                                dtype=np.int32))
        A = self._angles
        R = 19 + 2 * nshell + 2 * len(obo_ops) + 4 * len(obos_pairs)
        rec = np.zeros((A, R), np.complex128)
        rho = np.zeros((n, n), np.complex128, order="F")
        _check(load_library().tg_project_observables(self._h, len(obo_ops), _p(obo), len(obos_pairs),
                                                     _p(obs), int(nshell), _p(shell), _p(rec),
                                                     _p(rho)), "tg_project_observables")
        return rec, rho

    def enable_timing(self, on=True):
        _check(load_library().tg_enable_timing(self._h, 1 if on else 0), "tg_enable_timing")

    def reset_timing(self):
        _check(load_library().tg_reset_timing(self._h), "tg_reset_timing")

    def timing(self, key):
        """(summed ms, number of intervals) recorded under key."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        _check(load_library().tg_get_timing(self._h, key.encode(), ctypes.byref(ms), ctypes.byref(cnt)),
               "tg_get_timing")
        return ms.value, cnt.value


class Group:
    """Contexts (of any basis sizes) whose projections are issued together as
    one hipGraph launch (graph=True) or eagerly on their own streams."""

    def __init__(self, contexts, graph=True):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h for c in self.contexts])
        h = ctypes.c_void_p()
        _check(load_library().tg_group_create(ctypes.byref(h), len(self.contexts),
                                              ctypes.cast(arr, ctypes.c_void_p), 1 if graph else 0),
               "tg_group_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().tg_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def project(self):
        _check(load_library().tg_group_project(self._h), "tg_group_project")

    def iterate(self, gradient_eps=0.0):
        """tg_iterate of every context around one group projection; returns, per
        context, the list of per-state result dicts."""
        _check(load_library().tg_group_iterate(self._h, float(gradient_eps)), "tg_group_iterate")
        res = []
        for i, c in enumerate(self.contexts):
            bufs = c._iterate_buffers()
            _check(load_library().tg_group_iterate_results(self._h, i, *[_p(b) for b in bufs]),
                   "tg_group_iterate_results")
            res.append(c._iterate_dicts(*bufs))
        return res


def zskpfa(A):
    """ZSKPFA('U','P') on one matrix or a stack (count, n, n); returns pf (and info)."""
    A = np.asarray(A, dtype=np.complex128)
    single = A.ndim == 2
    if single:
        A = A[None]
    cnt, n, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))   # column-major per matrix
    pf = np.zeros(cnt, np.complex128)
    info = np.zeros(cnt, np.int32)
    _check(load_library().tg_zskpfa(n, cnt, _p(buf), max(1, n), _p(pf), _p(info)), "tg_zskpfa")
    return (pf[0], int(info[0])) if single else (pf, info)


def zskpfa_noalloc(A):
    """zskpfa through tg_zskpfa_noalloc (no device allocation: the library's static
    per-device scratch array; limited to TG_NOALLOC_SCRATCH_BYTES per call)."""
    A = np.asarray(A, dtype=np.complex128)
    single = A.ndim == 2
    if single:
        A = A[None]
    cnt, n, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    pf = np.zeros(cnt, np.complex128)
    info = np.zeros(cnt, np.int32)
    _check(load_library().tg_zskpfa_noalloc(n, cnt, _p(buf), max(1, n), _p(pf), _p(info)),
           "tg_zskpfa_noalloc")
    return (pf[0], int(info[0])) if single else (pf, info)


def inverse(A):
    """tg_inverse on one complex matrix or a stack (count, n, n): (inverses, det, info)."""
    A = np.asarray(A, dtype=np.complex128)
    single = A.ndim == 2
    if single:
        A = A[None]
    cnt, n, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    det = np.zeros(cnt, np.complex128)
    info = np.zeros(cnt, np.int32)
    _check(load_library().tg_inverse(n, cnt, _p(buf), max(1, n), _p(det), _p(info)), "tg_inverse")
    inv = np.transpose(buf, (0, 2, 1))
    return (inv[0], det[0], int(info[0])) if single else (inv, det, info)


def sym_eigenvalues(A):
    """Eigenvalues (ascending) of one real symmetric matrix or a stack (count, n, n),
    upper triangles read (tg_sym_eigenvalues)."""
    A = np.asarray(A, dtype=np.float64)
    single = A.ndim == 2
    if single:
        A = A[None]
    cnt, n, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    w = np.zeros(cnt * n)
    _check(load_library().tg_sym_eigenvalues(n, cnt, _p(buf), max(1, n), _p(w)),
           "tg_sym_eigenvalues")
    w = w.reshape(cnt, n)
    return w[0] if single else w


def sym_eigensystem(A):
    """Eigenvalues (ascending) and orthonormal eigenvectors (columns) of one real
    symmetric matrix or a stack (count, n, n), upper triangles read
    (tg_sym_eigensystem).  Returns (w, Z)."""
    A = np.asarray(A, dtype=np.float64)
    single = A.ndim == 2
    if single:
        A = A[None]
    cnt, n, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    w = np.zeros(cnt * n)
    Z = np.zeros((cnt, n, n))
    _check(load_library().tg_sym_eigensystem(n, cnt, _p(buf), max(1, n), _p(w), _p(Z)),
           "tg_sym_eigensystem")
    w = w.reshape(cnt, n)
    Z = np.transpose(Z, (0, 2, 1))
    return (w[0], Z[0]) if single else (w, Z)


def calculate_overlap(nocc0, nemp0, ovac0, voveru0, D0, nocc1, nemp1, ovac1, voveru1, D1, ROT):
    """calculate_overlap (src/module_projection.f90:782-937)."""
    n = np.asarray(D0).shape[0]
    D0, D1, ROT = (_c(x, np.complex128) for x in (D0, D1, ROT))
    v0 = np.ascontiguousarray(np.asarray(voveru0, dtype=np.float64).reshape(-1))
    v1 = np.ascontiguousarray(np.asarray(voveru1, dtype=np.float64).reshape(-1))
    if v0.size == 0:
        v0 = np.zeros(1)
    if v1.size == 0:
        v1 = np.zeros(1)
    out = np.zeros(1, np.complex128)
    _check(load_library().tg_calculate_overlap(n, int(nocc0), int(nemp0), float(ovac0), _p(v0), _p(D0),
                                               int(nocc1), int(nemp1), float(ovac1), _p(v1), _p(D1),
                                               _p(ROT), _p(out)), "tg_calculate_overlap")
    return complex(out[0])


def calculate_gradient(field_H20v, constraint_20, lagrange_lambda1, gradient_Zim1,
                       gradient_type, gradient_eta, gradient_mu, eigen_H11=None, eigen_hsp=None,
                       info_H11=0, info_hsp=0):
    """calculate_gradient (src/module_gradient.f90:79-143); returns (Zi, norm)."""
    H = np.ascontiguousarray(field_H20v, dtype=np.float64)
    n = int(round(np.sqrt(H.size)))
    lam = np.ascontiguousarray(np.asarray(lagrange_lambda1 if lagrange_lambda1 is not None else [],
                                          dtype=np.float64).reshape(-1))
    ncons = lam.size
    c20 = (np.ascontiguousarray(constraint_20, dtype=np.float64).reshape(-1) if ncons else np.zeros(1))
    if ncons == 0:
        lam = np.zeros(1)
    Zm = np.ascontiguousarray(gradient_Zim1, dtype=np.float64)
    Zi = np.zeros(n * n)
    norm = np.zeros(1)
    e11 = None if eigen_H11 is None else np.ascontiguousarray(eigen_H11, dtype=np.float64)
    ehs = None if eigen_hsp is None else np.ascontiguousarray(eigen_hsp, dtype=np.float64)
    _check(load_library().tg_calculate_gradient(n, int(gradient_type), float(gradient_eta),
                                                float(gradient_mu), int(info_H11), int(info_hsp),
                                                _p(e11), _p(ehs), _p(H), ncons, _p(c20), _p(lam),
                                                _p(Zm), _p(Zi), _p(norm)), "tg_calculate_gradient")
    return Zi, float(norm[0])
