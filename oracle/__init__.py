"""CPU oracle for rchem's RHF hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import it; the shipped library
(``rchem_amd``) never does, and fails loudly when its HIP library is missing
instead of falling back to anything here.

Contents
  * ctypes bindings to ``liboracle.so`` (``oracle.c``: the C restatement of the
    reference's integrals, Boys functions, JK_direct and build_I loops);
  * ``ref_lib()``: the reference's own ``libpyquante2`` compiled from
    ``/root/reference/libpyquante2/{cints,chgp}.c`` into ``oracle/_ref`` (only
    present where the reference tree was mounted at build time);
  * ``cgtos_from_bse``: rchem's ``Basis::new`` expansion of BSE-JSON shells into
    Cartesian contracted functions (basis.rs:182-212, shell.rs:1-12);
  * ``scf``: rchem's SCF driver restated in numpy (src/bin/rchem.rs:40-127).
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BOYS_EXACT = 0
BOYS_COMPAT = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

_lib = None
_ref = None


def lib():
    """liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} missing (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_fgamma_compat.restype = ctypes.c_double
        L.orc_fgamma_compat.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_boys_exact.restype = ctypes.c_double
        L.orc_boys_exact.argtypes = [ctypes.c_int, ctypes.c_double]
        for fn in ("orc_coulomb_repulsion", "orc_hgp_primitive"):
            f = getattr(L, fn)
            f.restype = ctypes.c_double
            f.argtypes = [_dp, _ip, _dp, _dp, ctypes.c_int]
        for fn in ("orc_overlap", "orc_kinetic"):
            f = getattr(L, fn)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_double, _ip, _dp, ctypes.c_double, _ip, _dp]
        L.orc_nuclear_attraction.restype = ctypes.c_double
        L.orc_nuclear_attraction.argtypes = [ctypes.c_double, _ip, _dp, ctypes.c_double, _ip,
                                             _dp, _dp, ctypes.c_int]
        L.orc_pgto_norm.restype = ctypes.c_double
        L.orc_pgto_norm.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        basis_args = [ctypes.c_int, _dp, _ip, _ip, _dp, _dp]
        L.orc_contracted_eri.restype = ctypes.c_double
        L.orc_contracted_eri.argtypes = basis_args + [ctypes.c_int] * 6
        L.orc_eri_tensor.restype = None
        L.orc_eri_tensor.argtypes = basis_args + [ctypes.c_int, _dp]
        L.orc_jk_direct_rows.restype = None
        L.orc_jk_direct_rows.argtypes = basis_args + [_dp, ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_jk_direct_rows_mt.restype = None
        L.orc_jk_direct_rows_mt.argtypes = basis_args + [_dp, ctypes.c_int, ctypes.c_void_p,
                                                          ctypes.c_int, ctypes.c_int,
                                                          ctypes.c_int, _dp, _dp]
        L.orc_jk_direct_sample_mt.restype = None
        L.orc_jk_direct_sample_mt.argtypes = basis_args + [_dp, ctypes.c_int, ctypes.c_void_p,
                                                            ctypes.c_int, _ip, ctypes.c_int,
                                                            _dp, _dp]
        L.orc_jk_direct_sample.restype = None
        L.orc_jk_direct_sample.argtypes = basis_args + [_dp, ctypes.c_int, ctypes.c_void_p,
                                                        ctypes.c_int, _ip, _dp, _dp]
        L.orc_one_electron.restype = None
        L.orc_one_electron.argtypes = basis_args + [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_nuclear_repulsion.restype = ctypes.c_double
        L.orc_nuclear_repulsion.argtypes = [ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def ref_lib():
    """The reference's libpyquante2 built under oracle/_ref, or None if absent."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libpyquante2_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.Fgamma.restype = ctypes.c_double
        R.Fgamma.argtypes = [ctypes.c_double, ctypes.c_double]
        d, i = ctypes.c_double, ctypes.c_int
        R.coulomb_repulsion.restype = d
        R.coulomb_repulsion.argtypes = [d, d, d, d, i, i, i, d] * 4
        R.hrr.restype = d
        R.hrr.argtypes = [d, d, d, d, i, i, i, d] * 4
        R.overlap.restype = d
        R.overlap.argtypes = [d, i, i, i, d, d, d] * 2
        R.kinetic.restype = d
        R.kinetic.argtypes = [d, i, i, i, d, d, d] * 2
        R.nuclear_attraction.restype = d
        R.nuclear_attraction.argtypes = [d, d, d, d, i, i, i, d] * 2 + [d, d, d]
        _ref = R
    return _ref


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


# --------------------------------------------------------------------------
# Scalar restatements.
# --------------------------------------------------------------------------

def fgamma_compat(m, x):
    """cints.c:302-373 Fgamma (NR incomplete gamma, EPS = 3e-7)."""
    return lib().orc_fgamma_compat(float(m), float(x))


def boys_exact(m, x):
    return lib().orc_boys_exact(int(m), float(x))


def coulomb_repulsion(centres, powers, alphas, norms, boys=BOYS_COMPAT, hgp=False):
    """Primitive (ab|cd) with norms; THO (cints.c:72-149) or HGP (chgp.c)."""
    r, rp = _d(np.asarray(centres).reshape(12))
    p, pp = _i(np.asarray(powers).reshape(12))
    a, ap = _d(alphas)
    n, np_ = _d(norms)
    f = lib().orc_hgp_primitive if hgp else lib().orc_coulomb_repulsion
    return f(rp, pp, ap, np_, boys)


def overlap(za, zb, ra, rb, c):
    """Unnormalised primitive overlap, os86::get_overlap semantics (os86.rs:629)."""
    p1, pp1 = _i(c[:3]); p2, pp2 = _i(c[3:])
    A, Ap = _d(ra); B, Bp = _d(rb)
    return lib().orc_overlap(za, pp1, Ap, zb, pp2, Bp)


def kinetic(za, zb, ra, rb, c):
    p1, pp1 = _i(c[:3]); p2, pp2 = _i(c[3:])
    A, Ap = _d(ra); B, Bp = _d(rb)
    return lib().orc_kinetic(za, pp1, Ap, zb, pp2, Bp)


def nuclear(za, zb, ra, rb, rc, c, boys=BOYS_EXACT):
    p1, pp1 = _i(c[:3]); p2, pp2 = _i(c[3:])
    A, Ap = _d(ra); B, Bp = _d(rb); C, Cp = _d(rc)
    return lib().orc_nuclear_attraction(za, pp1, Ap, zb, pp2, Bp, Cp, boys)


# --------------------------------------------------------------------------
# rchem's basis construction (Basis::new, basis.rs:182-212).
# --------------------------------------------------------------------------

def get_ijk_list(m):
    """Cartesian component order (shell.rs:1-12)."""
    out = []
    for a in range(1, m + 2):
        for b in range(1, a + 1):
            out.append((m + 1 - a, a - b, b - 1))
    return out


@dataclass
class CGTOBasis:
    """Cartesian contracted functions in rchem's order (basis.rs:152-173)."""
    centre: np.ndarray      # (n, 3)
    powers: np.ndarray      # (n, 3) int
    prim_off: np.ndarray    # (n + 1,) int
    exps: np.ndarray
    coefs: np.ndarray

    @property
    def n(self):
        return len(self.powers)

    def args(self):
        keep = []
        out = [self.n]
        for arr, conv in ((self.centre, _d), (self.powers, _i), (self.prim_off, _i),
                          (self.exps, _d), (self.coefs, _d)):
            a, p = conv(arr)
            keep.append(a)
            out.append(p)
        return out, keep


def load_bse_json(path):
    with open(path) as fh:
        return json.load(fh)


def cgtos_from_bse(bse, atomnos, coords):
    """rchem Basis::new: per atom, per BSE shell, per l in angular_momentum,
    per get_ijk_list(l) component, a CGTO with coefficients[l] taken raw
    (basis.rs:186-205).  Coordinates are used as given (rchem.rs:33-35)."""
    centre, powers, off, exps, coefs = [], [], [0], [], []
    for a, Z in enumerate(atomnos):
        el = bse["elements"][str(int(Z))]
        for shell in el["electron_shells"]:
            ex = [float(x) for x in shell["exponents"]]
            for l in shell["angular_momentum"]:
                cf = [float(x) for x in shell["coefficients"][l]]
                assert len(cf) == len(ex)
                for pw in get_ijk_list(l):
                    centre.append(coords[a])
                    powers.append(pw)
                    exps.extend(ex)
                    coefs.extend(cf)
                    off.append(len(exps))
    return CGTOBasis(np.array(centre, dtype=np.float64).reshape(-1, 3),
                     np.array(powers, dtype=np.int32).reshape(-1, 3),
                     np.array(off, dtype=np.int32), np.array(exps), np.array(coefs))


def cgtos_from_shells(centres, ls, nprims, exps, coefs):
    """The same expansion for explicit shells (centre, l, exponents, coefs)."""
    centre, powers, off, ex_all, cf_all = [], [], [0], [], []
    p = 0
    for s, l in enumerate(ls):
        ex = list(exps[p:p + nprims[s]])
        cf = list(coefs[p:p + nprims[s]])
        p += nprims[s]
        for pw in get_ijk_list(int(l)):
            centre.append(centres[s])
            powers.append(pw)
            ex_all.extend(ex)
            cf_all.extend(cf)
            off.append(len(ex_all))
    return CGTOBasis(np.array(centre, dtype=np.float64).reshape(-1, 3),
                     np.array(powers, dtype=np.int32).reshape(-1, 3),
                     np.array(off, dtype=np.int32), np.array(ex_all, dtype=np.float64),
                     np.array(cf_all, dtype=np.float64))


# --------------------------------------------------------------------------
# Matrix-level restatements.
# --------------------------------------------------------------------------

def one_electron(basis: CGTOBasis, atomnos, coords):
    """S, T, V (basis.rs:214-338)."""
    n = basis.n
    args, keep = basis.args()
    xyz, xp = _d(np.asarray(coords, dtype=np.float64).reshape(-1, 3))
    Z, Zp = _d(np.asarray(atomnos, dtype=np.float64))
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    lib().orc_one_electron(*args, len(atomnos), xp, Zp, S.ctypes.data_as(_dp),
                           T.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return S, T, V


def eri_tensor(basis: CGTOBasis, boys=BOYS_COMPAT):
    """build_I (basis.rs:430-460): full [mu,nu,la,si] tensor."""
    n = basis.n
    args, keep = basis.args()
    I = np.zeros((n, n, n, n))
    lib().orc_eri_tensor(*args, boys, I.ctypes.data_as(_dp))
    return I


def contracted_eri(basis: CGTOBasis, i, j, k, l, boys=BOYS_COMPAT, hgp=False):
    args, keep = basis.args()
    return lib().orc_contracted_eri(*args, i, j, k, l, boys, int(hgp))


def jk_inmem(I, D):
    """JK_inmem (basis.rs:462-484): J = (mu nu|la si) D_la,si, K = (mu la|nu si) D_la,si."""
    J = np.einsum("ijkl,kl->ij", I, D)
    K = np.einsum("ikjl,kl->ij", I, D)
    return J, K


def jk_direct(basis: CGTOBasis, D, boys=BOYS_COMPAT, rows=None, cints_fn=None, threads=1):
    """JK_direct (basis.rs:383-428), literally; ``rows`` limits mu to a range so
    a bounded sample can be timed.  ``cints_fn`` = address of the reference's
    compiled coulomb_repulsion (oracle/_ref) to use instead of the restatement.
    ``threads`` > 1 splits the mu loop over threads (each row is private)."""
    n = basis.n
    args, keep = basis.args()
    Dd, Dp = _d(D)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    r0, r1 = rows if rows is not None else (0, n)
    if threads > 1:
        lib().orc_jk_direct_rows_mt(*args, Dp, boys, cints_fn, r0, r1, int(threads),
                                    J.ctypes.data_as(_dp), K.ctypes.data_as(_dp))
    else:
        lib().orc_jk_direct_rows(*args, Dp, boys, cints_fn, r0, r1, J.ctypes.data_as(_dp),
                                 K.ctypes.data_as(_dp))
    return J, K


def jk_direct_sample(basis: CGTOBasis, D, quartets, boys=BOYS_COMPAT, cints_fn=None, threads=1):
    """JK_direct's inner body over an explicit list of function quartets
    (rows of ``quartets`` = mu, nu, la, si); returns partial (J, K)."""
    n = basis.n
    args, keep = basis.args()
    Dd, Dp = _d(D)
    q, qp = _i(np.asarray(quartets).reshape(-1, 4))
    J = np.zeros((n, n)); K = np.zeros((n, n))
    if threads > 1:
        lib().orc_jk_direct_sample_mt(*args, Dp, boys, cints_fn, len(q), qp, int(threads),
                                      J.ctypes.data_as(_dp), K.ctypes.data_as(_dp))
    else:
        lib().orc_jk_direct_sample(*args, Dp, boys, cints_fn, len(q), qp, J.ctypes.data_as(_dp),
                                   K.ctypes.data_as(_dp))
    return J, K


def cints_coulomb_repulsion_address():
    """Address of the reference's own coulomb_repulsion (oracle/_ref), or None."""
    R = ref_lib()
    if R is None:
        return None
    return ctypes.cast(R.coulomb_repulsion, ctypes.c_void_p).value


def nuclear_repulsion(atomnos, coords):
    xyz, xp = _d(np.asarray(coords, dtype=np.float64).reshape(-1, 3))
    Z, Zp = _d(np.asarray(atomnos, dtype=np.float64))
    return lib().orc_nuclear_repulsion(len(atomnos), xp, Zp)


# --------------------------------------------------------------------------
# SCF driver (src/bin/rchem.rs:40-127) in numpy.
# --------------------------------------------------------------------------

@dataclass
class SCFResult:
    e_elec: float
    iterations: int
    converged: bool
    D: np.ndarray
    energies: list


def scf(S, H, jk, nocc, thresh_e=1.0e-11, max_iter=1024, D0=None):
    """rchem's plain SCF: Loewdin X = U l^-1/2 U^T (rchem.rs:44-51), core guess
    (:56-61), loop F(D_old) -> F' = X^T F X -> eigh -> C = X C' ->
    D = C_occ C_occ^T -> E = sum D_new o (H + F(D_old)) until |dE| < thresh_e
    (:72-90).  ``jk(D) -> (J, K)``; F = H + 2J - K (:110-127)."""
    lam, U = np.linalg.eigh(S)
    X = U @ np.diag(1.0 / np.sqrt(lam)) @ U.T

    def density(F):
        _, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        return C[:, :nocc] @ C[:, :nocc].T

    D = density(H) if D0 is None else D0
    e_new = float(np.sum(D * (H + H)))
    energies = []
    for it in range(max_iter):
        J, K = jk(D)
        F = H + 2.0 * J - K
        D = density(F)
        e_old, e_new = e_new, float(np.sum(D * (H + F)))
        energies.append(e_new)
        if abs(e_new - e_old) < thresh_e:
            return SCFResult(e_new, it + 1, True, D, energies)
    return SCFResult(e_new, max_iter, False, D, energies)
