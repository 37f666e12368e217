"""rchem's operator API (src/basis.rs, src/bin/rchem.rs) on top of the C ABI.

Same names, argument meaning and error behaviour as the reference:

  * ``Basis.new(atomnos, atomcoords, basis_set_name)``  -- basis.rs:182-212
  * ``S(basis)``, ``T(basis)``, ``V(basis, atomcoords, atomnos)`` -- basis.rs:234, 273, 316
  * ``JK_direct(J, K, basis, D)`` fills caller-owned J and K -- basis.rs:383-428
  * ``build_I(basis)`` -- basis.rs:430-460
  * ``JK_inmem(I, D)`` -- basis.rs:462-484 (a host contraction of a tensor)

All integral work runs through ``librchem_gpu.so`` on the GPU; nothing here
computes an integral.  Where rchem panics (``unwrap``/``assert!``) these
raise ``RchemError`` with the library's status and message.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (BOYS_COMPAT, BOYS_EXACT, JKStats, Opts, RchemError, SCFOpts, SCFResult,
                   default_opts, lib)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Context:
    """One device context (HIP device + stream).  Not thread-safe."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().rchem_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise RchemError(rc, lib().rchem_last_error(None).decode())
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise RchemError(rc, lib().rchem_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            lib().rchem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Basis:
    """Cartesian contracted basis held by a device context (rchem's Basis,
    basis.rs:176-179)."""

    def __init__(self, ctx):
        self.ctx = ctx
        n = ctypes.c_int()
        ctx.check(lib().rchem_basis_nbf(ctx.h, ctypes.byref(n)))
        self.nbf = n.value

    @classmethod
    def new(cls, atomnos, atomcoords, basis_set_name, ctx=None):
        """Basis::new (basis.rs:182): embedded STO-3G / 6-31G data, coordinates
        used as given (bohr)."""
        ctx = ctx or default_context()
        Z = np.ascontiguousarray(atomnos, dtype=np.int32)
        xyz = np.ascontiguousarray(atomcoords, dtype=np.float64).reshape(-1, 3)
        ctx.check(lib().rchem_basis_from_atoms(ctx.h, len(Z), _iptr(Z), _dptr(xyz),
                                               basis_set_name.encode()))
        return cls(ctx)

    @classmethod
    def from_shells(cls, centres, ls, nprims, exps, coefs, ctx=None):
        ctx = ctx or default_context()
        xyz = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        l = np.ascontiguousarray(ls, dtype=np.int32)
        npr = np.ascontiguousarray(nprims, dtype=np.int32)
        ex = np.ascontiguousarray(exps, dtype=np.float64)
        cf = np.ascontiguousarray(coefs, dtype=np.float64)
        ctx.check(lib().rchem_basis_upload(ctx.h, len(l), _dptr(xyz), _iptr(l), _iptr(npr),
                                           _dptr(ex), _dptr(cf)))
        return cls(ctx)

    def shells(self):
        """(centres, l, nprim, exps, coefs) of the current basis."""
        ns, npt = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().rchem_basis_dims(self.ctx.h, ctypes.byref(ns), ctypes.byref(npt)))
        xyz = np.zeros((ns.value, 3))
        l = np.zeros(ns.value, dtype=np.int32)
        npr = np.zeros(ns.value, dtype=np.int32)
        ex = np.zeros(npt.value)
        cf = np.zeros(npt.value)
        self.ctx.check(lib().rchem_basis_shells(self.ctx.h, _dptr(xyz), _iptr(l), _iptr(npr),
                                                _dptr(ex), _dptr(cf)))
        return xyz, l, npr, ex, cf


def one_electron(basis: Basis, atomcoords, atomnos):
    n = basis.nbf
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    xyz = np.ascontiguousarray(atomcoords, dtype=np.float64).reshape(-1, 3)
    Z = np.ascontiguousarray(atomnos, dtype=np.float64)
    basis.ctx.check(lib().rchem_one_electron(basis.ctx.h, len(Z), _dptr(Z), _dptr(xyz),
                                             _dptr(S), _dptr(T), _dptr(V)))
    return S, T, V


def S(basis: Basis):
    """basis.rs:234."""
    return one_electron(basis, np.zeros((0, 3)), [])[0]


def T(basis: Basis):
    """basis.rs:273."""
    return one_electron(basis, np.zeros((0, 3)), [])[1]


def V(basis: Basis, atomcoords, atomnos):
    """basis.rs:316: sum over nuclei of Z_C <mu|-1/r_C|nu>."""
    return one_electron(basis, atomcoords, atomnos)[2]


def JK_direct(J, K, basis: Basis, D, opts: Opts | None = None):
    """basis.rs:383: overwrites caller-owned J and K (float64, C-contiguous)."""
    n = basis.nbf
    for name, a in (("J", J), ("K", K)):
        if a.shape != (n, n) or a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError(f"{name} must be a C-contiguous float64 ({n}, {n}) array")
    Dc = np.ascontiguousarray(D, dtype=np.float64)
    if Dc.shape != (n, n):
        raise ValueError(f"D must be ({n}, {n})")
    o = opts or default_opts()
    basis.ctx.check(lib().rchem_jk(basis.ctx.h, _dptr(Dc), _dptr(J), _dptr(K), ctypes.byref(o)))


def jk_last_stats(basis: Basis) -> JKStats:
    st = JKStats()
    basis.ctx.check(lib().rchem_jk_last_stats(basis.ctx.h, ctypes.byref(st)))
    return st


def build_I(basis: Basis, opts: Opts | None = None):
    """basis.rs:430: full [mu,nu,la,si] tensor."""
    n = basis.nbf
    I = np.zeros((n, n, n, n))
    o = opts or default_opts()
    basis.ctx.check(lib().rchem_eri_tensor(basis.ctx.h, _dptr(I), ctypes.byref(o)))
    return I


def eri_shell_quartets(basis: Basis, quartets, opts: Opts | None = None):
    """(AB|CD) blocks for rows (A, B, C, D) of shell indices; returns a list of
    arrays shaped (nA, nB, nC, nD)."""
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
    ls = basis.shells()[1]
    nc = (ls + 1) * (ls + 2) // 2
    sizes = [int(np.prod(nc[r])) for r in q]
    out = np.zeros(max(sum(sizes), 1))
    o = opts or default_opts()
    basis.ctx.check(lib().rchem_eri_shell_quartets(basis.ctx.h, len(q), _iptr(q), _dptr(out),
                                                   ctypes.byref(o)))
    blocks, off = [], 0
    for r, sz in zip(q, sizes):
        blocks.append(out[off:off + sz].reshape(tuple(nc[r])))
        off += sz
    return blocks


def JK_inmem(I, D):
    """basis.rs:462 (host contraction of a tensor already in memory)."""
    return np.einsum("ijkl,kl->ij", I, D), np.einsum("ikjl,kl->ij", I, D)


def schwarz(basis: Basis):
    ns = len(basis.shells()[1])
    Q = np.zeros((ns, ns))
    basis.ctx.check(lib().rchem_schwarz(basis.ctx.h, _dptr(Q), None))
    return Q


def scf(basis: Basis, atomcoords, atomnos, nocc, thresh_e=1e-11, max_iter=1024,
        opts: Opts | None = None, verbose=False, diis=0, graph=False, purify=False,
        damping=0.0, damping_until=0.0):
    """rchem.rs:10-90 on the device: returns (SCFResult, D).  diis = k > 0
    replaces rchem's plain iterations by DIIS over k Fock matrices; graph
    replays the Fock build as a hipGraph; purify forms each iteration's
    density by canonical purification (GEMMs) instead of the eigensolver."""
    o = SCFOpts()
    lib().rchem_scf_opts_default(ctypes.byref(o))
    o.diis = int(diis)
    o.graph = int(bool(graph))
    o.purify = int(bool(purify))
    o.nocc = int(nocc)
    o.max_iter = int(max_iter)
    o.thresh_e = float(thresh_e)
    o.jk = opts or default_opts()
    o.verbose = int(bool(verbose))
    o.damping = float(damping)
    o.damping_until = float(damping_until)
    xyz = np.ascontiguousarray(atomcoords, dtype=np.float64).reshape(-1, 3)
    Z = np.ascontiguousarray(atomnos, dtype=np.float64)
    res = SCFResult()
    D = np.zeros((basis.nbf, basis.nbf))
    rc = lib().rchem_scf(basis.ctx.h, len(Z), _dptr(Z), _dptr(xyz), ctypes.byref(o),
                         ctypes.byref(res), _dptr(D))
    if rc not in (0, 6):
        basis.ctx.check(rc)
    return res, D


def batch_scf(molecules, basis_set_name, thresh_e=1e-11, max_iter=1024, diis=0,
              opts: Opts | None = None, ctx=None, graph=False, damping=0.0, damping_until=0.0):
    """Many independent molecules [(atomnos, atomcoords), ...] solved together
    on the device; returns a list of SCFResult (the context's basis is cleared)."""
    ctx = ctx or default_context()
    natom = np.array([len(z) for z, _ in molecules], dtype=np.int32)
    Z = np.ascontiguousarray(np.concatenate([np.asarray(z) for z, _ in molecules]), dtype=np.int32)
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3)
                                               for _, x in molecules]))
    o = SCFOpts()
    lib().rchem_scf_opts_default(ctypes.byref(o))
    o.max_iter = int(max_iter)
    o.thresh_e = float(thresh_e)
    o.diis = int(diis)
    o.graph = int(bool(graph))
    o.jk = opts or default_opts()
    o.damping = float(damping)
    o.damping_until = float(damping_until)
    res = (SCFResult * len(molecules))()
    rc = lib().rchem_batch_scf(ctx.h, len(molecules), _iptr(natom), _iptr(Z), _dptr(xyz),
                               basis_set_name.encode(), ctypes.byref(o), res)
    if rc not in (0, 6):
        ctx.check(rc)
    return list(res)


def rhf_gradient(basis: Basis, atomcoords, atomnos, nocc, thresh_e=1e-11, max_iter=1024,
                 opts: Opts | None = None, diis=0, graph=False, purify=False):
    """Analytic RHF nuclear gradient (Eh/bohr) after rchem's SCF on the device
    (DIIS over `diis` Fock matrices when > 0; the Fock build replayed as a
    hipGraph when `graph`): returns (SCFResult, grad of shape (natom, 3))."""
    o = SCFOpts()
    lib().rchem_scf_opts_default(ctypes.byref(o))
    o.diis = int(diis)
    o.graph = int(bool(graph))
    o.purify = int(bool(purify))
    o.nocc = int(nocc)
    o.max_iter = int(max_iter)
    o.thresh_e = float(thresh_e)
    o.jk = opts or default_opts()
    xyz = np.ascontiguousarray(atomcoords, dtype=np.float64).reshape(-1, 3)
    Z = np.ascontiguousarray(atomnos, dtype=np.float64)
    res = SCFResult()
    g = np.zeros((len(Z), 3))
    rc = lib().rchem_rhf_gradient(basis.ctx.h, len(Z), _dptr(Z), _dptr(xyz), ctypes.byref(o),
                                  ctypes.byref(res), _dptr(g))
    if rc not in (0, 6):
        basis.ctx.check(rc)
    return res, g


__all__ = ["Context", "Basis", "S", "T", "V", "JK_direct", "JK_inmem", "build_I", "scf",
           "rhf_gradient", "batch_scf",
           "eri_shell_quartets",
           "one_electron", "schwarz", "jk_last_stats", "BOYS_EXACT", "BOYS_COMPAT"]
