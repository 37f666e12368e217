"""Error paths of the C ABI on a device.

rchem has no error returns: cints asserts or prints and returns an
uninitialised value (cints.c:313-314, 345-346), rchem's loop stops silently at
its iteration cap (rchem.rs:65, 72-90).  This library returns a status code
and leaves a message; these tests force each condition from the host (no
kernel is made to fault): invalid arguments, a non-finite density (host and
device entry points), an SCF that reaches max_iter, and a gradient requested
after such an SCF (refused, output untouched).
"""
import ctypes

import numpy as np
import pytest

from rchem_amd import basis as rb
from rchem_amd._lib import (BOYS_EXACT, PREC_MIXED, SCFOpts, SCFResult, default_opts, lib)
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z

pytestmark = pytest.mark.gpu

OK, INVALID, NO_DEVICE, UNSUPPORTED, NO_BASIS, NOT_CONVERGED, NONFINITE = 0, 1, 3, 4, 5, 6, 7


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _msg(ctx):
    return lib().rchem_last_error(ctx.h).decode()


def test_invalid_arguments(gpu_ctx):
    L = lib()
    h = gpu_ctx.h
    print("case: invalid arguments", flush=True)
    Z = np.array([8, 1, 1], dtype=np.int32)
    xyz = np.ascontiguousarray(np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3))
    assert L.rchem_basis_from_atoms(h, 3, _ip(Z), _dp(xyz), b"cc-pVQZ") == INVALID
    assert "unknown basis set" in _msg(gpu_ctx)
    Zfe = np.array([26], dtype=np.int32)
    assert L.rchem_basis_from_atoms(h, 1, _ip(Zfe), _dp(xyz), b"STO-3G") == INVALID
    bad = xyz.copy()
    bad[1, 2] = np.nan
    assert L.rchem_basis_from_atoms(h, 3, _ip(Z), _dp(bad), b"STO-3G") == INVALID
    # after a rejected basis the context has none
    D = np.zeros((7, 7))
    J = np.zeros((7, 7)); K = np.zeros((7, 7))
    assert L.rchem_jk(h, _dp(D), _dp(J), _dp(K), None) == NO_BASIS
    c = np.zeros((1, 3))
    ex = np.array([1.0]); cf = np.array([1.0])
    for l, npr, e in ((3, 1, 1.0), (0, 9, 1.0), (0, 0, 1.0), (0, 1, -1.0), (0, 1, np.inf)):
        lv = np.array([l], dtype=np.int32)
        nv = np.array([npr], dtype=np.int32)
        exv = np.full(max(npr, 1), e)
        cfv = np.ones(max(npr, 1))
        rc = L.rchem_basis_upload(h, 1, _dp(c), _ip(lv), _ip(nv), _dp(exv), _dp(cfv))
        assert rc in (INVALID, UNSUPPORTED), (l, npr, e, rc)
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    assert L.rchem_jk(h, None, _dp(J), _dp(K), None) == INVALID
    o = default_opts()
    o.boys_mode = 7
    assert L.rchem_jk(h, _dp(D), _dp(J), _dp(K), ctypes.byref(o)) == INVALID
    o = default_opts(precision=PREC_MIXED)
    o.schwarz_thresh = float("nan")
    assert L.rchem_jk(h, _dp(D), _dp(J), _dp(K), ctypes.byref(o)) == INVALID
    so = SCFOpts()
    L.rchem_scf_opts_default(ctypes.byref(so))
    so.nocc = B.nbf + 1
    res = SCFResult()
    zd = np.array(WATER_Z, dtype=float)
    assert L.rchem_scf(h, 3, _dp(zd), _dp(xyz), ctypes.byref(so), ctypes.byref(res), None) == INVALID
    badn = xyz.copy()
    badn[0, 0] = np.inf
    S = np.zeros((7, 7))
    assert L.rchem_one_electron(h, 3, _dp(zd), _dp(badn), _dp(S), None, None) == INVALID
    assert L.rchem_ctx_create(10_000, ctypes.byref(ctypes.c_void_p())) in (INVALID, NO_DEVICE)


def _hip():
    """The HIP runtime the library itself links (device buffers for the
    device-pointer entry point without a second runtime in the process)."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def test_nonfinite_density_host_and_device(gpu_ctx):
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    n = B.nbf
    print("case: NaN density through rchem_jk and rchem_jk_device", flush=True)
    D = np.eye(n)
    D[2, 3] = D[3, 2] = np.nan
    J = np.full((n, n), 5.0); K = np.full((n, n), 5.0)
    assert lib().rchem_jk(gpu_ctx.h, _dp(D), _dp(J), _dp(K), None) == NONFINITE
    assert "non-finite" in _msg(gpu_ctx)
    assert np.all(J == 5.0) and np.all(K == 5.0)  # outputs untouched
    hip = _hip()
    nb = 8 * n * n
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b in bufs:
        assert hip.hipMalloc(ctypes.byref(b), nb) == 0
    try:
        assert hip.hipMemcpy(bufs[0], D.ctypes.data, nb, 1) == 0  # host to device
        for prec in (0, PREC_MIXED):
            o = default_opts(precision=prec)
            assert lib().rchem_jk_device(gpu_ctx.h, bufs[0], bufs[1], bufs[2],
                                         ctypes.byref(o)) == OK
            assert hip.hipDeviceSynchronize() == 0
            Jh = np.zeros((n, n)); Kh = np.zeros((n, n))
            assert hip.hipMemcpy(Jh.ctypes.data, bufs[1], nb, 2) == 0  # device to host
            assert hip.hipMemcpy(Kh.ctypes.data, bufs[2], nb, 2) == 0
            # the NaN reaches J through every Coulomb sum and K through the exchange rows
            assert np.isnan(Jh[2, 3]) and np.isnan(Kh[2, 2]), prec
    finally:
        for b in bufs:
            hip.hipFree(b)


def test_scf_not_converged_and_gradient_refused(gpu_ctx):
    Z = np.array(WATER_Z)
    xyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    B = rb.Basis.new(Z, xyz, "STO-3G", ctx=gpu_ctx)
    print("case: SCF capped at 2 iterations; gradient after it", flush=True)
    res, D = rb.scf(B, xyz, Z, nocc=5, max_iter=2,
                    opts=default_opts(boys_mode=BOYS_EXACT))
    assert not res.converged and res.iterations == 2
    assert np.isfinite(res.e_total) and np.abs(D).max() > 0
    assert "did not converge" in _msg(gpu_ctx)
    for purify in (False, True):
        g0 = np.full((3, 3), 123.0)
        so = SCFOpts()
        lib().rchem_scf_opts_default(ctypes.byref(so))
        so.nocc, so.max_iter, so.purify = 5, 2, int(purify)
        so.jk = default_opts(boys_mode=BOYS_EXACT)
        r = SCFResult()
        zd = np.ascontiguousarray(Z, dtype=float)
        xc = np.ascontiguousarray(xyz)
        rc = lib().rchem_rhf_gradient(gpu_ctx.h, 3, _dp(zd), _dp(xc), ctypes.byref(so),
                                      ctypes.byref(r), _dp(g0))
        assert rc == NOT_CONVERGED and not r.converged
        assert np.all(g0 == 123.0)
    # and with enough iterations the same call succeeds
    res, g = rb.rhf_gradient(B, xyz, Z, nocc=5, diis=6)
    assert res.converged and np.isfinite(g).all()
