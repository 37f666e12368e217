"""Launch tuning of the Fock build: every compiled launch shape of the
LDS-digestion kernel, workgroup count, chunk size and chunk order (bra order
or sorted by first ket pair) gives the same J and K --
in FP64 to the order of floating-point atomics (1e-12 relative); in a mixed
build the shapes also order the FP32 arithmetic differently, so there the bar
is the FP32 quartets' share (threshold 1e-5 times FP32 rounding: 1e-9
relative) -- and the tuning-table reader takes well-formed lines and skips
the rest."""
import ctypes
import os

import numpy as np
import pytest

from rchem_amd import basis as rb
from rchem_amd._lib import PREC_MIXED, default_opts, lib
from rchem_amd.inputs import water_cluster

pytestmark = pytest.mark.gpu
SP_CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1)]


def _nvariants():
    n = 0
    while lib().rchem_lds_variant_info(n, None, None, None, None, None) == 0:
        n += 1
    return n


def test_every_launch_shape_gives_the_same_jk(gpu_ctx):
    L = lib()
    h = gpu_ctx.h
    Z, xyz = water_cluster(dims=(2, 2, 2), seed=8)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    n = B.nbf
    X = np.random.default_rng(4).standard_normal((n, n))
    D = 0.05 * (X + X.T)
    nv = _nvariants()
    assert nv >= 2
    try:
        for prec in (0, PREC_MIXED):
            o = default_opts(precision=prec, mixed_thresh=1e-5)
            assert L.rchem_load_tuning(h, None, None) == 0
            J0 = np.zeros((n, n)); K0 = np.zeros((n, n))
            rb.JK_direct(J0, K0, B, D, o)
            scale = max(np.abs(J0).max(), np.abs(K0).max())
            bar = (1e-9 if prec else 1e-12) * scale
            for v in range(nv):
                blocks, chunk = [(2048, 64), (300, 7), (4096, 128)][v % 3]
                order = (v // 3) % 2
                print(f"case: launch shape {v}, {blocks} workgroups, chunks of {chunk}, "
                      f"ket order {order}, precision {prec}", flush=True)
                for c in SP_CLASSES:
                    for fp32 in (0, 1):
                        assert L.rchem_set_class_tuning(h, *c, fp32, v, blocks, chunk) == 0
                        assert L.rchem_set_class_ket_order(h, *c, fp32, order) == 0
                J = np.zeros((n, n)); K = np.zeros((n, n))
                rb.JK_direct(J, K, B, D, o)
                assert np.max(np.abs(J - J0)) < bar, (v, prec, np.max(np.abs(J - J0)) / scale)
                assert np.max(np.abs(K - K0)) < bar, (v, prec, np.max(np.abs(K - K0)) / scale)
    finally:
        L.rchem_load_tuning(h, None, None)


def test_tuning_table_reader(gpu_ctx, tmp_path):
    L = lib()
    h = gpu_ctx.h
    p = tmp_path / "t.txt"
    p.write_text("# comment\n1 0 0 0 0 3 1024 32  # kept\n1 0 0 0 1 99 1024 32\n"
                 "0 0 0 0 1 2 5000 128\nnot a line\n2 2 2 2 0 1 0 64\n"
                 "1 1 0 0 0 1 512 16 0\n1 1 0 0 1 1 512 16 1\n1 1 1 0 0 1 512 16 2\n")
    cnt = ctypes.c_int()
    assert L.rchem_load_tuning(h, str(p).encode(), ctypes.byref(cnt)) == 0
    assert cnt.value == 4                              # the ket order 2 line is skipped
    o = ctypes.c_int()
    for cls, fp32, want in [((1, 1, 0, 0), 0, 0), ((1, 1, 0, 0), 1, 1),   # from the table
                            ((1, 0, 0, 0), 0, 1), ((1, 0, 0, 0), 1, 0)]:  # defaults
        assert L.rchem_get_class_ket_order(h, *cls, fp32, ctypes.byref(o)) == 0
        assert o.value == want, (cls, fp32)
    assert L.rchem_set_class_ket_order(h, 1, 0, 0, 0, 1, 1) == 0
    assert L.rchem_get_class_ket_order(h, 1, 0, 0, 0, 1, ctypes.byref(o)) == 0 and o.value == 1
    assert L.rchem_set_class_ket_order(h, 1, 0, 0, 0, 1, 2) == 1
    v, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.rchem_get_class_tuning(h, 1, 0, 0, 0, 0, ctypes.byref(v), ctypes.byref(b),
                                    ctypes.byref(c)) == 0
    assert (v.value, b.value, c.value) == (3, 1024, 32)
    assert L.rchem_get_class_tuning(h, 1, 0, 0, 0, 1, ctypes.byref(v), ctypes.byref(b),
                                    ctypes.byref(c)) == 0
    assert (v.value, b.value, c.value) == (0, 2048, 64)  # bad variant: skipped
    assert L.rchem_get_class_tuning(h, 0, 0, 0, 0, 1, ctypes.byref(v), ctypes.byref(b),
                                    ctypes.byref(c)) == 0
    assert (v.value, b.value, c.value) == (2, 5000, 128)
    assert L.rchem_set_class_tuning(h, 0, 0, 0, 0, 0, 1000, 1, 1) == 1
    assert L.rchem_load_tuning(h, None, ctypes.byref(cnt)) == 0 and cnt.value == 0
    # the batched SCF's table: same reader
    assert L.rchem_load_batch_tuning(h, str(p).encode(), ctypes.byref(cnt)) == 0
    assert cnt.value == 4
    assert L.rchem_load_batch_tuning(h, None, ctypes.byref(cnt)) == 0 and cnt.value == 0
    shipped = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "rchem_amd", "data", "tuning_gfx950.txt")
    if os.path.exists(shipped):
        assert L.rchem_load_tuning(h, shipped.encode(), ctypes.byref(cnt)) == 0
        assert cnt.value > 0
        L.rchem_load_tuning(h, None, None)
