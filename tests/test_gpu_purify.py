"""Densities by canonical purification (GEMMs on the matrix cores) in place of
the eigensolver: the projector onto the nocc lowest eigenvectors of
F' = X^T F X is rchem's D = C_occ C_occ^T (rchem.rs:74-78, 101-104), so every
iteration's energy must match the eigensolver run (plain iterations and
DIIS), and the gradient, whose W comes from the final eigensolve, too."""
import ctypes
import re

import numpy as np
import pytest

from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_EXACT, default_opts
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z, nocc_of, water_cluster

pytestmark = pytest.mark.gpu

_libc = ctypes.CDLL(None)


def _say(*a):
    print(*a, flush=True)


def _energies(capfd, fn):
    capfd.readouterr()
    out = fn()
    _libc.fflush(None)
    text = capfd.readouterr().out
    return out, [float(m.group(1)) for m in
                 re.finditer(r"^\s*\d+\s+(-?\d+\.\d+)", text, re.MULTILINE)]


@pytest.mark.parametrize("diis", [0, 8])
def test_purify_water_every_iteration(gpu_ctx, capfd, diis):
    xyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    B = rb.Basis.new(WATER_Z, xyz, "STO-3G", ctx=gpu_ctx)
    o = default_opts(boys_mode=BOYS_EXACT)
    with capfd.disabled():
        _say(f"case: water STO-3G, diis={diis}, purification vs eigensolver")
    run = lambda p: rb.scf(B, xyz, WATER_Z, nocc=5, thresh_e=1e-11, diis=diis, opts=o,
                           verbose=True, purify=p)
    (r0, D0), e0 = _energies(capfd, lambda: run(False))
    (r1, D1), e1 = _energies(capfd, lambda: run(True))
    assert r0.converged and r1.converged and r0.iterations == r1.iterations
    assert len(e0) == len(e1) >= 3
    assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-10
    assert abs(r0.e_total - r1.e_total) < 1e-10
    assert np.max(np.abs(D0 - D1)) < 1e-9


def test_purify_water_cluster_and_gradient(gpu_ctx, capfd):
    Z, xyz = water_cluster((2, 2, 2), seed=3)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    o = default_opts()
    with capfd.disabled():
        _say("case: (H2O)_8 6-31G DIIS(8), purification vs eigensolver, SCF and gradient")
    run = lambda p: rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=o,
                           verbose=True, purify=p)
    (r0, D0), e0 = _energies(capfd, lambda: run(False))
    (r1, D1), e1 = _energies(capfd, lambda: run(True))
    assert r0.converged and r1.converged
    m = min(len(e0), len(e1))
    assert m >= 5 and abs(r0.iterations - r1.iterations) <= 1
    assert np.max(np.abs(np.array(e0[:m]) - np.array(e1[:m]))) < 1e-8
    assert abs(r0.e_total - r1.e_total) < 1e-10
    with capfd.disabled():
        g0 = rb.rhf_gradient(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=o)[1]
        g1 = rb.rhf_gradient(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=o,
                             purify=True)[1]
    assert np.max(np.abs(g0 - g1)) < 1e-8
