"""hipGraph replay of the SCF iteration.  With graph=1 the Fock build
(planner, ERI class kernels, digestion, F = H + 2J - K) is captured once and
replayed on every later iteration, with the SCF's own eager work (energy,
DIIS, eigensolver, density, host synchronisations) between replays; with
DIIS and purified densities as well the whole iteration -- Fock build,
energy, commutator, DIIS dot products and solve, extrapolation, purification,
density, damping -- is one graph with no host step.  Every iteration's energy
must equal the kernel-by-kernel run's (each replay is checked, not only the
last), in FP64 and in mixed precision (a recapture at the switch to FP64),
for the single-molecule and the batched SCF (rchem.rs:72-90)."""
import ctypes
import re

import numpy as np
import pytest

from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_EXACT, PREC_MIXED, default_opts
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z, nocc_of, random_molecules, water_cluster

pytestmark = pytest.mark.gpu

_libc = ctypes.CDLL(None)


def _say(*a):
    print(*a, flush=True)


def _run(capfd, fn):
    capfd.readouterr()
    out = fn()
    _libc.fflush(None)
    text = capfd.readouterr().out
    energies = [float(m.group(1)) for m in
                re.finditer(r"^\s*\d+\s+(-?\d+\.\d+)", text, re.MULTILINE)]
    return out, energies


def _scf(B, xyz, Z, graph, diis, opts, purify=False, **kw):
    return rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=diis, opts=opts,
                  verbose=True, graph=graph, purify=purify, **kw)


@pytest.mark.parametrize("diis", [0, 8])
def test_graph_water_every_iteration(gpu_ctx, capfd, diis):
    xyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    B = rb.Basis.new(WATER_Z, xyz, "STO-3G", ctx=gpu_ctx)
    opts = default_opts(boys_mode=BOYS_EXACT)
    with capfd.disabled():
        _say(f"case: water STO-3G, diis={diis}, graph vs eager")
    (r0, _), e0 = _run(capfd, lambda: _scf(B, xyz, WATER_Z, False, diis, opts))
    (r1, _), e1 = _run(capfd, lambda: _scf(B, xyz, WATER_Z, True, diis, opts))
    assert r0.converged and r1.converged
    assert r0.iterations == r1.iterations and len(e0) == len(e1) >= 3
    assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-11
    assert abs(r0.e_total - r1.e_total) < 1e-11


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_water_cluster(gpu_ctx, capfd, mixed):
    Z, xyz = water_cluster((2, 2, 2), seed=3)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    opts = default_opts(precision=PREC_MIXED) if mixed else default_opts()
    with capfd.disabled():
        _say(f"case: (H2O)_8 6-31G DIIS(8), mixed={mixed}, graph vs eager")
    (r0, _), e0 = _run(capfd, lambda: _scf(B, xyz, Z, False, 8, opts))
    (r1, _), e1 = _run(capfd, lambda: _scf(B, xyz, Z, True, 8, opts))
    # a second graph SCF on the same context, after an eager one
    (r2, _), e2 = _run(capfd, lambda: _scf(B, xyz, Z, True, 8, opts))
    assert r0.converged and r1.converged and r2.converged
    if not mixed:
        assert r0.iterations == r1.iterations == r2.iterations and len(e0) == len(e1) >= 5
        assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-10
        assert np.max(np.abs(np.array(e0) - np.array(e2))) < 1e-10
        return
    # FP32 quartets accumulate in the order the atomics land, so the mixed
    # phase differs from run to run by ~1e-8 Eh (graph or not) and the switch
    # to FP64 (|dE| < 1e-7) can come an iteration earlier or later; past the
    # switch, DIIS over nearly parallel error vectors can amplify that noise
    # into a one-iteration excursion.  The bar: every replayed mixed-phase
    # iteration tracks the eager one to that noise, the FP64 graph (recaptured
    # at the switch) ends on the same energy.
    for r in (r1, r2):
        assert r.fp64_from > 0 and abs(r.iterations - r0.iterations) <= 3
        assert abs(r.e_total - r0.e_total) < 1e-9
    for r, e in ((r1, e1), (r2, e2)):
        m = min(r0.fp64_from, r.fp64_from)
        assert m >= 5
        assert np.max(np.abs(np.array(e0[:m]) - np.array(e[:m]))) < 1e-6


def test_graph_batch(gpu_ctx):
    """Water copies (each converges in a few dozen iterations, the same in
    every run) plus random molecules (some converge slowly and chaotically:
    digestion atomics land in run-dependent order, ~1e-16 relative, which can
    move a slow molecule's convergence by many iterations, graph or not)."""
    xyz0 = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    waters = [(list(WATER_Z), xyz0 + np.array(s)) for s in
              ([0, 0, 0], [30.0, 0, 0], [0, 5.0, 0], [0, 0, -7.5])]
    mols = waters + random_molecules(count=12, seed=512, nmin=3, nmax=10)
    _say(f"case: batch of {len(mols)} molecules, DIIS(8), graph vs eager")
    r0 = rb.batch_scf(mols, "STO-3G", thresh_e=1e-11, diis=8, ctx=gpu_ctx)
    r1 = rb.batch_scf(mols, "STO-3G", thresh_e=1e-11, diis=8, ctx=gpu_ctx, graph=True)
    for a, b in zip(r0[:len(waters)], r1[:len(waters)]):
        assert a.converged and b.converged
        assert abs(a.iterations - b.iterations) <= 1
        assert abs(a.e_total - b.e_total) < 1e-10
        assert abs(a.e_total - r0[0].e_total) < 1e-9
    both = 0
    for a, b in zip(r0[len(waters):], r1[len(waters):]):
        if a.converged and b.converged:
            assert abs(a.e_total - b.e_total) < 1e-8
            both += 1
    assert both >= 10


def test_whole_iteration_graph_water(gpu_ctx, capfd):
    xyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    B = rb.Basis.new(WATER_Z, xyz, "STO-3G", ctx=gpu_ctx)
    opts = default_opts(boys_mode=BOYS_EXACT)
    with capfd.disabled():
        _say("case: water STO-3G, DIIS(8), purified densities, whole-iteration graph vs eager")
    (r0, D0), e0 = _run(capfd, lambda: _scf(B, xyz, WATER_Z, False, 8, opts, purify=True))
    (r1, D1), e1 = _run(capfd, lambda: _scf(B, xyz, WATER_Z, True, 8, opts, purify=True))
    (r2, D2), e2 = _run(capfd, lambda: _scf(B, xyz, WATER_Z, False, 8, opts, purify=False))
    assert r0.converged and r1.converged and r2.converged
    assert r0.graph_iterations == 0 and r2.graph_iterations == 0
    # the first iteration plans and measures the purification; the rest replay
    assert r1.graph_iterations == r1.iterations - 1 >= 2
    assert r0.iterations == r1.iterations == r2.iterations and len(e0) == len(e1) >= 3
    assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-11
    assert np.max(np.abs(np.array(e0) - np.array(e2))) < 1e-10
    assert np.max(np.abs(D0 - D1)) < 1e-10 and np.max(np.abs(D0 - D2)) < 1e-9


@pytest.mark.parametrize("mixed", [False, True])
def test_whole_iteration_graph_water_cluster(gpu_ctx, capfd, mixed):
    Z, xyz = water_cluster((2, 2, 2), seed=3)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    opts = default_opts(precision=PREC_MIXED) if mixed else default_opts()
    with capfd.disabled():
        _say(f"case: (H2O)_8 6-31G DIIS(8), purified, mixed={mixed}, whole-iteration graph "
             "vs eager")
    (r0, D0), e0 = _run(capfd, lambda: _scf(B, xyz, Z, False, 8, opts, purify=True))
    (r1, D1), e1 = _run(capfd, lambda: _scf(B, xyz, Z, True, 8, opts, purify=True))
    assert r0.converged and r1.converged
    assert r1.graph_iterations >= r1.iterations - (3 if mixed else 1)
    if not mixed:
        assert r0.iterations == r1.iterations and len(e0) == len(e1) >= 5
        assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-10
        assert abs(r0.e_total - r1.e_total) < 1e-11
        assert np.max(np.abs(D0 - D1)) < 1e-8
        return
    # the FP32 quartets' atomics make the mixed phase differ by ~1e-8 Eh from
    # run to run (see test_graph_water_cluster)
    assert r1.fp64_from > 0 and abs(r1.iterations - r0.iterations) <= 3
    assert abs(r1.e_total - r0.e_total) < 1e-9
    m = min(r0.fp64_from, r1.fp64_from)
    assert m >= 5 and np.max(np.abs(np.array(e0[:m]) - np.array(e1[:m]))) < 1e-6


def test_whole_iteration_graph_damping_and_gradient(gpu_ctx, capfd):
    """Damping decided on the device per iteration (commutator above 1e-2),
    and the gradient from a graph-replayed SCF (W = D F D of its final F)."""
    Z, xyz = water_cluster((2, 1, 1), seed=5)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    opts = default_opts()
    with capfd.disabled():
        _say("case: (H2O)_2 6-31G DIIS(8), purified, damping 0.3 until 1e-2, graph vs eager, "
             "gradient")
    kw = dict(damping=0.3, damping_until=1e-2)
    (r0, D0), e0 = _run(capfd, lambda: _scf(B, xyz, Z, False, 8, opts, purify=True, **kw))
    (r1, D1), e1 = _run(capfd, lambda: _scf(B, xyz, Z, True, 8, opts, purify=True, **kw))
    (r2, _), e2 = _run(capfd, lambda: _scf(B, xyz, Z, True, 8, opts, purify=True))
    assert r0.converged and r1.converged and r2.converged and r1.graph_iterations > 0
    assert r0.iterations == r1.iterations and len(e0) == len(e1)
    assert np.max(np.abs(np.array(e0) - np.array(e1))) < 1e-10
    assert e1[1:3] != e2[1:3]  # the damped run took a different path
    assert abs(r1.e_total - r2.e_total) < 1e-9
    with capfd.disabled():
        g0 = rb.rhf_gradient(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=opts)[1]
        g1 = rb.rhf_gradient(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=opts,
                             purify=True, graph=True)[1]
    assert np.max(np.abs(g0 - g1)) < 1e-8
