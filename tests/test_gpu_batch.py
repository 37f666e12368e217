"""GPU batched SCF: many molecules solved together (molecule-restricted
quartets in one multi-molecule basis, batched SCF algebra) must give each
molecule the energy of its own single-molecule SCF on the same device path
(bar 1e-8 Eh), in plain rchem iterations and with DIIS, in one chunk and over
several chunks of molecules."""
import os

import numpy as np
import pytest

from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_EXACT, PREC_MIXED, default_opts
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z, nocc_of, random_molecules

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _single(ctx, Z, xyz, diis, opts=None):
    B = rb.Basis.new(Z, xyz, "STO-3G", ctx=ctx)
    res, _ = rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-11, diis=diis,
                    opts=opts or default_opts(boys_mode=BOYS_EXACT))
    return res


def test_batch_matches_single_molecule_scf_diis(gpu_ctx):
    mols = random_molecules(count=12, seed=512, nmin=3, nmax=10)
    _say(f"case: batch of {len(mols)} molecules, DIIS(8), one chunk")
    res = rb.batch_scf(mols, "STO-3G", thresh_e=1e-11, diis=8, ctx=gpu_ctx)
    compared = 0
    for (Z, xyz), r in zip(mols, res):
        s = _single(gpu_ctx, Z, xyz, diis=8)
        if not (s.converged and r.converged):
            continue
        assert abs(r.e_total - s.e_total) < 1e-8, (Z, r.e_total, s.e_total)
        assert abs(r.e_nuc - s.e_nuc) < 1e-10
        compared += 1
    assert compared >= 10


def test_batch_plain_iterations_water_copies(gpu_ctx):
    """rchem's plain iterations (diis = 0) on copies of water at different
    places: every copy converges to the single-molecule energy."""
    xyz0 = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    mols = [(list(WATER_Z), xyz0 + shift) for shift in ([0, 0, 0], [40.0, 0, 0], [0, 3.0, 0])]
    _say("case: batch of 3 water copies, plain iterations")
    res = rb.batch_scf(mols, "STO-3G", thresh_e=1e-11, diis=0, ctx=gpu_ctx)
    s = _single(gpu_ctx, WATER_Z, xyz0, diis=0)
    for r in res:
        assert r.converged
        assert abs(r.e_total - s.e_total) < 1e-9


def test_batch_several_chunks_and_mixed(gpu_ctx):
    """Random H/C/N/O random-walk molecules are often hard for RHF: under plain
    DIIS many oscillate.  With density damping (30 % of the previous density
    while the commutator exceeds 1e-2) at least 90 % converge; the energy bars
    are on those that do."""
    mols = random_molecules(count=120, seed=77, nmin=10, nmax=30)
    damp = dict(damping=0.3, damping_until=1e-2)
    _say(f"case: batch of {len(mols)} molecules over several chunks, FP64 and mixed, damped DIIS")
    # chunks of at most 2048 functions (the library reads the limit per call):
    # several chunks, each with its own retirement and slot compaction
    old = os.environ.get("RCHEM_BATCH_NBF")
    os.environ["RCHEM_BATCH_NBF"] = "2048"
    try:
        r64 = rb.batch_scf(mols, "STO-3G", thresh_e=1e-10, diis=8, max_iter=128, ctx=gpu_ctx,
                           **damp)
        rmx = rb.batch_scf(mols, "STO-3G", thresh_e=1e-10, diis=8, max_iter=128, ctx=gpu_ctx,
                           opts=default_opts(precision=PREC_MIXED), **damp)
    finally:
        if old is None:
            del os.environ["RCHEM_BATCH_NBF"]
        else:
            os.environ["RCHEM_BATCH_NBF"] = old
    ok = [i for i in range(len(mols)) if r64[i].converged and rmx[i].converged]
    _say(f"converged in both: {len(ok)} of {len(mols)}")
    assert len(ok) >= 0.9 * len(mols)
    assert max(abs(r64[i].e_total - rmx[i].e_total) for i in ok) < 1e-6
    for i in ok[:: max(1, len(ok) // 8)]:
        B = rb.Basis.new(mols[i][0], mols[i][1], "STO-3G", ctx=gpu_ctx)
        s, _ = rb.scf(B, mols[i][1], mols[i][0], nocc=nocc_of(mols[i][0]), thresh_e=1e-10,
                      diis=8, max_iter=128, opts=default_opts(boys_mode=BOYS_EXACT), **damp)
        if s.converged:
            assert abs(r64[i].e_total - s.e_total) < 1e-8
