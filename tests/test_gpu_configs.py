"""The gradient and batched-SCF workloads at their full sizes.

* (H2O)_32 6-31G (seed 32, 2x4x4 lattice, 416 functions, 96 atoms): the
  analytic RHF gradient against central differences (h = 1e-4 bohr) of the
  GPU energy on 12 coordinates spread over the cluster (seeded), and zero net
  force and torque over all 288 coordinates.  Bar 1e-6 Eh/bohr.
* 512 random H/C/N/O molecules (seed 512, 3-30 atoms, STO-3G) in one batched
  SCF call: a seeded sample against each molecule's own single-molecule SCF
  (1e-8 Eh), and the molecules of at most 20 functions against the oracle's
  restatement of rchem's SCF (rchem.rs:40-127) on the oracle's integrals
  (1e-8 Eh) wherever rchem's plain iterations converge there.
The reference solves one molecule with plain iterations (rchem.rs:72-90) and
has no derivative code; these are the additions' own bars.
"""
import os

import numpy as np
import pytest

import oracle as O
from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_EXACT, default_opts
from rchem_amd.inputs import nocc_of, random_molecules, water_cluster

pytestmark = pytest.mark.gpu
H = 1e-4
STO3G_NBF = {1: 1, 6: 5, 7: 5, 8: 5}
# random-walk molecules often oscillate under plain DIIS; mixing 30 % of the
# previous density while the commutator is large converges ~97 % of them
DAMP = dict(damping=0.3, damping_until=1e-2)
O_STO3G = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rchem_amd",
                       "data", "sto-3g.json")


def _say(*a):
    print(*a, flush=True)


def test_h2o32_gradient_vs_fd(gpu_ctx):
    Z, xyz = water_cluster(dims=(2, 4, 4), seed=32)
    nocc = nocc_of(Z)
    o = default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=1e-13)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    assert B.nbf == 416 and len(Z) == 96
    _say("case: (H2O)_32 6-31G analytic gradient, 288 coordinates")
    res, g = rb.rhf_gradient(B, xyz, Z, nocc=nocc, thresh_e=1e-11, diis=8, opts=o)
    assert res.converged
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8
    assert np.max(np.abs(np.cross(xyz, g).sum(axis=0))) < 1e-7

    def energy(p):
        Bp = rb.Basis.new(Z, p, "6-31G", ctx=gpu_ctx)
        r, _ = rb.scf(Bp, p, Z, nocc=nocc, thresh_e=1e-11, diis=8, opts=o)
        assert r.converged
        return r.e_total

    coords = sorted(np.random.default_rng(32).choice(3 * len(Z), 12, replace=False).tolist())
    worst = 0.0
    for c in coords:
        a, x = divmod(c, 3)
        xp, xm = xyz.copy(), xyz.copy()
        xp[a, x] += H
        xm[a, x] -= H
        fd = (energy(xp) - energy(xm)) / (2 * H)
        _say(f"  coordinate {c}: analytic {g[a, x]:+.9f} FD {fd:+.9f}")
        worst = max(worst, abs(fd - g[a, x]))
    assert worst < 1e-6


@pytest.fixture(scope="module")
def batch512(gpu_ctx):
    mols = random_molecules(count=512, seed=512, nmin=3, nmax=30)
    _say("case: 512 molecules (seed 512), one batched SCF call, DIIS(8) with density damping "
         "0.3 while the commutator exceeds 1e-2, max 128 iterations")
    res = rb.batch_scf(mols, "STO-3G", thresh_e=1e-10, diis=8, max_iter=128, ctx=gpu_ctx,
                       **DAMP)
    return mols, res


def test_batch512_vs_single_molecule_scf(gpu_ctx, batch512):
    mols, res = batch512
    conv = sum(r.converged for r in res)
    _say(f"  converged: {conv} of {len(res)}")
    assert conv >= 0.9 * len(res)
    sample = np.random.default_rng(512).choice(len(mols), 16, replace=False)
    compared = 0
    for i in sample:
        Z, xyz = mols[i]
        B = rb.Basis.new(Z, xyz, "STO-3G", ctx=gpu_ctx)
        s, _ = rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, max_iter=128,
                      opts=default_opts(boys_mode=BOYS_EXACT), **DAMP)
        if not (s.converged and res[i].converged):
            continue
        assert abs(res[i].e_total - s.e_total) < 1e-8, (int(i), res[i].e_total, s.e_total)
        assert abs(res[i].e_nuc - O.nuclear_repulsion(Z, xyz)) < 1e-9
        compared += 1
    assert compared >= 10


def test_batch512_small_molecules_vs_oracle(gpu_ctx, batch512):
    """For every molecule of at most 20 functions that converged in the batch:
    its single-molecule SCF reaches the same stationary point (1e-8 Eh), and
    that density, put through the oracle's integrals (THO ERIs, exact Boys,
    JK_inmem and F = H + 2J - K as rchem.rs:110-127), gives the same energy
    sum D o (H + F) (1e-8 Eh) and is stationary there (orthogonalised
    commutator below 1e-6).  Random molecules can have several SCF solutions,
    so the oracle is asked about the GPU's solution, not for one of its own."""
    mols, res = batch512
    small = [i for i, (Z, _) in enumerate(mols) if sum(STO3G_NBF[z] for z in Z) <= 20]
    _say(f"  molecules of <= 20 functions: {len(small)}")
    assert len(small) >= 5
    bse = O.load_bse_json(O_STO3G)
    compared = 0
    for i in small:
        Z, xyz = mols[i]
        if not res[i].converged:
            continue
        B = rb.Basis.new(Z, xyz, "STO-3G", ctx=gpu_ctx)
        s, D = rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, max_iter=128,
                      opts=default_opts(boys_mode=BOYS_EXACT), **DAMP)
        if not s.converged or abs(s.e_total - res[i].e_total) > 1e-6:
            continue  # the single SCF found another solution: nothing to compare
        assert abs(s.e_total - res[i].e_total) < 1e-8
        Bo = O.cgtos_from_bse(bse, Z, xyz)
        S, T, V = O.one_electron(Bo, Z, xyz)
        Hc = T + V
        Jo, Ko = O.jk_inmem(O.eri_tensor(Bo, boys=O.BOYS_EXACT), D)
        F = Hc + 2.0 * Jo - Ko
        assert abs(float(np.sum(D * (Hc + F))) - res[i].e_elec) < 1e-8, i
        lam, U = np.linalg.eigh(S)
        X = U @ np.diag(lam ** -0.5) @ U.T
        assert np.max(np.abs(X.T @ (F @ D @ S - S @ D @ F) @ X)) < 1e-6, i
        compared += 1
    _say(f"  compared with the oracle: {compared}")
    assert compared >= 0.6 * len(small)
