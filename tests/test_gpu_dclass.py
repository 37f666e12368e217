"""d-shell paths against the oracle (not against the library's own tensor).

* S, T, V of every (la, lb) class up to (d, d) from the per-class
  McMurchie-Davidson kernels against the oracle's THO one-electron
  restatement (os86.rs:629-773 defines the operators; the oracle follows
  cints.c's overlap/kinetic/nuclear_attraction).  Bar 1e-12 absolute.
* J and K of a basis with d shells on several centres from the direct build
  against J, K contracted from the oracle's ERI tensor (THO, exact Boys),
  JK_inmem (basis.rs:462-484).  Bar 1e-11 absolute.
* The same J and K with every class from total angular momentum 5 on the
  streaming kernel (RCHEM_ERI_STREAM=2: one lane per primitive quartet,
  targets summed over the wave) and with none on it (RCHEM_ERI_STREAM=0:
  per-lane and wave-cooperative kernels), and the Schwarz bounds of both.
* The analytic gradient with d shells on two atoms, so (dd|dp) and (dd|dd)
  quartets span two centres and their per-centre derivative programs cannot
  cancel on one atom, against central differences of the GPU energy.
  Bar 1e-6 Eh/bohr.
"""
import os

import numpy as np
import pytest

import oracle as O
from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_EXACT, default_opts
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z, synthetic_shells

pytestmark = pytest.mark.gpu
H = 1e-4


def _say(*a):
    print(*a, flush=True)


def _dbasis(nshell, seed, box):
    c, l, npr, ex, cf = synthetic_shells(nshell=nshell, seed=seed, box=box)
    l = l.copy()
    l[:3] = 2
    l[3:5] = 1
    return c, l, npr, ex, cf


def test_dclass_one_electron_vs_oracle(gpu_ctx):
    sh = _dbasis(nshell=9, seed=21, box=4.0)
    c, l, _, _, _ = sh
    B = rb.Basis.from_shells(*sh, ctx=gpu_ctx)
    # nuclei at three of the shell centres and one free point, fractional charges
    nuc = np.vstack([c[0], c[3], c[6], [0.3, -0.2, 0.9]])
    Z = np.array([8.0, 1.0, 6.0, 2.5])
    _say(f"case: S/T/V of a {B.nbf}-function basis with l = {sorted(set(l.tolist()))}, 4 nuclei")
    S, T, V = rb.one_electron(B, nuc, Z)
    Bo = O.cgtos_from_shells(*sh)
    So, To, Vo = O.one_electron(Bo, Z, nuc)
    # every (la, lb) class with la, lb <= 2 occurs
    lf = np.repeat(l, (l + 1) * (l + 2) // 2)
    assert {(a, b) for a in lf for b in lf} >= {(a, b) for a in range(3) for b in range(3)}
    assert np.max(np.abs(S - So)) < 1e-12
    assert np.max(np.abs(T - To)) < 1e-12
    assert np.max(np.abs(V - Vo)) < 1e-12
    assert np.max(np.abs(V)) > 1.0  # the comparison is not between near-zero matrices


def test_dbasis_jk_vs_oracle_tensor(gpu_ctx):
    sh = _dbasis(nshell=7, seed=5, box=3.0)
    B = rb.Basis.from_shells(*sh, ctx=gpu_ctx)
    n = B.nbf
    _say(f"case: d-basis J/K vs oracle tensor, {len(sh[1])} shells, nbf={n}")
    rng = np.random.default_rng(9)
    X = rng.standard_normal((n, n))
    D = 0.1 * (X + X.T)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    rb.JK_direct(J, K, B, D, default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=1e-14))
    Bo = O.cgtos_from_shells(*sh)
    Jo, Ko = O.jk_inmem(O.eri_tensor(Bo, boys=O.BOYS_EXACT), D)
    assert np.max(np.abs(J - Jo)) < 1e-11
    assert np.max(np.abs(K - Ko)) < 1e-11
    assert np.max(np.abs(Jo)) > 1e-2


def _ctx_with(var, value):
    os.environ[var] = value
    try:
        return rb.Context(0)
    finally:
        del os.environ[var]


@pytest.mark.parametrize("stream", ["0", "2"])
def test_dbasis_jk_engines_vs_oracle_tensor(stream):
    ctx = _ctx_with("RCHEM_ERI_STREAM", stream)
    try:
        sh = _dbasis(nshell=8, seed=11, box=3.5)
        B = rb.Basis.from_shells(*sh, ctx=ctx)
        n = B.nbf
        _say(f"case: d-basis J/K vs oracle tensor, RCHEM_ERI_STREAM={stream}, nbf={n}")
        rng = np.random.default_rng(17)
        X = rng.standard_normal((n, n))
        D = 0.1 * (X + X.T)
        J = np.zeros((n, n)); K = np.zeros((n, n))
        rb.JK_direct(J, K, B, D, default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=1e-14))
        Bo = O.cgtos_from_shells(*sh)
        Jo, Ko = O.jk_inmem(O.eri_tensor(Bo, boys=O.BOYS_EXACT), D)
        assert np.max(np.abs(J - Jo)) < 1e-11
        assert np.max(np.abs(K - Ko)) < 1e-11
        assert np.max(np.abs(Jo)) > 1e-2
        # the Schwarz bounds sqrt(max |(ab|ab)|) of the diagonal mode
        Q = rb.schwarz(B)
        c, l, npr, ex, cf = sh
        f0 = np.concatenate([[0], np.cumsum((l + 1) * (l + 2) // 2)])
        I = O.eri_tensor(Bo, boys=O.BOYS_EXACT)
        for a in range(len(l)):
            for b in range(a + 1):
                blk = I[f0[a]:f0[a + 1], f0[b]:f0[b + 1], f0[a]:f0[a + 1], f0[b]:f0[b + 1]]
                q = np.sqrt(np.max(np.abs(np.einsum("ijij->ij", blk))))
                assert abs(Q[a, b] - q) < 1e-12 * max(1.0, q)
    finally:
        ctx.close()


def test_dshell_gradient_two_centres_vs_fd(gpu_ctx):
    """Water STO-3G with a d shell on O (exponent 0.8) and one on H1
    (exponent 1.1): the (d_O d_H1|..) pairs survive screening, so (dd|dp) and
    (dd|dd) quartets with two centres reach the gradient."""
    Z = np.array(WATER_Z)
    xyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    base = rb.Basis.new(Z, xyz, "STO-3G", ctx=gpu_ctx)
    c0, l0, n0, e0, f0 = base.shells()
    owner = [int(np.argmin(np.sum((xyz - c) ** 2, axis=1))) for c in c0]

    def basis_at(p):
        cent = np.vstack([p[owner], p[0], p[1]])
        return rb.Basis.from_shells(cent, np.append(l0, [2, 2]), np.append(n0, [1, 1]),
                                    np.append(e0, [0.8, 1.1]), np.append(f0, [1.0, 1.0]),
                                    ctx=gpu_ctx)

    o = default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=1e-14)

    def energy(p):
        res, _ = rb.scf(basis_at(p), p, Z, nocc=5, thresh_e=1e-12, diis=6, opts=o)
        assert res.converged
        return res.e_total

    B = basis_at(xyz)
    Q = rb.schwarz(B)
    nsh = len(l0) + 2
    dO, dH = nsh - 2, nsh - 1
    # the two-centre d pair and its (dd|dd) quartet are far above the 1e-14 cut
    assert Q[dH, dO] ** 2 > 1e-6
    _say("case: water STO-3G + d on O and H1, analytic gradient vs GPU finite differences")
    res, g = rb.rhf_gradient(B, xyz, Z, nocc=5, thresh_e=1e-12, diis=6, opts=o)
    assert res.converged
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8
    assert np.max(np.abs(np.cross(xyz, g).sum(axis=0))) < 1e-8
    fd = np.zeros_like(g)
    for a in range(3):
        for x in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[a, x] += H
            xm[a, x] -= H
            fd[a, x] = (energy(xp) - energy(xm)) / (2 * H)
    assert np.max(np.abs(g - fd)) < 1e-6
