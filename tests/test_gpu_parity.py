"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden vectors.

Bars (written in each test):
  * ERIs in FP64 within 1e-12 absolute of the oracle / reference (compat Boys
    reproduces the reference's NR incomplete gamma; exact Boys is compared to
    the oracle's exact mode);
  * SCF electronic energy within 1e-8 Eh of the oracle's restatement of
    rchem's driver on the same inputs;
  * size-independent properties at the full (H2O)_64 size.
"""
import os

import numpy as np
import pytest

import oracle as O
from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_COMPAT, BOYS_EXACT, default_opts
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z, water_cluster

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DATA = os.path.join(os.path.dirname(os.path.dirname(__file__)), "rchem_amd", "data")


def _oracle_basis(gbasis):
    xyz, l, npr, ex, cf = gbasis.shells()
    return O.cgtos_from_shells(xyz, l, npr, ex, cf)


def _say(*a):
    print(*a, flush=True)


def test_water_eri_tensor_compat_vs_reference(gpu_ctx):
    _say("case: water STO-3G build_I, compat Boys, nbf=7")
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    assert B.nbf == 7
    I = rb.build_I(B, default_opts(boys_mode=BOYS_COMPAT, schwarz_thresh=0.0))
    g = np.load(os.path.join(GOLD, "water_sto3g.npz"))
    assert np.max(np.abs(I - g["eri"])) < 1e-12      # reference cints contr_coulomb
    Io = O.eri_tensor(_oracle_basis(B), boys=O.BOYS_COMPAT)
    assert np.max(np.abs(I - Io)) < 1e-12


def test_water_eri_tensor_exact_vs_oracle(gpu_ctx):
    _say("case: water STO-3G build_I, exact Boys")
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    I = rb.build_I(B, default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=0.0))
    Io = O.eri_tensor(_oracle_basis(B), boys=O.BOYS_EXACT)
    assert np.max(np.abs(I - Io)) < 1e-12


# every (la lb|lc ld) class with l <= 2: per-lane kernels up to L = 4, the
# wave-cooperative kernel above
ALL_CLASSES = {"0000", "1000", "1010", "1100", "1110", "1111", "2000", "2010", "2011", "2020",
               "2100", "2110", "2111", "2120", "2121", "2200", "2210", "2211", "2220", "2221",
               "2222"}


def _canonical_class(ls):
    """Class key of a shell quartet as the library orders it (higher l first
    within a pair, then the pair of higher pair-class index as bra)."""
    a, b = sorted(ls[:2], reverse=True)
    c, d = sorted(ls[2:], reverse=True)
    pc = lambda x, y: x * (x + 1) // 2 + y
    if pc(a, b) < pc(c, d):
        a, b, c, d = c, d, a, b
    return f"{a}{b}{c}{d}"


def test_contracted_samples_all_classes(gpu_ctx):
    """Every compiled (la lb|lc ld) class against the reference's contr_coulomb
    values (golden vectors from the compiled reference C), shell quartet by
    shell quartet, in the quartets' own (unsorted) shell order."""
    g = np.load(os.path.join(GOLD, "contracted_samples.npz"))
    ls = g["ls"]
    sel = list(range(len(g["quartets"])))
    classes = {_canonical_class([ls[s] for s in g["quartets"][qi]]) for qi in sel}
    _say(f"case: contracted samples, {len(sel)} quartets over {len(classes)} classes")
    assert classes == ALL_CLASSES
    B = rb.Basis.from_shells(g["centres"], ls, g["nprim"], g["exps"], g["coefs"], ctx=gpu_ctx)
    blocks = rb.eri_shell_quartets(B, g["quartets"][sel],
                                   default_opts(boys_mode=BOYS_COMPAT, schwarz_thresh=0.0))
    qof = g["quartet_of"]
    worst = 0.0
    for qi, blk in zip(sel, blocks):
        worst = max(worst, np.max(np.abs(blk.reshape(-1) - g["values"][qof == qi])))
    assert worst < 1e-12


def test_sampled_quartets_h2o64_vs_oracle(gpu_ctx):
    """Full (H2O)_64 6-31G size: 150 random shell quartets against the oracle's
    THO integrals with the reference's Boys function."""
    Z, xyz = water_cluster(dims=(4, 4, 4), seed=64)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    sh = B.shells()
    ns = len(sh[1])
    _say(f"case: (H2O)_64 sampled shell quartets, {ns} shells")
    rng = np.random.default_rng(2026)
    q = rng.integers(0, ns, (150, 4))
    q[:20, 1] = q[:20, 0]          # some same-shell pairs
    q[20:30, 2:] = q[20:30, :2]    # some (ab|ab) diagonals
    blocks = rb.eri_shell_quartets(B, q, default_opts(boys_mode=BOYS_COMPAT, schwarz_thresh=0.0))
    Bo = _oracle_basis(B)
    f0 = np.concatenate([[0], np.cumsum((sh[1] + 1) * (sh[1] + 2) // 2)])
    worst = 0.0
    for r, blk in zip(q, blocks):
        ref = np.zeros(blk.shape)
        for a in range(blk.shape[0]):
            for b in range(blk.shape[1]):
                for c in range(blk.shape[2]):
                    for d in range(blk.shape[3]):
                        ref[a, b, c, d] = O.contracted_eri(Bo, f0[r[0]] + a, f0[r[1]] + b,
                                                           f0[r[2]] + c, f0[r[3]] + d,
                                                           boys=O.BOYS_COMPAT)
        worst = max(worst, np.max(np.abs(blk - ref)))
    assert worst < 1e-12


def test_water_one_electron_vs_oracle(gpu_ctx):
    _say("case: water STO-3G S/T/V")
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    S, T, V = rb.one_electron(B, WATER_CRAWFORD, WATER_Z)
    So, To, Vo = O.one_electron(_oracle_basis(B), WATER_Z, WATER_CRAWFORD)
    assert np.max(np.abs(S - So)) < 1e-13
    assert np.max(np.abs(T - To)) < 1e-12
    assert np.max(np.abs(V - Vo)) < 1e-12
    g = np.load(os.path.join(GOLD, "water_sto3g.npz"))
    assert np.max(np.abs(S - g["S"])) < 1e-13
    assert np.max(np.abs(T - g["T"])) < 1e-12


@pytest.mark.parametrize("basis_name,boys", [("STO-3G", BOYS_COMPAT), ("6-31G", BOYS_EXACT)])
def test_jk_water_dimer_vs_oracle(gpu_ctx, basis_name, boys):
    Z, xyz = water_cluster(dims=(2, 1, 1), seed=11)
    B = rb.Basis.new(Z, xyz, basis_name, ctx=gpu_ctx)
    _say(f"case: (H2O)_2 {basis_name} JK_direct, nbf={B.nbf}, boys={boys}")
    n = B.nbf
    rng = np.random.default_rng(5)
    D = rng.normal(size=(n, n))
    D = 0.5 * (D + D.T)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    rb.JK_direct(J, K, B, D, default_opts(boys_mode=boys, schwarz_thresh=0.0))
    Io = O.eri_tensor(_oracle_basis(B), boys=O.BOYS_COMPAT if boys == BOYS_COMPAT else O.BOYS_EXACT)
    Jo, Ko = O.jk_inmem(Io, D)
    assert np.max(np.abs(J - Jo)) < 1e-11
    assert np.max(np.abs(K - Ko)) < 1e-11
    # Schwarz screening at 1e-12 changes J/K by at most ~1e-12 * sum|D| per element
    J2 = np.zeros((n, n)); K2 = np.zeros((n, n))
    rb.JK_direct(J2, K2, B, D, default_opts(boys_mode=boys, schwarz_thresh=1e-12))
    assert np.max(np.abs(J2 - J)) < 1e-12 * np.abs(D).sum()
    assert np.max(np.abs(K2 - K)) < 1e-12 * np.abs(D).sum()


def test_scf_water_energy_vs_oracle(gpu_ctx):
    _say("case: water STO-3G SCF (rchem driver), compat Boys")
    B = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=gpu_ctx)
    res, D = rb.scf(B, WATER_CRAWFORD, WATER_Z, nocc=5,
                    opts=default_opts(boys_mode=BOYS_COMPAT, schwarz_thresh=0.0))
    assert res.converged
    Bo = _oracle_basis(B)
    S, T, V = O.one_electron(Bo, WATER_Z, WATER_CRAWFORD)
    I = O.eri_tensor(Bo, boys=O.BOYS_COMPAT)
    r = O.scf(S, T + V, lambda Dm: O.jk_inmem(I, Dm), nocc=5)
    assert abs(res.e_elec - r.e_elec) < 1e-8
    assert abs(res.e_nuc - 8.00236706181077) < 1e-12
    assert np.max(np.abs(D - r.D)) < 1e-6


def test_jk_h2o64_properties(gpu_ctx):
    """Full (H2O)_64 6-31G size: symmetry, linearity in D, positivity of the
    Coulomb energy -- properties that do not need the oracle at this size."""
    Z, xyz = water_cluster(dims=(4, 4, 4), seed=64)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    n = B.nbf
    _say(f"case: (H2O)_64 6-31G JK properties, nbf={n}")
    assert n == 832
    rng = np.random.default_rng(64)
    C = rng.normal(size=(n, 320)) / np.sqrt(n)
    D1 = C @ C.T                      # positive semidefinite
    D2 = rng.normal(size=(n, n)) * 1e-3
    D2 = D2 + D2.T
    J1 = np.zeros((n, n)); K1 = np.zeros((n, n))
    J2 = np.zeros((n, n)); K2 = np.zeros((n, n))
    J3 = np.zeros((n, n)); K3 = np.zeros((n, n))
    rb.JK_direct(J1, K1, B, D1)
    rb.JK_direct(J2, K2, B, D2)
    rb.JK_direct(J3, K3, B, 2.0 * D1 - 3.0 * D2)
    st = rb.jk_last_stats(B)
    _say(f"  quartets={st.quartets:.4g} prim_quartets={st.primitive_quartets:.4g}")
    assert np.max(np.abs(J1 - J1.T)) == 0.0 and np.max(np.abs(K1 - K1.T)) == 0.0
    scale = np.abs(J1).max()
    assert np.max(np.abs(J3 - (2 * J1 - 3 * J2))) < 1e-10 * scale
    assert np.max(np.abs(K3 - (2 * K1 - 3 * K2))) < 1e-10 * np.abs(K1).max()
    assert np.sum(D1 * J1) > 0.0      # Coulomb self-energy of a PSD density
    assert np.sum(D1 * K1) > 0.0


def test_mixed_precision_h2o64(gpu_ctx):
    """FP32 quartets below Q_ab Q_cd max|D| < 1e-7: the two-electron energy
    moves by less than 1e-6 Eh from the all-FP64 build."""
    from rchem_amd._lib import PREC_MIXED
    Z, xyz = water_cluster(dims=(4, 4, 4), seed=64)
    B = rb.Basis.new(Z, xyz, "6-31G", ctx=gpu_ctx)
    n = B.nbf
    _say(f"case: (H2O)_64 6-31G mixed vs FP64, nbf={n}")
    S, T, V = rb.one_electron(B, xyz, Z)
    lam, U = np.linalg.eigh(S)
    X = U @ np.diag(lam ** -0.5) @ U.T
    _, Cp = np.linalg.eigh(X.T @ (T + V) @ X)
    C = X @ Cp
    D = C[:, :320] @ C[:, :320].T
    J = np.zeros((n, n)); K = np.zeros((n, n))
    Jm = np.zeros((n, n)); Km = np.zeros((n, n))
    rb.JK_direct(J, K, B, D)
    rb.JK_direct(Jm, Km, B, D, default_opts(precision=PREC_MIXED, mixed_thresh=1e-7))
    st = rb.jk_last_stats(B)
    _say(f"  quartets={st.quartets:.4g} fp32={st.quartets_fp32:.4g}")
    assert st.quartets_fp32 > 0.1 * st.quartets
    e64 = np.sum(D * (2 * J - K))
    e32 = np.sum(D * (2 * Jm - Km))
    _say(f"  E2(fp64)={e64:.10f} E2(mixed)={e32:.10f} diff={e32 - e64:.3e}")
    assert abs(e32 - e64) < 1e-6
    assert np.max(np.abs(Jm - J)) < 1e-6 and np.max(np.abs(Km - K)) < 1e-6


def test_microbench_basis_sampled_quartets_every_class(gpu_ctx):
    """The synthetic 2000-shell s/p/d microbenchmark basis: per (la lb|lc ld)
    class, shell quartets drawn at random among the Schwarz-surviving ones
    (Q_ab Q_cd >= 1e-12), against the oracle's THO integrals with the
    reference's Boys function (bar 1e-12 absolute)."""
    from rchem_amd.inputs import synthetic_shells
    sh = synthetic_shells(nshell=2000, seed=2000, box=20.0, nprim=3)
    B = rb.Basis.from_shells(*sh, ctx=gpu_ctx)
    Q = rb.schwarz(B)
    Q = np.maximum(Q, Q.T)
    ls = sh[1]
    rng = np.random.default_rng(99)
    picked = {}
    tries = 0
    while len(picked) < 21 * 2 and tries < 200000:
        tries += 1
        q = rng.integers(0, len(ls), 4)
        if Q[q[0], q[1]] * Q[q[2], q[3]] < 1e-12:
            continue
        key = _canonical_class([ls[s] for s in q])
        picked.setdefault(key, [])
        if len(picked[key]) < 2:
            picked[key].append(q)
    quartets = np.array([q for v in picked.values() for q in v])
    _say(f"case: microbench basis, {len(quartets)} quartets over {len(picked)} classes")
    assert set(picked) == ALL_CLASSES
    blocks = rb.eri_shell_quartets(B, quartets, default_opts(boys_mode=BOYS_COMPAT,
                                                             schwarz_thresh=0.0))
    Bo = O.cgtos_from_shells(*sh)
    f0 = np.concatenate([[0], np.cumsum((ls + 1) * (ls + 2) // 2)])
    worst = 0.0
    for r, blk in zip(quartets, blocks):
        ref = np.zeros(blk.shape)
        for a in range(blk.shape[0]):
            for b in range(blk.shape[1]):
                for c in range(blk.shape[2]):
                    for d in range(blk.shape[3]):
                        ref[a, b, c, d] = O.contracted_eri(Bo, f0[r[0]] + a, f0[r[1]] + b,
                                                           f0[r[2]] + c, f0[r[3]] + d,
                                                           boys=O.BOYS_COMPAT)
        worst = max(worst, np.max(np.abs(blk - ref)))
    assert worst < 1e-12


def test_boys_grid_edges_through_eris_vs_oracle(gpu_ctx):
    """The exact Boys function at its grid's edge cases, reached through
    integrals.  One-primitive s and p shells (exponent 1/2) share a centre at
    x_k = sqrt(2 T_k) on a line, an s shell sits at the origin: (s_k s_k|s_0 s_0),
    (p_k s_k|s_0 s_0) and (p_k p_k|s_0 s_0) then have zeta = eta = 1, P - Q = x_k
    and T = rho |P - Q|^2 = T_k to a few ulps, in the orders 0, 1 and 2 (the
    order-1 Taylor pass and the downward recursion).  T_k: midpoints between
    grid points (the ties of the rounding to the 1/8 grid, met from either
    side), grid points, 0, and either side of the switch to the asymptote at
    T = 36.  A row index that disagrees with its grid point T0, or a wrong
    switch, shows at once; an index one row off alone would not (the Taylor
    series still converges there), so this pins the grid's bookkeeping, not
    its rounding direction.  Bar: 1e-12 relative to the oracle's exact mode."""
    T = [0.0, 1 / 16, 3 / 16, 15 / 16, 17 / 16, 127 / 16, 129 / 16, 255 / 16, 401 / 16,
         573 / 16, 575 / 16, 0.125, 1.0, 12.0, 35.875, 36.0 - 1e-9, 36.0, 36.0 + 1e-9, 40.0,
         90.0]
    xs = np.sqrt(2.0 * np.array(T))
    centres = [[0.0, 0.0, 0.0]]
    ls = [0]
    for x in xs:
        centres += [[x, 0.0, 0.0], [x, 0.0, 0.0]]
        ls += [0, 1]
    centres = np.array(centres)
    ls = np.array(ls, dtype=np.int32)
    npr = np.ones(len(ls), dtype=np.int32)
    ex = np.full(len(ls), 0.5)
    cf = np.ones(len(ls))
    _say(f"case: Boys grid edges, {len(T)} values of T through (ss|ss), (ps|ss), (pp|ss)")
    B = rb.Basis.from_shells(centres, ls, npr, ex, cf, ctx=gpu_ctx)
    q = []
    for k in range(len(T)):
        s, p = 1 + 2 * k, 2 + 2 * k
        q += [(s, s, 0, 0), (p, p, 0, 0)]
        if T[k] > 0:  # at T = 0 every centre coincides and (p s|s s) vanishes by symmetry
            q.append((p, s, 0, 0))
    q = np.array(q, dtype=np.int32)
    blocks = rb.eri_shell_quartets(B, q, default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=0.0))
    Bo = O.cgtos_from_shells(centres, ls, npr, ex, cf)
    f0 = np.concatenate([[0], np.cumsum((ls + 1) * (ls + 2) // 2)])
    worst = 0.0
    for r, blk in zip(q, blocks):
        ref = np.zeros(blk.shape)
        for a in range(blk.shape[0]):
            for b in range(blk.shape[1]):
                ref[a, b, 0, 0] = O.contracted_eri(Bo, f0[r[0]] + a, f0[r[1]] + b, f0[r[2]],
                                                   f0[r[3]], boys=O.BOYS_EXACT)
        scale = np.max(np.abs(ref))
        assert scale > 1e-3  # every quartet is far from zero (F_0 >= 0.14 for T <= 40)
        worst = max(worst, np.max(np.abs(blk - ref)) / scale)
    assert worst < 1e-12
