"""Pins the CPU oracle before anything is compared against it (no GPU).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a swapped operand) fails it:
  * the reference's own golden constants (os86.rs / tho66.rs tests);
  * the reference's compiled C (golden vectors under tests/golden, written by
    tests/golden/make_golden.py from oracle/_ref, and oracle/_ref itself when
    present);
  * closed forms and invariants (Boys limits and recursion, mpmath, E_nuc,
    ERI permutational symmetry, Tr(DS) = nocc, idempotency, THO vs HGP).
"""
import os

import numpy as np
import pytest

import oracle as O
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DATA = os.path.join(os.path.dirname(os.path.dirname(__file__)), "rchem_amd", "data")

RA = [1.0, 0.0, 1.0]
RB = [0.0, 1.0, 2.0]
RC = [0.0, 0.0, 3.0]
RD = [0.0, 0.0, 4.0]


# ---- Boys functions -------------------------------------------------------

def test_fgamma_compat_matches_reference_grid():
    g = np.load(os.path.join(GOLD, "fgamma_grid.npz"))
    ours = np.array([O.fgamma_compat(m, x) for m, x in zip(g["m"], g["x"])])
    assert np.max(np.abs(ours - g["F"]) / np.maximum(np.abs(g["F"]), 1e-300)) < 1e-14


def test_boys_exact_closed_forms():
    from math import erf, exp, pi, sqrt
    for m in range(17):
        assert abs(O.boys_exact(m, 0.0) - 1.0 / (2 * m + 1)) < 1e-16
    for x in [1e-6, 0.1, 0.7, 1.0, 3.3, 10.0, 25.0, 35.9, 36.1, 60.0, 119.0, 121.0, 300.0]:
        f0 = 0.5 * sqrt(pi / x) * erf(sqrt(x))
        assert abs(O.boys_exact(0, x) - f0) <= 2e-16 * f0
        for m in range(16):
            # downward recursion identity F_m = (2x F_{m+1} + e^-x)/(2m+1)
            lhs = O.boys_exact(m, x)
            rhs = (2 * x * O.boys_exact(m + 1, x) + exp(-x)) / (2 * m + 1)
            assert abs(lhs - rhs) <= 4e-16 * lhs


def test_boys_exact_vs_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    for m in [0, 1, 2, 5, 8, 12, 16]:
        for x in [0.05, 0.9, 4.0, 17.0, 33.0, 50.0]:
            ref = mp.hyp1f1(m + 0.5, m + 1.5, -x) / (2 * m + 1)
            assert abs(O.boys_exact(m, x) - float(ref)) <= 3e-16 * float(ref)


def test_fgamma_compat_error_is_the_nr_truncation():
    # The reference's Boys agrees with the exact one only to the NR series
    # tolerance: the compat restatement must carry that error, not remove it.
    diffs = []
    for m in range(9):
        for x in np.linspace(0.01, 40, 97):
            e = O.boys_exact(m, x)
            diffs.append(abs(O.fgamma_compat(m, x) - e) / e)
    assert max(diffs) < 3e-6
    assert max(diffs) > 1e-10


# ---- primitive integrals against the reference's golden constants ---------

def test_coulomb_repulsion_golden_tho66():
    # tho66.rs:288-321 ((ss|ss) through cints.c, norms 1, tolerance 1e-12)
    v = O.coulomb_repulsion(RA + RB + RC + RD, [0] * 12, [1.1, 1.2, 1.3, 1.4], [1, 1, 1, 1],
                            boys=O.BOYS_COMPAT)
    assert abs(v - 0.08608517834596989) < 1e-12


@pytest.mark.parametrize("powers,ref", [
    ([0] * 12, 0.08608517834596989),
    ([1] + [0] * 11, -0.046888060557776973),
    ([2, 1, 0, 1, 0, 0, 1, 0, 0, 0, 1, 0], 1.71817807954e-05),
])
def test_eri_golden_os86(powers, ref):
    # os86.rs:882-937 with the exact Boys function (boys::micb25); the golden
    # 1.71817807954e-05 is printed to 12 significant digits only.
    for hgp in (False, True):
        v = O.coulomb_repulsion(RA + RB + RC + RD, powers, [1.1, 1.2, 1.3, 1.4], [1, 1, 1, 1],
                                boys=O.BOYS_EXACT, hgp=hgp)
        assert abs(v - ref) < 2e-15


OVL = [([0, 0, 0, 0, 0, 0], 0.20373275913014607), ([1, 0, 0, 0, 0, 0], 0.062005622343957505),
       ([1, 1, 0, 1, 1, 0], -0.00043801221837779696), ([2, 1, 0, 1, 1, 0], -0.0002385994651113168)]


@pytest.mark.parametrize("c,ref", OVL)
def test_overlap_golden(c, ref):
    # os86.rs:939-959 / tho66.rs:266-286 (threshold 1e-16 there; 2 ulp here)
    v = O.overlap(1.8, 2.8, [0, 0, 0], [0.5, 0.8, -0.2], c)
    assert abs(v - ref) <= 2 * np.spacing(abs(ref))


@pytest.mark.parametrize("c,ref", [([0, 0, 0, 0, 0, 0], 0.3652714583525358),
                                   ([1, 0, 0, 0, 0, 0], 0.2514265587836556),
                                   ([2, 2, 2, 2, 2, 2], -7.40057384314e-05)])
def test_kinetic_golden(c, ref):
    # os86.rs:961-978; the last value is printed to 12 digits
    v = O.kinetic(1.8, 2.0, [0, 0, 0], [0.5, 0.8, -0.2], c)
    assert abs(v - ref) <= max(4 * np.spacing(abs(ref)), 5e-17)


@pytest.mark.parametrize("c,ref", [([0, 0, 0, 0, 0, 0], -0.49742209545104593),
                                   ([1, 0, 0, 0, 0, 0], -0.15987439458254471),
                                   ([2, 2, 2, 0, 0, 0], -0.003801373531942607),
                                   ([1, 1, 1, 1, 1, 1], 8.8415484347060993e-5)])
def test_nuclear_golden(c, ref):
    # os86.rs:980-1002 (tolerance 1e-15 there)
    v = O.nuclear(1.8, 2.0, [0, 0, 0], [0.5, 0.8, -0.2], [0.5, 0.8, 0.2], c)
    assert abs(v - ref) < 1e-15


# ---- contracted integrals against the reference C (golden vectors) -------

def _sample_basis():
    g = np.load(os.path.join(GOLD, "contracted_samples.npz"))
    return g, O.cgtos_from_shells(g["centres"], g["ls"], g["nprim"], g["exps"], g["coefs"])


def test_contracted_samples_vs_reference():
    g, B = _sample_basis()
    # function index of shell s component k
    f0 = np.concatenate([[0], np.cumsum([(l + 1) * (l + 2) // 2 for l in g["ls"]])])
    got, got_hgp = [], []
    for q in g["quartets"]:
        rng = [range(f0[s], f0[s + 1]) for s in q]
        for a in rng[0]:
            for b in rng[1]:
                for c in rng[2]:
                    for d in rng[3]:
                        got.append(O.contracted_eri(B, a, b, c, d, boys=O.BOYS_COMPAT))
                        got_hgp.append(O.contracted_eri(B, a, b, c, d, boys=O.BOYS_COMPAT,
                                                        hgp=True))
    got, got_hgp = np.array(got), np.array(got_hgp)
    assert np.max(np.abs(got - g["values"])) < 1e-13
    # HRR at contracted vs primitive level: same sums, different rounding
    assert np.max(np.abs(got_hgp - g["values_hgp"])) < 2e-12
    # THO and HGP agree with each other (two independent formulas)
    assert np.max(np.abs(g["values"] - g["values_hgp"])) < 1e-12


def _water():
    bse = O.load_bse_json(os.path.join(DATA, "sto-3g.json"))
    return O.cgtos_from_bse(bse, WATER_Z, WATER_CRAWFORD)


def test_water_tensor_and_one_electron_vs_reference():
    g = np.load(os.path.join(GOLD, "water_sto3g.npz"))
    B = _water()
    I = O.eri_tensor(B, boys=O.BOYS_COMPAT)
    assert np.max(np.abs(I - g["eri"])) < 1e-13
    S, T, V = O.one_electron(B, WATER_Z, WATER_CRAWFORD)
    assert np.max(np.abs(S - g["S"])) < 1e-14
    assert np.max(np.abs(T - g["T"])) < 1e-13
    # 8-fold permutational symmetry
    for p in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.max(np.abs(I - I.transpose(p))) == 0.0


def test_ref_library_random_primitives():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    comps = [O.get_ijk_list(l) for l in range(3)]
    for _ in range(150):
        c = rng.uniform(-1, 1, 12)
        al = rng.uniform(0.1, 3, 4)
        nm = rng.uniform(0.5, 1.5, 4)
        pw = []
        for s in range(4):
            l = int(rng.integers(0, 3))
            pw += list(comps[l][int(rng.integers(0, len(comps[l])))])
        args = []
        for s in range(4):
            args += [c[3 * s], c[3 * s + 1], c[3 * s + 2], nm[s], int(pw[3 * s]),
                     int(pw[3 * s + 1]), int(pw[3 * s + 2]), al[s]]
        v_ref = R.coulomb_repulsion(*args)
        v = O.coulomb_repulsion(c, pw, al, nm, boys=O.BOYS_COMPAT)
        assert abs(v - v_ref) < 1e-14 * max(1.0, abs(v_ref))


# ---- matrix level: JK_direct, build_I/JK_inmem, SCF ------------------------

def test_jk_direct_equals_jk_inmem_water():
    B = _water()
    rng = np.random.default_rng(3)
    D = rng.normal(size=(B.n, B.n))
    D = D + D.T
    J1, K1 = O.jk_direct(B, D)
    J2, K2 = O.jk_inmem(O.eri_tensor(B), D)
    assert np.max(np.abs(J1 - J2)) < 1e-12
    assert np.max(np.abs(K1 - K2)) < 1e-12


def test_scf_water_invariants():
    B = _water()
    S, T, V = O.one_electron(B, WATER_Z, WATER_CRAWFORD)
    I = O.eri_tensor(B)
    r = O.scf(S, T + V, lambda D: O.jk_inmem(I, D), nocc=5)
    assert r.converged
    D = r.D
    assert abs(np.trace(D @ S) - 5.0) < 1e-10           # Tr(DS) = nocc
    assert np.max(np.abs(D @ S @ D - D)) < 1e-10         # idempotency in the S metric
    enuc = O.nuclear_repulsion(WATER_Z, WATER_CRAWFORD)
    assert abs(enuc - 8.00236706181077) < 1e-12          # closed form, this geometry
    # Crawford's published STO-3G energy for this geometry, to the accuracy
    # rchem's unnormalised contractions and NR Boys function allow
    assert abs(r.e_elec + enuc - (-74.942079928192)) < 5e-7
