"""Pins of the conventions every golden vector inherits (no GPU).

* Cartesian component order: the golden vectors and the kernel generators
  take their order from ``oracle.get_ijk_list`` and ``gen_eri.cart``; both are
  held here to the literal vectors of the reference's own tests
  (shell.rs:44-127: get_ijk_list(3), get_shell4(2,0,0,1), get_shell2(3,1)),
  so a wrong d or f ordering cannot pass by being shared.
* Embedded basis-set digits: rchem uses the published contraction
  coefficients raw (basis.rs:201) with per-primitive norms (basis.rs:140-149);
  the published sets are normalised, so every contracted function of every
  element has unit self-overlap.  A mistyped exponent or coefficient breaks
  that (the check is run per shell and per angular momentum).
"""
import importlib.util
import os

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "rchem_amd", "data")

# shell.rs:48-60
IJK3 = [[3, 0, 0], [2, 1, 0], [2, 0, 1], [1, 2, 0], [1, 1, 1], [1, 0, 2],
        [0, 3, 0], [0, 2, 1], [0, 1, 2], [0, 0, 3]]

# shell.rs:67-86: get_shell4(2, 0, 0, 1)
SHELL4_2001 = [
    [2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0], [2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0],
    [2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1], [1, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0],
    [1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0], [1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1],
    [1, 0, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0], [1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0],
    [1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1], [0, 2, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0],
    [0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0], [0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1],
    [0, 1, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0], [0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0],
    [0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1], [0, 0, 2, 0, 0, 0, 0, 0, 0, 1, 0, 0],
    [0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 1, 0], [0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1]]

# shell.rs:93-125: get_shell2(3, 1), the first three of each p triple
SHELL2_31_HEADS = [[3, 0, 0], [2, 1, 0], [2, 0, 1], [1, 2, 0], [1, 1, 1], [1, 0, 2],
                   [0, 3, 0], [0, 2, 1], [0, 1, 2], [0, 0, 3]]


def _gen_eri():
    spec = importlib.util.spec_from_file_location(
        "gen_eri", os.path.join(ROOT, "rchem_amd", "csrc", "gen_eri.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _shell4(order, a, b, c, d):
    return [list(p) + list(q) + list(r) + list(s)
            for p in order(a) for q in order(b) for r in order(c) for s in order(d)]


@pytest.mark.parametrize("name", ["oracle", "gen_eri"])
def test_component_order_matches_reference_vectors(name):
    order = O.get_ijk_list if name == "oracle" else _gen_eri().cart
    assert [list(t) for t in order(3)] == IJK3
    assert _shell4(order, 2, 0, 0, 1) == SHELL4_2001
    s2 = [list(p) + list(q) for p in order(3) for q in order(1)]
    assert len(s2) == 30
    assert [row[:3] for row in s2[::3]] == SHELL2_31_HEADS
    assert all([row[3:] for row in s2[k:k + 3]] == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
               for k in range(0, 30, 3))
    # lower l follow the same rule: s, then x y z, then xx xy xz yy yz zz
    assert [list(t) for t in order(0)] == [[0, 0, 0]]
    assert [list(t) for t in order(1)] == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    assert [list(t) for t in order(2)] == [[2, 0, 0], [1, 1, 0], [1, 0, 1], [0, 2, 0],
                                           [0, 1, 1], [0, 0, 2]]


def test_generator_order_equals_oracle_order_to_l4():
    cart = _gen_eri().cart
    for l in range(5):
        assert [tuple(t) for t in cart(l)] == [tuple(t) for t in O.get_ijk_list(l)]


@pytest.mark.parametrize("fname", ["sto-3g.json", "6-31g.json"])
def test_embedded_basis_sets_unit_self_overlap(fname):
    bse = O.load_bse_json(os.path.join(DATA, fname))
    checked = 0
    for zs, el in bse["elements"].items():
        Z = int(zs)
        B = O.cgtos_from_bse(bse, [Z], np.zeros((1, 3)))
        S, _, _ = O.one_electron(B, [Z], np.zeros((1, 3)))
        d = np.abs(np.diag(S) - 1.0)
        assert d.max() < 1e-8, (fname, Z, d.max())
        # one contracted function per (shell, l, component): all checked
        nfun = sum(len(O.get_ijk_list(l)) for sh in el["electron_shells"]
                   for l in sh["angular_momentum"])
        assert B.n == nfun
        checked += B.n
    assert checked > 0


def test_embedded_basis_shell_layout():
    """Shell structure rchem's function counts depend on: STO-3G H 1 / CNO 5,
    6-31G H 2 / CNO 9 functions, primitives per shell as published."""
    counts = {"sto-3g.json": {1: (1, [3]), 6: (5, [3, 3]), 7: (5, [3, 3]), 8: (5, [3, 3])},
              "6-31g.json": {1: (2, [3, 1]), 6: (9, [6, 3, 1]), 7: (9, [6, 3, 1]),
                             8: (9, [6, 3, 1])}}
    for fname, want in counts.items():
        bse = O.load_bse_json(os.path.join(DATA, fname))
        for Z, (nbf, nprims) in want.items():
            el = bse["elements"][str(Z)]
            shells = el["electron_shells"]
            assert [len(s["exponents"]) for s in shells] == nprims
            B = O.cgtos_from_bse(bse, [Z], np.zeros((1, 3)))
            assert B.n == nbf
