"""CPU checks of the wave-cooperative kernel's operation tables.

The tables (rchem_amd/csrc/gen_eri_prog.py) are executed here in numpy with
the kernel's semantics -- one dependency level at a time, every operation of a
level reading the state left by the levels before it, slots reused as the
generator allocated them -- on random shell quartets of every l <= 2 class,
with the library's primitive-pair formulas, and compared with the oracle's
contracted THO integrals (exact Boys).  Every slot, accumulator and output
index is also checked against the buffer sizes the kernel allocates, so a
table that would read or write outside its LDS buffer fails here first.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rchem_amd", "csrc")
sys.path.insert(0, CSRC)
from gen_eri import classes  # noqa: E402
from gen_eri_prog import eri_program, grad_program, lds_bytes  # noqa: E402


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def _prims(A, B, la, lb, ea, eb, ca, cb):
    """Primitive pairs as the library forms them (make_prims, radial norm of
    basis.rs:140-149 folded into the coefficients), with the first exponent."""
    nr = lambda a, l: math.sqrt(2.0 ** (2 * l + 1.5) * a ** (l + 1.5) / math.pi ** 1.5)
    AB2 = float(np.sum((A - B) ** 2))
    out = []
    for al, c1 in zip(ea, ca):
        for be, c2 in zip(eb, cb):
            z = al + be
            P = (al * A + be * B) / z
            kab = (c1 * nr(al, la)) * (c2 * nr(be, lb)) * math.sqrt(2.0) * math.pi ** 1.25 / z \
                * math.exp(-al * be / z * AB2)
            out.append((z, P, kab, 0.5 / z, al))
    return out


def _execute(pr, la, lb, lc, ld, shells):
    """Runs a program with the kernel's semantics; returns its outputs."""
    L, nacc, nv, nh = pr["L"], pr["nacc"], pr["nv"], pr["nh"]
    (A, ea, ca), (B, eb, cb), (C, ec, cc), (D, ed, cd) = shells
    bra = _prims(A, B, la, lb, ea, eb, ca, cb)
    ket = _prims(C, D, lc, ld, ec, ed, cc, cd)
    acc = np.zeros(nacc)
    for (zb, P, kb, oo2z, aa) in bra:
        for (zk, Q, kk, oo2e, ac) in ket:
            rs = 1.0 / math.sqrt(zb + zk)
            ooze = rs * rs
            roz, roe = zk * ooze, zb * ooze
            PQ = P - Q
            T = zb * roz * float(PQ @ PQ)
            coef = {0: (P - A, -roz * PQ, oo2z, roz), 1: (Q - C, roe * PQ, oo2e, roe)}
            oz = 0.5 * ooze
            wts = [1.0, 2.0 * aa, 2.0 * (zb - aa), 2.0 * ac]
            V = np.full(nv, np.nan)
            V[0] = 0.0
            for m in range(L + 1):
                V[1 + m] = kb * kk * rs * O.boys_exact(m, T)
            written = set()
            snap = V.copy()
            for g in pr["vgroups"]:
                b, e, flags = g[0], g[1], g[2]
                side, ax = flags & 1, (flags >> 1) & 3
                c0, c1, o, r = coef[side][0][ax], coef[side][1][ax], coef[side][2], coef[side][3]
                for w0, w1, w2, w3 in pr["vops"][b:e]:
                    dst, s0, s1, s2 = w0 & 0xFFFF, w0 >> 16, w1 & 0xFFFF, w1 >> 16
                    s3, s4 = w2 & 0xFFFF, w2 >> 16
                    n1, n2 = w3 & 15, (w3 >> 4) & 15
                    for s_ in (dst, s0, s1, s2, s3, s4):
                        assert 0 <= s_ < nv
                    assert dst >= L + 2, "an operation overwrites the zero or a G_m slot"
                    for s_ in (s0, s1, s2, s3, s4):
                        assert s_ not in written, "read of a slot written in the same level"
                    val = c0 * snap[s0] + c1 * snap[s1] + (n1 * o) * (snap[s2] - r * snap[s3]) \
                        + (n2 * oz) * snap[s4]
                    assert np.isfinite(val)
                    V[dst] = val
                    written.add(dst)
                if flags & 8:
                    written = set()
                    snap = V.copy()
            for x, ai in pr["accs"]:
                slot, st = x & 0xFFFF, x >> 16
                assert 0 <= slot < nv and 0 <= ai < nacc and 0 <= st < 4
                assert np.isfinite(V[slot])
                acc[ai] += wts[st] * V[slot]
    H = np.full(nh, np.nan)
    H[:nacc] = acc
    snap = H.copy()
    written = set()
    for g in pr["hgroups"]:
        b, e, flags = g[0], g[1], g[2]
        side, ax = flags & 1, (flags >> 1) & 3
        X = (A - B)[ax] if side == 0 else (C - D)[ax]
        for w0, w1 in pr["hops"][b:e]:
            dst, s0, s1 = w0 & 0xFFFF, w0 >> 16, w1
            for s_ in (dst, s0, s1):
                assert 0 <= s_ < nh
            assert s0 not in written and s1 not in written
            H[dst] = snap[s0] + X * snap[s1]
            written.add(dst)
        if flags & 8:
            written = set()
            snap = H.copy()
    out = np.array([nrm * (H[pl] - n * H[mi]) for pl, mi, n, nrm in pr["outs"]])
    assert np.all(np.isfinite(out))
    assert 0 < pr["P"] <= 32 and lds_bytes(L, nacc, nv, nh, pr["P"]) <= 64 * 1024
    return out


def _random_shells(ls, seed):
    rng = np.random.default_rng(seed)
    shells = []
    for l in ls:
        centre = rng.uniform(-1.0, 1.0, 3)
        exps = np.exp(rng.uniform(math.log(0.2), math.log(8.0), 2))
        coefs = rng.uniform(0.2, 1.0, 2)
        shells.append((centre, exps, coefs))
    return shells


def _oracle_block(ls, shells):
    basis = O.cgtos_from_shells(np.array([s[0] for s in shells]), np.array(ls, dtype=np.int32),
                                np.full(4, 2, dtype=np.int32),
                                np.concatenate([s[1] for s in shells]),
                                np.concatenate([s[2] for s in shells]))
    f0 = np.concatenate([[0], np.cumsum([_ncart(l) for l in ls])])
    ref = []
    for a in range(_ncart(ls[0])):
        for b in range(_ncart(ls[1])):
            for c in range(_ncart(ls[2])):
                for d in range(_ncart(ls[3])):
                    ref.append(O.contracted_eri(basis, f0[0] + a, f0[1] + b, f0[2] + c, f0[3] + d,
                                                boys=O.BOYS_EXACT))
    return np.array(ref)


@pytest.mark.parametrize("key", ["".join(map(str, c)) for c in classes(2)])
def test_program_matches_oracle(key):
    la, lb, lc, ld = (int(x) for x in key)
    ls = [la, lb, lc, ld]
    shells = _random_shells(ls, int(key) + 11)
    got = _execute(eri_program(la, lb, lc, ld), la, lb, lc, ld, shells)
    ref = _oracle_block(ls, shells)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("key", ["0000", "1000", "1010", "1100", "1110", "1111", "2000", "2110",
                                 "2111"])
def test_grad_program_matches_finite_differences(key):
    """Derivative integrals d/dA, d/dB, d/dC against central differences
    (h = 1e-4 bohr) of the oracle's contracted integrals."""
    la, lb, lc, ld = (int(x) for x in key)
    ls = [la, lb, lc, ld]
    shells = _random_shells(ls, int(key) + 29)
    pr = grad_program(la, lb, lc, ld)
    got = _execute(pr, la, lb, lc, ld, shells).reshape(3, 3, -1)
    h = 1e-4
    for centre in range(3):
        for i in range(3):
            sp = [list(s) for s in shells]
            sm = [list(s) for s in shells]
            sp[centre][0] = shells[centre][0].copy()
            sm[centre][0] = shells[centre][0].copy()
            sp[centre][0][i] += h
            sm[centre][0][i] -= h
            fd = (_oracle_block(ls, sp) - _oracle_block(ls, sm)) / (2 * h)
            scale = max(1.0, np.max(np.abs(fd)))
            assert np.max(np.abs(got[centre, i] - fd)) < 2e-7 * scale, (centre, i)


def _run_straight(g, la, lb, lc, ld, shells):
    """Executes the straight-line per-lane gradient code (gen_eri_grad.py),
    read as Python statements, over the same primitive loop as _execute."""
    def py(lines):
        out = []
        for s in lines:
            s = s.rstrip(";").replace("const R ", "").replace("R(", "float(")
            out.append(s)
        return compile("\n".join(out), "<generated>", "exec")
    vrr, hrr = py(g["vrr"]), py(g["hrr"])
    (A, ea, ca), (B, eb, cb), (C, ec, cc), (D, ed, cd) = shells
    bra = _prims(A, B, la, lb, ea, eb, ca, cb)
    ket = _prims(C, D, lc, ld, ec, ed, cc, cd)
    acc = [0.0] * g["nacc"]
    for (zb, P, kb, oo2z, aa) in bra:
        for (zk, Q, kk, oo2e, ac) in ket:
            rs = 1.0 / math.sqrt(zb + zk)
            ooze = rs * rs
            roz, roe = zk * ooze, zb * ooze
            PQ = P - Q
            T = zb * roz * float(PQ @ PQ)
            env = dict(G=[kb * kk * rs * O.boys_exact(m, T) for m in range(g["L"] + 1)],
                       acc=acc, oo2z=oo2z, oo2e=oo2e, oo2ze=0.5 * ooze, roz=roz, roe=roe,
                       wA=2.0 * aa, wB=2.0 * (zb - aa), wC=2.0 * ac, float=float)
            for i, x in enumerate("xyz"):
                env["PA" + x], env["WP" + x] = (P - A)[i], -roz * PQ[i]
                env["QC" + x], env["WQ" + x] = (Q - C)[i], roe * PQ[i]
            exec(vrr, env)
    out = [0.0] * g["nout"]
    env = dict(acc=acc, out=out, float=float)
    for i, x in enumerate("xyz"):
        env["AB" + x], env["CD" + x] = (A - B)[i], (C - D)[i]
    exec(hrr, env)
    return np.array(out)


@pytest.mark.parametrize("cls", __import__("gen_eri_grad").GRAD_CLASSES)
def test_straight_line_gradient_code_matches_tables(cls):
    """The per-lane gradient code is a translation of grad_program's tables
    (checked against finite differences above): same outputs."""
    from gen_eri_grad import translate
    la, lb, lc, ld = cls
    shells = _random_shells(list(cls), 97 + la + 2 * lb + 4 * lc + 8 * ld)
    ref = _execute(grad_program(la, lb, lc, ld), la, lb, lc, ld, shells)
    got = _run_straight(translate(la, lb, lc, ld), la, lb, lc, ld, shells)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("key", ["1100", "2111"])
def test_per_centre_grad_programs_match_full_program(key):
    """A per-centre gradient program (the form the kernel runs where the full
    program exceeds a workgroup's LDS) gives exactly that centre's slice of
    the full program's outputs."""
    la, lb, lc, ld = (int(x) for x in key)
    shells = _random_shells([la, lb, lc, ld], int(key) + 71)
    full = _execute(grad_program(la, lb, lc, ld), la, lb, lc, ld, shells).reshape(3, -1)
    for centre in range(3):
        part = _execute(grad_program(la, lb, lc, ld, centre), la, lb, lc, ld, shells)
        assert np.max(np.abs(part - full[centre])) < 1e-13 * max(1.0, np.max(np.abs(full)))


@pytest.mark.parametrize("key", ["2221", "2222"])
def test_per_centre_grad_programs_finite_differences(key):
    """(dd|dp) and (dd|dd): per-centre derivative integrals against central
    differences (h = 1e-4 bohr) of the oracle's contracted integrals."""
    la, lb, lc, ld = (int(x) for x in key)
    ls = [la, lb, lc, ld]
    shells = _random_shells(ls, int(key) + 29)
    h = 1e-4
    for centre in range(3):
        pr = grad_program(la, lb, lc, ld, centre)
        got = _execute(pr, la, lb, lc, ld, shells).reshape(3, -1)
        for i in range(3):
            sp = [list(s) for s in shells]
            sm = [list(s) for s in shells]
            sp[centre][0] = shells[centre][0].copy()
            sm[centre][0] = shells[centre][0].copy()
            sp[centre][0][i] += h
            sm[centre][0][i] -= h
            fd = (_oracle_block(ls, sp) - _oracle_block(ls, sm)) / (2 * h)
            scale = max(1.0, np.max(np.abs(fd)))
            assert np.max(np.abs(got[i] - fd)) < 2e-7 * scale, (centre, i)
