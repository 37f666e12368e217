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
from gen_eri_prog import NONE, hrr_program, vrr_program  # noqa: E402


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def _prims(A, B, la, lb, ea, eb, ca, cb):
    """Primitive pairs as the library forms them (make_prims, radial norm of
    basis.rs:140-149 folded into the coefficients)."""
    nr = lambda a, l: math.sqrt(2.0 ** (2 * l + 1.5) * a ** (l + 1.5) / math.pi ** 1.5)
    AB2 = float(np.sum((A - B) ** 2))
    out = []
    for al, c1 in zip(ea, ca):
        for be, c2 in zip(eb, cb):
            z = al + be
            P = (al * A + be * B) / z
            kab = (c1 * nr(al, la)) * (c2 * nr(be, lb)) * math.sqrt(2.0) * math.pi ** 1.25 / z \
                * math.exp(-al * be / z * AB2)
            out.append((z, P, kab, 0.5 / z))
    return out


def _run_program(la, lb, lc, ld, shells):
    v, h = vrr_program(la, lb, lc, ld), hrr_program(la, lb, lc, ld)
    L, nacc, nv, nh = v["L"], v["nacc"], v["nslots"], h["nslots"]
    (A, ea, ca), (B, eb, cb), (C, ec, cc), (D, ed, cd) = shells
    bra = _prims(A, B, la, lb, ea, eb, ca, cb)
    ket = _prims(C, D, lc, ld, ec, ed, cc, cd)
    acc = np.zeros(nacc)
    ops = v["ops"]
    for (zb, P, kb, oo2z) in bra:
        for (zk, Q, kk, oo2e) in ket:
            rs = 1.0 / math.sqrt(zb + zk)
            ooze = rs * rs
            roz, roe = zk * ooze, zb * ooze
            PQ = P - Q
            T = zb * roz * float(PQ @ PQ)
            coef = {0: (P - A, -roz * PQ, oo2z, roz), 1: (Q - C, roe * PQ, oo2e, roe)}
            oz = 0.5 * ooze
            V = np.full(nv, np.nan)
            V[0] = 0.0
            for m in range(L + 1):
                V[1 + m] = kb * kk * rs * O.boys_exact(m, T)
            written_in_level = set()
            snap = V.copy()
            for g in v["groups"]:
                b, e, flags = g[0], g[1], g[2]
                side, ax = flags & 1, (flags >> 1) & 3
                c0, c1, o, r = coef[side][0][ax], coef[side][1][ax], coef[side][2], coef[side][3]
                for w0, w1, w2, w3 in ops[b:e]:
                    dst, s0, s1, s2 = w0 & 0xFFFF, w0 >> 16, w1 & 0xFFFF, w1 >> 16
                    s3, s4, ai = w2 & 0xFFFF, w2 >> 16, w3 & 0xFFFF
                    n1, n2 = (w3 >> 16) & 15, (w3 >> 20) & 15
                    for s_ in (dst, s0, s1, s2, s3, s4):
                        assert 0 <= s_ < nv
                    assert dst >= L + 2, "an operation overwrites the zero or a G_m slot"
                    for s_ in (s0, s1, s2, s3, s4):
                        assert s_ not in written_in_level, "read of a slot written in the same level"
                    val = c0 * snap[s0] + c1 * snap[s1] + (n1 * o) * (snap[s2] - r * snap[s3]) \
                        + (n2 * oz) * snap[s4]
                    assert np.isfinite(val)
                    V[dst] = val
                    written_in_level.add(dst)
                    if ai != NONE:
                        assert 0 <= ai < nacc
                        acc[ai] += val
                if flags & 8:
                    written_in_level = set()
                    snap = V.copy()
            if not v["groups"]:
                acc[0] += V[1]
    H = np.full(nh, np.nan)
    H[:nacc] = acc
    snap = H.copy()
    written = set()
    for g in h["groups"]:
        b, e, flags = g[0], g[1], g[2]
        side, ax = flags & 1, (flags >> 1) & 3
        X = (A - B)[ax] if side == 0 else (C - D)[ax]
        for w0, w1 in h["ops"][b:e]:
            dst, s0, s1 = w0 & 0xFFFF, w0 >> 16, w1
            for s_ in (dst, s0, s1):
                assert 0 <= s_ < nh
            assert s0 not in written and s1 not in written
            H[dst] = snap[s0] + X * snap[s1]
            written.add(dst)
        if flags & 8:
            written = set()
            snap = H.copy()
    out = np.array([nrm * H[src] for src, nrm in h["out"]])
    assert len(out) == _ncart(la) * _ncart(lb) * _ncart(lc) * _ncart(ld)
    assert np.all(np.isfinite(out))
    # kernel LDS: records + max(nacc + nv, nh) doubles must stay below 64 KB
    assert 8 * (32 * (18 + L) + max(nacc + nv, nh)) <= 64 * 1024
    return out


@pytest.mark.parametrize("key", ["".join(map(str, c)) for c in classes(2)])
def test_program_matches_oracle(key):
    la, lb, lc, ld = (int(x) for x in key)
    rng = np.random.default_rng(int(key) + 11)
    ls = [la, lb, lc, ld]
    shells = []
    for l in ls:
        centre = rng.uniform(-1.0, 1.0, 3)
        exps = np.exp(rng.uniform(math.log(0.2), math.log(8.0), 2))
        coefs = rng.uniform(0.2, 1.0, 2)
        shells.append((centre, exps, coefs))
    got = _run_program(la, lb, lc, ld, shells)
    basis = O.cgtos_from_shells(np.array([s[0] for s in shells]), np.array(ls, dtype=np.int32),
                                np.full(4, 2, dtype=np.int32),
                                np.concatenate([s[1] for s in shells]),
                                np.concatenate([s[2] for s in shells]))
    f0 = np.concatenate([[0], np.cumsum([_ncart(l) for l in ls])])
    ref = []
    for a in range(_ncart(la)):
        for b in range(_ncart(lb)):
            for c in range(_ncart(lc)):
                for d in range(_ncart(ld)):
                    ref.append(O.contracted_eri(basis, f0[0] + a, f0[1] + b, f0[2] + c, f0[3] + d,
                                                boys=O.BOYS_EXACT))
    ref = np.array(ref)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
