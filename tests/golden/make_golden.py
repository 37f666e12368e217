#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference's own C code.

Runs in the container that has /root/reference mounted, after
``make -C oracle`` has compiled the reference's libpyquante2 (cints.c, chgp.c)
into oracle/_ref/.  Every value below is produced by a reference function:

  fgamma_grid.npz        Fgamma(m, x) (cints.c:302-373) on m = 0..12 and x in
                         {0, 1e-9, log-spaced 1e-6..60, m+1.5 +- 1e-9}.
  water_sto3g.npz        rchem's water (water_crawford.xyz, bohr) in the embedded
                         STO-3G data: the 7^4 ERI tensor from contr_coulomb
                         (cints.c:42-70) with rchem's per-primitive norms and raw
                         coefficients (basis.rs:140-149, 201), and S, T from the
                         reference's overlap/kinetic (cints.c:152-209).
  contracted_samples.npz random shells with l <= 2 (3 primitives each) and, for
                         24 shell quartets covering every (la lb|lc ld) class,
                         every Cartesian component from contr_coulomb (THO) and
                         contr_hrr (HGP, chgp.c:44-111).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z  # noqa: E402

R = oracle.ref_lib()
if R is None:
    sys.exit("oracle/_ref/libpyquante2_ref.so is missing: run make -C oracle with the reference mounted")

d, i = ctypes.c_double, ctypes.c_int
dp = ctypes.POINTER(ctypes.c_double)
R.contr_coulomb.restype = d
R.contr_coulomb.argtypes = ([i, dp, dp, dp, d, d, d, i, i, i] * 4)
R.contr_hrr.restype = d
R.contr_hrr.argtypes = ([i, d, d, d, dp, i, i, i, dp, dp] * 4)


def arr(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


def norms(exps, pw):
    return np.array([oracle.lib().orc_pgto_norm(float(a), int(pw[0]), int(pw[1]), int(pw[2]))
                     for a in exps])


def contr(funcs, hgp=False):
    """funcs: 4 tuples (centre, powers, exps, coefs)."""
    keep, args = [], []
    for (c, pw, ex, cf) in funcs:
        e, ep = arr(ex)
        k, kp = arr(cf)
        n, np_ = arr(norms(ex, pw))
        keep += [e, k, n]
        if hgp:
            args += [len(ex), c[0], c[1], c[2], np_, int(pw[0]), int(pw[1]), int(pw[2]), ep, kp]
        else:
            args += [len(ex), ep, kp, np_, c[0], c[1], c[2], int(pw[0]), int(pw[1]), int(pw[2])]
    return (R.contr_hrr if hgp else R.contr_coulomb)(*args)


def fgamma_grid():
    ms, xs = [], []
    base = [0.0, 1e-9] + list(np.logspace(-6, np.log10(60.0), 160))
    for m in range(13):
        for x in base + [m + 1.5 - 1e-9, m + 1.5 + 1e-9]:
            ms.append(m)
            xs.append(x)
    vals = [R.Fgamma(float(m), float(x)) for m, x in zip(ms, xs)]
    np.savez(os.path.join(HERE, "fgamma_grid.npz"), m=np.array(ms), x=np.array(xs),
             F=np.array(vals))


def water():
    bse = oracle.load_bse_json(os.path.join(ROOT, "rchem_amd", "data", "sto-3g.json"))
    B = oracle.cgtos_from_bse(bse, WATER_Z, WATER_CRAWFORD)
    n = B.n
    funcs = []
    for f in range(n):
        sl = slice(B.prim_off[f], B.prim_off[f + 1])
        funcs.append((B.centre[f], B.powers[f], B.exps[sl], B.coefs[sl]))
    I = np.zeros((n, n, n, n))
    for a in range(n):
        for b in range(n):
            for c in range(n):
                for e in range(n):
                    I[a, b, c, e] = contr([funcs[a], funcs[b], funcs[c], funcs[e]])
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            ca, pa, ea, fa = funcs[a]
            cb, pb, eb, fb = funcs[b]
            na, nb = norms(ea, pa), norms(eb, pb)
            for p in range(len(ea)):
                for q in range(len(eb)):
                    w = fa[p] * fb[q] * na[p] * nb[q]
                    S[a, b] += w * R.overlap(ea[p], int(pa[0]), int(pa[1]), int(pa[2]), ca[0], ca[1],
                                             ca[2], eb[q], int(pb[0]), int(pb[1]), int(pb[2]),
                                             cb[0], cb[1], cb[2])
                    T[a, b] += w * R.kinetic(ea[p], int(pa[0]), int(pa[1]), int(pa[2]), ca[0], ca[1],
                                             ca[2], eb[q], int(pb[0]), int(pb[1]), int(pb[2]),
                                             cb[0], cb[1], cb[2])
    np.savez(os.path.join(HERE, "water_sto3g.npz"), eri=I, S=S, T=T)


def samples():
    rng = np.random.default_rng(20261202)
    ns = 8
    centres = rng.uniform(-1.5, 1.5, (ns, 3))
    ls = np.array([0, 1, 2, 0, 1, 2, 1, 2], dtype=np.int32)
    nprim = np.full(ns, 3, dtype=np.int32)
    exps = np.exp(rng.uniform(np.log(0.2), np.log(20.0), ns * 3))
    coefs = rng.uniform(0.2, 1.0, ns * 3)
    by_l = {l: [s for s in range(ns) if ls[s] == l] for l in range(3)}
    quartets = []
    pairs = [(a, b) for a in range(3) for b in range(a + 1)]
    for bi, (la, lb) in enumerate(pairs):
        for (lc, ld) in pairs[:bi + 1]:
            q = [int(rng.choice(by_l[x])) for x in (la, lb, lc, ld)]
            quartets.append(q)
    # three more with shells shared, including a fully one-centre quartet
    quartets += [[2, 2, 2, 2], [1, 1, 4, 4], [5, 2, 5, 2]]
    vals, vals_hgp, index = [], [], []
    for qi, q in enumerate(quartets):
        comps = [oracle.get_ijk_list(int(ls[s])) for s in q]
        for ca in comps[0]:
            for cb in comps[1]:
                for cc in comps[2]:
                    for cd in comps[3]:
                        funcs = []
                        for s, pw in zip(q, (ca, cb, cc, cd)):
                            sl = slice(3 * s, 3 * s + 3)
                            funcs.append((centres[s], pw, exps[sl], coefs[sl]))
                        vals.append(contr(funcs))
                        vals_hgp.append(contr(funcs, hgp=True))
                        index.append(qi)
    np.savez(os.path.join(HERE, "contracted_samples.npz"), centres=centres, ls=ls, nprim=nprim,
             exps=exps, coefs=coefs, quartets=np.array(quartets, dtype=np.int32),
             values=np.array(vals), values_hgp=np.array(vals_hgp), quartet_of=np.array(index))


if __name__ == "__main__":
    fgamma_grid()
    water()
    samples()
    print("golden vectors written to", HERE)
