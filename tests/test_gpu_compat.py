"""GPU parity of the cints.h compatibility exports (include/rchem_cints_compat.h),
the functions rchem's tho66.rs binds: first the reference's own tests of those
bindings (tho66.rs:267-321: four overlaps at 1e-16, one ERI at 1e-12), then
random arguments against the oracle's restatement of cints.c (ERIs and nuclear
attraction with the reference's Boys function; bars written per check)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from rchem_amd._lib import lib

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_tho66_overlap_golden():
    _say("case: tho66.rs overlap goldens through overlap()")
    L = lib()
    za, zb = 1.8, 2.8
    ra, rb = (0.0, 0.0, 0.0), (0.5, 0.8, -0.2)
    cases = [((0, 0, 0, 0, 0, 0), 0.20373275913014607),
             ((1, 0, 0, 0, 0, 0), 0.062005622343957505),
             ((1, 1, 0, 1, 1, 0), -0.00043801221837779696),
             ((2, 1, 0, 1, 1, 0), -0.0002385994651113168)]
    for c, ref in cases:
        v = L.overlap(za, c[0], c[1], c[2], *ra, zb, c[3], c[4], c[5], *rb)
        assert abs(v - ref) < 1e-16, (c, v, ref)


def test_tho66_coulomb_repulsion_golden():
    _say("case: tho66.rs ERI golden through coulomb_repulsion()")
    L = lib()
    ra, rb, rc, rd = (1.0, 0.0, 1.0), (0.0, 1.0, 2.0), (0.0, 0.0, 3.0), (0.0, 0.0, 4.0)
    v = L.coulomb_repulsion(*ra, 1.0, 0, 0, 0, 1.1, *rb, 1.0, 0, 0, 0, 1.2,
                            *rc, 1.0, 0, 0, 0, 1.3, *rd, 1.0, 0, 0, 0, 1.4)
    assert abs(v - 0.08608517834596989) < 1e-12


def test_compat_functions_vs_oracle():
    L = lib()
    rng = np.random.default_rng(17)
    worst = {"eri": 0.0, "contr": 0.0, "S": 0.0, "T": 0.0, "V": 0.0, "F": 0.0}
    _say("case: compat exports vs oracle, random arguments")
    for _ in range(25):
        cen = rng.uniform(-1.5, 1.5, (4, 3))
        al = np.exp(rng.uniform(np.log(0.2), np.log(5.0), 4))
        nrm = rng.uniform(0.5, 2.0, 4)
        pw = []
        for _k in range(4):
            l = int(rng.integers(0, 3))
            p = [0, 0, 0]
            for _j in range(l):
                p[int(rng.integers(0, 3))] += 1
            pw.append(p)
        args = []
        for k in range(4):
            args += [*cen[k], nrm[k], *pw[k], al[k]]
        v = L.coulomb_repulsion(*args)
        ref = O.coulomb_repulsion(cen, np.array(pw), al, nrm, boys=O.BOYS_COMPAT)
        worst["eri"] = max(worst["eri"], abs(v - ref))
        # contracted: two primitives per function
        ex = [np.exp(rng.uniform(np.log(0.2), np.log(5.0), 2)) for _k in range(4)]
        cf = [rng.uniform(0.2, 1.0, 2) for _k in range(4)]
        nm = [rng.uniform(0.5, 2.0, 2) for _k in range(4)]
        cargs, ref = [], 0.0
        for k in range(4):
            cargs += [2, ex[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      cf[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      nm[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), *cen[k], *pw[k]]
        for i in range(2):
            for j in range(2):
                for kk in range(2):
                    for ll in range(2):
                        idx = (i, j, kk, ll)
                        ref += np.prod([cf[k][idx[k]] for k in range(4)]) * O.coulomb_repulsion(
                            cen, np.array(pw), np.array([ex[k][idx[k]] for k in range(4)]),
                            np.array([nm[k][idx[k]] for k in range(4)]), boys=O.BOYS_COMPAT)
        v = L.contr_coulomb(*cargs)
        worst["contr"] = max(worst["contr"], abs(v - ref))
        c6 = list(pw[0]) + list(pw[1])
        s = L.overlap(al[0], *pw[0], *cen[0], al[1], *pw[1], *cen[1])
        worst["S"] = max(worst["S"], abs(s - O.overlap(al[0], al[1], cen[0], cen[1], c6)))
        t = L.kinetic(al[0], *pw[0], *cen[0], al[1], *pw[1], *cen[1])
        worst["T"] = max(worst["T"], abs(t - O.kinetic(al[0], al[1], cen[0], cen[1], c6)))
        vn = L.nuclear_attraction(*cen[0], 1.0, *pw[0], al[0], *cen[1], 1.0, *pw[1], al[1],
                                  *cen[2])
        vref = O.nuclear(al[0], al[1], cen[0], cen[1], cen[2], c6, boys=O.BOYS_COMPAT)
        worst["V"] = max(worst["V"], abs(vn - vref))
        m, x = float(rng.integers(0, 9)), float(rng.uniform(0.0, 40.0))
        worst["F"] = max(worst["F"], abs(L.Fgamma(m, x) - O.fgamma_compat(m, x)))
    _say(worst)
    assert worst["eri"] < 1e-12 and worst["contr"] < 1e-12
    assert worst["S"] < 1e-14 and worst["T"] < 1e-13 and worst["V"] < 1e-12
    assert worst["F"] < 1e-14
