"""Operation tables for the wave-cooperative ERI kernel (classes with L >= 5).

For one (la lb|lc ld) class this builds the same Head-Gordon--Pople
recurrences as the straight-line generator (gen_eri.py), but as data: the
vertical recurrence for one primitive quartet becomes a list of operations
over slots of an LDS buffer, grouped into dependency levels and, inside a
level, into groups of equal (side, axis) so that every recurrence
coefficient is wave-uniform; the horizontal recurrence on the contracted
accumulators becomes a second list of the same shape.

VRR op (four 32-bit words):
  V[dst] = c0 V[s0] + c1 V[s1] + n1 o (V[s2] - r V[s3]) + n2 oo2ze V[s4]
  with (c0, c1, o, r) = (PA_i, WP_i, 1/2zeta, rho/zeta) on the bra side and
  (QC_i, WQ_i, 1/2eta, rho/eta) on the ket side; a target node [e0|f0]^(0)
  also adds its value into accumulator `acc` (0xffff: none).
  w0 = dst | s0 << 16, w1 = s1 | s2 << 16, w2 = s3 | s4 << 16,
  w3 = acc | n1 << 16 | n2 << 20.
HRR op (two 32-bit words): H[dst] = H[s0] + X_i H[s1], X = AB (bra) or CD (ket);
  w0 = dst | s0 << 16, w1 = s1.
Group record (four words): op_begin, op_end, side | axis << 1 | level_end << 3, 0.
The same recurrences as the reference's chgp.c:44-135 and chgp.c:412-586.
"""
from gen_eri import cart, comp_norm, dec, inc

NONE = 0xFFFF


def _vrr_graph(la, lb, lc, ld):
    """Nodes (e, f, m) -> (side, axis, srcs[5], n1, n2), built with the same
    axis choices as the straight-line generator."""
    E = [t for l in range(la, la + lb + 1) for t in cart(l)]
    F = [t for l in range(lc, lc + ld + 1) for t in cart(l)]
    nodes = {}

    def node(e, f, m):
        key = (e, f, m)
        if key in nodes:
            return key
        if sum(f) == 0:
            if sum(e) == 0:
                nodes[key] = None  # G_m
                return key
            i = min((k for k in range(3) if e[k] > 0), key=lambda k: (e[k] >= 2, k))
            e1 = dec(e, i)
            s = [node(e1, f, m), node(e1, f, m + 1), None, None, None]
            n1 = e1[i]
            if n1 > 0:
                e2 = dec(e1, i)
                s[2], s[3] = node(e2, f, m), node(e2, f, m + 1)
            nodes[key] = (0, i, s, n1, 0)
        else:
            i = min((k for k in range(3) if f[k] > 0),
                    key=lambda k: (2 * (f[k] >= 2) + (e[k] > 0), k))
            f1 = dec(f, i)
            s = [node(e, f1, m), node(e, f1, m + 1), None, None, None]
            n1 = f1[i]
            if n1 > 0:
                f2 = dec(f1, i)
                s[2], s[3] = node(e, f2, m), node(e, f2, m + 1)
            n2 = e[i]
            if n2 > 0:
                s[4] = node(dec(e, i), f1, m + 1)
            nodes[key] = (1, i, s, n1, n2)
        return key

    acc_of = {}
    for ie, e in enumerate(E):
        for jf, f in enumerate(F):
            acc_of[node(e, f, 0)] = ie * len(F) + jf
    return E, F, nodes, acc_of


def _levels(nodes):
    lev = {}

    def level(k):
        if k in lev:
            return lev[k]
        d = nodes[k]
        if d is None:
            lev[k] = 0
        else:
            lev[k] = 1 + max(level(s) for s in d[2] if s is not None)
        return lev[k]

    for k in nodes:
        level(k)
    return lev


def vrr_program(la, lb, lc, ld):
    L = la + lb + lc + ld
    E, F, nodes, acc_of = _vrr_graph(la, lb, lc, ld)
    lev = _levels(nodes)
    nlev = max(lev.values())
    last_use = {}
    for k, d in nodes.items():
        if d is None:
            continue
        for s in d[2]:
            if s is not None:
                last_use[s] = max(last_use.get(s, 0), lev[k])
    # slots: 0 holds zero, 1..L+1 hold G_0..G_L (loaded per primitive quartet)
    ZERO = 0
    slot = {}
    for k, d in nodes.items():
        if d is None:
            slot[k] = 1 + k[2]
    nfixed = L + 2
    free = []
    nslots = nfixed
    live = {}  # slot -> node
    ops_by_level = [[] for _ in range(nlev + 1)]
    for k in sorted((k for k in nodes if nodes[k] is not None), key=lambda k: lev[k]):
        ops_by_level[lev[k]].append(k)
    groups, ops = [], []
    for lv in range(1, nlev + 1):
        # release slots whose last reader is an earlier level
        for s_, k in list(live.items()):
            if last_use.get(k, 0) < lv:
                free.append(s_)
                del live[s_]
        free.sort(reverse=True)
        lvl_ops = ops_by_level[lv]
        for k in lvl_ops:
            if free:
                s_ = free.pop()
            else:
                s_ = nslots
                nslots += 1
            slot[k] = s_
        # a slot written at this level must not be one read at this level:
        # released slots had their last reader at an earlier level, so this holds.
        for k in lvl_ops:
            if k in last_use:
                live[slot[k]] = k
        keyed = sorted(lvl_ops, key=lambda k: (nodes[k][0], nodes[k][1]))
        gi = 0
        while gi < len(keyed):
            side, ax = nodes[keyed[gi]][0], nodes[keyed[gi]][1]
            gb = len(ops)
            while gi < len(keyed) and nodes[keyed[gi]][:2] == (side, ax):
                k = keyed[gi]
                _, _, s, n1, n2 = nodes[k]
                idx = [slot[x] if x is not None else ZERO for x in s]
                acc = acc_of.get(k, NONE)
                ops.append((slot[k] | idx[0] << 16, idx[1] | idx[2] << 16,
                            idx[3] | idx[4] << 16, acc | n1 << 16 | n2 << 20))
                gi += 1
            groups.append([gb, len(ops), side | ax << 1, 0])
        groups[-1][2] |= 1 << 3  # level end
    assert nslots < NONE and len(E) * len(F) < NONE
    # G slots that are themselves targets (the (ss|ss) class only) never occur for L >= 1
    return dict(L=L, nacc=len(E) * len(F), nslots=nslots, ops=ops, groups=groups,
                nlevels=nlev, E=E, F=F)


def hrr_program(la, lb, lc, ld):
    E = [t for l in range(la, la + lb + 1) for t in cart(l)]
    F = [t for l in range(lc, lc + ld + 1) for t in cart(l)]
    acc_index = {(e, f): ie * len(F) + jf for ie, e in enumerate(E) for jf, f in enumerate(F)}
    nodes = {}  # key -> (side, axis, s0, s1) or ("acc", index)

    def bra(a, b, f):
        key = ("b", a, b, f)
        if key not in nodes:
            if sum(b) == 0:
                nodes[key] = ("acc", acc_index[(a, f)])
            else:
                i = next(k for k in range(3) if b[k] > 0)
                b1 = dec(b, i)
                nodes[key] = (0, i, bra(inc(a, i), b1, f), bra(a, b1, f))
        return key

    def ket(a, b, c, d):
        if sum(d) == 0:
            return bra(a, b, c)
        key = ("k", a, b, c, d)
        if key not in nodes:
            i = next(k for k in range(3) if d[k] > 0)
            d1 = dec(d, i)
            nodes[key] = (1, i, ket(a, b, inc(c, i), d1), ket(a, b, c, d1))
        return key

    outs = []
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    k = ket(a, b, c, d)
                    outs.append((k, comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)))
    lev = {}

    def level(k):
        if k not in lev:
            d = nodes[k]
            lev[k] = 0 if d[0] == "acc" else 1 + max(level(d[2]), level(d[3]))
        return lev[k]

    for k in nodes:
        level(k)
    nacc = len(E) * len(F)
    INF = 1 << 30
    last_use = {}
    for k, d in nodes.items():
        if d[0] != "acc":
            for s_ in (d[2], d[3]):
                last_use[s_] = max(last_use.get(s_, 0), lev[k])
    for k, _ in outs:
        last_use[k] = INF
    # accumulators occupy slots [0, nacc); everything else reuses released slots
    slot = {k: d[1] for k, d in nodes.items() if d[0] == "acc"}
    live = {slot[k]: k for k in slot if k in last_use}
    free = sorted(set(range(nacc)) - set(live), reverse=True)
    nslots = nacc
    ops, groups = [], []
    nlev = max(lev.values())
    for lv in range(1, nlev + 1):
        for s_, k in list(live.items()):
            if last_use[k] < lv:
                free.append(s_)
                del live[s_]
        free.sort(reverse=True)
        lvl = sorted((k for k in nodes if lev[k] == lv), key=lambda k: (nodes[k][0], nodes[k][1]))
        for k in lvl:
            if free:
                slot[k] = free.pop()
            else:
                slot[k] = nslots
                nslots += 1
            live[slot[k]] = k
        gi = 0
        while gi < len(lvl):
            side, ax = nodes[lvl[gi]][0], nodes[lvl[gi]][1]
            gb = len(ops)
            while gi < len(lvl) and nodes[lvl[gi]][:2] == (side, ax):
                k = lvl[gi]
                ops.append((slot[k] | slot[nodes[k][2]] << 16, slot[nodes[k][3]]))
                gi += 1
            groups.append([gb, len(ops), side | ax << 1, 0])
        groups[-1][2] |= 1 << 3
    assert nslots < NONE
    return dict(nslots=nslots, ops=ops, groups=groups,
                out=[(slot[k], nrm) for k, nrm in outs], nlevels=nlev)


if __name__ == "__main__":
    import sys
    for key in sys.argv[1:] or ["2111", "2120", "2121", "2210", "2211", "2220", "2221", "2222"]:
        la, lb, lc, ld = (int(x) for x in key)
        v = vrr_program(la, lb, lc, ld)
        h = hrr_program(la, lb, lc, ld)
        print(key, "vrr ops", len(v["ops"]), "slots", v["nslots"], "groups", len(v["groups"]),
              "levels", v["nlevels"], "nacc", v["nacc"], "| hrr ops", len(h["ops"]), "slots",
              h["nslots"], "groups", len(h["groups"]), "nout", len(h["out"]))
