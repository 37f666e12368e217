"""Operation tables for the wave-cooperative ERI kernel.

For one (la lb|lc ld) class this builds the same Head-Gordon--Pople
recurrences as the straight-line generator (gen_eri.py), but as data that the
kernel (eri_generic.hip) interprets:

* VRR: for a batch of P primitive quartets, the vertical recurrence as a list
  of operations over slots of an LDS buffer V[slot][q], grouped into
  dependency levels and, inside a level, into groups of one (side, axis), so
  that a lane -- which always serves the same primitive quartet q of the batch
  -- loads the group's recurrence coefficients once:
    V[dst] = c0 V[s0] + c1 V[s1] + n1 o (V[s2] - r V[s3]) + n2 oo2ze V[s4]
  with (c0, c1, o, r) = (PA_i, WP_i, 1/2zeta, rho/zeta) on the bra side and
  (QC_i, WQ_i, 1/2eta, rho/eta) on the ket side.
    w0 = dst | s0 << 16, w1 = s1 | s2 << 16, w2 = s3 | s4 << 16, w3 = n1 | n2 << 4.
  Slot 0 holds zero, slots 1..L+1 hold G_0..G_L; target slots [e0|f0]^(0) stay
  live until the end of the batch, other slots are reused once their last
  reader's level has passed.
* Accumulation: (slot | set << 16, acc) -- acc[acc] += sum_q w_set(q) V[slot][q],
  with w_0 = 1 and, for the gradient, w_1 = 2 alpha_a, w_2 = 2 alpha_b,
  w_3 = 2 alpha_c of each primitive quartet.
* HRR on the contracted accumulators (the first slots of a buffer H), the same
  level/group layout: H[dst] = H[s0] + X_i H[s1], X = AB (bra) or CD (ket);
  w0 = dst | s0 << 16, w1 = s1.
* Outputs: value_k = norm_k (H[plus_k] - n_k H[minus_k]) -- the integrals of the
  class (n_k = 0), or for the gradient the derivative integrals
  d/dA_i (ab|cd) = (a+1_i b|cd)_{2 alpha_a} - a_i (a-1_i b|cd), and likewise for
  B and C, in the order (centre A, B, C) x (x, y, z) x components.
Group record: (op_begin, op_end, side | axis << 1 | level_end << 3, 0).
The recurrences are those of the reference's chgp.c:44-135 and chgp.c:412-586.
"""
from gen_eri import cart, comp_norm, dec, inc

NONE = 0xFFFF
LDS_BUDGET = 48 * 1024   # bytes of LDS per one-wave workgroup
PMAX = 32                # primitive quartets per batch at most
NREC_FIXED = 20          # record fields before G_0 (see eri_generic.hip)


def _build_vrr_nodes(targets):
    """Nodes (e, f, m) -> (side, axis, srcs[5], n1, n2) or None for G_m, with
    the straight-line generator's axis choices."""
    nodes = {}

    def node(e, f, m):
        key = (e, f, m)
        if key in nodes:
            return key
        if sum(f) == 0:
            if sum(e) == 0:
                nodes[key] = None
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

    for (e, f) in targets:
        node(e, f, 0)
    return nodes


def _levels(nodes, base):
    lev = {}

    def level(k):
        if k not in lev:
            d = base(k)
            lev[k] = 0 if d is None else 1 + max(level(s) for s in d if s is not None)
        return lev[k]

    for k in nodes:
        level(k)
    return lev


def vrr_program(L, targets):
    """VRR ops for one primitive quartet building [e0|f0]^(0) for every (e, f)
    in `targets` (total angular momentum up to L)."""
    nodes = _build_vrr_nodes(targets)
    lev = _levels(nodes, lambda k: None if nodes[k] is None else nodes[k][2])
    nlev = max(lev.values())
    INF = 1 << 30
    last_use = {}
    for k, d in nodes.items():
        if d is not None:
            for s in d[2]:
                if s is not None:
                    last_use[s] = max(last_use.get(s, 0), lev[k])
    for (e, f) in targets:
        last_use[(e, f, 0)] = INF
    slot = {k: 1 + k[2] for k, d in nodes.items() if d is None}
    assert max(slot.values()) <= L + 1
    nslots = L + 2
    free, live = [], {}
    ops, groups = [], []
    by_level = {}
    for k in nodes:
        if nodes[k] is not None:
            by_level.setdefault(lev[k], []).append(k)
    for lv in range(1, nlev + 1):
        for s_, k in list(live.items()):
            if last_use.get(k, 0) < lv:
                free.append(s_)
                del live[s_]
        free.sort(reverse=True)
        lvl = sorted(by_level.get(lv, []), key=lambda k: (nodes[k][0], nodes[k][1], k))
        for k in lvl:
            if free:
                slot[k] = free.pop()
            else:
                slot[k] = nslots
                nslots += 1
            if k in last_use:
                live[slot[k]] = k
        gi = 0
        while gi < len(lvl):
            side, ax = nodes[lvl[gi]][0], nodes[lvl[gi]][1]
            gb = len(ops)
            while gi < len(lvl) and nodes[lvl[gi]][:2] == (side, ax):
                k = lvl[gi]
                _, _, s, n1, n2 = nodes[k]
                idx = [slot[x] if x is not None else 0 for x in s]
                ops.append((slot[k] | idx[0] << 16, idx[1] | idx[2] << 16,
                            idx[3] | idx[4] << 16, n1 | n2 << 4))
                gi += 1
            groups.append([gb, len(ops), side | ax << 1, 0])
        if groups:
            groups[-1][2] |= 1 << 3
    assert nslots < NONE
    return dict(ops=ops, groups=groups, nslots=nslots,
                target_slot={(e, f): slot[(e, f, 0)] for (e, f) in targets})


def _hrr(accsets, outs_spec):
    """HRR nodes over accumulator sets.  accsets[s] = (E, F, base) with the
    contracted [e0|f0] of set s at accumulator base + ie * len(F) + jf.
    outs_spec: list of (plus, minus, n, norm), plus/minus = (set, a, b, c, d)
    or None.  Returns ops, groups, nslots, outs (slot_plus, slot_minus, n, norm)."""
    acc_index = {}
    for s, (E, F, base) in enumerate(accsets):
        for ie, e in enumerate(E):
            for jf, f in enumerate(F):
                acc_index[(s, e, f)] = base + ie * len(F) + jf
    nacc = sum(len(E) * len(F) for E, F, _ in accsets)
    nodes = {}

    def bra(s, a, b, f):
        key = ("b", s, a, b, f)
        if key not in nodes:
            if sum(b) == 0:
                nodes[key] = ("acc", acc_index[(s, a, f)])
            else:
                i = next(k for k in range(3) if b[k] > 0)
                b1 = dec(b, i)
                nodes[key] = (0, i, bra(s, inc(a, i), b1, f), bra(s, a, b1, f))
        return key

    def ket(s, a, b, c, d):
        if sum(d) == 0:
            return bra(s, a, b, c)
        key = ("k", s, a, b, c, d)
        if key not in nodes:
            i = next(k for k in range(3) if d[k] > 0)
            d1 = dec(d, i)
            nodes[key] = (1, i, ket(s, a, b, inc(c, i), d1), ket(s, a, b, c, d1))
        return key

    out_nodes = []
    for plus, minus, n, nrm in outs_spec:
        kp = ket(*plus)
        km = ket(*minus) if minus is not None else None
        out_nodes.append((kp, km, n, nrm))
    lev = _levels(nodes, lambda k: None if nodes[k][0] == "acc" else (nodes[k][2], nodes[k][3]))
    INF = 1 << 30
    last_use = {}
    for k, d in nodes.items():
        if d[0] != "acc":
            for s_ in (d[2], d[3]):
                last_use[s_] = max(last_use.get(s_, 0), lev[k])
    for kp, km, _, _ in out_nodes:
        last_use[kp] = INF
        if km is not None:
            last_use[km] = INF
    slot = {k: d[1] for k, d in nodes.items() if d[0] == "acc"}
    live = {slot[k]: k for k in slot if k in last_use}
    free = sorted(set(range(nacc)) - set(live), reverse=True)
    nslots = nacc
    ops, groups = [], []
    nlev = max(lev.values()) if lev else 0
    for lv in range(1, nlev + 1):
        for s_, k in list(live.items()):
            if last_use[k] < lv:
                free.append(s_)
                del live[s_]
        free.sort(reverse=True)
        lvl = sorted((k for k in nodes if lev[k] == lv), key=lambda k: (nodes[k][0], nodes[k][1], str(k)))
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
    outs = [(slot[kp], slot[km] if km is not None else slot[kp], n, nrm)
            for kp, km, n, nrm in out_nodes]
    return dict(ops=ops, groups=groups, nslots=nslots, outs=outs, nacc=nacc)


def _range(l0, l1):
    return [t for l in range(max(l0, 0), l1 + 1) for t in cart(l)]


def lds_bytes(L, nacc, nv, nh, p):
    return 8 * ((NREC_FIXED + L + 1) * p + max(nacc + nv * p, nh))


def _choose_p(L, nacc, nv, nh):
    """Largest batch within the LDS budget; 0 when not even one primitive
    quartet fits a workgroup's 64 KB."""
    p = PMAX
    while p > 1 and lds_bytes(L, nacc, nv, nh, p) > LDS_BUDGET:
        p //= 2
    return p if lds_bytes(L, nacc, nv, nh, p) <= 64 * 1024 else 0


def _assemble(L, accsets, vrr_targets, outs_spec, weights=None):
    """weights[s]: the accumulation weight of set s (0: 1, 1: 2 alpha_a,
    2: 2 alpha_b, 3: 2 alpha_c); by default the set's own index."""
    v = vrr_program(L, vrr_targets)
    acc_list = []
    for s, (E, F, base) in enumerate(accsets):
        w = s if weights is None else weights[s]
        for ie, e in enumerate(E):
            for jf, f in enumerate(F):
                acc_list.append((v["target_slot"][(e, f)] | w << 16, base + ie * len(F) + jf))
    h = _hrr(accsets, outs_spec)
    nacc = h["nacc"]
    P = _choose_p(L, nacc, v["nslots"], h["nslots"])
    return dict(L=L, nacc=nacc, nv=v["nslots"], nh=h["nslots"], P=P, vops=v["ops"],
                vgroups=v["groups"], accs=acc_list, hops=h["ops"], hgroups=h["groups"],
                outs=h["outs"])


def eri_program(la, lb, lc, ld):
    """The integrals (ab|cd) of the class, canonical component order."""
    E, F = _range(la, la + lb), _range(lc, lc + ld)
    outs = []
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    nrm = comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)
                    outs.append(((0, a, b, c, d), None, 0, nrm))
    return _assemble(la + lb + lc + ld, [(E, F, 0)], [(e, f) for e in E for f in F], outs)


def grad_program(la, lb, lc, ld, centre=None):
    """Derivatives of (ab|cd) with respect to the centres A, B and C (D follows
    from translational invariance), outputs ordered (A, B, C) x (x, y, z) x
    canonical components.  Sets: 0 unweighted, 1 by 2 alpha_a, 2 by 2 alpha_b,
    3 by 2 alpha_c.  With centre = 0, 1 or 2 only that centre's derivatives
    ((x, y, z) x components, sets 0 and 1 + centre): for the classes whose
    full program does not fit a workgroup's LDS, run once per centre."""
    plus = [
        (_range(la + 1, la + lb + 1), _range(lc, lc + ld)),   # A plus
        (_range(la, la + lb + 1), _range(lc, lc + ld)),       # B plus
        (_range(la, la + lb), _range(lc + 1, lc + ld + 1)),   # C plus
    ]
    minus = (_range(la - 1, la + lb), _range(lc - 1, lc + ld))
    centres = range(3) if centre is None else [centre]
    sets = [minus] + ([plus[c] for c in centres])
    weights = [0] + [1 + c for c in centres]
    accsets, base = [], 0
    for E, F in sets:
        accsets.append((E, F, base))
        base += len(E) * len(F)
    targets = sorted({(e, f) for E, F in sets for e in E for f in F})
    outs = []
    for ci, centre in enumerate(centres):
        for i in range(3):
            for a in cart(la):
                for b in cart(lb):
                    for c in cart(lc):
                        for d in cart(ld):
                            nrm = comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)
                            t = [a, b, c, d]
                            plus = list(t)
                            plus[centre] = inc(t[centre], i)
                            n = t[centre][i]
                            minus = None
                            if n > 0:
                                minus = list(t)
                                minus[centre] = dec(t[centre], i)
                                minus = (0, *minus)
                            outs.append(((1 + ci, *plus), minus, n, nrm))
    return _assemble(la + lb + lc + ld + 1, accsets, targets, outs, weights)


if __name__ == "__main__":
    import sys
    keys = sys.argv[1:] or ["0000", "1111", "2111", "2121", "2221", "2222"]
    for key in keys:
        la, lb, lc, ld = (int(x) for x in key)
        for name, fn in (("eri", eri_program), ("grad", grad_program)):
            p = fn(la, lb, lc, ld)
            lds = lds_bytes(p["L"], p["nacc"], p["nv"], p["nh"], p["P"])
            print(f"{key} {name}: L {p['L']} vops {len(p['vops'])} vgroups {len(p['vgroups'])} "
                  f"nv {p['nv']} nacc {p['nacc']} accs {len(p['accs'])} hops {len(p['hops'])} "
                  f"nh {p['nh']} outs {len(p['outs'])} P {p['P']} lds {lds}")
