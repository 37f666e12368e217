#!/usr/bin/env python3
"""Generates the per-class Head-Gordon--Pople recurrences as straight-line HIP code.

For every shell-quartet class (la lb|lc ld) with la >= lb, lc >= ld and
pair-class(la,lb) >= pair-class(lc,ld), l <= LMAX, it emits

  * ``vrr``: the Obara--Saika/HGP vertical recurrence for one primitive
    quartet, building [e0|f0]^(0) for l_e in [la, la+lb], l_f in [lc, lc+ld]
    from G_m = K_ab K_cd / sqrt(zeta+eta) * F_m(T) and adding them into the
    contracted accumulators ``acc`` (the same recurrences as the reference's
    chgp.c:412-586, with every intermediate computed once);
  * ``hrr``: the horizontal recurrence on the contracted accumulators,
    (a, b+1_i| = (a+1_i, b| + AB_i (a b|, then the same on the ket with CD
    (chgp.c:44-111), producing (ab|cd) in rchem's Cartesian order
    (shell.rs:1-12) times the per-component normalisation of basis.rs:140-149.

Intermediates are named locals, so after inlining they live in registers.
The generator also counts floating-point operations per class (add/mul = 1,
FMA = 2) for the roofline model, written to ``flops_per_class.json``.

It writes, into OUT_DIR: eri_classes.inc (the recurrences), one translation
unit eri_c_<la><lb><lc><ld>.hip per class instantiating the kernel variants
that class needs, eri_dispatch.inc (class -> launcher), eri_flops.inc and
flops_per_class.json (operation counts).

Usage: gen_eri.py OUT_DIR [LMAX [KEY ...]]   (KEYs: classes to dispatch, default all)
"""
import json
import math
import os
import re
import sys

AX = "xyz"
N_LDS_VARIANTS = 10  # entries of kLdsVariants (types.hpp)


def cart(l):
    """Cartesian components in rchem's order (shell.rs:1-12)."""
    out = []
    for a in range(1, l + 2):
        for b in range(1, a + 1):
            out.append((l + 1 - a, a - b, b - 1))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


def dec(t, i, k=1):
    t = list(t)
    t[i] -= k
    return tuple(t)


def inc(t, i):
    t = list(t)
    t[i] += 1
    return tuple(t)


def tname(t):
    return "".join(str(x) for x in t)


def comp_norm(t):
    """1/sqrt((2l-1)!!(2m-1)!!(2n-1)!!): the component-dependent factor of the
    primitive norm of basis.rs:140-149 (the exponent-dependent factor is folded
    into the contraction coefficients on the host)."""
    def f2(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return 1.0 / math.sqrt(f2(2 * t[0] - 1) * f2(2 * t[1] - 1) * f2(2 * t[2] - 1))


class Emitter:
    def __init__(self):
        self.lines = []
        self.flops = 0
        self.consts = {}

    def const(self, base, k):
        if k == 1:
            return base
        name = f"{base}_{k}"
        if name not in self.consts:
            self.consts[name] = f"const R {name} = R({k}) * {base};"
            self.flops += 1
        return name


class VRR:
    """Memoised VRR for one class; nodes are (e, f, m)."""

    def __init__(self, em):
        self.em = em
        self.memo = {}
        self.terms = {}  # node name -> its summands (the line is their sum)

    def node(self, e, f, m):
        key = (e, f, m)
        if key in self.memo:
            return self.memo[key]
        em = self.em
        if sum(f) == 0:
            if sum(e) == 0:
                name = f"G[{m}]"
                self.memo[key] = name
                return name
            # bra step on A: pick the axis that avoids the (e_i - 1) term if possible
            i = min((k for k in range(3) if e[k] > 0), key=lambda k: (e[k] >= 2, k))
            e1 = dec(e, i)
            t1, t2 = self.node(e1, f, m), self.node(e1, f, m + 1)
            terms = [f"PA{AX[i]} * {t1}", f"WP{AX[i]} * {t2}"]
            fl = 3
            if e1[i] > 0:
                e2 = dec(e1, i)
                t3, t4 = self.node(e2, f, m), self.node(e2, f, m + 1)
                c = em.const("oo2z", e1[i])
                terms.append(f"{c} * ({t3} - roz * {t4})")
                fl += 4
        else:
            # ket step on C: avoid the (f_i - 1) and electron-transfer terms when possible
            i = min((k for k in range(3) if f[k] > 0),
                    key=lambda k: (2 * (f[k] >= 2) + (e[k] > 0), k))
            f1 = dec(f, i)
            t1, t2 = self.node(e, f1, m), self.node(e, f1, m + 1)
            terms = [f"QC{AX[i]} * {t1}", f"WQ{AX[i]} * {t2}"]
            fl = 3
            if f1[i] > 0:
                f2 = dec(f1, i)
                t3, t4 = self.node(e, f2, m), self.node(e, f2, m + 1)
                c = em.const("oo2e", f1[i])
                terms.append(f"{c} * ({t3} - roe * {t4})")
                fl += 4
            if e[i] > 0:
                e1 = dec(e, i)
                t5 = self.node(e1, f1, m + 1)
                c = em.const("oo2ze", e[i])
                terms.append(f"{c} * {t5}")
                fl += 2
        name = f"v{tname(e)}_{tname(f)}_{m}"
        self.terms[name] = terms
        em.lines.append(f"const R {name} = {' + '.join(terms)};")
        em.flops += fl
        self.memo[key] = name
        return name


def fuse_accumulation(lines, terms, targets, L):
    """The VRR lines with each target added into its accumulator.  A target no
    other line reads is not formed: each of its summands is added into the
    accumulator on its own (acc += x * y contracts to one FMA, so the sum's
    final add is saved).  The others are formed and then added.  In the L = 1
    class the bra step's WP_i G_1 = -PQ_i (roz G_1) takes one product roz G_1
    for the three components in place of the three WP_i = -roz PQ_i."""
    used = set()
    for ln in lines:
        used.update(re.findall(r"\bv\w+", ln.split("=", 1)[1]))
    out = []
    if L == 1:
        out.append("const R H1 = roz * G[1];")
    for ln in lines:
        name = ln.split()[2]
        if name in targets and name not in used:
            acc = f"acc[{targets[name]}]"
            for t in terms[name]:
                m = re.fullmatch(r"WP([xyz]) \* G\[1\]", t)
                if L == 1 and m:
                    out.append(f"{acc} -= PQ{m.group(1)} * H1;")
                else:
                    out.append(f"{acc} += {t};")
        else:
            out.append(ln)
    out += [f"acc[{i}] += {name};" for name, i in targets.items()
            if name not in terms or name in used]
    return out


def gen_class(la, lb, lc, ld):
    E = [t for l in range(la, la + lb + 1) for t in cart(l)]
    F = [t for l in range(lc, lc + ld + 1) for t in cart(l)]
    acc_index = {}
    for ie, e in enumerate(E):
        for jf, f in enumerate(F):
            acc_index[(e, f)] = ie * len(F) + jf
    nacc = len(E) * len(F)
    L = la + lb + lc + ld

    # ---- VRR
    em = Emitter()
    v = VRR(em)
    targets = {}
    for e in E:
        for f in F:
            targets[v.node(e, f, 0)] = acc_index[(e, f)]
            em.flops += 1
    vrr_flops = em.flops
    vrr_body = list(em.consts.values()) + fuse_accumulation(em.lines, v.terms, targets, L)

    # ---- HRR (bra, then ket), memoised
    hl = []
    hmemo = {}
    hflops = [0]

    def bra(a, b, f):
        key = ("b", a, b, f)
        if key in hmemo:
            return hmemo[key]
        if sum(b) == 0:
            name = f"acc[{acc_index[(a, f)]}]"
        else:
            i = next(k for k in range(3) if b[k] > 0)
            b1 = dec(b, i)
            t1 = bra(inc(a, i), b1, f)
            t2 = bra(a, b1, f)
            name = f"hb{tname(a)}_{tname(b)}_{tname(f)}"
            hl.append(f"const R {name} = {t1} + AB{AX[i]} * {t2};")
            hflops[0] += 2
        hmemo[key] = name
        return name

    def ket(a, b, c, d):
        key = ("k", a, b, c, d)
        if key in hmemo:
            return hmemo[key]
        if sum(d) == 0:
            name = bra(a, b, c)
        else:
            i = next(k for k in range(3) if d[k] > 0)
            d1 = dec(d, i)
            t1 = ket(a, b, inc(c, i), d1)
            t2 = ket(a, b, c, d1)
            name = f"hk{tname(a)}_{tname(b)}_{tname(c)}_{tname(d)}"
            hl.append(f"const R {name} = {t1} + CD{AX[i]} * {t2};")
            hflops[0] += 2
        hmemo[key] = name
        return name

    out_lines = []
    idx = 0
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    name = ket(a, b, c, d)
                    cn = comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)
                    if abs(cn - 1.0) > 0:
                        out_lines.append(f"out[{idx}] = R({cn!r}) * {name};")
                        hflops[0] += 1
                    else:
                        out_lines.append(f"out[{idx}] = {name};")
                    idx += 1
    nout = idx
    return dict(L=L, nacc=nacc, nout=nout, vrr=vrr_body, hrr=hl + out_lines,
                vrr_flops=vrr_flops, hrr_flops=hflops[0])


def _stream_body(targets, acc_index):
    """Straight-line VRR for the given targets, each handed to acc.put<k>() as
    soon as it is formed; returns (lines, peak number of live VRR values)."""
    em = Emitter()
    v = VRR(em)
    body = []
    for (e, f) in targets:
        n0 = len(em.lines)
        name = v.node(e, f, 0)
        body.extend(em.lines[n0:])
        body.append(f"acc.template put<{acc_index[(e, f)]}>({name});")
    # peak live set of the emitted order (a target dies when handed over)
    import re
    defs = [re.match(r"const R (\w+) = (.*);", ln) for ln in body if ln.startswith("const R")]
    last = {}
    pos = 0
    for ln in body:
        if ln.startswith("acc."):
            last[re.search(r"\((\w+)\);", ln).group(1)] = pos
        else:
            for d in re.findall(r"\b(v\w+)", ln.split("=", 1)[1]):
                last[d] = pos
            pos += 1
    live, peak, pos = set(), 0, 0
    for ln in body:
        if ln.startswith("acc."):
            continue
        name = ln.split()[2]
        live.add(name)
        peak = max(peak, len(live))
        live = {x for x in live if last.get(x, -1) > pos}
        pos += 1
    return list(em.consts.values()) + body, peak


def stream_vrr(la, lb, lc, ld):
    """The VRR of the streaming kernel (eri_stream.hpp): one lane holds one
    primitive quartet and hands each target [e0|f0]^(0) to acc.put<k>() as
    soon as it is formed, k being its accumulator index (the wave-cooperative
    program's layout, ie * len(F) + jf), in increasing k: the lane keeps no
    accumulators, only the VRR values still needed.  Where that is more than
    STREAM_LIVE values, the targets are split into parts (boundaries on the
    16-target groups the kernel sums over the wave), each recomputing the
    intermediates it needs: more operations, no spills.  Returns the parts'
    bodies."""
    E = [t for l in range(la, la + lb + 1) for t in cart(l)]
    F = [t for l in range(lc, lc + ld + 1) for t in cart(l)]
    acc_index = {(e, f): ie * len(F) + jf for ie, e in enumerate(E) for jf, f in enumerate(F)}
    targets = [(e, f) for e in E for f in F]
    n = len(targets)
    for parts in range(1, 9):
        cuts = [0] + [16 * round(n * p / parts / 16) for p in range(1, parts)] + [n]
        bodies = [_stream_body(targets[cuts[p]:cuts[p + 1]], acc_index) for p in range(parts)]
        if max(pk for _, pk in bodies) <= STREAM_LIVE or parts == 8:
            return [b for b, _ in bodies]


# VRR values a streaming lane may keep live before the targets are split.  On
# gfx950, splitting (dd|dp) into 2 parts and (dd|dd) into 3 (live sets 102
# and 150 instead of 162 and 306 in the emitted order) left the compiler's
# register allocation and the kernel times unchanged, so every class streams
# in one part.
STREAM_LIVE = 1000
STREAM_L = 5  # classes from this total angular momentum get a streaming kernel
# L = 5 classes on the per-lane kernel (the streaming kernel is slower for
# them) that also get LDS digestion: 300-shell s/p/d microbenchmark, ms per
# FP64 build, global atomics -> LDS digestion: (dp|ds) 52.5 -> 28.5, (dd|ps)
# 17.3 -> 12.4.  (dp|pp) and (dd|ds) measured unchanged (127, 54 ms: their
# spills, not the atomics, bound them) and keep global atomics.
LDS_PER_LANE_BEYOND_4 = {"2120", "2210"}


def pair_class(l1, l2):
    return l1 * (l1 + 1) // 2 + l2


def classes(lmax):
    pairs = [(l1, l2) for l1 in range(lmax + 1) for l2 in range(l1 + 1)]
    pairs.sort(key=lambda p: pair_class(*p))
    out = []
    for bi, (la, lb) in enumerate(pairs):
        for (lc, ld) in pairs[:bi + 1]:
            out.append((la, lb, lc, ld))
    return out


def tu_source(la, lb, lc, ld):
    """Kernel variants of one class.  s/p classes carry everything (both Boys
    modes, FP32, LDS digestion); d classes the production and parity set, with
    LDS digestion up to L = 4 and for the L = 5, 6 classes the per-lane kernel
    serves (LDS_PER_LANE_BEYOND_4)."""
    key = f"{la}{lb}{lc}{ld}"
    sp = max(la, lb, lc, ld) <= 1
    A = f"{la}, {lb}, {lc}, {ld}"
    L = ["// Generated by gen_eri.py -- do not edit.", '#include "../eri_common.hpp"', "",
         "namespace rchem {", "",
         f"EriLaunchFn eri_launcher_c{key}(int mode, int bm, int fp32) {{",
         "  if (fp32)",
         f"    return (mode == kModeJK && bm == kBoysExact) ? eri_launch<{A}, kModeJK, kBoysExact, float>"
         " : nullptr;",
         "  if (bm == kBoysExact) {",
         f"    if (mode == kModeJK) return eri_launch<{A}, kModeJK, kBoysExact, double>;",
         f"    if (mode == kModeDiag) return eri_launch<{A}, kModeDiag, kBoysExact, double>;",
         f"    if (mode == kModeTensor) return eri_launch<{A}, kModeTensor, kBoysExact, double>;",
         f"    if (mode == kModeQuartet) return eri_launch<{A}, kModeQuartet, kBoysExact, double>;",
         "    return nullptr;",
         "  }"]
    if sp:
        L += [f"  if (mode == kModeJK) return eri_launch<{A}, kModeJK, kBoysCompat, double>;",
              f"  if (mode == kModeDiag) return eri_launch<{A}, kModeDiag, kBoysCompat, double>;"]
    L += [f"  if (mode == kModeTensor) return eri_launch<{A}, kModeTensor, kBoysCompat, double>;",
          f"  if (mode == kModeQuartet) return eri_launch<{A}, kModeQuartet, kBoysCompat, double>;",
          "  return nullptr;", "}", "",
          "// LDS-digestion kernel of launch shape kLdsVariants[v]; nullptr when the class",
          "// is not compiled in that shape (the caller falls back to shape 0).",
          f"EriLaunchFn eri_jk_lds_launcher_c{key}(int bm, int fp32, int v) {{"]
    if sp:
        # every shape for the s/p classes (the tuning table chooses), and the
        # compat-Boys parity kernel in shape 0
        L += ["  if (v == 0 && !fp32 && bm == kBoysCompat)",
              f"    return eri_jk_lds_launch<{A}, kBoysCompat, double>;",
              "  switch (v) {"]
        L += [f"    case {v}: return eri_jk_lds_variant<{A}, {v}>(bm, fp32);"
              for v in range(N_LDS_VARIANTS)]
        L += ["    default: return nullptr;", "  }"]
    elif la + lb + lc + ld <= 4 or key in LDS_PER_LANE_BEYOND_4:
        # d classes on the per-lane kernel: LDS digestion for the production
        # Boys, default shape only (compile time)
        L += [f"  return v == 0 ? eri_jk_lds_variant<{A}, 0>(bm, fp32) : nullptr;"]
    else:
        L += ["  (void)bm;", "  (void)fp32;", "  (void)v;", "  return nullptr;"]
    L += ["}", "", "}  // namespace rchem", ""]
    return "\n".join(L)


def stream_tu_source(la, lb, lc, ld):
    """Streaming-kernel launchers of one class (L >= STREAM_L)."""
    A = f"{la}, {lb}, {lc}, {ld}"
    key = f"{la}{lb}{lc}{ld}"
    return "\n".join([
        "// Generated by gen_eri.py -- do not edit.", '#include "../eri_stream.hpp"', "",
        "namespace rchem {", "",
        f"EriStreamFn eri_stream_launcher_c{key}(int mode) {{",
        f"  if (mode == kModeJK) return eri_stream_launch<{A}, kModeJK>;",
        f"  if (mode == kModeDiag) return eri_stream_launch<{A}, kModeDiag>;",
        f"  if (mode == kModeTensor) return eri_stream_launch<{A}, kModeTensor>;",
        f"  if (mode == kModeQuartet) return eri_stream_launch<{A}, kModeQuartet>;",
        "  return nullptr;", "}", "", "}  // namespace rchem", ""])


def write_if_changed(path, text):
    """Write text to path unless it already holds exactly that (an unchanged
    output keeps its timestamp, so make rebuilds only what changed)."""
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as fh:
            fh.write(text)


def write_programs(path, lmax):
    """Operation tables of the wave-cooperative kernel: the integrals of every
    class and the gradient's derivative integrals (grad = 1), split into one
    program per centre (grad = 2, 3, 4) where the full one exceeds its LDS."""
    from gen_eri_prog import eri_program, grad_program

    def arr(ctype, name, vals, per_line=8):
        vals = vals or ["0"]
        body = []
        for i in range(0, len(vals), per_line):
            body.append("    " + ", ".join(vals[i:i + per_line]) + ",")
        return [f"const {ctype} {name}[] = {{"] + body + ["};"]

    def words(rows):
        return [f"{x}u" for r in rows for x in r]

    out = ["// Generated by gen_eri.py -- do not edit.", ""]
    entries = []
    for (la, lb, lc, ld) in classes(lmax):
        variants = [("E", 0, eri_program(la, lb, lc, ld))]
        g = grad_program(la, lb, lc, ld)
        if g["P"] > 0:
            variants.append(("G", 1, g))
        else:  # too large for a workgroup's LDS: one program per centre (grad = 2 + centre)
            variants += [(f"G{c}_", 2 + c, grad_program(la, lb, lc, ld, c)) for c in range(3)]
        for kind, gflag, pr in variants:
            if pr["P"] == 0:
                continue  # does not fit a workgroup's LDS
            key = f"{kind}{la}{lb}{lc}{ld}"
            out += arr("uint32_t", f"kProgV_{key}", words(pr["vops"]))
            out += arr("uint32_t", f"kProgVG_{key}", words(pr["vgroups"]))
            out += arr("uint32_t", f"kProgA_{key}", words(pr["accs"]))
            out += arr("uint32_t", f"kProgH_{key}", words(pr["hops"]))
            out += arr("uint32_t", f"kProgHG_{key}", words(pr["hgroups"]))
            out += arr("int", f"kProgOut_{key}", [str(x) for o in pr["outs"] for x in (o[0], o[1], o[2], 0)])
            out += arr("double", f"kProgNorm_{key}", [repr(o[3]) for o in pr["outs"]], 4)
            entries.append(
                f"    {{{la}, {lb}, {lc}, {ld}, {gflag}, {pr['L']}, {pr['nacc']}, "
                f"{pr['nv']}, {pr['nh']}, {len(pr['outs'])}, {len(pr['vops'])}, "
                f"{len(pr['vgroups'])}, {len(pr['accs'])}, {len(pr['hops'])}, "
                f"{len(pr['hgroups'])}, {pr['P']}, kProgV_{key}, kProgVG_{key}, kProgA_{key}, "
                f"kProgH_{key}, kProgHG_{key}, kProgOut_{key}, kProgNorm_{key}}},")
            out.append("")
    out += ["const HostProgram kHostPrograms[] = {"] + entries + ["};", ""]
    write_if_changed(path, "\n".join(out))


def main():
    import os
    out_dir = sys.argv[1]
    lmax = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    only = set(sys.argv[3:])
    os.makedirs(out_dir, exist_ok=True)
    parts = ["// Generated by gen_eri.py -- do not edit.", ""]
    flops = {}
    for (la, lb, lc, ld) in classes(lmax):
        g = gen_class(la, lb, lc, ld)
        key = f"{la}{lb}{lc}{ld}"
        flops[key] = {"vrr_per_primitive_quartet": g["vrr_flops"],
                      "hrr_per_contracted_quartet": g["hrr_flops"],
                      "nacc": g["nacc"], "nout": g["nout"], "L": g["L"]}
        parts.append(f"template <> struct EriClass<{la}, {lb}, {lc}, {ld}> {{")
        parts.append(f"  static constexpr int L = {g['L']};")
        parts.append(f"  static constexpr int NACC = {g['nacc']};")
        parts.append(f"  static constexpr int NOUT = {g['nout']};")
        parts.append("  template <typename R>")
        parts.append("  __device__ static __forceinline__ void vrr(const R* __restrict__ G,"
                     " R PAx, R PAy, R PAz, R WPx, R WPy, R WPz,")
        parts.append("      R QCx, R QCy, R QCz, R WQx, R WQy, R WQz,"
                     " R oo2z, R oo2e, R oo2ze, R roz, R roe, R PQx, R PQy, R PQz,")
        parts.append("      R* __restrict__ acc) {")
        parts.append("    (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz;")
        parts.append("    (void)QCx; (void)QCy; (void)QCz; (void)WQx; (void)WQy; (void)WQz;")
        parts.append("    (void)oo2z; (void)oo2e; (void)oo2ze; (void)roz; (void)roe;")
        parts.append("    (void)PQx; (void)PQy; (void)PQz;")
        parts.extend("    " + s for s in g["vrr"])
        parts.append("  }")
        if g["L"] >= STREAM_L:
            bodies = stream_vrr(la, lb, lc, ld)
            parts.append(f"  static constexpr int NSTREAM = {len(bodies)};  // parts of the streamed VRR")
            parts.append("  template <int PART, typename SA>")
            parts.append("  __device__ static __forceinline__ void vrr_stream(const double* __restrict__ G,"
                         " double PAx, double PAy, double PAz, double WPx, double WPy, double WPz,")
            parts.append("      double QCx, double QCy, double QCz, double WQx, double WQy, double WQz,"
                         " double oo2z, double oo2e, double oo2ze, double roz, double roe, SA& acc) {")
            parts.append("    using R = double;")
            for pi, body in enumerate(bodies):
                parts.append(f"    {'if' if pi == 0 else '} else if'} constexpr (PART == {pi}) {{")
                parts.extend("      " + x for x in body)
            parts.append("    }")
            parts.append("  }")
        parts.append("  template <typename R>")
        parts.append("  __device__ static __forceinline__ void hrr(const R* __restrict__ acc,"
                     " R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, R* __restrict__ out) {")
        parts.append("    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")
        parts.extend("    " + s for s in g["hrr"])
        parts.append("  }")
        parts.append("};")
        parts.append("")
        tu = tu_source(la, lb, lc, ld)
        path = os.path.join(out_dir, f"eri_c_{key}.hip")
        old = open(path).read() if os.path.exists(path) else None
        if old != tu:  # keep timestamps of unchanged units
            with open(path, "w") as fh:
                fh.write(tu)
        if g["L"] >= STREAM_L:
            su = stream_tu_source(la, lb, lc, ld)
            path = os.path.join(out_dir, f"eri_s_{key}.hip")
            old = open(path).read() if os.path.exists(path) else None
            if old != su:
                with open(path, "w") as fh:
                    fh.write(su)
    write_if_changed(os.path.join(out_dir, "eri_classes.inc"), "\n".join(parts))
    write_programs(os.path.join(out_dir, "eri_programs.inc"), lmax)
    keys = [f"{a}{b}{c}{d}" for a, b, c, d in classes(lmax)]
    if only:
        keys = [k for k in keys if k in only]
    disp = ["// Generated by gen_eri.py -- do not edit.",
            f"static_assert(kNumLdsVariants == {N_LDS_VARIANTS}, \"gen_eri.py N_LDS_VARIANTS\");"]
    disp += [f"EriLaunchFn eri_launcher_c{k}(int mode, int bm, int fp32);" for k in keys]
    disp += [f"EriLaunchFn eri_jk_lds_launcher_c{k}(int bm, int fp32, int v);" for k in keys]
    disp += ["", "// nullptr when the class or the variant is not compiled in.",
             "EriLaunchFn eri_launcher(int la, int lb, int lc, int ld, int mode, int bm, int fp32) {"]
    for k in keys:
        disp.append(f"  if (la == {k[0]} && lb == {k[1]} && lc == {k[2]} && ld == {k[3]}) "
                    f"return eri_launcher_c{k}(mode, bm, fp32);")
    disp += ["  return nullptr;", "}", "",
             "EriLaunchFn eri_jk_lds_launcher(int la, int lb, int lc, int ld, int bm, int fp32,",
             "                                int v) {"]
    for k in keys:
        disp.append(f"  if (la == {k[0]} && lb == {k[1]} && lc == {k[2]} && ld == {k[3]}) "
                    f"return eri_jk_lds_launcher_c{k}(bm, fp32, v);")
    disp += ["  return nullptr;", "}", ""]
    skeys = [f"{a}{b}{c}{d}" for a, b, c, d in classes(lmax) if a + b + c + d >= STREAM_L]
    disp += [f"EriStreamFn eri_stream_launcher_c{k}(int mode);" for k in skeys]
    disp += ["", "// Streaming kernel of the class in `mode` (exact Boys); nullptr if none.",
             "EriStreamFn eri_stream_launcher(int la, int lb, int lc, int ld, int mode) {"]
    for k in skeys:
        disp.append(f"  if (la == {k[0]} && lb == {k[1]} && lc == {k[2]} && ld == {k[3]}) "
                    f"return eri_stream_launcher_c{k}(mode);")
    disp += ["  return nullptr;", "}", ""]
    write_if_changed(os.path.join(out_dir, "eri_dispatch.inc"), "\n".join(disp))
    write_if_changed(os.path.join(out_dir, "flops_per_class.json"),
                     json.dumps(flops, indent=1, sort_keys=True))
    # host-side copy of the operation counts: {la, lb, lc, ld, vrr/prim, hrr/contracted}
    rows = [f"  {{{k[0]}, {k[1]}, {k[2]}, {k[3]}, {v['vrr_per_primitive_quartet']}, "
            f"{v['hrr_per_contracted_quartet']}}}," for k, v in sorted(flops.items())]
    write_if_changed(os.path.join(out_dir, "eri_flops.inc"),
                     "// Generated by gen_eri.py -- do not edit.\n"
                     "static const int kEriFlops[][6] = {\n" + "\n".join(rows) + "\n};\n"
                     f"constexpr int kEriFlopsRows = {len(rows)};\n")


if __name__ == "__main__":
    main()
