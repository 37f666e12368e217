"""Autotuner of the Fock build's launch choices, per (la lb|lc ld) class and
precision, on the device it runs on.

For every class kernel the build launches it measures, with the library's own
per-class HIP-event timings:
  1. every compiled launch shape of the LDS-digestion kernel (waves per
     workgroup, register bound for occupancy, one-ahead ket fetch, bra
     primitive pairs per inner iteration, branch-free Boys function;
     rchem_lds_variant_info) at chunks of 64 and of 256 tiles, all classes
     switched together, one timed set of builds per setting;
  2. the chunk order: chunks sorted by their first ket pair (workgroups
     running together share ket records in L2) or left in bra order;
  3. persistent workgroup counts (512 ... 8192) at each class's best shape;
  4. chunk sizes (16 ... 256 tiles) and chunk order together, at the best
     shape and count; a precision whose ket-ordered classes together save
     less than sorting its chunk list costs (SORT_MS) stays in bra order.
A choice is kept for a class only where it beats the class's best so far by
more than the measured noise (1 %).  The FP64 kernels are timed in the
all-FP64 build and in the FP64 part of the mixed build and tuned on the sum;
the FP32 kernels in the mixed build.  The table is written in the format
rchem_load_tuning reads (data/tuning_gfx950.txt beside the library is the
default every context loads), and a JSON report of default vs tuned ms per
class goes to stdout.

    python -m rchem_amd.autotune [--workload h2o64|h2o32|eri300|batch512] [--out PATH]

The workload is a Fock build on the density the bench uses (core-Hamiltonian
guess) so that the mixed-precision split is the bench's; batch512 times the
first iterations of the 512-molecule batched SCF instead and writes the
batched SCF's own table.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "tuning_gfx950.txt")
DEFAULT_OUT_BATCH = os.path.join(HERE, "data", "tuning_gfx950_batch.txt")
BLOCKS = [512, 1024, 2048, 4096, 8192]
CHUNKS = [16, 32, 64, 128, 256]
NOISE = 0.01
# Sorting one precision's chunk list costs the planner ~0.13 ms per build
# ((H2O)_64): the kernels a precision's ket order speeds up together must save
# more than that, or that precision stays in bra order.
SORT_MS = 0.15


def _log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def core_guess_density(S, H, nocc):
    lam, U = np.linalg.eigh(S)
    X = U @ np.diag(lam ** -0.5) @ U.T
    _, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    return C[:, :nocc] @ C[:, :nocc].T


class Tuner:
    def __init__(self, ctx, basis, D, steps):
        from ._lib import PREC_FP64, PREC_MIXED, default_opts, lib
        from .device import DeviceMatrix
        self.L = lib()
        self.ctx = ctx
        self.B = basis
        n = basis.nbf
        self.dD, self.dJ, self.dK = DeviceMatrix(n), DeviceMatrix(n), DeviceMatrix(n)
        self.dD.upload(D)
        self.steps = steps
        self.opts = {"fp64": default_opts(precision=PREC_FP64, schwarz_thresh=1e-12),
                     "mixed": default_opts(precision=PREC_MIXED, schwarz_thresh=1e-12)}
        nv = 0
        while self.L.rchem_lds_variant_info(nv, None, None, None, None, None) == 0:
            nv += 1
        self.nvariants = nv

    def build(self, mode):
        self.ctx.check(self.L.rchem_jk_device(self.ctx.h, self.dD.ptr, self.dJ.ptr, self.dK.ptr,
                                              ctypes.byref(self.opts[mode])))

    def time_mode(self, mode):
        """Per-class ms per build {(la, lb, lc, ld, fp32): ms} in one precision mode."""
        from ._lib import ClassTiming
        from .device import synchronize
        self.build(mode)  # re-plans after a tuning change, warms up
        synchronize()
        L, h = self.L, self.ctx.h
        self.ctx.check(L.rchem_reset_class_timings(h))
        self.ctx.check(L.rchem_set_profiling(h, 1))
        for _ in range(self.steps):
            self.build(mode)
        synchronize()
        self.ctx.check(L.rchem_set_profiling(h, 0))
        t = (ClassTiming * 128)()
        cnt = ctypes.c_int()
        self.ctx.check(L.rchem_jk_class_timings(h, t, 128, ctypes.byref(cnt)))
        return {(x.la, x.lb, x.lc, x.ld, x.fp32): x.ms / self.steps for x in t[:cnt.value]}

    def time_all(self):
        """Per (class, fp32) ms summed over the FP64 and the mixed build."""
        out = {}
        for mode in ("fp64", "mixed"):
            for k, v in self.time_mode(mode).items():
                out[k] = out.get(k, 0.0) + v
        return out

    def set(self, key, variant, blocks, chunk, order):
        la, lb, lc, ld, fp32 = key
        self.ctx.check(self.L.rchem_set_class_tuning(self.ctx.h, la, lb, lc, ld, fp32, variant,
                                                     blocks, chunk))
        self.ctx.check(self.L.rchem_set_class_ket_order(self.ctx.h, la, lb, lc, ld, fp32, order))

    def order(self, key):
        out = ctypes.c_int()
        self.ctx.check(self.L.rchem_get_class_ket_order(self.ctx.h, *key, ctypes.byref(out)))
        return out.value


class BatchTuner(Tuner):
    """The batched SCF's Fock builds: the first iterations of rchem_batch_scf
    over the molecules (multi-molecule bases), class kernels timed over all
    their builds; the batch runs on the context's table while it tunes."""

    def __init__(self, ctx, molecules, iters):
        from ._lib import lib
        self.L = lib()
        self.ctx = ctx
        self.mols = molecules
        self.iters = iters
        nv = 0
        while self.L.rchem_lds_variant_info(nv, None, None, None, None, None) == 0:
            nv += 1
        self.nvariants = nv

    def time_mode(self, mode):
        from . import basis as rb
        from ._lib import PREC_FP64, PREC_MIXED, ClassTiming, default_opts
        L, h = self.L, self.ctx.h
        o = default_opts(precision=PREC_MIXED if mode == "mixed" else PREC_FP64)
        self.ctx.check(L.rchem_reset_class_timings(h))
        self.ctx.check(L.rchem_set_profiling(h, 1))
        rb.batch_scf(self.mols, "STO-3G", thresh_e=1e-10, diis=8, max_iter=self.iters, opts=o,
                     ctx=self.ctx, damping=0.3, damping_until=1e-2)
        self.ctx.check(L.rchem_set_profiling(h, 0))
        t = (ClassTiming * 128)()
        cnt = ctypes.c_int()
        self.ctx.check(L.rchem_jk_class_timings(h, t, 128, ctypes.byref(cnt)))
        return {(x.la, x.lb, x.lc, x.ld, x.fp32): x.ms for x in t[:cnt.value]}


def tune(tuner):
    from ._lib import lib
    L = lib()
    L.rchem_load_tuning(tuner.ctx.h, None, None)  # start from the defaults
    L.rchem_load_batch_tuning(tuner.ctx.h, None, None)  # a batch runs on those too
    base = tuner.time_all()
    keys = sorted(base)
    _log(f"defaults: {len(keys)} class kernels, {sum(base.values()):.2f} ms "
         f"(FP64 + mixed build)")
    best = {k: {"variant": 0, "blocks": 2048, "chunk": 64, "order": tuner.order(k), "ms": base[k]}
            for k in keys}

    def sweep(name, values, apply):
        for val in values:
            for k in keys:
                tuner.set(k, *apply(best[k], val))
            t = tuner.time_all()
            for k in keys:
                if k in t and t[k] < best[k]["ms"] * (1.0 - NOISE):
                    v, b, c, o = apply(best[k], val)
                    best[k] = {"variant": v, "blocks": b, "chunk": c, "order": o, "ms": t[k]}
            _log(f"{name} {val}: {sum(t.get(k, 0.0) for k in keys):.2f} ms")
        for k in keys:
            tuner.set(k, best[k]["variant"], best[k]["blocks"], best[k]["chunk"], best[k]["order"])

    # shape and chunk size interact (the chunk is the workgroup's share of one
    # bra pair): every shape at a short and at a long chunk first
    for chunk in (64, 256):
        sweep(f"shape at chunk {chunk}:", range(tuner.nvariants),
              lambda cur, v, c=chunk: (v, cur["blocks"], c, cur["order"]))
    sweep("ket order", (0, 1), lambda cur, o: (cur["variant"], cur["blocks"], cur["chunk"], o))
    sweep("workgroups", BLOCKS, lambda cur, b: (cur["variant"], b, cur["chunk"], cur["order"]))
    # chunk length and order together (a chunk's length sets how many kets its
    # workgroup walks through, and so how much ket order can share)
    sweep("chunk, order", [(c, o) for o in (0, 1) for c in CHUNKS],
          lambda cur, co: (cur["variant"], cur["blocks"], co[0], co[1]))
    # a precision whose ket-ordered classes save less than its list's sort
    # costs goes back to bra order
    t_bra = {}
    for k in keys:
        tuner.set(k, best[k]["variant"], best[k]["blocks"], best[k]["chunk"], 0)
    t_bra = tuner.time_all()
    for fp32 in (0, 1):
        on = [k for k in keys if k[4] == fp32 and best[k]["order"] == 1 and k in t_bra]
        saved = sum(t_bra[k] - best[k]["ms"] for k in on)
        if on and saved < SORT_MS:
            _log(f"ket order of the {'FP32' if fp32 else 'FP64'} lists saves {saved:.3f} ms "
                 f"< {SORT_MS} ms: bra order")
            for k in on:
                best[k]["order"] = 0
                best[k]["ms"] = t_bra[k]
    for k in keys:
        tuner.set(k, best[k]["variant"], best[k]["blocks"], best[k]["chunk"], best[k]["order"])
    final = tuner.time_all()
    return base, best, final


def write_table(path, best, device, workload):
    lines = ["# Fock-build launch tuning, written by rchem_amd/autotune.py",
             f"# device: {device}; workload: {workload}",
             "# la lb lc ld fp32 variant blocks chunk_tiles ket_order   "
             "(ms per FP64 + mixed build)"]
    for k in sorted(best):
        b = best[k]
        lines.append(f"{k[0]} {k[1]} {k[2]} {k[3]} {k[4]} {b['variant']} {b['blocks']} "
                     f"{b['chunk']} {b['order']}   # {b['ms']:.3f}")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=["h2o64", "h2o32", "eri300", "batch512"],
                    default="h2o64")
    ap.add_argument("--out", default=None,
                    help="table to write (default: data/tuning_gfx950.txt, for batch512 "
                         "data/tuning_gfx950_batch.txt, beside the library)")
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args(argv)
    from . import basis as rb
    from .inputs import nocc_of, random_molecules, synthetic_shells, water_cluster

    if args.out is None:
        args.out = DEFAULT_OUT_BATCH if args.workload == "batch512" else DEFAULT_OUT
    t0 = time.perf_counter()
    ctx = rb.Context(0)
    if args.workload == "batch512":
        mols = random_molecules(count=512, seed=512)
        tuner = BatchTuner(ctx, mols, args.steps)
        workload = (f"batched SCF of 512 random H/C/N/O molecules (seed 512), STO-3G, "
                    f"the first {args.steps} iterations")
        _log(f"{workload}: {tuner.nvariants} launch shapes")
    elif args.workload in ("h2o64", "h2o32"):
        dims = (4, 4, 4) if args.workload == "h2o64" else (2, 4, 4)
        Z, xyz = water_cluster(dims=dims, seed=64 if args.workload == "h2o64" else 32)
        B = rb.Basis.new(Z, xyz, "6-31G", ctx=ctx)
        S, T, V = rb.one_electron(B, xyz, Z)
        D = core_guess_density(S, T + V, nocc_of(Z))
        workload = f"{args.workload} 6-31G Fock build, core-guess density, Schwarz 1e-12"
    else:
        sh = synthetic_shells(nshell=300, seed=2000, box=20.0 * (300 / 2000.0) ** (1 / 3), nprim=3)
        B = rb.Basis.from_shells(*sh, ctx=ctx)
        X = np.random.default_rng(2000).standard_normal((B.nbf, B.nbf)) / B.nbf
        D = X + X.T
        workload = "300-shell s/p/d microbenchmark, Schwarz 1e-12"
    if args.workload != "batch512":
        tuner = Tuner(ctx, B, D, args.steps)
        _log(f"{workload}: nbf {B.nbf}, {tuner.nvariants} launch shapes")
    base, best, final = tune(tuner)
    dev = os.environ.get("RCHEM_DEVICE_NAME", "gfx950")
    write_table(args.out, best, dev, workload)
    report = {"workload": workload, "table": os.path.relpath(args.out),
              "seconds": time.perf_counter() - t0,
              "default_ms": sum(base.values()), "tuned_ms": sum(final.values()),
              "classes": [{"class": f"({k[0]}{k[1]}|{k[2]}{k[3]})" + (" fp32" if k[4] else ""),
                           "default_ms": base[k], "tuned_ms": final.get(k),
                           "variant": best[k]["variant"], "blocks": best[k]["blocks"],
                           "chunk_tiles": best[k]["chunk"], "ket_order": best[k]["order"]}
                          for k in sorted(best)]}
    print(json.dumps(report), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
