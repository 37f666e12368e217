#!/usr/bin/env python3
"""Fock-build benchmark: (H2O)_64, 6-31G (832 functions), one GPU.

A step is one Fock build (J and K for a fixed density) through
``rchem_jk_device`` with D, J, K resident in HBM.  The default run times the
mixed-precision build (FP32 for quartets whose Schwarz bound times max|D| is
below 1e-7, FP64 for the rest) and the all-FP64 build on the same density,
then runs the SCF to convergence both ways (DIIS, all-FP64 final iterations
after the convergence-triggered switch) and reports the energy difference.
Prints one JSON line:

  value        unique contracted shell quartets per second (after Schwarz
               screening at 1e-12 and 8-fold symmetry) of the mixed build;
  ms_per_step  mixed Fock-build milliseconds; "fp64" holds the all-FP64 build,
               "speedup_mixed_vs_fp64" their ratio, "scf" the two SCF energies;
  roofline     the dominant (la lb|lc ld) class kernel of the mixed build:
               algorithmic flops per launch (generator operation counts,
               FMA = 2) over its launch time from HIP events on the library's
               stream, HBM traffic from the committed rocprofv3 counter runs;
  cpu_baseline the reference algorithm (rchem's JK_direct loop over the
               reference's own cints coulomb_repulsion, oracle/_ref) on one
               host core over a bounded random sample of function quartets,
               expressed as the same Fock build's quartets per second.

Options: --steps K --warmup W --gpus N (N must be 1) --no-cpu-baseline
         --cpu-seconds S --no-scf --precision both|fp64|mixed
         --workload h2o64|eri2000 [--shells N]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector
HBM_PEAK_GBS = 8000.0


def core_guess_density(S, H, nocc):
    lam, U = np.linalg.eigh(S)
    X = U @ np.diag(lam ** -0.5) @ U.T
    _, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    return C[:, :nocc] @ C[:, :nocc].T


def host_cores():
    """Cores this process may use, capped at the 16 a one-GPU box gives it."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _sample_rate(O, B, D, addr, seconds, threads):
    """Function quartets/s of the JK_direct body over random quartets."""
    rng = np.random.default_rng(1234)
    done, t_total, batch = 0, 0.0, 256 * threads
    while t_total < seconds:
        q = rng.integers(0, B.n, (batch, 4))
        t0 = time.perf_counter()
        O.jk_direct_sample(B, D, q, boys=O.BOYS_COMPAT, cints_fn=addr, threads=threads)
        t_total += time.perf_counter() - t0
        done += batch
        batch = min(batch * 2, 8192 * threads)
    return done, t_total


def jk_rates(Bo, D, seconds):
    """rchem's JK_direct loop (basis.rs:383-428) over the reference's cints
    coulomb_repulsion (oracle/_ref; the C restatement where that is not
    built): function quartets/s on one core and on every core, each from a
    bounded random sample."""
    import oracle as O
    addr = O.cints_coulomb_repulsion_address()
    cores = host_cores()
    d1, t1 = _sample_rate(O, Bo, D, addr, seconds, 1)
    dm, tm = _sample_rate(O, Bo, D, addr, seconds, cores)
    return {"kind": "reference" if addr else "port",
            "src": "the reference cints.c coulomb_repulsion" if addr else "the C restatement",
            "rate_1": d1 / t1, "done_1": d1, "t_1": t1,
            "rate_all": dm / tm, "done_all": dm, "t_all": tm, "cores": cores}


def cpu_baseline(Z, xyz, D, gpu_quartets, seconds, water_gpu_scf_s):
    """The reference algorithm on the host: rchem's JK_direct loop (basis.rs:383-428)
    over the reference's own cints coulomb_repulsion (oracle/_ref) -- on one core
    (rchem is single-threaded) and on every core (OpenMP over mu, each row
    private, basis.rs:422-423).  (H2O)_64: a bounded random sample of function
    quartets, extrapolated to the n^4 quartets one build visits.  Water STO-3G
    (7 functions): rchem's whole SCF end to end (rchem.rs:40-90)."""
    import oracle as O
    from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z
    bse = O.load_bse_json(os.path.join(ROOT, "rchem_amd", "data", "6-31g.json"))
    B = O.cgtos_from_bse(bse, Z, xyz)
    addr = O.cints_coulomb_repulsion_address()
    kind = "reference" if addr else "port"
    src = "the reference cints.c coulomb_repulsion" if addr else "the C restatement"
    cores = host_cores()
    done, t_total = _sample_rate(O, B, D, addr, seconds, 1)
    rate = done / t_total                       # function quartets / s (each: J and K ERI)
    fock_s = B.n ** 4 / rate                    # JK_direct visits n^4 function quartets
    done_mt, t_mt = _sample_rate(O, B, D, addr, seconds, cores)
    rate_mt = done_mt / t_mt
    fock_mt = B.n ** 4 / rate_mt
    # water STO-3G, rchem's SCF end to end on the reference integrals
    sto = O.load_bse_json(os.path.join(ROOT, "rchem_amd", "data", "sto-3g.json"))
    wxyz = np.array(WATER_CRAWFORD, dtype=float).reshape(-1, 3)
    Bw = O.cgtos_from_bse(sto, WATER_Z, wxyz)
    water = {}
    for th in (1, cores):
        t0 = time.perf_counter()
        S, T, V = O.one_electron(Bw, WATER_Z, wxyz)
        r = O.scf(S, T + V, lambda Dm: O.jk_direct(Bw, Dm, boys=O.BOYS_COMPAT, cints_fn=addr,
                                                   threads=th), nocc=5)
        water[th] = (time.perf_counter() - t0, r)
    return {
        "value": gpu_quartets / fock_s,
        "unit": "shell quartets/s (same Fock build)",
        "cores": 1,
        "kind": kind,
        "sample": (f"{done} random (mu,nu,la,si) function quartets of the (H2O)_64 6-31G "
                   f"JK_direct loop in {t_total:.1f} s ({rate:.0f}/s, 2 contracted ERIs each via "
                   f"{src}); extrapolated Fock build {fock_s:.3g} s"),
        "fock_build_s_extrapolated": fock_s,
        "all_cores": {
            "value": gpu_quartets / fock_mt, "cores": cores,
            "sample": (f"{done_mt} random function quartets in {t_mt:.1f} s on {cores} threads "
                       f"({rate_mt:.0f}/s); extrapolated Fock build {fock_mt:.3g} s"),
            "fock_build_s_extrapolated": fock_mt},
        "water_sto3g_scf": {
            "what": "rchem's whole SCF (plain iterations, |dE| < 1e-11) on water STO-3G, "
                    "JK_direct over the reference integrals every iteration",
            "iterations": water[1][1].iterations, "e_elec": water[1][1].e_elec,
            "seconds_1_core": water[1][0], "seconds_all_cores": water[cores][0],
            "cores": cores, "gpu_seconds": water_gpu_scf_s},
    }


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    from the committed rocprofv3 counter runs of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"eri_jk_h2o64_631g_{precision}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        k = json.load(fh)["kernels"].get(kernel)
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def batch_workload(args, ctx, t_start):
    """512 random H/C/N/O molecules (3-30 atoms, seed 512), STO-3G, solved
    together: FP64 and mixed precision, DIIS(8) with density damping (30 % of
    the previous density while the commutator exceeds 1e-2), |dE| < 1e-10, at
    most 128 iterations."""
    from rchem_amd import basis as rb
    from rchem_amd._lib import PREC_FP64, PREC_MIXED, default_opts
    from rchem_amd.inputs import random_molecules
    import torch

    def say(msg):
        print(f"[bench] {msg} ({time.perf_counter() - t_start:.1f} s)", file=sys.stderr, flush=True)

    mols = random_molecules(count=args.molecules, seed=512, nmin=3, nmax=30)
    nbf = [sum(1 if z == 1 else 5 for z in Z) for Z, _ in mols]
    # warm-up on a few molecules (programs, kernels, solver workspaces)
    rb.batch_scf(mols[:8], "STO-3G", thresh_e=1e-10, diis=8, max_iter=16, ctx=ctx)
    out_runs = {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = rb.batch_scf(mols, "STO-3G", thresh_e=1e-10, diis=8, max_iter=128, ctx=ctx,
                           opts=default_opts(precision=prec, schwarz_thresh=1e-12), graph=True,
                           damping=0.3, damping_until=1e-2)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out_runs[name] = (res, dt)
        say(f"batch {name}: {sum(r.converged for r in res)}/{len(res)} converged, {dt:.2f} s")
    r64, t64 = out_runs["fp64"]
    rmx, tmx = out_runs["mixed"]
    both = [i for i in range(len(mols)) if r64[i].converged and rmx[i].converged]
    out = {
        "metric": "Batched SCF: molecules solved per second (512 random H/C/N/O molecules, STO-3G)",
        "value": len(mols) / tmx,
        "unit": "molecules/s (mixed precision)",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{len(mols)} random-walk molecules, 3-30 atoms from H/C/N/O, seed "
                               f"512, STO-3G ({min(nbf)}-{max(nbf)} functions), DIIS(8), damping 0.3 "
                               f"while max|commutator| > 1e-2, "
                               f"|dE| < 1e-10 Eh, max 128 iterations, chunks of <= 8192 functions"},
        "fp64_seconds": t64,
        "mixed_seconds": tmx,
        "fp64_molecules_per_s": len(mols) / t64,
        "speedup_mixed_vs_fp64": t64 / tmx,
        "converged_fp64": int(sum(r.converged for r in r64)),
        "converged_mixed": int(sum(r.converged for r in rmx)),
        "max_abs_delta_e_mixed_vs_fp64": max((abs(r64[i].e_total - rmx[i].e_total)
                                              for i in both), default=None),
        "mean_iterations_fp64": float(np.mean([r.iterations for r in r64 if r.converged]))
        if any(r.converged for r in r64) else None,
        "total_functions": int(sum(nbf)),
    }
    if not args.no_cpu_baseline:
        # rchem solves one molecule at a time: per molecule, its SCF's Fock
        # builds (the iterations the GPU run took) of n^4 function quartets at
        # the reference loop's rate, sampled on a few of the molecules
        import oracle as O
        bse = O.load_bse_json(os.path.join(ROOT, "rchem_amd", "data", "sto-3g.json"))
        rates, rates_all, done, secs = [], [], 0, 0.0
        pick = np.random.default_rng(512).choice(len(mols), 4, replace=False)
        for i in pick:
            Zm, xm = mols[i]
            Bo = O.cgtos_from_bse(bse, Zm, np.asarray(xm, dtype=float).reshape(-1, 3))
            Dm = np.eye(Bo.n) * 0.5
            r = jk_rates(Bo, Dm, args.cpu_seconds / 4)
            rates.append(r["rate_1"])
            rates_all.append(r["rate_all"])
            done += r["done_1"]
            secs += r["t_1"]
        rate, rate_all = float(np.mean(rates)), float(np.mean(rates_all))
        quartets = sum(res_i.iterations * float(nb) ** 4 for res_i, nb in zip(rmx, nbf))
        out["cpu_baseline"] = {
            "value": len(mols) / (quartets / rate), "unit": "molecules/s", "cores": 1,
            "kind": r["kind"],
            "sample": (f"{done} random function quartets of the JK_direct loop on molecules "
                       f"{sorted(pick.tolist())} in {secs:.1f} s (mean {rate:.0f}/s via "
                       f"{r['src']}); the batch's Fock builds (each molecule's GPU iteration "
                       f"count x n^4) = {quartets:.3g} function quartets"),
            "batch_s_extrapolated": quartets / rate,
            "all_cores": {"value": len(mols) / (quartets / rate_all), "cores": r["cores"],
                          "batch_s_extrapolated": quartets / rate_all}}
    print(json.dumps(out), flush=True)


def gradient_workload(args, ctx, t_start):
    """(H2O)_32 6-31G (416 functions, 288 coordinates): forward SCF time, the
    analytic gradient's time, and central differences (h = 1e-4 bohr) of the
    GPU energy on sampled coordinates."""
    from rchem_amd import basis as rb
    from rchem_amd._lib import default_opts
    from rchem_amd.inputs import nocc_of, water_cluster
    import torch

    def say(msg):
        print(f"[bench] {msg} ({time.perf_counter() - t_start:.1f} s)", file=sys.stderr, flush=True)

    Z, xyz = water_cluster(dims=(2, 4, 4), seed=32)
    nocc = nocc_of(Z)
    o = default_opts(schwarz_thresh=1e-12)

    def energy(p):
        B = rb.Basis.new(Z, p, "6-31G", ctx=ctx)
        res, _ = rb.scf(B, p, Z, nocc=nocc, thresh_e=1e-11, diis=8, opts=o)
        assert res.converged
        return res.e_total

    B = rb.Basis.new(Z, xyz, "6-31G", ctx=ctx)
    energy(xyz)  # warm-up (pair layout, programs, kernels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, D_gpu = rb.scf(B, xyz, Z, nocc=nocc, thresh_e=1e-11, diis=8, opts=o)
    t_scf = time.perf_counter() - t0
    say(f"forward SCF {res.iterations} iterations, {t_scf:.3f} s")
    t0 = time.perf_counter()
    resg, g = rb.rhf_gradient(B, xyz, Z, nocc=nocc, thresh_e=1e-11, diis=8, opts=o)
    t_grad = time.perf_counter() - t0
    say(f"gradient call {t_grad:.3f} s (includes its own SCF, {resg.iterations} iterations)")
    ncoord = 3 * len(Z)
    if args.fd_coords == "all":
        coords = list(range(ncoord))
    else:
        coords = sorted(np.random.default_rng(32).choice(ncoord, int(args.fd_coords),
                                                         replace=False).tolist())
    h = 1e-4
    worst = 0.0
    for c in coords:
        a, x = divmod(c, 3)
        xp, xm = xyz.copy(), xyz.copy()
        xp[a, x] += h
        xm[a, x] -= h
        fd = (energy(xp) - energy(xm)) / (2 * h)
        worst = max(worst, abs(fd - g[a, x]))
        say(f"coordinate {c}: analytic {g[a, x]:+.9f} FD {fd:+.9f}")
    out = {
        "metric": "RHF nuclear gradient, (H2O)_32 6-31G: analytic adjoint time vs forward SCF, "
                  "max |analytic - central FD| (Eh/bohr)",
        "value": t_grad - t_scf,
        "unit": "s (gradient beyond its SCF)",
        "higher_is_better": False,
        "n_gpus": 1,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "(H2O)_32 6-31G, seed 32, 2x4x4 lattice 3.1 A, DIIS(8), "
                               "Schwarz 1e-12 (gradient quartets 1e-14)", "nbf": B.nbf,
                   "atoms": len(Z)},
        "forward_scf_s": t_scf,
        "scf_iterations": res.iterations,
        "gradient_call_s": t_grad,
        "gradient_over_scf": (t_grad - t_scf) / t_scf,
        "e_total": res.e_total,
        "max_abs_gradient": float(np.max(np.abs(g))),
        "fd_coordinates": len(coords),
        "fd_step_bohr": h,
        "max_abs_fd_error": worst,
        "net_force": float(np.max(np.abs(g.sum(axis=0)))),
    }
    if not args.no_cpu_baseline:
        # rchem has no gradient: the reference CPU path to one is central
        # differences of its SCF energy, 2 x 288 SCFs of JK_direct builds
        import oracle as O
        bse = O.load_bse_json(os.path.join(ROOT, "rchem_amd", "data", "6-31g.json"))
        Bo = O.cgtos_from_bse(bse, Z, xyz)
        r = jk_rates(Bo, D_gpu, args.cpu_seconds)
        n4 = float(Bo.n) ** 4
        scf_s = res.iterations * n4 / r["rate_1"]
        out["cpu_baseline"] = {
            "value": 2 * ncoord * scf_s, "unit": "s (gradient by central differences)",
            "cores": 1, "kind": r["kind"],
            "sample": (f"{r['done_1']} random function quartets of the (H2O)_32 6-31G JK_direct "
                       f"loop in {r['t_1']:.1f} s ({r['rate_1']:.0f}/s via {r['src']}); one SCF "
                       f"= {res.iterations} builds = {scf_s:.3g} s, a gradient = 2 x {ncoord} "
                       f"SCFs"),
            "forward_scf_s_extrapolated": scf_s,
            "all_cores": {"value": 2 * ncoord * res.iterations * n4 / r["rate_all"],
                          "cores": r["cores"],
                          "forward_scf_s_extrapolated": res.iterations * n4 / r["rate_all"]}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-scf", action="store_true", help="skip the two full SCF runs")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="Fock builds to time (profiling runs take one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--fd-coords", default="6",
                    help="grad32 workload: coordinates checked by central differences "
                         "(a count sampled with seed 32, or 'all')")
    ap.add_argument("--molecules", type=int, default=512, help="batch512 workload: molecule count")
    ap.add_argument("--workload", choices=["h2o64", "eri2000", "grad32", "batch512"],
                    default="h2o64",
                    help="h2o64: (H2O)_64 6-31G Fock build (default); eri2000: the synthetic "
                         "2000-shell s/p/d ERI microbenchmark, per class; grad32: analytic "
                         "RHF gradient of (H2O)_32 6-31G against the forward SCF")
    ap.add_argument("--shells", type=int, default=2000,
                    help="eri2000 workload: number of shells (the cube grows to keep the density)")
    ap.add_argument("--no-tuning-compare", dest="tuning_compare", action="store_false",
                    help="skip timing the builds again without the tuning table")
    ap.add_argument("--mixed-sweep", type=float, nargs="*", default=[1e-7, 1e-6, 1e-5, 1e-4, 1e-3],
                    help="h2o64: mixed-precision thresholds to sweep after the SCF")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library does not shard the Fock build across "
              "GPUs (that would split the quartet list across ranks and all-reduce J and K once "
              "per SCF iteration); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from rchem_amd import basis as rb
    from rchem_amd._lib import (PREC_FP64, PREC_MIXED, ClassTiming, default_opts, lib)
    from rchem_amd.inputs import nocc_of, synthetic_shells, water_cluster

    t_start = time.perf_counter()
    ctx = rb.Context(0)
    if args.workload == "grad32":
        gradient_workload(args, ctx, t_start)
        ctx.close()
        return
    if args.workload == "batch512":
        batch_workload(args, ctx, t_start)
        ctx.close()
        return
    if args.workload == "h2o64":
        Z, xyz = water_cluster(dims=(4, 4, 4), seed=64)
        B = rb.Basis.new(Z, xyz, "6-31G", ctx=ctx)
        n = B.nbf
        S, T, V = rb.one_electron(B, xyz, Z)
        D = core_guess_density(S, T + V, nocc_of(Z))
        workload = ("(H2O)_64 6-31G Fock build (J and K), seed 64, 4x4x4 lattice 3.1 A, "
                    "Schwarz 1e-12")
    else:
        box = 20.0 * (args.shells / 2000.0) ** (1.0 / 3.0)
        sh = synthetic_shells(nshell=args.shells, seed=2000, box=box, nprim=3)
        B = rb.Basis.from_shells(*sh, ctx=ctx)
        n = B.nbf
        rng = np.random.default_rng(2000)
        X = rng.standard_normal((n, n)) / n
        D = X + X.T
        workload = (f"synthetic ERI microbenchmark: {args.shells} shells, seed 2000, "
                    f"{box:.1f}-bohr cube, l 0/1/2 at 0.5/0.35/0.15, 3 primitives, every quartet "
                    f"surviving Schwarz 1e-12, digested into J and K")
    dev = torch.device("cuda:0")
    dD = torch.from_numpy(D).to(dev)
    dJ = torch.zeros((n, n), dtype=torch.float64, device=dev)
    dK = torch.zeros((n, n), dtype=torch.float64, device=dev)
    L = lib()

    def say(msg):
        print(f"[bench] {msg} ({time.perf_counter() - t_start:.1f} s)", file=sys.stderr, flush=True)

    if args.copy_calibration_mb > 0:
        nel = args.copy_calibration_mb * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def measure(prec_name):
        prec = PREC_FP64 if prec_name == "fp64" else PREC_MIXED
        opts = default_opts(precision=prec, schwarz_thresh=1e-12)

        def step():
            ctx.check(L.rchem_jk_device(ctx.h, ctypes.c_void_p(dD.data_ptr()),
                                        ctypes.c_void_p(dJ.data_ptr()),
                                        ctypes.c_void_p(dK.data_ptr()), ctypes.byref(opts)))

        for i in range(args.warmup):
            step()
            torch.cuda.synchronize()
            say(f"{prec_name} warmup step {i + 1}/{args.warmup} done")
        ctx.check(L.rchem_reset_class_timings(ctx.h))
        ctx.check(L.rchem_set_profiling(ctx.h, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        say(f"{prec_name}: {args.steps} timed steps done")
        ctx.check(L.rchem_set_profiling(ctx.h, 0))
        st = rb.jk_last_stats(B)
        timings = (ClassTiming * 64)()
        cnt = ctypes.c_int()
        ctx.check(L.rchem_jk_class_timings(ctx.h, timings, 64, ctypes.byref(cnt)))
        classes = []
        for i in range(cnt.value):
            t = timings[i]
            avg = t.ms / args.steps   # per build (a large build launches a class once per pass)
            classes.append({"class": f"({t.la}{t.lb}|{t.lc}{t.ld})" + (" fp32" if t.fp32 else ""),
                            "avg_ms": avg, "launches_per_build": t.launches / args.steps,
                            "quartets": t.quartets, "prim_quartets": t.primitive_quartets,
                            "flops": t.flops,
                            "tflops": t.flops / (avg * 1e-3) / 1e12 if avg > 0 else 0.0})
        ms = elapsed / args.steps * 1e3
        return {"ms": ms, "st": st, "classes": classes,
                "quartets_per_s": st.quartets / (ms * 1e-3),
                "tflops": st.flops / (ms * 1e-3) / 1e12}

    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    if args.workload == "eri2000":
        precisions = ["fp64"]
    runs = {p: measure(p) for p in precisions}
    # the same builds with every class at its default launch choices (no
    # tuning table), against the table the context loaded
    tuning = None
    table = os.environ.get("RCHEM_TUNING_FILE",
                           os.path.join(ROOT, "rchem_amd", "data", "tuning_gfx950.txt"))
    if args.tuning_compare and table and os.path.exists(table):
        ctx.check(L.rchem_load_tuning(ctx.h, None, None))
        base = {p: measure(p) for p in precisions}
        ctx.check(L.rchem_load_tuning(ctx.h, table.encode(), None))
        tuning = {"table": os.path.relpath(table, ROOT)}
        for p in precisions:
            tc = {c["class"]: c["avg_ms"] for c in runs[p]["classes"]}
            tuning[p] = {"default_ms": base[p]["ms"], "tuned_ms": runs[p]["ms"],
                         "classes": [{"class": c["class"], "default_ms": c["avg_ms"],
                                      "tuned_ms": tc.get(c["class"])}
                                     for c in base[p]["classes"]]}
    main_run = runs["mixed"] if "mixed" in runs else runs["fp64"]
    st = main_run["st"]
    classes = main_run["classes"]
    ms = main_run["ms"]
    dom = max(classes, key=lambda c: c["avg_ms"])
    dom_fp32 = dom["class"].endswith("fp32")
    dom_peak = FP32_PEAK_TFLOPS if dom_fp32 else FP64_PEAK_TFLOPS
    traffic = measured_traffic(dom["class"], "mixed" if "mixed" in runs else "fp64")
    out = {
        "metric": "Fock-build ms & ERI quartets/s, mixed vs FP64 on GPU at |dE|<1e-6 Eh",
        "value": main_run["quartets_per_s"],
        "unit": "ERI shell quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if "mixed" in runs else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "nbf": n, "shells": len(B.shells()[1]),
                   "precision": "mixed" if "mixed" in runs else "fp64"},
        "quartets": st.quartets,
        "quartets_fp32": st.quartets_fp32,
        "tiles": st.tiles,
        "lane_fill": st.quartets / max(st.tiles * 64.0, 1.0),
        "primitive_quartets": st.primitive_quartets,
        "integrals": st.integrals,
        "fock_tflops": main_run["tflops"],
        "roofline": {"bound": "mfma", "achieved": dom["tflops"], "peak": dom_peak,
                     "unit": "TFLOP/s", "frac": dom["tflops"] / dom_peak,
                     "traffic": traffic, "kernel": dom["class"],
                     "note": ("VALU kernel; peak is MI355X's FP32 vector peak" if dom_fp32 else
                              "FP64 VALU kernel; MI355X FP64 vector peak = FP64 matrix peak")},
        "classes": classes,
    }
    if args.workload == "h2o64":
        # the same mixed build through the host-buffer entry point (rchem_jk:
        # D copied in, J and K copied out over PCIe), for reference only
        Jh, Kh = np.zeros((n, n)), np.zeros((n, n))
        oh = default_opts(precision=PREC_MIXED if "mixed" in runs else PREC_FP64,
                          schwarz_thresh=1e-12)
        rb.JK_direct(Jh, Kh, B, D, oh)
        t0 = time.perf_counter()
        for _ in range(3):
            rb.JK_direct(Jh, Kh, B, D, oh)
        out["host_buffers_ms_per_step"] = (time.perf_counter() - t0) / 3 * 1e3
    if tuning is not None:
        out["tuning"] = tuning
    if "mixed" in runs and "fp64" in runs:
        f = runs["fp64"]
        out["fp64"] = {"ms_per_step": f["ms"], "value": f["quartets_per_s"],
                       "fock_tflops": f["tflops"], "classes": f["classes"]}
        out["speedup_mixed_vs_fp64"] = f["ms"] / ms
    if args.workload == "eri2000":
        out["metric"] = "ERI shell quartets/s per (la lb|lc ld) class, synthetic 2000-shell microbench"
        out["roofline"]["traffic"] = None
        if not args.no_cpu_baseline:
            import oracle as O
            r = jk_rates(O.cgtos_from_shells(*sh), D, args.cpu_seconds)
            n4 = float(n) ** 4  # JK_direct visits every (mu, nu, la, si): no screening
            out["cpu_baseline"] = {
                "value": st.quartets / (n4 / r["rate_1"]),
                "unit": "shell quartets/s (same Fock build)", "cores": 1, "kind": r["kind"],
                "sample": (f"{r['done_1']} random (mu,nu,la,si) function quartets of this "
                           f"basis's JK_direct loop in {r['t_1']:.1f} s ({r['rate_1']:.0f}/s, "
                           f"2 contracted ERIs each via {r['src']}); extrapolated Fock build "
                           f"{n4 / r['rate_1']:.3g} s"),
                "fock_build_s_extrapolated": n4 / r["rate_1"],
                "all_cores": {"value": st.quartets / (n4 / r["rate_all"]), "cores": r["cores"],
                              "fock_build_s_extrapolated": n4 / r["rate_all"],
                              "sample": (f"{r['done_all']} random function quartets in "
                                         f"{r['t_all']:.1f} s on {r['cores']} threads")}}
    if args.workload == "h2o64" and not args.no_scf:
        scf = {}
        for p in ("fp64", "mixed"):
            o = default_opts(precision=PREC_FP64 if p == "fp64" else PREC_MIXED,
                             schwarz_thresh=1e-12)
            secs, energies = {}, {}
            for graph, purify in ((False, False), (True, False), (False, True), (True, True)):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res, Dc = rb.scf(B, xyz, Z, nocc=nocc_of(Z), thresh_e=1e-10, diis=8, opts=o,
                                 graph=graph, purify=purify)
                secs[(graph, purify)] = time.perf_counter() - t0
                energies[(graph, purify)] = res.e_total
            scf[p] = {"e_total": res.e_total, "iterations": res.iterations,
                      "converged": bool(res.converged), "fp64_from": res.fp64_from,
                      "seconds": secs[(True, True)],
                      "graph_iterations": res.graph_iterations,
                      "seconds_kernel_by_kernel_purified": secs[(False, True)],
                      "seconds_graph_eigensolver": secs[(True, False)],
                      "seconds_kernel_by_kernel_eigensolver": secs[(False, False)],
                      "max_delta_e_between_variants": max(energies.values()) - min(energies.values())}
            say(f"SCF {p}: {res.iterations} iterations, E = {res.e_total:.10f}, "
                f"{secs[(True, True)]:.2f} s (whole iteration one hipGraph, {res.graph_iterations} "
                f"replays, purified densities), {secs[(False, True)]:.2f} s (kernel by kernel, "
                f"purified), {secs[(True, False)]:.2f} s (hipGraph Fock build, eigensolver), "
                f"{secs[(False, False)]:.2f} s (kernel by kernel, eigensolver)")
            if p == "fp64":
                D_conv = Dc
        scf["delta_e"] = abs(scf["mixed"]["e_total"] - scf["fp64"]["e_total"])
        # the mixed Fock build itself at the converged density:
        # E = sum D o (H + F), F = H + 2J - K, mixed vs all-FP64 J and K
        Hc = T + V
        e_build = {}
        for p in ("fp64", "mixed"):
            Jm, Km = np.zeros((n, n)), np.zeros((n, n))
            rb.JK_direct(Jm, Km, B, D_conv, default_opts(
                precision=PREC_FP64 if p == "fp64" else PREC_MIXED, schwarz_thresh=1e-12))
            e_build[p] = float(np.sum(D_conv * (2.0 * Hc + 2.0 * Jm - Km)))
        scf["delta_e_fock_build"] = abs(e_build["mixed"] - e_build["fp64"])
        # mixed-precision threshold sweep: the build's energy error at the
        # converged density against the all-FP64 build, and its time on the
        # bench density
        sweep = []
        for thr in args.mixed_sweep:
            om = default_opts(precision=PREC_MIXED, schwarz_thresh=1e-12, mixed_thresh=thr)
            Jm, Km = np.zeros((n, n)), np.zeros((n, n))
            rb.JK_direct(Jm, Km, B, D_conv, om)
            de = abs(float(np.sum(D_conv * (2.0 * Hc + 2.0 * Jm - Km))) - e_build["fp64"])
            st_m = rb.jk_last_stats(B)

            def step_m():
                ctx.check(L.rchem_jk_device(ctx.h, ctypes.c_void_p(dD.data_ptr()),
                                            ctypes.c_void_p(dJ.data_ptr()),
                                            ctypes.c_void_p(dK.data_ptr()), ctypes.byref(om)))
            step_m()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step_m()
            torch.cuda.synchronize()
            ms_m = (time.perf_counter() - t0) / args.steps * 1e3
            sweep.append({"mixed_thresh": thr, "ms": ms_m,
                          "speedup_vs_fp64": runs["fp64"]["ms"] / ms_m if "fp64" in runs else None,
                          "fp32_quartet_fraction": st_m.quartets_fp32 / max(st_m.quartets, 1.0),
                          "delta_e_fock_build": de})
            say(f"mixed threshold {thr:g}: {ms_m:.2f} ms, |dE| {de:.2e} Eh")
        scf["mixed_sweep"] = sweep
        ok = [x for x in sweep if x["delta_e_fock_build"] < 1e-6]
        if ok:
            best = max(ok, key=lambda x: x["mixed_thresh"])
            scf["largest_threshold_within_1e-6"] = best["mixed_thresh"]
            scf["speedup_at_largest_threshold"] = best["speedup_vs_fp64"]
        scf["note"] = ("DIIS(8); mixed switches to all-FP64 J/K once |dE| < 1e-7 Eh; "
                       "convergence |dE| < 1e-10 Eh and max|FDS-SDF| < 1e-8; 'seconds' "
                       "replays the captured Fock build (hipGraph) on every iteration and "
                       "forms each iteration's density by canonical purification (DGEMM) "
                       "instead of the eigensolver")
        out["scf"] = scf
    if not args.no_cpu_baseline and args.workload == "h2o64":
        from rchem_amd.inputs import WATER_CRAWFORD, WATER_Z
        Bw = rb.Basis.new(WATER_Z, WATER_CRAWFORD, "STO-3G", ctx=ctx)
        ow = default_opts(boys_mode=1, schwarz_thresh=0.0)  # compat Boys, as the reference
        rb.scf(Bw, WATER_CRAWFORD, WATER_Z, nocc=5, opts=ow)
        t0 = time.perf_counter()
        rw, _ = rb.scf(Bw, WATER_CRAWFORD, WATER_Z, nocc=5, opts=ow)
        water_gpu = time.perf_counter() - t0
        out["cpu_baseline"] = cpu_baseline(Z, xyz, D, st.quartets, args.cpu_seconds, water_gpu)
        out["cpu_baseline"]["water_sto3g_scf"]["gpu_e_elec"] = rw.e_elec
        say("CPU baseline done")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
