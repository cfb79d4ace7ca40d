#!/usr/bin/env python3
"""Benchmark: Richards Newton-Krylov time stepping on one MI355X.

Metric: cell-timesteps/s on the 16.8M-cell heterogeneous workload
(512x512x64, dt = 0.5 h), solved with the mixed-precision preconditioner
(flexible GMRES(30), FP32 semicoarsening V-cycle with 2+2 red/black sweeps;
residual, J.v, Krylov basis and Gram-Schmidt in FP64).  The FP64 plain path
(the reference's algorithm: GMRES with modified Gram-Schmidt and an FP64
V-cycle) runs the same steps beside it; the JSON line reports its
throughput, the per-step pressure error of the mixed solve against it, and
the Newton/GMRES iteration counts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2]
    python bench.py --config c4 [--members 512] [--ensemble-stride 16]

With --config c4 it advances the hillslope ensemble with batched launches
and a sample of its members one at a time, and reports both rates.

Warmup runs W steps on a scratch copy of the state; the timed region is
then exactly the workload's first K steps from its initial condition, with
inputs resident in HBM and a device synchronisation on both sides.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _field_reads(gp):
    """Distinct FP64 input fields of the closures: alpha, n, K (aliased
    Kx/Ky/Kz count once), plus s_res/s_sat when they are per-cell fields."""
    d = gp.desc
    n_k = len({d.kx, d.ky, d.kz})
    n_vg = sum(1 for v in (d.sres, d.ssat) if v)
    return 2 + n_k + n_vg


def kernel_bytes_per_cell(gp):
    """Compulsory HBM bytes per interior cell of one launch of each measured
    kernel: every array it reads or writes counted once (neighbour values
    come from LDS or cache).
      residual (F form):   p, alpha, n, K, porosity, Ss, p_old, S_old read,
                           F written
      residual (J.v form): u and v instead of p, and F(u) read in addition
      residual (line-search form): u and the step instead of p, the trial
                           point written in addition
      Jacobian (PC matrix): p, alpha, n, K, porosity, Ss read; C, E, N, U
                           written in the multigrid's FP32
      smoother pair (finest level, FP32): C, E, N, U, b, x_in read, x_out
                           written (the V-cycle's last pair, which writes
                           the FP64 result instead, and its first, which
                           reads the FP64 right-hand side and stores its
                           FP32 copy, are not in the timed set)"""
    f = _field_reads(gp)
    res = 8 * (1 + f + 2 + 2 + 1)
    return {
        "residual": res,
        "residual_jv": res + 16,
        "residual_step": res + 16,
        "jacobian": 8 * (1 + f + 2) + 4 * 4,
        "rbgs_pair": 4 * 7,
    }


PROF_KERNELS = (
    # (key, api prof id name, kernel as rocprofv3 names it)
    ("residual", "PROF_RESIDUAL", "k_residual_pf<64, 8, 16, *, 0> (NlFunctionEval)"),
    ("residual_jv", "PROF_RESIDUAL_JV",
     "k_residual_pf<64, 8, 16, *, 1> (difference-quotient J.v)"),
    ("residual_step", "PROF_RESIDUAL_STEP",
     "k_residual_pf<64, 8, 16, *, 2> (line-search trial point)"),
    ("jacobian", "PROF_JACOBIAN", "k_jacobian_pf<64, 8, 16, *, true, float> (PC matrix)"),
    ("rbgs_pair", "PROF_RBGS",
     "k_rbgs_pair2<float, TY, KC, *, false> (finest-level red/black pair, FP32 result, "
     "autotuned tile)"),
)


def load_traffic():
    """HBM traffic per launch from the committed rocprofv3 --pmc summary:
    FETCH_SIZE and WRITE_SIZE of each kernel scaled by what the same counters
    read for a plain copy kernel of known size."""
    path = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if not os.path.exists(path):
        return {}
    d = json.load(open(path))
    cp = d["copy"]
    fscale = cp["bytes_read"] / (cp["fetch_kib"] * 1024.0)
    wscale = cp["bytes_written"] / (cp["write_kib"] * 1024.0)
    out = {}
    for key, v in d["kernels"].items():
        out[key] = v["fetch_kib"] * 1024.0 * fscale + v["write_kib"] * 1024.0 * wscale
    return out


def run_gpu(args):
    from parflow_amd import api
    from parflow_amd.decks import deck

    L = api.lib()
    t0 = time.time()
    prob, opts, p0, dt, nsteps_cfg, extras = deck(args.config)
    _log(f"deck {args.config}: {prob.nx}x{prob.ny}x{prob.nz} "
         f"({prob.ncells} cells) generated in {time.time() - t0:.1f}s")
    gp = api.GPUProblem(prob)
    ncell = prob.ncells
    K, W = args.steps, args.warmup

    def make(mode):
        if mode == "mixed":
            return api.GPUSolver(gp, opts, precision=32, krylov=1)
        return api.GPUSolver(gp, opts, precision=64, krylov=0)

    results = {}
    snaps = {}
    shapes = []
    for mode in ("mixed", "plain"):
        if mode == "plain" and args.no_plain:
            continue
        sv = make(mode)
        # warmup on a scratch state
        if W > 0:
            dp, ds = gp.vec(p0), gp.vec()
            sv.advance_device(dp, ds, 0.0, dt, W)
            del dp, ds
        dp, ds = gp.vec(p0), gp.vec()
        snap = [gp.vec() for _ in range(K)]
        if mode == "mixed":
            for _, pid, _ in PROF_KERNELS:
                api.check(L.pfhip_prof_enable(getattr(api, pid), 200000), "prof")
        newton, gmres, fevals = [], [], []
        t = 0.0
        api.check(L.pfhip_synchronize(), "sync")
        tic = time.perf_counter()
        for s in range(K):
            lg = sv.advance_device(dp, ds, t, dt, 1)
            t += float(lg["dt"][0])
            newton.append(int(lg["newton"][0]))
            gmres.append(int(lg["gmres"][0]))
            fevals.append(int(lg["fevals"][0]))
            api.check(L.pfhip_vec_linear_sum(gp.h, 1.0, dp.ptr, 0.0, dp.ptr,
                                             snap[s].ptr, None), "snapshot")
        api.check(L.pfhip_synchronize(), "sync")
        elapsed = time.perf_counter() - tic
        prof = None
        if mode == "mixed":
            prof = {}
            for key, pid, _ in PROF_KERNELS:
                ms = api.C.c_double()
                n = api.C.c_long()
                api.check(L.pfhip_prof_read(getattr(api, pid), api.C.byref(ms),
                                            api.C.byref(n)), "prof_read")
                api.check(L.pfhip_prof_disable(getattr(api, pid)), "prof_off")
                prof[key] = (ms.value, n.value)
        if mode == "mixed":
            # smoother tile shape the multigrid's autotuner chose per grid level
            # (at its first V-cycle)
            L.pfhip_solver_mg.restype = api.C.c_void_p
            L.pfhip_solver_mg.argtypes = [api.C.c_void_p]
            L.pfhip_mg_pair_shapes.argtypes = [api.C.c_void_p, api.C.c_void_p, api.C.c_int]
            arr = (api.C.c_int * 64)()
            nl = L.pfhip_mg_pair_shapes(L.pfhip_solver_mg(sv.h), arr, 64)
            shapes = list(arr[:nl])
            _log(f"smoother tile shape per grid level (autotuned): {shapes}")
        results[mode] = dict(elapsed=elapsed, newton=newton, gmres=gmres, fevals=fevals,
                             value=ncell * K / elapsed, prof=prof)
        snaps[mode] = snap
        _log(f"{mode}: {K} steps in {elapsed:.3f}s = {ncell * K / elapsed:.4e} cell-steps/s "
             f"newton={newton} gmres={gmres}")
        del sv

    errs = []
    if "plain" in snaps:
        for s in range(K):
            a = prob.interior(gp.get(snaps["mixed"][s]))
            b = prob.interior(gp.get(snaps["plain"][s]))
            errs.append(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))))
    results["mixed"]["pair_shapes"] = shapes
    return prob, gp, results, errs


def run_cpu_baseline(args, cores):
    """The oracle (C restatement, OpenMP over cells) on a bounded sample of
    the same workload: the deck's fields cropped to the first
    (nx_s x ny_s x nz) columns, one time step."""
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    from oracle import oracle as O
    from parflow_amd.decks import deck
    from parflow_amd.problem import Problem

    prob, opts, p0, dt, nsteps_cfg, extras = deck(args.config)
    n = args.cpu_sample
    I = lambda a: prob.interior(a)[:, :n, :n]
    sub = Problem(nx=n, ny=n, nz=prob.nz, dx=prob.dx, dy=prob.dy, dz=prob.dz,
                  x0=prob.x0, y0=prob.y0, z0=prob.z0, kx=I(prob.g_kx), poro=I(prob.g_poro),
                  ss=I(prob.g_ss), alpha=I(prob.g_alpha), nvg=I(prob.g_nvg),
                  sres=I(prob.g_sres), ssat=I(prob.g_ssat), bc=prob.bc)
    op = O.OracleProblem(sub)
    ps = sub.hydrostatic(0.0, prob.bc[4].z_ref)
    tic = time.perf_counter()
    r, ph, sh, log = O.advance(op, opts, ps, 0.0, dt, 1, keep_history=False)
    el = time.perf_counter() - tic
    return dict(value=sub.ncells / el, unit="cell-timesteps/s", cores=int(cores),
                kind="port",
                sample=f"oracle (C restatement, OpenMP) on the {n}x{n}x{prob.nz} "
                       f"corner of the {args.config} fields, 1 step of dt={dt} h, "
                       f"{el:.1f}s, newton={log.newton[0]}, status={r}")


def kernel_probe(args):
    """Each measured kernel once on the workload's fields, for hardware
    counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the residual in
    both forms, the PC-matrix assembly into the FP32 multigrid, one V-cycle,
    and a plain copy kernel (z = 1.0 x over the interior cells) that
    calibrates the counters."""
    import ctypes as C
    from parflow_amd import api
    from parflow_amd.decks import deck
    L = api.lib()
    prob, opts, p0, dt, _, _ = deck(args.config)
    gp = api.GPUProblem(prob)
    p = gp.vec(p0)
    pold, sold, F, Z, v = gp.vec(p0), gp.vec(), gp.vec(), gp.vec(), gp.vec()
    rng = np.random.default_rng(7)
    vh = np.zeros(prob.shape)
    vh[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    v = gp.vec(vh)
    api.check(L.pfhip_saturation(gp.h, pold.ptr, sold.ptr, 0, None), "S")
    gp.bc_pressure(dt)
    mg = api.GPUMultigrid(gp, opts.num_pre_relax, opts.num_post_relax, 0, 32)
    L.pfhip_residual_jv.argtypes = [C.c_void_p] * 3 + [C.c_double, C.c_void_p, C.c_double] + \
        [C.c_void_p] * 3 + [C.c_double, C.c_void_p]
    L.pfhip_mg_setup_jacobian.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    L.pfhip_vec_scale.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    for _ in range(2):   # the first round warms up; a counter run reads both
        api.check(L.pfhip_nl_function_eval(gp.h, p.ptr, F.ptr, pold.ptr, sold.ptr,
                                           C.c_double(dt), None), "F")
        api.check(L.pfhip_residual_jv(gp.h, p.ptr, v.ptr, 1e-6, F.ptr, 1e6, Z.ptr, pold.ptr,
                                      sold.ptr, dt, None), "Jv")
        api.check(L.pfhip_mg_setup_jacobian(mg.h, p.ptr, dt, None), "PC")
        api.check(L.pfhip_mg_apply(mg.h, F.ptr, Z.ptr, None), "V-cycle")
        api.check(L.pfhip_vec_scale(gp.h, 1.0, F.ptr, Z.ptr, None), "copy")
    api.check(L.pfhip_synchronize(), "sync")
    print(json.dumps({"kernel_probe": args.config, "cells": prob.ncells,
                      "copy_bytes_read": 8 * prob.ncells,
                      "copy_bytes_written": 8 * prob.ncells}), flush=True)


def run_ensemble(args):
    """The hillslope ensemble (deck c4): every member advanced together with
    batched launches, then a sample of the same members advanced one at a
    time on the single-problem path; prints one JSON line."""
    from parflow_amd import api
    from parflow_amd.decks import c4_members

    L = api.lib()
    K = args.steps if args.steps is not None else 24
    W = args.warmup
    tic = time.perf_counter()
    members = c4_members(args.members)
    ncells = sum(p.ncells for p, _, _, _ in members)
    _log(f"c4: {len(members)} members, {ncells} cells, generated in "
         f"{time.perf_counter() - tic:.1f}s")
    tic = time.perf_counter()
    ens = api.GPUEnsemble([(p, o) for p, o, _, _ in members])
    _log(f"ensemble created in {time.perf_counter() - tic:.1f}s")
    dts = [m[3] for m in members]
    if W > 0:
        dps = [ens.problems[i].vec(m[2]) for i, m in enumerate(members)]
        dss = [ens.problems[i].vec() for i in range(len(members))]
        ens.advance_device(dps, dss, 0.0, dts, W)
        del dps, dss
    dps = [ens.problems[i].vec(m[2]) for i, m in enumerate(members)]
    dss = [ens.problems[i].vec() for i in range(len(members))]
    api.check(L.pfhip_synchronize(), "sync")
    tic = time.perf_counter()
    lg = ens.advance_device(dps, dss, 0.0, dts, K)
    api.check(L.pfhip_synchronize(), "sync")
    el = time.perf_counter() - tic
    rounds, launches = ens.stats()
    value = ncells * K / el
    _log(f"ensemble: {K} steps in {el:.3f}s = {value:.4e} cell-steps/s, {rounds} rounds, "
         f"{launches} launches, newton/step {lg['newton'].mean():.2f}")
    # one at a time: every stride-th member on the single-problem mixed path
    sample = members[::args.ensemble_stride]
    solvers = [(api.GPUProblem(p), o, p0, dt) for p, o, p0, dt in sample]
    solvers = [(gp, api.GPUSolver(gp, o, precision=32, krylov=1), p0, dt)
               for gp, o, p0, dt in solvers]
    scells = sum(gp.prob.ncells for gp, _, _, _ in solvers)
    seq = 0.0
    snewton = []
    for gp, sv, p0, dt in solvers:
        dp, ds = gp.vec(p0), gp.vec()
        api.check(L.pfhip_synchronize(), "sync")
        t1 = time.perf_counter()
        lg1 = sv.advance_device(dp, ds, 0.0, dt, K)
        api.check(L.pfhip_synchronize(), "sync")
        seq += time.perf_counter() - t1
        snewton.append(float(np.mean(lg1["newton"])))
    seq_value = scells * K / seq
    _log(f"one at a time: {len(solvers)} members, {K} steps in {seq:.3f}s = "
         f"{seq_value:.4e} cell-steps/s")
    line = {
        "metric": "cell-timesteps/s (512-member hillslope ensemble, batched)",
        "value": value,
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * el / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 preconditioner)",
        "data": "synthetic (seeded per-member hillslopes)",
        "config": {
            "workload": f"c4: {len(members)} hillslopes, nx, ny in [8, 64], nz in [20, 50], "
                        "FGMRES(30) + FP32 V-cycle (2+2 RB sweeps), dt=0.25 h",
            "members": len(members),
            "cells": ncells,
            "newton_per_member_step": float(lg["newton"].mean()),
            "gmres_per_member_step": float(lg["gmres"].mean()),
        },
        "ensemble": {
            "rounds": rounds,
            "launches": launches,
            "one_at_a_time": {
                "value": seq_value,
                "sample": f"every {args.ensemble_stride}th member ({len(solvers)} members, "
                          f"{scells} cells), single-problem mixed path, {K} steps",
                "newton_per_member_step": float(np.mean(snewton)),
            },
            "speedup": value / seq_value,
        },
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for c2, 24 for c4)")
    ap.add_argument("--members", type=int, default=512, help="c4 ensemble size")
    ap.add_argument("--ensemble-stride", type=int, default=16,
                    help="c4: every stride-th member runs one at a time for the comparison")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=384)
    ap.add_argument("--kernel-probe", action="store_true",
                    help="run each measured kernel once (for counter runs) and exit")
    args = ap.parse_args()
    if args.kernel_probe:
        kernel_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU domain decomposition is design-only).", file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        run_ensemble(args)
        return
    if args.steps is None:
        args.steps = 20

    prob, gp, res, errs = run_gpu(args)
    mixed = res["mixed"]
    K = args.steps
    bpc_all = kernel_bytes_per_cell(gp)
    traffic = load_traffic()
    kernels = []
    for key, _, kname in PROF_KERNELS:
        ms_tot, nlaunch = mixed["prof"][key]
        if nlaunch == 0:
            continue
        avg_s = (ms_tot / 1e3) / nlaunch
        bpl = bpc_all[key] * prob.ncells
        ach = bpl / avg_s / 1e9
        kernels.append({"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "traffic": traffic.get(key), "bytes_per_cell": bpc_all[key],
                        "avg_launch_us": avg_s * 1e6, "launches": nlaunch,
                        "total_ms": ms_tot})
    # the dominant kernel: most total time in the timed region
    dom = max(kernels, key=lambda r: r["total_ms"])
    line = {
        "metric": "cell-timesteps/s (16.8M cells), mixed-precond vs FP64 plain at 1e-8 rel err",
        "value": mixed["value"],
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["elapsed"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 preconditioner)",
        "data": "synthetic (seeded FFT Gaussian random field, seeded van Genuchten fields)",
        "config": {
            "workload": f"{args.config}: {prob.nx}x{prob.ny}x{prob.nz} heterogeneous Richards, "
                        "FGMRES(30) + FP32 V-cycle (2+2 RB sweeps), dt=0.5 h",
            "cells": prob.ncells,
            "newton_per_step": mixed["newton"],
            "gmres_per_step": mixed["gmres"],
            "fevals_per_step": mixed["fevals"],
            "mg_pair_shape_per_level": mixed["pair_shapes"],
        },
        "roofline": {k: dom[k] for k in ("bound", "achieved", "peak", "unit", "frac",
                                          "traffic", "kernel", "bytes_per_cell",
                                          "avg_launch_us", "launches")},
        "roofline_kernels": kernels,
    }
    if "plain" in res:
        plain = res["plain"]
        line["plain_fp64"] = {
            "value": plain["value"],
            "ms_per_step": 1e3 * plain["elapsed"] / K,
            "newton_per_step": plain["newton"],
            "gmres_per_step": plain["gmres"],
            "mixed_over_plain": mixed["value"] / plain["value"],
            "max_rel_err_vs_plain_per_step": errs,
        }
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = run_cpu_baseline(args, cores)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
