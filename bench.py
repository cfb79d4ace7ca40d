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
                           written; the V-cycle's last pair writes the FP64
                           result instead (4 more bytes) and, where the
                           difference-quotient J.v follows (no overland),
                           reads the Newton iterate u for the J.v's sums (8
                           more); its first (zero initial guess: no x_in)
                           reads C, E, N, U and the FP64 right-hand side and
                           writes the level's FP32 copy of it and x_out"""
    f = _field_reads(gp)
    res = 8 * (1 + f + 2 + 2 + 1)
    return {
        "residual": res,
        "residual_jv": res + 16,
        "residual_step": res + 16,
        "jacobian": 8 * (1 + f + 2) + 4 * 4,
        "rbgs_pair": 4 * 7,
        "rbgs_pair_dout": 4 * 6 + 8 + (0 if gp.prob.overland else 8),
        "rbgs_pair_din": 4 * 4 + 8 + 4 + 4,
    }


PROF_KERNELS = (
    # (key, api prof id name, kernel as rocprofv3 names it)
    ("residual", "PROF_RESIDUAL", "k_residual_pf<64, 8, 16, *, 0> (NlFunctionEval)"),
    ("residual_jv", "PROF_RESIDUAL_JV",
     "k_residual_pf<64, 8, 16, *, 1> (difference-quotient J.v)"),
    ("residual_step", "PROF_RESIDUAL_STEP",
     "k_residual_pf<64, 8, 16, *, 2> (line-search trial point)"),
    ("jacobian", "PROF_JACOBIAN", "k_jacobian_pf<64, 8, 16, *, true, float> (PC matrix)"),
    ("rbgs_pair", "PROF_RBGS", "k_rbgs_pair2<float, ...> FP32 operands"),
    ("rbgs_pair_dout", "PROF_RBGS_DOUT", "k_rbgs_pair2<float, ...> FP64 result"),
    ("rbgs_pair_din", "PROF_RBGS_DIN", "k_rbgs_pair2<float, ...> FP64 right-hand side"),
)
# kernels reported together as one roofline row: the finest-level smoother
# pairs in all three operand forms
FAMILIES = {
    "rbgs_pair_family": (("rbgs_pair", "rbgs_pair_dout", "rbgs_pair_din"),
                         "k_rbgs_pair2<float, TY, KC, ...> (finest-level red/black pairs: "
                         "FP32 operands, FP64 result, FP64 right-hand side; autotuned tile)"),
}

# solver phases timed in a separate instrumented pass, named after the
# reference's timers (PL/timing.c, the modules' own timers)
PHASES = (
    ("NlFunctionEval", "PROF_PH_NL_F_EVAL"),
    ("KINSpgmrAtimes (J.v)", "PROF_PH_JV"),
    ("KinsolPC setup (J + multigrid)", "PROF_PH_PC_SETUP"),
    ("PFMG (V-cycle)", "PROF_PH_PC_SOLVE"),
    ("Gram-Schmidt", "PROF_PH_ORTHO"),
    ("KINSol", "PROF_PH_KINSOL"),
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


def _workload_label(name, prob, opts, dt, krylov_label, vcycle="FP32"):
    kind = ("tilted-V catchment with kinematic overland flow" if prob.overland
            else "heterogeneous Richards")
    return (f"{name}: {prob.nx}x{prob.ny}x{prob.nz} {kind}, {krylov_label}({opts.krylov_dim}) "
            f"+ {vcycle} V-cycle ({opts.num_pre_relax}+{opts.num_post_relax} RB sweeps), "
            f"dt={dt} h")


def _timed_run(api, gp, sv, p0, dt, K, W, snapshot=False):
    """W untimed warmup steps on a scratch state, then exactly K steps from
    p0 between two device synchronisations, in one pfhip_advance_richards
    call (the time loop as a caller runs it).  With snapshot, the K steps
    are run one call at a time instead and the pressure after every step is
    copied into a device snapshot: for the untimed passes that supply the
    per-step comparisons (the solver is deterministic, so their states are
    the timed run's)."""
    L = api.lib()
    if W > 0:
        dp, ds = gp.vec(p0), gp.vec()
        sv.advance_device(dp, ds, 0.0, dt, W)
        del dp, ds
    dp, ds = gp.vec(p0), gp.vec()
    snap = [gp.vec() for _ in range(K)] if snapshot else None
    newton, gmres, fevals = [], [], []
    api.check(L.pfhip_synchronize(), "sync")
    tic = time.perf_counter()
    if snapshot:
        t = 0.0
        for s in range(K):
            lg = sv.advance_device(dp, ds, t, dt, 1)
            t += float(lg["dt"][0])
            newton.append(int(lg["newton"][0]))
            gmres.append(int(lg["gmres"][0]))
            fevals.append(int(lg["fevals"][0]))
            api.check(L.pfhip_vec_linear_sum(gp.h, 1.0, dp.ptr, 0.0, dp.ptr, snap[s].ptr,
                                             None), "snapshot")
    else:
        lg = sv.advance_device(dp, ds, 0.0, dt, K)
        newton = [int(x) for x in lg["newton"]]
        gmres = [int(x) for x in lg["gmres"]]
        fevals = [int(x) for x in lg["fevals"]]
    api.check(L.pfhip_synchronize(), "sync")
    elapsed = time.perf_counter() - tic
    return dict(elapsed=elapsed, newton=newton, gmres=gmres, fevals=fevals,
                value=gp.prob.ncells * K / elapsed), snap


def _read_prof(api, pid_names):
    L = api.lib()
    out = {}
    for key, pid in pid_names:
        ms, n = api.C.c_double(), api.C.c_long()
        api.check(L.pfhip_prof_read(getattr(api, pid), api.C.byref(ms), api.C.byref(n)),
                  "prof_read")
        api.check(L.pfhip_prof_disable(getattr(api, pid)), "prof_off")
        out[key] = (ms.value, n.value)
    return out


def _phase_pass(api, gp, sv, p0, dt, K):
    """The same K steps again with the phase timers on (not the timed run):
    device milliseconds per step of each phase."""
    for _, pid in PHASES:
        api.check(api.lib().pfhip_prof_enable(getattr(api, pid), 20000), "prof")
    _timed_run(api, gp, sv, p0, dt, K, 0)
    prof = _read_prof(api, [(name, pid) for name, pid in PHASES])
    return {name: {"ms_per_step": ms / K, "calls": n} for name, (ms, n) in prof.items()}


def _water_balance(gp, prob, snaps, p0, dt):
    """Per-step |change of stored water - budget| / storage change of a run's
    pressures (device snapshots): stored water as the residual's accumulation
    terms book it, differenced cell by cell, the budget from the device's
    boundary, overland-top, runoff and source terms.  The storage change is
    the sum of the cells' absolute changes: the net change under rain or a
    top flux, and the volume moved between cells in a recession step whose
    net change is zero to round-off (the tests' measure)."""
    I = prob.interior
    vol = prob.dx * prob.dy * prob.dz
    poro, ss = I(prob.g_poro), I(prob.g_ss)
    p_prev = p0
    s_prev = gp.saturation(p0)
    out = []
    for k, d in enumerate(snaps):
        p1 = gp.get(d)
        s1 = gp.saturation(p1)
        dc = (poro * (I(s1) - I(s_prev)) * prob.rho * vol
              + ss * vol * (I(p1) * I(s1) - I(p_prev) * I(s_prev)) * prob.rho)
        dm, gross = float(np.sum(dc)), float(np.sum(np.abs(dc)))
        if prob.overland:
            ds = prob.dx * prob.dy * (np.maximum(I(p1)[-1], 0.0) - np.maximum(I(p_prev)[-1], 0.0))
            dm += float(np.sum(ds))
            gross += float(np.sum(np.abs(ds)))
        f = gp.water_balance_fluxes(p1, p_prev, dt, (k + 1) * dt)
        out.append(abs(dm - (f[0] + f[1] - f[2] + f[3])) / gross)
        p_prev, s_prev = p1, s1
    return out


def run_gpu(args):
    import dataclasses
    from parflow_amd import api
    from parflow_amd.decks import deck

    L = api.lib()
    t0 = time.time()
    prob, opts, p0, dt, nsteps_cfg, extras = deck(args.config)
    _log(f"deck {args.config}: {prob.nx}x{prob.ny}x{prob.nz} "
         f"({prob.ncells} cells) generated in {time.time() - t0:.1f}s")
    gp = api.GPUProblem(prob)
    K, W = args.steps, args.warmup
    res = {"opts": opts, "dt": dt}

    # mixed: the timed run (no timers), then the same steps with the
    # per-launch kernel timers on (their events bracket each launch, and the
    # V-cycle then launches directly instead of replaying its graph)
    sv = api.GPUSolver(gp, opts, precision=32, krylov=1)
    mixed, _ = _timed_run(api, gp, sv, p0, dt, K, W)
    for _, pid, _ in PROF_KERNELS:
        api.check(L.pfhip_prof_enable(getattr(api, pid), 20000), "prof")
    # the kernel-timer pass also takes the per-step pressure snapshots
    timer_pass, snap_m = _timed_run(api, gp, sv, p0, dt, K, 0, snapshot=True)
    assert timer_pass["newton"] == mixed["newton"], "mixed passes diverged"
    mixed["prof"] = _read_prof(api, [(k, pid) for k, pid, _ in PROF_KERNELS])
    L.pfhip_solver_mg.restype = api.C.c_void_p
    L.pfhip_solver_mg.argtypes = [api.C.c_void_p]
    L.pfhip_mg_pair_shapes.argtypes = [api.C.c_void_p, api.C.c_void_p, api.C.c_int]
    arr = (api.C.c_int * 64)()
    nl = L.pfhip_mg_pair_shapes(L.pfhip_solver_mg(sv.h), arr, 64)
    mixed["pair_shapes"] = list(arr[:nl])
    din, dout = api.C.c_int(), api.C.c_int()
    L.pfhip_mg_pair_shapes_fp64.argtypes = [api.C.c_void_p, api.C.c_void_p, api.C.c_void_p]
    api.check(L.pfhip_mg_pair_shapes_fp64(L.pfhip_solver_mg(sv.h), api.C.byref(din),
                                          api.C.byref(dout)), "pair_shapes_fp64")
    mixed["pair_shapes_fp64"] = [din.value, dout.value]
    _log(f"smoother tile shape per grid level (autotuned): {mixed['pair_shapes']}")
    _log(f"mixed: {K} steps in {mixed['elapsed']:.3f}s = {mixed['value']:.4e} cell-steps/s "
         f"newton={mixed['newton']} gmres={mixed['gmres']}")
    if not args.no_phases:
        mixed["phases"] = _phase_pass(api, gp, sv, p0, dt, K)
    del sv
    res["mixed"] = mixed
    if not args.no_plain:
        # plain: the reference's algorithm in FP64 with the deck's sweeps
        sv = api.GPUSolver(gp, opts, precision=64, krylov=0)
        plain, _ = _timed_run(api, gp, sv, p0, dt, K, W)
        _log(f"plain: {K} steps in {plain['elapsed']:.3f}s = {plain['value']:.4e} "
             f"cell-steps/s newton={plain['newton']} gmres={plain['gmres']}")
        # an untimed pass for the per-step snapshots
        again, snap_p = _timed_run(api, gp, sv, p0, dt, K, 0, snapshot=True)
        assert again["newton"] == plain["newton"], "plain passes diverged"
        if not args.no_phases:
            plain["phases"] = _phase_pass(api, gp, sv, p0, dt, K)
        del sv
        plain["err"] = []
        for s in range(K):
            a = prob.interior(gp.get(snap_m[s]))
            b = prob.interior(gp.get(snap_p[s]))
            plain["err"].append(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))))
        res["plain"] = plain
        # plain at the reference's PFMG settings: 1 pre- and 1 post-sweep
        # (PL/pf_pfmg.c:294-297 defaults)
        ref_opts = dataclasses.replace(opts, num_pre_relax=1, num_post_relax=1)
        sv = api.GPUSolver(gp, ref_opts, precision=64, krylov=0)
        pref, _ = _timed_run(api, gp, sv, p0, dt, K, W)
        _log(f"plain (1+1 sweeps): {K} steps in {pref['elapsed']:.3f}s = "
             f"{pref['value']:.4e} cell-steps/s newton={pref['newton']} gmres={pref['gmres']}")
        del sv
        res["plain_ref"] = pref
        # the device plain state after step 10: the CPU baseline's start
        ks = min(10, K - 1)
        res["cpu_start"] = (ks, gp.get(snap_p[ks - 1]) if ks > 0 else p0,
                            gp.get(snap_p[ks]), plain["newton"][ks])
        del snap_p
    if not args.no_water_balance:
        res["water_balance"] = _water_balance(gp, prob, snap_m, p0, dt)
        _log(f"water balance of the mixed run (ResidualTol {opts.residual_tol}): "
             f"worst {max(res['water_balance']):.2e}")
    del snap_m
    return prob, gp, res


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(args, cores, prob, opts, dt, start):
    """The oracle (C restatement, OpenMP over cells) on a bounded sample of
    the same workload: one time step of the full fields, step ks + 1 taken
    from the plain device path's state after step ks (a step with the
    run's typical Newton count; step 1 needs the most).  Its result is
    checked against the device's step ks + 1 as well."""
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    from oracle import oracle as O
    ks, p_start, p_dev_next, newton_dev = start
    op = O.OracleProblem(prob)
    tic = time.perf_counter()
    r, ph, sh, log = O.advance(op, opts, p_start, ks * dt, dt, 1, keep_history=False)
    el = time.perf_counter() - tic
    a, b = prob.interior(ph), prob.interior(p_dev_next)
    err = float(np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-3)))
    return dict(value=prob.ncells / el, unit="cell-timesteps/s", cores=int(cores),
                kind="port", cpu=_cpu_model(),
                sample=f"oracle (C restatement, OpenMP) on the full {prob.nx}x{prob.ny}x"
                       f"{prob.nz} {args.config} fields: step {ks + 1} (dt={dt} h) from the "
                       f"plain device state after step {ks}, {el:.1f} s, newton="
                       f"{log.newton[0]} (device {newton_dev}), status={r}, device vs oracle "
                       f"max rel p diff {err:.1e}",
                newton=int(log.newton[0]), seconds=el, device_rel_diff=err)


def kernel_probe(args):
    """Each measured kernel once on the workload's fields, for hardware
    counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the residual in
    its three forms, the PC-matrix assembly into the FP32 multigrid, one V-cycle,
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
    L.pfhip_residual_step.argtypes = [C.c_void_p] * 3 + [C.c_double] + [C.c_void_p] * 4 + \
        [C.c_double, C.c_void_p]
    unew = gp.vec()
    L.pfhip_vec_scale.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    for _ in range(2):   # the first round warms up; a counter run reads both
        api.check(L.pfhip_nl_function_eval(gp.h, p.ptr, F.ptr, pold.ptr, sold.ptr,
                                           C.c_double(dt), None), "F")
        api.check(L.pfhip_residual_jv(gp.h, p.ptr, v.ptr, 1e-6, F.ptr, 1e6, Z.ptr, pold.ptr,
                                      sold.ptr, dt, None), "Jv")
        api.check(L.pfhip_residual_step(gp.h, p.ptr, v.ptr, 1e-3, unew.ptr, Z.ptr, pold.ptr,
                                        sold.ptr, dt, None), "line-search point")
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
    time on the single-problem path; returns the JSON line's fields."""
    from parflow_amd import api
    from parflow_amd.decks import c4_members

    L = api.lib()
    K = args.ens_steps if args.config != "c4" else (args.steps if args.steps is not None else 24)
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
    return line


def kernel_rows(prob, gp, prof):
    """Roofline rows of the timed steps (repeated with the per-launch timers
    on): achieved = algorithmic bytes of a launch / its average HIP-event
    duration; one row per kernel, and one per family (its forms' bytes and
    times summed)."""
    bpc_all = kernel_bytes_per_cell(gp)
    traffic = load_traffic()
    rows = {}
    for key, _, kname in PROF_KERNELS:
        ms_tot, nlaunch = prof[key]
        if nlaunch == 0:
            continue
        avg_s = (ms_tot / 1e3) / nlaunch
        bpl = bpc_all[key] * prob.ncells
        ach = bpl / avg_s / 1e9
        rows[key] = {"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic.get(key),
                     "bytes_per_cell": bpc_all[key], "avg_launch_us": avg_s * 1e6,
                     "launches": nlaunch, "total_ms": ms_tot}
    out = []
    members = set()
    for fam, (keys, fname) in FAMILIES.items():
        parts = [rows[k] for k in keys if k in rows]
        if not parts:
            continue
        members.update(keys)
        n = sum(r["launches"] for r in parts)
        ms = sum(r["total_ms"] for r in parts)
        nbytes = sum(r["bytes_per_cell"] * prob.ncells * r["launches"] for r in parts)
        avg_s = ms / 1e3 / n
        ach = nbytes / n / avg_s / 1e9
        tr = ([r["traffic"] for r in parts] if all(r["traffic"] for r in parts) else None)
        out.append({"kernel": fname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                    "traffic": (sum(t * r["launches"] for t, r in zip(tr, parts)) / n
                                if tr else None),
                    "bytes_per_cell": nbytes / n / prob.ncells, "avg_launch_us": avg_s * 1e6,
                    "launches": n, "total_ms": ms,
                    "forms": {k: {f: rows[k][f] for f in ("launches", "avg_launch_us", "frac",
                                                          "bytes_per_cell", "traffic")}
                              for k in keys if k in rows}})
    out += [r for k, r in rows.items() if k not in members]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for c2, 24 for c4)")
    ap.add_argument("--members", type=int, default=512, help="c4 ensemble size")
    ap.add_argument("--ensemble-stride", type=int, default=16,
                    help="c4: every stride-th member runs one at a time for the comparison")
    ap.add_argument("--ens-steps", type=int, default=24,
                    help="steps of the ensemble extra of the c1-c3 line")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-phases", action="store_true")
    ap.add_argument("--no-water-balance", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
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
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    if args.steps is None:
        args.steps = 20

    prob, gp, res = run_gpu(args)
    mixed, opts, dt = res["mixed"], res["opts"], res["dt"]
    K = args.steps
    kernels = kernel_rows(prob, gp, mixed["prof"])
    # the dominant kernel: most total time over the timed steps
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
        "data": "synthetic (seeded FFT Gaussian random field, seeded van Genuchten fields)"
                if not prob.overland else "synthetic (seeded tilted-V slopes)",
        "config": {
            "workload": _workload_label(args.config, prob, opts, dt, "FGMRES"),
            "cells": prob.ncells,
            "newton_per_step": mixed["newton"],
            "gmres_per_step": mixed["gmres"],
            "fevals_per_step": mixed["fevals"],
            "mg_pair_shape_per_level": mixed["pair_shapes"],
            "mg_pair_shape_level0_fp64_in_out": mixed["pair_shapes_fp64"],
        },
        "roofline": {k: dom[k] for k in ("bound", "achieved", "peak", "unit", "frac",
                                          "traffic", "kernel", "bytes_per_cell",
                                          "avg_launch_us", "launches")},
        "roofline_kernels": kernels,
    }
    if "phases" in mixed:
        line["phases_ms_per_step"] = mixed["phases"]
    if "water_balance" in res:
        line["water_balance"] = {
            "residual_over_storage_change_per_step": res["water_balance"],
            "residual_tol": opts.residual_tol,
            "note": "mixed run at the deck's ResidualTol; the 1e-10 bar is tested at "
                    "ResidualTol 1e-12 (tests/test_gpu_whole_configs.py)"}
    if "plain" in res:
        plain, pref = res["plain"], res["plain_ref"]
        line["plain_fp64"] = {
            "value": plain["value"],
            "workload": _workload_label(args.config, prob, opts, dt, "GMRES", "FP64"),
            "ms_per_step": 1e3 * plain["elapsed"] / K,
            "newton_per_step": plain["newton"],
            "gmres_per_step": plain["gmres"],
            "mixed_over_plain": mixed["value"] / plain["value"],
            "max_rel_err_vs_plain_per_step": plain["err"],
        }
        if "phases" in plain:
            line["plain_fp64"]["phases_ms_per_step"] = plain["phases"]
        line["plain_fp64_reference_pfmg_settings"] = {
            "value": pref["value"],
            "workload": _workload_label(args.config, prob, dataclasses_replace(
                opts, num_pre_relax=1, num_post_relax=1), dt, "GMRES", "FP64"),
            "ms_per_step": 1e3 * pref["elapsed"] / K,
            "newton_per_step": pref["newton"],
            "gmres_per_step": pref["gmres"],
            "mixed_over_this": mixed["value"] / pref["value"],
        }
    if not args.no_cpu and "cpu_start" in res:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = run_cpu_baseline(args, cores, prob, opts, dt, res["cpu_start"])
    if not args.no_ensemble:
        del gp
        e = run_ensemble(args)
        line["ensemble"] = {"metric": e["metric"], "value": e["value"], "steps": e["steps"],
                            "workload": e["config"]["workload"],
                            "newton_per_member_step": e["config"]["newton_per_member_step"],
                            "gmres_per_member_step": e["config"]["gmres_per_member_step"],
                            **e["ensemble"]}
    print(json.dumps(line), flush=True)


def dataclasses_replace(obj, **kw):
    import dataclasses
    return dataclasses.replace(obj, **kw)


if __name__ == "__main__":
    main()
