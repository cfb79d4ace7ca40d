"""ctypes binding of the oracle (CPU restatement of ParFlow's Richards
Newton-Krylov path).  TEST INFRASTRUCTURE ONLY: used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker,
never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MAXI = 8


class BCSpecC(C.Structure):
    _fields_ = [("type", C.c_int), ("n_intervals", C.c_int), ("repeat", C.c_int),
                ("lengths", C.c_int * MAXI), ("values", C.c_double * MAXI),
                ("z_ref", C.c_double)]


_dp = C.POINTER(C.c_double)


class ProblemC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("rho", C.c_double), ("mu", C.c_double), ("gravity", C.c_double),
                ("kx", _dp), ("ky", _dp), ("kz", _dp), ("poro", _dp), ("ss", _dp),
                ("alpha", _dp), ("nvg", _dp), ("sres", _dp), ("ssat", _dp),
                ("src", _dp),
                ("slope_x", _dp), ("slope_y", _dp), ("mann", _dp),
                ("ov_eps", C.c_double),
                ("bc", BCSpecC * 6),
                ("base_unit", C.c_double), ("start_time", C.c_double),
                ("bc_val", _dp * 6),
                ("vgt", _dp), ("vgt_points", C.c_int), ("vgt_method", C.c_int),
                ("vgt_min_head", C.c_double)]


class OptsC(C.Structure):
    _fields_ = [("residual_tol", C.c_double), ("step_tol", C.c_double),
                ("max_iter", C.c_int), ("krylov_dim", C.c_int),
                ("max_restarts", C.c_int), ("eta_choice", C.c_int),
                ("eta_value", C.c_double), ("eta_alpha", C.c_double),
                ("eta_gamma", C.c_double), ("use_jacobian", C.c_int),
                ("deriv_eps", C.c_double), ("globalization", C.c_int),
                ("precond", C.c_int), ("mg_pre", C.c_int), ("mg_post", C.c_int),
                ("mg_symmetric", C.c_int), ("mg_max_levels", C.c_int),
                ("pc_matrix", C.c_int)]


class KinStatsC(C.Structure):
    _fields_ = [("ret", C.c_int), ("nni", C.c_int), ("nli", C.c_int),
                ("nfe", C.c_int), ("npe", C.c_int), ("nps", C.c_int),
                ("nbcf", C.c_int), ("nbktrk", C.c_int), ("ncfl", C.c_int),
                ("fnorm", C.c_double)]


class RunLogC(C.Structure):
    _fields_ = [("nsteps_done", C.c_int), ("newton", C.c_int * 4096),
                ("gmres", C.c_int * 4096), ("fevals", C.c_int * 4096),
                ("dt_used", C.c_double * 4096)]


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.pfo_kinsol.restype = C.c_int
        _lib.pfo_vg_table_lookup.restype = C.c_double
        _lib.pfo_advance.restype = C.c_int
        for f in ("pfo_dot", "pfo_l2norm", "pfo_maxnorm", "pfo_l1norm"):
            getattr(_lib, f).restype = C.c_double
        _lib.pfo_mg_create.restype = C.c_void_p
        _lib.pfo_mg_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int, C.c_int]
        _lib.pfo_mg_create_full.restype = C.c_void_p
        _lib.pfo_mg_create_full.argtypes = _lib.pfo_mg_create.argtypes + [C.c_int]
        _lib.pfo_mg_level_matrix.argtypes = [C.c_void_p, C.c_int, C.POINTER(_dp)]
        _lib.pfo_mg_setup.argtypes = [C.c_void_p, C.POINTER(_dp)]
        _lib.pfo_mg_apply.argtypes = [C.c_void_p, _dp, _dp]
        _lib.pfo_mg_destroy.argtypes = [C.c_void_p]
        _lib.pfo_mg_num_levels.argtypes = [C.c_void_p]
        _lib.pfo_mg_level_dims.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class OracleProblem:
    """Holds a ProblemC view of a parflow_amd.problem.Problem."""

    def __init__(self, prob, src=None):
        self.prob = prob
        self.keep = []
        c = ProblemC()
        c.nx, c.ny, c.nz = prob.nx, prob.ny, prob.nz
        c.dx, c.dy, c.dz = prob.dx, prob.dy, prob.dz
        c.x0, c.y0, c.z0 = prob.x0, prob.y0, prob.z0
        c.rho, c.mu, c.gravity = prob.rho, prob.mu, prob.gravity
        for name in ("kx", "ky", "kz", "poro", "ss", "alpha", "nvg", "sres",
                     "ssat", "slope_x", "slope_y", "mann"):
            a = np.ascontiguousarray(getattr(prob, "g_" + name))
            self.keep.append(a)
            setattr(c, name, _ptr(a))
        self.src = None
        if src is not None:
            s = np.ascontiguousarray(src, dtype=np.float64)
            self.keep.append(s)
            self.src = s
            c.src = _ptr(s)
        c.ov_eps = prob.ov_eps
        for f, b in enumerate(prob.bc):
            cb = c.bc[f]
            cb.type = b.type
            cb.n_intervals = len(b.values)
            cb.repeat = b.repeat
            for i, (L, v) in enumerate(zip(b.lengths, b.values)):
                cb.lengths[i] = int(L)
                cb.values[i] = float(v)
            cb.z_ref = b.z_ref
            buf = np.zeros(prob.face_ncells(f))
            self.keep.append(buf)
            c.bc_val[f] = _ptr(buf)
        c.base_unit = prob.base_unit
        c.start_time = prob.start_time
        tp = prob.vg_table_params
        self.vg_table = None
        if tp is not None:
            method, pts, head, alpha, n = tp
            self.vg_table = vg_table_build(method, pts, head, alpha, n)
            c.vgt = _ptr(self.vg_table)
            c.vgt_points, c.vgt_method, c.vgt_min_head = pts, method, head
        self.c = c
        self.bc_bufs = [k for k in self.keep[-6:]]

    @property
    def ref(self):
        return C.byref(self.c)

    def set_time(self, t):
        lib().pfo_bc_pressure(self.ref, C.c_double(t))


def vg_table_build(method, points, min_head, alpha, n):
    """pfo_vg_table_build: 7 x (points+1) array x | a | d | a_der | d_der |
    slope | slope_der."""
    tab = np.zeros(7 * (points + 1))
    lib().pfo_vg_table_build(C.c_int(method), C.c_int(points), C.c_double(min_head),
                             C.c_double(alpha), C.c_double(n), _ptr(tab))
    return tab


def vg_table_lookup(tab, method, points, min_head, head, fcn=0):
    return lib().pfo_vg_table_lookup(_ptr(tab), C.c_int(method), C.c_int(points),
                                     C.c_double(min_head), C.c_double(head), C.c_int(fcn))


def opts_c(o):
    pre, post, sym = o.mg_params()
    c = OptsC()
    c.residual_tol = o.residual_tol
    c.step_tol = o.step_tol
    c.max_iter = o.max_iter
    c.krylov_dim = o.krylov_dim
    c.max_restarts = o.max_restarts
    c.eta_choice = o.eta_choice
    c.eta_value = o.eta_value
    c.eta_alpha = o.eta_alpha
    c.eta_gamma = o.eta_gamma
    c.use_jacobian = int(o.use_jacobian)
    c.deriv_eps = o.deriv_eps
    c.globalization = o.globalization
    c.precond = 0 if o.precond == "NoPC" else 1
    c.mg_pre, c.mg_post, c.mg_symmetric = pre, post, sym
    c.mg_max_levels = o.max_levels
    c.pc_matrix = o.pc_matrix_code()
    return c


# ---------------------------------------------------------------------------
def saturation(op, p, fcn=0):
    out = np.zeros_like(p)
    lib().pfo_saturation(op.ref, _ptr(p), _ptr(out), C.c_int(fcn))
    return out


def relperm(op, p, fcn=0):
    out = np.zeros_like(p)
    lib().pfo_relperm(op.ref, _ptr(p), _ptr(out), C.c_int(fcn))
    return out


def nl_function_eval(op, p, p_old, s_old, dt, t):
    op.set_time(t)
    F = np.zeros_like(p)
    lib().pfo_nl_function_eval(op.ref, _ptr(p), _ptr(p_old), _ptr(s_old),
                               C.c_double(dt), _ptr(F))
    return F


def jacobian_eval(op, p, dt, t, symm):
    op.set_time(t)
    J = [np.zeros_like(p) for _ in range(7)]
    arr = (_dp * 7)(*[_ptr(a) for a in J])
    lib().pfo_jacobian_eval(op.ref, _ptr(p), C.c_double(dt), C.c_int(symm), arr)
    return J


def matvec(op, J, symm, x):
    y = np.zeros_like(x)
    arr = (_dp * 7)(*[_ptr(a) for a in J])
    lib().pfo_matvec(op.ref, arr, C.c_int(symm), _ptr(x), _ptr(y))
    return y


def face_velocity(op, p, p_old, dt, t):
    """vx (nz, ny, nx+1), vy (nz, ny+1, nx), vz (nz+1, ny, nx): the face
    velocities NlFunctionEval writes (PL/nl_function_eval.c:846-848, 1230,
    1432, 2023) at pressure p, BC values at time t."""
    op.set_time(t)
    nx, ny, nz = op.prob.nx, op.prob.ny, op.prob.nz
    vx = np.zeros((nz, ny, nx + 1))
    vy = np.zeros((nz, ny + 1, nx))
    vz = np.zeros((nz + 1, ny, nx))
    lib().pfo_face_velocity(op.ref, _ptr(p), _ptr(p_old), C.c_double(dt),
                            _ptr(vx), _ptr(vy), _ptr(vz))
    return vx, vy, vz


def water_balance_fluxes(op, p, p_old, dt, t):
    """[boundary inflow, overland-top flux inflow, runoff, sources] of one
    step, volumes over the step."""
    op.set_time(t)
    out = np.zeros(4)
    lib().pfo_water_balance_fluxes(op.ref, _ptr(p), _ptr(p_old), C.c_double(dt), _ptr(out))
    return out


def overland_kin(op, p):
    n = (op.prob.nx + 2) * (op.prob.ny + 2)
    qx = np.zeros(n)
    qy = np.zeros(n)
    lib().pfo_overland_kin(op.ref, _ptr(p), _ptr(qx), _ptr(qy))
    return qx, qy


class MG:
    """pfo_mg: full=True keeps all seven couplings per row (the FullJacobian
    PC matrix); setup() then reads J[1], J[3], J[5] (W, S, L) as well."""

    def __init__(self, prob, pre, post, symmetric, max_levels=100, full=False):
        self.h = lib().pfo_mg_create_full(prob.nx, prob.ny, prob.nz, prob.dx,
                                          prob.dy, prob.dz, max_levels, pre, post,
                                          symmetric, int(full))
        self.shape = prob.shape
        self.full = full

    def level_matrix(self, l):
        """(C, E, N, U[, W, S, L]) ghosted arrays of level l."""
        dims, _ = self.levels()[l]
        shp = (dims[2] + 2, dims[1] + 2, dims[0] + 2)
        out = [np.zeros(shp) for _ in range(7 if self.full else 4)]
        arr = (_dp * 7)(*([_ptr(a) for a in out] + [None] * (7 - len(out))))
        lib().pfo_mg_level_matrix(self.h, l, arr)
        return out

    def levels(self):
        out = []
        for l in range(lib().pfo_mg_num_levels(self.h)):
            d = (C.c_int * 3)()
            dr = C.c_int()
            lib().pfo_mg_level_dims(self.h, l, d, C.byref(dr))
            out.append((tuple(d), dr.value))
        return out

    def setup(self, J):
        arr = (_dp * 7)(*[_ptr(a) for a in J])
        lib().pfo_mg_setup(self.h, arr)

    def apply(self, b):
        x = np.zeros(self.shape)
        lib().pfo_mg_apply(self.h, _ptr(np.ascontiguousarray(b)), _ptr(x))
        return x

    def __del__(self):
        try:
            lib().pfo_mg_destroy(self.h)
        except Exception:
            pass


def kinsol(op, opts, p, p_old, s_old, dt, t):
    st = KinStatsC()
    oc = opts_c(opts)
    r = lib().pfo_kinsol(op.ref, C.byref(oc), _ptr(p), _ptr(p_old), _ptr(s_old),
                         C.c_double(dt), C.c_double(t), C.byref(st))
    return r, st


def advance(op, opts, p0, t0, dt, nsteps, keep_history=True):
    """Returns (status, p_hist, s_hist, log) with histories of shape
    (nsteps+1, nz+2, ny+2, nx+2)."""
    prob = op.prob
    p = np.ascontiguousarray(p0.copy())
    s = np.zeros_like(p)
    ph = np.zeros((nsteps + 1,) + prob.shape) if keep_history else None
    sh = np.zeros((nsteps + 1,) + prob.shape) if keep_history else None
    log = RunLogC()
    oc = opts_c(opts)
    r = lib().pfo_advance(op.ref, C.byref(oc), _ptr(p), _ptr(s), C.c_double(t0),
                          C.c_double(dt), C.c_int(nsteps),
                          C.c_int(opts.max_convergence_failures),
                          _ptr(ph) if keep_history else None,
                          _ptr(sh) if keep_history else None, C.byref(log))
    return r, (ph if keep_history else p), (sh if keep_history else s), log


def storage(op, p, s):
    a = C.c_double()
    b = C.c_double()
    lib().pfo_storage(op.ref, _ptr(p), _ptr(s), C.byref(a), C.byref(b))
    return a.value, b.value


# ---------------------------------------------------------------------------
# Time-step control: SelectTimeStep (PL/select_time_step.c:58-156),
# TimeCycleDataComputeNextTransition (PL/time_cycle_data.c:206-299) and the
# AdvanceRichards time loop (PL/solver_richards.c:2081-2090, 2970-3240,
# 3380-3410, 4603-4612), restated in Python: scalar logic only, the solves
# go through pfo_kinsol.
# ---------------------------------------------------------------------------
TIME_EPSILON = float(np.finfo(np.float32).eps) * 10      # PL/general.h:257
TIME_CYCLE_SUBDIVISIONS = 1000                            # PL/time_cycle_data.c:150


def _c_div(a, b):
    """C integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _c_mod(a, b):
    return a - b * _c_div(a, b)


def _pfround(x):
    return int(x - 0.5) if x < 0.0 else int(x + 0.5)     # PL/general.h:176


def next_transition(cycles, base_unit, start_time, stop_time, time):
    """cycles: list of (lengths, repeat).  Time to the next interval change
    of any multi-interval cycle, or -1.0."""
    deltat = -1.0
    assigned = False
    for lengths, repeat in cycles:
        division = len(lengths)
        cycle_length = int(sum(lengths))
        time_defined = float(repeat * cycle_length * base_unit) if repeat > 0 else stop_time
        if not (time < time_defined and division > 1):
            continue
        sub = TIME_CYCLE_SUBDIVISIONS
        dtime = _c_div(_pfround(time / (base_unit / sub)), sub)
        dstart = _c_div(_pfround(start_time / (base_unit / sub)), sub)
        n = _c_div(dtime - dstart, cycle_length)
        completed = _c_mod(dtime - dstart - 1, cycle_length)
        k, total = 0, int(lengths[0])
        while completed >= total:
            k += 1
            total += int(lengths[k])
        if (k + 1) % division == 0:
            n += 1
            transition = start_time + float(n * cycle_length) * base_unit
        else:
            transition = start_time + float(n * cycle_length + total) * base_unit
        if transition < time_defined:
            deltat = min(deltat, transition - time) if assigned else transition - time
            assigned = True
    if (time + deltat) - time <= TIME_EPSILON:
        deltat = -1.0
    return deltat


def select_time_step(tc, cycles, time, dt):
    """(dt, dt_info) of SelectTimeStep; tc a parflow_amd.problem.TimeControl."""
    if tc.ts_type == "Constant":
        dt = tc.value
    else:
        if dt == 0.0:
            dt = tc.initial_step
        else:
            dt = dt * tc.growth_factor
            if dt < tc.min_step:
                dt = tc.min_step
            if dt > tc.max_step:
                dt = tc.max_step
    bc_dt = next_transition(cycles, tc.base_unit, tc.start_time, tc.stop_time, time)
    if bc_dt <= 0.0:
        bc_dt = dt
    if dt > bc_dt:
        return bc_dt, "b"
    return dt, "p"


def bc_cycles(prob):
    return [(tuple(b.lengths), b.repeat) for b in prob.bc]


def time_loop(tc, cycles, solve, max_steps=10 ** 9, max_conv_failures=3):
    """The AdvanceRichards loop.  solve(dt, t, retry) -> (ok, (newton, gmres,
    fevals)) solves the step ending at t (retry: after a failed attempt, the
    caller restores the step's initial state first).  Returns the step log:
    rows (time, dt, dt_info, dumped, newton, gmres, fevals, failures)."""
    return list(time_loop_iter(tc, cycles, solve, max_steps, max_conv_failures))


def surface_predictor(op, p, s, vel, q, dt, pressure_value, lateral):
    """PL/solver_richards.c:3236-3362 on the top layer (z_mult = 1): returns
    p with unsaturated top cells whose free pore volume is below the step's
    inflow set to the ponding pressure of the excess.  vel = (vx, vy, vz) as
    face_velocity returns them; q = (qx, qy) ghosted surface fluxes (ke, kn)
    or None; the problem's source field stands for evap_trans."""
    prob = op.prob
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    dx, dy, dz = prob.dx, prob.dy, prob.dz
    k = nz - 1
    p = p.copy()
    top = p[nz, 1:-1, 1:-1]
    po = prob.g_poro[nz, 1:-1, 1:-1]
    sat = s[nz, 1:-1, 1:-1]
    et = op.src[nz, 1:-1, 1:-1] if op.src is not None else np.zeros((ny, nx))
    vx, vy, vz = vel
    vol = dx * dy * dz * 1.0 * po * sat
    flux_in = dx * dy * dz * 1.0 * et * dt
    vol_max = dx * dy * dz * 1.0 * po
    flux_darcy = (vz[k] * dx * dy * dt + (-vx[k, :, 1:] + vx[k, :, :-1]) * dy * dz * 1.0 * dt
                  + (-vy[k, 1:, :] + vy[k, :-1, :]) * dx * dz * 1.0 * dt)
    if lateral:
        lat = np.zeros((ny, nx))
        if q is not None:
            qx, qy = (np.asarray(a).reshape(ny + 2, nx + 2) for a in q)
            qx_west = qx[1:-1, :-2].copy()
            qx_west[:, 0] = 0.0                      # (i > 0) ? qx[i-1] : 0
            qy_south = qy[:-2, 1:-1].copy()
            qy_south[0, :] = 0.0
            lat = dt * ((qx[1:-1, 1:-1] - qx_west) * dy + (qy[1:-1, 1:-1] - qy_south) * dx)
        total = flux_in + flux_darcy + lat
    else:
        total = flux_in + flux_darcy
    pred = (total - (vol_max - vol)) / (dx * dy * po)
    mask = (total > (vol_max - vol)) & (top < 0.0)
    if pressure_value > 0.0:
        pred = np.full_like(pred, pressure_value)
    top[mask] = pred[mask]
    return p


def advance_ts(op, opts, p0, tc, max_steps=10 ** 9):
    """time_loop with oracle solves; returns (p_hist, s_hist, rows).  With
    tc.surface_predictor the predictor adjusts each attempt's initial guess:
    zero velocities before the first solve, else those of the step's initial
    pressure with the BC values of the last solve (the previous step's, or
    the failed attempt's after a failure)."""
    p = np.ascontiguousarray(p0.copy())
    s = saturation(op, p)
    ph, sh = [p.copy()], [s.copy()]
    state = {"solved": 0}

    def solve(dt, t, retry):
        nonlocal p, s
        if retry:
            p[...] = state["p_old"]
            s[...] = state["s_old"]
        else:
            state["p_old"] = p.copy()
            state["s_old"] = s.copy()
        if tc.surface_predictor:
            prob = op.prob
            if state["solved"] == 0 and not retry:
                vel = (np.zeros((prob.nz, prob.ny, prob.nx + 1)),
                       np.zeros((prob.nz, prob.ny + 1, prob.nx)),
                       np.zeros((prob.nz + 1, prob.ny, prob.nx)))
                q = None
            else:
                t_bc = t if retry else state["t_prev"]
                dt_v = state.get("dt_prev", dt)
                vel = face_velocity(op, p, p, dt_v, t_bc)
                q = overland_kin(op, p) if prob.overland else None
            p[...] = surface_predictor(op, p, s, vel, q, dt, tc.surface_predictor_pressure,
                                       tc.surface_lateral_flows)
        r, st = kinsol(op, opts, p, state["p_old"], state["s_old"], dt, t)
        if r == 0:
            state["solved"] += 1
            state["t_prev"], state["dt_prev"] = t, dt
        return r == 0, (st.nni, st.nli, st.nfe)

    rows = []
    for row in time_loop_iter(tc, bc_cycles(op.prob), solve, max_steps,
                              opts.max_convergence_failures):
        s = saturation(op, p)
        ph.append(p.copy())
        sh.append(s.copy())
        rows.append(row)
    return np.array(ph), np.array(sh), rows


def time_loop_iter(tc, cycles, solve, max_steps=10 ** 9, max_conv_failures=3):
    """Generator form of time_loop: yields each step's row after its solve."""
    t, dt = tc.start_time, 0.0
    iteration, dump_index, n = 0, 1, 0
    while n < max_steps and iteration < tc.max_iterations and t < tc.stop_time:
        converged, failures = True, 0
        counts = [0, 0, 0]
        info = "p"
        while True:
            if converged:
                dt, info = select_time_step(tc, cycles, t, dt)
            else:
                t = t - dt
                new_dt = 0.5 * dt
                if (t + new_dt) - t > TIME_EPSILON:
                    dt = new_dt
                else:
                    raise RuntimeError("Time increment is too small; solver has failed")
            dump = False
            if tc.dump_interval > 0:
                print_dt = tc.start_time + dump_index * tc.dump_interval - t
                if (dt + TIME_EPSILON) > print_dt:
                    if abs(dt - print_dt) > TIME_EPSILON:
                        dt = print_dt
                    info = "p"
                    dump = True
            elif tc.dump_interval < 0:
                dump = iteration % (-int(tc.dump_interval)) == 0
            if (t + dt) >= tc.stop_time:
                dt = tc.stop_time - t
                info = "f"
            t += dt
            ok, c = solve(dt, t, not converged)
            counts = [a + b for a, b in zip(counts, c)]
            if ok:
                break
            converged = False
            failures += 1
            if failures >= max_conv_failures:
                raise RuntimeError(f"Time step failed for time {t:12.4e}")
        iteration += 1
        if dump:
            dump_index += 1
        n += 1
        yield (t, dt, info, dump, counts[0], counts[1], counts[2], failures)
