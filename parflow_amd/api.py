"""Python binding of libpfhip's C-ABI (include/pfhip.h).

Mirrors the reference's module interface names for the Richards path:
``NlFunctionEval``, ``RichardsJacobianEval``, ``Saturation``,
``PhaseRelPerm``, ``Matvec``, ``OverlandFlowEvalKin``, the multigrid
``Precond`` and ``KinsolNonlinSolver`` / ``AdvanceRichards``.  Device
memory is allocated through the library (hipMalloc); numpy arrays are the
host side.  There is no CPU fallback: every call runs the HIP library and
raises if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .problem import Problem, SolverOptions, TimeControl

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfhip.so")

_lib = None
_dp = C.POINTER(C.c_double)
MAXI = 8

# kernel and phase timer ids of pfhip_prof_* (include/pfhip.h)
(PROF_RESIDUAL, PROF_CLOSURE, PROF_JACOBIAN, PROF_RBGS, PROF_RESIDUAL_JV,
 PROF_RESIDUAL_STEP, PROF_RBGS_DOUT, PROF_RBGS_DIN) = range(8)
(PROF_PH_NL_F_EVAL, PROF_PH_JV, PROF_PH_PC_SETUP, PROF_PH_PC_SOLVE, PROF_PH_ORTHO,
 PROF_PH_KINSOL) = range(8, 14)


class PFHIPError(RuntimeError):
    pass


class BCSpecC(C.Structure):
    _fields_ = [("type", C.c_int), ("n_intervals", C.c_int), ("repeat", C.c_int),
                ("lengths", C.c_int * MAXI), ("values", C.c_double * MAXI),
                ("z_ref", C.c_double)]


class ProblemDescC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("rho", C.c_double), ("mu", C.c_double), ("gravity", C.c_double),
                ("kx", C.c_void_p), ("ky", C.c_void_p), ("kz", C.c_void_p),
                ("poro", C.c_void_p), ("ss", C.c_void_p),
                ("alpha", C.c_void_p), ("nvg", C.c_void_p),
                ("sres", C.c_void_p), ("ssat", C.c_void_p),
                ("sres_value", C.c_double), ("ssat_value", C.c_double),
                ("src", C.c_void_p),
                ("slope_x", C.c_void_p), ("slope_y", C.c_void_p), ("mann", C.c_void_p),
                ("ov_eps", C.c_double),
                ("bc", BCSpecC * 6),
                ("base_unit", C.c_double), ("start_time", C.c_double),
                ("vg_table_points", C.c_int), ("vg_table_method", C.c_int),
                ("vg_table_min_head", C.c_double), ("vg_table_alpha", C.c_double),
                ("vg_table_n", C.c_double)]


class SolverOptsC(C.Structure):
    _fields_ = [("residual_tol", C.c_double), ("step_tol", C.c_double),
                ("max_iter", C.c_int), ("krylov_dim", C.c_int),
                ("max_restarts", C.c_int), ("eta_choice", C.c_int),
                ("eta_value", C.c_double), ("eta_alpha", C.c_double),
                ("eta_gamma", C.c_double), ("use_jacobian", C.c_int),
                ("deriv_eps", C.c_double), ("globalization", C.c_int),
                ("precond", C.c_int), ("mg_pre", C.c_int), ("mg_post", C.c_int),
                ("mg_symmetric", C.c_int), ("mg_max_levels", C.c_int),
                ("mg_precision", C.c_int), ("krylov", C.c_int),
                ("pc_matrix", C.c_int)]


def _fill_bc(cb, b):
    cb.type = b.type
    cb.n_intervals = len(b.values)
    cb.repeat = b.repeat
    for i, (L_, v) in enumerate(zip(b.lengths, b.values)):
        cb.lengths[i] = int(L_)
        cb.values[i] = float(v)
    cb.z_ref = b.z_ref


def bc_specs_c(prob: Problem):
    """The six faces' BC specs as a pfhip_bc_spec[6] array."""
    arr = (BCSpecC * 6)()
    for f, b in enumerate(prob.bc):
        _fill_bc(arr[f], b)
    return arr


class TimeStepC(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double), ("initial_step", C.c_double),
                ("factor", C.c_double), ("min_step", C.c_double), ("max_step", C.c_double)]


class TimeControlC(C.Structure):
    _fields_ = [("start_time", C.c_double), ("stop_time", C.c_double),
                ("base_unit", C.c_double), ("dump_interval", C.c_double),
                ("max_iterations", C.c_int), ("ts", TimeStepC),
                ("surface_predictor", C.c_int), ("surface_predictor_pressure", C.c_double),
                ("surface_lateral_flows", C.c_int)]


class TimeStateC(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("iteration_number", C.c_int),
                ("dump_index", C.c_int)]


class StepLogC(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("dt_info", C.c_char),
                ("dumped", C.c_int), ("newton", C.c_int), ("gmres", C.c_int),
                ("fevals", C.c_int), ("conv_failures", C.c_int)]


def time_control_c(tc: TimeControl):
    c = TimeControlC()
    c.start_time, c.stop_time = tc.start_time, tc.stop_time
    c.base_unit, c.dump_interval = tc.base_unit, tc.dump_interval
    c.max_iterations = int(tc.max_iterations)
    c.ts.type = 0 if tc.ts_type == "Constant" else 1
    c.ts.value = tc.value
    c.ts.initial_step, c.ts.factor = tc.initial_step, tc.growth_factor
    c.ts.min_step, c.ts.max_step = tc.min_step, tc.max_step
    c.surface_predictor = int(bool(tc.surface_predictor))
    c.surface_predictor_pressure = tc.surface_predictor_pressure
    c.surface_lateral_flows = int(bool(tc.surface_lateral_flows))
    return c


def next_transition(prob: Problem, tc: TimeControl, time):
    """pfhip_next_transition over prob's face cycles (host only)."""
    return lib().pfhip_next_transition(bc_specs_c(prob), 6, C.c_double(tc.base_unit),
                                       C.c_double(tc.start_time),
                                       C.c_double(tc.stop_time), C.c_double(time))


def select_time_step(prob: Problem, tc: TimeControl, time, dt):
    """(dt, dt_info) of pfhip_select_time_step (host only)."""
    d = C.c_double(dt)
    info = C.c_char(b"?")
    check(lib().pfhip_select_time_step(C.byref(time_control_c(tc)), bc_specs_c(prob), 6,
                                       C.c_double(time), C.byref(d), C.byref(info)),
          "select_time_step")
    return d.value, info.value.decode()


class KinStatsC(C.Structure):
    _fields_ = [("ret", C.c_int), ("nni", C.c_int), ("nli", C.c_int),
                ("nfe", C.c_int), ("npe", C.c_int), ("nps", C.c_int),
                ("nbcf", C.c_int), ("nbktrk", C.c_int), ("ncfl", C.c_int),
                ("fnorm", C.c_double)]


# exported symbols of include/pfhip.h
EXPORTS = [
    "pfhip_last_error", "pfhip_device_count", "pfhip_malloc", "pfhip_free",
    "pfhip_memcpy_h2d", "pfhip_memcpy_d2h", "pfhip_memset", "pfhip_synchronize",
    "pfhip_problem_create", "pfhip_problem_destroy", "pfhip_bc_pressure",
    "pfhip_saturation", "pfhip_phase_rel_perm", "pfhip_nl_function_eval",
    "pfhip_face_velocity", "pfhip_water_balance_fluxes", "pfhip_evap_trans_sum",
    "pfhip_next_transition", "pfhip_select_time_step", "pfhip_advance_richards_ts",
    "pfhip_surface_predictor",
    "pfhip_richards_jacobian_eval", "pfhip_matvec", "pfhip_overland_flow_eval_kin",
    "pfhip_vec_linear_sum", "pfhip_vec_const", "pfhip_vec_dot", "pfhip_vec_max_norm",
    "pfhip_vec_l1_norm", "pfhip_vec_prod", "pfhip_vec_div", "pfhip_vec_scale", "pfhip_vec_abs",
    "pfhip_vec_inv", "pfhip_vec_add_const", "pfhip_vec_wrms_norm", "pfhip_vec_wl2_norm",
    "pfhip_vec_min", "pfhip_vec_constr_prod_pos", "pfhip_vec_compare", "pfhip_vec_inv_test",
    "pfhip_mg_create", "pfhip_mg_create_full", "pfhip_mg_destroy",
    "pfhip_mg_num_levels",
    "pfhip_mg_setup", "pfhip_mg_apply", "pfhip_solver_create", "pfhip_solver_destroy",
    "pfhip_kinsol_nonlin_solver", "pfhip_advance_richards", "pfhip_water_balance_storage",
    "pfhip_ensemble_create", "pfhip_ensemble_destroy", "pfhip_ensemble_advance_richards",
    "pfhip_ensemble_stats",
]


def lib():
    """Loads libpfhip.so (built in-tree by parflow_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PFHIPError(f"{LIB_PATH} not built: run python -m parflow_amd.build")
        L = C.CDLL(LIB_PATH)
        L.pfhip_last_error.restype = C.c_char_p
        vp = C.c_void_p
        L.pfhip_malloc.argtypes = [C.POINTER(vp), C.c_long]
        L.pfhip_free.argtypes = [vp]
        L.pfhip_memcpy_h2d.argtypes = [vp, vp, C.c_long]
        L.pfhip_memcpy_d2h.argtypes = [vp, vp, C.c_long]
        L.pfhip_memset.argtypes = [vp, C.c_int, C.c_long]
        L.pfhip_problem_create.argtypes = [C.POINTER(ProblemDescC), C.POINTER(vp)]
        L.pfhip_problem_destroy.argtypes = [vp]
        L.pfhip_bc_pressure.argtypes = [vp, C.c_double, vp]
        L.pfhip_saturation.argtypes = [vp, vp, vp, C.c_int, vp]
        L.pfhip_phase_rel_perm.argtypes = [vp, vp, vp, C.c_int, vp]
        L.pfhip_nl_function_eval.argtypes = [vp, vp, vp, vp, vp, C.c_double, vp]
        L.pfhip_face_velocity.argtypes = [vp, vp, vp, C.c_double, vp, vp, vp, vp]
        L.pfhip_water_balance_fluxes.argtypes = [vp, vp, vp, C.c_double,
                                                 C.POINTER(C.c_double), vp]
        L.pfhip_evap_trans_sum.argtypes = [vp, C.c_double, vp, vp]
        L.pfhip_next_transition.restype = C.c_double
        L.pfhip_next_transition.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_double]
        L.pfhip_select_time_step.argtypes = [vp, vp, C.c_int, C.c_double,
                                             C.POINTER(C.c_double), C.POINTER(C.c_char)]
        L.pfhip_advance_richards_ts.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp,
                                                C.POINTER(C.c_int)]
        L.pfhip_surface_predictor.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                              C.c_double, C.c_int, vp]
        L.pfhip_richards_jacobian_eval.argtypes = [vp, vp, C.c_double, C.c_int,
                                                   C.POINTER(vp), vp]
        L.pfhip_matvec.argtypes = [vp, C.POINTER(vp), C.c_int, vp, vp, vp]
        L.pfhip_overland_flow_eval_kin.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, vp]
        L.pfhip_vec_linear_sum.argtypes = [vp, C.c_double, vp, C.c_double, vp, vp, vp]
        L.pfhip_vec_const.argtypes = [vp, C.c_double, vp, vp]
        for f in ("pfhip_vec_dot",):
            getattr(L, f).argtypes = [vp, vp, vp, C.POINTER(C.c_double), vp]
        for f in ("pfhip_vec_max_norm", "pfhip_vec_l1_norm"):
            getattr(L, f).argtypes = [vp, vp, C.POINTER(C.c_double), vp]
        for f in ("pfhip_vec_prod", "pfhip_vec_div"):
            getattr(L, f).argtypes = [vp, vp, vp, vp, vp]
        for f in ("pfhip_vec_abs", "pfhip_vec_inv"):
            getattr(L, f).argtypes = [vp, vp, vp, vp]
        L.pfhip_vec_scale.argtypes = [vp, C.c_double, vp, vp, vp]
        L.pfhip_vec_compare.argtypes = [vp, C.c_double, vp, vp, vp]
        L.pfhip_vec_add_const.argtypes = [vp, vp, C.c_double, vp, vp]
        for f in ("pfhip_vec_wrms_norm", "pfhip_vec_wl2_norm"):
            getattr(L, f).argtypes = [vp, vp, vp, C.POINTER(C.c_double), vp]
        L.pfhip_vec_min.argtypes = [vp, vp, C.POINTER(C.c_double), vp]
        L.pfhip_vec_constr_prod_pos.argtypes = [vp, vp, vp, C.POINTER(C.c_int), vp]
        L.pfhip_vec_inv_test.argtypes = [vp, vp, vp, C.POINTER(C.c_int), vp]
        L.pfhip_water_balance_storage.argtypes = [vp, vp, vp, C.POINTER(C.c_double),
                                                  C.POINTER(C.c_double), vp]
        L.pfhip_mg_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(vp)]
        L.pfhip_mg_create_full.argtypes = L.pfhip_mg_create.argtypes
        L.pfhip_mg_destroy.argtypes = [vp]
        L.pfhip_mg_num_levels.argtypes = [vp, C.POINTER(C.c_int)]
        L.pfhip_mg_setup.argtypes = [vp, C.POINTER(vp), vp]
        L.pfhip_mg_apply.argtypes = [vp, vp, vp, vp]
        L.pfhip_solver_create.argtypes = [vp, C.POINTER(SolverOptsC), C.POINTER(vp)]
        L.pfhip_solver_destroy.argtypes = [vp]
        L.pfhip_kinsol_nonlin_solver.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double,
                                                 C.POINTER(KinStatsC)]
        L.pfhip_advance_richards.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_int,
                                             C.c_int, vp, vp, vp, vp, C.POINTER(C.c_int)]
        L.pfhip_ensemble_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp)]
        L.pfhip_ensemble_destroy.argtypes = [vp]
        L.pfhip_ensemble_advance_richards.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), vp, vp,
                                                      C.c_int, C.c_int, vp, vp, vp, vp]
        L.pfhip_ensemble_stats.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.pfhip_prof_enable.argtypes = [C.c_int, C.c_int]
        L.pfhip_prof_reset.argtypes = [C.c_int]
        L.pfhip_prof_read.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]
        L.pfhip_prof_disable.argtypes = [C.c_int]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().pfhip_last_error().decode(errors="replace")
        raise PFHIPError(f"{what}: error {rc}: {msg}")


# ---------------------------------------------------------------------------
class DeviceArray:
    """A device allocation of float64 (or float32) values."""

    def __init__(self, n, dtype=np.float64, zero=True):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().pfhip_malloc(C.byref(p), self.nbytes), "pfhip_malloc")
        self.ptr = p
        if zero:
            check(lib().pfhip_memset(self.ptr, 0, self.nbytes), "pfhip_memset")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, zero=False)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n, (a.size, self.n)
        check(lib().pfhip_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes),
              "h2d")

    def download(self, shape=None):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().pfhip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes),
              "d2h")
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().pfhip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def value(self):
        return self.ptr.value


def solver_opts_c(o: SolverOptions, precision=64, krylov=0):
    pre, post, sym = o.mg_params()
    c = SolverOptsC()
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
    c.mg_precision = precision
    c.krylov = krylov
    c.pc_matrix = o.pc_matrix_code()
    return c


class GPUProblem:
    """Device-resident ProblemData for one Problem (box domain)."""

    def __init__(self, prob: Problem, src=None):
        self.prob = prob
        self._keep = {}
        d = ProblemDescC()
        d.nx, d.ny, d.nz = prob.nx, prob.ny, prob.nz
        d.dx, d.dy, d.dz = prob.dx, prob.dy, prob.dz
        d.x0, d.y0, d.z0 = prob.x0, prob.y0, prob.z0
        d.rho, d.mu, d.gravity = prob.rho, prob.mu, prob.gravity
        fields = ["kx", "ky", "kz", "poro", "ss", "alpha", "nvg"]
        # region-constant residual/saturated saturations travel as scalars
        for name in ("sres", "ssat"):
            if getattr(prob, "uniform_" + name):
                setattr(d, name, None)
                setattr(d, name + "_value", float(getattr(prob, name)))
            else:
                fields.append(name)
        # isotropic permeability shares one device array
        cache = {}
        for name in fields:
            a = np.ascontiguousarray(getattr(prob, "g_" + name))
            key = None
            for kname, (arr, darr) in cache.items():
                if arr.shape == a.shape and np.array_equal(arr, a):
                    key = kname
                    break
            if key is None:
                darr = DeviceArray.from_host(a)
                cache[name] = (a, darr)
                self._keep[name] = darr
            else:
                darr = cache[key][1]
            setattr(d, name, darr.value)
        if src is not None:
            self._keep["src"] = DeviceArray.from_host(np.ascontiguousarray(src))
            d.src = self._keep["src"].value
        if prob.overland:
            for name in ("slope_x", "slope_y", "mann"):
                darr = DeviceArray.from_host(np.ascontiguousarray(getattr(prob, "g_" + name)))
                self._keep[name] = darr
                setattr(d, name, darr.value)
        d.ov_eps = prob.ov_eps
        for f, b in enumerate(prob.bc):
            _fill_bc(d.bc[f], b)
        d.base_unit = prob.base_unit
        d.start_time = prob.start_time
        tp = prob.vg_table_params
        if tp is not None:
            (d.vg_table_method, d.vg_table_points, d.vg_table_min_head, d.vg_table_alpha,
             d.vg_table_n) = tp
        self.desc = d
        h = C.c_void_p()
        check(lib().pfhip_problem_create(C.byref(d), C.byref(h)), "pfhip_problem_create")
        self.h = h
        self.ng = int(np.prod(prob.shape))

    def __del__(self):
        try:
            if self.h:
                lib().pfhip_problem_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # -- helpers -------------------------------------------------------------
    def vec(self, a=None):
        if a is None:
            return DeviceArray(self.ng)
        return DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float64).ravel())

    def get(self, d):
        return d.download(self.prob.shape)

    def bc_pressure(self, t):
        check(lib().pfhip_bc_pressure(self.h, C.c_double(t), None), "bc_pressure")

    # -- module entry points (host arrays in, host arrays out) ---------------
    def saturation(self, p, fcn=0):
        dp, out = self.vec(p), self.vec()
        check(lib().pfhip_saturation(self.h, dp.ptr, out.ptr, fcn, None), "saturation")
        return self.get(out)

    def phase_rel_perm(self, p, fcn=0):
        dp, out = self.vec(p), self.vec()
        check(lib().pfhip_phase_rel_perm(self.h, dp.ptr, out.ptr, fcn, None), "relperm")
        return self.get(out)

    def nl_function_eval(self, p, p_old, s_old, dt, t):
        self.bc_pressure(t)
        dp, dpo, dso, F = self.vec(p), self.vec(p_old), self.vec(s_old), self.vec()
        check(lib().pfhip_nl_function_eval(self.h, dp.ptr, F.ptr, dpo.ptr, dso.ptr,
                                           C.c_double(dt), None), "nl_function_eval")
        return self.get(F)

    def face_velocity(self, p, p_old, dt, t):
        """(vx, vy, vz) face velocities at pressure p, shaped (nz, ny, nx+1),
        (nz, ny+1, nx), (nz+1, ny, nx) like the reference's velx/vely/velz."""
        self.bc_pressure(t)
        nx, ny, nz = self.prob.nx, self.prob.ny, self.prob.nz
        dp, dpo = self.vec(p), self.vec(p_old)
        shp = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
        outs = [DeviceArray(int(np.prod(s))) for s in shp]
        check(lib().pfhip_face_velocity(self.h, dp.ptr, dpo.ptr, C.c_double(dt),
                                        *[o.ptr for o in outs], None), "face_velocity")
        return [o.download(s) for o, s in zip(outs, shp)]

    def richards_jacobian_eval(self, p, dt, t, symm):
        self.bc_pressure(t)
        dp = self.vec(p)
        J = [self.vec() for _ in range(7)]
        arr = (C.c_void_p * 7)(*[j.value for j in J])
        check(lib().pfhip_richards_jacobian_eval(self.h, dp.ptr, C.c_double(dt), int(symm),
                                                 arr, None), "jacobian_eval")
        return [self.get(j) for j in J]

    def matvec(self, J, symm, x):
        dJ = [self.vec(j) for j in J]
        arr = (C.c_void_p * 7)(*[j.value for j in dJ])
        dx_, y = self.vec(x), self.vec()
        check(lib().pfhip_matvec(self.h, arr, int(symm), dx_.ptr, y.ptr, None), "matvec")
        return self.get(y)

    def overland_flow_eval_kin(self, p, fcn=0):
        n2 = (self.prob.nx + 2) * (self.prob.ny + 2)
        dp = self.vec(p)
        outs = [DeviceArray(n2) for _ in range(4)]
        check(lib().pfhip_overland_flow_eval_kin(self.h, dp.ptr, *[o.ptr for o in outs],
                                                 fcn, None), "overland")
        return [o.download() for o in outs]

    def vec_dot(self, x, y):
        dx_, dy_ = self.vec(x), self.vec(y)
        r = C.c_double()
        check(lib().pfhip_vec_dot(self.h, dx_.ptr, dy_.ptr, C.byref(r), None), "dot")
        return r.value

    def vec_max_norm(self, x):
        dx_ = self.vec(x)
        r = C.c_double()
        check(lib().pfhip_vec_max_norm(self.h, dx_.ptr, C.byref(r), None), "maxnorm")
        return r.value

    def vec_l1_norm(self, x):
        dx_ = self.vec(x)
        r = C.c_double()
        check(lib().pfhip_vec_l1_norm(self.h, dx_.ptr, C.byref(r), None), "l1norm")
        return r.value

    def water_balance_storage(self, p, s):
        """(subsurface, surface) storage of pressure p and saturation s
        (host arrays or DeviceArrays)."""
        dp = p if isinstance(p, DeviceArray) else self.vec(p)
        ds = s if isinstance(s, DeviceArray) else self.vec(s)
        a, b = C.c_double(), C.c_double()
        check(lib().pfhip_water_balance_storage(self.h, dp.ptr, ds.ptr, C.byref(a),
                                                C.byref(b), None), "storage")
        return a.value, b.value

    def water_balance_fluxes(self, p, p_old, dt, t):
        """[boundary inflow, overland-top flux inflow, runoff, sources] of a
        step from p_old to p of length dt ending at time t (volumes)."""
        self.bc_pressure(t)
        dp = p if isinstance(p, DeviceArray) else self.vec(p)
        dpo = p_old if isinstance(p_old, DeviceArray) else self.vec(p_old)
        out = (C.c_double * 4)()
        check(lib().pfhip_water_balance_fluxes(self.h, dp.ptr, dpo.ptr, C.c_double(dt), out,
                                               None), "water_balance_fluxes")
        return np.array(out[:])

    def surface_predictor(self, p, sat, vel, q, dt, pressure_value, lateral):
        """p after pfhip_surface_predictor; vel = (vx, vy, vz) face arrays,
        q = (qx, qy) surface fluxes or None."""
        dp, ds = self.vec(p), self.vec(sat)
        dv = [DeviceArray.from_host(np.ascontiguousarray(v, dtype=np.float64).ravel())
              for v in vel]
        dq = ([DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float64).ravel())
               for a in q] if q is not None else None)
        check(lib().pfhip_surface_predictor(self.h, dp.ptr, ds.ptr, *[v.ptr for v in dv],
                                            dq[0].ptr if dq else None,
                                            dq[1].ptr if dq else None, C.c_double(dt),
                                            C.c_double(pressure_value), int(bool(lateral)),
                                            None), "surface_predictor")
        return self.get(dp)

    def evap_trans_sum(self, acc, dt):
        """acc (DeviceArray, ghosted) += src * dx dy dz * dt (EvapTransSum)."""
        check(lib().pfhip_evap_trans_sum(self.h, C.c_double(dt), acc.ptr, None),
              "evap_trans_sum")


def water_balance_report(gp, p_hist, s_hist, dts, t0):
    """Per-step water budget of a run, the bookkeeping of the reference's
    test/python/water_balance_x.py:520-620 on the device: storages
    (pftools hydrology.py:136-205) and the step's boundary, overland-top,
    runoff and source volumes.  p_hist/s_hist: host histories (step 0 = the
    initial state); dts: the step lengths.  Each row holds the step's terms,
    the total water (subsurface + surface storage), the expected total
    (previous total + inflows - runoff) and percent_diff =
    |total - expected| / |expected| * 100, the quantity the reference test
    bounds by 0.005 %."""
    rows = []
    sub, sur = gp.water_balance_storage(p_hist[0], s_hist[0])
    prev = sub + sur
    t = t0
    for n in range(1, len(p_hist)):
        dt = float(dts[n - 1])
        t += dt
        sub, sur = gp.water_balance_storage(p_hist[n], s_hist[n])
        f = gp.water_balance_fluxes(p_hist[n], p_hist[n - 1], dt, t)
        total = sub + sur
        expected = prev + f[0] + f[1] - f[2] + f[3]
        rows.append(dict(step=n, time=t, subsurface_storage=sub, surface_storage=sur,
                         boundary_inflow=f[0], overland_flux_inflow=f[1],
                         surface_runoff=f[2], sources=f[3], total=total,
                         expected=expected,
                         percent_diff=abs(total - expected) / abs(expected) * 100.0))
        prev = total
    return rows


class GPUMultigrid:
    """The multigrid preconditioner module (PrecondInvoke).  full=True keeps
    all seven couplings per row (PCMatrixType FullJacobian); setup() then
    reads W, S, L (J[1], J[3], J[5]) as well."""

    def __init__(self, gp: GPUProblem, pre, post, symmetric, precision=64, max_levels=100,
                 full=False):
        self.gp = gp
        h = C.c_void_p()
        create = lib().pfhip_mg_create_full if full else lib().pfhip_mg_create
        check(create(gp.h, max_levels, pre, post, symmetric, precision, C.byref(h)),
              "mg_create")
        self.h = h

    def levels(self):
        n = C.c_int()
        check(lib().pfhip_mg_num_levels(self.h, C.byref(n)), "mg_levels")
        return n.value

    def setup(self, Jsym):
        self._J = [self.gp.vec(j) for j in Jsym]
        arr = (C.c_void_p * 7)(*[j.value for j in self._J])
        check(lib().pfhip_mg_setup(self.h, arr, None), "mg_setup")

    def apply(self, b):
        db, x = self.gp.vec(b), self.gp.vec()
        check(lib().pfhip_mg_apply(self.h, db.ptr, x.ptr, None), "mg_apply")
        return self.gp.get(x)

    def __del__(self):
        try:
            if self.h:
                lib().pfhip_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GPUSolver:
    """KinsolNonlinSolver + AdvanceRichards on the device."""

    def __init__(self, gp: GPUProblem, opts: SolverOptions, precision=64, krylov=0):
        self.gp = gp
        self.opts = opts
        self.c_opts = solver_opts_c(opts, precision, krylov)
        h = C.c_void_p()
        check(lib().pfhip_solver_create(gp.h, C.byref(self.c_opts), C.byref(h)),
              "solver_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().pfhip_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def kinsol(self, p, p_old, s_old, dt, t):
        dp, dpo, dso = self.gp.vec(p), self.gp.vec(p_old), self.gp.vec(s_old)
        st = KinStatsC()
        rc = lib().pfhip_kinsol_nonlin_solver(self.h, dp.ptr, dpo.ptr, dso.ptr,
                                              C.c_double(dt), C.c_double(t), C.byref(st))
        if rc not in (0, -4):
            check(rc, "kinsol")
        return rc, self.gp.get(dp), st

    def advance_ts_device(self, dp, ds, tc: TimeControl, state: TimeStateC, max_steps):
        """pfhip_advance_richards_ts on device arrays: up to max_steps steps
        of the time loop from state (updated in place); returns the rows."""
        log = (StepLogC * max(int(max_steps), 1))()
        done = C.c_int()
        tcc = time_control_c(tc)
        rc = lib().pfhip_advance_richards_ts(
            self.h, dp.ptr, ds.ptr, C.byref(tcc), C.byref(state), int(max_steps),
            int(self.opts.max_convergence_failures), log, C.byref(done))
        check(rc, "advance_richards_ts")
        return [dict(time=r.time, dt=r.dt, dt_info=r.dt_info.decode(), dumped=bool(r.dumped),
                     newton=r.newton, gmres=r.gmres, fevals=r.fevals,
                     conv_failures=r.conv_failures) for r in log[:done.value]]

    def run(self, p0, tc: TimeControl, max_steps=None, history=True):
        """The whole time loop from p0 at tc.start_time.  With history the
        loop returns to the host after every step: (p_hist, s_hist, rows),
        step 0 the initial state; otherwise (p, s, rows) at the end."""
        dp, ds = self.gp.vec(p0), self.gp.vec()
        state = TimeStateC()
        state.t, state.dt, state.iteration_number, state.dump_index = tc.start_time, 0.0, 0, 1
        limit = tc.max_iterations if max_steps is None else int(max_steps)
        if not history:
            rows = self.advance_ts_device(dp, ds, tc, state, limit)
            return self.gp.get(dp), self.gp.get(ds), rows
        check(lib().pfhip_saturation(self.gp.h, dp.ptr, ds.ptr, 0, None), "saturation")
        ph, sh, rows = [self.gp.get(dp)], [self.gp.get(ds)], []
        while len(rows) < limit:
            r = self.advance_ts_device(dp, ds, tc, state, 1)
            if not r:
                break
            rows += r
            ph.append(self.gp.get(dp))
            sh.append(self.gp.get(ds))
        return np.array(ph), np.array(sh), rows

    def advance_device(self, dp, ds, t0, dt, nsteps):
        """Runs nsteps on device arrays dp, ds; returns per-step counts."""
        newton = np.zeros(nsteps, dtype=np.int32)
        gmres = np.zeros(nsteps, dtype=np.int32)
        fevals = np.zeros(nsteps, dtype=np.int32)
        dtu = np.zeros(nsteps, dtype=np.float64)
        done = C.c_int()
        rc = lib().pfhip_advance_richards(
            self.h, dp.ptr, ds.ptr, C.c_double(t0), C.c_double(dt), int(nsteps),
            int(self.opts.max_convergence_failures),
            newton.ctypes.data_as(C.c_void_p), gmres.ctypes.data_as(C.c_void_p),
            fevals.ctypes.data_as(C.c_void_p), dtu.ctypes.data_as(C.c_void_p),
            C.byref(done))
        check(rc, "advance_richards")
        return dict(newton=newton, gmres=gmres, fevals=fevals, dt=dtu, steps=done.value)

    def advance(self, p0, t0, dt, nsteps, history=True):
        """Host-array convenience: returns (p_hist, s_hist, log)."""
        dp, ds = self.gp.vec(p0), self.gp.vec()
        ph, sh = [], []
        logs = []
        if history:
            check(lib().pfhip_saturation(self.gp.h, dp.ptr, ds.ptr, 0, None), "saturation")
            ph.append(self.gp.get(dp))
            sh.append(self.gp.get(ds))
            t = t0
            for n in range(nsteps):
                lg = self.advance_device(dp, ds, t, dt, 1)
                t += lg["dt"][0]
                logs.append(lg)
                ph.append(self.gp.get(dp))
                sh.append(self.gp.get(ds))
            log = {k: np.concatenate([l_[k] for l_ in logs]) for k in ("newton", "gmres",
                                                                        "fevals", "dt")}
            return np.array(ph), np.array(sh), log
        lg = self.advance_device(dp, ds, t0, dt, nsteps)
        return self.gp.get(dp), self.gp.get(ds), lg


class GPUEnsemble:
    """Independent problems advanced together, their device work batched
    (pfhip_ensemble_*).  members: list of (Problem, SolverOptions); every
    member uses the mixed form (FP32 V-cycle inside flexible GMRES)."""

    def __init__(self, members, precision=32):
        self.problems = [GPUProblem(p) for p, _ in members]
        self.solvers = [GPUSolver(gp, o, precision=precision, krylov=1)
                        for gp, (_, o) in zip(self.problems, members)]
        self.n = len(members)
        arr = (C.c_void_p * self.n)(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        check(lib().pfhip_ensemble_create(self.n, arr, C.byref(h)), "ensemble_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().pfhip_ensemble_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def advance_device(self, dps, dss, t0, dts, nsteps, max_conv_failures=3):
        """nsteps steps for every member on device arrays dps[m], dss[m];
        returns per-member per-step counts (arrays n x nsteps)."""
        n = self.n
        P = (C.c_void_p * n)(*[d.value for d in dps])
        S = (C.c_void_p * n)(*[d.value for d in dss])
        t0a = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
        dta = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (n,)))
        newton = np.zeros((n, nsteps), dtype=np.int32)
        gmres = np.zeros((n, nsteps), dtype=np.int32)
        fevals = np.zeros((n, nsteps), dtype=np.int32)
        done = np.zeros(n, dtype=np.int32)
        rc = lib().pfhip_ensemble_advance_richards(
            self.h, P, S, t0a.ctypes.data_as(C.c_void_p), dta.ctypes.data_as(C.c_void_p),
            int(nsteps), int(max_conv_failures), newton.ctypes.data_as(C.c_void_p),
            gmres.ctypes.data_as(C.c_void_p), fevals.ctypes.data_as(C.c_void_p),
            done.ctypes.data_as(C.c_void_p))
        check(rc, "ensemble_advance_richards")
        return dict(newton=newton, gmres=gmres, fevals=fevals, steps=done)

    def stats(self):
        r, l_ = C.c_long(), C.c_long()
        check(lib().pfhip_ensemble_stats(self.h, C.byref(r), C.byref(l_)), "ensemble_stats")
        return r.value, l_.value
