"""Problem decks for the Richards Newton-Krylov path.

A :class:`Problem` carries what ParFlow's ProblemData and BCPressure
modules hand to the nonlinear solver for a box domain:

* the computational grid (``ComputationalGrid.*`` keys),
* per-cell permeability (``Perm.*`` times ``TensorVal*``), porosity,
  specific storage and van Genuchten parameters (``Phase.Saturation.*`` /
  ``Phase.RelPerm.*``, region constants or per-cell ``File`` fields,
  pfsimulator/parflow_lib/problem_saturation.c:313-416),
* constant density / viscosity / gravity (problem_phase_density.c:119),
* the six box-face boundary conditions with their time cycles
  (``BCPressure.*``, ``Cycle.*``; bc_pressure.h:97-141),
* surface slopes and Manning's coefficient for the kinematic overland top
  boundary (``TopoSlopesX/Y``, ``Mannings``),
* the solver keys of kinsol_nonlin_solver.c:766-963 and pf_pfmg.c:278-310.

Arrays are stored ghosted, x fastest -- the Subvector layout of
pfsimulator/parflow_lib/vector.h:133-137 -- as numpy ``float64`` arrays of
shape ``(nz+2, ny+2, nx+2)`` (surface fields ``(ny+2, nx+2)``).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

BC_FLUX = 0
BC_DIR_EQUIL = 1
BC_OVERLAND = 2

FACE_NAMES = ("x_lower", "x_upper", "y_lower", "y_upper", "z_lower", "z_upper")


@dataclass
class BCSpec:
    """One box-face boundary condition with a ParFlow time cycle.

    ``type``: ``BC_FLUX`` (FluxConst), ``BC_DIR_EQUIL`` (DirEquilRefPatch:
    ``values`` are reference pressures at elevation ``z_ref``) or
    ``BC_OVERLAND`` (OverlandKinematic, top face only; ``values`` are the
    rain/evaporation fluxes).  ``lengths`` are interval lengths in
    ``TimingInfo.BaseUnit`` units; ``repeat`` -1 repeats forever
    (time_cycle_data.c:152-197).
    """

    type: int = BC_FLUX
    values: Sequence[float] = (0.0,)
    lengths: Sequence[int] = (1,)
    repeat: int = -1
    z_ref: float = 0.0

    def __post_init__(self):
        if len(self.values) != len(self.lengths):
            raise ValueError("BCSpec: one value per cycle interval required")
        if not 1 <= len(self.values) <= 8:
            raise ValueError("BCSpec: 1..8 cycle intervals supported")


@dataclass
class SolverOptions:
    """Solver keys (defaults of kinsol_nonlin_solver.c:783-963)."""

    residual_tol: float = 1e-7          # Solver.Nonlinear.ResidualTol
    step_tol: float = 1e-7              # Solver.Nonlinear.StepTol
    max_iter: int = 15                  # Solver.Nonlinear.MaxIter
    krylov_dim: int = 10                # Solver.Linear.KrylovDimension
    max_restarts: int = 0               # Solver.Linear.MaxRestarts
    eta_choice: int = 2                 # 0 EtaConstant, 1 Walker1, 2 Walker2
    eta_value: float = 1e-4             # Solver.Nonlinear.EtaValue
    eta_alpha: float = 2.0              # Solver.Nonlinear.EtaAlpha
    eta_gamma: float = 0.9              # Solver.Nonlinear.EtaGamma
    use_jacobian: bool = False          # Solver.Nonlinear.UseJacobian
    deriv_eps: float = 1e-7             # Solver.Nonlinear.DerivativeEpsilon
    globalization: int = 1              # 0 InexactNewton, 1 LineSearch
    precond: str = "MGSemi"             # NoPC | MGSemi | PFMG
    num_pre_relax: int = 1              # PFMG NumPreRelax
    num_post_relax: int = 1             # PFMG NumPostRelax
    max_levels: int = 100               # MGSemi MaxLevels
    max_convergence_failures: int = 3   # Solver.MaxConvergenceFailures
    pc_matrix: str = "PFSymmetric"      # Solver.Linear.Preconditioner.PCMatrixType

    def pc_matrix_code(self):
        """0 PFSymmetric (symmetric part, C E N U), 1 FullJacobian (all seven
        couplings); kinsol_pc.c:215-240.  SymmetricPart and Picard are not
        part of this build."""
        codes = {"PFSymmetric": 0, "FullJacobian": 1}
        if self.pc_matrix not in codes:
            raise ValueError(f"unsupported PCMatrixType {self.pc_matrix!r}")
        return codes[self.pc_matrix]

    def mg_params(self):
        """(pre, post, symmetric) sweeps of the V-cycle for ``precond``.

        MGSemi uses the RedBlackGSPoint defaults (MaxIter 1, Symmetric True:
        red-black-red, rb_GS_point.c:440-470).  PFMG uses its NumPre/PostRelax
        red-black sweeps (pf_pfmg.c:292-310)."""
        if self.precond == "MGSemi":
            return 1, 1, 1
        if self.precond == "PFMG":
            return self.num_pre_relax, self.num_post_relax, 0
        if self.precond == "NoPC":
            return 0, 0, 0
        raise ValueError(f"unknown preconditioner {self.precond!r}")


@dataclass
class TimeControl:
    """``TimingInfo.*``, ``TimeStep.*`` and ``Solver.MaxIter``: what the
    AdvanceRichards time loop and SelectTimeStep read
    (pfsimulator/parflow_lib/select_time_step.c:200-245; the TimingInfo
    keys read at problem.c:87-98).

    ``ts_type`` "Constant" uses ``value``; "Growth" starts at
    ``initial_step`` and multiplies the previous step by ``growth_factor``,
    clipped to [``min_step``, ``max_step``].  ``dump_interval`` > 0 is a time
    interval the steps are clipped to; < 0 dumps every ``-dump_interval``
    steps.  ``surface_predictor`` turns on the top-cell ponding predictor
    applied to each step's initial guess (solver_richards.c:3236-3362)."""

    start_time: float = 0.0
    stop_time: float = 1.0
    base_unit: float = 1.0
    dump_interval: float = -1.0
    max_iterations: int = 1000000
    ts_type: str = "Constant"
    value: float = 1.0
    initial_step: float = 0.0
    growth_factor: float = 1.0
    min_step: float = 0.0
    max_step: float = 0.0
    # Solver.SurfacePredictor, .PressureValue, .LateralFlows
    # (solver_richards.c:6959-6976)
    surface_predictor: bool = False
    surface_predictor_pressure: float = 0.00001
    surface_lateral_flows: bool = False

    def __post_init__(self):
        if self.ts_type not in ("Constant", "Growth"):
            raise ValueError(f"TimeControl: TimeStep.Type {self.ts_type!r} is not supported")
        if self.ts_type == "Growth" and not (self.initial_step > 0 and self.max_step > 0):
            raise ValueError("TimeControl: Growth needs InitialStep and MaxStep > 0")


def _ghost3(a, nx, ny, nz, fill_edge=True):
    out = np.zeros((nz + 2, ny + 2, nx + 2), dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 0:
        out[...] = float(a) if fill_edge else 0.0
        if not fill_edge:
            out[1:-1, 1:-1, 1:-1] = float(a)
        return out
    if a.shape != (nz, ny, nx):
        raise ValueError(f"field shape {a.shape} != {(nz, ny, nx)}")
    out[1:-1, 1:-1, 1:-1] = a
    if fill_edge:
        out[1:-1, 1:-1, 0] = a[:, :, 0]
        out[1:-1, 1:-1, -1] = a[:, :, -1]
        out[1:-1, 0, :] = out[1:-1, 1, :]
        out[1:-1, -1, :] = out[1:-1, -2, :]
        out[0] = out[1]
        out[-1] = out[-2]
    return out


def _ghost2(a, nx, ny):
    a = np.asarray(a, dtype=np.float64)
    out = np.zeros((ny + 2, nx + 2), dtype=np.float64)
    if a.ndim == 0:
        out[1:-1, 1:-1] = float(a)
        return out
    if a.shape != (ny, nx):
        raise ValueError(f"surface field shape {a.shape} != {(ny, nx)}")
    out[1:-1, 1:-1] = a
    return out


@dataclass
class Problem:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    rho: float = 1.0
    mu: float = 1.0
    gravity: float = 1.0
    kx: object = 1.0
    ky: object = None
    kz: object = None
    poro: object = 0.4
    ss: object = 1e-4
    alpha: object = 2.0
    nvg: object = 2.0
    sres: object = 0.2
    ssat: object = 1.0
    slope_x: object = 0.0
    slope_y: object = 0.0
    mann: object = 1.0
    ov_eps: float = 1e-5
    bc: List[BCSpec] = field(default_factory=lambda: [BCSpec() for _ in range(6)])
    base_unit: float = 1.0
    start_time: float = 0.0
    name: str = "problem"
    # van Genuchten kr lookup table (Geom.*.RelPerm.NumSamplePoints,
    # .MinPressureHead, .InterpolationMethod; problem_phase_rel_perm.c
    # :1894-1915): (points, min_pressure_head, "Spline" | "Linear"), or None
    # for the closed form.  The table is built from alpha and n, which must
    # then be uniform.
    vg_table: Optional[tuple] = None

    def __post_init__(self):
        if self.ky is None:
            self.ky = self.kx
        if self.kz is None:
            self.kz = self.kx
        if self.vg_table is not None:
            pts, head, method = self.vg_table
            if int(pts) < 2 or not float(head) < 0.0 or method not in ("Spline", "Linear"):
                raise ValueError(f"vg_table {self.vg_table!r}: need points >= 2, "
                                 "min pressure head < 0, Spline or Linear")
            for name in ("alpha", "nvg"):
                a = np.asarray(getattr(self, name), dtype=np.float64)
                if a.ndim and not np.all(a == a.flat[0]):
                    raise ValueError(f"vg_table needs a uniform {name}")
        if len(self.bc) != 6:
            raise ValueError("six face boundary conditions required")
        for f, b in enumerate(self.bc):
            if b.type == BC_OVERLAND and f != 5:
                raise ValueError("OverlandKinematic is a top-face (z_upper) BC")
        self._build()

    def _build(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        # permeability ghosts stay zero (vector.c:321 ctalloc), the other
        # fields repeat their edge value
        self.g_kx = _ghost3(self.kx, nx, ny, nz, fill_edge=False)
        self.g_ky = _ghost3(self.ky, nx, ny, nz, fill_edge=False)
        self.g_kz = _ghost3(self.kz, nx, ny, nz, fill_edge=False)
        self.g_poro = _ghost3(self.poro, nx, ny, nz)
        self.g_ss = _ghost3(self.ss, nx, ny, nz)
        self.g_alpha = _ghost3(self.alpha, nx, ny, nz)
        self.g_nvg = _ghost3(self.nvg, nx, ny, nz)
        self.g_sres = _ghost3(self.sres, nx, ny, nz)
        self.g_ssat = _ghost3(self.ssat, nx, ny, nz)
        self.g_slope_x = _ghost2(self.slope_x, nx, ny)
        self.g_slope_y = _ghost2(self.slope_y, nx, ny)
        self.g_mann = _ghost2(self.mann, nx, ny)
        self.uniform_sres = bool(np.ndim(self.sres) == 0)
        self.uniform_ssat = bool(np.ndim(self.ssat) == 0)

    # ------------------------------------------------------------------
    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def vg_table_params(self):
        """(method 0 spline / 1 linear, points, min head, alpha, n) or None."""
        if self.vg_table is None:
            return None
        pts, head, method = self.vg_table
        alpha = float(np.asarray(self.alpha, dtype=np.float64).flat[0])
        n = float(np.asarray(self.nvg, dtype=np.float64).flat[0])
        return (0 if method == "Spline" else 1, int(pts), float(head), alpha, n)

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    @property
    def overland(self):
        return self.bc[5].type == BC_OVERLAND

    def zeros(self):
        return np.zeros(self.shape, dtype=np.float64)

    def interior(self, a):
        return np.asarray(a).reshape(self.shape)[1:-1, 1:-1, 1:-1]

    def face_ncells(self, f):
        if f < 2:
            return self.ny * self.nz
        if f < 4:
            return self.nx * self.nz
        return self.nx * self.ny

    # -- boundary values (problem_bc_pressure.c:143-290, 582-612) -------
    def interval_number(self, b: BCSpec, time: float) -> int:
        """TimeCycleDataComputeIntervalNumber (time_cycle_data.c:152-197)."""
        cycle = int(sum(b.lengths))
        bu = self.base_unit
        if b.repeat < 0 or time < b.repeat * cycle * bu:
            dt_ = int(np.rint(time / (bu / 1000))) // 1000
            ds_ = int(np.rint(self.start_time / (bu / 1000))) // 1000
            completed = (dt_ - ds_ - 1)
            # C remainder truncates toward zero
            completed = int(np.fmod(completed, cycle))
            iv, total = 0, b.lengths[0]
            while completed >= total:
                iv += 1
                total += b.lengths[iv]
            return iv
        return len(b.lengths) - 1

    def bc_values(self, time: float):
        """Per-face value arrays at ``time``: list of 6 float64 arrays."""
        out = []
        for f, b in enumerate(self.bc):
            n = self.face_ncells(f)
            val = float(b.values[self.interval_number(b, time)])
            if b.type != BC_DIR_EQUIL:
                out.append(np.full(n, val))
                continue
            if f < 2:
                k = np.repeat(np.arange(self.nz), self.ny)
            elif f < 4:
                k = np.repeat(np.arange(self.nz), self.nx)
            else:
                k = np.full(n, 0 if f == 4 else self.nz - 1)
            z = self.z0 + (k + 0.5) * self.dz
            if f == 4:
                z = z - 0.5 * self.dz
            if f == 5:
                z = z + 0.5 * self.dz
            out.append(val - self.rho * self.gravity * (z - b.z_ref))
        return out

    def hydrostatic(self, value: float, z_ref: float):
        """HydroStaticPatch initial pressure (problem_ic_phase_pressure.c
        :440-673 with constant density)."""
        p = self.zeros()
        zc = self.z0 + (np.arange(self.nz) + 0.5) * self.dz
        prof = value - self.rho * self.gravity * (zc - z_ref)
        p[1:-1, 1:-1, 1:-1] = prof[:, None, None]
        return p


def with_solver(opts: Optional[SolverOptions] = None, **kw) -> SolverOptions:
    opts = dataclasses.replace(opts or SolverOptions(), **kw)
    return opts
