"""ParFlow input database (.pfidb) and decks built from it.

``read_pfidb`` / ``write_pfidb`` handle the key/value file ParFlow's driver
reads at start (pfsimulator/parflow_lib/input_database.c:101-131; written by
pftools/python/parflow/tools/io.py:1322-1331): the number of entries, then
per entry the key length, the key, the value length and the value, each
followed by a newline.

``problem_from_pfidb`` maps the keys the Richards Newton-Krylov path
consumes to ``(Problem, SolverOptions, p0, dt, nsteps)`` for a box domain:

* ``ComputationalGrid.*``; ``GeomInput.*`` Box geometries and their
  ``Patches``; ``Domain.GeomName``;
* ``Perm`` (Constant or PFBFile per geometry, ``TensorByGeom`` factors),
  ``Porosity``, ``SpecificStorage``, van Genuchten ``Phase.Saturation`` /
  ``Phase.RelPerm`` (region constants or ``VanGenuchten.File`` fields; the
  region kr lookup tables ``NumSamplePoints`` / ``MinPressureHead`` /
  ``InterpolationMethod``),
  constant density, viscosity and gravity;
* ``TimingInfo`` with a constant ``TimeStep``; ``Cycle.*``;
  ``BCPressure`` FluxConst, DirEquilRefPatch and OverlandKinematic;
  ``ICPressure`` HydroStaticPatch or PFBFile;
* ``TopoSlopesX/Y`` and ``Mannings`` (Constant per geometry or PFBFile) with
  ``Solver.OverlandKinematic.Epsilon``;
* the ``Solver.Nonlinear.*`` / ``Solver.Linear.*`` keys of
  pfsimulator/parflow_lib/kinsol_nonlin_solver.c:766-963 and pf_pfmg.c.

A property given per geometry is assigned to the cells whose centres lie in
the geometry's box, geometry by geometry in the order of its ``GeomNames``
list, so a later geometry overrides an earlier one -- the order of
ParFlow's geometry-based property modules.  Keys that select features
outside this path (non-box geometries, wells, sources, terrain-following
grids, variable dz, other BC or curve types, a full-Jacobian
preconditioner matrix) raise ``ValueError``.

``advance_deck`` runs a deck on the device and writes ParFlow-named PFB
outputs (``<run>.out.press.NNNNN.pfb``, ``<run>.out.satur.NNNNN.pfb`` and,
with ``Solver.PrintVelocities``, ``<run>.out.velx/vely/velz.NNNNN.pfb``).
"""
from __future__ import annotations

import os

import numpy as np

from .pfb import read_pfb, write_pfb
from .problem import (BC_DIR_EQUIL, BC_FLUX, BC_OVERLAND, BCSpec, Problem, TimeControl,
                      SolverOptions)


# ---------------------------------------------------------------------------
# file format
# ---------------------------------------------------------------------------
def read_pfidb(path):
    """The key -> value (str) dict of a .pfidb file."""
    with open(path, "r") as f:
        text = f.read()
    pos = 0

    def line():
        nonlocal pos
        end = text.index("\n", pos)
        s = text[pos:end]
        pos = end + 1
        return s

    def chars(n):
        nonlocal pos
        s = text[pos:pos + n]
        if len(s) != n:
            raise ValueError(f"{path}: truncated entry")
        pos += n
        if pos < len(text) and text[pos] == "\n":
            pos += 1
        return s

    db = {}
    n = int(line().strip())
    for _ in range(n):
        klen = int(line().strip())
        key = chars(klen)
        vlen = int(line().strip())
        db[key] = chars(vlen)
    return db


def write_pfidb(db, path):
    """Writes a key -> value dict as a .pfidb file (values via str())."""
    with open(path, "w") as f:
        f.write(f"{len(db)}\n")
        for key, value in db.items():
            v = _str(value)
            f.write(f"{len(key)}\n{key}\n{len(v)}\n{v}\n")


def _str(v):
    if isinstance(v, bool):
        return "True" if v else "False"
    return str(v)


# ---------------------------------------------------------------------------
# typed access
# ---------------------------------------------------------------------------
class _Keys:
    def __init__(self, db):
        self.db = dict(db)

    def has(self, key):
        return key in self.db

    def s(self, key, default=None):
        if key in self.db:
            return self.db[key].strip()
        if default is None:
            raise ValueError(f"pfidb: missing key {key}")
        return default

    def f(self, key, default=None):
        return float(self.s(key, None if default is None else str(default)))

    def i(self, key, default=None):
        return int(float(self.s(key, None if default is None else str(default))))

    def b(self, key, default=False):
        v = self.s(key, "True" if default else "False").lower()
        if v in ("true", "1", "yes"):
            return True
        if v in ("false", "0", "no"):
            return False
        raise ValueError(f"pfidb: {key} = {v!r} is not a boolean")

    def names(self, key):
        return self.db[key].split() if key in self.db else []


class _Grid:
    def __init__(self, K):
        self.nx = K.i("ComputationalGrid.NX")
        self.ny = K.i("ComputationalGrid.NY")
        self.nz = K.i("ComputationalGrid.NZ")
        self.dx = K.f("ComputationalGrid.DX")
        self.dy = K.f("ComputationalGrid.DY")
        self.dz = K.f("ComputationalGrid.DZ")
        self.x0 = K.f("ComputationalGrid.Lower.X")
        self.y0 = K.f("ComputationalGrid.Lower.Y")
        self.z0 = K.f("ComputationalGrid.Lower.Z")
        self.xc = self.x0 + (np.arange(self.nx) + 0.5) * self.dx
        self.yc = self.y0 + (np.arange(self.ny) + 0.5) * self.dy
        self.zc = self.z0 + (np.arange(self.nz) + 0.5) * self.dz


class _Geoms:
    """Box geometries: name -> (lower, upper), plus patch names."""

    def __init__(self, K):
        self.box = {}
        self.patches = {}
        for inp in K.names("GeomInput.Names"):
            kind = K.s(f"GeomInput.{inp}.InputType")
            name = K.s(f"GeomInput.{inp}.GeomName")
            if kind != "Box":
                raise ValueError(f"pfidb: geometry input {inp} is {kind}; only Box is supported")
            lo = [K.f(f"Geom.{name}.Lower.{a}") for a in "XYZ"]
            up = [K.f(f"Geom.{name}.Upper.{a}") for a in "XYZ"]
            self.box[name] = (lo, up)
            if K.has(f"Geom.{name}.Patches"):
                self.patches[name] = K.names(f"Geom.{name}.Patches")

    def mask3(self, name, G):
        if name not in self.box:
            raise ValueError(f"pfidb: unknown geometry {name}")
        lo, up = self.box[name]
        mx = (G.xc >= lo[0]) & (G.xc <= up[0])
        my = (G.yc >= lo[1]) & (G.yc <= up[1])
        mz = (G.zc >= lo[2]) & (G.zc <= up[2])
        return mz[:, None, None] & my[None, :, None] & mx[None, None, :]

    def mask2(self, name, G):
        return self.mask3(name, G).any(axis=0)


def _read_field(K, base, key, G, two_d=False):
    path = K.s(key)
    if not os.path.isabs(path):
        path = os.path.join(base, path)
    _, a = read_pfb(path)
    want = (1, G.ny, G.nx) if two_d else (G.nz, G.ny, G.nx)
    if a.shape != want:
        raise ValueError(f"pfidb: {path} has shape {a.shape}, expected {want}")
    return a[0] if two_d else a


def _by_geom(K, geoms, G, names_key, value_key, file_type_key=None, file_key=None,
             base=".", init=0.0):
    """A cell field assigned geometry by geometry (later names override)."""
    out = np.full((G.nz, G.ny, G.nx), init, dtype=np.float64)
    for g in K.names(names_key):
        m = geoms.mask3(g, G)
        kind = K.s(file_type_key.format(g=g), "Constant") if file_type_key else "Constant"
        if kind == "Constant":
            out[m] = K.f(value_key.format(g=g))
        elif kind == "PFBFile":
            out[m] = _read_field(K, base, file_key.format(g=g), G)[m]
        else:
            raise ValueError(f"pfidb: {file_type_key.format(g=g)} = {kind} is not supported")
    return out


def _surface(K, geoms, G, prefix, base):
    kind = K.s(f"{prefix}.Type", "Constant")
    if kind == "PFBFile":
        return _read_field(K, base, f"{prefix}.FileName", G, two_d=True)
    if kind != "Constant":
        raise ValueError(f"pfidb: {prefix}.Type = {kind} is not supported")
    out = np.zeros((G.ny, G.nx), dtype=np.float64)
    for g in K.names(f"{prefix}.GeomNames"):
        out[geoms.mask2(g, G)] = K.f(f"{prefix}.Geom.{g}.Value")
    return out


def _patch_elevation(geoms, geom, patch):
    """Elevation of a flat z patch of a box geometry (DirEquilRefPatch and
    HydroStaticPatch references)."""
    names = geoms.patches.get(geom)
    if not names or patch not in names:
        raise ValueError(f"pfidb: geometry {geom} has no patch {patch}")
    face = names.index(patch)
    lo, up = geoms.box[geom]
    if face == 4:
        return lo[2]
    if face == 5:
        return up[2]
    raise ValueError(f"pfidb: reference patch {patch} is not a z face of {geom}")


def _unique(a):
    """A region-constant field as its scalar, else the cell array."""
    v = np.unique(a)
    return float(v[0]) if v.size == 1 else a


# ---------------------------------------------------------------------------
# deck
# ---------------------------------------------------------------------------
_ETA = {"EtaConstant": 0, "Walker1": 1, "Walker2": 2}
_GLOB = {"InexactNewton": 0, "LineSearch": 1}


def time_control_from_pfidb(db):
    """TimeControl of a key -> value dict (or a .pfidb path): TimingInfo
    (problem.c:87-98), TimeStep Constant or Growth
    (select_time_step.c:200-245), Solver.MaxIter and the
    Solver.SurfacePredictor keys (solver_richards.c:6959-6976)."""
    if isinstance(db, str):
        db = read_pfidb(db)
    K = _Keys(db)
    ts = K.s("TimeStep.Type", "Constant")
    if ts not in ("Constant", "Growth"):
        raise ValueError(f"pfidb: TimeStep.Type {ts} is not supported")
    kw = dict(start_time=K.f("TimingInfo.StartTime", 0.0),
              stop_time=K.f("TimingInfo.StopTime"),
              base_unit=K.f("TimingInfo.BaseUnit"),
              dump_interval=K.f("TimingInfo.DumpInterval", -1.0),
              max_iterations=K.i("Solver.MaxIter", 1000000), ts_type=ts)
    kw.update(surface_predictor=K.b("Solver.SurfacePredictor"),
              surface_predictor_pressure=K.f("Solver.SurfacePredictor.PressureValue", 0.00001),
              surface_lateral_flows=K.b("Solver.SurfacePredictor.LateralFlows"))
    if ts == "Constant":
        kw["value"] = K.f("TimeStep.Value")
    else:
        kw.update(initial_step=K.f("TimeStep.InitialStep"),
                  growth_factor=K.f("TimeStep.GrowthFactor"),
                  min_step=K.f("TimeStep.MinStep"), max_step=K.f("TimeStep.MaxStep"))
    return TimeControl(**kw)


def problem_from_pfidb(db, base_dir=".", name=None):
    """(Problem, SolverOptions, p0, dt, nsteps) of a key -> value dict (or a
    .pfidb path); file names in the keys are relative to base_dir."""
    if isinstance(db, str):
        base_dir = os.path.dirname(os.path.abspath(db))
        db = read_pfidb(db)
    K = _Keys(db)
    # features outside the path
    if K.b("Solver.TerrainFollowingGrid"):
        raise ValueError("pfidb: terrain-following grids are not supported")
    if K.b("Solver.Nonlinear.VariableDz"):
        raise ValueError("pfidb: variable dz is not supported")
    if K.names("Wells.Names"):
        raise ValueError("pfidb: wells are not supported")
    if K.s("Solver", "Richards") != "Richards":
        raise ValueError("pfidb: only the Richards solver is supported")
    pc_matrix = K.s("Solver.Linear.Preconditioner.PCMatrixType", "PFSymmetric")
    if pc_matrix not in ("PFSymmetric", "FullJacobian"):
        raise ValueError(f"pfidb: PCMatrixType {pc_matrix} is not supported")
    G = _Grid(K)
    geoms = _Geoms(K)
    dom = K.s("Domain.GeomName")
    if dom not in geoms.box:
        raise ValueError(f"pfidb: domain geometry {dom} is not a Box input")
    faces = geoms.patches.get(dom)
    if not faces or len(faces) != 6:
        raise ValueError(f"pfidb: domain {dom} needs six patches")

    # sources must vanish
    for g in K.names("PhaseSources.water.GeomNames"):
        if K.f(f"PhaseSources.water.Geom.{g}.Value", 0.0) != 0.0:
            raise ValueError("pfidb: phase sources are not supported")

    # permeability
    perm = _by_geom(K, geoms, G, "Geom.Perm.Names", "Geom.{g}.Perm.Value",
                    "Geom.{g}.Perm.Type", "Geom.{g}.Perm.FileName", base_dir)
    ttype = K.s("Perm.TensorType", "TensorByGeom")
    if ttype != "TensorByGeom":
        raise ValueError(f"pfidb: Perm.TensorType {ttype} is not supported")
    tens = []
    for ax in "XYZ":
        t = _by_geom(K, geoms, G, "Geom.Perm.TensorByGeom.Names",
                     "Geom.{g}.Perm.TensorVal" + ax, init=1.0)
        tens.append(t)
    kx, ky, kz = (perm * t for t in tens)
    poro = _by_geom(K, geoms, G, "Geom.Porosity.GeomNames", "Geom.{g}.Porosity.Value",
                    "Geom.{g}.Porosity.Type", "Geom.{g}.Porosity.FileName", base_dir)
    ss = _by_geom(K, geoms, G, "SpecificStorage.GeomNames",
                  "Geom.{g}.SpecificStorage.Value")

    # van Genuchten closures: one (alpha, n) serves S and kr here
    for key in ("Phase.Saturation.Type", "Phase.RelPerm.Type"):
        if K.s(key) != "VanGenuchten":
            raise ValueError(f"pfidb: {key} = {K.s(key)} is not supported")
    sfile = K.b("Phase.Saturation.VanGenuchten.File")
    rfile = K.b("Phase.RelPerm.VanGenuchten.File")

    def vg(prefix, param, use_file):
        names = f"Phase.{prefix}.GeomNames"
        if use_file:
            out = np.zeros((G.nz, G.ny, G.nx))
            for g in K.names(names):
                m = geoms.mask3(g, G)
                out[m] = _read_field(K, base_dir, f"Geom.{g}.{prefix}.{param}.Filename", G)[m]
            return out
        return _by_geom(K, geoms, G, names, "Geom.{g}." + prefix + "." + param)

    alpha, nvg = vg("Saturation", "Alpha", sfile), vg("Saturation", "N", sfile)
    sres, ssat = vg("Saturation", "SRes", sfile), vg("Saturation", "SSat", sfile)
    ra, rn = vg("RelPerm", "Alpha", rfile), vg("RelPerm", "N", rfile)
    if not (np.array_equal(ra, alpha) and np.array_equal(rn, nvg)):
        raise ValueError("pfidb: saturation and relative permeability need the same "
                         "van Genuchten alpha and n")
    # kr lookup tables (problem_phase_rel_perm.c:1894-1915): every RelPerm
    # region must ask for the same table, from a uniform alpha and n
    vg_table = None
    if not rfile:
        specs = set()
        for g in K.names("Phase.RelPerm.GeomNames"):
            pts = K.i(f"Geom.{g}.RelPerm.NumSamplePoints", 0)
            if pts:
                specs.add((pts, K.f(f"Geom.{g}.RelPerm.MinPressureHead"),
                           K.s(f"Geom.{g}.RelPerm.InterpolationMethod", "Spline")))
            else:
                specs.add(None)
        if specs != {None}:
            if len(specs) != 1:
                raise ValueError("pfidb: RelPerm lookup tables must be the same in every "
                                 "region (and present in all or none)")
            vg_table = specs.pop()

    rho = K.f("Phase.water.Density.Value", 1.0)
    mu = K.f("Phase.water.Viscosity.Value", 1.0)
    grav = K.f("Gravity", 1.0)

    # time
    tc = time_control_from_pfidb(db)
    base_unit, t0 = tc.base_unit, tc.start_time
    if tc.ts_type == "Constant":
        dt = tc.value
        # Solver.MaxIter caps the number of time steps (solver_richards.c)
        nsteps = min(int(round((tc.stop_time - t0) / dt)), tc.max_iterations)
    else:
        dt, nsteps = tc.initial_step, tc.max_iterations

    # cycles and boundary conditions
    cycles = {}
    for c in K.names("Cycle.Names"):
        ivs = K.names(f"Cycle.{c}.Names")
        cycles[c] = (ivs, [K.i(f"Cycle.{c}.{iv}.Length") for iv in ivs],
                     K.i(f"Cycle.{c}.Repeat", -1))
    bc = [None] * 6
    for p in K.names("BCPressure.PatchNames"):
        if p not in faces:
            raise ValueError(f"pfidb: patch {p} is not a face of {dom}")
        face = faces.index(p)
        kind = K.s(f"Patch.{p}.BCPressure.Type")
        ivs, lengths, repeat = cycles[K.s(f"Patch.{p}.BCPressure.Cycle")]
        vals = tuple(K.f(f"Patch.{p}.BCPressure.{iv}.Value") for iv in ivs)
        if kind == "FluxConst":
            bc[face] = BCSpec(BC_FLUX, vals, tuple(lengths), repeat)
        elif kind == "DirEquilRefPatch":
            zr = _patch_elevation(geoms, K.s(f"Patch.{p}.BCPressure.RefGeom"),
                                  K.s(f"Patch.{p}.BCPressure.RefPatch"))
            bc[face] = BCSpec(BC_DIR_EQUIL, vals, tuple(lengths), repeat, z_ref=zr)
        elif kind == "OverlandKinematic":
            if face != 5:
                raise ValueError("pfidb: OverlandKinematic only on the top face")
            bc[face] = BCSpec(BC_OVERLAND, vals, tuple(lengths), repeat)
        else:
            raise ValueError(f"pfidb: Patch.{p}.BCPressure.Type {kind} is not supported")
    if any(b is None for b in bc):
        raise ValueError("pfidb: every face of the domain needs a BCPressure")
    overland = bc[5].type == BC_OVERLAND
    slope_x = _surface(K, geoms, G, "TopoSlopesX", base_dir) if overland else 0.0
    slope_y = _surface(K, geoms, G, "TopoSlopesY", base_dir) if overland else 0.0
    mann = _surface(K, geoms, G, "Mannings", base_dir) if overland else 1.0

    prob = Problem(nx=G.nx, ny=G.ny, nz=G.nz, dx=G.dx, dy=G.dy, dz=G.dz,
                   x0=G.x0, y0=G.y0, z0=G.z0, rho=rho, mu=mu, gravity=grav,
                   kx=_unique(kx), ky=_unique(ky), kz=_unique(kz),
                   poro=_unique(poro), ss=_unique(ss),
                   alpha=_unique(alpha), nvg=_unique(nvg),
                   sres=_unique(sres), ssat=_unique(ssat),
                   slope_x=slope_x, slope_y=slope_y,
                   mann=_unique(mann) if overland else 1.0,
                   ov_eps=K.f("Solver.OverlandKinematic.Epsilon", 1e-5), bc=bc,
                   base_unit=base_unit, start_time=t0, vg_table=vg_table,
                   name=name or "pfidb")

    # initial pressure
    ic = K.s("ICPressure.Type")
    if ic == "HydroStaticPatch":
        gs = K.names("ICPressure.GeomNames")
        if len(gs) != 1:
            raise ValueError("pfidb: HydroStaticPatch over one geometry only")
        g = gs[0]
        zr = _patch_elevation(geoms, K.s(f"Geom.{g}.ICPressure.RefGeom"),
                              K.s(f"Geom.{g}.ICPressure.RefPatch"))
        p0 = prob.hydrostatic(K.f(f"Geom.{g}.ICPressure.Value"), zr)
    elif ic == "PFBFile":
        p0 = prob.zeros()
        p0[1:-1, 1:-1, 1:-1] = _read_field(K, base_dir, f"Geom.{dom}.ICPressure.FileName", G)
    else:
        raise ValueError(f"pfidb: ICPressure.Type {ic} is not supported")

    # solver
    pc = K.s("Solver.Linear.Preconditioner", "MGSemi")
    if pc not in ("PFMG", "MGSemi", "NoPC"):
        raise ValueError(f"pfidb: preconditioner {pc} is not supported")
    opts = SolverOptions(
        residual_tol=K.f("Solver.Nonlinear.ResidualTol", 1e-7),
        step_tol=K.f("Solver.Nonlinear.StepTol", 1e-7),
        max_iter=K.i("Solver.Nonlinear.MaxIter", 15),
        krylov_dim=K.i("Solver.Linear.KrylovDimension", 10),
        max_restarts=K.i("Solver.Linear.MaxRestarts", 0),
        eta_choice=_ETA[K.s("Solver.Nonlinear.EtaChoice", "Walker2")],
        eta_value=K.f("Solver.Nonlinear.EtaValue", 1e-4),
        eta_alpha=K.f("Solver.Nonlinear.EtaAlpha", 2.0),
        eta_gamma=K.f("Solver.Nonlinear.EtaGamma", 0.9),
        use_jacobian=K.b("Solver.Nonlinear.UseJacobian"),
        deriv_eps=K.f("Solver.Nonlinear.DerivativeEpsilon", 1e-7),
        globalization=_GLOB[K.s("Solver.Nonlinear.Globalization", "LineSearch")],
        precond=pc,
        num_pre_relax=K.i("Solver.Linear.Preconditioner.PFMG.NumPreRelax", 1),
        num_post_relax=K.i("Solver.Linear.Preconditioner.PFMG.NumPostRelax", 1),
        max_levels=K.i("Solver.Linear.Preconditioner.MGSemi.MaxLevels", 100),
        max_convergence_failures=K.i("Solver.MaxConvergenceFailures", 3),
        pc_matrix=pc_matrix)
    return prob, opts, p0, dt, nsteps


def advance_deck(path, out_dir, run_name=None, nsteps=None, precision=64, krylov=0):
    """Runs a .pfidb deck on the device (plain form by default) through the
    AdvanceRichards time loop (TimeStep Constant or Growth, BC-cycle, dump
    and stop-time cuts) and writes ParFlow-named PFB outputs: file 00000 the
    initial state, then one file number per dump step
    (pfsimulator/parflow_lib/solver_richards.c:3141-3176, 4603-4612):
    <run>.out.press.NNNNN.pfb and .satur.NNNNN.pfb and, with
    ``Solver.PrintVelocities`` True, <run>.out.velx/vely/velz.NNNNN.pfb --
    the face velocities of the step's last residual evaluation, i.e. at the
    step's pressure (:1669-1700; zero in file 00000, :1034-1044).
    ``nsteps`` caps the steps taken.  Returns the per-step log rows."""
    from . import api
    db = read_pfidb(path)
    prob, opts, p0, dt, n = problem_from_pfidb(db, os.path.dirname(os.path.abspath(path)))
    tc = time_control_from_pfidb(db)
    print_vel = _Keys(db).b("Solver.PrintVelocities")
    run_name = run_name or os.path.splitext(os.path.basename(path))[0]
    os.makedirs(out_dir, exist_ok=True)
    gp = api.GPUProblem(prob)
    sv = api.GPUSolver(gp, opts, precision, krylov)
    ph, sh, rows = sv.run(p0, tc, max_steps=nsteps)
    kw = dict(x=prob.x0, y=prob.y0, z=prob.z0, dx=prob.dx, dy=prob.dy, dz=prob.dz)
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    file_number = 0
    for s in range(len(rows) + 1):
        if s > 0 and not rows[s - 1]["dumped"]:
            continue
        write_pfb(os.path.join(out_dir, f"{run_name}.out.press.{file_number:05d}.pfb"),
                  prob.interior(ph[s]), **kw)
        write_pfb(os.path.join(out_dir, f"{run_name}.out.satur.{file_number:05d}.pfb"),
                  prob.interior(sh[s]), **kw)
        if print_vel:
            if s == 0:
                vel = [np.zeros((nz, ny, nx + 1)), np.zeros((nz, ny + 1, nx)),
                       np.zeros((nz + 1, ny, nx))]
            else:
                r = rows[s - 1]
                vel = gp.face_velocity(ph[s], ph[s - 1], r["dt"], r["time"])
            for c, v in zip("xyz", vel):
                write_pfb(os.path.join(out_dir, f"{run_name}.out.vel{c}.{file_number:05d}.pfb"),
                          v, **kw)
        file_number += 1
    return rows
