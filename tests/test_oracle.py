"""CPU tests: the oracle against the reference's own golden outputs and
against what the mathematics fixes (closed forms, finite differences,
invariants, brute force on tiny systems)."""
import dataclasses
import os

import numpy as np
import pytest

from oracle import oracle as O
import cases
import ref_decks

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _build_oracle():
    O.build()


def _rel(a, b, floor=0.0):
    return np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor))


# --- reference goldens ---------------------------------------------------------
def test_default_richards_golden():
    """test/python/default_richards.py: 50 steps, DirEquilRefPatch sides.

    The reference solved with hypre PFMG; both runs stop at ResidualTol 1e-9,
    so they agree to the tolerance-limited level (<= 6e-7 relative), i.e.
    the reference test's own 6-significant-digit bar up to its last digit."""
    prob, opts, p0, dt, n = ref_decks.default_richards()
    op = O.OracleProblem(prob)
    r, ph, sh, log = O.advance(op, opts, p0, 0.0, dt, n)
    assert r == 0
    g = np.load(os.path.join(HERE, "golden", "default_richards.npz"))
    for i, s in enumerate(g["steps"]):
        p = prob.interior(ph[s])
        q = prob.interior(sh[s])
        assert _rel(p, g["press"][i]) < 6e-7, s
        assert _rel(q, g["satur"][i]) < 5e-7, s
    # "This run ... should take 3 nonlinear iterations" (default_richards.py:3-4)
    assert log.newton[0] == 3


def test_tiltedv_overlandkin_golden():
    """test/python/overland_tiltedV_KWE.py run TiltedV_OverlandKin: kinematic
    overland flow with a rain/recession cycle; dumps every 2 steps
    (DumpInterval -2, dump k after step 2k-1).  Ponded depths are 1e-5..1e-3
    m against an absolute residual tolerance of 1e-9, so agreement is
    tolerance limited (<= 5e-6 relative)."""
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    op = O.OracleProblem(prob)
    r, ph, sh, log = O.advance(op, opts, p0, 0.0, dt, n)
    assert r == 0
    g = np.load(os.path.join(HERE, "golden", "tiltedv_overlandkin.npz"))
    for d in range(21):
        s = max(2 * d - 1, 0)
        assert _rel(prob.interior(ph[s]), g["press"][d]) < 5e-6, d
        assert _rel(prob.interior(sh[s]), g["satur"][d]) < 1e-12, d


def test_tiltedv_overlandkin_jacobian_true_golden():
    """The same deck's "Jacobian True" variant (overland_tiltedV_KWE.py
    :511-545: UseJacobian True, PFSymmetric preconditioner), checked against
    the same correct_output: the analytic J.v includes the kinematic surface
    couplings."""
    import dataclasses
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    opts = dataclasses.replace(opts, use_jacobian=True)
    op = O.OracleProblem(prob)
    r, ph, sh, log = O.advance(op, opts, p0, 0.0, dt, n)
    assert r == 0
    g = np.load(os.path.join(HERE, "golden", "tiltedv_overlandkin.npz"))
    for d in range(21):
        s = max(2 * d - 1, 0)
        assert _rel(prob.interior(ph[s]), g["press"][d], 1e-3) < 1e-7, d


def test_tiltedv_overlandkin_jacobian_true_nonsymmetric_pc_golden():
    """The "Jacobian True Nonsymmetric Preconditioner" variant
    (overland_tiltedV_KWE.py:417-440: UseJacobian True, PCMatrixType
    FullJacobian), checked against the same correct_output: the
    preconditioner matrix is the full Jacobian with the kinematic surface
    couplings (PL/pf_hypre.c:355-520) in a full-storage hierarchy."""
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    opts = dataclasses.replace(opts, use_jacobian=True, pc_matrix="FullJacobian")
    op = O.OracleProblem(prob)
    r, ph, sh, log = O.advance(op, opts, p0, 0.0, dt, n)
    assert r == 0
    g = np.load(os.path.join(HERE, "golden", "tiltedv_overlandkin.npz"))
    for d in range(21):
        s = max(2 * d - 1, 0)
        assert _rel(prob.interior(ph[s]), g["press"][d], 1e-3) < 1e-7, d


# --- closures --------------------------------------------------------------------
def test_vg_closed_forms():
    prob = cases.hetero(4, 3, 3, seed=3)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=4)
    S = O.saturation(op, p)
    kr = O.relperm(op, p)
    a, n = prob.g_alpha, prob.g_nvg
    m = 1 - 1 / n
    h = np.abs(p)
    se = 1.0 / (1 + (a * h) ** n) ** m
    S_ref = np.where(p >= 0, 1.0, 0.8 * se + 0.2)
    kr_ref = np.where(p >= 0, 1.0, np.sqrt(se) * (1 - (1 - se ** (1 / m)) ** m) ** 2)
    I = prob.interior
    assert np.allclose(I(S), I(S_ref), rtol=1e-13, atol=0)
    # the reference's kr expression equals the Mualem form sqrt(Se)(1-(1-Se^(1/m))^m)^2
    assert np.allclose(I(kr), I(kr_ref), rtol=1e-10, atol=1e-14)


def test_vg_derivatives_fd():
    prob = cases.hetero(4, 3, 3, seed=5)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=6) - 0.5
    eps = 1e-6
    for fcn_fun in (O.saturation, O.relperm):
        d = fcn_fun(op, p, 1)
        fd = (fcn_fun(op, p + eps) - fcn_fun(op, p - eps)) / (2 * eps)
        I = prob.interior
        assert np.allclose(I(d), I(fd), rtol=1e-5, atol=1e-9)


# --- residual / Jacobian -----------------------------------------------------------
def test_hydrostatic_equilibrium_zero_residual():
    """test/python/richards_hydrostatic_equilibrium.py idea: a hydrostatic
    state with no-flux sides and a consistent Dirichlet bottom is a
    steady solution, so F(p) = 0 with p = p_old."""
    prob = cases.hetero(6, 5, 7, seed=7)
    from parflow_amd.problem import BCSpec, BC_FLUX
    prob.bc[5] = BCSpec(BC_FLUX, (0.0,))
    op = O.OracleProblem(prob)
    p = prob.hydrostatic(0.0, 0.0)
    s = O.saturation(op, p)
    F = O.nl_function_eval(op, p, p, s, 0.5, 0.5)
    assert np.max(np.abs(prob.interior(F))) < 1e-14


def _phys(prob, rho, gravity, mu):
    return dataclasses.replace(prob, rho=rho, gravity=gravity, mu=mu)


def test_hydrostatic_equilibrium_zero_residual_with_density_and_gravity():
    """With rho g != 1 (PhaseDensity constant, PL/problem_phase_density.c:119;
    Gravity, PL/problem.c:152) the hydrostatic profile p = -rho g (z - z_ref),
    with the bottom DirEquilRefPatch value referenced the same way, is a
    steady state: the z face's lower/upper conditions
    p/dz -+ rho g / 2 (PL/nl_function_eval.c:817-836) balance exactly and
    F = 0.  The same profile built with g alone (rho dropped) is not."""
    from parflow_amd.problem import BCSpec, BC_FLUX
    base = cases.hetero(6, 5, 7, seed=7)
    base.bc[5] = BCSpec(BC_FLUX, (0.0,))
    prob = _phys(base, 1.7, 2.3, 0.6)
    op = O.OracleProblem(prob)
    p = prob.hydrostatic(0.0, 0.0)
    s = O.saturation(op, p)
    F = O.nl_function_eval(op, p, p, s, 0.5, 0.5)
    assert np.max(np.abs(prob.interior(F))) < 1e-13
    wrong = _phys(base, 1.0, 2.3, 0.6).hydrostatic(0.0, 0.0)
    Fw = O.nl_function_eval(op, wrong, wrong, O.saturation(op, wrong), 0.5, 0.5)
    assert np.max(np.abs(prob.interior(Fw))) > 1e-3


def _zero_flux_tops(prob):
    from parflow_amd.problem import BCSpec, BC_FLUX
    bc = [b if b.type != BC_FLUX else BCSpec(BC_FLUX, (0.0,)) for b in prob.bc]
    return dataclasses.replace(prob, bc=bc)


@pytest.mark.parametrize("dirs", [False, True])
def test_residual_and_jacobian_scale_as_inverse_viscosity(dirs):
    """Every face flux of the residual, interior and Dirichlet, divides by
    the viscosity (PL/nl_function_eval.c:767-836, :1018-1232) and the
    accumulation terms do not, while dt multiplies only the fluxes; with
    FluxConst values of zero (a flux value is volumetric, not divided by
    mu) F(p; mu, dt) = F(p; 1, dt / mu) and likewise J."""
    base = _zero_flux_tops(cases.hetero(6, 5, 4, seed=8, dir_sides=dirs))
    rng = np.random.default_rng(9)
    mu, dt = 0.6, 0.1
    pa, pb = _phys(base, 1.0, 1.0, mu), _phys(base, 1.0, 1.0, 1.0)
    oa, ob = O.OracleProblem(pa), O.OracleProblem(pb)
    p_old = cases.random_pressure(base, seed=10)
    s_old = O.saturation(oa, p_old)
    p = np.ascontiguousarray(p_old + rng.normal(0, 0.05, p_old.shape))
    Fa = O.nl_function_eval(oa, p, p_old, s_old, dt, dt)
    Fb = O.nl_function_eval(ob, p, p_old, s_old, dt / mu, dt)
    I = base.interior
    scale = np.max(np.abs(I(Fa)))
    assert np.max(np.abs(I(Fa) - I(Fb))) < 1e-13 * scale
    assert np.max(np.abs(I(Fa) - I(O.nl_function_eval(ob, p, p_old, s_old, dt, dt)))) > \
        1e-3 * scale
    Ja = O.jacobian_eval(oa, p, dt, dt, 0)
    Jb = O.jacobian_eval(ob, p, dt / mu, dt, 0)
    for q in range(7):
        assert np.max(np.abs(I(Ja[q]) - I(Jb[q]))) <= 1e-13 * np.max(np.abs(I(Ja[0]))), q


def test_residual_density_identities():
    """(rho, g) -> (c rho, g / c) leaves every head p / (rho g) and so S and kr
    unchanged (PL/problem_saturation.c:265-310).  In the reference the
    accumulation terms rho phi S vol and Ss vol p S rho and the vertical face
    flux K (kr rho) (p_a/dz - p_b/dz - rho g) / mu then scale by c, and the
    lateral face flux K (kr rho) g cos(atan(slope)) (p_a - p_b) / (dx mu)
    (PL/nl_function_eval.c:767-785, x_dir_g_c) does not.  So on a column
    (one cell in x and y: accumulation and vertical faces only)
    F(c rho, g / c, dt) = c F(rho, g, dt), and on a single layer with no-flux
    faces (accumulation and lateral faces only) F(c rho, g / c, c dt) =
    c F(rho, g, dt).  A kr that missed its density factor, a gravity term
    with g but not rho, or a lateral flux without the g factor fails one."""
    from parflow_amd.problem import BCSpec, BC_FLUX
    rho, g, c = 1.7, 2.3, 3.0
    col = _phys(_zero_flux_tops(cases.hetero(1, 1, 9, seed=11)), rho, g, 1.0)
    lay = cases.hetero(7, 6, 1, seed=12)
    lay = _phys(dataclasses.replace(lay, bc=[BCSpec(BC_FLUX, (0.0,))] * 6), rho, g, 1.0)
    for prob, dt_scale in ((col, 1.0), (lay, c)):
        scaled = _phys(prob, c * rho, g / c, 1.0)
        oa, ob = O.OracleProblem(prob), O.OracleProblem(scaled)
        p_old = cases.random_pressure(prob, seed=13)
        s_old = O.saturation(oa, p_old)
        assert np.max(np.abs(O.saturation(ob, p_old) - s_old)) < 1e-15
        p = np.ascontiguousarray(p_old + np.random.default_rng(14).normal(0, 0.2, p_old.shape))
        dt = 0.3
        Fa = O.nl_function_eval(oa, p, p_old, s_old, dt, dt)
        Fb = O.nl_function_eval(ob, p, p_old, s_old, dt * dt_scale, dt)
        I = prob.interior
        scale = np.max(np.abs(I(Fa)))
        assert scale > 0
        assert np.max(np.abs(c * I(Fa) - I(Fb))) < 1e-13 * c * scale, prob.shape
        # the fluxes are a sizeable part of F (the identity is not only the
        # accumulation's)
        F0 = O.nl_function_eval(oa, p, p_old, s_old, 0.0, dt)
        assert np.max(np.abs(I(Fa) - I(F0))) > 0.05 * scale


def test_jacobian_with_density_gravity_viscosity():
    """The Jacobian's rho, g and mu factors.

    The reference's closure derivatives are taken with respect to the head
    h = |p| / (rho g) (PL/problem_saturation.c:306-309,
    PL/problem_phase_rel_perm.c:900-912: no 1/(rho g) chain factor), so they
    equal rho g dS/dp.  With rho g = 1 but rho, g, mu != 1 separately the
    analytic J is therefore the derivative of F (central differences); with
    rho g != 1, (rho, g) -> (c rho, g / c) scales J by c on a column and
    J(dt) by c on a no-flux layer with dt -> c dt, as for F."""
    from parflow_amd.problem import BCSpec, BC_FLUX
    rng = np.random.default_rng(16)
    prob = _phys(cases.hetero(5, 4, 4, seed=15, dir_sides=True), 2.0, 0.5, 0.6)
    op = O.OracleProblem(prob)
    p = np.ascontiguousarray(prob.hydrostatic(-1.0, 2.0) + rng.normal(0, 0.05, prob.shape))
    s_old = O.saturation(op, p)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 0)
    I = prob.interior
    for trial in range(3):
        v = np.zeros_like(p)
        v[1:-1, 1:-1, 1:-1] = rng.normal(0, 1, (prob.nz, prob.ny, prob.nx))
        jv = O.matvec(op, J, 0, v)
        fd = _fd_jv(op, p, p, s_old, 0.1, 0.1, v)
        assert np.max(np.abs(I(jv) - I(fd))) / np.max(np.abs(I(fd))) < 1e-5
    # the derivative convention: dS/dp returned = rho g d S / d p
    pg = _phys(prob, 1.7, 2.3, 0.6)
    og = O.OracleProblem(pg)
    q = cases.random_pressure(pg, seed=17)
    eps = 1e-5
    fd = (O.saturation(og, q + eps) - O.saturation(og, q - eps)) / (2 * eps)
    d = O.saturation(og, q, 1)
    wet = I(q) < -1e-3
    assert np.allclose(I(d)[wet], 1.7 * 2.3 * I(fd)[wet], rtol=1e-6, atol=1e-12)
    fd = (O.relperm(og, q + eps) - O.relperm(og, q - eps)) / (2 * eps)
    d = O.relperm(og, q, 1)
    assert np.allclose(I(d)[wet], 1.7 * 2.3 * I(fd)[wet], rtol=1e-5, atol=1e-12)
    # (rho, g) -> (c rho, g / c)
    rho, g, c = 1.7, 2.3, 3.0
    col = _phys(_zero_flux_tops(cases.hetero(1, 1, 9, seed=11)), rho, g, 0.6)
    lay = cases.hetero(7, 6, 1, seed=12)
    lay = _phys(dataclasses.replace(lay, bc=[BCSpec(BC_FLUX, (0.0,))] * 6), rho, g, 0.6)
    for pr, dt_scale in ((col, 1.0), (lay, c)):
        oa, ob = O.OracleProblem(pr), O.OracleProblem(_phys(pr, c * rho, g / c, 0.6))
        q = np.ascontiguousarray(cases.random_pressure(pr, seed=18))
        Ja = O.jacobian_eval(oa, q, 0.3, 0.3, 0)
        Jb = O.jacobian_eval(ob, q, 0.3 * dt_scale, 0.3, 0)
        Ip = pr.interior
        for k in range(7):
            assert np.max(np.abs(c * Ip(Ja[k]) - Ip(Jb[k]))) <= \
                1e-13 * c * np.max(np.abs(Ip(Ja[0]))), (pr.shape, k)


def test_residual_mass_balance():
    """Interior fluxes cancel: sum(F) = storage change - dt * boundary inflow."""
    prob = cases.hetero(6, 5, 4, seed=8)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=9)
    s_old = O.saturation(op, p_old)
    p = p_old + 0.01
    dt = 0.1
    F = O.nl_function_eval(op, p, p_old, s_old, dt, dt)
    # rebuild with fluxes removed: K -> 0 leaves accumulation + BC flux terms
    s = O.saturation(op, p)
    vol = prob.dx * prob.dy * prob.dz
    I = prob.interior
    acc = (I(s) - I(s_old)) * I(prob.g_poro) * vol + I(prob.g_ss) * vol * (I(p) * I(s) - I(p_old) * I(s_old))
    # top flux BC -0.01 over the top face area enters the domain
    top_in = dt * prob.dx * prob.dy * prob.nx * prob.ny * (-0.01)
    # bottom Dirichlet flux: compute from F's residual balance instead
    bal = F[1:-1, 1:-1, 1:-1].sum() - acc.sum()
    # the balance equals dt*(sum of boundary outflows); top contributes top_in
    F_nobot = bal - top_in
    # bottom outflow computed independently from the Dirichlet face formula
    kr = O.relperm(op, p)
    pb = p[1, 1:-1, 1:-1]
    lc = 0.0 / prob.dz - 0.25
    uc = pb / prob.dz + 0.25
    diff = lc - uc
    krg = 1.0  # value 0 at the bottom face -> kr = 1
    up = np.where(lc - uc >= 0, krg, kr[1, 1:-1, 1:-1])
    u = up * prob.dx * prob.dy * prob.g_kz[1, 1:-1, 1:-1] * 2.0 * diff
    bottom = -dt * u.sum()
    assert abs(F_nobot - bottom) < 1e-10 * max(1.0, abs(bottom))


def test_default_richards_face_velocity_golden():
    """Face velocities computed from the reference's own step-n pressure
    reproduce its velx/vely/velz outputs of step n bit for bit
    (test/python/default_richards.py: DirEquilRefPatch x faces, FluxConst 0
    elsewhere).  KINSOL's last residual evaluation of a step is at the
    accepted iterate, so the velocities written after the step are a
    function of the step's output pressure (and BC values) alone."""
    prob, opts, p0, dt, n = ref_decks.default_richards()
    op = O.OracleProblem(prob)
    g = np.load(os.path.join(HERE, "golden", "default_richards.npz"))
    gv = np.load(os.path.join(HERE, "golden", "default_richards_vel.npz"))
    steps = [int(x) for x in g["steps"]]
    for i, s in enumerate(gv["steps"]):
        p = prob.zeros()
        prob.interior(p)[...] = g["press"][steps.index(int(s))]
        vx, vy, vz = O.face_velocity(op, p, p, dt, s * dt)
        assert np.array_equal(vx, gv["velx"][i]), s
        assert np.array_equal(vy, gv["vely"][i]), s
        assert np.array_equal(vz, gv["velz"][i]), s


@pytest.mark.parametrize("overland,dirs", [(False, False), (False, True), (True, True)])
def test_face_velocity_divergence_is_residual_flux(overland, dirs):
    """F = accumulation + dt * (net outflow through the cell's six faces),
    an outflow being velocity x face area: velocities and residual see the
    same face fluxes, on interior faces and on Dirichlet, flux and kinematic
    overland domain faces alike (PL/nl_function_eval.c:846-853, :1225-1230,
    :1428-1432, :2015-2023).  No reference output holds an overland-top
    velocity; this identity pins it to F, which the overland goldens pin."""
    # dx != dy so that a face area taken from the wrong direction shows
    prob = dataclasses.replace(cases.hetero(7, 6, 5, seed=31, overland=overland,
                                            dir_sides=dirs), dx=2.0, dy=1.5)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=32)
    s_old = O.saturation(op, p_old)
    p = p_old + np.random.default_rng(33).normal(0, 0.01, p_old.shape)
    dt, t = 0.1, 0.35
    F = O.nl_function_eval(op, p, p_old, s_old, dt, t)
    vx, vy, vz = O.face_velocity(op, p, p_old, dt, t)
    assert vx.shape == (prob.nz, prob.ny, prob.nx + 1)
    assert vy.shape == (prob.nz, prob.ny + 1, prob.nx)
    assert vz.shape == (prob.nz + 1, prob.ny, prob.nx)
    I = prob.interior
    s = O.saturation(op, p)
    rho = prob.rho
    vol = prob.dx * prob.dy * prob.dz
    acc = ((I(s) * rho - I(s_old) * rho) * I(prob.g_poro) * vol
           + I(prob.g_ss) * vol * (I(p) * I(s) * rho - I(p_old) * I(s_old) * rho))
    out = (prob.dy * prob.dz * (vx[:, :, 1:] - vx[:, :, :-1])
           + prob.dx * prob.dz * (vy[:, 1:, :] - vy[:, :-1, :])
           + prob.dx * prob.dy * (vz[1:] - vz[:-1]))
    lhs = I(F) - acc
    scale = max(np.max(np.abs(I(F))), np.max(np.abs(acc)), dt * np.max(np.abs(out)))
    assert np.max(np.abs(lhs - dt * out)) < 1e-12 * scale
    # the domain faces carry their BC: FluxConst faces the flux value
    if not dirs:
        assert np.all(vx[:, :, 0] == 0.0) and np.all(vx[:, :, -1] == 0.0)
    assert np.all(vy[:, 0, :] == 0.0)
    if not overland:
        assert np.all(vz[-1] == -0.01)


@pytest.mark.parametrize("case", ["tiltedv", "dir_src"])
def test_water_balance_budget_closes(case):
    """After each converged step the change of stored water equals the
    step's boundary inflow + overland-top flux inflow - runoff + sources
    (the budget of test/python/water_balance_x.py:546-620), to the Newton
    tolerance: summing F over the cells cancels every interior face flux."""
    if case == "tiltedv":
        prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
        src, n = None, 8          # two rain steps, then recession with runoff
    else:
        prob = cases.hetero(6, 5, 4, seed=41, dir_sides=True)
        src = cases.hetero_src(prob, 42)
        opts, p0, dt, n = cases.solver(), cases.random_pressure(prob, seed=43), 0.1, 4
    op = O.OracleProblem(prob, src=src)
    r, ph, sh, log = O.advance(op, opts, p0, 0.0, dt, n)
    assert r == 0
    t = 0.0
    seen = np.zeros(4)
    for s in range(1, n + 1):
        d = log.dt_used[s - 1]
        t += d
        f = O.water_balance_fluxes(op, ph[s], ph[s - 1], d, t)
        dm = cases.stored_mass(prob, ph[s], sh[s]) - cases.stored_mass(prob, ph[s - 1], sh[s - 1])
        assert abs(dm - (f[0] + f[1] - f[2] + f[3])) < 1e-8 * max(1.0, np.sum(np.abs(f))), s
        seen += np.abs(f)
    # every term the case exercises is non-trivial
    want = [False, True, True, False] if case == "tiltedv" else [True, False, False, True]
    assert [bool(x > 1e-3) for x in seen] == want


def _fd_jv(op, p, p_old, s_old, dt, t, v, eps=1e-7):
    Fp = O.nl_function_eval(op, p + eps * v, p_old, s_old, dt, t)
    Fm = O.nl_function_eval(op, p - eps * v, p_old, s_old, dt, t)
    return (Fp - Fm) / (2 * eps)


@pytest.mark.parametrize("overland", [False, True])
def test_jacobian_matches_finite_difference(overland):
    """Full analytic J.v vs a central difference of F at a point away from
    upwind switches and p = 0 kinks."""
    prob = cases.hetero(5, 4, 4, seed=10, overland=overland, dir_sides=True)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(11)
    p = prob.hydrostatic(-1.0, 2.0) + rng.normal(0, 0.05, prob.shape)
    if overland:
        p[-2] = rng.uniform(0.01, 0.03, p[-2].shape)
    p = np.ascontiguousarray(p)
    s_old = O.saturation(op, p)
    dt, t = 0.1, 0.1
    J = O.jacobian_eval(op, p, dt, t, 0)
    for trial in range(3):
        v = np.zeros_like(p)
        v[1:-1, 1:-1, 1:-1] = rng.normal(0, 1, (prob.nz, prob.ny, prob.nx))
        jv = O.matvec(op, J, 0, v)
        fd = _fd_jv(op, p, p, s_old, dt, t, v)
        I = prob.interior
        err = np.max(np.abs(I(jv) - I(fd))) / np.max(np.abs(I(fd)))
        assert err < 1e-5, err


def test_symmetric_pc_matrix():
    """PFSymmetric storage: upper couplings equal the full J's symmetric
    part; the diagonal is the full one (PL/richards_jacobian_eval.c:594-776)."""
    prob = cases.hetero(5, 4, 4, seed=12, dir_sides=True)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=13)
    Jf = O.jacobian_eval(op, p, 0.1, 0.1, 0)
    Js = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    assert np.array_equal(Jf[0], Js[0])
    # couplings: symmetric storage is a symmetric operator
    rng = np.random.default_rng(0)
    x = np.zeros_like(p); y = np.zeros_like(p)
    x[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    y[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    Ax = O.matvec(op, Js, 1, x)
    Ay = O.matvec(op, Js, 1, y)
    assert abs((y * Ax).sum() - (x * Ay).sum()) < 1e-10 * abs((y * Ax).sum())


# --- multigrid -----------------------------------------------------------------------
def test_mg_levels_follow_mgsemi_rule():
    """PL/mg_semi.c:832-900: coarsen the smallest spacing with >= 2 points."""
    from parflow_amd.problem import Problem
    prob = Problem(nx=8, ny=6, nz=16, dx=10, dy=10, dz=1)
    mg = O.MG(prob, 1, 1, 1)
    lv = mg.levels()
    assert lv[0] == ((8, 6, 16), 2)
    assert lv[1] == ((8, 6, 8), 2)
    assert lv[-1][0] == (1, 1, 1)


def test_mg_is_convergent_solver_on_poisson():
    """Repeated V-cycles on an SPD 7-point operator converge with a
    textbook contraction factor (< 0.3 per cycle)."""
    from parflow_amd.problem import Problem
    nx, ny, nz = 16, 12, 10
    prob = Problem(nx=nx, ny=ny, nz=nz, dx=1, dy=1, dz=0.5)
    shp = prob.shape
    C = np.zeros(shp); E = np.zeros(shp); N = np.zeros(shp); U = np.zeros(shp)
    ax, ay, az = 1 / 1.0 ** 2, 1 / 1.0 ** 2, 1 / 0.5 ** 2
    sl = (slice(1, -1),) * 3
    C[sl] = 2 * (ax + ay + az) + 0.01
    E[sl] = -ax; N[sl] = -ay; U[sl] = -az
    E[:, :, -2] = 0; N[:, -2, :] = 0; U[-2, :, :] = 0
    J = [C, None, E, None, N, None, U]
    J = [a if a is not None else np.zeros(shp) for a in J]
    op = O.OracleProblem(prob)
    mg = O.MG(prob, 1, 1, 1)
    mg.setup(J)
    rng = np.random.default_rng(1)
    b = np.zeros(shp)
    b[sl] = rng.normal(size=(nz, ny, nx))
    x = np.zeros(shp)
    rn = [np.linalg.norm(b)]
    for it in range(6):
        r = b - O.matvec(op, J, 1, x)
        r[0, :, :] = r[-1, :, :] = 0
        x = x + mg.apply(r)
        rr = b - O.matvec(op, J, 1, x)
        rn.append(np.linalg.norm(rr[sl]))
    factors = np.array(rn[1:]) / np.array(rn[:-1])
    assert np.all(factors[1:] < 0.3), factors


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("n", [17, 32, 33])
@pytest.mark.parametrize("sym,sweeps", [(1, 1), (0, 1), (0, 2)])
def test_mg_vcycle_is_exact_on_a_line(d, n, sym, sweeps):
    """On a line the semicoarsening V-cycle is cyclic reduction: red = odd
    points are the fine points, the operator-induced weights make the
    coarse operator the Schur complement, and the 1-point coarsest level is
    solved exactly, so ONE V-cycle from a zero guess solves A x = b to
    round-off, for both sweep orders.  A dropped coarse-operator term, a
    wrong weight or a wrong prolongation index leaves a residual of O(1)."""
    from parflow_amd.problem import Problem
    dims = [1, 1, 1]
    dims[d] = n
    prob = Problem(nx=dims[0], ny=dims[1], nz=dims[2], dx=1.0, dy=1.0, dz=1.0)
    A = cases.full_operator(prob, 60 + d, symmetric=True)
    mg = O.MG(prob, sweeps, sweeps, sym)
    mg.setup(A)
    b = np.zeros(prob.shape)
    b[1:-1, 1:-1, 1:-1] = np.random.default_rng(5).normal(size=(prob.nz, prob.ny, prob.nx))
    r = b - O.matvec(O.OracleProblem(prob), A, 0, mg.apply(b))
    assert np.linalg.norm(r[1:-1, 1:-1, 1:-1]) < 1e-14 * np.linalg.norm(b)


def test_mg_vcycle_with_palindromic_sweeps_is_symmetric():
    """MGSemi's smoother (Symmetric True) sweeps red-black-red before and
    after the coarse correction, a palindrome, so with restriction = the
    transpose of prolongation and symmetric coarse operators the V-cycle
    is a symmetric operator: u.M(w) = w.M(u) to round-off on a 3-D SPD
    matrix.  Swapped restriction weights break it (7e-3 in a planted
    test); the red-black sweeps of PFMG's nonsymmetric smoother are not
    palindromic, and there it does not hold."""
    from parflow_amd.problem import Problem
    prob = Problem(nx=9, ny=7, nz=6, dx=1.0, dy=1.5, dz=0.7)
    A = cases.full_operator(prob, 70, symmetric=True)
    rng = np.random.default_rng(5)
    sl = (slice(1, -1),) * 3
    u = np.zeros(prob.shape)
    w = np.zeros(prob.shape)
    u[sl] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    w[sl] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    mg = O.MG(prob, 1, 1, 1)
    mg.setup(A)
    a, b = (u * mg.apply(w)).sum(), (w * mg.apply(u)).sum()
    assert abs(a - b) < 1e-13 * abs(a)
    mg = O.MG(prob, 1, 1, 0)
    mg.setup(A)
    a, b = (u * mg.apply(w)).sum(), (w * mg.apply(u)).sum()
    assert abs(a - b) > 1e-3 * abs(a)


def test_full_mg_reduces_to_symmetric_mg():
    """A symmetric matrix in full storage gives the symmetric hierarchy: the
    same level operators (lower couplings = the neighbours' upper ones) and
    the same V-cycle output."""
    from parflow_amd.problem import Problem
    prob = Problem(nx=9, ny=7, nz=6, dx=1.0, dy=1.5, dz=0.7)
    A = cases.full_operator(prob, 21, symmetric=True)
    ms = O.MG(prob, 1, 1, 0)
    mf = O.MG(prob, 1, 1, 0, full=True)
    ms.setup(A)
    mf.setup(A)
    assert len(ms.levels()) > 4
    for l in range(len(ms.levels())):
        Ls, Lf = ms.level_matrix(l), mf.level_matrix(l)
        for q in range(4):
            assert np.allclose(Lf[q], Ls[q], rtol=1e-13, atol=1e-13), (l, q)
        assert np.allclose(Lf[4], np.roll(Ls[1], 1, axis=2), rtol=1e-13, atol=1e-13), l
        assert np.allclose(Lf[5], np.roll(Ls[2], 1, axis=1), rtol=1e-13, atol=1e-13), l
        assert np.allclose(Lf[6], np.roll(Ls[3], 1, axis=0), rtol=1e-13, atol=1e-13), l
    rng = np.random.default_rng(22)
    b = np.zeros(prob.shape)
    b[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    assert _rel(mf.apply(b), ms.apply(b), 1e-3) < 1e-12


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("full", [False, True])
def test_mg_coarse_operator_is_galerkin_on_a_line(d, full):
    """On a line (one coarsening direction, no off-direction couplings) the
    coarse operator is exactly P^T A P with the operator-induced
    prolongation P (injection at even points, -a_L/a_0 and -a_U/a_0 at odd
    ones): brute force with dense matrices, a non-symmetric A in full
    storage, a symmetric one otherwise."""
    from parflow_amd.problem import Problem
    n = 9
    dims = [1, 1, 1]
    dims[d] = n
    prob = Problem(nx=dims[0], ny=dims[1], nz=dims[2], dx=1.0, dy=1.0, dz=1.0)
    A = cases.full_operator(prob, 30 + d, symmetric=not full)
    line = lambda a: a[1:-1, 1:-1, 1:-1].reshape(-1)
    c, lo, up = line(A[0]), line(A[1 + 2 * d]), line(A[2 + 2 * d])
    Ad = np.diag(c) + np.diag(up[:-1], 1) + np.diag(lo[1:], -1)
    nc = (n + 1) // 2
    P = np.zeros((n, nc))
    for f in range(n):
        if f % 2 == 0:
            P[f, f // 2] = 1.0
        else:
            P[f, (f - 1) // 2] = -lo[f] / c[f]
            if f + 1 < n:
                P[f, (f + 1) // 2] = -up[f] / c[f]
    Ac = P.T @ Ad @ P
    mg = O.MG(prob, 1, 1, 0, full=full)
    mg.setup(A)
    assert mg.levels()[1][0][d] == nc
    L1 = [line(a) for a in mg.level_matrix(1)]
    assert np.allclose(L1[0], np.diag(Ac), rtol=1e-13)
    assert np.allclose(L1[1 + d][:-1], np.diag(Ac, 1), rtol=1e-13)
    assert L1[1 + d][-1] == 0.0
    if full:
        assert np.allclose(L1[4 + d][1:], np.diag(Ac, -1), rtol=1e-13)
        assert L1[4 + d][0] == 0.0
    # everything outside the tridiagonal band of Ac is zero
    assert np.count_nonzero(np.triu(Ac, 2)) == 0 and np.count_nonzero(np.tril(Ac, -2)) == 0


def test_full_mg_off_direction_collapse_orientation():
    """Off-direction couplings are collapsed from the three fine rows a
    coarse point stands for (weights 1, 1/2, 1/2), each side from its own
    side: with couplings that do not vary along the coarsening direction,
    an interior coarse point's north and south couplings are exactly twice
    the fine ones, and they stay apart when north != south."""
    from parflow_amd.problem import Problem
    prob = Problem(nx=9, ny=5, nz=1, dx=1.0, dy=4.0, dz=1.0)
    A = cases.full_operator(prob, 40)
    rng = np.random.default_rng(41)
    nrow = rng.uniform(-1.5, -0.5, size=prob.ny)
    srow = rng.uniform(-3.0, -2.0, size=prob.ny)
    A[4][1, 1:-1, 1:-1] = nrow[:, None]
    A[3][1, 1:-1, 1:-1] = srow[:, None]
    A[4][:, -2, :] = 0
    A[3][:, 1, :] = 0
    mg = O.MG(prob, 1, 1, 0, full=True)
    mg.setup(A)
    assert mg.levels()[0][1] == 0 and mg.levels()[1][0] == (5, 5, 1)
    L1 = mg.level_matrix(1)
    inner = (1, slice(1, -1), slice(2, -2))       # coarse i = 1 .. 3
    assert np.allclose(L1[2][inner], 2 * A[4][1, 1:-1, 1:4], rtol=1e-14)
    assert np.allclose(L1[5][inner], 2 * A[3][1, 1:-1, 1:4], rtol=1e-14)


def test_full_mg_is_convergent_on_convection_diffusion():
    """Repeated V-cycles of the full-storage hierarchy on a non-symmetric
    operator (random couplings plus upwinded convection) contract the
    residual per cycle, while the hierarchy of its symmetric part (what
    PFSymmetric storage keeps of it) makes the same iteration diverge."""
    from parflow_amd.problem import Problem
    prob = Problem(nx=16, ny=12, nz=10, dx=1.0, dy=1.0, dz=0.5)
    A = cases.full_operator(prob, 50, vel=(0.25, -0.125, 0.0625))
    S = [A[0], None, 0.5 * (A[2] + np.roll(A[1], -1, axis=2)), None,
         0.5 * (A[4] + np.roll(A[3], -1, axis=1)), None,
         0.5 * (A[6] + np.roll(A[5], -1, axis=0))]
    S = [a if a is not None else np.zeros(prob.shape) for a in S]
    op = O.OracleProblem(prob)
    sl = (slice(1, -1),) * 3
    rng = np.random.default_rng(51)
    b = np.zeros(prob.shape)
    b[sl] = rng.normal(size=(prob.nz, prob.ny, prob.nx))

    def factors(mg):
        x = np.zeros(prob.shape)
        rn = [np.linalg.norm(b)]
        for it in range(6):
            x = x + mg.apply(b - O.matvec(op, A, 0, x))
            rn.append(np.linalg.norm((b - O.matvec(op, A, 0, x))[sl]))
        return np.array(rn[1:]) / np.array(rn[:-1])

    mf = O.MG(prob, 1, 1, 0, full=True)
    mf.setup(A)
    f = factors(mf)
    assert np.all(f < 0.5), f
    ms = O.MG(prob, 1, 1, 0)
    ms.setup(S)
    assert factors(ms)[-1] > 1.0


def test_gmres_matches_direct_solve():
    """One Newton step on a tiny problem: the oracle's preconditioned GMRES
    solve equals a dense numpy solve with the assembled Jacobian."""
    prob = cases.hetero(3, 3, 3, seed=20, dir_sides=True)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=21)
    s_old = O.saturation(op, p_old)
    opts = cases.solver(use_jacobian=True, max_iter=1, krylov_dim=27,
                        residual_tol=1e-300, eta_choice=0, eta_value=1e-12,
                        globalization=0)
    p = p_old.copy()
    r, st = O.kinsol(op, opts, p, p_old, s_old, 0.1, 0.1)
    # dense Newton step
    J = O.jacobian_eval(op, p_old, 0.1, 0.1, 0)
    n = prob.ncells
    A = np.zeros((n, n))
    idx = np.flatnonzero(np.zeros(prob.shape, bool)[...] | True).reshape(prob.shape)[1:-1, 1:-1, 1:-1].ravel()
    for c in range(n):
        e = np.zeros(prob.shape).ravel()
        e[idx[c]] = 1
        A[:, c] = O.matvec(op, J, 0, e.reshape(prob.shape)).ravel()[idx]
    F0 = O.nl_function_eval(op, p_old, p_old, s_old, 0.1, 0.1).ravel()[idx]
    dp = np.linalg.solve(A, -F0)
    assert np.allclose(p.ravel()[idx] - p_old.ravel()[idx], dp, rtol=1e-8, atol=1e-12)


def test_time_cycle_interval_numbers():
    """TimeCycleDataComputeIntervalNumber: rain for base-unit intervals
    1..2 of a (2, 300) cycle, recession after (time_cycle_data.c:152-197)."""
    prob, _, _, _, _ = ref_decks.tiltedv_kin()
    b = prob.bc[5]
    got = [prob.interval_number(b, t) for t in (0.0, 0.05, 0.1, 0.15, 0.5, 15.1, 15.15)]
    assert got == [0, 0, 0, 1, 1, 1, 0]
    op = O.OracleProblem(prob)
    import ctypes as C
    for t in (0.05, 0.1, 0.15, 15.1, 15.15):
        assert O.lib().pfo_interval_number(op.ref, C.byref(op.c.bc[5]), C.c_double(t)) == \
            prob.interval_number(b, t)


def test_storage_closed_form():
    """Water-balance storage (hydrology.py:136-205) on fields whose sums are
    known from counts: constant porosity 0.4 and Ss 1e-4, every cell dry
    (p = -2, S = 0.3) except the top layer, where the first half of each row
    is ponded (p = 0.5, S = 1) and the rest dry.  Dry cells add no
    compressible storage and negative top pressure no surface storage."""
    prob = cases.hetero(8, 6, 5, seed=3)
    op = O.OracleProblem(prob)
    p = np.full(prob.shape, -2.0)
    s = np.full(prob.shape, 0.3)
    top = p[prob.nz]                       # ghosted plane of the top layer
    top[1:-1, 1:1 + prob.nx // 2] = 0.5
    s[prob.nz][1:-1, 1:1 + prob.nx // 2] = 1.0
    n_pond = prob.ny * (prob.nx // 2)
    n_all = prob.nx * prob.ny * prob.nz
    vol = prob.dx * prob.dy * prob.dz
    sub, sur = O.storage(op, p, s)
    sum_s = 0.3 * (n_all - n_pond) + 1.0 * n_pond
    expect_sub = 0.4 * vol * sum_s + 0.5 * 1.0 * 1e-4 * vol * n_pond
    assert abs(sub - expect_sub) < 1e-12 * expect_sub
    assert abs(sur - 0.5 * prob.dx * prob.dy * n_pond) < 1e-12 * sur
    # all dry: only the incompressible part remains, no surface water
    sub2, sur2 = O.storage(op, np.full(prob.shape, -2.0), np.full(prob.shape, 0.3))
    assert abs(sub2 - 0.4 * vol * 0.3 * n_all) < 1e-12 * sub2
    assert sur2 == 0.0


# --- van Genuchten lookup tables ------------------------------------------------
def _kr_closed(h, alpha, n):
    """van Genuchten-Mualem kr(h), h >= 0 the pressure head magnitude."""
    m = 1.0 - 1.0 / n
    t = (alpha * h) ** n
    return (1.0 - (alpha * h) ** (n - 1) * (1.0 + t) ** (-m)) ** 2 * (1.0 + t) ** (-m / 2)


@pytest.mark.parametrize("n", [2.0, 1.6, 3.5])
def test_vg_table_nodes_and_accuracy(n):
    """PL/problem_phase_rel_perm.c:116-354: at its nodes the spline table
    returns kr exactly as the closed form gives it; between them the
    monotone cubic Hermite interpolant converges (error ratio >= 3.5 when the
    spacing halves: the averaged-secant slopes make it second order), stays
    monotone (Fritsch-Carlson),
    and the linear table interpolates the nodes; 0 beyond |MinPressureHead|."""
    alpha, head_min = 1.5, -10.0
    errs = []
    for pts in (200, 400):
        tab = O.vg_table_build(0, pts, head_min, alpha, n)
        x = tab[:pts + 1]
        h = abs(head_min) / (pts - 1)
        assert np.allclose(x, np.arange(pts + 1) * h, rtol=0, atol=1e-12)
        for i in (0, 1, 7, pts // 2, pts - 2):
            got = O.vg_table_lookup(tab, 0, pts, head_min, x[i], 0)
            assert abs(got - _kr_closed(x[i], alpha, n)) <= 1e-14 * max(1.0, got), (i, got)
        fine = np.linspace(0.0, 9.999, 4001)
        vals = np.array([O.vg_table_lookup(tab, 0, pts, head_min, v, 0) for v in fine])
        assert np.all(np.diff(vals) <= 1e-15)             # monotone non-increasing
        errs.append(np.max(np.abs(vals - _kr_closed(fine, alpha, n))))
        assert O.vg_table_lookup(tab, 0, pts, head_min, 10.0, 0) == 0.0
        assert O.vg_table_lookup(tab, 0, pts, head_min, 12.5, 1) == 0.0
        lin = O.vg_table_build(1, pts, head_min, alpha, n)
        for i in (3, pts // 3):
            xa, xb = lin[i], lin[i + 1]
            ka, kb = _kr_closed(xa, alpha, n), _kr_closed(xb, alpha, n)
            assert abs(O.vg_table_lookup(lin, 1, pts, head_min, xa, 0) - ka) <= 1e-14
            mid = O.vg_table_lookup(lin, 1, pts, head_min, 0.5 * (xa + xb), 0)
            assert abs(mid - 0.5 * (ka + kb)) <= 1e-14
    if n >= 2.0:
        assert errs[0] / errs[1] >= 3.5, errs
    else:
        # 1 < n < 2: kr has unbounded curvature at h = 0, so only the
        # interpolation error shrinks, not at a fixed order
        assert errs[1] < errs[0], errs


def test_vg_table_derivative_values():
    """The derivative table holds the reference's CALCDER formula: it equals
    -dkr/dh of the closed form (central differences) at interior nodes, and
    is set to 0 at head 0 when n < 2 (PL/problem_phase_rel_perm.c:181-184)."""
    alpha, n, pts, head_min = 0.8, 2.4, 300, -20.0
    tab = O.vg_table_build(0, pts, head_min, alpha, n)
    x, ader = tab[:pts + 1], tab[3 * (pts + 1):4 * (pts + 1)]
    for i in (5, 60, 150, 290):
        e = 1e-6
        fd = -(_kr_closed(x[i] + e, alpha, n) - _kr_closed(x[i] - e, alpha, n)) / (2 * e)
        assert abs(ader[i] - fd) <= 1e-6 * max(abs(fd), 1e-3), (i, ader[i], fd)
        assert O.vg_table_lookup(tab, 0, pts, head_min, x[i], 1) == ader[i]
    tab15 = O.vg_table_build(0, pts, head_min, alpha, 1.5)
    assert tab15[3 * (pts + 1)] == 0.0


def test_vg_table_in_the_residual():
    """A problem with a kr table: relperm and the residual take kr from the
    table (spline and linear differ from the closed form between nodes)."""
    import dataclasses
    base = cases.hetero(6, 5, 4, seed=71)
    kw = dict(alpha=1.2, nvg=2.2)
    p0 = dataclasses.replace(base, **kw)
    op0 = O.OracleProblem(p0)
    p = cases.random_pressure(p0, seed=72)
    kr0 = O.relperm(op0, p)
    for method in ("Spline", "Linear"):
        pt = dataclasses.replace(base, vg_table=(12, -8.0, method), **kw)
        opt = O.OracleProblem(pt)
        krt = O.relperm(opt, p)
        I = pt.interior
        wet = I(p) >= 0.0
        assert np.all(I(krt)[wet] == 1.0)
        d = np.abs(I(krt) - I(kr0))[~wet]
        assert 0.0 < d.max() < 0.2, (method, d.max())   # a coarse 12-point table
        # every cell's kr is the table's value at its head
        heads = np.abs(I(p)[~wet]) / (pt.rho * pt.gravity)
        tab = opt.vg_table
        ref = [O.vg_table_lookup(tab, 0 if method == "Spline" else 1, 12, -8.0, h, 0)
               for h in heads]
        assert np.array_equal(I(krt)[~wet], np.array(ref))
        s_old = O.saturation(opt, p)
        F0 = O.nl_function_eval(op0, p, p, s_old, 0.1, 0.1)
        Ft = O.nl_function_eval(opt, p, p, s_old, 0.1, 0.1)
        assert np.max(np.abs(I(F0) - I(Ft))) > 0.0
