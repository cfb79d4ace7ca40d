"""GPU parity tests: the HIP library, called through its C-ABI, against the
oracle on identical seeded inputs.

Bars: FP64 kernels agree with the oracle to round-off (1e-12 relative to
the magnitude of the terms involved; device pow and FMA contraction differ
from glibc by a few ulp); the plain (reference-algorithm) Newton solve
agrees per cell within 1e-9 relative after every step with Newton counts
within +-1.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
import cases
import ref_decks

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _api():
    from parflow_amd import api
    return api


def _I(prob, a):
    return prob.interior(a)


def _relmax(a, b, scale=None):
    d = np.max(np.abs(a - b))
    s = scale if scale is not None else max(np.max(np.abs(a)), np.max(np.abs(b)), 1e-300)
    return d / s


@pytest.mark.parametrize("seed", [0, 1])
def test_closures(gpu, seed):
    api = _api()
    prob = cases.hetero(13, 7, 5, seed=seed)
    print(f"closures seed={seed} shape={prob.shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=seed + 10)
    for fcn in (0, 1):
        a = gp.saturation(p, fcn)
        b = O.saturation(op, p, fcn)
        assert _relmax(_I(prob, a), _I(prob, b)) < 1e-13
        a = gp.phase_rel_perm(p, fcn)
        b = O.relperm(op, p, fcn)
        assert _relmax(_I(prob, a), _I(prob, b)) < 1e-12


@pytest.mark.parametrize("overland,dirs", [(False, False), (False, True), (True, True)])
def test_residual(gpu, overland, dirs):
    api = _api()
    prob = cases.hetero(11, 9, 6, seed=3, overland=overland, dir_sides=dirs)
    print(f"residual overland={overland} dir={dirs}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=4)
    s_old = O.saturation(op, p_old)
    p = p_old + np.random.default_rng(5).normal(0, 0.01, p_old.shape)
    for t in (0.1, 0.35):
        Fg = gp.nl_function_eval(p, p_old, s_old, 0.1, t)
        Fo = O.nl_function_eval(op, p, p_old, s_old, 0.1, t)
        # scale: magnitude of the accumulation terms
        scale = np.max(np.abs(_I(prob, Fo))) + 1e-30
        assert _relmax(_I(prob, Fg), _I(prob, Fo), scale) < 1e-12


@pytest.mark.parametrize("shape", [(11, 9, 6), (70, 13, 5)])
@pytest.mark.parametrize("overland,dirs", [(False, False), (False, True), (True, True)])
def test_face_velocity(gpu, overland, dirs, shape):
    """pfhip_face_velocity against the oracle; the second shape spans two
    workgroups in x and four in y."""
    api = _api()
    prob = cases.hetero(*shape, seed=3, overland=overland, dir_sides=dirs)
    print(f"face_velocity overland={overland} dir={dirs} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=4)
    p = p_old + np.random.default_rng(5).normal(0, 0.01, p_old.shape)
    for t in (0.1, 0.35):
        got = gp.face_velocity(p, p_old, 0.1, t)
        ref = O.face_velocity(op, p, p_old, 0.1, t)
        scale = max(np.max(np.abs(r)) for r in ref)
        for a, b in zip(got, ref):
            assert a.shape == b.shape
            assert np.max(np.abs(a - b)) < 1e-12 * scale


def test_face_velocity_golden_default_richards(gpu):
    """The device velocities from the reference's step-n pressure against
    its velx/vely/velz outputs of step n (which the oracle reproduces bit for
    bit); FMA contraction on the device leaves round-off differences."""
    api = _api()
    prob, opts, p0, dt, n = ref_decks.default_richards()
    gp = api.GPUProblem(prob)
    g = np.load(os.path.join(HERE, "golden", "default_richards.npz"))
    gv = np.load(os.path.join(HERE, "golden", "default_richards_vel.npz"))
    steps = [int(x) for x in g["steps"]]
    for i, s in enumerate(gv["steps"]):
        p = prob.zeros()
        prob.interior(p)[...] = g["press"][steps.index(int(s))]
        got = gp.face_velocity(p, p, dt, s * dt)
        ref = [gv["velx"][i], gv["vely"][i], gv["velz"][i]]
        scale = max(np.max(np.abs(r)) for r in ref)
        for a, b in zip(got, ref):
            assert np.max(np.abs(a - b)) < 1e-13 * scale, s


def _table_case(method, overland=False, dirs=True, shape=(11, 9, 6)):
    import dataclasses
    base = cases.hetero(*shape, seed=3, overland=overland, dir_sides=dirs)
    return dataclasses.replace(base, alpha=1.3, nvg=2.3, vg_table=(40, -6.0, method))


@pytest.mark.parametrize("method", ["Spline", "Linear"])
def test_vg_table_closures_residual_jacobian(gpu, method):
    """A problem with a van Genuchten kr table (region-uniform alpha, n):
    relperm and its derivative, the residual (Dirichlet sides, and the
    tiled kernel across several tiles) and the symmetric and full Jacobians
    against the oracle, which takes kr from its own table."""
    api = _api()
    for overland, shape in ((False, (11, 9, 6)), (True, (70, 13, 5))):
        prob = _table_case(method, overland, shape=shape)
        print(f"vg table {method} overland={overland} shape={shape}", flush=True)
        gp = api.GPUProblem(prob)
        op = O.OracleProblem(prob)
        p_old = cases.random_pressure(prob, seed=4)
        p = p_old + np.random.default_rng(5).normal(0, 0.01, p_old.shape)
        for fcn in (0, 1):
            a, b = gp.phase_rel_perm(p, fcn), O.relperm(op, p, fcn)
            assert _relmax(_I(prob, a), _I(prob, b)) < 1e-13, fcn
        s_old = O.saturation(op, p_old)
        Fg = gp.nl_function_eval(p, p_old, s_old, 0.1, 0.35)
        Fo = O.nl_function_eval(op, p, p_old, s_old, 0.1, 0.35)
        scale = np.max(np.abs(_I(prob, Fo))) + 1e-30
        assert _relmax(_I(prob, Fg), _I(prob, Fo), scale) < 1e-12
        for symm in (0, 1):
            Jg = gp.richards_jacobian_eval(p, 0.1, 0.35, symm)
            Jo = O.jacobian_eval(op, p, 0.1, 0.35, symm)
            for q in ((0, 2, 4, 6) if symm else range(7)):
                sc = np.max(np.abs(_I(prob, Jo[q]))) + 1e-30
                assert _relmax(_I(prob, Jg[q]), _I(prob, Jo[q]), sc) < 1e-12, (symm, q)


def test_vg_table_newton_matches_oracle(gpu):
    """The plain Newton solve of a problem with a kr table against the oracle
    (per-cell 1e-9, Newton counts +-1)."""
    prob = _table_case("Spline")
    _compare_runs(prob, cases.solver(), cases.random_pressure(prob, seed=9), 0.1, 3)


@pytest.mark.parametrize("symm", [0, 1])
@pytest.mark.parametrize("overland", [False, True])
def test_jacobian_and_matvec(gpu, symm, overland):
    api = _api()
    prob = cases.hetero(9, 8, 7, seed=6, overland=overland, dir_sides=True)
    print(f"jacobian symm={symm} overland={overland}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=7)
    Jg = gp.richards_jacobian_eval(p, 0.2, 0.2, symm)
    Jo = O.jacobian_eval(op, p, 0.2, 0.2, symm)
    idx = [0, 2, 4, 6] if symm else range(7)
    scale = max(np.max(np.abs(Jo[s])) for s in idx)
    for s in idx:
        assert _relmax(_I(prob, Jg[s]), _I(prob, Jo[s]), scale) < 1e-12, s
    x = np.zeros(prob.shape)
    x[1:-1, 1:-1, 1:-1] = np.random.default_rng(8).normal(size=(prob.nz, prob.ny, prob.nx))
    yg = gp.matvec(Jo, symm, x)
    yo = O.matvec(op, Jo, symm, x)
    assert _relmax(_I(prob, yg), _I(prob, yo)) < 1e-13


def test_overland_fluxes(gpu):
    api = _api()
    prob = cases.hetero(10, 7, 3, seed=9, overland=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=10)
    qx, qy, _, _ = gp.overland_flow_eval_kin(p, 0)
    qxo, qyo = O.overland_kin(op, p)
    n2 = (prob.nx + 2) * (prob.ny + 2)
    sl = np.zeros((prob.ny + 2, prob.nx + 2), bool)
    sl[1:-1, 0:-1] = True        # faces read by the residual (incl. west outflow)
    m = sl.ravel()
    assert np.allclose(qx[m], qxo[m], rtol=1e-13, atol=1e-300)
    sl = np.zeros((prob.ny + 2, prob.nx + 2), bool)
    sl[0:-1, 1:-1] = True
    m = sl.ravel()
    assert np.allclose(qy[m], qyo[m], rtol=1e-13, atol=1e-300)


def test_vector_reductions(gpu):
    api = _api()
    prob = cases.hetero(37, 21, 9, seed=11)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(12)
    x = np.zeros(prob.shape); y = np.zeros(prob.shape)
    x[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    y[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    x[0] = 1e9   # ghost values must not be reduced
    lib = O.lib()
    ref_dot = float(np.sum(_I(prob, x) * _I(prob, y)))
    assert abs(gp.vec_dot(x, y) - ref_dot) < 1e-12 * np.sum(np.abs(_I(prob, x) * _I(prob, y)))
    assert gp.vec_max_norm(x) == np.max(np.abs(_I(prob, x)))
    assert abs(gp.vec_l1_norm(x) - np.sum(np.abs(_I(prob, x)))) < 1e-12 * np.sum(np.abs(_I(prob, x)))
    # determinism: identical bits on repeat
    assert gp.vec_dot(x, y) == gp.vec_dot(x, y)


def test_n_vector_operations(gpu):
    """Every N_Vector operation of PL/n_vector.h at the boundary against numpy
    on the interior cells: elementwise results exactly (one IEEE operation
    per cell), reductions to round-off, the ghost layer neither read (it is
    poisoned) nor written."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(37, 21, 9, seed=13)
    gp = api.GPUProblem(prob)
    rng = np.random.default_rng(14)
    I = prob.interior

    def field(lo=-2.0, hi=2.0):
        a = np.full(prob.shape, 1e9)
        a[1:-1, 1:-1, 1:-1] = rng.uniform(lo, hi, (prob.nz, prob.ny, prob.nx))
        return a
    x, y, w = field(), field(0.5, 1.5), field()
    dx_, dy_, dw = gp.vec(x), gp.vec(y), gp.vec(w)
    z0 = np.full(prob.shape, -7.0)

    def ew(call, ref):
        dz = gp.vec(z0)
        api.check(call(dz), "N_Vector op")
        got = gp.get(dz)
        assert np.array_equal(I(got), ref)
        assert np.all(got[0] == -7.0) and np.all(got[:, 0] == -7.0)
    ew(lambda dz: L.pfhip_vec_prod(gp.h, dx_.ptr, dy_.ptr, dz.ptr, None), I(x) * I(y))
    ew(lambda dz: L.pfhip_vec_div(gp.h, dx_.ptr, dy_.ptr, dz.ptr, None), I(x) / I(y))
    ew(lambda dz: L.pfhip_vec_scale(gp.h, -1.5, dx_.ptr, dz.ptr, None), -1.5 * I(x))
    ew(lambda dz: L.pfhip_vec_abs(gp.h, dx_.ptr, dz.ptr, None), np.abs(I(x)))
    ew(lambda dz: L.pfhip_vec_inv(gp.h, dy_.ptr, dz.ptr, None), 1.0 / I(y))
    ew(lambda dz: L.pfhip_vec_add_const(gp.h, dx_.ptr, 0.25, dz.ptr, None), I(x) + 0.25)
    ew(lambda dz: L.pfhip_vec_compare(gp.h, 1.0, dx_.ptr, dz.ptr, None),
       np.where(np.abs(I(x)) >= 1.0, 1.0, 0.0))
    r = C.c_double()
    api.check(L.pfhip_vec_wl2_norm(gp.h, dx_.ptr, dw.ptr, C.byref(r), None), "wl2")
    ref = np.sqrt(np.sum((I(x) * I(w)) ** 2))
    assert abs(r.value - ref) < 1e-13 * ref
    api.check(L.pfhip_vec_wrms_norm(gp.h, dx_.ptr, dw.ptr, C.byref(r), None), "wrms")
    assert abs(r.value - ref / np.sqrt(prob.ncells)) < 1e-13 * ref
    api.check(L.pfhip_vec_min(gp.h, dx_.ptr, C.byref(r), None), "min")
    assert r.value == np.min(I(x))
    pos = field(0.1, 1.0)
    api.check(L.pfhip_vec_min(gp.h, gp.vec(pos).ptr, C.byref(r), None), "min")
    assert r.value == np.min(I(pos)) and r.value > 0
    ok = C.c_int()
    c = np.zeros(prob.shape)
    c[1:-1, 1:-1, 1:-1] = np.where(rng.uniform(size=(prob.nz, prob.ny, prob.nx)) < 0.3,
                                   np.sign(I(x)), 0.0)
    dc = gp.vec(c)
    api.check(L.pfhip_vec_constr_prod_pos(gp.h, dc.ptr, dx_.ptr, C.byref(ok), None), "constr")
    assert ok.value == 1
    c[5, 7, 9] = -np.sign(x[5, 7, 9])
    api.check(L.pfhip_vec_constr_prod_pos(gp.h, gp.vec(c).ptr, dx_.ptr, C.byref(ok), None),
              "constr")
    assert ok.value == 0
    ew(lambda dz: L.pfhip_vec_inv_test(gp.h, dy_.ptr, dz.ptr, C.byref(ok), None), 1.0 / I(y))
    assert ok.value == 1
    xz = x.copy()
    xz[3, 4, 5] = 0.0
    dz = gp.vec(z0)
    api.check(L.pfhip_vec_inv_test(gp.h, gp.vec(xz).ptr, dz.ptr, C.byref(ok), None), "invtest")
    got = gp.get(dz)
    assert ok.value == 0 and got[3, 4, 5] == -7.0
    mask = I(xz) != 0
    assert np.array_equal(I(got)[mask], 1.0 / I(xz)[mask])


@pytest.mark.parametrize("m", [1, 7, 16, 31])
def test_block_gram_schmidt_products(gpu, m):
    """CGS2 block products: the FP64 multi-dot h = V^T w and w -= V h with
    the fused ||w||^2, against numpy on the interior cells (ragged nx,
    ghosts poisoned)."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(37, 21, 9, seed=3)
    gp = api.GPUProblem(prob)
    rng = np.random.default_rng(100 + m)

    def field():
        a = np.full(prob.shape, 1e9)
        a[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
        return a
    Vh = [field() for _ in range(m)]
    wh = field()
    V = [gp.vec(a) for a in Vh]
    w = gp.vec(wh)
    ptrs = (C.c_void_p * m)(*[v.value for v in V])
    L.pfhip_vec_multi_dot.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    L.pfhip_vec_multi_axpy.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
    h = (C.c_double * m)()
    api.check(L.pfhip_vec_multi_dot(gp.h, ptrs, m, w.ptr, h, None), "multi_dot")
    Wi = _I(prob, wh)
    for q in range(m):
        ref = float(np.sum(_I(prob, Vh[q]) * Wi))
        scale = float(np.sum(np.abs(_I(prob, Vh[q]) * Wi)))
        assert abs(h[q] - ref) <= 1e-13 * scale, (q, h[q], ref)
    h2 = list(h)
    again = (C.c_double * m)()
    api.check(L.pfhip_vec_multi_dot(gp.h, ptrs, m, w.ptr, again, None), "multi_dot")
    assert list(again) == h2          # fixed reduction order: identical bits
    n2 = C.c_double()
    api.check(L.pfhip_vec_multi_axpy(gp.h, ptrs, m, h, w.ptr, C.byref(n2), None), "multi_axpy")
    wref = Wi - sum(h2[q] * _I(prob, Vh[q]) for q in range(m))
    got = _I(prob, gp.get(w))
    assert np.max(np.abs(got - wref)) <= 1e-12 * np.max(np.abs(wref))
    assert abs(n2.value - float(np.sum(wref * wref))) <= 1e-12 * float(np.sum(wref * wref))


@pytest.mark.parametrize("k", [1, 4, 5, 8, 9, 17])
def test_cgs2_matches_separate_passes(gpu, k):
    """pfhip_vec_cgs2 (for k <= 8 its middle passes fused into one) against
    the same projection done by separate multi-dot / multi-axpy calls: the
    same products summed in the same order."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(37, 21, 9, seed=5)
    gp = api.GPUProblem(prob)
    rng = np.random.default_rng(200 + k)

    def field():
        a = np.full(prob.shape, 1e9)
        a[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
        return a
    Vh = [field() for _ in range(k)]
    wh = field() + 0.3 * Vh[0]
    V = [gp.vec(a) for a in Vh]
    ptrs = (C.c_void_p * k)(*[v.value for v in V])
    vp = C.c_void_p
    L.pfhip_vec_cgs2.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp]
    L.pfhip_vec_multi_dot.argtypes = [vp, vp, C.c_int, vp, vp, vp]
    L.pfhip_vec_multi_axpy.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
    w1 = gp.vec(wh)
    h1, h2, n2 = (C.c_double * k)(), (C.c_double * k)(), C.c_double()
    api.check(L.pfhip_vec_cgs2(gp.h, ptrs, k, w1.ptr, h1, h2, C.byref(n2), None), "cgs2")
    w2 = gp.vec(wh)
    a1, a2, m2 = (C.c_double * k)(), (C.c_double * k)(), C.c_double()
    api.check(L.pfhip_vec_multi_dot(gp.h, ptrs, k, w2.ptr, a1, None), "multi_dot")
    api.check(L.pfhip_vec_multi_axpy(gp.h, ptrs, k, a1, w2.ptr, None, None), "multi_axpy")
    api.check(L.pfhip_vec_multi_dot(gp.h, ptrs, k, w2.ptr, a2, None), "multi_dot")
    api.check(L.pfhip_vec_multi_axpy(gp.h, ptrs, k, a2, w2.ptr, C.byref(m2), None), "axpy")
    assert list(h1) == list(a1)
    s2 = max(abs(x) for x in a2) + 1e-300
    assert max(abs(x - y) for x, y in zip(h2, a2)) <= 1e-13 * s2
    g1, g2 = _I(prob, gp.get(w1)), _I(prob, gp.get(w2))
    assert np.max(np.abs(g1 - g2)) <= 1e-13 * np.max(np.abs(g2))
    assert abs(n2.value - m2.value) <= 1e-13 * m2.value
    print("bit-identical:", list(h2) == list(a2) and np.array_equal(g1, g2))


def test_water_balance_storage(gpu):
    api = _api()
    prob = cases.hetero(37, 21, 9, seed=27)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=28)
    # pond part of the top layer
    p[prob.nz][1:-1, 1:-1] = np.random.default_rng(29).uniform(-0.2, 0.3, (prob.ny, prob.nx))
    s = O.saturation(op, p)
    sub_o, sur_o = O.storage(op, p, s)
    sub_g, sur_g = gp.water_balance_storage(p, s)
    assert abs(sub_g - sub_o) < 1e-13 * abs(sub_o)
    assert abs(sur_g - sur_o) < 1e-13 * max(abs(sur_o), 1e-300)
    assert sur_o > 0.0
    assert gp.water_balance_storage(p, s) == (sub_g, sur_g)


@pytest.mark.parametrize("overland,dirs,with_src", [(False, True, True), (True, True, False),
                                                    (True, False, True)])
def test_water_balance_fluxes(gpu, overland, dirs, with_src):
    """pfhip_water_balance_fluxes against the oracle (which takes the
    boundary terms from the face velocities instead of the BC formulas)."""
    api = _api()
    prob = cases.hetero(37, 21, 9, seed=27, overland=overland, dir_sides=dirs)
    print(f"water_balance_fluxes overland={overland} dir={dirs} src={with_src}", flush=True)
    src = cases.hetero_src(prob, 28) if with_src else None
    gp = api.GPUProblem(prob, src=src)
    op = O.OracleProblem(prob, src=src)
    p_old = cases.random_pressure(prob, seed=29)
    p = p_old + np.random.default_rng(30).normal(0, 0.01, p_old.shape)
    for t in (0.1, 0.35):
        got = gp.water_balance_fluxes(p, p_old, 0.1, t)
        ref = O.water_balance_fluxes(op, p, p_old, 0.1, t)
        scale = np.sum(np.abs(ref)) + 1e-300
        assert np.max(np.abs(got - ref)) < 1e-12 * scale, (got, ref)
        # the overland top rains over [0, 0.2) (cases.hetero's cycle)
        assert (got[1] != 0.0) == (overland and t < 0.2)
        assert (got[2] != 0.0) == overland
        assert (got[3] != 0.0) == with_src


def test_evap_trans_sum(gpu):
    """EvapTransSum: the running sum grows by src dx dy dz dt per call."""
    api = _api()
    prob = cases.hetero(13, 7, 5, seed=2)
    src = cases.hetero_src(prob, 3)
    gp = api.GPUProblem(prob, src=src)
    acc = gp.vec(np.zeros(prob.shape))
    vol = prob.dx * prob.dy * prob.dz
    want = np.zeros(prob.shape)
    for dt in (0.1, 0.25, 0.05):
        gp.evap_trans_sum(acc, dt)
        want[1:-1, 1:-1, 1:-1] += src[1:-1, 1:-1, 1:-1] * (vol * dt)
    got = gp.get(acc)
    # the device fuses the multiply-add: agreement to round-off
    w = want[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(got[1:-1, 1:-1, 1:-1] - w)) <= 1e-15 * np.max(np.abs(w))
    # a problem without a source field adds nothing
    gp0 = api.GPUProblem(prob)
    z = gp0.vec(np.zeros(prob.shape))
    gp0.evap_trans_sum(z, 0.1)
    assert np.all(gp0.get(z) == 0.0)


@pytest.mark.parametrize("ss", [1.0e-4, 0.0])
def test_water_balance_budget_on_gpu(gpu, ss):
    """The device solve of the reference's TiltedV_OverlandKin deck (rain,
    then recession with runoff): every step's stored-water change equals
    the device budget terms to the Newton tolerance.  With Ss = 0 pftools'
    storages are the residual's own, so water_balance_report's percent
    difference stays far inside the reference test's 0.005 % bar."""
    import dataclasses
    api = _api()
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    prob = dataclasses.replace(prob, ss=ss)
    n = 8
    print(f"water balance tiltedv ss={ss} {n} steps", flush=True)
    gp = api.GPUProblem(prob)
    sv = api.GPUSolver(gp, opts)
    ph, sh, log = sv.advance(p0, 0.0, dt, n)
    rows = api.water_balance_report(gp, ph, sh, log["dt"], 0.0)
    assert len(rows) == n
    for s, row in enumerate(rows, start=1):
        dm = cases.stored_mass(prob, ph[s], sh[s]) - cases.stored_mass(prob, ph[s - 1], sh[s - 1])
        terms = [row["boundary_inflow"], row["overland_flux_inflow"], row["surface_runoff"],
                 row["sources"]]
        budget = terms[0] + terms[1] - terms[2] + terms[3]
        assert abs(dm - budget) < 1e-8 * max(1.0, np.sum(np.abs(terms))), s
        if ss == 0.0:
            assert row["percent_diff"] < 1e-5, row
    assert rows[0]["overland_flux_inflow"] > 0.0 and rows[-1]["surface_runoff"] > 0.0


@pytest.mark.parametrize("precision", [64, 32])
def test_multigrid(gpu, precision):
    api = _api()
    prob = cases.hetero(16, 12, 10, seed=13, dir_sides=True)
    print(f"mg precision={precision}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=14)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    for pre, post, sym in ((1, 1, 1), (2, 2, 0)):
        mgo = O.MG(prob, pre, post, sym)
        mgo.setup(J)
        mgg = api.GPUMultigrid(gp, pre, post, sym, precision)
        assert mgg.levels() == len(mgo.levels())
        mgg.setup(J)
        b = np.zeros(prob.shape)
        b[1:-1, 1:-1, 1:-1] = np.random.default_rng(15).normal(size=(prob.nz, prob.ny, prob.nx))
        xo = mgo.apply(b)
        xg = mgg.apply(b)
        tol = 1e-11 if precision == 64 else 2e-5
        assert _relmax(_I(prob, xg), _I(prob, xo)) < tol


@pytest.mark.parametrize("d,n", [(0, 32), (0, 33), (1, 64), (2, 40)])
def test_multigrid_exact_on_a_line(gpu, d, n):
    """The device V-cycle (FP64, both sweep orders) solves a symmetric line
    operator exactly in one cycle, as cyclic reduction does; the same
    hierarchy's V-cycle on a 3-D SPD matrix with MGSemi's palindromic
    sweeps is a symmetric operator."""
    api = _api()
    from parflow_amd.problem import Problem
    dims = [1, 1, 1]
    dims[d] = n
    prob = Problem(nx=dims[0], ny=dims[1], nz=dims[2], dx=1.0, dy=1.0, dz=1.0)
    print(f"mg line d={d} n={n}", flush=True)
    A = cases.full_operator(prob, 80 + d, symmetric=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    b = np.zeros(prob.shape)
    b[1:-1, 1:-1, 1:-1] = np.random.default_rng(6).normal(size=(prob.nz, prob.ny, prob.nx))
    for sym, sweeps in ((1, 1), (0, 2)):
        mg = api.GPUMultigrid(gp, sweeps, sweeps, sym, 64)
        mg.setup(A)
        r = b - O.matvec(op, A, 0, mg.apply(b))
        assert np.linalg.norm(r[1:-1, 1:-1, 1:-1]) < 1e-13 * np.linalg.norm(b), (sym, sweeps)
    prob = Problem(nx=40, ny=18, nz=12, dx=1.0, dy=1.5, dz=0.7)
    A = cases.full_operator(prob, 90, symmetric=True)
    gp = api.GPUProblem(prob)
    mg = api.GPUMultigrid(gp, 1, 1, 1, 64)
    mg.setup(A)
    rng = np.random.default_rng(7)
    u = np.zeros(prob.shape)
    w = np.zeros(prob.shape)
    u[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    w[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    a, c = (u * mg.apply(w)).sum(), (w * mg.apply(u)).sum()
    assert abs(a - c) < 1e-12 * abs(a)


@pytest.mark.parametrize("shape", [(16, 12, 10), (70, 40, 24)])
@pytest.mark.parametrize("precision", [64, 32])
def test_multigrid_full_storage(gpu, precision, shape):
    """pfhip_mg_create_full (PCMatrixType FullJacobian) on the full, non-
    symmetric Jacobian of an overland problem against the oracle's full-
    storage hierarchy: the level-0 grid kernels and the one-workgroup
    bottom, plus grid-kernel coarse levels at the larger shape."""
    api = _api()
    prob = cases.hetero(*shape, seed=17, overland=True)
    print(f"mg full precision={precision} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=18)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 0)
    # the PC matrix is non-symmetric: W(i) != E(i - 1) somewhere
    assert np.max(np.abs(J[1][1:-1, 1:-1, 2:-1] - J[2][1:-1, 1:-1, 1:-2])) > \
        1e-3 * np.max(np.abs(J[2]))
    for pre, post, sym in ((1, 1, 1), (2, 2, 0)):
        mgo = O.MG(prob, pre, post, sym, full=True)
        mgo.setup(J)
        mgg = api.GPUMultigrid(gp, pre, post, sym, precision, full=True)
        assert mgg.levels() == len(mgo.levels())
        mgg.setup(J)
        b = np.zeros(prob.shape)
        b[1:-1, 1:-1, 1:-1] = np.random.default_rng(19).normal(size=(prob.nz, prob.ny, prob.nx))
        xo = mgo.apply(b)
        xg = mgg.apply(b)
        tol = 1e-11 if precision == 64 else 2e-5
        assert _relmax(_I(prob, xg), _I(prob, xo)) < tol, (pre, post, sym)


@pytest.mark.parametrize("shape", [(70, 40, 24), (130, 20, 18)])
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("fuse", [0, 1])
def test_multigrid_grid_levels(gpu, precision, fuse, shape, monkeypatch):
    """A hierarchy whose upper levels run as grid kernels and whose bottom
    runs in one workgroup, with the red -> black pairs fused or not: levels
    of even x extent take the two-points-per-lane pair (ragged 124 x 14 x 16
    tiles, one or two tiles across x), odd ones the one-point pair (62 x 14
    x 16), and the 64 x 4 tiles of the other kernels are ragged too.  fuse
    also selects the level residual fused with the z restriction (z-chunks
    of 16 fine planes, ragged and odd plane counts on the coarser levels)
    or the two separate kernels, and the z-marching coarse operators of
    the z-coarsened levels or the point-wise ones."""
    api = _api()
    monkeypatch.setenv("PFHIP_MG_SMALL", "300")
    monkeypatch.setenv("PFHIP_MG_FUSE", str(fuse))
    monkeypatch.setenv("PFHIP_MG_RRZ", str(fuse))
    monkeypatch.setenv("PFHIP_MG_CZ", str(fuse))
    prob = cases.hetero(*shape, seed=17, dir_sides=True)
    print(f"mg grid levels precision={precision} fuse={fuse} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=18)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    rng = np.random.default_rng(19)
    for pre, post, sym in ((1, 1, 1), (2, 2, 0), (3, 1, 0)):
        mgo = O.MG(prob, pre, post, sym)
        mgo.setup(J)
        mgg = api.GPUMultigrid(gp, pre, post, sym, precision)
        assert mgg.levels() == len(mgo.levels())
        mgg.setup(J)
        for rep in range(2):
            b = np.zeros(prob.shape)
            b[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
            xo = mgo.apply(b)
            xg = mgg.apply(b)
            tol = 1e-11 if precision == 64 else 2e-5
            err = _relmax(_I(prob, xg), _I(prob, xo))
            assert err < tol, (pre, post, sym, rep, err)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("variant", [0, 7])
def test_multigrid_graph_replay(gpu, precision, variant, monkeypatch):
    """The V-cycle replayed from its captured hipGraph against the same
    hierarchy launching directly (PFHIP_MG_GRAPH=0): six V-cycles on new
    right-hand sides, the direct hierarchy's V-cycle and a host round trip
    between the replays, every replay identical to the last bit; and the
    oracle bar.  The tile shape is forced, so both hierarchies run the same
    kernels."""
    api = _api()
    monkeypatch.setenv("PFHIP_RBGS_VARIANT", str(variant))
    prob = cases.hetero(70, 40, 24, seed=41, dir_sides=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=42)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    mgo = O.MG(prob, 2, 2, 0)
    mgo.setup(J)
    mgg = api.GPUMultigrid(gp, 2, 2, 0, precision)
    monkeypatch.setenv("PFHIP_MG_GRAPH", "0")
    mgd = api.GPUMultigrid(gp, 2, 2, 0, precision)
    monkeypatch.delenv("PFHIP_MG_GRAPH")
    mgg.setup(J)
    mgd.setup(J)
    rng = np.random.default_rng(43)
    tol = 1e-11 if precision == 64 else 2e-5
    for rep in range(6):
        b = np.zeros(prob.shape)
        b[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
        xg = mgg.apply(b)
        xd = mgd.apply(b)
        assert np.array_equal(_I(prob, xg), _I(prob, xd)), (rep, _relmax(_I(prob, xg),
                                                                         _I(prob, xd)))
        assert _relmax(_I(prob, xg), _I(prob, mgo.apply(b))) < tol, rep


@pytest.mark.parametrize("shape", [(70, 40, 24), (64, 64, 40), (9, 7, 30)])
@pytest.mark.parametrize("precision", [64, 32])
def test_multigrid_bottom_in_lds(gpu, precision, shape, monkeypatch):
    """The bottom of the V-cycle with every level's arrays in LDS (the
    default where they fit) against the oracle and against the bottom in
    global memory (PFHIP_MG_LDS=0): the same per-point arithmetic in the
    same order, so the two agree to the last bit; three sweep schedules."""
    api = _api()
    prob = cases.hetero(*shape, seed=31, dir_sides=True)
    print(f"mg bottom in LDS precision={precision} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=32)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    rng = np.random.default_rng(33)
    for pre, post, sym in ((1, 1, 1), (2, 2, 0), (3, 1, 0)):
        mgo = O.MG(prob, pre, post, sym)
        mgo.setup(J)
        mgl = api.GPUMultigrid(gp, pre, post, sym, precision)
        monkeypatch.setenv("PFHIP_MG_LDS", "0")
        mgg = api.GPUMultigrid(gp, pre, post, sym, precision)
        monkeypatch.delenv("PFHIP_MG_LDS")
        mgl.setup(J)
        mgg.setup(J)
        for rep in range(2):
            b = np.zeros(prob.shape)
            b[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
            xo = mgo.apply(b)
            xl = mgl.apply(b)
            xg = mgg.apply(b)
            tol = 1e-11 if precision == 64 else 2e-5
            err = _relmax(_I(prob, xl), _I(prob, xo))
            assert err < tol, (pre, post, sym, rep, err)
            assert np.array_equal(_I(prob, xl), _I(prob, xg)), (
                pre, post, sym, rep, _relmax(_I(prob, xl), _I(prob, xg)))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("shape", [(70, 40, 24), (131, 20, 18)])
@pytest.mark.parametrize("variant", list(range(17)))
def test_multigrid_pair_shapes(gpu, variant, shape, precision, monkeypatch):
    """Every smoother tile shape the autotuner chooses from (5 tile shapes,
    each with the y neighbours loaded or taken through LDS, and the LDS form
    loading two or three planes ahead), forced on all grid levels (even and
    odd x extents), against the oracle V-cycle."""
    api = _api()
    monkeypatch.setenv("PFHIP_MG_SMALL", "300")
    monkeypatch.setenv("PFHIP_RBGS_VARIANT", str(variant))
    prob = cases.hetero(*shape, seed=23, dir_sides=True)
    print(f"mg pair shape {variant} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p = cases.random_pressure(prob, seed=24)
    J = O.jacobian_eval(op, p, 0.1, 0.1, 1)
    rng = np.random.default_rng(25)
    mgo = O.MG(prob, 2, 2, 1)
    mgo.setup(J)
    mgg = api.GPUMultigrid(gp, 2, 2, 1, precision)
    mgg.setup(J)
    b = np.zeros(prob.shape)
    b[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
    err = _relmax(_I(prob, mgg.apply(b)), _I(prob, mgo.apply(b)))
    assert err < (1e-11 if precision == 64 else 2e-5), err


@pytest.mark.parametrize("precision", [32, 64])
def test_multigrid_fp64_vector_alignment(gpu, precision):
    """The V-cycle's first pair reads the FP64 right-hand side and its last
    pair writes the FP64 result with one 16-byte access per lane when the
    vectors' even-i values are 16-byte aligned (vectors starting 8 bytes
    past a 16-byte boundary, as the solver allocates its Krylov vectors),
    and with two 8-byte accesses otherwise.  Both forms, and a captured
    V-cycle graph called with vectors of the other alignment (launched
    directly then), give the same bits, call after call."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(70, 40, 24, seed=45, dir_sides=True)
    print(f"mg fp64 vector alignment precision={precision}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    J = O.jacobian_eval(op, cases.random_pressure(prob, seed=46), 0.1, 0.1, 1)
    mg = api.GPUMultigrid(gp, 2, 2, 0, precision)
    mg.setup(J)
    mgo = O.MG(prob, 2, 2, 0)
    mgo.setup(J)
    n = gp.ng
    # two allocations with one spare element: views at +0 and +8 bytes
    rawb, rawx = api.DeviceArray(n + 1), api.DeviceArray(n + 1)
    b_al, x_al = rawb.ptr.value, rawx.ptr.value
    b_od, x_od = b_al + 8, x_al + 8
    assert b_al % 16 == 0 and x_al % 16 == 0
    rng = np.random.default_rng(47)
    tol = 1e-11 if precision == 64 else 2e-5

    def run(bp, xp, bh):
        buf = np.ascontiguousarray(bh, dtype=np.float64).ravel()
        api.check(L.pfhip_memcpy_h2d(C.c_void_p(bp), buf.ctypes.data_as(C.c_void_p),
                                     buf.nbytes), "h2d")
        api.check(L.pfhip_mg_apply(mg.h, C.c_void_p(bp), C.c_void_p(xp), None), "apply")
        out = np.empty(n)
        api.check(L.pfhip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(xp),
                                     out.nbytes), "d2h")
        return _I(prob, out.reshape(prob.shape))

    for rep, (bp, xp) in enumerate([(b_od, x_od), (b_al, x_al), (b_od, x_od), (b_al, x_od),
                                    (b_od, x_al), (b_od, x_od), (b_al, x_al)]):
        bh = np.zeros(prob.shape)
        bh[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
        xa = run(b_al if bp != b_al else b_od, x_al if xp != x_al else x_od, bh)
        xb = run(bp, xp, bh)
        assert np.array_equal(xa, xb), (rep, _relmax(xa, xb))
        assert _relmax(xb, _I(prob, mgo.apply(bh))) < tol, rep


@pytest.mark.parametrize("shape", [(70, 40, 24), (128, 30, 37)])
def test_multigrid_pair_prefetch_depth_bitwise(gpu, shape, monkeypatch):
    """The FP32 smoother pair loading its planes two or three planes ahead
    (variants 10-16) performs the arithmetic of the one-plane-ahead form of
    the same tile shape (variants 5-9) in the same order: the V-cycles agree
    to the last bit."""
    api = _api()
    monkeypatch.setenv("PFHIP_MG_SMALL", "300")
    prob = cases.hetero(*shape, seed=27, dir_sides=True)
    print(f"mg pair prefetch depth shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    J = O.jacobian_eval(op, cases.random_pressure(prob, seed=28), 0.1, 0.1, 1)
    b = np.zeros(prob.shape)
    b[1:-1, 1:-1, 1:-1] = np.random.default_rng(29).normal(size=(prob.nz, prob.ny, prob.nx))
    out = {}
    for v in (5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16):
        monkeypatch.setenv("PFHIP_RBGS_VARIANT", str(v))
        mg = api.GPUMultigrid(gp, 2, 2, 0, 32)
        mg.setup(J)
        out[v] = _I(prob, mg.apply(b))
    for deep, base in ((10, 5), (11, 6), (12, 7), (13, 8), (14, 9), (15, 5), (16, 8)):
        assert np.array_equal(out[deep], out[base]), (deep, base)


@pytest.mark.parametrize("shape,dirs", [((11, 9, 6), True), ((150, 37, 40), False),
                                        ((70, 13, 33), True)])
def test_line_search_stats_match_separate_passes(gpu, shape, dirs):
    """The fused line-search entry points against the passes they replace,
    bit for bit: pfhip_residual_step_stats (the trial point's residual with
    F.F and KINStop's max |F| and scaled step length gathered by the
    residual launches) against pfhip_residual_step + pfhip_vec_dot +
    pfhip_kin_stop_norms, and pfhip_vec_dot_scsteplength against
    pfhip_vec_dot + pfhip_vec_scsteplength (KS/kinsol.c:1228-1405,
    1507-1580), and pfhip_vec_kinsol_init_norms against pfhip_vec_max_norm
    + two pfhip_vec_dot."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(*shape, seed=31, dir_sides=dirs)
    print(f"line-search stats shape={shape} dirs={dirs}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(32)
    p_old = cases.random_pressure(prob, seed=33)
    s_old = O.saturation(op, p_old)
    uu_h = p_old + rng.normal(0, 0.02, p_old.shape)
    pp_h = rng.normal(0, 0.05, p_old.shape)
    gp.bc_pressure(0.3)
    dpo, dso, uu, pp = gp.vec(p_old), gp.vec(s_old), gp.vec(uu_h), gp.vec(pp_h)
    vp, dbl = C.c_void_p, C.c_double
    L.pfhip_residual_step.argtypes = [vp, vp, vp, dbl, vp, vp, vp, vp, dbl, vp]
    L.pfhip_residual_step_stats.argtypes = [vp, vp, vp, dbl, vp, vp, vp, vp, dbl, vp, vp, vp,
                                            vp]
    L.pfhip_kin_stop_norms.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.pfhip_vec_dot_scsteplength.argtypes = [vp, vp, vp, vp, vp, vp]
    L.pfhip_vec_scsteplength.argtypes = [vp, vp, vp, vp, vp]
    L.pfhip_vec_dot.argtypes = [vp, vp, vp, vp, vp]
    for rl in (1.0, 0.37):
        un_a, F_a, un_b, F_b = gp.vec(), gp.vec(), gp.vec(), gp.vec()
        ff, fmax, step = dbl(), dbl(), dbl()
        api.check(L.pfhip_residual_step_stats(gp.h, uu.ptr, pp.ptr, rl, un_a.ptr, F_a.ptr,
                                              dpo.ptr, dso.ptr, 0.1, C.byref(ff),
                                              C.byref(fmax), C.byref(step), None), "stats")
        api.check(L.pfhip_residual_step(gp.h, uu.ptr, pp.ptr, rl, un_b.ptr, F_b.ptr, dpo.ptr,
                                        dso.ptr, 0.1, None), "step")
        ff_b, fmax_b, step_b = dbl(), dbl(), dbl()
        api.check(L.pfhip_vec_dot(gp.h, F_b.ptr, F_b.ptr, C.byref(ff_b), None), "dot")
        api.check(L.pfhip_kin_stop_norms(gp.h, F_b.ptr, un_b.ptr, uu.ptr, C.byref(fmax_b),
                                         C.byref(step_b), None), "kin_stop_norms")
        assert np.array_equal(_I(prob, gp.get(F_a)), _I(prob, gp.get(F_b)))
        assert np.array_equal(_I(prob, gp.get(un_a)), _I(prob, gp.get(un_b)))
        assert (ff.value, fmax.value, step.value) == (ff_b.value, fmax_b.value, step_b.value)
        # the maxima are those of the host arrays
        Fh, uh = _I(prob, gp.get(F_b)), _I(prob, gp.get(un_b))
        assert fmax.value == np.max(np.abs(Fh))
        assert step.value == np.max(np.abs((uh - _I(prob, uu_h)) / (1.0 + np.abs(uh))))
    pp2, rlen = dbl(), dbl()
    api.check(L.pfhip_vec_dot_scsteplength(gp.h, uu.ptr, pp.ptr, C.byref(pp2), C.byref(rlen),
                                           None), "dot_scsteplength")
    pp2_b, rlen_b = dbl(), dbl()
    api.check(L.pfhip_vec_dot(gp.h, pp.ptr, pp.ptr, C.byref(pp2_b), None), "dot")
    api.check(L.pfhip_vec_scsteplength(gp.h, uu.ptr, pp.ptr, C.byref(rlen_b), None), "scstep")
    assert (pp2.value, rlen.value) == (pp2_b.value, rlen_b.value)
    # KINSOL's start: max |F|, F.F and u.u from one pass against the three
    # separate reductions (KS/kinsol.c:518-522, 640-644)
    L.pfhip_vec_kinsol_init_norms.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.pfhip_vec_max_norm.argtypes = [vp, vp, vp, vp]
    fm, f2, u2 = dbl(), dbl(), dbl()
    api.check(L.pfhip_vec_kinsol_init_norms(gp.h, uu.ptr, F_a.ptr, C.byref(fm), C.byref(f2),
                                            C.byref(u2), None), "init_norms")
    fm_b, f2_b, u2_b = dbl(), dbl(), dbl()
    api.check(L.pfhip_vec_max_norm(gp.h, F_a.ptr, C.byref(fm_b), None), "max_norm")
    api.check(L.pfhip_vec_dot(gp.h, F_a.ptr, F_a.ptr, C.byref(f2_b), None), "dot F")
    api.check(L.pfhip_vec_dot(gp.h, uu.ptr, uu.ptr, C.byref(u2_b), None), "dot u")
    assert (fm.value, f2.value, u2.value) == (fm_b.value, f2_b.value, u2_b.value)
    assert fm.value == np.max(np.abs(_I(prob, gp.get(F_a))))


@pytest.mark.parametrize("m", [1, 5, 16, 23])
def test_combine_step_matches_separate_passes(gpu, m):
    """The flexible GMRES's solution update with the line search's terms:
    pfhip_vec_combine_step writing x = sum y_k Z_k (x not read) or adding
    to x, with x.x and max |x| / (1 + |u|), against pfhip_vec_const +
    pfhip_vec_combine + pfhip_vec_dot + pfhip_vec_scsteplength, bit for bit
    (KS/spgmr.c:420-450, KS/kinsol.c:1240-1262); m > 16 takes two chunks."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(37, 21, 11, seed=51, dir_sides=True)
    print(f"combine_step m={m}", flush=True)
    gp = api.GPUProblem(prob)
    rng = np.random.default_rng(52 + m)
    Z = [gp.vec(rng.normal(size=prob.shape)) for _ in range(m)]
    u = gp.vec(rng.normal(0, 3, size=prob.shape))
    y = (C.c_double * m)(*rng.normal(size=m))
    zp = (C.c_void_p * m)(*[z.ptr.value for z in Z])
    vp, dbl = C.c_void_p, C.c_double
    L.pfhip_vec_combine_step.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp, vp]
    L.pfhip_vec_combine.argtypes = [vp, vp, C.c_int, vp, vp, vp]
    L.pfhip_vec_const.argtypes = [vp, dbl, vp, vp]
    L.pfhip_vec_dot.argtypes = [vp, vp, vp, vp, vp]
    L.pfhip_vec_scsteplength.argtypes = [vp, vp, vp, vp, vp]
    x0 = rng.normal(size=prob.shape)
    for zero in (1, 0):
        # x holds garbage (zero = 1, never read) or a start value (zero = 0)
        xa = gp.vec(x0 if not zero else rng.normal(size=prob.shape) * 1e300)
        xb = gp.vec(x0)
        if zero:
            api.check(L.pfhip_vec_const(gp.h, 0.0, xb.ptr, None), "const")
        d, st = dbl(), dbl()
        api.check(L.pfhip_vec_combine_step(gp.h, zp, m, y, xa.ptr, u.ptr, zero, C.byref(d),
                                           C.byref(st), None), "combine_step")
        api.check(L.pfhip_vec_combine(gp.h, zp, m, y, xb.ptr, None), "combine")
        d2, st2 = dbl(), dbl()
        api.check(L.pfhip_vec_dot(gp.h, xb.ptr, xb.ptr, C.byref(d2), None), "dot")
        api.check(L.pfhip_vec_scsteplength(gp.h, u.ptr, xb.ptr, C.byref(st2), None), "scstep")
        ia, ib = _I(prob, gp.get(xa)), _I(prob, gp.get(xb))
        assert np.array_equal(ia, ib), (zero, _relmax(ia, ib))
        assert (d.value, st.value) == (d2.value, st2.value), zero
        assert st.value == np.max(np.abs(ib / (1.0 + np.abs(_I(prob, gp.get(u))))))


@pytest.mark.parametrize("k", [1, 5, 9, 20])
def test_cgs2_scaled_basis_matches_scaled_vectors(gpu, k):
    """CGS2 with the basis kept unnormalised (vector q read as sc[q] * V[q],
    pfhip_vec_cgs2_scaled) against CGS2 on the vectors scaled first by
    pfhip_vec_scale (the flexible GMRES's former normalisation pass): h1,
    h2, ||w||^2 and w agree bit for bit (k <= 8 takes the fused middle
    pass, k > 16 two chunks)."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(37, 21, 11, seed=61, dir_sides=True)
    print(f"cgs2 scaled k={k}", flush=True)
    gp = api.GPUProblem(prob)
    rng = np.random.default_rng(62 + k)
    raw = [rng.normal(size=prob.shape) * rng.uniform(0.1, 10.0) for _ in range(k)]
    sc = rng.uniform(-3.0, 3.0, size=k)
    Vr = [gp.vec(a) for a in raw]
    Vs = [gp.vec(a) for a in raw]
    vp, dbl = C.c_void_p, C.c_double
    L.pfhip_vec_scale.argtypes = [vp, dbl, vp, vp, vp]
    for q in range(k):
        api.check(L.pfhip_vec_scale(gp.h, sc[q], Vs[q].ptr, Vs[q].ptr, None), "scale")
    w0 = rng.normal(size=prob.shape)
    wa, wb = gp.vec(w0), gp.vec(w0)
    L.pfhip_vec_cgs2_scaled.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp, vp]
    L.pfhip_vec_cgs2.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp]
    pa = (vp * k)(*[v.ptr.value for v in Vr])
    pb_ = (vp * k)(*[v.ptr.value for v in Vs])
    sca = (dbl * k)(*sc)
    h1a, h2a, h1b, h2b = (dbl * k)(), (dbl * k)(), (dbl * k)(), (dbl * k)()
    na, nb = dbl(), dbl()
    api.check(L.pfhip_vec_cgs2_scaled(gp.h, pa, sca, k, wa.ptr, h1a, h2a, C.byref(na), None),
              "cgs2_scaled")
    api.check(L.pfhip_vec_cgs2(gp.h, pb_, k, wb.ptr, h1b, h2b, C.byref(nb), None), "cgs2")
    assert list(h1a) == list(h1b) and list(h2a) == list(h2b) and na.value == nb.value
    assert np.array_equal(_I(prob, gp.get(wa)), _I(prob, gp.get(wb)))


@pytest.mark.parametrize("precision,shape", [(32, (70, 40, 24)), (64, (70, 40, 24)),
                                             (32, (71, 30, 20))])
def test_mg_apply_scaled_matches_scaled_vector(gpu, precision, shape, monkeypatch):
    """pfhip_mg_apply_scaled(b, c) applies the factor where the V-cycle reads
    b (the first pair's FP64 right-hand side, or the conversion pass for odd
    x extents): the result equals pfhip_mg_apply of the vector scaled by
    pfhip_vec_scale, bit for bit, from the captured graph (several
    factors in a row, with 1.0 between) and launching directly."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(*shape, seed=63, dir_sides=True)
    print(f"mg apply scaled precision={precision} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    J = O.jacobian_eval(op, cases.random_pressure(prob, seed=64), 0.1, 0.1, 1)
    vp, dbl = C.c_void_p, C.c_double
    L.pfhip_mg_apply_scaled.argtypes = [vp, vp, dbl, vp, vp]
    L.pfhip_vec_scale.argtypes = [vp, dbl, vp, vp, vp]
    rng = np.random.default_rng(65)
    for graph in ("1", "0"):
        monkeypatch.setenv("PFHIP_MG_GRAPH", graph)
        mg = api.GPUMultigrid(gp, 2, 2, 0, precision)
        mg.setup(J)
        for c in (0.37, 1.0, -2.5e-3, 7.0):
            bh = np.zeros(prob.shape)
            bh[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
            b, b2, xa, xb = gp.vec(bh), gp.vec(bh), gp.vec(), gp.vec()
            api.check(L.pfhip_mg_apply_scaled(mg.h, b.ptr, c, xa.ptr, None), "apply_scaled")
            api.check(L.pfhip_vec_scale(gp.h, c, b2.ptr, b2.ptr, None), "scale")
            api.check(L.pfhip_mg_apply(mg.h, b2.ptr, xb.ptr, None), "apply")
            ia, ib = _I(prob, gp.get(xa)), _I(prob, gp.get(xb))
            assert np.array_equal(ia, ib), (graph, c, _relmax(ia, ib))


@pytest.mark.parametrize("precision,shape", [(32, (70, 40, 24)), (64, (70, 40, 24)),
                                             (32, (71, 30, 20)), (32, (128, 30, 37))])
def test_mg_apply_dq_sums(gpu, precision, shape, monkeypatch):
    """pfhip_mg_apply_dq: the V-cycle's result is that of
    pfhip_mg_apply_scaled bit for bit, and the sums it leaves for the
    difference-quotient J.v (u.x, x.x, sum |x| over the grid's cells,
    gathered by the last pair where it writes x in FP64, by a pass of their
    own for odd x extents) agree with numpy's to 1e-13 (summation order
    differs), from the captured graph and launching directly; the J.v of
    pfhip_residual_jv_ready with them agrees with pfhip_residual_jv_dq
    (KS/kinspgmr.c:481-531) to 1e-12."""
    import ctypes as C
    api = _api()
    L = api.lib()
    prob = cases.hetero(*shape, seed=71, dir_sides=True)
    print(f"mg apply dq precision={precision} shape={shape}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=72)
    s_old = O.saturation(op, p_old)
    J = O.jacobian_eval(op, p_old, 0.1, 0.1, 1)
    vp, dbl = C.c_void_p, C.c_double
    L.pfhip_mg_apply_scaled.argtypes = [vp, vp, dbl, vp, vp]
    L.pfhip_mg_apply_dq.argtypes = [vp, vp, dbl, vp, vp, vp]
    L.pfhip_dq_slots.argtypes = [vp, vp, vp]
    L.pfhip_residual_jv_ready.argtypes = [vp, vp, vp, dbl, vp, vp, vp, vp, dbl, vp]
    L.pfhip_residual_jv_dq.argtypes = [vp, vp, vp, dbl, vp, vp, vp, vp, dbl, vp]
    rng = np.random.default_rng(73)
    gp.bc_pressure(0.3)
    dpo, dso = gp.vec(p_old), gp.vec(s_old)
    uh = p_old + rng.normal(0, 0.02, p_old.shape)
    u = gp.vec(uh)
    F = gp.vec(rng.normal(size=prob.shape))
    for graph in ("1", "0"):
        monkeypatch.setenv("PFHIP_MG_GRAPH", graph)
        mga = api.GPUMultigrid(gp, 2, 2, 0, precision)
        mgb = api.GPUMultigrid(gp, 2, 2, 0, precision)
        mga.setup(J)
        mgb.setup(J)
        for rep, c in enumerate((1.0, 0.3, -4.0)):
            bh = np.zeros(prob.shape)
            bh[1:-1, 1:-1, 1:-1] = rng.normal(size=(prob.nz, prob.ny, prob.nx))
            b, xa, xb = gp.vec(bh), gp.vec(), gp.vec()
            api.check(L.pfhip_mg_apply_dq(mga.h, b.ptr, c, xa.ptr, u.ptr, None), "apply_dq")
            sums = (dbl * 3)()
            api.check(L.pfhip_dq_slots(gp.h, sums, None), "dq_slots")
            api.check(L.pfhip_mg_apply_scaled(mgb.h, b.ptr, c, xb.ptr, None), "apply_scaled")
            ia, ib = _I(prob, gp.get(xa)), _I(prob, gp.get(xb))
            assert np.array_equal(ia, ib), (graph, rep, _relmax(ia, ib))
            ui = _I(prob, uh)
            ref = (np.sum(ui * ia), np.sum(ia * ia), np.sum(np.abs(ia)))
            for q in range(3):
                assert abs(sums[q] - ref[q]) <= 1e-13 * max(abs(ref[q]), 1e-300) + \
                    1e-13 * np.sum(np.abs(ui * ia)), (graph, rep, q, sums[q], ref[q])
            # the J.v from the sums left in the slots against the separate pass
            za, zb = gp.vec(), gp.vec()
            api.check(L.pfhip_residual_jv_ready(gp.h, u.ptr, xa.ptr, 1.49e-8, F.ptr, za.ptr,
                                                dpo.ptr, dso.ptr, 0.1, None), "jv_ready")
            api.check(L.pfhip_residual_jv_dq(gp.h, u.ptr, xa.ptr, 1.49e-8, F.ptr, zb.ptr,
                                             dpo.ptr, dso.ptr, 0.1, None), "jv_dq")
            za_, zb_ = _I(prob, gp.get(za)), _I(prob, gp.get(zb))
            assert _relmax(za_, zb_, np.max(np.abs(zb_)) + 1e-300) < 1e-12, (graph, rep)


@pytest.mark.parametrize("overland,dirs", [(False, True), (True, False)])
def test_residual_multi_tile(gpu, overland, dirs):
    """Residual on a grid spanning several 64-wide tiles, ragged tile rows
    and z-chunks."""
    api = _api()
    prob = cases.hetero(150, 37, 40, seed=23, overland=overland, dir_sides=dirs)
    print(f"residual multi-tile overland={overland} dir={dirs}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=24)
    s_old = O.saturation(op, p_old)
    p = p_old + np.random.default_rng(25).normal(0, 0.01, p_old.shape)
    Fg = gp.nl_function_eval(p, p_old, s_old, 0.1, 0.2)
    Fo = O.nl_function_eval(op, p, p_old, s_old, 0.1, 0.2)
    scale = np.max(np.abs(_I(prob, Fo))) + 1e-30
    assert _relmax(_I(prob, Fg), _I(prob, Fo), scale) < 1e-12


def _compare_runs(prob, opts, p0, dt, nsteps, precision=64, krylov=0, ptol=1e-9,
                  newton_tol=1):
    api = _api()
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    r, pho, sho, logo = O.advance(op, opts, p0, 0.0, dt, nsteps)
    assert r == 0
    sv = api.GPUSolver(gp, opts, precision, krylov)
    phg, shg, logg = sv.advance(p0, 0.0, dt, nsteps)
    worst = 0.0
    for s in range(nsteps + 1):
        a, b = _I(prob, phg[s]), _I(prob, pho[s])
        e = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))
        es = _relmax(_I(prob, shg[s]), _I(prob, sho[s]))
        worst = max(worst, e, es)
        assert e < ptol, (s, e)
        assert es < ptol, (s, es)
    newton_o = np.array(logo.newton[:nsteps])
    assert np.all(np.abs(logg["newton"] - newton_o) <= newton_tol), (logg["newton"], newton_o)
    return worst, logg, logo


def test_time_loop_growth_matches_oracle(gpu):
    """pfhip_advance_richards_ts on the device solve of TiltedV_OverlandKin
    with growing steps cut at the rain's end, at dump times and at StopTime:
    the same steps as the oracle's loop and the same pressures."""
    from parflow_amd.problem import TimeControl
    api = _api()
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    tc = TimeControl(start_time=0.0, stop_time=0.5, base_unit=0.05, dump_interval=0.1,
                     ts_type="Growth", initial_step=0.01, growth_factor=1.5, min_step=0.005,
                     max_step=0.05)
    print("time loop growth tiltedv", flush=True)
    gp = api.GPUProblem(prob)
    sv = api.GPUSolver(gp, opts)
    ph, sh, rows = sv.run(p0, tc)
    op = O.OracleProblem(prob)
    pho, sho, rowso = O.advance_ts(op, opts, p0, tc)
    assert len(rows) == len(rowso) == 18
    for k, (r, ro) in enumerate(zip(rows, rowso)):
        assert (r["time"], r["dt"], r["dt_info"], r["dumped"]) == ro[:4], k
        assert abs(r["newton"] - ro[4]) <= 1, k
        a, b = _I(prob, ph[k + 1]), _I(prob, pho[k + 1])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-9, k
    assert [r["dt_info"] for r in rows][-1] == "f" and rows[-1]["time"] == 0.5
    # history=False runs the whole loop in one call: same end state
    p_end, s_end, rows2 = sv.run(p0, tc, history=False)
    assert [(r["time"], r["dt"]) for r in rows2] == [(r["time"], r["dt"]) for r in rows]
    assert np.max(np.abs(_I(prob, p_end) - _I(prob, ph[-1]))) < 1e-12


def test_time_loop_schedule_matches_reference_log(gpu):
    """The device time loop with LW_var_dz's timing keys (constant step 10,
    DumpInterval 20, StopTime 200, the rain/recession cycle on the top flux
    face) steps, dumps and ends exactly as the reference's own run log
    (test/correct_output/LW_var_dz.out.log)."""
    import dataclasses
    from parflow_amd.problem import BC_FLUX, BCSpec, TimeControl
    api = _api()
    prob = cases.hetero(6, 5, 4, seed=51)
    bc = list(prob.bc)
    bc[5] = BCSpec(BC_FLUX, (-0.01, 0.0), (10, 20), 14)
    prob = dataclasses.replace(prob, bc=bc, base_unit=10.0)
    tc = TimeControl(start_time=0.0, stop_time=200.0, base_unit=10.0, dump_interval=20.0,
                     ts_type="Constant", value=10.0)
    print("time loop LW_var_dz schedule", flush=True)
    gp = api.GPUProblem(prob)
    sv = api.GPUSolver(gp, cases.solver())
    _, _, rows = sv.run(prob.hydrostatic(1.0, 0.0), tc, history=False)
    g = np.load(os.path.join(HERE, "golden", "time_logs.npz"))
    time, dtl, info, dump = (g["LW_var_dz_time"], g["LW_var_dz_dt"], g["LW_var_dz_info"],
                             g["LW_var_dz_dump"])
    assert len(rows) == len(time) - 1
    fileno = 0
    for k, r in enumerate(rows, start=1):
        assert abs(r["time"] - time[k]) <= 5e-7 * time[k] and abs(r["dt"] - dtl[k]) <= 5e-7 * dtl[k]
        assert r["dt_info"] == chr(info[k]) and r["conv_failures"] == 0
        fileno += r["dumped"]
        assert (fileno if r["dumped"] else -1) == dump[k], k


@pytest.mark.parametrize("pv,lateral", [(0.0, False), (0.0, True), (1e-5, True)])
def test_surface_predictor(gpu, pv, lateral):
    """pfhip_surface_predictor against the oracle on random top-layer
    states: about half the unsaturated top cells are predicted."""
    api = _api()
    prob = cases.hetero(37, 21, 6, seed=61, overland=True)
    src = cases.hetero_src(prob, 62) * 40.0
    gp = api.GPUProblem(prob, src=src)
    op = O.OracleProblem(prob, src=src)
    rng = np.random.default_rng(63)
    p = cases.random_pressure(prob, seed=64)
    s = O.saturation(op, p)
    nx, ny, nz = prob.nx, prob.ny, prob.nz
    vel = (rng.normal(0, 0.02, (nz, ny, nx + 1)), rng.normal(0, 0.02, (nz, ny + 1, nx)),
           rng.normal(0, 0.02, (nz + 1, ny, nx)))
    q = (rng.normal(0, 0.05, (ny + 2, nx + 2)), rng.normal(0, 0.05, (ny + 2, nx + 2)))
    print(f"surface_predictor pv={pv} lateral={lateral}", flush=True)
    got = gp.surface_predictor(p, s, vel, q, 0.1, pv, lateral)
    ref = O.surface_predictor(op, p, s, vel, q, 0.1, pv, lateral)
    changed = ref != p
    assert 0.1 < changed[nz].mean() < 0.9 and not changed[:nz].any()
    assert np.array_equal(got != p, changed)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref[changed]))


@pytest.mark.parametrize("pv", [0.0, 1e-5])
def test_time_loop_surface_predictor_matches_oracle(gpu, pv):
    """The time loop with Solver.SurfacePredictor on TiltedV_OverlandKin with
    a strong top source: the predictor changes the first step's initial
    guess (5 Newton iterations without it, 11 or 6 with it, in the oracle),
    and the device loop follows the oracle's step for step."""
    from parflow_amd.problem import TimeControl
    api = _api()
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    src = np.zeros(prob.shape)
    src[1:-1, 1:-1, 1:-1] = 0.3
    tc = TimeControl(start_time=0.0, stop_time=0.3, base_unit=0.05, dump_interval=0.1,
                     ts_type="Constant", value=0.05, surface_predictor=True,
                     surface_predictor_pressure=pv, surface_lateral_flows=True)
    print(f"time loop surface predictor pv={pv}", flush=True)
    gp = api.GPUProblem(prob, src=src)
    sv = api.GPUSolver(gp, opts)
    ph, sh, rows = sv.run(p0, tc)
    op = O.OracleProblem(prob, src=src)
    pho, sho, rowso = O.advance_ts(op, opts, p0, tc)
    assert rowso[0][4] == (11 if pv == 0.0 else 6)
    assert len(rows) == len(rowso)
    for k, (r, ro) in enumerate(zip(rows, rowso)):
        assert (r["time"], r["dt"], r["dt_info"], r["dumped"]) == ro[:4], k
        assert abs(r["newton"] - ro[4]) <= 1 and r["conv_failures"] == ro[7], k
        a, b = _I(prob, ph[k + 1]), _I(prob, pho[k + 1])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-9, k


def test_tiltedv_jacobian_true_on_gpu(gpu):
    """overland_tiltedV_KWE.py's TiltedV_OverlandKin "Jacobian True" variant
    (analytic J.v with the kinematic surface couplings, PFSymmetric PC) on
    the plain device path: against the oracle per step and against the
    reference's correct_output."""
    import dataclasses
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    opts = dataclasses.replace(opts, use_jacobian=True)
    print("tiltedv UseJacobian True", flush=True)
    worst, lg, lo = _compare_runs(prob, opts, p0, dt, n)
    api = _api()
    gp = api.GPUProblem(prob)
    ph, sh, log = api.GPUSolver(gp, opts).advance(p0, 0.0, dt, n)
    g = np.load(os.path.join(HERE, "golden", "tiltedv_overlandkin.npz"))
    for d in range(21):
        s = max(2 * d - 1, 0)
        a, b = prob.interior(ph[s]), g["press"][d]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-7, d


def test_tiltedv_jacobian_true_nonsymmetric_pc_on_gpu(gpu):
    """overland_tiltedV_KWE.py's "Jacobian True Nonsymmetric Preconditioner"
    variant (:417-440: UseJacobian True, PCMatrixType FullJacobian): the
    full-storage hierarchy built from the full Jacobian on the device,
    against the oracle per step and against the reference's correct_output
    (the same files every variant of the deck is checked against)."""
    import dataclasses
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    opts = dataclasses.replace(opts, use_jacobian=True, pc_matrix="FullJacobian")
    print("tiltedv UseJacobian True FullJacobian", flush=True)
    worst, lg, lo = _compare_runs(prob, opts, p0, dt, n)
    # the full PC matrix cuts the linear iterations of the PFSymmetric run
    sym = dataclasses.replace(opts, pc_matrix="PFSymmetric")
    r, _, _, los = O.advance(O.OracleProblem(prob), sym, p0, 0.0, dt, n)
    assert int(np.sum(lg["gmres"])) < 0.5 * int(np.sum(los.gmres[:n]))
    api = _api()
    gp = api.GPUProblem(prob)
    ph, sh, log = api.GPUSolver(gp, opts).advance(p0, 0.0, dt, n)
    g = np.load(os.path.join(HERE, "golden", "tiltedv_overlandkin.npz"))
    for d in range(21):
        s = max(2 * d - 1, 0)
        a, b = prob.interior(ph[s]), g["press"][d]
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-7, d


@pytest.mark.parametrize("use_jac", [False, True])
def test_kinsol_full_jacobian_pc_hetero_overland(gpu, use_jac):
    """PCMatrixType FullJacobian on a heterogeneous overland problem: the
    plain device path against the oracle step for step, and the mixed path
    (FGMRES + FP32 full-storage V-cycle) within the solver tolerance of
    the plain one."""
    api = _api()
    prob = cases.hetero(20, 16, 8, seed=23, overland=True)
    opts = cases.solver(use_jacobian=use_jac, residual_tol=1e-8, krylov_dim=25,
                        precond="PFMG", pc_matrix="FullJacobian")
    p0 = prob.hydrostatic(0.5, 0.0)
    print(f"advance hetero overland FullJacobian use_jac={use_jac}", flush=True)
    _compare_runs(prob, opts, p0, 0.1, 4)
    gp = api.GPUProblem(prob)
    tight = cases.solver(use_jacobian=use_jac, residual_tol=1e-12, krylov_dim=25,
                         precond="PFMG", pc_matrix="FullJacobian")
    pp, _, _ = api.GPUSolver(gp, tight, 64, 0).advance(p0, 0.0, 0.1, 3)
    pm, _, _ = api.GPUSolver(gp, tight, 32, 1).advance(p0, 0.0, 0.1, 3)
    for s in range(4):
        a, b = _I(prob, pm[s]), _I(prob, pp[s])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-8, s


def test_kinsol_plain_matches_oracle_default_richards(gpu):
    """Reference deck default_richards (UseJacobian True) on the plain path."""
    prob, opts, p0, dt, n = ref_decks.default_richards()
    print("advance default_richards", flush=True)
    worst, lg, lo = _compare_runs(prob, opts, p0, dt, 10)
    print("worst rel diff", worst)


def test_kinsol_plain_matches_oracle_tiltedv(gpu):
    """Reference deck TiltedV_OverlandKin (DQ J.v, overland BC)."""
    prob, opts, p0, dt, n = ref_decks.tiltedv_kin()
    print("advance tiltedv_kin", flush=True)
    _compare_runs(prob, opts, p0, dt, 12)


@pytest.mark.parametrize("use_jac", [False, True])
def test_kinsol_plain_matches_oracle_hetero(gpu, use_jac):
    prob = cases.hetero(14, 12, 10, seed=21, dir_sides=True)
    opts = cases.solver(use_jacobian=use_jac, residual_tol=1e-8, krylov_dim=25,
                        precond="PFMG")
    p0 = prob.hydrostatic(0.5, 0.0)
    print(f"advance hetero use_jac={use_jac}", flush=True)
    _compare_runs(prob, opts, p0, 0.2, 4)


def test_golden_default_richards_on_gpu(gpu):
    """The GPU solve against the reference's own outputs."""
    api = _api()
    prob, opts, p0, dt, n = ref_decks.default_richards()
    gp = api.GPUProblem(prob)
    sv = api.GPUSolver(gp, opts)
    ph, sh, log = sv.advance(p0, 0.0, dt, 20)
    g = np.load(os.path.join(HERE, "golden", "default_richards.npz"))
    for i, s in enumerate(g["steps"]):
        if s > 20:
            break
        a = prob.interior(ph[s])
        e = np.max(np.abs(a - g["press"][i]) / np.maximum(np.abs(a), np.abs(g["press"][i])))
        assert e < 6e-7, (s, e)
    assert log["newton"][0] == 3


def test_pfidb_deck_to_pfb_outputs(gpu, tmp_path):
    """A .pfidb deck in, ParFlow-named PFB outputs out (advance_deck),
    compared with the reference's correct_output for default_richards."""
    from parflow_amd import pfidb
    from parflow_amd.pfb import read_pfb
    import test_pfidb
    keys = test_pfidb.default_richards_keys()
    deck = tmp_path / "default_richards.pfidb"
    pfidb.write_pfidb(keys, str(deck))
    print("advance_deck default_richards.pfidb 10 steps", flush=True)
    rows = pfidb.advance_deck(str(deck), str(tmp_path / "out"), nsteps=10)
    g = np.load(os.path.join(HERE, "golden", "default_richards.npz"))
    for i, s in enumerate(g["steps"]):
        if s > 10:
            break
        hdr, a = read_pfb(str(tmp_path / "out" / f"default_richards.out.press.{s:05d}.pfb"))
        assert (hdr["nx"], hdr["ny"], hdr["nz"]) == (18, 15, 8)
        e = np.max(np.abs(a - g["press"][i]) / np.maximum(np.abs(a), np.abs(g["press"][i])))
        assert e < 6e-7, (s, e)
        _, sat = read_pfb(str(tmp_path / "out" / f"default_richards.out.satur.{s:05d}.pfb"))
        assert np.all((sat > 0.2) & (sat <= 0.99))
    assert rows[0]["newton"] == 3 and len(rows) == 10 and all(r["dumped"] for r in rows)
    # Solver.PrintVelocities: zero at step 0, then the reference's own
    # velocities to the level the pressure agrees at
    gv = np.load(os.path.join(HERE, "golden", "default_richards_vel.npz"))
    for c in "xyz":
        _, v = read_pfb(str(tmp_path / "out" / f"default_richards.out.vel{c}.00000.pfb"))
        assert np.all(v == 0.0)
    for i, s in enumerate(gv["steps"]):
        if s > 10:
            break
        ref = [gv["velx"][i], gv["vely"][i], gv["velz"][i]]
        scale = max(np.max(np.abs(r)) for r in ref)
        for c, r in zip("xyz", ref):
            hdr, v = read_pfb(str(tmp_path / "out" / f"default_richards.out.vel{c}.{s:05d}.pfb"))
            assert v.shape == r.shape
            assert np.max(np.abs(v - r)) < 1e-6 * scale, (s, c)


def test_pfidb_growth_deck_dump_files(gpu, tmp_path):
    """A Growth deck with a time DumpInterval: files are numbered by dump
    (00000 the initial state), each holding the state at its dump time,
    against the oracle's time loop on the same keys."""
    from parflow_amd import pfidb
    from parflow_amd.pfb import read_pfb
    import test_pfidb
    keys = test_pfidb.default_richards_keys()
    keys.update({"TimeStep.Type": "Growth", "TimeStep.InitialStep": 0.0005,
                 "TimeStep.GrowthFactor": 1.5, "TimeStep.MinStep": 0.0001,
                 "TimeStep.MaxStep": 0.002, "TimingInfo.DumpInterval": 0.003,
                 "TimingInfo.StopTime": 0.01})
    deck = tmp_path / "growth.pfidb"
    pfidb.write_pfidb(keys, str(deck))
    print("advance_deck growth deck", flush=True)
    rows = pfidb.advance_deck(str(deck), str(tmp_path / "out"))
    prob, opts, p0, dt, n = pfidb.problem_from_pfidb(str(deck))
    tc = pfidb.time_control_from_pfidb(str(deck))
    op = O.OracleProblem(prob)
    pho, sho, rowso = O.advance_ts(op, opts, p0, tc)
    assert [(r["time"], r["dt"], r["dt_info"], r["dumped"]) for r in rows] == \
        [ro[:4] for ro in rowso]
    dumps = [0] + [k + 1 for k, r in enumerate(rows) if r["dumped"]]
    # dumps at t = .003 .006 .009; the final step to StopTime .01 falls
    # short of the next dump time and writes nothing (no forced last dump)
    assert len(dumps) == 4 and rows[-1]["dt_info"] == "f" and not rows[-1]["dumped"]
    for fileno, s in enumerate(dumps):
        _, a = read_pfb(str(tmp_path / "out" / f"growth.out.press.{fileno:05d}.pfb"))
        b = prob.interior(pho[s])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-9, fileno
    assert not os.path.exists(tmp_path / "out" / f"growth.out.press.{len(dumps):05d}.pfb")


def test_mixed_precision_close_to_plain(gpu):
    """FGMRES + FP32 V-cycle vs the plain FP64 path on a tight tolerance."""
    api = _api()
    prob = cases.hetero(16, 14, 12, seed=31, dir_sides=True)
    opts = cases.solver(residual_tol=1e-12, krylov_dim=30, precond="PFMG",
                        num_pre_relax=2, num_post_relax=2)
    p0 = prob.hydrostatic(0.5, 0.0)
    gp = api.GPUProblem(prob)
    plain = api.GPUSolver(gp, opts, 64, 0)
    mixed = api.GPUSolver(gp, opts, 32, 1)
    pp, sp_, lp = plain.advance(p0, 0.0, 0.2, 3)
    pm, sm, lm = mixed.advance(p0, 0.0, 0.2, 3)
    for s in range(4):
        a, b = _I(prob, pm[s]), _I(prob, pp[s])
        e = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))
        assert e < 1e-8, (s, e)


def test_ensemble_matches_one_at_a_time(gpu):
    """Hillslope ensemble members advanced together (batched launches, one
    coroutine per member) against the same members advanced one at a time
    on the single-problem mixed path, and the smallest member against the
    oracle's plain FP64 solve (the mixed-vs-plain bar)."""
    api = _api()
    from parflow_amd.decks import c4_member
    nm, nsteps = 6, 3
    members = [c4_member(m) for m in range(nm)]
    for m, (prob, _, _, _) in enumerate(members):
        print(f"member {m}: {prob.nx}x{prob.ny}x{prob.nz}", flush=True)
    ens = api.GPUEnsemble([(p, o) for p, o, _, _ in members])
    dps = [ens.problems[i].vec(members[i][2]) for i in range(nm)]
    dss = [ens.problems[i].vec() for i in range(nm)]
    print("ensemble advance", flush=True)
    lg = ens.advance_device(dps, dss, 0.0, [m[3] for m in members], nsteps)
    rounds, launches = ens.stats()
    print("rounds", rounds, "launches", launches, "newton", lg["newton"].tolist(), flush=True)
    assert np.all(lg["steps"] == nsteps)
    for i, (prob, opts, p0, dt) in enumerate(members):
        gp = api.GPUProblem(prob)
        sv = api.GPUSolver(gp, opts, 32, 1)
        pg, sg, l1 = sv.advance(p0, 0.0, dt, nsteps, history=False)
        pe = _I(prob, ens.problems[i].get(dps[i]))
        se = _I(prob, ens.problems[i].get(dss[i]))
        e = np.max(np.abs(pe - _I(prob, pg)) / np.maximum(np.abs(_I(prob, pg)), 1e-3))
        es = _relmax(se, _I(prob, sg))
        print(f"member {i}: rel diff p {e:.3e} S {es:.3e}", flush=True)
        assert e < 1e-10 and es < 1e-10, (i, e, es)
        assert np.all(np.abs(lg["newton"][i] - l1["newton"]) <= 1), (i, lg["newton"][i],
                                                                      l1["newton"])
    small = min(range(nm), key=lambda i: members[i][0].ncells)
    prob, opts, p0, dt = members[small]
    op = O.OracleProblem(prob)
    r, pho, sho, logo = O.advance(op, opts, p0, 0.0, dt, nsteps)
    assert r == 0
    pe = _I(prob, ens.problems[small].get(dps[small]))
    b = _I(prob, pho[nsteps])
    e = np.max(np.abs(pe - b) / np.maximum(np.abs(b), 1e-3))
    print(f"member {small} vs oracle plain: {e:.3e}", flush=True)
    assert e < 1e-8


def test_ensemble_grid_levels_and_graph_replay(gpu, monkeypatch):
    """The ensemble multigrid's three execution forms give the same bits:
    (a) the members' large levels as grid-level phase launches replayed from
    a captured hipGraph (the default), (b) the same launches issued
    directly, (c) every level of every member inside one workgroup per
    member.  The per-point arithmetic and the sweep order are the same in
    all three, so pressures, saturations and iteration counts agree
    exactly; every replay of the graph (dozens per step, with the
    ensemble's other launches and host synchronisations between them) is
    checked through the state it leaves."""
    api = _api()
    from parflow_amd.decks import c4_member
    members = [c4_member(m) for m in range(10)]
    big = [p.ncells for p, _, _, _ in members]
    print("member cells", big, flush=True)
    assert max(big) > 4 * 4096          # some members have grid levels
    nsteps = 2

    def run(env):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ens = api.GPUEnsemble([(p, o) for p, o, _, _ in members])
        dps = [ens.problems[i].vec(members[i][2]) for i in range(len(members))]
        dss = [ens.problems[i].vec() for i in range(len(members))]
        lg = ens.advance_device(dps, dss, 0.0, [m[3] for m in members], nsteps)
        out = ([ens.problems[i].get(d) for i, d in enumerate(dps)],
               [ens.problems[i].get(d) for i, d in enumerate(dss)], lg, ens.stats())
        for k in env:
            monkeypatch.delenv(k)
        return out

    pg, sg, lgg, stg = run({"PFHIP_ENS_GRAPH": "1"})
    pd, sd, lgd, std = run({"PFHIP_ENS_GRAPH": "0"})
    pb, sb, lgb, stb = run({"PFHIP_ENS_GRAPH": "0", "PFHIP_ENS_MG_SMALL": "100000000"})
    print("rounds/launches graph", stg, "direct", std, "one workgroup", stb, flush=True)
    for i in range(len(members)):
        assert np.array_equal(pg[i], pd[i]) and np.array_equal(sg[i], sd[i]), i
        assert np.array_equal(pg[i], pb[i]) and np.array_equal(sg[i], sb[i]), i
    for key in ("newton", "gmres"):
        assert np.array_equal(lgg[key], lgd[key]) and np.array_equal(lgg[key], lgb[key])


def _phys(prob, rho, gravity, mu):
    import dataclasses
    return dataclasses.replace(prob, rho=rho, gravity=gravity, mu=mu)


@pytest.mark.parametrize("overland,dirs", [(False, True), (True, False)])
def test_density_gravity_viscosity_kernels(gpu, overland, dirs):
    """rho, g, mu != 1 (PL/problem_phase_density.c:119, PL/problem.c:152,
    the phase viscosity): closures and their derivatives (taken with respect
    to the head p / (rho g), as the reference does), the residual, the full
    and PFSymmetric Jacobians, the face velocities and the storage against
    the oracle, whose placement of rho, g and mu the CPU identities pin.
    The second shape spans several residual tiles."""
    api = _api()
    shape = (11, 9, 6) if not overland else (70, 13, 5)
    prob = _phys(cases.hetero(*shape, seed=41, overland=overland, dir_sides=dirs), 1.7, 2.3, 0.6)
    print(f"rho g mu kernels overland={overland} dirs={dirs}", flush=True)
    gp = api.GPUProblem(prob)
    op = O.OracleProblem(prob)
    p_old = cases.random_pressure(prob, seed=42)
    s_old = O.saturation(op, p_old)
    p = p_old + np.random.default_rng(43).normal(0, 0.01, p_old.shape)
    for fcn in (0, 1):
        assert _relmax(_I(prob, gp.saturation(p, fcn)), _I(prob, O.saturation(op, p, fcn))) < 1e-13
        assert _relmax(_I(prob, gp.phase_rel_perm(p, fcn)), _I(prob, O.relperm(op, p, fcn))) < 1e-12
    Fg = gp.nl_function_eval(p, p_old, s_old, 0.1, 0.35)
    Fo = O.nl_function_eval(op, p, p_old, s_old, 0.1, 0.35)
    assert _relmax(_I(prob, Fg), _I(prob, Fo), np.max(np.abs(_I(prob, Fo)))) < 1e-12
    for symm in (0, 1):
        Jg = gp.richards_jacobian_eval(p, 0.1, 0.35, symm)
        Jo = O.jacobian_eval(op, p, 0.1, 0.35, symm)
        idx = [0, 2, 4, 6] if symm else range(7)
        sc = max(np.max(np.abs(Jo[q])) for q in idx)
        for q in idx:
            assert _relmax(_I(prob, Jg[q]), _I(prob, Jo[q]), sc) < 1e-12, (symm, q)
    got = gp.face_velocity(p, p_old, 0.1, 0.35)
    ref = O.face_velocity(op, p, p_old, 0.1, 0.35)
    scale = max(np.max(np.abs(r)) for r in ref)
    for a, b in zip(got, ref):
        assert np.max(np.abs(a - b)) < 1e-12 * scale
    a, b = gp.water_balance_storage(p, s_old), O.storage(op, p, s_old)
    assert abs(a[0] - b[0]) <= 1e-13 * abs(b[0]) and abs(a[1] - b[1]) <= 1e-13 * max(abs(b[1]), 1)


@pytest.mark.parametrize("use_jac,rgm", [(False, (1.7, 2.3, 0.6)), (True, (2.0, 0.5, 0.6))])
def test_density_gravity_viscosity_newton(gpu, use_jac, rgm):
    """The plain Newton solve with rho, g, mu != 1 against the oracle per
    cell per step, Newton +-1.  With the analytic J.v (UseJacobian True)
    rho g is kept at 1 (rho and g separately != 1): the reference's closure
    derivatives are taken with respect to the head p / (rho g), so with
    rho g != 1 its analytic Jacobian is off by that factor and the oracle's
    Newton iteration (like the reference's) does not converge."""
    prob = _phys(cases.hetero(14, 12, 10, seed=44, dir_sides=True), *rgm)
    opts = cases.solver(use_jacobian=use_jac, residual_tol=1e-8, krylov_dim=25,
                        precond="PFMG")
    p0 = prob.hydrostatic(0.5, 0.0)
    print(f"advance rho g mu use_jac={use_jac}", flush=True)
    _compare_runs(prob, opts, p0, 0.2, 4)


def test_time_loop_halving_matches_oracle(gpu):
    """The dt-halving branch on the device (PL/solver_richards.c:3001-3027):
    with Solver.Nonlinear.MaxIter 8 the first Growth step of 1.0 fails
    (it needs 10 Newton iterations), the state is restored and the step
    retried at 0.5, and growth continues from the halved step.  The device
    loop takes the oracle's steps, failures and sizes exactly, its
    pressures within 1e-9."""
    from parflow_amd.problem import TimeControl
    api = _api()
    prob = cases.hetero(14, 12, 10, seed=21, dir_sides=True)
    p0 = prob.hydrostatic(0.5, 0.0)
    opts = cases.solver(residual_tol=1e-8, krylov_dim=25, precond="PFMG", max_iter=8)
    tc = TimeControl(start_time=0.0, stop_time=4.0, base_unit=1.0, dump_interval=-1,
                     ts_type="Growth", initial_step=1.0, growth_factor=1.5, min_step=1.0 / 16,
                     max_step=4.0)
    print("time loop halving", flush=True)
    gp = api.GPUProblem(prob)
    ph, sh, rows = api.GPUSolver(gp, opts).run(p0, tc)
    pho, sho, rowso = O.advance_ts(O.OracleProblem(prob), opts, p0, tc)
    print([(r["dt"], r["newton"], r["conv_failures"]) for r in rows], flush=True)
    assert rowso[0][7] == 1 and rowso[0][1] == 0.5
    assert len(rows) == len(rowso)
    for k, (r, ro) in enumerate(zip(rows, rowso)):
        assert (r["time"], r["dt"], r["dt_info"], r["dumped"]) == ro[:4], k
        assert r["conv_failures"] == ro[7] and abs(r["newton"] - ro[4]) <= 1, k
        a, b = _I(prob, ph[k + 1]), _I(prob, pho[k + 1])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-9, k


def test_fgmres_without_preconditioner_matches_oracle(gpu):
    """The flexible GMRES variant with NoPC (its search directions are the
    basis vectors themselves) against the oracle's plain NoPC solve."""
    api = _api()
    prob = cases.hetero(12, 10, 8, seed=45, dir_sides=True)
    opts = cases.solver(residual_tol=1e-12, krylov_dim=40, precond="NoPC")
    p0 = prob.hydrostatic(0.5, 0.0)
    print("FGMRES NoPC", flush=True)
    gp = api.GPUProblem(prob)
    pg, sg, lg = api.GPUSolver(gp, opts, 64, 1).advance(p0, 0.0, 0.1, 3)
    r, pho, sho, lo = O.advance(O.OracleProblem(prob), opts, p0, 0.0, 0.1, 3)
    assert r == 0
    for s in range(4):
        a, b = _I(prob, pg[s]), _I(prob, pho[s])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-8, s
    assert np.all(np.abs(lg["newton"] - np.array(lo.newton[:3])) <= 1)


@pytest.mark.parametrize("precision", [64, 32])
def test_fgmres_with_restarts_matches_oracle(gpu, precision):
    """Solver.Linear.MaxRestarts > 0 with a Krylov dimension small enough
    that GMRES restarts: the flexible variant (FP64 or FP32 V-cycle) against
    the oracle's plain solve with the same restarts."""
    api = _api()
    prob = cases.hetero(16, 14, 12, seed=46, dir_sides=True)
    opts = cases.solver(residual_tol=1e-12, krylov_dim=1, max_restarts=8, precond="PFMG")
    p0 = prob.hydrostatic(0.5, 0.0)
    print(f"FGMRES restarts precision={precision}", flush=True)
    gp = api.GPUProblem(prob)
    pg, sg, lg = api.GPUSolver(gp, opts, precision, 1).advance(p0, 0.0, 0.2, 3)
    r, pho, sho, lo = O.advance(O.OracleProblem(prob), opts, p0, 0.0, 0.2, 3)
    assert r == 0
    # more linear iterations than Newton iterations: with a Krylov dimension
    # of 1 every further iteration is a restart cycle
    assert np.max(lg["gmres"] / lg["newton"]) > opts.krylov_dim
    for s in range(4):
        a, b = _I(prob, pg[s]), _I(prob, pho[s])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-8, s


def test_mixed_precision_matches_oracle(gpu):
    """The mixed form (flexible GMRES, FP32 V-cycle, CGS2, Arnoldi ||J p||)
    against the oracle's plain FP64 solve directly: per cell within 1e-8 at
    a tight tolerance on a heterogeneous deck with Dirichlet sides, Newton
    +-1."""
    api = _api()
    prob = cases.hetero(16, 14, 12, seed=31, dir_sides=True)
    opts = cases.solver(residual_tol=1e-12, krylov_dim=30, precond="PFMG",
                        num_pre_relax=2, num_post_relax=2)
    p0 = prob.hydrostatic(0.5, 0.0)
    print("mixed vs oracle", flush=True)
    gp = api.GPUProblem(prob)
    pm, sm, lm = api.GPUSolver(gp, opts, 32, 1).advance(p0, 0.0, 0.2, 3)
    r, pho, sho, lo = O.advance(O.OracleProblem(prob), opts, p0, 0.0, 0.2, 3)
    assert r == 0
    for s in range(4):
        a, b = _I(prob, pm[s]), _I(prob, pho[s])
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)) < 1e-8, s
        assert _relmax(_I(prob, sm[s]), _I(prob, sho[s])) < 1e-8, s
    assert np.all(np.abs(lm["newton"] - np.array(lo.newton[:3])) <= 1)
