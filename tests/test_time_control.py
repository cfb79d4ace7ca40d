"""Time-step control: SelectTimeStep and the AdvanceRichards time loop.

The oracle's restatement (oracle/oracle.py) is pinned by the time tables of
the reference's own runs (test/correct_output/LW_var_dz.out.log: a dump
interval of two steps and the final 'f' step; var.dz.1d.out.log: dumps
every 100 iterations counted from iteration 0) and by hand-derived growth
sequences.  The library's host functions (pfhip_select_time_step,
pfhip_next_transition) are checked against the same sequences; the full
device loop is checked in test_gpu_parity.py."""
import os

import numpy as np
import pytest

from oracle import oracle as O
import ref_decks
from parflow_amd.problem import BC_FLUX, BCSpec, TimeControl

HERE = os.path.dirname(os.path.abspath(__file__))


def _dry(dt, t, retry):
    return True, (0, 0, 0)


def _log(name):
    g = np.load(os.path.join(HERE, "golden", "time_logs.npz"))
    return (g[name + "_seq"], g[name + "_time"], g[name + "_dt"],
            [chr(c) for c in g[name + "_info"]], g[name + "_dump"])


def lw_var_dz_timing():
    """test/python/LW_var_dz.py:174-241: BaseUnit 10, StopTime 200,
    DumpInterval 20, constant step 10; cycles "constant" (one interval)
    and "rainrec" (10 + 20 units, repeated 14 times)."""
    tc = TimeControl(start_time=0.0, stop_time=200.0, base_unit=10.0, dump_interval=20.0,
                     ts_type="Constant", value=10.0)
    return tc, [((10000000,), -1), ((10, 20), 14)]


def var_dz_1d_timing():
    """test/python/var_dz_1D.py:175-230, 326: StopTime 50, DumpInterval
    -100, constant step 0.01, Solver.MaxIter 2500; one constant cycle."""
    tc = TimeControl(start_time=0.0, stop_time=50.0, base_unit=1.0, dump_interval=-100.0,
                     max_iterations=2500, ts_type="Constant", value=0.01)
    return tc, [((1,), -1)]


def _check_against_log(rows, log, nrows):
    seq, time, dt, info, dump = log
    assert info[0] == "i" and dump[0] == 0       # the initial dump, file 0
    file_number = 0
    for k in range(1, nrows):
        t, d, inf, dumped = rows[k - 1][:4]
        assert abs(t - time[k]) <= 5e-7 * abs(time[k]), k   # printed as %e
        assert abs(d - dt[k]) <= 5e-7 * abs(dt[k]), k
        assert inf == info[k], k
        if dumped:
            file_number += 1
        assert (file_number if dumped else -1) == dump[k], k


def test_oracle_time_loop_reproduces_lw_var_dz_log():
    tc, cycles = lw_var_dz_timing()
    rows = O.time_loop(tc, cycles, _dry)
    log = _log("LW_var_dz")
    assert len(rows) == len(log[0]) - 1 == 20
    _check_against_log(rows, log, len(log[0]))


def test_oracle_time_loop_reproduces_var_dz_1d_log():
    tc, cycles = var_dz_1d_timing()
    log = _log("var_dz_1d")
    rows = O.time_loop(tc, cycles, _dry, max_steps=len(log[0]) - 1)
    _check_against_log(rows, log, len(log[0]))
    # Solver.MaxIter ends the run at iteration 2500, before StopTime 50
    rows = O.time_loop(tc, cycles, _dry)
    assert len(rows) == 2500 and abs(rows[-1][0] - 25.0) < 1e-9


# Growth: InitialStep 0.01, GrowthFactor 1.5, MaxStep 0.05, with the
# reference TiltedV_OverlandKin rain cycle (2 + 300 base units of 0.05): the
# rain ends at t = 0.1, so the step that would cross it is cut to end there
# ('b'), and growth resumes from the cut step.  Worked by hand:
#   0.01, 0.015, 0.0225, 0.03375 (t = 0.08125), cut to 0.01875 (t = 0.1),
#   0.028125, 0.0421875, then 0.0632... clipped to MaxStep 0.05.
GROWTH_EXPECTED = [(0.01, "p"), (0.015, "p"), (0.0225, "p"), (0.03375, "p"),
                   (0.01875, "b"), (0.028125, "p"), (0.0421875, "p"), (0.05, "p"),
                   (0.05, "p")]


def _growth_tc():
    return TimeControl(start_time=0.0, stop_time=2.0, base_unit=0.05, ts_type="Growth",
                       initial_step=0.01, growth_factor=1.5, min_step=0.005, max_step=0.05)


def test_oracle_growth_sequence():
    prob = ref_decks.tiltedv_kin()[0]
    tc = _growth_tc()
    rows = O.time_loop(tc, O.bc_cycles(prob), _dry, max_steps=len(GROWTH_EXPECTED))
    for (t, d, info, *_), (dw, iw) in zip(rows, GROWTH_EXPECTED):
        assert abs(d - dw) < 1e-15 and info == iw
    assert abs(rows[4][0] - 0.1) < 1e-15


def test_oracle_growth_min_step_and_stop_time():
    """A shrinking growth factor settles on MinStep; the last step is cut
    to land on StopTime ('f')."""
    tc = TimeControl(start_time=0.0, stop_time=0.0312, base_unit=1.0, ts_type="Growth",
                     initial_step=0.01, growth_factor=0.5, min_step=0.004, max_step=1.0)
    rows = O.time_loop(tc, [((1,), -1)], _dry)
    dts = [r[1] for r in rows]
    assert dts[:4] == [0.01, 0.005, 0.004, 0.004]
    assert rows[-1][2] == "f" and abs(rows[-1][0] - 0.0312) < 1e-15
    assert abs(sum(dts) - 0.0312) < 1e-15


def test_oracle_halves_failed_steps():
    """A failed solve retries the step at half the size; growth continues
    from the halved step (PL/solver_richards.c:2998-3020)."""
    tc = _growth_tc()
    tc.stop_time = 0.2
    calls = []

    def solve(dt, t, retry):
        calls.append((dt, t, retry))
        return not (len(calls) == 2), (1, 1, 1)     # the second step fails once
    rows = O.time_loop(tc, [((1,), -1)], solve)
    assert [c[2] for c in calls[:4]] == [False, False, True, False]
    assert calls[1][0] == 0.015 and calls[2][0] == 0.0075
    assert rows[1][1] == 0.0075 and rows[1][7] == 1
    assert rows[2][1] == 0.0075 * 1.5


def test_library_select_time_step_matches():
    """pfhip_select_time_step / pfhip_next_transition (host functions of the
    library) along the growth sequence and against the oracle on a sweep of
    times through several cycles."""
    from parflow_amd import api
    prob = ref_decks.tiltedv_kin()[0]
    tc = _growth_tc()
    t, d = 0.0, 0.0
    for dw, iw in GROWTH_EXPECTED:
        d, info = api.select_time_step(prob, tc, t, d)
        assert abs(d - dw) < 1e-15 and info == iw
        t += d
    # cycles with several intervals and repeat counts on different faces
    bc = [BCSpec(BC_FLUX, (0.0, 1.0, 2.0), (3, 1, 2), 4),
          BCSpec(BC_FLUX, (0.0, 1.0), (5, 7), -1)] + [BCSpec(BC_FLUX, (0.0,))] * 4
    import dataclasses
    prob2 = dataclasses.replace(prob, bc=bc, nx=3, ny=3, nz=2, kx=1.0, poro=0.3,
                                slope_x=0.0, slope_y=0.0, mann=1.0)
    tc2 = TimeControl(start_time=0.25, stop_time=40.0, base_unit=0.5, ts_type="Constant",
                      value=0.3)
    cyc = O.bc_cycles(prob2)
    for t in np.linspace(0.25, 41.0, 997):
        a = api.next_transition(prob2, tc2, t)
        b = O.next_transition(cyc, tc2.base_unit, tc2.start_time, tc2.stop_time, t)
        assert a == b, t
        for dprev in (0.0, 0.3):
            assert api.select_time_step(prob2, tc2, t, dprev) == \
                O.select_time_step(tc2, cyc, t, dprev)


def test_next_transition_hand_worked():
    """TimeCycleDataComputeNextTransition worked by hand from
    PL/time_cycle_data.c:239-298 for one cycle of 2 + 3 base units (start 0,
    base 1, forever).  The whole base units elapsed are counted minus one
    (:250), so at a transition instant the interval just ended is still
    'current': its end (now) is no step away and reports -1, and at the
    end of a whole cycle the next transition found is the cycle after."""
    cyc = [((2, 3), -1)]
    want = {0.0: 2.0, 1.0: 1.0, 1.5: 0.5, 2.0: -1.0, 2.5: -1.0, 3.0: 2.0, 4.9: 0.1,
            5.0: 5.0, 6.0: 1.0, 7.0: -1.0, 8.0: 2.0}
    for t, w in want.items():
        got = O.next_transition(cyc, 1.0, 0.0, 100.0, t)
        assert abs(got - w) < 1e-12, (t, got, w)
    # a cycle with a repeat count defines transitions only before its end
    # (repeat * length * base = 10): after t = 8 none is left
    assert O.next_transition([((2, 3), 2)], 1.0, 0.0, 100.0, 8.0) == -1.0
    assert O.next_transition([((2, 3), 2)], 1.0, 0.0, 100.0, 6.0) == 1.0
    # single-interval cycles never bind
    assert O.next_transition([((7,), -1)], 1.0, 0.0, 100.0, 3.0) == -1.0


def _column_problem():
    from parflow_amd.problem import Problem
    bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(6)]
    return Problem(nx=2, ny=1, nz=2, dx=2.0, dy=1.0, dz=0.5, kx=1.0, poro=0.4, ss=1e-4,
                   alpha=1.0, nvg=2.0, sres=0.2, ssat=1.0, bc=bc, base_unit=1.0)


def test_surface_predictor_hand_worked():
    """PL/solver_richards.c:3296-3345 worked by hand on a 2 x 1 x 2 box
    (dx 2, dy 1, dz 0.5, porosity 0.4, dt 1):
      top cell 0: S 0.9 -> free pore volume 0.4 - 0.36 = 0.04; source 0.05
        gives 0.05, the lower face (vz 0.01) 0.02, the right face (vx 0.01,
        outflow) -0.005: 0.065 > 0.04, so p = 0.025 / (2 * 1 * 0.4) = 0.03125
        (or PressureValue when > 0);
      top cell 1: S 0.5 -> free 0.2 > 0.05 + 0.005: unchanged;
      a saturated or ponded cell (p >= 0) is never changed."""
    prob = _column_problem()
    src = np.zeros(prob.shape)
    src[2, 1, 1:3] = 0.05
    op = O.OracleProblem(prob, src=src)
    p = np.full(prob.shape, -0.5)
    s = np.zeros(prob.shape)
    s[2, 1, 1], s[2, 1, 2] = 0.9, 0.5
    vx = np.zeros((2, 1, 3))
    vx[1, 0, 1] = 0.01
    vy = np.zeros((2, 2, 2))
    vz = np.zeros((3, 1, 2))
    vz[1, 0, 0] = 0.01
    out = O.surface_predictor(op, p, s, (vx, vy, vz), None, 1.0, 0.0, False)
    assert abs(out[2, 1, 1] - 0.03125) < 1e-15
    assert out[2, 1, 2] == -0.5
    assert np.array_equal(np.delete(out.ravel(), np.ravel_multi_index((2, 1, 1), out.shape)),
                          np.delete(p.ravel(), np.ravel_multi_index((2, 1, 1), p.shape)))
    assert O.surface_predictor(op, p, s, (vx, vy, vz), None, 1.0, 1e-5, False)[2, 1, 1] == 1e-5
    p2 = p.copy()
    p2[2, 1, 1] = 0.0
    assert O.surface_predictor(op, p2, s, (vx, vy, vz), None, 1.0, 0.0, False)[2, 1, 1] == 0.0
    # lateral term, as :3312-3321 write it: dt ((qx_i - qx_(i-1)) dy +
    # (qy_j - qy_(j-1)) dx) is ADDED to the inflow, i.e. a cell's own east
    # flux ke counts toward its total.  ke = 0.3 in cell 0, 0.1 in cell 1:
    # cell 0 gets 0.065 + 0.3 = 0.365 > 0.04 -> p = 0.325 / 0.8;
    # cell 1 gets 0.055 + (0.1 - 0.3) < 0.2 -> unchanged
    qx = np.zeros((3, 4))
    qx[1, 1], qx[1, 2] = 0.3, 0.1
    qy = np.zeros((3, 4))
    out = O.surface_predictor(op, p, s, (vx, vy, vz), (qx, qy), 1.0, 0.0, True)
    assert abs(out[2, 1, 1] - 0.325 / 0.8) < 1e-15 and out[2, 1, 2] == -0.5
    # the west flux of the first column is not counted (i > 0 test, :3316)
    qx2 = np.zeros((3, 4))
    qx2[1, 0] = 5.0
    out = O.surface_predictor(op, p, s, (vx, vy, vz), (qx2, qy), 1.0, 0.0, True)
    assert abs(out[2, 1, 1] - 0.03125) < 1e-15
