"""The reference's own regression decks restated as parflow_amd Problems.

default_richards  -- test/python/default_richards.py
tiltedv_kin       -- test/python/overland_tiltedV_KWE.py, run "TiltedV_OverlandKin"
"""
import numpy as np

from parflow_amd.problem import (BC_DIR_EQUIL, BC_FLUX, BC_OVERLAND, BCSpec,
                                 Problem, SolverOptions)


def default_richards():
    """test/python/default_richards.py:17-260."""
    nx, ny, nz = 18, 15, 8
    bc = [BCSpec(BC_DIR_EQUIL, (5.0,), (1,), -1, z_ref=1.0),   # left
          BCSpec(BC_DIR_EQUIL, (3.0,), (1,), -1, z_ref=1.0),   # right
          BCSpec(BC_FLUX, (0.0,)), BCSpec(BC_FLUX, (0.0,)),   # front/back
          BCSpec(BC_FLUX, (0.0,)), BCSpec(BC_FLUX, (0.0,))]   # bottom/top
    prob = Problem(nx=nx, ny=ny, nz=nz, dx=8.8888888888888893,
                   dy=10.666666666666666, dz=1.0, x0=-10.0, y0=10.0, z0=1.0,
                   kx=4.0, poro=1.0, ss=1.0e-4, alpha=0.005, nvg=2.0,
                   sres=0.2, ssat=0.99, bc=bc, base_unit=1.0,
                   name="default_richards")
    opts = SolverOptions(residual_tol=1e-9, max_iter=10, eta_choice=0,
                         eta_value=1e-5, use_jacobian=True, deriv_eps=1e-2,
                         krylov_dim=10, precond="PFMG")
    p0 = prob.hydrostatic(3.0, 1.0)
    return prob, opts, p0, 0.001, 50


def tiltedv_kin():
    """test/python/overland_tiltedV_KWE.py:14-180 with the
    TiltedV_OverlandKin settings of :262-282."""
    nx, ny, nz = 5, 5, 1
    dx = dy = 10.0
    xc = (np.arange(nx) + 0.5) * dx
    sx = np.where(xc < 20.0, -0.01, 0.01)          # left -0.01, right/channel 0.01
    slope_x = np.tile(sx, (ny, 1))
    slope_y = np.full((ny, nx), 0.01)
    bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(5)]
    bc.append(BCSpec(BC_OVERLAND, (-0.01, 0.0), (2, 300), -1))
    prob = Problem(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=0.05,
                   kx=1e-7, poro=0.01, ss=1.0e-4, alpha=6.0, nvg=2.0,
                   sres=0.2, ssat=1.0, slope_x=slope_x, slope_y=slope_y,
                   mann=3.0e-6, ov_eps=1e-5, bc=bc, base_unit=0.05,
                   name="TiltedV_OverlandKin")
    opts = SolverOptions(residual_tol=1e-9, step_tol=1e-20, max_iter=100,
                         eta_choice=0, eta_value=0.01, use_jacobian=False,
                         deriv_eps=1e-15, krylov_dim=50, precond="PFMG")
    p0 = prob.hydrostatic(-3.0, 0.05)
    return prob, opts, p0, 0.05, 40
