"""Seeded synthetic decks for the Richards benchmark workloads.

c0  32x32x32 homogeneous box, hydrostatic IC (water table at z = 8 m),
    no-flux sides, constant head at the bottom (DirEquilRefPatch with zero
    pressure head referenced at the water-table elevation z = 8 m, so the
    bottom face holds the initial hydrostatic pressure), top flux -0.01 m/h,
    dt = 0.1 h, 10 steps, GMRES(25) + multigrid.
c1  256x256x64 (4.2M cells), dx = dy = 10 m, dz = 1 m; ln Ksat a Gaussian
    random field (mean ln 0.1 m/h, variance 1, exponential covariance,
    correlation length 8 cells) made by FFT circulant embedding; per-cell
    van Genuchten alpha lognormal (median 1 1/m, sigma 0.3) and n uniform
    in [1.5, 3]; top flux -0.005 m/h, dt = 0.5 h, 20 steps.
c2  512x512x64 (16.8M cells), the c1 recipe with a new seed; flexible
    GMRES(30) with the FP32 V-cycle (2 pre + 2 post red/black sweeps).
c3  512x512x32, dx = dy = 10 m, dz = 0.5 m, tilted-V catchment with
    kinematic overland flow, rain 0.05 m/h for 1 h then 4 h recession,
    dt = 0.05 h.
c4  an ensemble of independent hillslopes (``c4_members``): nx, ny
    log-uniform in [8, 64], nz log-uniform in [20, 50], dx = dy = 1 m,
    dz = 0.25 m; per member Ksat lognormal (median 0.05 m/h, sigma 1),
    alpha lognormal (median 1.5 1/m, sigma 0.4), n uniform in [1.5, 3],
    surface slope uniform in [0.005, 0.1] (the initial water table is tilted
    along x by it), top flux uniform in [-0.02, 0] m/h; 24 steps of 0.25 h
    with the mixed solver of c2.

Every field is generated from a fixed seed so a CPU run and a GPU run of
the same deck see bit-identical inputs.
"""
from __future__ import annotations

import os

import numpy as np

from .problem import (BC_DIR_EQUIL, BC_FLUX, BC_OVERLAND, BCSpec, Problem,
                      SolverOptions)

#: elevation of the initial water table (m) for every deck
WATER_TABLE = 8.0


def gaussian_random_field(shape, corr_len, variance, seed, pad=None):
    """Stationary Gaussian field with covariance variance*exp(-r/corr_len)
    (r in cells) by circulant embedding on a padded periodic grid."""
    import scipy.fft as sfft
    nz, ny, nx = shape
    if pad is None:
        pad = int(4 * corr_len)
    Nz, Ny, Nx = nz + pad, ny + pad, nx + pad
    z = np.minimum(np.arange(Nz), Nz - np.arange(Nz)).astype(np.float64)
    y = np.minimum(np.arange(Ny), Ny - np.arange(Ny)).astype(np.float64)
    x = np.minimum(np.arange(Nx), Nx - np.arange(Nx)).astype(np.float64)
    r = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
    cov = variance * np.exp(-r / corr_len)
    del r
    workers = min(16, os.cpu_count() or 1)
    lam = sfft.rfftn(cov, workers=workers).real
    del cov
    lam = np.sqrt(np.maximum(lam, 0.0))
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((Nz, Ny, Nx))
    W = sfft.rfftn(w, workers=workers)
    del w
    W *= lam
    f = sfft.irfftn(W, s=(Nz, Ny, Nx), workers=workers)
    return np.ascontiguousarray(f[:nz, :ny, :nx])


def _vg_fields(shape, seed):
    rng = np.random.default_rng(seed + 1)
    alpha = np.exp(rng.normal(np.log(1.0), 0.3, shape))
    n = rng.uniform(1.5, 3.0, shape)
    return alpha, n


def deck(name: str, seed: int = None):
    """Returns (Problem, SolverOptions, p0, dt, nsteps, extras)."""
    name = name.lower()
    if name == "c0":
        nx = ny = nz = 32
        bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(4)]
        bc.append(BCSpec(BC_DIR_EQUIL, (0.0,), (1,), -1, z_ref=WATER_TABLE))
        bc.append(BCSpec(BC_FLUX, (-0.01,)))
        prob = Problem(nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.5, kx=1.0, poro=0.4,
                       ss=1e-4, alpha=2.0, nvg=2.0, sres=0.2, ssat=1.0, bc=bc,
                       name="c0")
        opts = SolverOptions(krylov_dim=25, precond="PFMG")
        p0 = prob.hydrostatic(0.0, WATER_TABLE)
        return prob, opts, p0, 0.1, 10, {}
    if name in ("c1", "c2"):
        if name == "c1":
            nx, ny, nz = 256, 256, 64
            seed = 101 if seed is None else seed
        else:
            nx, ny, nz = 512, 512, 64
            seed = 202 if seed is None else seed
        shape = (nz, ny, nx)
        lnK = np.log(0.1) + gaussian_random_field(shape, 8.0, 1.0, seed)
        K = np.exp(lnK)
        del lnK
        alpha, n = _vg_fields(shape, seed)
        bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(4)]
        bc.append(BCSpec(BC_DIR_EQUIL, (0.0,), (1,), -1, z_ref=WATER_TABLE))
        bc.append(BCSpec(BC_FLUX, (-0.005,)))
        prob = Problem(nx=nx, ny=ny, nz=nz, dx=10.0, dy=10.0, dz=1.0, kx=K, poro=0.4,
                       ss=1e-4, alpha=alpha, nvg=n, sres=0.2, ssat=1.0, bc=bc,
                       name=name)
        if name == "c1":
            opts = SolverOptions(krylov_dim=25, precond="PFMG")
        else:
            opts = SolverOptions(krylov_dim=30, precond="PFMG", num_pre_relax=2,
                                 num_post_relax=2)
        p0 = prob.hydrostatic(0.0, WATER_TABLE)
        return prob, opts, p0, 0.5, 20, {"seed": seed}
    if name == "c3":
        nx, ny, nz = 512, 512, 32
        seed = 303 if seed is None else seed
        rng = np.random.default_rng(seed)
        dx = 10.0
        xc = (np.arange(nx) + 0.5) * dx
        mid = nx * dx / 2
        dist = np.abs(xc - mid) / mid
        mag = 0.01 + 0.04 * dist
        sx = np.where(xc < mid, -mag, mag)
        slope_x = np.tile(sx, (ny, 1)) + rng.normal(0.0, 0.002, (ny, nx))
        slope_y = np.full((ny, nx), 0.005) + rng.normal(0.0, 0.002, (ny, nx))
        bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(4)]
        bc.append(BCSpec(BC_DIR_EQUIL, (0.0,), (1,), -1, z_ref=WATER_TABLE))
        bc.append(BCSpec(BC_OVERLAND, (-0.05, 0.0), (20, 80), -1))
        prob = Problem(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=0.5, kx=0.01, poro=0.4,
                       ss=1e-4, alpha=2.0, nvg=2.0, sres=0.2, ssat=1.0,
                       slope_x=slope_x, slope_y=slope_y, mann=5.52e-6, bc=bc,
                       base_unit=0.05, name="c3")
        opts = SolverOptions(krylov_dim=30, precond="PFMG", num_pre_relax=2,
                             num_post_relax=2)
        p0 = prob.hydrostatic(0.0, WATER_TABLE)
        return prob, opts, p0, 0.05, 100, {"seed": seed}
    raise ValueError(f"unknown deck {name!r}")


def c4_member(m: int, seed: int = 404):
    """Member m of the hillslope ensemble: (Problem, SolverOptions, p0, dt).

    Each member draws from its own generator (seed, m).  The water table
    starts at mid-depth at the centre of the slope and tilts along x with
    the member's surface slope (higher upslope, clipped to 10-90 % of the
    depth); the bottom holds the hydrostatic head of the mid-depth table."""
    rng = np.random.default_rng([seed, m])
    nx = int(round(np.exp(rng.uniform(np.log(8), np.log(64)))))
    ny = int(round(np.exp(rng.uniform(np.log(8), np.log(64)))))
    nz = int(round(np.exp(rng.uniform(np.log(20), np.log(50)))))
    ksat = float(np.exp(rng.normal(np.log(0.05), 1.0)))
    alpha = float(np.exp(rng.normal(np.log(1.5), 0.4)))
    nvg = float(rng.uniform(1.5, 3.0))
    slope = float(rng.uniform(0.005, 0.1))
    qtop = float(rng.uniform(-0.02, 0.0))
    dx = dy = 1.0
    dz = 0.25
    depth = nz * dz
    wt0 = 0.5 * depth
    bc = [BCSpec(BC_FLUX, (0.0,)) for _ in range(4)]
    bc.append(BCSpec(BC_DIR_EQUIL, (0.0,), (1,), -1, z_ref=wt0))
    bc.append(BCSpec(BC_FLUX, (qtop,)))
    prob = Problem(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, kx=ksat, poro=0.4, ss=1e-4,
                   alpha=alpha, nvg=nvg, sres=0.2, ssat=1.0, bc=bc, name=f"c4m{m}")
    opts = SolverOptions(krylov_dim=30, precond="PFMG", num_pre_relax=2, num_post_relax=2)
    xc = (np.arange(nx) + 0.5) * dx
    zwt = np.clip(wt0 + slope * (0.5 * nx * dx - xc), 0.1 * depth, 0.9 * depth)
    zc = prob.z0 + (np.arange(nz) + 0.5) * dz
    p0 = prob.zeros()
    p0[1:-1, 1:-1, 1:-1] = (-prob.rho * prob.gravity *
                            (zc[:, None, None] - zwt[None, None, :]))
    return prob, opts, p0, 0.25


def c4_members(n: int = 512, seed: int = 404):
    """The first n members of the hillslope ensemble."""
    return [c4_member(m, seed) for m in range(n)]
