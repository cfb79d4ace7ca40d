"""pfhip_water_balance_storage_noalloc: the storage sums with the reduction
scratch in the library's static per-GPU array (no allocation in the call)
against pfhip_water_balance_storage (pool scratch) on the same inputs.

Both run the same kernels with the same block count, so the sums agree
bit for bit.  Two calls in a row, the second on a smaller problem, so the
second call reads only what it wrote itself (the array keeps the last
call's contents); a third call repeats the first problem after the smaller
one.
"""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _noalloc(api, gp, dp, ds):
    a, b = C.c_double(), C.c_double()
    L = api.lib()
    L.pfhip_water_balance_storage_noalloc.argtypes = [C.c_void_p] * 3 + \
        [C.POINTER(C.c_double)] * 2 + [C.c_void_p]
    api.check(L.pfhip_water_balance_storage_noalloc(gp.h, dp.ptr, ds.ptr, C.byref(a),
                                                    C.byref(b), None), "storage_noalloc")
    return a.value, b.value


def _case(api, nx, ny, nz, seed):
    prob = cases.hetero(nx, ny, nz, seed=seed)
    gp = api.GPUProblem(prob)
    p = cases.random_pressure(prob, seed=seed + 1)
    # pond part of the top layer, so the surface sum is not zero
    p[prob.nz][1:-1, 1:-1] = np.random.default_rng(seed + 2).uniform(
        -0.2, 0.3, (prob.ny, prob.nx))
    dp = gp.vec(p)
    ds = gp.vec()
    api.check(api.lib().pfhip_saturation(gp.h, dp.ptr, ds.ptr, 0, None), "saturation")
    return gp, dp, ds


def test_storage_noalloc_matches_pool_scratch(gpu):
    from parflow_amd import api
    # 96*70 = 6720 (j, k) rows: the reduction uses its full 1024 blocks;
    # then a problem of 9*5 = 45 rows (45 blocks)
    big = _case(api, 61, 96, 70, 31)
    small = _case(api, 13, 9, 5, 41)
    for gp, dp, ds in (big, small, big):
        ref = gp.water_balance_storage(dp, ds)
        got = _noalloc(api, gp, dp, ds)
        print(f"{gp.prob.nx}x{gp.prob.ny}x{gp.prob.nz}: pool {ref} no-alloc {got}", flush=True)
        assert got == ref
        assert got[1] > 0.0
