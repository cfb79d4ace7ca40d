"""CPU tests of the drop-in boundary: libpfhip.so loads and exports every
entry point include/pfhip.h declares (no compute calls without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "parflow_amd", "libpfhip.so")


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "pfhip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(pfhip_[a-z0-9_]+)\s*\(", hdr)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from parflow_amd import build
        build.build()
    return LIB


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("pfhip_nl_function_eval", "pfhip_richards_jacobian_eval",
              "pfhip_saturation", "pfhip_phase_rel_perm", "pfhip_matvec",
              "pfhip_overland_flow_eval_kin", "pfhip_mg_apply",
              "pfhip_kinsol_nonlin_solver", "pfhip_advance_richards"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_with_ctypes(lib):
    from parflow_amd import api
    L = api.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    for s in api.EXPORTS:
        assert s in declared_symbols(), s


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    """The product path fails loudly without the HIP library."""
    from parflow_amd import api
    monkeypatch.setattr(api, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(api, "_lib", None)
    with pytest.raises(api.PFHIPError):
        api.lib()


def test_problem_create_rejects_grids_beyond_the_offset_limit(lib):
    """pfhip_problem_create refuses a ghosted grid of 2^29 cells or more (the
    residual kernels' 32-bit byte offsets would wrap) with an argument error,
    before any device call."""
    import ctypes as C
    from parflow_amd import api
    L = api.lib()
    d = api.ProblemDescC()
    d.dx = d.dy = d.dz = 1.0
    cell = C.c_double(0.0)
    dummy = C.c_void_p(C.addressof(cell))
    for f in ("kx", "ky", "kz", "poro", "ss", "alpha", "nvg"):
        setattr(d, f, dummy)
    for b in d.bc:
        b.n_intervals = 1
        b.lengths[0] = 1
    out = C.c_void_p()
    d.nx, d.ny, d.nz = 1022, 1022, 510          # (nx+2)(ny+2)(nz+2) = 2^29
    rc = L.pfhip_problem_create(C.byref(d), C.byref(out))
    assert rc == -2 and not out.value
    msg = L.pfhip_last_error().decode()
    assert "exceeds the limit of 536870912" in msg, msg

