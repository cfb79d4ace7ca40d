"""Extracts golden vectors from the reference's own regression outputs
(test/correct_output/*.pfb, data files written by the reference CI) into
small .npz fixtures.  Run in the development container only, where the
reference tree is mounted:

    python tests/golden/make_golden.py /root/reference

Fixtures written:
  default_richards.npz     -- test/python/default_richards.py (18x15x8,
                              DirEquilRefPatch + FluxConst, 50 steps):
                              pressure and saturation at selected steps
  default_richards_vel.npz -- the same run's face velocities (velx/vely/velz,
                              (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1)) at steps 1,
                              10, 50
  time_logs.npz            -- the time-step tables of LW_var_dz.out.log and
                              var.dz.1d.out.log (sequence, time, dt, dt_info,
                              dump file number)
  tiltedv_overlandkin.npz  -- test/python/overland_tiltedV_KWE.py run
                              "TiltedV_OverlandKin" (5x5x1, OverlandKinematic
                              rain/recession cycle): all 21 dumps (one every 2 steps)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from parflow_amd.pfb import read_pfb  # noqa: E402


def main(ref):
    co = os.path.join(ref, "test", "correct_output")
    steps = [0, 1, 2, 3, 5, 10, 20, 30, 40, 50]
    press, satur = [], []
    for s in steps:
        _, p = read_pfb(os.path.join(co, f"default_richards.out.press.{s:05d}.pfb"))
        _, q = read_pfb(os.path.join(co, f"default_richards.out.satur.{s:05d}.pfb"))
        press.append(p)
        satur.append(q)
    np.savez_compressed(os.path.join(HERE, "default_richards.npz"),
                        steps=np.array(steps), press=np.array(press),
                        satur=np.array(satur))
    # face velocities at steps whose pressure default_richards.npz holds
    vsteps = [1, 10, 50]
    vel = {c: [] for c in "xyz"}
    for s in vsteps:
        for c in "xyz":
            _, v = read_pfb(os.path.join(co, f"default_richards.out.vel{c}.{s:05d}.pfb"))
            vel[c].append(v)
    np.savez_compressed(os.path.join(HERE, "default_richards_vel.npz"),
                        steps=np.array(vsteps), velx=np.array(vel["x"]),
                        vely=np.array(vel["y"]), velz=np.array(vel["z"]))
    d = co
    press, satur = [], []
    for s in range(21):
        _, p = read_pfb(os.path.join(d, f"TiltedV_OverlandKin.out.press.{s:05d}.pfb"))
        _, q = read_pfb(os.path.join(d, f"TiltedV_OverlandKin.out.satur.{s:05d}.pfb"))
        press.append(p)
        satur.append(q)
    np.savez_compressed(os.path.join(HERE, "tiltedv_overlandkin.npz"),
                        dumps=np.arange(21), press=np.array(press),
                        satur=np.array(satur))


def parse_time_log(path, max_rows=None):
    """The "Sequence # / Time / Delta t / Dumpfile #" table of a reference
    run's <run>.out.log (written by PL/solver_richards.c:4985-5030): rows
    (seq, time, dt, dt_info, dumpfile or -1), the initial row included."""
    rows = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 5 or not parts[0].isdigit() or len(parts[0]) != 6:
                continue
            if parts[3] not in ("i", "p", "b", "f", "w"):
                continue
            dump = int(parts[4]) if len(parts) == 6 else -1
            rows.append((int(parts[0]), float(parts[1]), float(parts[2]), parts[3], dump))
            if max_rows and len(rows) >= max_rows:
                break
    return rows


def time_logs(ref):
    """LW_var_dz (TimingInfo.DumpInterval 20 = two steps, StopTime 200) and
    var.dz.1d (DumpInterval -100, Solver.MaxIter 2500; first 301 rows)."""
    co = os.path.join(ref, "test", "correct_output")
    out = {}
    for name, cap in (("LW_var_dz", None), ("var.dz.1d", 301)):
        rows = parse_time_log(os.path.join(co, f"{name}.out.log"), cap)
        key = name.replace(".", "_")
        out[key + "_seq"] = np.array([r[0] for r in rows])
        out[key + "_time"] = np.array([r[1] for r in rows])
        out[key + "_dt"] = np.array([r[2] for r in rows])
        out[key + "_info"] = np.array([ord(r[3]) for r in rows], dtype=np.int8)
        out[key + "_dump"] = np.array([r[4] for r in rows])
    np.savez_compressed(os.path.join(HERE, "time_logs.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    time_logs(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
