"""Writes tests/golden/c1_reference.json: the energies of the reference's own
C1 run (4 x 8 cylinder, t = 1, J = 0.4, 28 electrons, S = 0, m = 500, 6
sweeps), from the reference's routines compiled here.

TEST INFRASTRUCTURE ONLY.  The reference fixes the lattice at compile time
(code/tjsq/common.f90:10: Nx, Ny) and the particle label of its site
operators (up_bias): the build directory gets a copy of common.f90 with
Nx = 8, Ny = 4, up_bias = 1 (the canonical ensemble of 28 electrons), the
other sources (block, get_sys, eigen, den_mat, dmrg_sub) are compiled where
they lie, and c1_driver.f90 takes the place of main.f90 (its parameters and
the Get_Ground_State schedule, main.f90:13-187).  Nothing of the build
directory is committed; only the energies are.

    python tests/golden/make_c1.py [build_dir]      (default oracle/_ref/c1)
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/code/tjsq"
FLANG = "/opt/rocm/lib/llvm/bin/amdflang"
MKLDIR = "/opt/conda/lib"
SOURCES = ["block", "get_sys", "eigen", "den_mat", "dmrg_sub"]
PARAMS = dict(ny=4, nx=8, t=1.0, tp=0.0, J=0.4, electrons=28, kept_max=500, kept_maxf=500, sweeps=6,
              lanerr=1e-10, trun_err=1e-18)


def patched_common(text: str) -> str:
    out, n1 = re.subn(r"Nx=\d+,Ny=\d+,", "Nx=%d,Ny=%d," % (PARAMS["nx"], PARAMS["ny"]), text, count=1)
    out, n2 = re.subn(r"up_bias=\d+,", "up_bias=1,", out, count=1)
    if n1 != 1 or n2 != 1:
        raise RuntimeError("common.f90: the Nx / Ny / up_bias parameters were not found")
    return out


def build(bdir: str) -> str:
    os.makedirs(bdir, exist_ok=True)
    with open(os.path.join(REF, "common.f90")) as f:
        text = patched_common(f.read())
    with open(os.path.join(bdir, "common.f90"), "w") as f:
        f.write(text)
    fl = [FLANG, "-O2", "-module-dir", bdir, "-I", bdir]
    objs = []
    for name, src in [("common", os.path.join(bdir, "common.f90"))] + [(s, os.path.join(REF, s + ".f90"))
                                                                       for s in SOURCES] + \
            [("c1_driver", os.path.join(HERE, "c1_driver.f90"))]:
        obj = os.path.join(bdir, name + ".o")
        subprocess.run(fl + ["-c", src, "-o", obj], check=True, capture_output=True)
        objs.append(obj)
    exe = os.path.join(bdir, "c1_run")
    subprocess.run([FLANG, "-o", exe] + objs + ["-L" + MKLDIR, "-Wl,-rpath," + MKLDIR, "-lmkl_rt"], check=True)
    return exe


def main():
    bdir = os.path.abspath(sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "oracle", "_ref", "c1"))
    sys.path.insert(0, ROOT)
    from grand_sql_tj_amd.lattice import TJModel
    m = TJModel(ny=PARAMS["ny"], nx=PARAMS["nx"], t=PARAMS["t"], tp=PARAMS["tp"], J=PARAMS["J"], mu=0.0, up_bias=1)
    exe = build(bdir)
    run = os.path.join(bdir, "run")
    os.makedirs(run, exist_ok=True)
    for name in os.listdir(run):
        os.remove(os.path.join(run, name))
    with open(os.path.join(run, "c1.in"), "w") as f:
        f.write("%r %r %r %r %r %r %r\n" % (m.jt[0], m.jt[1], m.jd[0], m.jd[1], m.jn[0], m.jn[1], m.esite))
        f.write("%r %r\n" % (PARAMS["lanerr"], PARAMS["trun_err"]))
        f.write("%d %d %d %d\n" % (PARAMS["kept_max"], PARAMS["kept_maxf"], PARAMS["sweeps"], PARAMS["electrons"]))
    env = dict(os.environ, MKL_NUM_THREADS="1")
    t0 = time.time()
    with open(os.path.join(run, "stdout.txt"), "w") as out:
        subprocess.run([exe], cwd=run, env=env, stdout=out, stderr=subprocess.STDOUT, check=True)
    seconds = time.time() - t0
    steps = []
    with open(os.path.join(run, "Energy_All.dat")) as f:
        for line in f:
            if line.startswith("Sys_len="):
                # format (A8,I3,2X,A3,2X,F18.12,2X,A5,I4,2X,A5,2X,F12.6), eigen.f90:174
                steps.append({"sys_len": int(line[8:11]), "e0": float(line[18:36]), "iterations": int(line[43:47])})
    fixture = {"config": "C1: t-J 4x8 cylinder (y periodic, x open), t=1, J=0.4, 28 electrons, S=0, kept_max = "
                         "kept_maxf = 500, 6 sweeps, lanerr 1e-10, trun_err 1e-18, Get_Ground_State schedule",
               "params": PARAMS, "reference_seconds_1_core": seconds, "steps": steps,
               "final_e0": steps[-1]["e0"] if steps else None,
               "min_e0": min(s["e0"] for s in steps) if steps else None}
    with open(os.path.join(HERE, "c1_reference.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print("C1: %d solves, final E0 %.12f, %.0f s" % (len(steps), fixture["final_e0"], seconds))


if __name__ == "__main__":
    main()
