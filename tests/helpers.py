"""Shared inputs for the tests."""
import os

import numpy as np

from grand_sql_tj_amd.lattice import TJModel
from grand_sql_tj_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# the reference build is compiled for Ny = 6 and a 24-column lattice (common.f90:10)
REF_NY, REF_NX = 6, 24

# geometries (sys_len, env_len): symmetric, asymmetric, and ones where the
# cross-cut bonds are carried by sys sites or by env sites
GEOMS = [(5, 5), (6, 5), (7, 9), (11, 11), (12, 10), (13, 17)]


def model(up_bias=1, t=1.0, tp=-0.2, J=0.4, mu=0.3, ny=REF_NY, nx=REF_NX):
    return TJModel(ny=ny, nx=nx, t=t, tp=tp, J=J, mu=mu, up_bias=up_bias)


def small_superblock(m, sys_len, env_len, seed, msys=24, menv=28, median=3, nsec=8):
    return synth.superblock(m, sys_len, env_len, msys, menv, seed=seed, median=median, nsec=nsec)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
