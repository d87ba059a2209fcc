"""The Fortran drop-in shim on the CPU: what it checks before any GPU call."""
import os
import subprocess
import sys

import pytest

import oracle
from helpers import ROOT

needs = pytest.mark.skipif(not (oracle.ref_available() and os.path.exists(oracle.dropin_path())),
                           reason="reference build / drop-in shim not built")


@needs
def test_dropin_rejects_nonuniform_couplings():
    # the library takes one coupling per bond class (common.f90:36-39 holds
    # them per direction and site): a site with another jt must end the
    # program with a message, not be read as the uniform value
    code = """
import sys
sys.path[:0] = ['tests', 'oracle']
import numpy as np, oracle
from helpers import model, small_superblock
from grand_sql_tj_amd import synth
m = model()
sb = small_superblock(m, 7, 9, seed=3)
oracle._ref_init(m)
oracle.dropin_set_jt(2, 5, 0.123)
y = np.zeros(sb.dim)
oracle.dropin_lib().dropin_model_wave(sb.sys_len, sb.env_len, *oracle._m(sb.sys), *oracle._m(sb.env),
    *oracle._bs(sb.sys_bs), *oracle._bs(sb.env_bs), int(sb.wsec.shape[0]), np.ascontiguousarray(sb.wsec),
    np.ascontiguousarray(synth.random_vector(sb, 1)), y)
print('returned')
"""
    p = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert p.returncode != 0
    assert "returned" not in p.stdout
    assert "not uniform per bond class" in p.stdout + p.stderr
