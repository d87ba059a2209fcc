"""The Fortran drop-in shim (reference signatures) against the reference."""
import os

import numpy as np
import pytest

import oracle
from helpers import model, rel, small_superblock
from grand_sql_tj_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
needs = pytest.mark.skipif(not (oracle.ref_available() and os.path.exists(oracle.dropin_path())),
                           reason="reference build / drop-in shim not built")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@needs
@pytest.mark.parametrize("geom", [(7, 9), (11, 11)])
def test_dropin_model_wave_matches_reference(geom):
    m = model()
    sb = small_superblock(m, *geom, seed=50 + geom[0])
    x = synth.random_vector(sb, 3)
    assert rel(oracle.dropin_model_wave(m, sb, x), oracle.ref_model_wave(m, sb, x)) < 1e-13


@needs
def test_dropin_lanczos_matches_reference(tmp_path):
    m = model()
    sb = small_superblock(m, 11, 11, seed=4, msys=30, menv=30, median=4)
    x0 = synth.random_vector(sb, 1)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        e_ref, psi_ref = oracle.ref_lanczos(m, sb, x0, 1e-10)
        e_gpu, psi_gpu = oracle.dropin_lanczos(m, sb, x0, 1e-10)
    finally:
        os.chdir(cwd)
    assert abs(e_gpu - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(abs(float(np.dot(psi_gpu, psi_ref))) - 1.0) < 1e-8
