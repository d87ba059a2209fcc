"""The Fortran drop-in shim (reference signatures) against the reference:
the Lanczos step (lanczos_memo_new, model_wave) and the step after it
(Get_density_matrix_*_new, Get_truncation_operator, Get_Hamiltonian_Sys /
_Env, wave_trans_LR / _RL), each called with the reference's derived types
and compared with the reference routine on the same inputs."""
import os

import numpy as np
import pytest

import oracle
from helpers import GEOMS, model, random_trun, rel, small_superblock, wave_args, wave_transform_case
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


@needs
def test_dropin_lanczos_writes_the_reference_energy_line(tmp_path):
    # Energy_All.dat (eigen.f90:166-171): Sys_len, E0, iterations, time
    m = model()
    sb = small_superblock(m, 11, 11, seed=4, msys=30, menv=30, median=4)
    x0 = synth.random_vector(sb, 1)
    cwd = os.getcwd()
    lines = {}
    for who in ("ref", "gpu"):
        (tmp_path / who).mkdir()
        os.chdir(tmp_path / who)
        try:
            (oracle.ref_lanczos if who == "ref" else oracle.dropin_lanczos)(m, sb, x0, 1e-10)
        finally:
            os.chdir(cwd)
        lines[who] = (tmp_path / who / "Energy_All.dat").read_text().splitlines()
    assert len(lines["gpu"]) == 1 and len(lines["ref"]) == 1
    g, r = lines["gpu"][0], lines["ref"][0]
    # format (A8,I3,2X,A3,2X,F18.12,2X,A5,I4,2X,A5,2X,F12.6)
    assert len(g) == len(r) and g[:18] == r[:18]                 # "Sys_len= 12  E0=  "
    assert abs(float(g[18:36]) - float(r[18:36])) <= 1e-10 * abs(float(r[18:36]))
    assert g[38:43] == "Iter=" and int(g[43:47]) > 0 and g[49:54] == "time="


@needs
def test_dropin_plan_cache_compiles_once_per_bond():
    # lanczos_memo_new calls model_wave once per iteration at one bond: the
    # drop-in compiles one plan there, then only uploads the models' data
    import dataclasses
    from grand_sql_tj_amd.blocks import OpSet
    m = model()
    sb = small_superblock(m, 11, 11, seed=61)
    x = synth.random_vector(sb, 3)
    oracle.dropin_release_plan()
    n0 = oracle.dropin_plan_builds()
    ys = [oracle.dropin_model_wave(m, sb, x) for _ in range(3)]
    assert oracle.dropin_plan_builds() == n0 + 1
    assert all(np.array_equal(y, ys[0]) for y in ys)
    assert rel(ys[0], oracle.ref_model_wave(m, sb, x)) < 1e-13
    # the same bond with rebuilt models (new data, same layouts): no compile
    scaled = lambda o, f: OpSet(o.length, o.hdr, o.sub, o.off, o.data * f)
    sb2 = dataclasses.replace(sb, sys=scaled(sb.sys, 0.7), env=scaled(sb.env, 1.3))
    y2 = oracle.dropin_model_wave(m, sb2, x)
    assert oracle.dropin_plan_builds() == n0 + 1
    assert rel(y2, oracle.ref_model_wave(m, sb2, x)) < 1e-13
    # another bond: a new plan replaces the cached one
    sb3 = small_superblock(m, 7, 9, seed=62)
    x3 = synth.random_vector(sb3, 1)
    assert rel(oracle.dropin_model_wave(m, sb3, x3), oracle.ref_model_wave(m, sb3, x3)) < 1e-13
    assert oracle.dropin_plan_builds() == n0 + 2
    oracle.dropin_release_plan()


def _same_layout(a, b):
    assert np.asarray(a.hdr).shape == np.asarray(b.hdr).shape
    assert (np.asarray(a.hdr) == np.asarray(b.hdr)).all()
    assert (np.asarray(a.sub) == np.asarray(b.sub)).all()


@needs
@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("side", [0, 1])
def test_dropin_density_matrix_matches_reference(side, levels):
    m = model()
    sb = small_superblock(m, 7, 9, seed=5)
    psi = np.concatenate([synth.random_vector(sb, 3 + l) for l in range(levels)])
    d = oracle.dropin_density_matrix(side, 8, 10, sb.wsec, psi, levels)
    r = oracle.ref_density_matrix(side, 8, 10, sb.wsec, psi, levels)
    _same_layout(d, r)
    assert rel(d.data, r.data) < 1e-13


@needs
@pytest.mark.parametrize("kmin,kmax,terr,sysblk", [(20, 60, 1e-6, True), (5, 200, 1e-3, False)])
def test_dropin_truncation_matches_reference(kmin, kmax, terr, sysblk, tmp_path):
    m = model()
    sb = small_superblock(m, 11, 11, seed=9, msys=160, menv=150, median=14, nsec=10)
    den = oracle.ref_density_matrix(0, 12, 12, sb.wsec, synth.random_vector(sb, 4), 1)
    (tmp_path / "ref").mkdir()
    (tmp_path / "gpu").mkdir()
    r, ns_r, ent_r = oracle.ref_truncation(den, kmin, kmax, terr, str(tmp_path / "ref"), system_block=sysblk)
    d, ns_d, ent_d = oracle.dropin_truncation(den, kmin, kmax, terr, str(tmp_path / "gpu"), system_block=sysblk)
    assert ns_d == ns_r
    # the reference's three log files (den_mat.f90:299-317), line for line
    for name in ("Entropy.dat", "entropy1.dat", "Num_State.dat"):
        lr = (tmp_path / "ref" / name).read_text().split()
        lg = (tmp_path / "gpu" / name).read_text().split()
        assert len(lg) == len(lr) and lg, name
        for a, b in zip(lg, lr):
            try:
                fa, fb = float(a), float(b)
            except ValueError:
                assert a == b, name
                continue
            assert abs(fa - fb) <= 1e-7 * max(1.0, abs(fb)) + 1e-12, (name, a, b)
    assert abs(ent_d - ent_r) <= 1e-12 * max(1.0, abs(ent_r))
    _same_layout(d, r)
    for s in range(np.asarray(d.sub).shape[0]):
        n, c = int(d.sub[s, 3]), int(d.sub[s, 4])
        t = np.asarray(d.data[d.off[s]:d.off[s] + n * c]).reshape(c, n).T
        tr = np.asarray(r.data[r.off[s]:r.off[s] + n * c]).reshape(c, n).T
        assert np.abs(t.T @ t - np.eye(c)).max() < 1e-12              # orthonormal kept columns
        assert np.abs(t @ t.T - tr @ tr.T).max() < 1e-8                # the reference's kept subspace


# esite = 0: with esite != 0 the reference's own on-site term matches
# sub-blocks by an unset down_dif (block.f90:464-469); the device path's
# on-site term is pinned against the oracle in the update tests
@needs
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("geom", GEOMS[:3])
def test_dropin_get_hamiltonian_matches_reference(geom, side):
    m = model(mu=0.0)
    sl, el = geom
    sb = small_superblock(m, sl, el, seed=31 + sl + el)
    pre, bs = (sb.sys, sb.sys_bs) if side == 0 else (sb.env, sb.env_bs)
    tr = random_trun(bs, 7 + sl)
    d = oracle.dropin_get_hamiltonian(m, side, sb.sys_len, sb.env_len, pre, bs, tr)
    r = oracle.ref_get_hamiltonian(m, side, sb.sys_len, sb.env_len, pre, bs, tr)
    _same_layout(d, r)
    assert rel(d.data, r.data) < 1e-13


@needs
@pytest.mark.parametrize("direction,seed", [(0, 2), (1, 3), (0, 7)])
def test_dropin_wave_trans_matches_reference(direction, seed):
    c = wave_transform_case(direction, seed)
    d = oracle.dropin_wave_trans(*wave_args(c))
    r = oracle.ref_wave_trans(*wave_args(c))
    assert d.shape == r.shape and d.size > 0
    assert rel(d, r) < 1e-13


@needs
@pytest.mark.parametrize("sys_len,env_len", [(5, 5), (6, 8), (9, 7)])
def test_dropin_measure_density_matches_reference(tmp_path, sys_len, env_len):
    """measure_operator_dia through the drop-in (gpu_measure_operator_dia)
    against the reference's own, on synthetic chains with random
    truncations (trun_idx = 1: every truncation applies)."""
    from test_measure_cpu import case
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 3 + sys_len)
    args = (model(up_bias=1), tmp_path, sys_len, env_len, sb, eb, st, et, wsec, x)
    got = oracle.dropin_measure_density(*args)
    want = oracle.ref_measure_density(*args)
    assert np.abs(got - want).max() < 1e-13
    assert abs(got.sum() - N) < 1e-12 * N
    assert os.path.exists(tmp_path / "num_elec.dat")
