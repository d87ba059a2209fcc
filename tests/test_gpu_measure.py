"""Density measurement on the device (grand_sql_tj_amd.measure: the
transport programs tjsq_dia_transport_create and the contraction
tjsq_measure_create) against the numpy oracle of measure_operator_dia
(oracle/measure_oracle.py, pinned against the reference on the CPU) on
synthetic chains with random truncations, and against exact
diagonalisation through the device DMRG: untruncated, <n_i> of the DMRG
ground state is the exact ground state's density, site by site."""
import json
import os

import numpy as np
import pytest

import ed
import measure_oracle as mo
from test_measure_cpu import GEOMS, case
from grand_sql_tj_amd.lattice import TJModel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "ed_energies.json")) as f:
    ED = json.load(f)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


class _Trun:
    def __init__(self, t):
        from grand_sql_tj_amd.blocks import OpSet
        self.ops = OpSet(int(t.length), np.asarray(t.hdr), np.asarray(t.sub), np.asarray(t.off),
                         np.zeros(max(1, np.asarray(t.data).size)))
        self.data = torch.from_numpy(np.ascontiguousarray(t.data, np.float64)).cuda()


@pytest.mark.parametrize("sys_len,env_len", GEOMS)
def test_device_density_matches_oracle(sys_len, env_len):
    from grand_sql_tj_amd import measure
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 3 + sys_len)
    want = mo.density_profile(1, sys_len + env_len, sys_len, env_len, sb, eb, st, et, wsec, x)
    psi = torch.from_numpy(x).cuda()
    got = measure.density_profile(1, sys_len, env_len, sb, eb, {L: _Trun(t) for L, t in st.items()},
                                  {L: _Trun(t) for L, t in et.items()}, wsec, psi)
    assert np.abs(got - want).max() < 1e-13
    assert abs(got.sum() - N) < 1e-12 * N


def _ed_density(model, n_elec):
    import scipy.sparse.linalg as spla
    H, states = ed.hamiltonian(model, n_elec)
    if H.shape[0] <= 2000:
        w, v = np.linalg.eigh(H.toarray())
        g = v[:, 0]
    else:
        w, v = spla.eigsh(H, k=1, which="SA", tol=1e-13)
        g = v[:, 0]
    n = model.nsite
    up, dn = states >> n, states & ((1 << n) - 1)
    occ = np.array([((up >> b) & 1) + ((dn >> b) & 1) for b in range(n)], dtype=np.float64)
    return occ @ (g * g)


@pytest.mark.parametrize("case_", [c for c in ED if c["nx"] * c["ny"] <= 12][:3],
                         ids=lambda c: f"{c['ny']}x{c['nx']}N{c['N']}")
def test_untruncated_dmrg_density_is_exact(case_):
    from grand_sql_tj_amd import sweep
    m = TJModel(ny=case_["ny"], nx=case_["nx"], t=case_["t"], tp=case_["tp"], J=case_["J"], mu=case_["mu"],
                up_bias=1)
    d = sweep.DMRG(m, total_up=case_["N"], kept_max=1000, trun_err=1e-12, tol=1e-12)
    d.run(sweeps=1)
    got = d.measure_density()
    want = _ed_density(m, case_["N"])
    assert abs(got.sum() - case_["N"]) < 1e-10
    # the state converges to the Lanczos tolerance on the energy (1e-12), so to
    # ~1e-6 in norm: densities agree to that order
    assert np.abs(got - want).max() < 2e-5


@pytest.mark.parametrize("sys_len,env_len", GEOMS)
def test_device_cross_correlation_matches_oracle(sys_len, env_len):
    from grand_sql_tj_amd import measure
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 3 + sys_len)
    want = mo.cross_correlation(1, sys_len, env_len, sb, eb, st, et, wsec, x)
    psi = torch.from_numpy(x).cuda()
    got = measure.cross_correlation(1, sys_len, env_len, sb, eb, {L: _Trun(t) for L, t in st.items()},
                                    {L: _Trun(t) for L, t in et.items()}, wsec, psi)
    assert got.shape == (sys_len, env_len)
    assert np.abs(got - want).max() < 1e-13


def _ed_pair_density(model, n_elec):
    H, states = ed.hamiltonian(model, n_elec)
    import scipy.sparse.linalg as spla
    g = spla.eigsh(H, k=1, which="SA", tol=1e-13)[1][:, 0]
    n = model.nsite
    up, dn = states >> n, states & ((1 << n) - 1)
    occ = np.array([((up >> b) & 1) + ((dn >> b) & 1) for b in range(n)], dtype=np.float64)
    return (occ * (g * g)) @ occ.T


def test_untruncated_dmrg_cross_correlation_is_exact():
    from grand_sql_tj_amd import sweep
    c = next(c for c in ED if c["ny"] == 2 and c["nx"] == 6 and c["N"] == 10 and c["mu"] == 0.0)
    m = TJModel(ny=c["ny"], nx=c["nx"], t=c["t"], tp=c["tp"], J=c["J"], mu=c["mu"], up_bias=1)
    d = sweep.DMRG(m, total_up=c["N"], kept_max=1000, trun_err=1e-12, tol=1e-12)
    d.run(sweeps=1)
    got = d.measure_cross_correlation()
    sl, el = d.bond[0] + 1, d.bond[1] + 1
    nn = _ed_pair_density(m, c["N"])
    n = m.nsite
    want = np.array([[nn[i, n - j] for j in range(1, el + 1)] for i in range(sl)])
    assert got.shape == (sl, el)
    assert np.abs(got - want).max() < 2e-5


@pytest.mark.parametrize("case_", [c for c in ED if c["nx"] * c["ny"] <= 12][:3],
                         ids=lambda c: f"{c['ny']}x{c['nx']}N{c['N']}")
def test_untruncated_dmrg_correlation_matrix_is_exact(case_):
    """The whole <n_i n_j> matrix (same-side products formed with their SU(2)
    factor, cross-bond pairs as a Gram matrix) against exact
    diagonalisation, entry by entry."""
    from grand_sql_tj_amd import measure, sweep
    m = TJModel(ny=case_["ny"], nx=case_["nx"], t=case_["t"], tp=case_["tp"], J=case_["J"], mu=case_["mu"],
                up_bias=1)
    d = sweep.DMRG(m, total_up=case_["N"], kept_max=1000, trun_err=1e-12, tol=1e-12)
    d.run(sweeps=1)
    sl, el = d.bond[0] + 1, d.bond[1] + 1
    got = measure.correlation_matrix(1, sl, el, d.sys_basis, d.env_basis, d.sys_trun, d.env_trun, d.wsec, d.psi)
    want = _ed_pair_density(m, case_["N"])
    assert np.abs(got - want).max() < 2e-5
    assert abs(got.sum() - case_["N"] ** 2) < 1e-9 * case_["N"] ** 2


def test_device_pair_vectors_match_tables():
    """side_pair_vectors on the device against the same programs' tables run
    on the CPU, on a truncated synthetic chain."""
    from grand_sql_tj_amd import measure
    from test_measure_cpu import _tables_pairs
    sb, eb, st, et, wsec, x, N = case(6, 8, 9)
    psi = torch.from_numpy(x).cuda()
    for side, bases, truns, length in ((0, sb, st, 6), (1, eb, et, 8)):
        dev_truns = {L: _Trun(t) for L, t in truns.items()}
        Y, labels = measure.side_pair_vectors(side, measure.num_elec(1), bases, dev_truns, length, wsec, psi)
        Yc, lc = _tables_pairs(side, bases, truns, length, wsec, x)
        assert labels == lc
        assert np.abs((Y @ psi).cpu().numpy() - Yc @ x).max() < 1e-13


def test_measure_from_checkpoint_files(tmp_path):
    """Measurement's flow: the density profile and correlation matrix from a
    checkpoint directory (written here by DMRG.save_checkpoint in the
    reference's format) equal the ones measured on the live run."""
    from grand_sql_tj_amd import measure, sweep
    c = next(c for c in ED if c["ny"] == 2 and c["nx"] == 8)
    m = TJModel(ny=c["ny"], nx=c["nx"], t=c["t"], tp=c["tp"], J=c["J"], mu=c["mu"], up_bias=1)
    d = sweep.DMRG(m, total_up=c["N"], kept_max=48, trun_err=1e-12, tol=1e-12)
    d.run(sweeps=1)
    d.save_checkpoint(tmp_path)
    dens, cor = measure.measure_run(tmp_path)
    assert np.abs(dens - d.measure_density()).max() < 1e-13
    assert np.abs(cor - d.measure_correlations()).max() < 1e-13
    assert abs(dens.sum() - c["N"]) < 1e-10 and abs(cor.sum() - c["N"] ** 2) < 1e-8


def test_measure_run_on_reference_written_files(tmp_path):
    """Interop end to end: the reference's own writers (truns_to_disk,
    basis_to_disk, wave_to_disk) store a synthetic run's files; the device
    measurement of that directory equals the reference's
    measure_operator_dia of the same data."""
    import oracle
    from helpers import model
    from grand_sql_tj_amd import checkpoint as ck, measure
    if not oracle.ref_available():
        pytest.skip("reference build not present")
    sys_len, env_len = 6, 8
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 9)
    run = tmp_path / "run"
    run.mkdir()
    for side, bases, truns in ((0, sb, st), (1, eb, et)):
        for L, b in bases.items():
            oracle.ref_basis_to_disk(run, b, L, side == 0)
        for L, t in truns.items():
            oracle.ref_truns_to_disk(run, t, int(np.asarray(t.sub)[:, 4].sum()), L, side == 0)
    oracle.ref_wave_to_disk(run, ck.wave_file(1), sys_len, env_len, wsec, x)
    dens, cor = measure.measure_run(run)
    log = tmp_path / "log"
    log.mkdir()
    want = oracle.ref_measure_density(model(up_bias=1), log, sys_len, env_len, sb, eb, st, et, wsec, x)
    assert np.abs(dens - want).max() < 1e-13
    assert abs(cor.sum() - N * N) < 1e-10 * N * N
