"""Two-point functions of rank-k site operators on the device
(measure.pair_correlations: Get_density_cor, code/tjsq/measure.f90:
3110-3223, through tjsq_ndia_transport_create, tjsq_pair_ndia_create and
tjsq_cor_ndia_create) against the reference's own routines on synthetic
chains with random truncations, against the same programs' task tables run
on the CPU, by the sum rules of fixed-N singlet wavefunctions, and, through
the untruncated device DMRG, against exact diagonalisation: <n_i n_j>,
<S_i . S_j> and sum over s of <c^+_is c_js> of the exact ground state,
pair by pair."""
import json
import os

import numpy as np
import pytest

import ed
import measure_oracle as mo
import oracle
from helpers import model
from test_cor_cpu import COR_GEOMS, tables_cor
from test_measure_cpu import case
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


def device_cor(kind, sid1, sys_len, env_len, sb, eb, st, et, wsec, x):
    from grand_sql_tj_amd import measure
    return measure.pair_correlations(kind, 1, sid1, sys_len, env_len, sb, eb, {L: _Trun(t) for L, t in st.items()},
                                     {L: _Trun(t) for L, t in et.items()}, wsec, torch.from_numpy(x).cuda())


@pytest.mark.parametrize("sys_len,env_len", COR_GEOMS)
@pytest.mark.parametrize("kind", ["density", "spin", "hopping"])
def test_device_matches_reference_and_tables(tmp_path, kind, sys_len, env_len):
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 3 + sys_len)
    for sid1 in sorted({1, 3, sys_len - 1}):
        got = device_cor(kind, sid1, sys_len, env_len, sb, eb, st, et, wsec, x)
        tab = tables_cor(kind, sid1, sys_len, env_len, sb, eb, st, et, wsec, x)
        m = ~np.isnan(tab)
        assert np.array_equal(m, ~np.isnan(got))
        assert np.abs(got[m] - tab[m]).max() < 1e-13
        if oracle.ref_available():
            ref = oracle.ref_density_cor(model(up_bias=1), tmp_path, kind, sid1, sys_len, env_len, sb, eb, st, et,
                                         wsec, x)
            assert np.abs(got[m] - ref[m]).max() < 1e-13


@pytest.mark.parametrize("sys_len,env_len", COR_GEOMS)
def test_device_sum_rules(sys_len, env_len):
    sb, eb, st, et, wsec, x, N = case(sys_len, env_len, 3 + sys_len)
    n1 = mo.density_profile(1, sys_len + env_len, sys_len, env_len, sb, eb, st, et, wsec, x)[0]
    dens = device_cor("density", 1, sys_len, env_len, sb, eb, st, et, wsec, x)
    spin = device_cor("spin", 1, sys_len, env_len, sb, eb, st, et, wsec, x)
    assert abs(np.nansum(dens) - (N - 1) * n1) < 1e-12 * N
    assert abs(np.nansum(spin) + 0.75 * n1) < 1e-13


@pytest.mark.parametrize("case_", [c for c in ED if c["nx"] * c["ny"] <= 12][:2],
                         ids=lambda c: f"{c['ny']}x{c['nx']}N{c['N']}")
def test_untruncated_dmrg_matches_exact_diagonalisation(case_):
    from grand_sql_tj_amd import sweep
    m = TJModel(ny=case_["ny"], nx=case_["nx"], t=case_["t"], tp=case_["tp"], J=case_["J"], mu=case_["mu"],
                up_bias=1)
    d = sweep.DMRG(m, total_up=case_["N"], kept_max=1000, trun_err=1e-12, tol=1e-12)
    d.run(sweeps=1)
    _, g, states = ed.ground_state(m, case_["N"])
    want = {"density": ed.density_correlation(m, g, states), "spin": ed.spin_correlation(m, g, states),
            "hopping": ed.hopping_correlation(m, g, states)}
    sl = d.bond[0] + 1
    n = m.nsite
    for kind, w in want.items():
        for sid1 in range(1, sl):
            got = d.measure_pair_correlations(kind, sid1)
            print(f"{kind} sid1={sid1}", flush=True)
            # the state converges to the Lanczos tolerance on the energy, so
            # to ~1e-6 in norm: two-point functions agree to that order
            assert np.abs(got[sid1:] - w[sid1 - 1, sid1:n]).max() < 2e-5, (kind, sid1)
