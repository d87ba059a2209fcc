"""The pairing measurements on the device (measure.sc_order: Get_SC_Order,
code/tjsq/measure.f90:3226-3376; measure.sc_pair_cor0:
Get_mSCOP_Operator_Cor0, :3377-3507) against the same programs' task tables
run on the CPU and the reference's own routines on synthetic chains, and,
through the untruncated device DMRG, against exact diagonalisation: the
pair-pair correlations <Delta^+_ij Delta_kl> of the singlet pair
Delta^+_ij = (c^+_i,up c^+_j,dn - c^+_i,dn c^+_j,up) / sqrt(2) in the exact
ground state, bond by bond."""
import numpy as np
import pytest

import ed
import oracle
from helpers import model
from test_pairing_cpu import COR_GEOMS, SC_GEOMS, pairing_case, tables_cor0, tables_sc_order
from grand_sql_tj_amd.lattice import TJModel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


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


@pytest.mark.parametrize("ny,sys_len,env_len", SC_GEOMS)
def test_device_sc_order_matches_tables_and_reference(tmp_path, ny, sys_len, env_len):
    from grand_sql_tj_amd import measure
    sb, eb, st, et, wsec, x, neib = pairing_case(ny, sys_len, env_len, 5, 0)
    n = sys_len + env_len
    print(f"sc_order ny={ny} {sys_len}|{env_len} D={x.size}", flush=True)
    got = measure.sc_order(0, neib, ny, 1, n - ny, sys_len, env_len, sb, eb, {L: _Trun(t) for L, t in st.items()},
                           {L: _Trun(t) for L, t in et.items()}, wsec, torch.from_numpy(x).cuda())
    tab = tables_sc_order(ny, sys_len, env_len, sb, eb, st, et, wsec, x, neib)
    g, t = np.array([r[2:] for r in got]), np.array([r[2:] for r in tab])
    assert [r[:2] for r in got] == [r[:2] for r in tab]
    assert np.abs(g - t).max() < 1e-13
    if oracle.ref_available():
        ref = oracle.ref_sc_order(model(up_bias=0), tmp_path, neib, ny, 1, n - ny, sys_len, env_len, sb, eb, st, et,
                                  wsec, x)
        assert np.abs(g - np.array([r[2:] for r in ref])).max() < 1e-13


@pytest.mark.parametrize("up_bias", [0, 1])
@pytest.mark.parametrize("ny,sys_len,env_len", COR_GEOMS)
def test_device_pair_cor0_matches_tables_and_reference(tmp_path, up_bias, ny, sys_len, env_len):
    from grand_sql_tj_amd import measure
    sb, eb, st, et, wsec, x, neib = pairing_case(ny, sys_len, env_len, 5, up_bias)
    for ij in (1, 3):
        print(f"pair_cor0 ub={up_bias} ny={ny} {sys_len}|{env_len} ij={ij} D={x.size}", flush=True)
        try:
            tab = tables_cor0(up_bias, ny, ij, sys_len, env_len, sb, eb, st, et, wsec, x, neib)
        except ValueError:
            continue
        got = measure.sc_pair_cor0(up_bias, neib, ny, 1, ij, sys_len, env_len, sb, eb,
                                   {L: _Trun(t) for L, t in st.items()}, {L: _Trun(t) for L, t in et.items()}, wsec,
                                   torch.from_numpy(x).cuda())
        assert [r[:4] for r in got] == [r[:4] for r in tab]
        g = np.array([r[4:] for r in got])
        assert np.abs(g - np.array([r[4:] for r in tab])).max() < 1e-13
        if oracle.ref_available():
            ref = oracle.ref_sc_pair_cor0(model(up_bias=up_bias), tmp_path, neib, ny, 1, ij, sys_len, env_len, sb,
                                          eb, st, et, wsec, x)
            assert np.abs(g - np.array([r[4:] for r in ref])).max() < 1e-13


def test_untruncated_dmrg_pair_correlations_match_exact_diagonalisation():
    """2 x 8 cylinder, N = 14, every multiplet kept: the rows of
    Get_mSCOP_Operator_Cor0 for the y bond (1, 2) and the x bond (1, 3) equal
    the exact ground state's <Delta^+ Delta> of the same bonds, sign
    included, to the state's convergence (the reference writes their
    absolute values, code/tjsq/readme.md:66-75)."""
    from grand_sql_tj_amd import sweep
    m = TJModel(ny=2, nx=8, t=1.0, tp=0.0, J=0.4, mu=0.0, up_bias=1)
    N = 14
    d = sweep.DMRG(m, total_up=N, kept_max=4000, trun_err=1e-14, tol=1e-12)
    d.run(sweeps=1)
    d.left_to_right(1, m.nsite // 2)          # to the centre bond, 8 | 8
    _, g, states = ed.ground_state(m, N)
    neib = d._neighbours()
    sl, el = d.bond[0] + 1, d.bond[1] + 1
    print(f"bond {sl}|{el}", flush=True)
    got, want = [], []
    for ij in (1, 3):
        s1, s2 = 1, int(neib[1, ij - 1])
        for row in d.measure_sc_pair_cor0(1, ij):
            s3 = row[1]
            for k, i in enumerate((1, 3, 5)):
                s4 = int(neib[s3, i - 1])
                if s4 == 0 or row[4 + k] == 0.0:
                    continue
                got.append(row[4 + k])
                want.append(ed.pair_correlation(m, g, states, N, s1 - 1, s2 - 1, s3 - 1, s4 - 1))
    got, want = np.array(got), np.array(want)
    print(np.c_[got, want], flush=True)
    assert got.size >= 10 and np.abs(want).max() > 1e-3
    assert np.abs(got - want).max() < 2e-5


def test_dmrg_pair_amplitude_matches_exact_diagonalisation():
    """4 x 3 cylinder without the particle label (up_bias 0), the chemical
    potential set where the N = 8 and N = 10 ground states are degenerate
    (mu = (E_10 - E_8) / 2 from exact diagonalisation): the untruncated
    device DMRG converges to a|g_8> + b|g_10>, whose singlet pair amplitude
    on every bond is <Delta^+_ij> = ab <g_10|Delta^+_ij|g_8> (Delta^+ only
    raises N).  a^2 and b^2 follow from the measured <N>; the amplitude of
    every bond of all four legs (all branches of Get_SC_Order) must equal
    that, up to the sign of ab."""
    from grand_sql_tj_amd import sweep
    base = TJModel(ny=4, nx=3, t=1.0, tp=0.0, J=0.4, mu=0.0, up_bias=1)
    E = {N: ed.ground_energy(base, N) for N in (6, 8, 10, 12)}
    mu = 0.5 * (E[10] - E[8])
    shifted = {N: E[N] - mu * N for N in E}
    assert min(shifted[6], shifted[12]) - shifted[8] > 1e-2          # only N = 8, 10 are degenerate
    m = TJModel(ny=4, nx=3, t=1.0, tp=0.0, J=0.4, mu=mu, up_bias=0)
    d = sweep.DMRG(m, total_up=0, kept_max=4000, trun_err=1e-14, tol=1e-13)
    d.run(sweeps=1)
    d.left_to_right(1, m.nsite // 2)                  # centre bond, 6 | 6
    e0 = d.log[-1].energy
    print(f"E0 {e0:.12f} exact {shifted[8]:.12f}", flush=True)
    assert abs(e0 - shifted[8]) < 1e-9
    nbar = float(np.sum(d.measure_density()))
    a2, b2 = (10.0 - nbar) / 2.0, (nbar - 8.0) / 2.0
    amp = np.sqrt(max(a2, 0.0) * max(b2, 0.0))
    print(f"<N> = {nbar:.10f}, |ab| = {amp:.6f}", flush=True)
    assert amp > 0.05
    lo_base = TJModel(ny=4, nx=3, t=1.0, tp=0.0, J=0.4, mu=0.0, up_bias=1)
    _, g8, s8 = ed.ground_state(lo_base, 8)
    _, g10, s10 = ed.ground_state(lo_base, 10)
    got, want = [], []
    neib = d._neighbours()
    for xc1 in range(1, 5):
        for sid1, _, vy, vx in d.measure_sc_order(xc1, m.nsite - m.ny):
            for ij, v in ((1, vy), (3, vx)):
                sid2 = int(neib[sid1, ij - 1])
                got.append(v)
                want.append(amp * ed.pair_amplitude(lo_base, g8, s8, g10, s10, sid1 - 1, sid2 - 1))
    got, want = np.array(got), np.array(want)
    print(np.c_[got, want], flush=True)
    k = int(np.argmax(np.abs(want)))
    assert abs(want[k]) > 1e-2
    sign = np.sign(got[k] * want[k])
    assert np.abs(got - sign * want).max() < 2e-5
