"""The pairing measurements of Measurement (code/tjsq/main.f90:243-259) on
the CPU: Get_SC_Order (code/tjsq/measure.f90:3226-3376), the singlet pair
amplitude of the y and x bonds along a leg, and Get_mSCOP_Operator_Cor0
(:3377-3507), the pair-pair correlations of one bond with the bonds further
along.  The device programs' task tables (pair chains, rank-1/2 and rank-0
products, the bond-site contractions) are executed by plan_eval, which
asserts every read and write stays inside its buffer, and compared with the
reference's own routines compiled into oracle/_ref (refb_sc_order,
refb_sc_pair_cor0: the reference's transports, products and contractions in
the order and branches of those routines) on synthetic chains with random
truncations that keep every sector.  The exact-diagonalisation pair
correlator used by the GPU tests is pinned here on states whose pair
content is known."""
import numpy as np
import pytest

import ed
import oracle
from helpers import model
from test_cor_cpu import TableRunner, _T
from test_measure_cpu import chain
from grand_sql_tj_amd import measure
from grand_sql_tj_amd.blocks import get_super_basis
from grand_sql_tj_amd.lattice import TJModel, square_neighbours

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")


def pairing_case(ny, sys_len, env_len, seed, up_bias):
    """Chains of both sides (every sector kept at each truncation), the
    wavefunction's sectors and a random normalised wavefunction, and the
    neighbour table of an ny-leg cylinder of sys_len + env_len sites."""
    n = sys_len + env_len
    sb, st = chain(up_bias, sys_len, seed, drop=0.0)
    eb, et = chain(up_bias, env_len, seed + 7, drop=0.0)
    N = int(round(0.875 * n))
    N -= N % 2
    wsec = get_super_basis(sb[sys_len], eb[env_len], N if up_bias else 0, 0)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(int((wsec[:, 4].astype(np.int64) * wsec[:, 5]).sum()))
    x /= np.linalg.norm(x)
    return sb, eb, st, et, wsec, x, square_neighbours(ny, n // ny, n)


def tables_sc_order(ny, sys_len, env_len, sb, eb, st, et, wsec, x, neib, xc1=1):
    n = sys_len + env_len
    return measure.sc_order(0, neib, ny, xc1, n - ny, sys_len, env_len, sb, eb, {L: _T(t) for L, t in st.items()},
                            {L: _T(t) for L, t in et.items()}, wsec, x, runner=TableRunner())


def tables_cor0(up_bias, ny, ij, sys_len, env_len, sb, eb, st, et, wsec, x, neib):
    return measure.sc_pair_cor0(up_bias, neib, ny, 1, ij, sys_len, env_len, sb, eb,
                                {L: _T(t) for L, t in st.items()}, {L: _T(t) for L, t in et.items()}, wsec, x,
                                runner=TableRunner())


# (ny, sys_len, env_len): the bonds of the first leg fall in every branch of
# Get_SC_Order across the three chains (system pair, environment pair, system
# block with the system / environment bond site, bond site with bond site,
# bond site with environment block, environment bond site with block, system
# block with environment block)
SC_GEOMS = [(2, 6, 8), (2, 7, 7), (3, 8, 7)]
# Get_mSCOP_Operator_Cor0: system-side four-fermion operators (ny 2, 3), the
# environment pairs of the first bond (ny 2, 3) and of the bond of site 5 (ny 6)
COR_GEOMS = [(2, 9, 9), (3, 8, 7), (6, 8, 10)]


@needs_ref
@pytest.mark.parametrize("ny,sys_len,env_len", SC_GEOMS)
def test_sc_order_tables_match_reference(tmp_path, ny, sys_len, env_len):
    sb, eb, st, et, wsec, x, neib = pairing_case(ny, sys_len, env_len, 5, 0)
    n = sys_len + env_len
    ref = oracle.ref_sc_order(model(up_bias=0), tmp_path, neib, ny, 1, n - ny, sys_len, env_len, sb, eb, st, et,
                              wsec, x)
    got = tables_sc_order(ny, sys_len, env_len, sb, eb, st, et, wsec, x, neib)
    assert [r[:2] for r in got] == [r[:2] for r in ref]
    g, r = np.array([v[2:] for v in got]), np.array([v[2:] for v in ref])
    assert np.abs(r).max() > 1e-4            # the chains carry pair amplitude
    assert np.abs(g - r).max() < 1e-14 * max(1.0, np.abs(r).max())


@needs_ref
@pytest.mark.parametrize("up_bias", [0, 1])
@pytest.mark.parametrize("ny,sys_len,env_len", COR_GEOMS)
def test_pair_cor0_tables_match_reference(tmp_path, up_bias, ny, sys_len, env_len):
    sb, eb, st, et, wsec, x, neib = pairing_case(ny, sys_len, env_len, 5, up_bias)
    for ij in (1, 3):
        try:
            got = tables_cor0(up_bias, ny, ij, sys_len, env_len, sb, eb, st, et, wsec, x, neib)
        except ValueError:            # the bond leaves the system block: the reference's call is ill-defined
            continue
        ref = oracle.ref_sc_pair_cor0(model(up_bias=up_bias), tmp_path, neib, ny, 1, ij, sys_len, env_len, sb, eb,
                                      st, et, wsec, x)
        assert [r[:4] for r in got] == [r[:4] for r in ref]
        g, r = np.array([v[4:] for v in got]), np.array([v[4:] for v in ref])
        assert np.abs(r).max() > 1e-6
        assert np.abs(g - r).max() < 1e-14 * max(1.0, np.abs(r).max())


def test_sc_order_needs_su2_only_basis():
    sb, eb, st, et, wsec, x, neib = pairing_case(2, 6, 8, 5, 1)
    with pytest.raises(ValueError):
        measure.sc_order(1, neib, 2, 1, 12, 6, 8, sb, eb, st, et, wsec, x, runner=TableRunner())


def test_backward_bond_rejected():
    """A pair bond of Get_mSCOP_Operator_Cor0 must point forward inside the
    system block (the reference carries it from its second site)."""
    sb, eb, st, et, wsec, x, neib = pairing_case(2, 6, 8, 5, 1)
    with pytest.raises(ValueError):
        measure.sc_pair_cor0(1, neib, 2, 2, 1, 6, 8, sb, eb, {L: _T(t) for L, t in st.items()},
                             {L: _T(t) for L, t in et.items()}, wsec, x, runner=TableRunner())


# ---- the exact-diagonalisation pair correlator -------------------------------------------------

def test_ed_pair_correlator_on_a_singlet_pair():
    """|psi> = Delta^+_01 |0> on two sites: <Delta^+_01 Delta_01> = 1, and on
    four sites the resonating state (Delta^+_01 + Delta^+_23)|0>/sqrt(2)
    gives <Delta^+_01 Delta_23> = 1/2, <Delta^+_02 Delta_13> = 0."""
    m2 = TJModel(ny=2, nx=1, up_bias=1)
    st = ed._states(2, 1, 1)
    # Delta^+_01 |0> = (c+_0up c+_1dn - c+_0dn c+_1up)/sqrt(2) |0>; with the up
    # modes ordered first, c+_0dn c+_1up |0> = -c+_1up c+_0dn |0> = -|u1 d0>, so
    # the singlet is (|u0 d1> + |u1 d0>)/sqrt(2) (bits up << n | dn)
    g = np.zeros(st.size)
    g[np.searchsorted(st, (0b01 << 2) | 0b10)] = 1 / np.sqrt(2)
    g[np.searchsorted(st, (0b10 << 2) | 0b01)] = 1 / np.sqrt(2)
    assert abs(ed.pair_correlation(m2, g, st, 2, 0, 1, 0, 1) - 1.0) < 1e-14
    # the triplet (m = 0) has no singlet pair
    g[np.searchsorted(st, (0b10 << 2) | 0b01)] = -1 / np.sqrt(2)
    assert abs(ed.pair_correlation(m2, g, st, 2, 0, 1, 0, 1)) < 1e-14
    m4 = TJModel(ny=2, nx=2, up_bias=1)
    st4 = ed._states(4, 1, 1)
    g4 = np.zeros(st4.size)
    for a, b in ((0, 1), (2, 3)):
        g4[np.searchsorted(st4, ((1 << a) << 4) | (1 << b))] += 0.5
        g4[np.searchsorted(st4, ((1 << b) << 4) | (1 << a))] += 0.5
    assert abs(np.linalg.norm(g4) - 1.0) < 1e-14
    assert abs(ed.pair_correlation(m4, g4, st4, 2, 0, 1, 2, 3) - 0.5) < 1e-14
    assert abs(ed.pair_correlation(m4, g4, st4, 2, 0, 1, 0, 1) - 0.5) < 1e-14
    assert abs(ed.pair_correlation(m4, g4, st4, 2, 0, 2, 1, 3)) < 1e-14
    # symmetric in the two sites of a pair
    assert abs(ed.pair_correlation(m4, g4, st4, 2, 1, 0, 3, 2) - 0.5) < 1e-14

