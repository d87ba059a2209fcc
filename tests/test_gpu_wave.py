"""Wavefunction transformation between bonds on the GPU through the C-ABI
(wave_trans_LR / wave_trans_RL, code/tjsq/dmrg_sub.f90:1093-1384), against
the oracle: the normalised start vector of the next solve."""
import numpy as np
import pytest

import oracle
from helpers import wave_args, wave_transform_case
from grand_sql_tj_amd import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda")


def _device(c, runs=1):
    pg = capi.Program.wave_transform(c["direction"], c["sys_bs"], c["env_bs"], c["sys_trun"], c["env_trun"],
                                     c["wsec_old"], c["wsec_new"])
    ops = _cuda(np.concatenate([c["sys_trun"].data, c["env_trun"].data]))
    x = _cuda(c["x"])
    out = []
    for _ in range(runs):
        y = torch.full((max(1, pg.sizes["ndata"]),), float("nan"), dtype=torch.float64, device="cuda")
        pg.run(ops, x, y)
        torch.cuda.synchronize()
        out.append(y.cpu().numpy()[:pg.sizes["ndata"]])
    return out


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("ub", [0, 1])
def test_wave_transform_matches_oracle(direction, ub):
    c = wave_transform_case(direction, 40 + direction, up_bias=ub)
    y = _device(c)[0]
    ref = oracle.wave_trans(*wave_args(c))
    assert np.isfinite(y).all()
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    assert abs(np.linalg.norm(y) - 1.0) < 1e-13


@pytest.mark.parametrize("direction", [0, 1])
def test_wave_transform_large_blocks(direction):
    """Blocks large enough for 64x64 work items and split K."""
    c = wave_transform_case(direction, 7, up_bias=1, ls=9, le=10, m=400, median=40, nsec=10)
    y1, y2 = _device(c, runs=2)
    ref = oracle.wave_trans(*wave_args(c))
    assert np.linalg.norm(y1 - ref) / np.linalg.norm(ref) < 1e-13
    assert np.array_equal(y1, y2)          # repeat runs are bitwise the same


def test_wave_transform_round_trip():
    c = wave_transform_case(0, 12, up_bias=1, frac=1.0, drop=0.0, m=120, median=12)
    y = _device(c)[0]
    back = {"direction": 1, "sys_bs": c["sys_new"], "env_bs": c["env_old"], "sys_trun": c["sys_trun"],
            "env_trun": c["env_trun"], "wsec_old": c["wsec_new"], "x": y, "wsec_new": c["wsec_old"]}
    z = _device(back)[0]
    assert np.linalg.norm(z - c["x"]) < 1e-12
