"""Device DMRG (warm-up and sweeps assembled from the library's device calls,
grand_sql_tj_amd.sweep) against exact diagonalisation of the t-J model
(tests/golden/ed_energies.json, written by tests/golden/make_ed.py).

With kept_max above every block dimension nothing is truncated and the
energy is exact; with truncation it stays above the exact energy and
approaches it as kept_max grows.  This pins the whole chain (hmerge and the
14 apply kernels, Lanczos, density matrix, truncation, block update and the
wave transforms between bonds) against the physics, independently of the
reference."""
import json
import os

import pytest

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


def _model(c):
    return TJModel(ny=c["ny"], nx=c["nx"], t=c["t"], tp=c["tp"], J=c["J"], mu=c["mu"], up_bias=1)


def _run(c, kept_max, sweeps=1, precision="fp64"):
    from grand_sql_tj_amd import sweep
    d = sweep.DMRG(_model(c), total_up=c["N"], kept_max=kept_max, trun_err=1e-12, tol=1e-12, precision=precision)
    return d.run(sweeps=sweeps), d


@pytest.mark.parametrize("case", [c for c in ED if c["nx"] * c["ny"] <= 12], ids=lambda c: f"{c['ny']}x{c['nx']}N{c['N']}tp{c['tp']}mu{c['mu']}")
def test_untruncated_dmrg_is_exact(case):
    e, d = _run(case, kept_max=1000)
    assert abs(e - case["E0"]) < 1e-9 * abs(case["E0"])
    # every step of the last sweep holds the same state (the transforms carry it exactly)
    last = [r.energy for r in d.log[-(2 * (d.nsite - 3)):]]
    assert max(last) - min(last) < 1e-9


@pytest.mark.parametrize("ny,nx", [(2, 8), (4, 4)])
def test_truncated_dmrg_converges_to_exact(ny, nx):
    case = next(c for c in ED if c["ny"] == ny and c["nx"] == nx)
    e_small, _ = _run(case, kept_max=48, sweeps=2)
    e_large, d = _run(case, kept_max=400, sweeps=2)
    assert e_small >= case["E0"] - 1e-9          # variational
    assert e_large >= case["E0"] - 1e-9
    assert e_large <= e_small + 1e-12
    assert abs(e_large - case["E0"]) < 1e-6 * abs(case["E0"])
    assert any(r.kept < 400 and r.trunc_err > 0 for r in d.log) or e_small > case["E0"] + 1e-9


def test_mixed_precision_sweep():
    """BF16x3 Lanczos with the FP64 Rayleigh quotient inside the sweep."""
    case = next(c for c in ED if c["ny"] == 2 and c["nx"] == 6 and c["N"] == 10 and c["tp"] == 0.0 and c["mu"] == 0.0)
    e, _ = _run(case, kept_max=1000, precision="bf16x3")
    assert abs(e - case["E0"]) < 1e-7
