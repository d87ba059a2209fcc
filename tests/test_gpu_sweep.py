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


def test_checkpoint_resume(tmp_path):
    """save_checkpoint writes the run as the reference's checkpoint files
    (grand_sql_tj_amd.checkpoint, pinned byte for byte against the
    reference's writers on the CPU); a new run that loads them continues the
    sweep exactly as the run that wrote them."""
    import numpy as np
    from grand_sql_tj_amd import checkpoint as ck
    from grand_sql_tj_amd import sweep
    case = next(c for c in ED if c["ny"] == 2 and c["nx"] == 8)
    m = _model(case)
    kw = dict(total_up=case["N"], kept_max=48, trun_err=1e-12, tol=1e-12)
    a = sweep.DMRG(m, **kw)
    N = a.nsite
    a.warm_up()
    a.left_to_right(N // 2 - 1, N - 2)
    a.right_to_left(N - 2, 1)
    a.save_checkpoint(tmp_path)
    for side, blocks in ((0, a.sys), (1, a.env)):
        for L, blk in blocks.items():
            ops, dims = ck.load_model(tmp_path / ck.file_name(side, "blocks", L))
            assert ops.length == L and np.array_equal(ops.sub, blk.ops.sub) and np.array_equal(ops.off, blk.ops.off)
            assert np.array_equal(ops.data, blk.data[:ops.data.size].cpu().numpy())
            assert np.array_equal(dims, ck.default_dims(ops))
    assert any(t.ops.sub[:, 4].sum() < t.ops.sub[:, 3].sum() for t in a.env_trun.values())   # truncated
    b = sweep.DMRG(m, **kw)
    b.load_checkpoint(tmp_path)
    assert sorted(b.sys) == sorted(a.sys) and sorted(b.env) == sorted(a.env)
    assert sorted(b.env_trun) == sorted(a.env_trun) and sorted(b.env_basis) == sorted(a.env_basis)
    assert b.bond == a.bond and np.array_equal(b.wsec, a.wsec)
    n0 = len(a.log)
    a.left_to_right(1, N - 2)
    b.left_to_right(1, N - 2)
    ea = [r.energy for r in a.log[n0:]]
    eb = [r.energy for r in b.log]
    assert len(ea) == len(eb) == N - 3
    assert max(abs(x - y) for x, y in zip(ea, eb)) <= 1e-12 * abs(case["E0"])
    assert min(eb) >= case["E0"] - 1e-9
