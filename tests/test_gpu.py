"""HIP path parity on the GPU: the C-ABI against the oracle and the reference."""
import os

import numpy as np
import pytest

import oracle
import lanczos as olanczos
from helpers import GEOMS, model, rel, small_superblock
from grand_sql_tj_amd import capi, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _cuda(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return t if dtype is None else t.to(dtype)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("geom", GEOMS)
def test_apply_fp64_matches_oracle(ub, geom):
    m = model(up_bias=ub)
    sb = small_superblock(m, *geom, seed=7 * geom[0] + geom[1] + ub, msys=40, menv=44, median=5)
    x = synth.random_vector(sb, 9)
    p = capi.Plan(m, sb)
    y = p.apply(_cuda(x)).cpu().numpy()
    assert rel(y, oracle.model_wave(m, sb, x)) < 1e-13


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 2e-5)])
def test_apply_mixed_precision_tolerance(prec, tol):
    m = model()
    sb = small_superblock(m, 11, 11, seed=21, msys=120, menv=120, median=12, nsec=10)
    x = synth.random_vector(sb, 9)
    p = capi.Plan(m, sb)
    y = p.apply(_cuda(x, torch.float32), precision=prec).double().cpu().numpy()
    assert rel(y, oracle.model_wave(m, sb, x)) < tol


def test_apply_larger_blocks_fp64():
    # sub-blocks larger than one 32x32 tile and K larger than one LDS chunk
    m = model()
    sb = small_superblock(m, 11, 11, seed=23, msys=400, menv=380, median=40, nsec=8)
    x = synth.random_vector(sb, 4)
    p = capi.Plan(m, sb)
    y = p.apply(_cuda(x)).cpu().numpy()
    assert rel(y, oracle.model_wave(m, sb, x)) < 1e-13


def test_apply_symmetric_and_deterministic():
    m = model()
    sb = small_superblock(m, 12, 10, seed=3, msys=80, menv=80, median=8)
    p = capi.Plan(m, sb)
    x, z = _cuda(synth.random_vector(sb, 1)), _cuda(synth.random_vector(sb, 2))
    hx, hz = p.apply(x), p.apply(z)
    assert abs(float(torch.dot(z, hx) - torch.dot(x, hz))) < 1e-12 * float(hx.norm() * z.norm())
    assert torch.equal(p.apply(x), hx)   # pull-style apply: bitwise repeatable


def test_lanczos_fp64_matches_reference_and_oracle(tmp_path):
    m = model()
    sb = small_superblock(m, 11, 11, seed=3 + 1, msys=30, menv=30, median=4)
    x0 = synth.random_vector(sb, 1)
    e_o, psi_o, k_o, _ = olanczos.lanczos(lambda v: oracle.model_wave(m, sb, v), x0, 1e-10, 110)
    p = capi.Plan(m, sb)
    psi = _cuda(x0)
    r = p.lanczos(psi, tol=1e-10, max_iter=110, precision="fp64")
    assert abs(r["e0"] - e_o) <= 1e-10 * abs(e_o)
    assert abs(abs(float(np.dot(psi.cpu().numpy(), psi_o))) - 1.0) < 1e-8
    if oracle.ref_available():
        cwd = os.getcwd()
        os.chdir(tmp_path)
        try:
            e_ref, _ = oracle.ref_lanczos(m, sb, x0, 1e-10)
        finally:
            os.chdir(cwd)
        assert abs(r["e0"] - e_ref) <= 1e-10 * abs(e_ref)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_lanczos_mixed_precision_energy(prec):
    m = model()
    sb = small_superblock(m, 11, 11, seed=8, msys=120, menv=120, median=12, nsec=10)
    x0 = synth.random_vector(sb, 5)
    p = capi.Plan(m, sb)
    ref = _cuda(x0)
    r64 = p.lanczos(ref, tol=1e-10, max_iter=80, precision="fp64")
    psi = _cuda(x0)
    r = p.lanczos(psi, tol=1e-10, max_iter=80, precision=prec, refine_tol=1e-6, max_refine=2)
    nsite = sb.sys_len + sb.env_len + 2
    assert abs(r["e0"] - r64["e0"]) / nsite <= 1e-8
    assert r["residual"] <= 1e-5
    # energy-targeted refinement: restart in FP64 only if |r|^2/gap > target
    psi2 = _cuda(x0)
    target = 1e-9 * nsite
    r2 = p.lanczos(psi2, tol=1e-10, max_iter=80, precision=prec, max_refine=2, energy_tol=target)
    assert abs(r2["e0"] - r64["e0"]) / nsite <= 1e-8
    assert r2["energy_err_est"] <= target or r2["refinements"] == 2


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_lanczos_graph_replay_is_bitwise_the_launched_solve(prec):
    """Each iteration captured as a hipGraph (first solve) and replayed
    (second solve) gives the launched form's result bit for bit."""
    m = model()
    sb = small_superblock(m, 11, 11, seed=21, msys=120, menv=120, median=12, nsec=10)
    x = _cuda(synth.random_vector(sb, 9))
    p = capi.Plan(m, sb)
    s = torch.cuda.Stream()
    outs = []
    with torch.cuda.stream(s):
        for graphs in (False, True, True):
            p.graphs(graphs)
            psi = x.clone()
            r = p.lanczos(psi, tol=1e-10, max_iter=30, precision=prec, energy_tol=1e-7 if prec != "fp64" else 0.0)
            s.synchronize()
            outs.append((r["e0"], r["iterations"], psi.cpu()))
    for e0, it, psi in outs[1:]:
        assert e0 == outs[0][0] and it == outs[0][1]
        assert torch.equal(psi, outs[0][2])
    assert outs[0][1] >= 3


@pytest.mark.parametrize("shape", [None, {"merge_right": 1}, {"merge_right": 1, "big": 32, "xcds": 8}])
def test_apply_work_item_shapes_match_oracle(shape):
    m = model()
    sb = small_superblock(m, 11, 11, seed=23, msys=400, menv=380, median=40, nsec=8)
    x = synth.random_vector(sb, 4)
    p = capi.Plan(m, sb, shape=shape)
    ref = oracle.model_wave(m, sb, x)
    assert rel(p.apply(_cuda(x)).cpu().numpy(), ref) < 1e-13
    y = p.apply(_cuda(x, torch.float32), precision="bf16x3").double().cpu().numpy()
    assert rel(y, ref) < 2e-5


def test_plan_set_ops_equals_a_new_plan():
    # new block data with the same layouts (a later sweep at the same bond):
    # every precision's apply bitwise that of a plan compiled for the new data
    import dataclasses
    from grand_sql_tj_amd.blocks import OpSet
    m = model()
    sb = small_superblock(m, 11, 11, seed=13, msys=80, menv=90, median=8)
    scaled = lambda o, f: OpSet(o.length, o.hdr, o.sub, o.off, o.data * f)
    sb2 = dataclasses.replace(sb, sys=scaled(sb.sys, 0.6), env=scaled(sb.env, -1.2))
    p = capi.Plan(m, sb)
    p.set_ops(sb2.sys, sb2.env)
    q = capi.Plan(m, sb2)
    x = _cuda(synth.random_vector(sb, 2))
    for prec in ("fp64", "fp32", "bf16x3"):
        xx = x if prec == "fp64" else x.float()
        assert torch.equal(p.apply(xx, precision=prec), q.apply(xx, precision=prec)), prec
    other = small_superblock(m, 7, 9, seed=14)
    with pytest.raises(RuntimeError, match="layouts differ"):
        p.set_ops(other.sys, other.env)
