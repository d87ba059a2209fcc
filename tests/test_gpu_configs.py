"""The batched-apply and serving configurations at their full sizes on the
GPU: 64 independent applies of width-4/6 cylinders with m ~ U[1000, 6000]
(block dims log-normal, median 32, max 512), and a serving scan of widths
4/6, lengths 8-16, m ~ U[500, 6000] with continuous admission.  No CPU
oracle finishes at these sizes; the checks are the size-independent ones:
bitwise equality of grouped and single launches, of graph replays and of
results alone and in the mix, linearity, agreement of the precisions."""
import concurrent.futures as cf
import time

import numpy as np
import pytest

import bench
from grand_sql_tj_amd import capi, serve, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


def _synth(fn, n, seed, workers=16):
    """bench's workload generators with FastNormals entries, in host threads
    (numpy releases the GIL), yielded in order"""
    with cf.ThreadPoolExecutor(workers) as ex:
        for k in range(0, n, workers):
            yield from ex.map(fn, [(i, seed, True) for i in range(k, min(n, k + workers))])


@pytest.fixture(scope="module")
def batch64():
    plans, xs, meta = [], [], []
    for model, sb, m in _synth(bench.batch_member, 64, 2024):
        plans.append(capi.Plan(model, sb, precisions=("fp64", "fp32", "bf16x3")))
        xs.append(torch.from_numpy(synth.random_vector(sb, 11 + len(xs))).cuda())
        meta.append(m)
        del sb
    return plans, xs, meta


@pytest.mark.timeout(600)
def test_batch64_workload_shape(batch64):
    plans, xs, meta = batch64
    assert len(plans) == 64
    ms = [m["m"] for m in meta]
    assert min(ms) >= 1000 and max(ms) <= 6000 and {m["ny"] for m in meta} == {4, 6}
    assert all(p.stats["packed"] == 1 for p in plans)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_batch64_equals_members_and_replays(batch64, prec):
    plans, xs, _ = batch64
    b = capi.Batch(plans, precisions=(prec,))
    x = torch.cat(xs).float()
    y = b.apply(x, precision=prec)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        lo, hi = int(b.offsets[i]), int(b.offsets[i + 1])
        yi = p.apply(xs[i].float(), precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(y[lo:hi], yi), i
    # the grouped launch captured as a hipGraph: every replay, with eager work
    # between them, equals the eager result
    b.graph(True)
    for r in range(3):
        yg = b.apply(x, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(yg, y), r
        _ = (x * 2.0).sum().item()
    b.graph(False)
    del b


@pytest.mark.timeout(600)
def test_batch64_members_linear_and_precisions_agree(batch64):
    plans, xs, _ = batch64
    for i, p in enumerate(plans):
        x = xs[i]
        z = x.flip(0)
        y = p.apply(x)
        ylin = p.apply(2.0 * x - 0.5 * z)
        ref = 2.0 * y - 0.5 * p.apply(z)
        assert float((ylin - ref).norm() / ref.norm()) < 1e-13, i
        for prec, tol in (("fp32", 2e-6), ("bf16x3", 2e-5)):
            yp = p.apply(x.float(), precision=prec).double()
            assert float((yp - y).norm() / y.norm()) < tol, (i, prec)


@pytest.mark.timeout(600)
def test_serving_scan_same_alone_and_in_the_mix():
    reqs = []
    for k, (model, sb, psi0, meta) in enumerate(_synth(bench.serve_request, 12, 7)):
        nsite = sb.sys_len + sb.env_len + 2
        # random start vectors: up to the reference's loop_in = 110 iterations
        # (eigen.f90:12), so the solves converge
        reqs.append(serve.Request(k, model, sb, psi0, precision="bf16x3", tol=1e-10, max_iter=110,
                                  energy_tol=0.1 * 1e-8 * nsite))
    ms = [r.sb.sys.data.size for r in reqs]
    assert len({r.model.ny for r in reqs}) == 2 and len(ms) == 12
    alone = [serve.solve_alone(r) for r in reqs]
    with serve.Server(slots=4) as srv:
        futs = []
        for i, r in enumerate(reqs):
            futs.append(srv.submit(r))
            if i % 3 == 2:
                time.sleep(0.05)       # later requests arrive while earlier ones run
        res = [f.wait(300) for f in futs]
        assert srv.peak_in_use <= srv.budget
    assert len({r.slot for r in res}) > 1
    for (e0, psi), r in zip(alone, res):
        assert r.error is None, r.error
        assert r.e0 == e0
        assert np.array_equal(r.psi, psi)
    # the mixed-precision energies against FP64 solves of three of them
    for r, (e0, _) in list(zip(reqs, alone))[:3]:
        nsite = r.sb.sys_len + r.sb.env_len + 2
        p = capi.Plan(r.model, r.sb, precisions=("fp64",))
        e64 = p.lanczos(torch.from_numpy(r.psi0).cuda(), tol=1e-12, max_iter=250)["e0"]
        assert abs(e0 - e64) / nsite < 1e-8, (r.id, e0, e64)
