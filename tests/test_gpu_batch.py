"""Batched grouped launch, hipGraph replay and work-item shapes on the GPU."""
import pytest

import oracle
from helpers import model, small_superblock, rel
from grand_sql_tj_amd import capi, synth, autotune

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


def _members():
    out = []
    for ub, geom, seed, ms, me, med in [(1, (11, 11), 21, 160, 150, 14), (0, (7, 9), 3, 40, 44, 5),
                                        (1, (12, 10), 8, 240, 220, 30), (1, (5, 5), 4, 24, 28, 3),
                                        (0, (13, 17), 9, 120, 100, 12)]:
        m = model(up_bias=ub)
        sb = small_superblock(m, *geom, seed=seed, msys=ms, menv=me, median=med, nsec=10)
        out.append((m, sb))
    return out


@pytest.fixture(scope="module")
def members():
    mem = _members()
    plans = [capi.Plan(m, sb) for m, sb in mem]
    xs = [torch.from_numpy(synth.random_vector(sb, 5 + i)).cuda() for i, (m, sb) in enumerate(mem)]
    return mem, plans, xs


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_batch_apply_bit_identical_to_members(members, prec):
    mem, plans, xs = members
    b = capi.Batch(plans)
    dt = torch.float64 if prec == "fp64" else torch.float32
    x = torch.cat(xs).to(dt)
    y = b.apply(x, precision=prec)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        lo, hi = int(b.offsets[i]), int(b.offsets[i + 1])
        yi = p.apply(xs[i].to(dt), precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(y[lo:hi], yi), i
    if prec == "fp64":
        m, sb = mem[1]
        lo, hi = int(b.offsets[1]), int(b.offsets[2])
        assert rel(y[lo:hi].cpu().numpy(), oracle.model_wave(m, sb, xs[1].cpu().numpy())) < 1e-13


def test_batch_graph_replays_match_eager(members):
    mem, plans, xs = members
    b = capi.Batch(plans)
    x = torch.cat(xs)
    eager = b.apply(x).clone()
    b.graph(True)
    y = torch.full_like(x, float("nan"))
    for rep in range(4):
        b.apply(x, y)
        torch.cuda.synchronize()
        assert torch.equal(y, eager), rep
        # eager work and a synchronisation between the replays
        plans[0].apply(xs[0])
        torch.cuda.synchronize()
        y.fill_(float("nan"))
    # new vectors: the graph is captured again
    x2 = torch.cat([v.flip(0) for v in xs])
    b.graph(False)
    e2 = b.apply(x2).clone()
    b.graph(True)
    y2 = b.apply(x2)
    torch.cuda.synchronize()
    assert torch.equal(y2, e2)
    xf = x.float()
    b.graph(False)
    ef = b.apply(xf, precision="fp32").clone()
    b.graph(True)
    for _ in range(3):
        yf = b.apply(xf, precision="fp32")
        torch.cuda.synchronize()
        assert torch.equal(yf, ef)


def test_reshape_and_autotune_keep_results(members):
    mem, plans, xs = members
    m, sb = mem[2]
    p = capi.Plan(m, sb)
    ref = oracle.model_wave(m, sb, xs[2].cpu().numpy())
    for shape in autotune.CANDIDATES:
        p.reshape(shape)
        assert rel(p.apply(xs[2]).cpu().numpy(), ref) < 1e-13, shape
    res = autotune.tune(p, m, sb, precision="fp32", reps=2)
    assert res["shape"] in autotune.CANDIDATES and len(res["results"]) == len(autotune.CANDIDATES)
    assert rel(p.apply(xs[2]).cpu().numpy(), ref) < 1e-13
