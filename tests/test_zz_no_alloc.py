"""The apply that allocates nothing (tjsq_heff_apply_noalloc) against the
allocating entry points, through the C-ABI, on two calls in a row (the
second plan smaller)."""
import ctypes as C

import numpy as np
import pytest

from helpers import model, small_superblock
from grand_sql_tj_amd import capi, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


def _noalloc(plan, precision, x, y):
    lib = capi.lib()
    s = torch.cuda.current_stream().cuda_stream
    return lib.tjsq_heff_apply_noalloc(plan.h, capi.PRECISIONS[precision], C.c_void_p(x.data_ptr()),
                                       C.c_void_p(y.data_ptr()), C.c_void_p(s))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_apply_equals_allocating_apply(precision):
    m = model(up_bias=1)
    # the larger plan first, then a smaller one: the second call reuses the
    # array the first one left behind
    cases = [small_superblock(m, 11, 11, seed=61, msys=160, menv=170, median=20, nsec=8),
             small_superblock(m, 7, 9, seed=62, msys=40, menv=44, median=5)]
    assert cases[1].dim < cases[0].dim
    dt = torch.float64 if precision == "fp64" else torch.float32
    for sb in cases:
        p = capi.Plan(m, sb, precisions=("fp64", "fp32"))
        x = torch.from_numpy(synth.random_vector(sb, 5)).cuda().to(dt)
        want = p.apply(x, precision=precision)
        got = torch.full_like(x, float("nan"))
        assert _noalloc(p, precision, x, got) == capi.TJSQ_OK, capi.lib().tjsq_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        # again into the same output, after other work on the device
        _ = (x * 3.0).sum().item()
        got2 = torch.zeros_like(x)
        assert _noalloc(p, precision, x, got2) == capi.TJSQ_OK
        torch.cuda.synchronize()
        assert torch.equal(got2, want)


def test_noalloc_apply_rejects_what_it_cannot_run():
    m = model(up_bias=1)
    sb = small_superblock(m, 7, 9, seed=62, msys=40, menv=44, median=5)
    p = capi.Plan(m, sb, precisions=("fp64", "bf16x3"))
    x = torch.from_numpy(synth.random_vector(sb, 5)).cuda()
    y = torch.empty_like(x)
    assert _noalloc(p, "bf16x3", x.float(), y.float()) == capi.TJSQ_ERR_ARG
    assert _noalloc(p, "fp32", x.float(), y.float()) == capi.TJSQ_ERR_ARG    # no FP32 operands
    p.profile(True)
    assert _noalloc(p, "fp64", x, y) == capi.TJSQ_ERR_ARG
    p.profile(False)
    assert _noalloc(p, "fp64", x, y) == capi.TJSQ_OK
    # a plan needing more scratch than the array holds
    big = small_superblock(m, 11, 11, seed=3, msys=900, menv=900, median=60, nsec=12)
    pb = capi.Plan(m, big, precisions=("fp64",))
    need = 8 * pb.stats["scratch_elems"] + 8 * 1024 * pb.stats["n_slots"]
    assert need > 64 * 1024 * 1024
    xb = torch.zeros(big.dim, dtype=torch.float64, device="cuda")
    assert _noalloc(pb, "fp64", xb, torch.empty_like(xb)) == capi.TJSQ_ERR_ARG
    assert b"static scratch holds" in capi.lib().tjsq_last_error()
