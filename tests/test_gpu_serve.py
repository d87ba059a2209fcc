"""Solve server (continuous admission, one stream per slot) on the GPU."""
import time

import numpy as np
import pytest

from helpers import model, small_superblock
from grand_sql_tj_amd import capi, serve, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()


def _requests():
    reqs = []
    specs = [(1, (11, 11), 0.4, 160, 14, "bf16x3"), (0, (7, 9), 0.3, 40, 5, "fp64"),
             (1, (12, 10), 0.5, 200, 20, "fp32"), (1, (5, 5), 0.4, 24, 3, "bf16x3"),
             (0, (13, 17), 0.6, 120, 12, "fp64"), (1, (6, 5), 0.2, 60, 6, "fp32")]
    for i, (ub, geom, J, ms, med, prec) in enumerate(specs):
        m = model(up_bias=ub, J=J)
        sb = small_superblock(m, *geom, seed=40 + i, msys=ms, menv=ms, median=med, nsec=10)
        reqs.append(serve.Request(i, m, sb, synth.random_vector(sb, 3 + i), precision=prec, max_iter=40,
                                  energy_tol=1e-7 if prec != "fp64" else 0.0))
    return reqs


def test_results_same_alone_and_in_the_mix():
    reqs = _requests()
    alone = [serve.solve_alone(r) for r in reqs]
    with serve.Server(slots=3) as srv:
        futs = []
        for i, r in enumerate(reqs):
            futs.append(srv.submit(r))
            if i % 2:
                time.sleep(0.01)       # some arrive while others run
        res = [f.wait(120) for f in futs]
    assert len({r.slot for r in res}) > 1          # the work was spread over slots
    for (e0, psi), r in zip(alone, res):
        assert r.e0 == e0
        assert np.array_equal(r.psi, psi)


def test_admission_holds_the_budget():
    reqs = _requests()[:4]
    fp = [serve.footprint(r.model, r.sb, r.precision, r.max_iter) for r in reqs]
    with serve.Server(slots=4, hbm_budget=max(fp)) as srv:   # one request at a time fits
        futs = [srv.submit(r) for r in reqs]
        res = [f.wait(120) for f in futs]
        assert srv.peak_in_use <= srv.budget
    assert all(r.e0 is not None for r in res)
    # a request larger than the whole budget is refused, not queued forever
    with serve.Server(slots=1, hbm_budget=min(fp) // 2) as srv:
        r = srv.submit(reqs[0])
        with pytest.raises(RuntimeError, match="budget"):
            r.wait(10)
