"""Serving loop for Lanczos ground-state solves with continuous admission.

A request is one lanczos_memo_new-style solve: the block models of one
superblock, a start vector and the solver options.  `slots` workers, each
with its own HIP stream (default priority), take admitted requests, compile
and upload the plan on their stream, solve, and return the energy, the state
and the timestamps.  A request is admitted when a worker is free and its
device footprint fits in the HBM budget left by the requests in flight, so
requests that arrive while others run join as slots and memory free up, and
their kernels overlap on the device.  Every solve is deterministic and
independent of the others: a result is bit-identical alone and in the mix.

The footprint is exact for the task tables (a host-only compile of the plan)
and counts the operators in FP64 and in the mixed precision, the stage-1
scratch and partial tiles of both, the Krylov basis and the work vectors.
"""
from __future__ import annotations

import collections
import threading
import time

import numpy as np

from . import capi


class Request:
    """device_ops: optional (sys, env) torch float64 CUDA tensors holding the
    block operators already on the device (the environments a sweep keeps
    resident); the plan then copies them device to device."""

    def __init__(self, rid, model, sb, psi0, precision="bf16x3", tol=1e-10, max_iter=40, energy_tol=0.0,
                 device_ops=None):
        self.id = rid
        self.model = model
        self.sb = sb
        self.device_ops = device_ops
        self.psi0 = np.ascontiguousarray(psi0, np.float64)
        self.precision = precision
        self.tol = tol
        self.max_iter = max_iter
        self.energy_tol = energy_tol
        self.footprint = None
        self.t_arrival = None


class Result:
    def __init__(self, req):
        self.id = req.id
        self.t_arrival = req.t_arrival
        self.footprint = req.footprint
        self.t_admit = self.t_start = self.t_plan = self.t_solved = self.t_done = None
        self.e0 = None
        self.psi = None
        self.info = None
        self.error = None
        self.slot = None
        self._ev = threading.Event()

    def wait(self, timeout=None):
        if not self._ev.wait(timeout):
            raise TimeoutError(f"request {self.id} not done")
        if self.error:
            raise RuntimeError(f"request {self.id} failed: {self.error}")
        return self

    @property
    def latency(self):
        return self.t_done - self.t_arrival


def footprint(model, sb, precision: str, max_iter: int) -> int:
    """Device bytes a request holds while it is served."""
    hp = capi.HostPlan(model, sb)
    D = sb.dim
    slots = hp.nslots * 1024
    tables = 40 * (hp.seg1.size + hp.seg2.size) + 48 * (hp.tile1.size + hp.tile2.size + hp.rtile.size)
    b = 8 * (hp.arena_elems + hp.scratch_elems + slots) + tables
    w = 8
    pk = hp.packed() if precision == "bf16x3" else None
    if pk is not None:   # the packed BF16x3 operands, products, partial tiles and tables
        b += (pk["ops_bytes"] + pk["x_bytes"] + pk["t_bytes"] + 4 * 4096 * pk["nslots"]
              + 48 * (pk["seg1"].size + pk["seg2"].size) + 56 * (pk["tile1"].size + pk["tile2"].size))
        w = 4
    elif precision != "fp64":
        b += 4 * (hp.arena_elems + hp.scratch_elems + slots)
        w = 4
    b += (max_iter + 2) * D * w + 6 * D * 8
    return int(b)


class Server:
    """Continuous-admission solve server (one HIP stream per slot)."""

    # two slots: on the serving scan (m 500-6000) one request already fills
    # the GPU, so 1, 2 and 4 slots serve at the same rate; two gave the
    # lowest p50 / p99 at 0.8 load (0.97 / 2.27 s against 1.12 / 2.36 s for
    # one and 1.42 / 2.71 s for four)
    def __init__(self, slots: int = 2, hbm_budget: int | None = None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("the solve server needs a GPU")
        self.torch = torch
        capi.lib()
        free, _total = torch.cuda.mem_get_info()
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * free)
        self.in_use = 0
        self.peak_in_use = 0
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.closed = False
        self.busy = 0
        self.busy_since = None
        self.busy_time = 0.0
        self.workers = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(slots)]
        for t in self.workers:
            t.start()

    def submit(self, req: Request) -> Result:
        if req.t_arrival is None:
            req.t_arrival = time.perf_counter()
        if req.footprint is None:
            req.footprint = footprint(req.model, req.sb, req.precision, req.max_iter)
        res = Result(req)
        if req.footprint > self.budget:
            res.error = f"footprint {req.footprint} B exceeds the HBM budget {self.budget} B"
            res.t_done = time.perf_counter()
            res._ev.set()
            return res
        with self.cv:
            self.pending.append((req, res))
            self.cv.notify_all()
        return res

    def _take(self):
        """Head of the queue once it fits in the budget (FIFO admission)."""
        with self.cv:
            while True:
                if self.closed and not self.pending:
                    return None
                if self.pending and self.in_use + self.pending[0][0].footprint <= self.budget:
                    req, res = self.pending.popleft()
                    self.in_use += req.footprint
                    self.peak_in_use = max(self.peak_in_use, self.in_use)
                    now = time.perf_counter()
                    if self.busy == 0:
                        self.busy_since = now
                    self.busy += 1
                    res.t_admit = now
                    self.cv.notify_all()
                    return req, res
                self.cv.wait()

    def _release(self, req):
        with self.cv:
            self.in_use -= req.footprint
            self.busy -= 1
            if self.busy == 0:
                self.busy_time += time.perf_counter() - self.busy_since
            self.cv.notify_all()

    def _worker(self, slot):
        torch = self.torch
        stream = torch.cuda.Stream()
        while True:
            item = self._take()
            if item is None:
                return
            req, res = item
            res.slot = slot
            try:
                with torch.cuda.stream(stream):
                    res.t_start = time.perf_counter()
                    plan = capi.Plan(req.model, req.sb, precisions=("fp64", req.precision),
                                     device_ops=req.device_ops)
                    res.t_plan = time.perf_counter()
                    psi = torch.from_numpy(req.psi0).to("cuda")
                    info = plan.lanczos(psi, tol=req.tol, max_iter=req.max_iter, precision=req.precision,
                                        refine_tol=1e-6, max_refine=1, energy_tol=req.energy_tol)
                    res.t_solved = time.perf_counter()
                    res.psi = psi.cpu().numpy()
                    res.e0 = info["e0"]
                    res.info = info
                    # the result is complete: publish it before the plan is
                    # freed (hipFree waits for the other slots' queued work);
                    # the request's HBM is returned to the budget after the free
                    res.t_done = time.perf_counter()
                    res._ev.set()
                    del plan, psi
                    stream.synchronize()
            except Exception as e:  # reported through the result
                if not res._ev.is_set():
                    res.error = repr(e)
            finally:
                self._release(req)
                if not res._ev.is_set():
                    res.t_done = time.perf_counter()
                    res._ev.set()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in self.workers:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def solve_alone(req: Request):
    """The same solve outside the server (the reference for the equality
    check): returns (e0, psi)."""
    import torch
    plan = capi.Plan(req.model, req.sb, precisions=("fp64", req.precision), device_ops=req.device_ops)
    psi = torch.from_numpy(req.psi0).to("cuda")
    info = plan.lanczos(psi, tol=req.tol, max_iter=req.max_iter, precision=req.precision, refine_tol=1e-6,
                        max_refine=1, energy_tol=req.energy_tol)
    return info["e0"], psi.cpu().numpy()
