"""Work-item shape autotuner for the grouped H_eff GEMM.

For one plan it recompiles the task tables for each candidate shape
(tjsq_plan_reshape: the operators stay on the device), times applies with
HIP events recorded on the launching stream, and leaves the plan on the
fastest shape.  Results can be cached in a JSON file keyed by a workload
signature (the block structure, not the operator values), so later plans of
the same structure skip the search.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

# the default shape first: a candidate replaces it only when faster
CANDIDATES = [
    {"big": 64, "k_split": 256, "item_chunks": 32, "xcds": 1, "merge_right": 1},
    {"big": 64, "k_split": 512, "item_chunks": 32, "xcds": 1, "merge_right": 1},
    {"big": 64, "k_split": 128, "item_chunks": 16, "xcds": 1, "merge_right": 1},
    {"big": 64, "k_split": 1024, "item_chunks": 64, "xcds": 1, "merge_right": 1},
    {"big": 64, "k_split": 256, "item_chunks": 32, "xcds": 8, "merge_right": 1},
    {"big": 64, "k_split": 256, "item_chunks": 32, "xcds": 1, "merge_right": 0},
    {"big": 32, "k_split": 256, "item_chunks": 32, "xcds": 1, "merge_right": 1},
    {"big": 32, "k_split": 512, "item_chunks": 64, "xcds": 1, "merge_right": 1},
]


def signature(model, sb, precision: str) -> str:
    """Hash of everything the task tables depend on."""
    h = hashlib.sha256()
    for a in (sb.sys.hdr, sb.sys.sub, sb.env.hdr, sb.env.sub, sb.sys_bs.sec, sb.sys_bs.sub, sb.env_bs.sec,
              sb.env_bs.sub, sb.wsec):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr((model.ny, model.table_columns, model.up_bias, sb.sys_len, sb.env_len, precision)).encode())
    return h.hexdigest()[:24]


def time_applies(plan, precision: str, reps: int, torch) -> float:
    """Median milliseconds of one apply (events on the current stream)."""
    dt = torch.float64 if precision == "fp64" else torch.float32
    x = torch.randn(plan.dim, dtype=torch.float64, device="cuda").to(dt)
    y = torch.empty_like(x)
    plan.apply(x, y, precision=precision)
    st = torch.cuda.current_stream()
    out = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        plan.apply(x, y, precision=precision)
        b.record(st)
        b.synchronize()
        out.append(a.elapsed_time(b))
    return float(np.median(out))


def tune(plan, model, sb, precision: str = "fp32", reps: int = 5, candidates=None, cache: str | None = None):
    """Leave `plan` on its fastest candidate shape; returns
    {"shape", "ms", "results": [(shape, ms), ...], "cached"}."""
    import torch
    key = signature(model, sb, precision)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        best = table[key]
        plan.reshape(best["shape"])
        return {"shape": best["shape"], "ms": best["ms"], "results": [], "cached": True}
    results = []
    for shape in candidates or CANDIDATES:
        plan.reshape(shape)
        results.append((shape, time_applies(plan, precision, reps, torch)))
    shape, ms = min(results, key=lambda r: r[1])
    plan.reshape(shape)
    if cache:
        table[key] = {"shape": shape, "ms": ms}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1)
        os.replace(tmp, cache)
    return {"shape": shape, "ms": ms, "results": results, "cached": False}
