"""Lanczos ground-state solves/s on one MI355X: mixed precision vs plain FP64.

Workload (one "step" = one Lanczos ground-state solve at one bond, the unit of
lanczos_memo_new): t-J model on a 6 x 16 cylinder (t = 1, t' = -0.2,
J = 0.4, J' = J (t'/t)^2, N = 84, doping 1/8, S = 0), centre bond
[47 sites | site | site | 47 sites], m = 8000 SU(2) multiplets per block.
The starting blocks are synthetic and symmetric (grand_sql_tj_amd.synth,
SplitMix64-seeded): sector dimensions log-normal (median 40, sigma 0.8,
max 600), reduced operators with the reference's selection rules and
hermiticity, so H_eff is symmetric.  Inputs are resident in HBM before the
timed region.  The bond is reached the way a sweep reaches it: from those
blocks, one right-to-left step at [47'|s|s|47] makes the 48-site
environment (density-matrix truncation to m) and one left-to-right step at
[46|s|s|48] makes the 47-site system block; every timed solve starts from
wave_trans_LR of that step's ground state, the vector lanczos_memo_new
receives in a sweep (dmrg_sub.f90:845).

The printed JSON line carries solves/s of the mixed path (value) and its
ratio to the FP64 path at the same stopping rule (speedup_vs_fp64, with the
dE/site of both against a converged FP64 solve), the FP64 rate at the
reference's own stopping rule beside it, p50/p99 latency, the roofline of the
dominant kernel (HIP events on the launching stream) and the CPU baseline:
the reference's own model_wave (compiled Fortran + MKL, one core) timed on a
bounded sample of applies of the same bond.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 3.0}
HBM_PEAK = 8.0e12
# kernel kinds of Plan.profile_read, and the kernel symbol of each
KERNEL_KINDS = ("stage1_64", "stage1_32", "stage2_64", "stage2_32", "reduce")
KERNEL_SYMBOL = {"stage1_64": "k_tiles<{mode}, 64, 1>", "stage1_32": "k_tiles<{mode}, 32, 1>",
                 "stage2_64": "k_tiles<{mode}, 64, 2>", "stage2_32": "k_tiles<{mode}, 32, 2>", "reduce": "k_reduce"}
# the packed BF16x3 apply: its kernels in the same profiling slots (slot 1 =
# the per-apply packing of x, slot 3 unused; stage 1 runs in quads of tiles
# that share their x lines)
PACKED_SYMBOL = {"stage1_64": "k_packed_group<4>", "stage1_32": "k_pack<float>", "stage2_64": "k_packed<2>",
                 "stage2_32": "-", "reduce": "k_preduce"}
MODE_ID = {"fp64": 0, "fp32": 1, "bf16x3": 2}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "heff_apply_m8000_traffic.json")


def kernel_flops(hp):
    """Useful flops per launch of each kernel kind: 2 mt nt K summed over the
    GEMM segments of its work items."""
    out = {}
    for stage, tiles, segs, n64 in ((1, hp.tile1, hp.seg1, hp.n64_1), (2, hp.tile2, hp.seg2, hp.n64_2)):
        kk = np.concatenate([[0.0], np.cumsum(segs["K"].astype(np.float64) * (segs["r_space"] != 3))])
        per = 2.0 * tiles["mt"].astype(np.float64) * tiles["nt"] * (kk[tiles["seg_end"]] - kk[tiles["seg_begin"]])
        out[f"stage{stage}_64"] = float(per[:n64].sum())
        out[f"stage{stage}_32"] = float(per[n64:].sum())
    out["reduce"] = 0.0
    return out


def measured_traffic(kernel, precision, dim):
    """HBM bytes per launch of `kernel` at this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (calibrated by a copy of
    known size); None when that measurement is for another workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("D") != dim:
        return None
    return t.get("bytes_per_launch", {}).get(precision, {}).get(kernel)


BLOCK_SHAPE = dict(median=40, sigma=0.8, maxdim=600, nsec=140, smax=12, filling=0.875)


def workload(m_mult: int, seed: int):
    from grand_sql_tj_amd import synth
    from grand_sql_tj_amd.lattice import TJModel
    model = TJModel(ny=6, nx=16, t=1.0, tp=-0.2, J=0.4, mu=0.0, up_bias=1)
    sb = synth.superblock(model, 47, 47, m_mult, m_mult, seed=seed, total_up=84, **BLOCK_SHAPE)
    return model, sb


def swept_bond(m_mult: int, seed: int, max_iter: int = 250):
    """The centre bond [47 | s | s | 47] as a left-to-right sweep reaches it,
    with its start vector from wave_trans_LR of the bond before.

    From synthetic blocks S47', E47 (workload()) and S46 (synthetic, same
    sector statistics), on the device (grand_sql_tj_amd.sweep.DMRG):

      right_to_left at [S47' | s | s | E47]: ground state (FP64, random start),
          env density matrix -> truncation (m kept) -> E48 = Get_Hamiltonian_Env
      left_to_right at [S46 | s | s | E48]: ground state psi_A (FP64, random
          start), sys density matrix -> truncation (m kept) -> S47
      start vector of [S47 | s | s | E47] = wave_trans_LR(psi_A) with the S47
          truncation and the E48 truncation, as left_to_right passes it on
          (dmrg_sub.f90:845, 1093-1384)

    Returns (model, sb, psi0, setup): sb holds S47 and E47 on the host, psi0
    the transformed vector (CUDA float64), setup the two solves' records."""
    import torch
    from grand_sql_tj_amd import synth
    from grand_sql_tj_amd.blocks import Superblock, get_super_basis
    from grand_sql_tj_amd.sweep import DMRG
    model, sb0 = workload(m_mult, seed)
    s46 = synth.block(model, 46, m_mult, seed + 46, **BLOCK_SHAPE)
    total_up = int(sb0.meta["total_up"])
    d = DMRG(model, total_up=total_up, kept_max=m_mult, kept_min=m_mult, trun_err=1e-8, tol=1e-10,
             max_iter=max_iter, precision="fp64", seed=seed)
    d.sys[47], d.env[47], d.sys[46] = d._block(sb0.sys), d._block(sb0.env), d._block(s46)
    d.right_to_left(48, 47)     # one step, bond 47|s|s|47: makes E48
    d.left_to_right(46, 47)     # one step, bond 46|s|s|48: makes S47
    sys_bs, env_bs = d._basis(d.sys[47]), d._basis(d.env[47])
    wsec = get_super_basis(sys_bs, env_bs, total_up, 0)
    psi0 = d._transform(0, sys_bs, d._basis(d.env[48]), d.sys_trun[47], d.env_trun[48], wsec).clone()
    psi0 /= psi0.norm()
    sb = Superblock(47, 47, d._host(d.sys[47]), sb0.env, sys_bs, env_bs, wsec,
                    meta={"seed": seed, "total_up": total_up, "model": model})
    setup = [{"phase": r.phase, "sys_len": r.sys_len, "env_len": r.env_len, "D": r.dim, "e0": r.energy,
              "iterations": r.iterations, "kept": r.kept, "trunc_err": r.trunc_err} for r in d.log]
    del d
    torch.cuda.synchronize()
    return model, sb, psi0, setup


def timed_solves(plan, x0_t, steps, warmup, precision, tol, max_iter, torch, energy_tol=0.0, rtol=0.0, stream=None,
                 max_refine=1):
    """Wall time of each solve (the solve synchronises its stream at the end),
    the last solve's result and its ground state.  On a side stream with the
    plan's graphs enabled, each Lanczos iteration is replayed from its
    captured hipGraph (captured during the warm-up)."""
    lat = []
    res = None
    psi = None
    ctx = torch.cuda.stream(stream) if stream is not None else None
    if ctx is not None:
        ctx.__enter__()
    try:
        for i in range(warmup + steps):
            psi = x0_t.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = plan.lanczos(psi, tol=tol, max_iter=max_iter, precision=precision, refine_tol=1e-6, max_refine=max_refine,
                               energy_tol=energy_tol, rtol=rtol)
            torch.cuda.synchronize()
            if i >= warmup:
                lat.append(time.perf_counter() - t0)
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    return np.array(lat), res, psi


def block_update_step(model, sb, psi, m_keep, torch, reps=5):
    """The step after each solve on the device: Get_density_matrix_sys ->
    Get_truncation_operator (kept_min = m/2, kept_max = m, trun_err 1e-8) ->
    Get_Hamiltonian_Sys of the system block at this bond -> wave_trans_LR
    of the ground state to the next bond.  Times each part with events on
    the launching stream (the current torch stream)."""
    from grand_sql_tj_amd import capi
    st = torch.cuda.current_stream()

    def timed(fn, n):
        out = []
        for _ in range(n):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            fn()
            b.record(st)
            torch.cuda.synchronize()
            out.append(a.elapsed_time(b))
        return float(np.median(out))

    pgd = capi.Program.density(0, sb.sys_len + 1, sb.wsec, 1)
    _, dsub, doff = pgd.layout()
    rho = torch.empty(pgd.sizes["ndata"], dtype=torch.float64, device="cuda")
    pgd.run(None, psi, rho)
    ms_density = timed(lambda: pgd.run(None, psi, rho), reps)
    rho0 = rho.clone()
    res = {}

    def trunc():
        rho.copy_(rho0)
        res["t"] = capi.truncation_operator(dsub, doff, rho, m_keep // 2, m_keep, 1e-8)
    # the copy back is a small part of the timed region
    ms_trunc = timed(trunc, 3)
    tsub, toff, _, info = res["t"]
    from grand_sql_tj_amd.blocks import OpSet
    trun = OpSet(sb.sys_len + 1, np.array([[sb.sys_len + 1, 0, 0, tsub.shape[0]]], np.int32), tsub, toff,
                 np.zeros(rho.numel()))
    t0 = time.perf_counter()
    pgu = capi.Program.block_update(model, 0, sb.sys_len, sb.env_len, sb.sys, sb.sys_bs, trun)
    build_s = time.perf_counter() - t0
    pre = torch.from_numpy(sb.sys.data).cuda()
    out = torch.empty(pgu.sizes["ndata"], dtype=torch.float64, device="cuda")
    pgu.run(rho, pre, out)
    ms_update = timed(lambda: pgu.run(rho, pre, out), reps)
    fl = pgu.stats()["flops_exec"]
    # wave_trans_LR to the next bond: the system truncation just computed,
    # an orthogonal env truncation per env sector (square: the synthetic
    # environment's previous block has the same sector dims), psi as input
    from grand_sql_tj_amd.blocks import Basis, get_basis, get_super_basis, site_sectors
    kept = np.stack([tsub[:, 0], tsub[:, 1], tsub[:, 4]], axis=1).astype(np.int32)
    sys_new_bs = get_basis(sb.sys_len + 1, kept, site_sectors(model.up_bias))
    env_secs = sb.env.sectors()
    env_new = Basis(sb.env_len - 1, np.column_stack([env_secs, np.zeros(len(env_secs), np.int32)]).astype(np.int32),
                    np.zeros((0, 6), np.int32))
    wsec_new = get_super_basis(sys_new_bs, env_new, int(sb.meta["total_up"]), 0)
    rng = np.random.default_rng(5)
    e_sub, e_off, e_dat, nd = [], [], [], 0
    for u, d, n in env_secs:
        q, _ = np.linalg.qr(rng.normal(size=(int(n), int(n))))
        e_sub.append((u, d, 0, n, n))
        e_off.append(nd)
        e_dat.append(q.T.reshape(-1))
        nd += int(n) * int(n)
    env_trun = OpSet(sb.env_len, np.array([[sb.env_len, 0, 0, len(e_sub)]], np.int32), np.array(e_sub, np.int32),
                     np.array(e_off, np.int64), np.concatenate(e_dat))
    nt = int((toff + tsub[:, 3].astype(np.int64) * tsub[:, 4]).max()) if tsub.shape[0] else 0
    sys_trun = OpSet(sb.sys_len + 1, trun.hdr, tsub, toff, np.zeros(nt))
    pgw = capi.Program.wave_transform(0, sys_new_bs, sb.env_bs, sys_trun, env_trun, sb.wsec, wsec_new)
    wops = torch.cat([rho[:nt], torch.from_numpy(env_trun.data).cuda()])
    psi_new = torch.empty(max(1, pgw.sizes["ndata"]), dtype=torch.float64, device="cuda")
    pgw.run(wops, psi, psi_new)
    ms_wave = timed(lambda: pgw.run(wops, psi, psi_new), reps)
    wfl = pgw.stats()["flops_exec"]
    return {"density_ms": ms_density, "truncation_ms": ms_trunc, "update_ms": ms_update,
            "wave_transform_ms": ms_wave, "wave_transform_flops": wfl, "wave_new_dim": pgw.sizes["ndata"],
            "wave_transform_tflops": wfl / (ms_wave * 1e-3) / 1e12,
            "update_build_s": build_s, "kept": info["num_state"], "rho_dim": info["dim"],
            "trunc_err": info["trunc_err"], "update_flops": fl,
            "update_tflops": fl / (ms_update * 1e-3) / 1e12, "update_frac_fp64_peak": fl / (ms_update * 1e-3) / PEAK["fp64"],
            "density_flops": pgd.stats()["flops_exec"], "new_ops": pgu.sizes["nops"], "new_elems": pgu.sizes["ndata"]}


def batch_member(args):
    """One member of the batched-apply workload: a bond in the middle of a
    width-4 (4 x 16) or width-6 (6 x 12) cylinder, m multiplets per block,
    sector dims log-normal (median 32, max 512).  args = (i, seed) or
    (i, seed, fast): fast draws the operator entries with FastNormals."""
    i, seed = args[0], args[1]
    fast = len(args) > 2 and bool(args[2])
    from grand_sql_tj_amd import synth
    from grand_sql_tj_amd.lattice import TJModel
    r = np.random.default_rng(seed + 7919 * i)
    ny = int(r.choice([4, 6]))
    nx = 16 if ny == 4 else 12
    n = ny * nx
    m = int(r.integers(1000, 6001))
    sys_len = int(r.integers(n // 2 - n // 6, n // 2 + n // 6))
    env_len = n - 2 - sys_len
    model = TJModel(ny=ny, nx=nx, t=1.0, tp=float(r.choice([0.0, -0.2])), J=float(r.uniform(0.2, 0.6)), mu=0.0,
                    up_bias=1)
    nsec = int(r.integers(4, 9))
    sb = synth.superblock(model, sys_len, env_len, m, m, seed=seed + i, median=32, sigma=0.8, maxdim=512,
                          nsec=nsec, smax=10, filling=0.875, fast=fast)
    return model, sb, {"ny": ny, "m": m, "sys_len": sys_len, "env_len": env_len}


def batch_bench(args):
    """Batched grouped launch of --batch independent FP32 applies."""
    import multiprocessing as mp
    t0 = time.perf_counter()
    # synthesise on host worker processes before the GPU is initialised
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        members = []
        for k, res in enumerate(pool.imap(batch_member, [(i, args.seed) for i in range(args.batch)])):
            members.append(res)
            print(f"[bench] member {k}: {res[2]} D={res[1].dim}", file=sys.stderr, flush=True)
    synth_s = time.perf_counter() - t0
    import torch
    from grand_sql_tj_amd import capi, autotune
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    prec = args.batch_precision
    plans, xs, tuned = [], [], []
    for k, (model, sb, meta) in enumerate(members):
        p = capi.Plan(model, sb, precisions=(prec,))
        if not args.no_tune:
            tuned.append(autotune.tune(p, model, sb, precision=prec, reps=3)["shape"])
        plans.append(p)
        xs.append(torch.randn(sb.dim, device="cuda", dtype=torch.float32))
        print(f"[bench] plan {k} contributions={p.stats['n_contrib']} shape={tuned[-1] if tuned else None}",
              file=sys.stderr, flush=True)
    del members
    st = torch.cuda.current_stream()

    def timed(fn, reps):
        out = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            fn()
            b.record(st)
            b.synchronize()
            out.append(a.elapsed_time(b))
        return float(np.median(out))

    ys = [torch.empty_like(x) for x in xs]

    def seq():
        for p, x, y in zip(plans, xs, ys):
            p.apply(x, y, precision=prec)
    seq()
    ms_seq = timed(seq, args.steps)
    b = capi.Batch(plans, precisions=(prec,))
    del plans, p            # the batch holds its own operand copies
    x = torch.cat(xs)
    y = torch.empty_like(x)
    b.apply(x, y, precision=prec)
    ms_batch = timed(lambda: b.apply(x, y, precision=prec), args.steps)
    same = all(torch.equal(y[int(b.offsets[i]):int(b.offsets[i + 1])], ys[i]) for i in range(len(xs)))
    b.graph(True)
    b.apply(x, y, precision=prec)
    ms_graph = timed(lambda: b.apply(x, y, precision=prec), args.steps)
    b.graph(False)
    b.profile(True)
    for _ in range(5):
        b.apply(x, y, precision=prec)
    torch.cuda.synchronize()
    ms_k, n_k = b.profile_read()
    b.profile(False)
    stb = b.stats
    fl = stb["flops_exec"]
    best = min(ms_batch, ms_graph)
    line = {
        "metric": "batched H_eff applies/s (grouped launch per stage)", "value": len(xs) / (best * 1e-3),
        "unit": "applies/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": best,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": {"fp32": "f32",
                                                                                   "bf16x3": "bf16x3->f32"}[prec],
        "data": "synthetic",
        "config": {"workload": f"{len(xs)} independent H_eff applies, bonds of width-4/6 t-J cylinders, "
                               "m ~ U[1000, 6000], sector dims log-normal(median 32, max 512)",
                   "D_total": b.dim, "contributions": stb["n_contrib"], "precision": prec},
        "sequential_ms": ms_seq, "batch_ms": ms_batch, "batch_graph_ms": ms_graph,
        "speedup_vs_sequential": ms_seq / best, "bit_identical_to_members": bool(same),
        "kernel_ms": {k: float(ms_k[i] / max(1, n_k[i])) for i, k in enumerate(KERNEL_KINDS)},
        "roofline": {"bound": "mfma", "achieved": fl / (best * 1e-3) / 1e12, "peak": PEAK[prec] / 1e12,
                     "unit": "TFLOP/s", "frac": fl / (best * 1e-3) / PEAK[prec], "traffic": None},
        "autotuned_shapes": tuned, "synthesis_s": synth_s,
    }
    print(json.dumps(line), flush=True)


def serve_request(args):
    """One request of the serving scan: the centre bond of a width-4/6,
    length 8-16 t-J cylinder, J/t ~ U[0.2, 0.6], t'/t in {0, -0.2}, doping in
    {1/12, 1/8, 1/6}, m ~ U[500, 6000].  args = (i, seed) or (i, seed, fast)
    (fast: operator entries drawn with FastNormals)."""
    i, seed = args[0], args[1]
    fast = len(args) > 2 and bool(args[2])
    from grand_sql_tj_amd import synth
    from grand_sql_tj_amd.lattice import TJModel
    r = np.random.default_rng(seed + 104729 * i)
    ny = int(r.choice([4, 6]))
    nx = int(r.integers(8, 17))
    n = ny * nx
    dop = float(r.choice([1 / 12, 1 / 8, 1 / 6]))
    ne = int(round(n * (1 - dop)))
    ne -= ne % 2
    m = int(r.integers(500, 6001))
    model = TJModel(ny=ny, nx=nx, t=1.0, tp=float(r.choice([0.0, -0.2])), J=float(r.uniform(0.2, 0.6)), mu=0.0,
                    up_bias=1)
    sys_len = n // 2 - 1
    env_len = n - 2 - sys_len
    sb = synth.superblock(model, sys_len, env_len, m, m, seed=seed + i, median=32, sigma=0.8, maxdim=512,
                          nsec=int(np.clip(m // 400, 4, 14)), smax=10, filling=ne / n, total_up=ne, fast=fast)
    psi0 = synth.random_vector(sb, seed + i)
    return model, sb, psi0, {"ny": ny, "nx": nx, "m": m, "N": ne, "J": model.J, "D": sb.dim}


def serve_bench(args):
    """Poisson arrivals of solve requests into the continuous-admission server."""
    import multiprocessing as mp
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        reqs = []
        for k, res in enumerate(pool.imap(serve_request, [(i, args.seed) for i in range(args.serve + 1)])):
            reqs.append(res)
            print(f"[bench] request {k}: {res[3]}", file=sys.stderr, flush=True)
    synth_s = time.perf_counter() - t0
    import torch
    from grand_sql_tj_amd import capi, serve
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    prec = args.batch_precision if args.precision == "fp64" else args.precision
    # the environments are device-resident before a request arrives, as in a
    # sweep where the previous block update leaves them in HBM
    dev_ops = [(torch.from_numpy(sb.sys.data).cuda(), torch.from_numpy(sb.env.data).cuda()) for _, sb, _, _ in reqs]
    torch.cuda.synchronize()

    def make(k, rid):
        model, sb, psi0, meta = reqs[k]
        nsite = sb.sys_len + sb.env_len + 2
        return serve.Request(rid, model, sb, psi0, precision=prec, tol=args.tol, max_iter=args.max_iter,
                             energy_tol=0.1 * 1e-8 * nsite, device_ops=dev_ops[k])
    # warm-up (library, solver handles, allocator), then every request alone,
    # one at a time: the one-at-a-time capacity that sets the arrival rate and
    # the reference results of the equality check
    serve.solve_alone(make(args.serve, -1))
    alone, cal = [], []
    for k in range(args.serve):
        t1 = time.perf_counter()
        alone.append(serve.solve_alone(make(k, k)))
        torch.cuda.synchronize()
        cal.append(time.perf_counter() - t1)
    s_alone = float(np.mean(cal))
    lam = args.load / s_alone
    rng = np.random.default_rng(args.seed)
    gaps = rng.exponential(1.0 / lam, size=args.serve)
    print(f"[bench] serving {args.serve} requests, {args.slots} slots, lambda={lam:.2f}/s "
          f"(alone {s_alone * 1e3:.0f} ms)", file=sys.stderr, flush=True)
    results = []
    with serve.Server(slots=args.slots) as srv:
        t_begin = time.perf_counter()
        t_next = t_begin
        for k in range(args.serve):
            t_next += gaps[k]
            while time.perf_counter() < t_next:
                time.sleep(min(0.002, max(0.0, t_next - time.perf_counter())))
            rq = make(k, k)
            rq.t_arrival = time.perf_counter()
            results.append(srv.submit(rq))
        for r in results:
            r.wait(600)
            print(f"[bench] done {r.id} slot={r.slot} latency={r.latency * 1e3:.0f} ms", file=sys.stderr, flush=True)
        t_end = max(r.t_done for r in results)
    # (read after close(): a request's result is published before its plan
    # is freed, and its busy interval ends with the free)
    busy = min(srv.busy_time, t_end - t_begin)
    peak = srv.peak_in_use
    budget = srv.budget
    lat = np.array([r.latency for r in results])
    wait = np.array([r.t_admit - r.t_arrival for r in results])
    t_plan = np.array([r.t_plan - r.t_start for r in results])
    t_solve = np.array([r.t_solved - r.t_plan for r in results])
    # every request against its result alone: must be bit-identical
    same = all(alone[k][0] == results[k].e0 and bool(np.array_equal(alone[k][1], results[k].psi))
               for k in range(args.serve))
    check = list(range(args.serve))
    wall = t_end - t_begin
    line = {
        "metric": "serving: solves/s with p50/p99 latency (Poisson arrivals, continuous admission)",
        "value": args.serve / wall, "unit": "solves/s", "n_gpus": 1, "steps": args.serve, "warmup": 1,
        "ms_per_step": 1e3 * wall / args.serve, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16x3": "bf16x3->f32"}[prec] + "+f64", "data": "synthetic",
        "config": {"workload": "ground-state solves at the centre bond of width-4/6, length 8-16 t-J cylinders; "
                               "J/t~U[0.2,0.6], t'/t in {0,-0.2}, doping in {1/12,1/8,1/6}, m~U[500,6000]",
                   "slots": args.slots, "arrival_rate_per_s": lam, "precision": prec},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "mean_queue_wait_ms": 1e3 * float(wait.mean()), "mean_plan_ms": 1e3 * float(t_plan.mean()),
        "mean_solve_ms": 1e3 * float(t_solve.mean()), "alone_service_ms": 1e3 * s_alone,
        "busy_fraction": busy / wall, "offered_load_vs_alone_rate": args.load,
        "alone_solves_per_s": 1.0 / s_alone, "p50_ms_alone": 1e3 * float(np.percentile(cal, 50)),
        "p99_ms_alone": 1e3 * float(np.percentile(cal, 99)),
        "peak_admitted_bytes": peak, "hbm_budget_bytes": budget,
        "bit_identical_alone_and_in_mix": same, "requests_checked_alone": len(check), "synthesis_s": synth_s,
    }
    print(json.dumps(line), flush=True)


def sweep_bench(args):
    """Whole ground-state runs on the device (grand_sql_tj_amd.sweep): the
    reference's warm-up and `--sweeps` pairs of finite half sweeps on a t-J
    cylinder (t = 1, J = 0.4, t' = -0.2 unless --sweep-tp), every bond's
    plan, solve, truncation, block update and wave transform through the
    library.  Reports bonds per second over the whole run, the time per
    phase, the energy at the end of every sweep, and the time of the final
    state's density profile (DMRG.measure_density)."""
    import torch
    from grand_sql_tj_amd import sweep
    from grand_sql_tj_amd.lattice import TJModel
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    ny, nx, n_up, m = (int(v) for v in args.sweep_shape.split(","))
    model = TJModel(ny=ny, nx=nx, t=1.0, tp=args.sweep_tp, J=0.4, up_bias=1)

    def run():
        d = sweep.DMRG(model, total_up=n_up, kept_max=m, trun_err=1e-9, tol=1e-9, precision=args.precision)
        t0 = time.perf_counter()
        e = d.run(sweeps=args.sweeps)
        return d, e, time.perf_counter() - t0

    run_small = sweep.DMRG(TJModel(ny=2, nx=4, t=1.0, J=0.4, up_bias=1), total_up=6, kept_max=64)
    run_small.run(sweeps=1)          # loads the library and the solver handles
    d, e, wall = run()
    bonds = len(d.log)
    # the density profile of the final state (measure_operator_dia), timed
    # after the run: every site operator carried through the run's
    # truncations on the device, then contracted with the ground state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dens = d.measure_density()
    torch.cuda.synchronize()
    t_meas = time.perf_counter() - t0
    t0 = time.perf_counter()
    cor = d.measure_cross_correlation()
    torch.cuda.synchronize()
    t_cor = time.perf_counter() - t0
    t0 = time.perf_counter()
    full = d.measure_correlations()
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0
    ends = [r.energy for i, r in enumerate(d.log)
            if r.phase == "right-to-left" and (i + 1 == bonds or d.log[i + 1].phase != "right-to-left")]
    line = {
        "metric": "device DMRG: bonds/s over warm-up + sweeps (plan, solve, truncation, block update, transform)",
        "value": bonds / wall, "unit": "bonds/s", "n_gpus": 1, "steps": bonds, "warmup": 1,
        "ms_per_step": 1e3 * wall / bonds, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16x3": "bf16x3->f32+f64"}[args.precision], "data": "synthetic",
        "config": {"workload": f"t-J {ny}x{nx} cylinder (t=1, t'={args.sweep_tp}, J=0.4), N={n_up}, kept_max={m}, "
                               f"trun_err 1e-9, Lanczos tol 1e-9, warm-up + {args.sweeps} sweep(s)",
                   "precision": args.precision},
        "energy": e, "energy_per_site": e / (ny * nx), "energy_after_each_sweep": ends, "wall_s": wall,
        "phase_s": {k: round(v, 4) for k, v in d.phase_s.items()},
        "host_other_s": wall - sum(d.phase_s.values()),
        "max_superblock_dim": max(r.dim for r in d.log), "max_kept": max(r.kept for r in d.log),
        "last_trunc_err": d.log[-1].trunc_err,
        "density_profile_s": t_meas, "density_sum": float(dens.sum()),
        "density_min_max": [float(dens.min()), float(dens.max())],
        "cross_correlation_s": t_cor, "cross_correlation_shape": list(cor.shape),
        "correlation_matrix_s": t_full, "correlation_matrix_sum": float(full.sum()),
    }
    print(json.dumps(line), flush=True)


C1 = dict(ny=4, nx=8, t=1.0, tp=0.0, J=0.4, electrons=28, kept_max=500, kept_maxf=500, sweeps=6, lanerr=1e-10,
          trun_err=1e-18)


def c1_bench(args):
    """C1 on the device: t-J 4 x 8 cylinder, t = 1, J = 0.4, 28 electrons,
    S = 0, kept_max = kept_maxf = 500, 6 sweeps with Get_Ground_State's
    schedule (sweep.DMRG.ground_state), FP64.  Reports the wall time and the
    energies beside those of the reference's own C1 run (its routines built
    for the 4 x 8 lattice, tests/golden/c1_reference.json, timed on one core
    where that fixture was written)."""
    import torch
    from grand_sql_tj_amd import sweep
    from grand_sql_tj_amd.lattice import TJModel
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    P = C1
    model = TJModel(ny=P["ny"], nx=P["nx"], t=P["t"], tp=P["tp"], J=P["J"], mu=0.0, up_bias=1)
    warm = sweep.DMRG(TJModel(ny=2, nx=4, t=1.0, J=0.4, up_bias=1), total_up=6, kept_max=64)
    warm.run(sweeps=1)               # loads the library and the solver handles
    d = sweep.DMRG(model, total_up=P["electrons"], kept_max=P["kept_max"], max_iter=110, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = d.ground_state(P["lanerr"], P["trun_err"], P["kept_max"], P["kept_maxf"], P["sweeps"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ref = None
    path = os.path.join(ROOT, "tests", "golden", "c1_reference.json")
    if os.path.exists(path):
        with open(path) as f:
            ref = json.load(f)
    energies = [r.energy for r in d.log]
    line = {
        "metric": "C1 ground-state run on the device (warm-up + 6 sweeps + closing pass, Get_Ground_State schedule)",
        "value": len(d.log) / wall, "unit": "bonds/s", "n_gpus": 1, "steps": len(d.log), "warmup": 1,
        "ms_per_step": 1e3 * wall / len(d.log), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C1: t-J 4x8 cylinder (y periodic, x open), t=1, J=0.4, N=28, S=0, kept_max = "
                               "kept_maxf = 500, 6 sweeps, lanerr 1e-10, trun_err 1e-18"},
        "wall_s": wall, "final_e0": e, "min_e0": min(energies), "solves": len(d.log),
        "phase_s": {k: round(v, 4) for k, v in d.phase_s.items()},
        "reference_final_e0": ref["final_e0"] if ref else None, "reference_min_e0": ref["min_e0"] if ref else None,
        "reference_solves": len(ref["steps"]) if ref else None,
        "reference_seconds_1_core": ref["reference_seconds_1_core"] if ref else None,
        "rel_diff_final": abs(e - ref["final_e0"]) / abs(ref["final_e0"]) if ref else None,
        "max_superblock_dim": max(r.dim for r in d.log),
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(model, sb, x0, n_iter, threads=1, budget_s=20.0):
    """The reference's own model_wave (hmerge + 14 kernels, compiled from its
    sources) timed on this host on a bounded sample: a few applies of the same
    workload, scaled by the applies of one FP64 solve.  The reference builds
    with threaded MKL (code/tjsq/makefile:3, -qmkl=parallel) and has no
    OpenMP of its own; its DGEMMs are small, and on the MI355X box's host one
    MKL thread was fastest (an m=2000 apply: 0.136 s on 1 thread, 0.155 /
    0.316 / 0.330 / 0.371 s on 2 / 4 / 8 / 16), so one thread is the default."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.ref_available():
        return None
    cwd = os.getcwd()
    os.chdir(os.environ.get("TMPDIR", "/tmp"))
    try:
        oracle.ref_threads(threads)
        times = []
        while sum(times) < budget_s and len(times) < 3:
            t0 = time.perf_counter()
            oracle.ref_model_wave(model, sb, x0)
            times.append(time.perf_counter() - t0)
    finally:
        oracle.ref_threads(1)
        os.chdir(cwd)
    t_apply = float(np.median(times))
    return {"value": 1.0 / (t_apply * n_iter), "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": f"{len(times)} reference model_wave applies (hmerge + 14 kernels, MKL on {threads} thread(s)) "
                      f"of this workload, {t_apply:.3f} s each, scaled by the {n_iter} applies of one FP64 solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=8000)
    ap.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--max-iter", type=int, default=110, help="Lanczos iterations per pass (eigen.f90 loop_in = 110)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=1, help="MKL threads of the reference CPU baseline")
    ap.add_argument("--no-update", action="store_true", help="skip the density/truncation/update step")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the solve benchmark: N independent applies in one grouped launch")
    ap.add_argument("--no-tune", action="store_true", help="batch mode: keep the default work-item shape")
    ap.add_argument("--tune", action="store_true", help="autotune the work-item shape of the solve plan first")
    ap.add_argument("--no-graphs", dest="graphs", action="store_false",
                    help="launch every Lanczos iteration instead of replaying its captured hipGraph")
    ap.add_argument("--batch-precision", default="fp32", choices=["fp32", "bf16x3"])
    ap.add_argument("--serve", type=int, default=0,
                    help="instead of the solve benchmark: N solve requests with Poisson arrivals")
    ap.add_argument("--slots", type=int, default=2, help="--serve: solve slots (HIP streams)")
    ap.add_argument("--load", type=float, default=0.8,
                    help="arrival rate as a fraction of the one-at-a-time service rate")
    ap.add_argument("--sweep", action="store_true",
                    help="instead of the solve benchmark: a whole device DMRG run (warm-up + one sweep)")
    ap.add_argument("--sweep-shape", default="4,12,44,1500", help="ny,nx,N,kept_max of the --sweep run")
    ap.add_argument("--sweeps", type=int, default=1, help="finite sweeps of the --sweep run after the warm-up")
    ap.add_argument("--sweep-tp", type=float, default=-0.2, help="t'/t of the --sweep run")
    ap.add_argument("--c1", action="store_true",
                    help="instead of the solve benchmark: the C1 ground-state run (4x8, N=28, m=500, 6 sweeps, "
                         "Get_Ground_State schedule) on the device, against the reference run's energies")
    ap.add_argument("--apply-only", type=int, default=0,
                    help="profiling mode: a 512 MiB device copy (counter calibration) then N isolated applies")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solve is not sharded across GPUs in this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.batch:
        batch_bench(args)
        return
    if args.serve:
        serve_bench(args)
        return
    if args.sweep:
        sweep_bench(args)
        return
    if args.c1:
        c1_bench(args)
        return
    import torch
    from grand_sql_tj_amd import capi
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    print(f"[bench] building the bond m={args.m}: synthetic blocks, one right-to-left and one left-to-right step",
          file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    model, sb, psi0, setup = swept_bond(args.m, args.seed)
    t_setup = time.perf_counter() - t_setup
    for r in setup:
        print(f"[bench] setup {r}", file=sys.stderr, flush=True)
    from grand_sql_tj_amd import synth
    if args.apply_only:
        plan = capi.Plan(model, sb)
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")   # 512 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        x = psi0.clone()
        if args.precision != "fp64":
            x = x.float()
        y = torch.empty_like(x)
        for _ in range(args.apply_only):
            plan.apply(x, y, precision=args.precision)
        torch.cuda.synchronize()
        print(json.dumps({"apply_only": args.apply_only, "precision": args.precision, "D": sb.dim,
                          "stats": plan.stats, "copy_bytes": 512 * 1024 * 1024}), flush=True)
        return
    plan = capi.Plan(model, sb)
    tuned = None
    if args.tune:
        from grand_sql_tj_amd import autotune
        tuned = autotune.tune(plan, model, sb, precision=args.precision, reps=5)
        for shp, ms in tuned["results"]:
            print(f"[bench] shape {shp}: {ms:.3f} ms/apply", file=sys.stderr, flush=True)
    st = plan.stats
    nsite = sb.sys_len + sb.env_len + 2
    print(f"[bench] D={sb.dim} contributions={st['n_contrib']} F={st['flops_ref']:.3e}", file=sys.stderr, flush=True)
    # the energy every dE is measured against: an untimed FP64 solve from the
    # same start vector to a 1e-12 change of the Ritz value
    g = psi0.clone()
    conv = plan.lanczos(g, tol=1e-12, max_iter=250, precision="fp64")
    x0_t = psi0
    x0 = x0_t.cpu().numpy()
    print(f"[bench] converged E={conv['e0']:.12f} ({conv['iterations']} it, residual {conv['residual']:.2e})",
          file=sys.stderr, flush=True)

    side = None
    if args.graphs:
        side = torch.cuda.Stream()
        plan.graphs(True)
    # Both precisions stop at the reference's rule, a Ritz-value change of at
    # most tol (eigen.f90:93-96); the mixed solve restarts in FP64 only when
    # its estimated energy error |r|^2 / gap exceeds a tenth of the 1e-8
    # per-site target.  A tiny rtol keeps the mixed pass on that same rule.
    etol = 0.1 * 1e-8 * nsite
    lat64, r64, psi64 = timed_solves(plan, x0_t, args.steps, args.warmup, "fp64", args.tol, args.max_iter, torch,
                                     stream=side)
    print(f"[bench] fp64 solve {np.median(lat64)*1e3:.1f} ms E0={r64['e0']:.12f} it={r64['iterations']}",
          file=sys.stderr, flush=True)
    latmp, rmp, psimp = timed_solves(plan, x0_t, args.steps, args.warmup, args.precision, args.tol, args.max_iter,
                                     torch, energy_tol=etol, rtol=1e-14, stream=side)
    print(f"[bench] {args.precision} solve {np.median(latmp)*1e3:.1f} ms E0={rmp['e0']:.12f} it={rmp['iterations']} "
          f"refine={rmp['refinements']}", file=sys.stderr, flush=True)
    # looser stopping rules |dE| <= rtol |E| for both precisions, each with
    # its dE/site, so the cheapest rule that still meets the target can be
    # read off; like the FP64 solve, the mixed one then ends at the rule
    # (no FP64 restart)
    rules = []
    for rtol in (1e-9, 1e-7):
        for prec in ("fp64", args.precision):
            lt, rr, _ = timed_solves(plan, x0_t, max(2, args.steps // 2), 1, prec, args.tol, args.max_iter, torch,
                                     rtol=rtol, stream=side, max_refine=0)
            rules.append({"precision": prec, "rtol": rtol, "solves_per_s": 1.0 / float(np.mean(lt)),
                          "iterations": rr["iterations"], "refinements": rr["refinements"],
                          "dE_per_site": abs(rr["e0"] - conv["e0"]) / nsite})
            print(f"[bench] rule {rules[-1]}", file=sys.stderr, flush=True)
    plan.graphs(False)

    # the dominant kernel timed on isolated applies of the mixed precision
    # (inside a solve the FP64 residual apply shares the stage counters)
    xf = x0_t.float()
    yf = torch.empty_like(xf)
    plan.profile(True)
    for _ in range(10):
        plan.apply(xf, yf, precision=args.precision)
    torch.cuda.synchronize()
    ms_k, n_k = plan.profile_read()
    plan.profile(False)
    kavg = {k: float(ms_k[i] / max(1, n_k[i])) for i, k in enumerate(KERNEL_KINDS)}   # ms per launch
    packed = args.precision == "bf16x3" and bool(st.get("packed"))
    # (the packed apply's slot 1 is the packing of x, not a GEMM stage)
    dom = max(("stage1_64", "stage2_64") if packed else KERNEL_KINDS[:4], key=lambda k: kavg[k])

    value = 1.0 / float(np.mean(latmp))
    v64 = 1.0 / float(np.mean(lat64))
    e_conv = conv["e0"]
    de_site = abs(rmp["e0"] - e_conv) / nsite
    # wavefunction infidelity of the mixed-precision ground state against the
    # FP64 one (both normalised; the sign of a ground state is arbitrary)
    ov = float(torch.dot(psimp.double(), psi64.double()) / (psimp.double().norm() * psi64.double().norm()))
    infidelity = 1.0 - abs(ov)
    ok = [r for r in rules if r["dE_per_site"] <= 1e-8]
    best = {p: max([r["solves_per_s"] for r in ok if r["precision"] == p] +
                   [value if p == args.precision else v64]) for p in ("fp64", args.precision)}
    line = {
        "metric": "Lanczos solves/s vs plain FP64 GPU at dE/site<=1e-8; p50/p99 solve latency",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.mean(latmp)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"fp32": "f32", "bf16x3": "bf16x3->f32"}[args.precision] + "+f64",
        "data": "synthetic",
        "config": {"workload": "tjsq Lanczos ground-state solve, t-J 6x16 cylinder (t=1, t'=-0.2, J=0.4, N=84, S=0), "
                               f"centre bond 47|s|s|47 reached by one right-to-left and one left-to-right step from "
                               f"synthetic blocks, m={args.m} SU(2) multiplets kept, start vector wave_trans_LR of "
                               "the previous bond's ground state",
                   "D": sb.dim, "precision": args.precision, "lanczos_tol": args.tol, "max_iter": args.max_iter},
        # the headline ratio: both precisions at the reference's stopping rule
        "fp64_solves_per_s": v64, "speedup_vs_fp64": value / v64,
        "dE_per_site": de_site, "dE_per_site_fp64": abs(r64["e0"] - e_conv) / nsite,
        "infidelity_vs_fp64": infidelity,
        "iterations_mixed": rmp["iterations"], "iterations_fp64": r64["iterations"],
        # each precision at its cheapest measured rule with dE/site <= 1e-8
        "stopping_rules": rules,
        "speedup_vs_fp64_cheapest_rules": best[args.precision] / best["fp64"],
        "e0_converged": e_conv, "iterations_converged": conv["iterations"],
        "e0_fp64": r64["e0"], "e0_mixed": rmp["e0"], "residual_mixed": rmp["residual"],
        "refinements": rmp["refinements"], "energy_err_est": rmp["energy_err_est"],
        "setup": setup, "setup_s": t_setup,
        "p50_ms": 1e3 * float(np.percentile(latmp, 50)), "p99_ms": 1e3 * float(np.percentile(latmp, 99)),
        "p50_ms_fp64": 1e3 * float(np.percentile(lat64, 50)), "p99_ms_fp64": 1e3 * float(np.percentile(lat64, 99)),
        "apply_kernel_ms": kavg,
        "apply_kernels": PACKED_SYMBOL if packed else {k: v.format(mode=MODE_ID[args.precision])
                                                       for k, v in KERNEL_SYMBOL.items()},
        "apply_ms": float(sum(ms_k[:5]) / max(1, n_k[2])),
        "apply_tflops_reference_order": st["flops_ref"] / (float(sum(ms_k[:5]) / max(1, n_k[2])) * 1e-3) / 1e12,
    }
    # roofline of the dominant kernel: the useful flops of its own task table
    # (sum of 2 mt nt K over its GEMM segments; no MFMA padding, no
    # recomputation) over its average launch time (events on the launching
    # stream, isolated applies of the mixed precision)
    fl = kernel_flops(capi.HostPlan(model, sb))
    if packed:   # one packed kernel per stage does both tile sizes' work
        fl = {"stage1_64": fl["stage1_64"] + fl["stage1_32"], "stage2_64": fl["stage2_64"] + fl["stage2_32"]}
    achieved = fl[dom] / (kavg[dom] * 1e-3)
    peak = PEAK[args.precision]
    kname = (PACKED_SYMBOL if packed else KERNEL_SYMBOL)[dom].format(mode=MODE_ID[args.precision])
    line["roofline"] = {"bound": "mfma", "kernel": kname, "achieved": achieved / 1e12, "peak": peak / 1e12,
                        "unit": "TFLOP/s", "frac": achieved / peak,
                        "traffic": measured_traffic(kname, args.precision, sb.dim),
                        "algorithmic_flops_per_launch": fl[dom], "avg_launch_ms": kavg[dom]}
    if tuned is not None:
        line["work_item_shape"] = tuned["shape"]
    line["lanczos_iteration_graphs"] = bool(args.graphs)
    if not args.no_update:
        line["block_update"] = block_update_step(model, sb, g, args.m, torch)   # g: the converged ground state
        print(f"[bench] block update {line['block_update']}", file=sys.stderr, flush=True)
    # the CPU baseline last, so MKL's worker threads cannot compete with the
    # host side of the GPU measurements
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(model, sb, x0, max(1, r64["applies"]),
                                                                 threads=args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
