"""numpy evaluation of the task tables a plan uploads (test helper).

Executes merge recipes, stage 1 and stage 2 exactly as the device kernels
define them, so the plan compiler can be checked on the CPU against the
oracle.  This is a checker of the tables, not a product path.
"""
import numpy as np

SP_OPS, SP_X, SP_T, SP_IDENT = 0, 1, 2, 3


def _mat(buf, off, ld, rows, cols, trans):
    """op(stored) restricted to rows x cols: stored column-major with ld."""
    if not trans:
        idx = off + np.arange(rows)[:, None] + np.arange(cols)[None, :] * ld
    else:
        idx = off + np.arange(rows)[:, None] * ld + np.arange(cols)[None, :]
    if idx.size:
        assert idx.min() >= 0 and idx.max() < buf.size, "table reads outside its buffer"
    return buf[idx]


def _store(buf, idx, val):
    if idx.size:
        assert idx.min() >= 0 and idx.max() < buf.size, "table writes outside its buffer"
    buf[idx] = val


def evaluate(hp, sb, x):
    arena = np.zeros(hp.arena_elems)
    arena[hp.sys_base:hp.sys_base + sb.sys.data.size] = sb.sys.data
    arena[hp.env_base:hp.env_base + sb.env.data.size] = sb.env.data
    for m in hp.merge:   # recipes accumulate in table order per destination
        d, s, n = int(m["dst_off"]), int(m["src_off"]), int(m["n"])
        arena[d:d + n] = arena[d:d + n] + m["coef"] * arena[s:s + n]
    y = np.zeros(hp.dim)
    tables = {"tile1": hp.tile1, "seg1": hp.seg1, "tile2": hp.tile2, "seg2": hp.seg2, "rtile": hp.rtile}
    return run_tables(tables, hp.n64_1, hp.n64_2, hp.scratch_elems, hp.nslots, arena, x, y)


def evaluate_program(prog, ops, x):
    """Output of a tjsq_program (capi.Program) on host arrays ops / x."""
    z = prog.sizes
    y = np.zeros(z["ndata"])
    ops = np.zeros(1) if ops is None else np.asarray(ops, np.float64)
    return run_tables(prog.tables(), z["n64_1"], z["n64_2"], z["scratch_elems"], z["n_slots"], ops,
                      np.asarray(x, np.float64), y)


def run_tables(tb, n64_1, n64_2, scratch_elems, nslots, arena, x, y):
    T = np.zeros(max(1, scratch_elems))
    part = np.zeros(max(1, nslots) * 1024)
    n64 = {1: n64_1, 2: n64_2}
    spaces = {SP_OPS: arena, SP_X: x, SP_T: T}
    for stage, tiles, segs, out in ((1, tb["tile1"], tb["seg1"], T), (2, tb["tile2"], tb["seg2"], y)):
        for ti, tl in enumerate(tiles):
            edge = 64 if ti < n64[stage] else 32
            m0, n0, mt, nt = int(tl["m0"]), int(tl["n0"]), int(tl["mt"]), int(tl["nt"])
            acc = np.zeros((mt, nt))
            for s in segs[int(tl["seg_begin"]):int(tl["seg_end"])]:
                L = spaces[int(s["l_space"])]
                if int(s["r_space"]) == SP_IDENT:
                    acc += s["c"] * _mat(L, int(s["l_off"]) + m0 + n0 * int(s["l_ld"]), int(s["l_ld"]), mt, nt, 0)
                    continue
                R = spaces[int(s["r_space"])]
                K = int(s["K"])
                lt, rt = int(s["l_t"]), int(s["r_t"])
                lo = int(s["l_off"]) + (m0 * int(s["l_ld"]) if lt else m0)
                ro = int(s["r_off"]) + (n0 if rt else n0 * int(s["r_ld"]))
                Lm = _mat(L, lo, int(s["l_ld"]), mt, K, lt)
                Rm = _mat(R, ro, int(s["r_ld"]), K, nt, rt)
                acc += (s["c"] * Lm) @ Rm
            if int(tl["slot"]) >= 0:
                idx = int(tl["slot"]) * 1024 + np.arange(mt)[:, None] + np.arange(nt)[None, :] * edge
                _store(part, idx, acc)
                continue
            o, ld = int(tl["out_off"]), int(tl["out_ld"])
            idx = o + (m0 + np.arange(mt))[:, None] + (n0 + np.arange(nt))[None, :] * ld
            _store(out, idx, acc)
    for tl in tb["rtile"]:   # ordered sums of the partial tiles
        m0, n0, mt, nt = int(tl["m0"]), int(tl["n0"]), int(tl["mt"]), int(tl["nt"])
        acc = np.zeros((mt, nt))
        units, pld = int(tl["slot"]), int(tl["out_space"])
        for q in range(int(tl["seg_begin"]), int(tl["seg_end"]), units):
            acc += _mat(part, q * 1024, pld, mt, nt, 0)
        o, ld = int(tl["out_off"]), int(tl["out_ld"])
        _store(y, o + (m0 + np.arange(mt))[:, None] + (n0 + np.arange(nt))[None, :] * ld, acc)
    return y
