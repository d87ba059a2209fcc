"""numpy evaluation of the packed BF16x3 tables (test helper).

Packs the operators and x into the byte layout the device uses (lines of k
in groups of eight: eight bf16 high halves, then eight low halves), runs
stage 1, stage 2 and the partial-slot reduction as the packed tile kernel
defines them (a . b = a_hi b_hi + a_hi b_lo + a_lo b_hi), and asserts that
every read and write stays inside its buffer.  A checker of the host
compiler's tables, not a product path.
"""
import numpy as np

PK_OPS, PK_X, PK_T, PK_XF = 0, 1, 2, 3


def bf16_split(v64):
    """(hi, lo) bf16 bit patterns of v: hi = rne(float(v)), lo = rne(float(v - hi))."""
    def rne(f32):
        u = np.ascontiguousarray(f32, np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
        return (u >> 16).astype(np.uint16)
    v64 = np.asarray(v64, np.float64)
    hi = rne(v64.astype(np.float32))
    hif = (hi.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    lo = rne((v64 - hif).astype(np.float32))
    return hi, lo


def bf16_value(h):
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _group_index(base_half, line, stride_bytes, k):
    """uint16 indices of the high halves of elements k of `line`."""
    return base_half + line * (stride_bytes // 2) + (k // 8) * 16 + (k % 8)


def pack(jobs, src, nbytes):
    buf = np.zeros(nbytes // 2, np.uint16)
    for j in jobs:
        lines, kpad, R, C, ld = int(j["lines"]), int(j["kpad"]), int(j["rows"]), int(j["cols"]), int(j["src_ld"])
        l = np.arange(lines)[:, None]
        k = np.arange(kpad)[None, :]
        r, c = (l, k) if j["trans"] else (k, l)
        ok = (r < R) & (c < C)
        v = np.zeros((lines, kpad))
        idx = int(j["src_off"]) + r + c * ld
        v[ok] = src[idx[ok]]
        hi, lo = bf16_split(v)
        hidx = _group_index(int(j["dst_off"]) // 2, l, kpad * 4, k)
        assert hidx.min() >= 0 and hidx.max() + 8 < buf.size
        buf[hidx] = hi
        buf[hidx + 8] = lo
    return buf


def _lines(buf, off, stride, l0, nl, K):
    """(hi, lo) values of lines l0..l0+nl, k < K, of the copy at byte offset off."""
    l = np.arange(l0, l0 + nl)[:, None]
    k = np.arange(K)[None, :]
    hidx = _group_index(off // 2, l, stride, k)
    assert hidx.min() >= 0 and hidx.max() + 8 < buf.size, "packed read outside its buffer"
    return bf16_value(buf[hidx]), bf16_value(buf[hidx + 8])


def _tile(pk, spaces, tl, segs, xf):
    mt, nt = 16 * int(tl["mb"]), 16 * int(tl["nb"])
    m0, n0 = int(tl["m0"]), int(tl["n0"])
    acc = np.zeros((mt, nt))
    for s in segs[tl["seg_begin"]:tl["seg_end"]]:
        ls, rs = int(s["spaces"]) & 0xff, (int(s["spaces"]) >> 8) & 0xff
        if ls == PK_XF:
            rows = np.arange(m0, m0 + mt)[:, None]
            cols = np.arange(n0, n0 + nt)[None, :]
            ok = (rows < tl["mlim"]) & (cols < tl["nlim"])
            idx = int(s["l_off"]) + rows + cols * int(s["l_stride"])
            acc[ok] += s["c"] * xf[idx[ok]]
            continue
        K = int(s["K"])
        assert K % 16 == 0 and K > 0
        lh, ll = _lines(spaces[ls], int(s["l_off"]), int(s["l_stride"]), m0, mt, K)
        rh, rl = _lines(spaces[rs], int(s["r_off"]), int(s["r_stride"]), n0, nt, K)
        acc += s["c"] * (lh @ rh.T + lh @ rl.T + ll @ rh.T)
    return acc.astype(np.float32).astype(np.float64)


def evaluate(hp, sb, x):
    """y = H_eff x through the packed tables of HostPlan hp (x: FP32 values)."""
    pk = hp.packed()
    assert pk is not None, "packed layout does not apply"
    arena = np.zeros(hp.arena_elems)
    arena[hp.sys_base:hp.sys_base + sb.sys.data.size] = sb.sys.data
    arena[hp.env_base:hp.env_base + sb.env.data.size] = sb.env.data
    for m in hp.merge:
        d, s, n = int(m["dst_off"]), int(m["src_off"]), int(m["n"])
        arena[d:d + n] = arena[d:d + n] + m["coef"] * arena[s:s + n]
    xf = np.asarray(x, np.float32).astype(np.float64)
    spaces = {PK_OPS: pack(pk["ops_jobs"], arena, pk["ops_bytes"]), PK_X: pack(pk["x_jobs"], xf, pk["x_bytes"]),
              PK_T: np.zeros(pk["t_bytes"] // 2, np.uint16)}
    for tl in pk["tile1"]:
        if tl["mb"] == 0:          # an XCD bin's padding: reads and writes nothing
            assert tl["nb"] == 0 and tl["seg_begin"] == tl["seg_end"]
            continue
        D = _tile(pk, spaces, tl, pk["seg1"], xf)
        u = np.arange(int(tl["m0"]), int(tl["m0"]) + D.shape[0])[:, None]
        v = np.arange(int(tl["n0"]), int(tl["n0"]) + D.shape[1])[None, :]
        hi, lo = bf16_split(D)
        hidx = _group_index(int(tl["out_off"]) // 2, v, int(tl["out_ld"]), u)
        assert hidx.min() >= 0 and hidx.max() + 8 < spaces[PK_T].size, "stage-1 write outside T"
        spaces[PK_T][hidx] = hi
        spaces[PK_T][hidx + 8] = lo
    y = np.zeros(hp.dim)
    written = np.zeros(hp.dim, bool)
    part = np.zeros((max(1, pk["nslots"]), 64, 64))
    for tl in pk["tile2"]:
        if tl["mb"] == 0:
            assert tl["nb"] == 0 and tl["seg_begin"] == tl["seg_end"] and tl["slot"] == -1
            continue
        A = _tile(pk, spaces, tl, pk["seg2"], xf)
        if tl["slot"] >= 0:
            part[tl["slot"], :A.shape[0], :A.shape[1]] = A
            continue
        m0, n0 = int(tl["m0"]), int(tl["n0"])
        rows = np.arange(m0, m0 + A.shape[0])[:, None]
        cols = np.arange(n0, n0 + A.shape[1])[None, :]
        ok = (rows < tl["mlim"]) & (cols < tl["nlim"])
        idx = int(tl["out_off"]) + rows + cols * int(tl["out_ld"])
        assert not written[idx[ok]].any(), "output element written twice"
        y[idx[ok]] = A[ok]
        written[idx[ok]] = True
    for r in pk["rtile"]:
        mt, nt = int(r["mt"]), int(r["nt"])
        s = part[int(r["s0"]):int(r["s1"]), :mt, :nt].sum(axis=0)
        rows = np.arange(int(r["m0"]), int(r["m0"]) + mt)[:, None]
        cols = np.arange(int(r["n0"]), int(r["n0"]) + nt)[None, :]
        idx = int(r["out_off"]) + rows + cols * int(r["out_ld"])
        assert not written[idx].any(), "output element written twice"
        y[idx] = s
        written[idx] = True
    assert written.all(), "output elements never written"
    return y
