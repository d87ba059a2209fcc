"""The packed BF16x3 tables on the CPU: evaluated byte for byte in numpy
(packed_eval.py, every read and write bounds-checked) against the oracle."""
import numpy as np
import pytest

import oracle
import packed_eval
from helpers import GEOMS, model, rel, small_superblock
from grand_sql_tj_amd import capi, synth


@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("geom", GEOMS)
def test_packed_tables_reproduce_oracle(ub, geom):
    m = model(up_bias=ub)
    sb = small_superblock(m, *geom, seed=7 * geom[0] + geom[1] + ub, msys=40, menv=44, median=5)
    x = synth.random_vector(sb, 3).astype(np.float32)
    hp = capi.HostPlan(m, sb)
    y = packed_eval.evaluate(hp, sb, x)
    yo = oracle.model_wave(m, sb, x.astype(np.float64))
    # BF16x3: each product to ~2^-16 relative, sums of O(100) terms
    assert rel(y, yo) < 2e-5


def test_packed_tables_split_long_blocks():
    # blocks wider than a tile and K longer than a piece: several tiles per
    # block, partial slots and their ordered reduction
    m = model(up_bias=1)
    sb = small_superblock(m, 11, 11, seed=5, msys=200, menv=210, median=90, nsec=3)
    hp = capi.HostPlan(m, sb)
    pk = hp.packed()
    assert pk["nslots"] > 0 and len(pk["rtile"]) > 0
    assert (pk["tile2"]["mb"] == 4).any() and (pk["tile1"]["mb"] == 4).any()
    x = synth.random_vector(sb, 4).astype(np.float32)
    assert rel(packed_eval.evaluate(hp, sb, x), oracle.model_wave(m, sb, x.astype(np.float64))) < 2e-5


def test_packed_layout_geometry():
    m = model(up_bias=1)
    sb = small_superblock(m, 7, 9, seed=2, msys=60, menv=60, median=6)
    pk = capi.HostPlan(m, sb).packed()
    for jobs in (pk["ops_jobs"], pk["x_jobs"]):
        assert (jobs["lines"] % 16 == 0).all() and (jobs["kpad"] % 16 == 0).all()
        assert (jobs["dst_off"] % 64 == 0).all()
        end = jobs["dst_off"] + jobs["lines"].astype(np.int64) * jobs["kpad"] * 4
        order = np.argsort(jobs["dst_off"])
        assert (end[order][:-1] <= jobs["dst_off"][order][1:]).all()   # copies do not overlap
    for segs in (pk["seg1"], pk["seg2"]):
        g = segs[(segs["spaces"] & 0xff) != 3]
        assert (g["K"] % 16 == 0).all() and (g["K"] > 0).all()
        assert (g["l_off"] % 64 == 0).all() and (g["r_off"] % 64 == 0).all()
    # stage 2 runs column strips on one XCD each (workgroup g on XCD g mod 8):
    # its table is 8 interleaved bins, padded with empty tiles that read and
    # write nothing
    t2 = pk["tile2"]
    pad = t2["mb"] == 0
    assert t2.size % 8 == 0
    assert (t2["nb"][pad] == 0).all() and (t2["seg_begin"][pad] == t2["seg_end"][pad]).all()
    assert (t2["slot"][pad] == -1).all()
    t1 = pk["tile1"]
    pad1 = t1["mb"] == 0
    assert t1.size % 32 == 0      # quads of 4 tiles in 8 XCD bins
    assert (t1["nb"][pad1] == 0).all() and (t1["seg_begin"][pad1] == t1["seg_end"][pad1]).all()
    for t in (t1[~pad1], t2[~pad]):
        assert ((t["mb"] >= 1) & (t["mb"] <= 4) & (t["nb"] >= 1) & (t["nb"] <= 4)).all()
        assert (t["m0"] % 64 == 0).all() and (t["n0"] % 64 == 0).all()
    # every output tile of every block is there once (or split into partial
    # slots that one reduction sums)
    real = t2[~pad]
    whole = real[real["slot"] < 0]
    keys = set(zip(whole["out_off"].tolist(), whole["m0"].tolist(), whole["n0"].tolist()))
    assert len(keys) == whole.size
    red = set(zip(pk["rtile"]["out_off"].tolist(), pk["rtile"]["m0"].tolist(), pk["rtile"]["n0"].tolist()))
    assert not keys & red and len(red) == pk["rtile"].size
    assert sorted(real["slot"][real["slot"] >= 0].tolist()) == list(range(pk["nslots"]))
    # stage 1 runs in quads (tiles 4q..4q+3, one workgroup) that share the
    # x lines: every tile of a quad reads the lead tile's x rows and x
    # segments; the quads of one operator copy of their lead (bin q mod 8)
    # run on one XCD
    seg1 = pk["seg1"]
    home = {}
    for q in range(t1.size // 4):
        lead = t1[4 * q]
        quad = t1[4 * q:4 * q + 4]
        if lead["mb"] == 0:
            assert (quad["mb"] == 0).all()      # a whole padding quad of an XCD bin
            continue
        ls = seg1[lead["seg_begin"]:lead["seg_end"]]
        for tl in quad[quad["mb"] > 0]:
            assert tl["m0"] == lead["m0"] and tl["mb"] == lead["mb"]
            s_ = seg1[tl["seg_begin"]:tl["seg_end"]]
            assert len(s_) == len(ls)
            for f in ("l_off", "l_stride", "l_lim", "K"):
                assert (s_[f] == ls[f]).all()
        r = int(ls["r_off"][0])
        assert home.setdefault(r, q % 8) == q % 8
