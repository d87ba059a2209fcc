"""The checkpoint files (grand_sql_tj_amd.checkpoint) against the
reference's own writers and readers (model/truns/basis/wave _to_disk and
_from_disk, code/tjsq/block.f90:2760-3097, compiled into oracle/_ref):

- what the reference writes, this module reads back exactly, `dim` fields
  included;
- what this module writes is byte for byte the file the reference writes
  for the same data, and the reference's readers load it back exactly;
- the `dim` fields that default_dims fills in are the ones the reference's
  Get_Hamiltonian_Sys / _Env leave in a new block model and its
  Get_truncation_operator in a truncation operator.

The record framing is pinned on its own by known byte strings, so a framing
mistake shows without the reference build."""
import os
import struct

import numpy as np
import pytest

import oracle
from helpers import GEOMS, model, opset, random_trun, small_superblock
from grand_sql_tj_amd import checkpoint as ck
from grand_sql_tj_amd import synth

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")


def _rec(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def _same_ops(a, b):
    assert a.length == b.length
    assert np.array_equal(np.asarray(a.hdr), np.asarray(b.hdr))
    assert np.array_equal(np.asarray(a.sub), np.asarray(b.sub))
    assert np.array_equal(np.asarray(a.off), np.asarray(b.off))
    assert np.array_equal(np.asarray(a.data), np.asarray(b.data))


def _with_len(ops, length):
    hdr = np.array(ops.hdr, np.int32)
    hdr[:, 0] = length
    return opset(length, hdr, ops.sub, ops.off, ops.data)


# ---- framing and names (no reference needed) -------------------------------------------

def test_record_bytes_known_answer(tmp_path):
    p = tmp_path / "eigvec1"
    x = np.array([1.5, -2.0, 0.25])
    ck.save_wave(p, 3, 4, [[1, 1, 2, 0, 1, 3]], x)
    want = (_rec(struct.pack("<i", 3)) + _rec(struct.pack("<i", 4)) + _rec(struct.pack("<i", 1))
            + _rec(struct.pack("<i", 3)) + _rec(b"") + _rec(struct.pack("<2i", 1, 1)) + _rec(struct.pack("<2i", 2, 0))
            + _rec(struct.pack("<2i", 1, 3)) + _rec(struct.pack("<3d", 1.5, -2.0, 0.25)))
    assert p.read_bytes() == want
    w = ck.load_wave(p)
    assert (w.sys_len, w.env_len) == (3, 4)
    assert w.wsec.tolist() == [[1, 1, 2, 0, 1, 3]] and np.array_equal(w.x, x)


def test_subrecords_split_and_join(tmp_path):
    p = tmp_path / "rec"
    payload = bytes(range(10))
    with open(p, "wb") as f:
        ck.RecordWriter(f, max_subrecord=4).record(payload)
    raw = p.read_bytes()
    # 4 + 4 + 2 bytes: leading markers -4, -4, 2; trailing markers 4, -4, -2
    want = (struct.pack("<i", -4) + payload[:4] + struct.pack("<i", 4)
            + struct.pack("<i", -4) + payload[4:8] + struct.pack("<i", -4)
            + struct.pack("<i", 2) + payload[8:] + struct.pack("<i", -2))
    assert raw == want
    with open(p, "rb") as f:
        assert ck.RecordReader(f).record() == payload


def test_bad_files_raise(tmp_path):
    p = tmp_path / "bad"
    p.write_bytes(struct.pack("<i", 8) + b"\0" * 8 + struct.pack("<i", 7))
    with pytest.raises(ck.CheckpointError):
        ck.load_trun(p)
    p.write_bytes(_rec(struct.pack("<i", 2)) + struct.pack("<i", 4))
    with pytest.raises(ck.CheckpointError):
        ck.load_wave(p)
    ck.save_wave(p, 1, 1, [[0, 0, 0, 0, 1, 2]], [1.0, 2.0])
    raw = p.read_bytes()
    p.write_bytes(raw[:-6])
    with pytest.raises(ck.CheckpointError):
        ck.load_wave(p)
    with pytest.raises(ValueError):
        ck.save_wave(p, 1, 1, [[0, 0, 0, 0, 1, 2]], [1.0])
    with pytest.raises(ValueError):
        ck.file_name(2, "blocks", 3)
    with pytest.raises(ValueError):
        ck.file_name(0, "models", 3)


def test_names():
    assert ck.file_name(0, "blocks", 7) == "sysblocks7"
    assert ck.file_name(1, "truns", 12) == "envtruns12"
    assert ck.file_name(0, "bs", 140) == "sysbs140"
    assert ck.wave_file(1) == "eigvec1"


def test_basis_text_known_answer(tmp_path):
    from grand_sql_tj_amd.blocks import Basis
    b = Basis(2, np.array([[1, 1, 2, 2]], np.int32), np.array([[0, 0, 1, 1, 0, 1], [1, 1, 0, 0, 1, 1]], np.int32))
    p = tmp_path / "sysbs2"
    ck.save_basis(p, b)
    assert p.read_text() == ("len=     2\nnum=     1\ndim=     2\n\nidx=     2\nnum=     2\ndim=     2\n"
                             "new_num_up=   1\nnew_num_down=   1\n"
                             "pos=     0\ndim=     1\nbl_num_up=   0  bl_num_down=   0\nst_num_up=   1  st_num_down=   1\n"
                             "pos=     1\ndim=     1\nbl_num_up=   1  bl_num_down=   1\nst_num_up=   0  st_num_down=   0\n")
    r, idx = ck.load_basis(p)
    assert r.length == 2 and np.array_equal(r.sec, b.sec) and np.array_equal(r.sub, b.sub)
    assert idx.tolist() == [2]


def test_model_round_trip_without_reference(tmp_path):
    m = model()
    sb = small_superblock(m, 7, 9, seed=5)
    ops = _with_len(sb.sys, sb.sys_len)
    p = tmp_path / "sysblocks7"
    ck.save_model(p, ops)
    back, dims = ck.load_model(p)
    _same_ops(back, ops)
    assert np.array_equal(dims, ck.default_dims(ops))


# ---- against the reference's writers and readers -----------------------------------------

@needs_ref
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("geom", GEOMS[:4])
def test_model_files_match_reference(tmp_path, geom, side):
    m = model(up_bias=1, mu=0.0)
    sl, el = geom
    sb = small_superblock(m, sl, el, seed=31 + sl + el)
    pre, bs = (sb.sys, sb.sys_bs) if side == 0 else (sb.env, sb.env_bs)
    tr = random_trun(bs, 7 + sl)
    # a block model as the reference's Get_Hamiltonian leaves it, dim fields included
    new = oracle.ref_get_hamiltonian(m, side, sb.sys_len, sb.env_len, pre, bs, tr)
    dims = oracle.ref_result_dims()
    length = bs.length
    ops = _with_len(new, length)
    assert np.array_equal(np.asarray(new.hdr)[:, 0], np.full(new.nops, length))
    assert np.array_equal(ck.default_dims(ops), dims)
    idx, system = length, side == 0
    name = ck.file_name(side, "blocks", idx)
    # reference writes, we read
    refdir, ourdir = tmp_path / "ref", tmp_path / "ours"
    refdir.mkdir()
    ourdir.mkdir()
    oracle.ref_model_to_disk(m, refdir, ops, dims, idx, system)
    got, gdims = ck.load_model(refdir / name)
    _same_ops(got, ops)
    assert np.array_equal(gdims, dims)
    # we write, the file is the reference's, and the reference reads it
    ck.save_model(ourdir / name, ops)
    assert (ourdir / name).read_bytes() == (refdir / name).read_bytes()
    back, bdims = oracle.ref_model_from_disk(m, ourdir, idx, system)
    _same_ops(back, ops)
    assert np.array_equal(bdims, dims)


@needs_ref
@pytest.mark.parametrize("side", [0, 1])
def test_truns_and_basis_files_match_reference(tmp_path, side):
    m = model(up_bias=1)
    sb = small_superblock(m, 11, 11, seed=3)
    bs = sb.sys_bs if side == 0 else sb.env_bs
    # a truncation operator as the reference's Get_truncation_operator leaves it
    refdir, ourdir, logdir = tmp_path / "ref", tmp_path / "ours", tmp_path / "log"
    for d in (refdir, ourdir, logdir):
        d.mkdir()
    psi = synth.random_vector(sb, 9)
    psi /= np.linalg.norm(psi)
    den = oracle.ref_density_matrix(side, sb.sys_len + 1, sb.env_len + 1, sb.wsec, psi)
    tr, _, _ = oracle.ref_truncation(den, 4, 12, 1e-6, logdir, system_block=side == 0)
    dim = int(oracle.ref_result_dims()[0])
    assert ck.default_dims(tr).tolist() == [dim] and 4 <= dim <= 12
    tr = opset(bs.length, np.array([[bs.length, 0, 0, tr.hdr[0, 3]]], np.int32), tr.sub, tr.off, tr.data)
    idx = bs.length
    tname, bname = ck.file_name(side, "truns", idx), ck.file_name(side, "bs", idx)
    oracle.ref_truns_to_disk(refdir, tr, dim, idx, side == 0)
    oracle.ref_basis_to_disk(refdir, bs, idx, side == 0)
    t, d = ck.load_trun(refdir / tname)
    _same_ops(t, tr)
    assert d == dim
    b, bidx = ck.load_basis(refdir / bname)
    assert b.length == bs.length and np.array_equal(b.sec, bs.sec) and np.array_equal(b.sub, bs.sub)
    assert np.array_equal(bidx, np.asarray(bs.sec)[:, 3])
    ck.save_trun(ourdir / tname, tr)
    ck.save_basis(ourdir / bname, bs)
    assert (ourdir / tname).read_bytes() == (refdir / tname).read_bytes()
    assert (ourdir / bname).read_bytes() == (refdir / bname).read_bytes()
    rt, rd = oracle.ref_truns_from_disk(ourdir, idx, side == 0)
    _same_ops(rt, tr)
    assert rd == dim
    ln, sec, sub, ridx = oracle.ref_basis_from_disk(ourdir, idx, side == 0)
    assert ln == bs.length and np.array_equal(sec, bs.sec) and np.array_equal(sub, bs.sub)
    assert np.array_equal(ridx, np.asarray(bs.sec)[:, 3])


@needs_ref
@pytest.mark.parametrize("geom", [(5, 5), (7, 9), (13, 17)])
def test_wave_files_match_reference(tmp_path, geom):
    m = model(up_bias=1)
    sb = small_superblock(m, *geom, seed=11)
    x = synth.random_vector(sb, 4)
    refdir, ourdir = tmp_path / "ref", tmp_path / "ours"
    refdir.mkdir()
    ourdir.mkdir()
    name = ck.wave_file(1)
    oracle.ref_wave_to_disk(refdir, name, sb.sys_len + 1, sb.env_len + 1, sb.wsec, x)
    w = ck.load_wave(refdir / name)
    assert (w.sys_len, w.env_len) == (sb.sys_len + 1, sb.env_len + 1)
    assert np.array_equal(w.wsec, sb.wsec) and np.array_equal(w.x, x)
    ck.save_wave(ourdir / name, sb.sys_len + 1, sb.env_len + 1, sb.wsec, x)
    assert (ourdir / name).read_bytes() == (refdir / name).read_bytes()
    sl, el, wsec, y = oracle.ref_wave_from_disk(ourdir, name)
    assert (sl, el) == (sb.sys_len + 1, sb.env_len + 1)
    assert np.array_equal(wsec, sb.wsec) and np.array_equal(y, x)
    assert not os.path.exists(os.path.join(os.getcwd(), name))
