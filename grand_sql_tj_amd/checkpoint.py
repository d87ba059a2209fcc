"""Checkpoint files of the tjsq DMRG in the reference's on-disk format.

The reference keeps every block model, truncation operator, fused basis and
ground state of a run on disk between the warm-up and the sweeps, and a
restarted run reads them back (code/tjsq/block.f90:2760-3097, called from
warm_up_point / left_to_right / right_to_left, dmrg_sub.f90:590-866).  This
module reads and writes those files, so that a run of the device driver can
start from a reference checkpoint and the reference can resume from one
written here.

Files and names (Get_model_name, dmrg_sub.f90:294-331; Get_wave_name,
dmrg_sub.f90:335-346), each indexed by the block length:

    sysblocks<L> / envblocks<L>   model_to_disk   unformatted sequential
    systruns<L>  / envtruns<L>    truns_to_disk   unformatted sequential
    sysbs<L>     / envbs<L>       basis_to_disk   formatted text
    eigvec<i>                     wave_to_disk    unformatted sequential

Unformatted sequential files are the usual Fortran record stream: each
record is a little-endian int32 byte count, the payload and the count again
(default INTEGER is 4 bytes, REAL(8) 8 bytes).  A record above 2**31 - 9
bytes is split into subrecords whose leading count is negative when another
subrecord follows and whose trailing count is negative when one precedes;
the reader accepts that form and the writer produces it.

Record sequences (one line per record):

    block  (block_to_disk, block.f90:2826-2848):
        len | num | dim | up_dif | down_dif
        then per sub block: <empty> | num_up, num_down, down_dif |
                            row_dim, sdim | mat(row_dim, sdim), column-major
    model  (model_to_disk, block.f90:2873-2908):
        len | block(ham) | per active site: block(elec_up), block(elec_down),
                                            block(spin_sd), block(num_sn)
    truns  (truns_to_disk, block.f90:2956-2975): one block
    wave   (wave_to_disk, block.f90:2760-2787):
        sys_len | env_len | num | dim
        then per sector: <empty> | sys_num_up, sys_num_down |
                         env_num_up, env_num_down | sys_dim, env_dim |
                         vec(sys_dim, env_dim), column-major

The basis file (basis_to_disk, block.f90:3000-3045) is text with the
formats (A4,I6), (A11,I4), (A13,I4) and (A10,I4,2X,A12,I4), one blank
record before each sector.  The reader parses the fixed columns of those
formats, as the reference's basis_from_disk does, and accepts a blank line
of any width.

The flat layouts are the library's: an OpSet holds per operator the header
(len, up_dif, down_dif, num) and per sub block (num_up, num_down, down_dif,
row_dim, sdim) with an offset into one column-major data array; a Basis
holds per sector (new_num_up, new_num_down, dim, num) and per sub sector
(bl_num_up, bl_num_down, st_num_up, st_num_down, spos, sdim).  The block's
`dim` field, which no layout column holds, is returned beside the data on
reading and taken from the caller (or the rule of `default_dims`) on
writing.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from .blocks import OPS_PER_SITE, Basis, OpSet

MAX_SUBRECORD = 2 ** 31 - 9      # largest payload one record marker describes


class CheckpointError(ValueError):
    """A file that does not hold the record sequence its reader expects."""


# ---------------------------------------------------------------------------
# File names
# ---------------------------------------------------------------------------

_KINDS = {"blocks": 1, "truns": 2, "bs": 3}


def file_name(side: int, kind: str, length: int) -> str:
    """sysname(length, k) / envname(length, k) of Get_model_name
    (dmrg_sub.f90:294-331): side 0 = system, 1 = environment; kind is
    "blocks", "truns" or "bs"."""
    if kind not in _KINDS:
        raise ValueError(f"unknown checkpoint kind {kind!r}")
    if side not in (0, 1):
        raise ValueError("side must be 0 (system) or 1 (environment)")
    if not 1 <= length <= 999999:
        raise ValueError("block length out of range")
    return ("sys" if side == 0 else "env") + kind + str(int(length))


def wave_file(level: int) -> str:
    """wave_name(level) of Get_wave_name (dmrg_sub.f90:335-346), level 1..8."""
    if not 1 <= level <= 8:
        raise ValueError("wave level must be 1..8")
    return "eigvec" + str(int(level))


# ---------------------------------------------------------------------------
# Fortran sequential unformatted records
# ---------------------------------------------------------------------------

class RecordWriter:
    def __init__(self, f, max_subrecord: int = MAX_SUBRECORD):
        self.f = f
        self.max_sub = int(max_subrecord)

    def record(self, payload=b""):
        view = memoryview(payload).cast("B")
        n = len(view)
        if n <= self.max_sub:
            self.f.write(struct.pack("<i", n))
            self.f.write(view)
            self.f.write(struct.pack("<i", n))
            return
        pos = 0
        while pos < n:
            k = min(self.max_sub, n - pos)
            more = pos + k < n
            self.f.write(struct.pack("<i", -k if more else k))
            self.f.write(view[pos:pos + k])
            self.f.write(struct.pack("<i", -k if pos > 0 else k))
            pos += k

    def ints(self, *values):
        self.record(np.asarray(values, dtype="<i4").tobytes())

    def reals(self, a):
        self.record(np.ascontiguousarray(a, dtype="<f8"))


class RecordReader:
    def __init__(self, f, name: str = "checkpoint"):
        self.f = f
        self.name = name

    def record(self) -> bytes:
        parts = []
        first = True
        while True:
            head = self.f.read(4)
            if len(head) < 4:
                if first and not head:
                    raise CheckpointError(f"{self.name}: end of file where a record was expected")
                raise CheckpointError(f"{self.name}: truncated record marker")
            (n,) = struct.unpack("<i", head)
            k = abs(n)
            body = self.f.read(k)
            tail = self.f.read(4)
            if len(body) < k or len(tail) < 4:
                raise CheckpointError(f"{self.name}: truncated record")
            (t,) = struct.unpack("<i", tail)
            if abs(t) != k or (t < 0) != (not first):
                raise CheckpointError(f"{self.name}: record markers disagree ({n}, {t})")
            parts.append(body)
            first = False
            if n >= 0:
                return b"".join(parts)

    def at_end(self) -> bool:
        pos = self.f.tell()
        more = self.f.read(1)
        self.f.seek(pos)
        return not more

    def ints(self, count: int):
        rec = self.record()
        if len(rec) < 4 * count:
            raise CheckpointError(f"{self.name}: record of {len(rec)} bytes where {count} integers were expected")
        return [int(v) for v in np.frombuffer(rec, dtype="<i4", count=count)]

    def int1(self) -> int:
        return self.ints(1)[0]

    def empty(self):
        self.record()

    def reals(self, count: int) -> np.ndarray:
        rec = self.record()
        if len(rec) < 8 * count:
            raise CheckpointError(f"{self.name}: record of {len(rec)} bytes where {count} reals were expected")
        return np.frombuffer(rec, dtype="<f8", count=count).astype(np.float64)


# ---------------------------------------------------------------------------
# Blocks, models and truncation operators
# ---------------------------------------------------------------------------

def _op_blocks(ops: OpSet, op: int):
    f = ops.first_sub(op)
    for s in range(f, f + int(ops.hdr[op, 3])):
        r, c, o = int(ops.sub[s, 3]), int(ops.sub[s, 4]), int(ops.off[s])
        yield ops.sub[s], ops.data[o:o + r * c]


def default_dims(ops: OpSet) -> np.ndarray:
    """The `dim` field of each operator as the reference's routines leave
    it: the sum of sdim over the operator's sub blocks (the operators of a
    new block model come out of update_trun_dia / _ndia and
    block_mul_block_ndia, which count it so, block.f90:1365-1430 and
    1956-2075, and Get_truncation_operator's result holds the same)."""
    out = np.zeros(ops.nops, dtype=np.int64)
    f = 0
    for op in range(ops.nops):
        n = int(ops.hdr[op, 3])
        out[op] = int(ops.sub[f:f + n, 4].astype(np.int64).sum())
        f += n
    return out


def _write_block(w: RecordWriter, ops: OpSet, op: int, dim: int):
    length, up_dif, down_dif, num = (int(v) for v in ops.hdr[op])
    for v in (length, num, dim, up_dif, down_dif):
        w.ints(v)
    for s, mat in _op_blocks(ops, op):
        w.record()
        w.ints(int(s[0]), int(s[1]), int(s[2]))
        w.ints(int(s[3]), int(s[4]))
        w.reals(mat)


def _read_block(r: RecordReader, hdr, sub, off, chunks, nd):
    length = r.int1()
    num = r.int1()
    dim = r.int1()
    up_dif = r.int1()
    down_dif = r.int1()
    if num < 0:
        raise CheckpointError(f"{r.name}: negative sub block count {num}")
    hdr.append((length, up_dif, down_dif, num))
    for _ in range(num):
        r.empty()
        u, d, dd = r.ints(3)
        rows, cols = r.ints(2)
        if rows < 0 or cols < 0:
            raise CheckpointError(f"{r.name}: negative block shape ({rows}, {cols})")
        n = rows * cols
        sub.append((u, d, dd, rows, cols))
        off.append(nd)
        chunks.append(r.reals(n))
        nd += n
    return dim, nd


def _opset(length, hdr, sub, off, chunks) -> OpSet:
    return OpSet(int(length), np.asarray(hdr, np.int32).reshape(-1, 4), np.asarray(sub, np.int32).reshape(-1, 5),
                 np.asarray(off, np.int64),
                 np.concatenate(chunks) if chunks else np.zeros(0, np.float64))


def _dims_for(ops: OpSet, dims):
    if dims is None:
        return default_dims(ops)
    dims = np.asarray(dims, dtype=np.int64).reshape(-1)
    if dims.shape[0] != ops.nops:
        raise ValueError(f"dims holds {dims.shape[0]} entries for {ops.nops} operators")
    return dims


def _check_opset(ops: OpSet):
    if ops.hdr.ndim != 2 or ops.hdr.shape[1] != 4 or ops.sub.ndim != 2 or ops.sub.shape[1] != 5:
        raise ValueError("operator set tables have the wrong shape")
    if int(ops.hdr[:, 3].sum()) != ops.sub.shape[0] or ops.off.shape[0] != ops.sub.shape[0]:
        raise ValueError("operator set header counts do not match its sub block table")
    end = ops.off + ops.sub[:, 3].astype(np.int64) * ops.sub[:, 4]
    if end.size and (int(end.max()) > ops.data.shape[0] or int(ops.off.min()) < 0):
        raise ValueError("operator set offsets run outside its data")


def save_model(path, ops: OpSet, dims=None):
    """model_to_disk (block.f90:2873-2908): the model's length, its
    Hamiltonian, then elec_up, elec_down, spin_sd and num_sn of each active
    site in order.  `dims` gives each operator's `dim` field (default:
    default_dims)."""
    _check_opset(ops)
    if (ops.nops - 1) % OPS_PER_SITE:
        raise ValueError("a model holds the Hamiltonian and four operators per active site")
    dims = _dims_for(ops, dims)
    with open(path, "wb") as f:
        w = RecordWriter(f)
        w.ints(ops.length)
        for op in range(ops.nops):
            _write_block(w, ops, op, int(dims[op]))


def load_model(path):
    """model_from_disk (block.f90:2910-2950) without its deletion of the
    next file: returns (OpSet, dims).  The number of active sites is read
    from the file (the reference takes it from nleg11(len) of the lattice)."""
    with open(path, "rb") as f:
        r = RecordReader(f, os.fspath(path))
        length = r.int1()
        hdr, sub, off, chunks, dims, nd = [], [], [], [], [], 0
        while True:
            dim, nd = _read_block(r, hdr, sub, off, chunks, nd)
            dims.append(dim)
            if r.at_end():
                break
    if (len(hdr) - 1) % OPS_PER_SITE:
        raise CheckpointError(f"{path}: {len(hdr)} blocks do not form a Hamiltonian and four operators per site")
    return _opset(length, hdr, sub, off, chunks), np.asarray(dims, np.int64)


def save_trun(path, trun: OpSet, dim=None):
    """truns_to_disk (block.f90:2956-2975): one block.  `dim` defaults to
    default_dims."""
    _check_opset(trun)
    if trun.nops != 1:
        raise ValueError("a truncation operator is one block")
    d = int(default_dims(trun)[0]) if dim is None else int(dim)
    with open(path, "wb") as f:
        _write_block(RecordWriter(f), trun, 0, d)


def load_trun(path):
    """truns_from_disk (block.f90:2977-2997): returns (OpSet, dim)."""
    with open(path, "rb") as f:
        r = RecordReader(f, os.fspath(path))
        hdr, sub, off, chunks = [], [], [], []
        dim, _ = _read_block(r, hdr, sub, off, chunks, 0)
        if not r.at_end():
            raise CheckpointError(f"{path}: data after the truncation operator")
    return _opset(hdr[0][0], hdr, sub, off, chunks), dim


# ---------------------------------------------------------------------------
# Wavefunctions
# ---------------------------------------------------------------------------

@dataclass
class Wave:
    sys_len: int
    env_len: int
    wsec: np.ndarray     # int32 (nw, 6): sys_up, sys_down, env_up, env_down, sys_dim, env_dim
    x: np.ndarray        # float64 (dim,), sector blocks back to back, each column-major


def save_wave(path, sys_len: int, env_len: int, wsec, x):
    """wave_to_disk (block.f90:2760-2787)."""
    wsec = np.asarray(wsec, dtype=np.int64).reshape(-1, 6)
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    sizes = wsec[:, 4] * wsec[:, 5]
    dim = int(sizes.sum())
    if dim != x.shape[0]:
        raise ValueError(f"the sector list describes {dim} amplitudes, x holds {x.shape[0]}")
    with open(path, "wb") as f:
        w = RecordWriter(f)
        for v in (sys_len, env_len, wsec.shape[0], dim):
            w.ints(int(v))
        o = 0
        for k in range(wsec.shape[0]):
            w.record()
            w.ints(int(wsec[k, 0]), int(wsec[k, 1]))
            w.ints(int(wsec[k, 2]), int(wsec[k, 3]))
            w.ints(int(wsec[k, 4]), int(wsec[k, 5]))
            w.reals(x[o:o + int(sizes[k])])
            o += int(sizes[k])


def load_wave(path) -> Wave:
    """wave_from_disk (block.f90:2791-2820)."""
    with open(path, "rb") as f:
        r = RecordReader(f, os.fspath(path))
        sys_len, env_len, num, dim = r.int1(), r.int1(), r.int1(), r.int1()
        if num < 0:
            raise CheckpointError(f"{path}: negative sector count {num}")
        wsec, chunks = [], []
        for _ in range(num):
            r.empty()
            su, sd = r.ints(2)
            eu, ed = r.ints(2)
            ns, ne = r.ints(2)
            if ns < 0 or ne < 0:
                raise CheckpointError(f"{path}: negative sector shape ({ns}, {ne})")
            wsec.append((su, sd, eu, ed, ns, ne))
            chunks.append(r.reals(ns * ne))
    x = np.concatenate(chunks) if chunks else np.zeros(0, np.float64)
    if x.shape[0] != dim:
        raise CheckpointError(f"{path}: header dim {dim} but the sectors hold {x.shape[0]} amplitudes")
    return Wave(sys_len, env_len, np.asarray(wsec, np.int32).reshape(-1, 6), x)


# ---------------------------------------------------------------------------
# Fused bases (text)
# ---------------------------------------------------------------------------

def _field(value: int, width: int) -> str:
    s = str(int(value))
    if len(s) > width:
        raise ValueError(f"{value} does not fit the reference's I{width} field")
    return s.rjust(width)


def save_basis(path, basis: Basis):
    """basis_to_disk (block.f90:3000-3045).  The sector's `idx` field is
    written equal to its sub sector count, which is what Get_basis leaves
    in it (block.f90:80-97)."""
    sec = np.asarray(basis.sec).reshape(-1, 4)
    sub = np.asarray(basis.sub).reshape(-1, 6)
    if int(sec[:, 3].sum()) != sub.shape[0]:
        raise ValueError("basis sector counts do not match its sub sector table")
    lines = ["len=" + _field(basis.length, 6), "num=" + _field(sec.shape[0], 6),
             "dim=" + _field(int(sec[:, 2].astype(np.int64).sum()), 6)]
    q = 0
    for up, down, dim, num in (tuple(int(v) for v in row) for row in sec):
        lines.append("")
        lines += ["idx=" + _field(num, 6), "num=" + _field(num, 6), "dim=" + _field(dim, 6),
                  "new_num_up=" + _field(up, 4), "new_num_down=" + _field(down, 4)]
        for bu, bd, su, sd, spos, sdim in (tuple(int(v) for v in row) for row in sub[q:q + num]):
            lines += ["pos=" + _field(spos, 6), "dim=" + _field(sdim, 6),
                      "bl_num_up=" + _field(bu, 4) + "  bl_num_down=" + _field(bd, 4),
                      "st_num_up=" + _field(su, 4) + "  st_num_down=" + _field(sd, 4)]
        q += num
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _int_at(line: str, lo: int, hi: int, path, what) -> int:
    s = line[lo:hi].strip()
    if not s:
        return 0             # a blank numeric field reads as zero (BN editing)
    try:
        return int(s)
    except ValueError:
        raise CheckpointError(f"{path}: cannot read {what} from {line!r}") from None


def load_basis(path):
    """basis_from_disk (block.f90:3047-3097): returns (Basis, idx), idx the
    sectors' `idx` fields."""
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 0

    def nxt():
        nonlocal pos
        if pos >= len(lines):
            raise CheckpointError(f"{path}: end of file inside the basis")
        pos += 1
        return lines[pos - 1]

    length = _int_at(nxt(), 4, 10, path, "len")
    nsec = _int_at(nxt(), 4, 10, path, "num")
    _int_at(nxt(), 4, 10, path, "dim")
    if nsec < 0:
        raise CheckpointError(f"{path}: negative sector count {nsec}")
    sec, sub, idx = [], [], []
    for _ in range(nsec):
        if nxt().strip():
            raise CheckpointError(f"{path}: expected the blank record before a sector")
        idx.append(_int_at(nxt(), 4, 10, path, "idx"))
        num = _int_at(nxt(), 4, 10, path, "num")
        dim = _int_at(nxt(), 4, 10, path, "dim")
        up = _int_at(nxt(), 11, 15, path, "new_num_up")
        down = _int_at(nxt(), 13, 17, path, "new_num_down")
        if num < 0:
            raise CheckpointError(f"{path}: negative sub sector count {num}")
        sec.append((up, down, dim, num))
        for _ in range(num):
            spos = _int_at(nxt(), 4, 10, path, "pos")
            sdim = _int_at(nxt(), 4, 10, path, "dim")
            b = nxt()
            s = nxt()
            sub.append((_int_at(b, 10, 14, path, "bl_num_up"), _int_at(b, 28, 32, path, "bl_num_down"),
                        _int_at(s, 10, 14, path, "st_num_up"), _int_at(s, 28, 32, path, "st_num_down"),
                        spos, sdim))
    return (Basis(length, np.asarray(sec, np.int32).reshape(-1, 4), np.asarray(sub, np.int32).reshape(-1, 6)),
            np.asarray(idx, np.int32))
