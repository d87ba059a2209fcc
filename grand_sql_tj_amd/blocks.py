"""Block-sparse SU(2) containers in the flat layout the C-ABI takes.

The layout mirrors the reference's derived types (code/tjsq/common.f90):

* ``OpSet`` holds a ``Total_Model`` (common.f90:139-149) as a list of operators
  (``Total_Block``, common.f90:114-122).  Operator 0 is the block Hamiltonian;
  operator ``1 + 4*(i-1) + q`` is active site ``i``'s ``elec_up`` (q=0),
  ``elec_down`` (1), ``spin_sd`` (2) or ``num_sn`` (3).
  ``hdr[op] = (len, up_dif, down_dif, nsub)``; ``sub[s] = (num_up, num_down,
  down_dif, row_dim, sdim)`` with the KET quantum numbers; ``off[s]`` is the
  offset, in elements, of the column-major ``row_dim x sdim`` matrix in
  ``data``.
* ``Basis`` holds a ``Total_Basis`` (common.f90:65-82): ``sec[k] = (new_num_up,
  new_num_down, dim, nsub)`` and ``sub[q] = (bl_num_up, bl_num_down, st_num_up,
  st_num_down, spos, sdim)``.
* ``Superblock`` carries the sys/env models, their enlarged bases and the
  wavefunction sector list ``wsec[i] = (sys_num_up, sys_num_down, env_num_up,
  env_num_down, sys_dim, env_dim)`` (Super_Basis, common.f90:88-97).  A
  wavefunction is one contiguous vector: sectors back to back, each
  column-major ``sys_dim x env_dim``.

Quantum numbers: ``num_down`` is 2S; ``num_up`` is N when particle number is
conserved (``up_bias = 1``) and 0 in the reference's shipped SU(2)-only mode.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

OPS_PER_SITE = 4
ELEC_UP, ELEC_DOWN, SPIN_SD, NUM_SN = range(4)


@dataclass
class OpSet:
    length: int
    hdr: np.ndarray      # int32 (nops, 4)
    sub: np.ndarray      # int32 (nsub, 5)
    off: np.ndarray      # int64 (nsub,)
    data: np.ndarray     # float64 (ndata,)

    @property
    def nops(self) -> int:
        return int(self.hdr.shape[0])

    @property
    def nact(self) -> int:
        return (self.nops - 1) // OPS_PER_SITE

    def first_sub(self, op: int) -> int:
        return int(self.hdr[:op, 3].sum())

    def blocks(self, op: int):
        """Yield (sub_row, matrix view) of operator ``op``."""
        f = self.first_sub(op)
        for s in range(f, f + int(self.hdr[op, 3])):
            r, c = int(self.sub[s, 3]), int(self.sub[s, 4])
            o = int(self.off[s])
            yield self.sub[s], self.data[o:o + r * c].reshape(c, r).T

    def sectors(self):
        """Sector list (num_up, num_down, dim) from the Hamiltonian's blocks."""
        f = self.first_sub(0)
        n = int(self.hdr[0, 3])
        s = self.sub[f:f + n]
        return np.stack([s[:, 0], s[:, 1], s[:, 4]], axis=1).astype(np.int32)


@dataclass
class Basis:
    length: int
    sec: np.ndarray      # int32 (nsec, 4)
    sub: np.ndarray      # int32 (nsub, 6)

    @property
    def nsec(self) -> int:
        return int(self.sec.shape[0])


@dataclass
class Superblock:
    sys_len: int
    env_len: int
    sys: OpSet
    env: OpSet
    sys_bs: Basis
    env_bs: Basis
    wsec: np.ndarray     # int32 (nw, 6)
    meta: dict = field(default_factory=dict)

    @property
    def dim(self) -> int:
        return int((self.wsec[:, 4].astype(np.int64) * self.wsec[:, 5]).sum())

    def sector_offsets(self) -> np.ndarray:
        sz = self.wsec[:, 4].astype(np.int64) * self.wsec[:, 5]
        return np.concatenate([[0], np.cumsum(sz)])


# ---------------------------------------------------------------------------
# Fusion bases (host restatement used by the generators; the C-ABI exports the
# same construction as tjsq_get_basis / tjsq_get_super_basis).
# ---------------------------------------------------------------------------

def site_sectors(up_bias: int) -> np.ndarray:
    """Single-site sectors: hole (0, 0) and electron (up_bias, 1)
    (Get_single_site_basis, code/tjsq/dmrg_sub.f90:521-551)."""
    return np.array([[0, 0, 1], [up_bias, 1, 1]], dtype=np.int32)


def _sort_qn(pairs):
    """QuickSort_Integer (code/tjsq/block.f90:212-239): exchange sort, 2S
    descending, then N descending among equal 2S."""
    a = [list(p) for p in pairs]
    n = len(a)
    for i in range(n):
        for x in range(i + 1, n):
            if a[i][1] < a[x][1]:
                a[i], a[x] = a[x], a[i]
    for i in range(n):
        for x in range(i + 1, n):
            if a[i][1] == a[x][1] and a[i][0] < a[x][0]:
                a[i], a[x] = a[x], a[i]
    return a


def get_basis(block_len: int, blk: np.ndarray, site: np.ndarray) -> Basis:
    """Get_basis (code/tjsq/block.f90:9-104) on (num_up, num_down, dim) lists."""
    pre = []
    for bu, bd, _ in blk:
        for su_, sd, _ in site:
            for ss in range(abs(bd - sd), bd + sd + 1, 2):
                pre.append((bu + su_, ss))
    pre = _sort_qn(pre)
    uniq = []
    for p in pre:           # Get_Unique_QN keeps the first of each equal run
        if not uniq or uniq[-1] != p:
            if p not in uniq:
                uniq.append(p)
    secs, subs = [], []
    for up, down in uniq:
        dim = 0
        cnt = 0
        for bu, bd, bdim in blk:
            for su_, sd, sdim in site:
                if bu + su_ != up:
                    continue
                if abs(bd - sd) <= down <= bd + sd:
                    subs.append((bu, bd, su_, sd, dim, bdim * sdim))
                    dim += bdim * sdim
                    cnt += 1
        secs.append((up, down, dim, cnt))
    return Basis(block_len + 1, np.array(secs, dtype=np.int32).reshape(-1, 4),
                 np.array(subs, dtype=np.int32).reshape(-1, 6))


def get_super_basis(sys_bs: Basis, env_bs: Basis, total_up: int, total_down: int) -> np.ndarray:
    """Get_super_basis (code/tjsq/block.f90:110-161)."""
    out = []
    for su_, sd, sdim, _ in sys_bs.sec:
        for eu, ed, edim, _ in env_bs.sec:
            if eu != total_up - su_:
                continue
            if abs(ed - sd) <= total_down <= ed + sd:
                out.append((su_, sd, eu, ed, sdim, edim))
    return np.array(out, dtype=np.int32).reshape(-1, 6)


# ---------------------------------------------------------------------------
# Operator construction helpers
# ---------------------------------------------------------------------------

class OpSetBuilder:
    """Accumulates operators into the flat OpSet layout."""

    def __init__(self, length: int):
        self.length = length
        self.hdr = []
        self.sub = []
        self.off = []
        self.chunks = []
        self.ndata = 0

    def add(self, up_dif: int, down_dif: int, blocks):
        """blocks: iterable of (num_up, num_down, down_dif, matrix)."""
        n = 0
        for up, down, dd, m in blocks:
            m = np.asarray(m, dtype=np.float64)
            r, c = m.shape
            self.sub.append((up, down, dd, r, c))
            self.off.append(self.ndata)
            self.chunks.append(np.asfortranarray(m).ravel(order="F"))
            self.ndata += r * c
            n += 1
        self.hdr.append((self.length, up_dif, down_dif, n))

    def build(self) -> OpSet:
        data = np.concatenate(self.chunks) if self.chunks else np.zeros(0)
        return OpSet(self.length, np.array(self.hdr, dtype=np.int32).reshape(-1, 4),
                     np.array(self.sub, dtype=np.int32).reshape(-1, 5),
                     np.array(self.off, dtype=np.int64), np.ascontiguousarray(data))


def spin_hermitian_factor(bra_2s: int, ket_2s: int) -> float:
    """Factor f with <a||S||b> = f * <b||S||a>^T for a hermitian rank-1 block
    operator in the reduced-matrix convention of the apply kernels (bra 2S = a,
    ket 2S = b): f = (-1)^((a-b)/2).  With it the spin-spin terms of H_eff
    (get_sys.f90:2016-2981, one pass without an explicit conjugate) are
    symmetric; the tests check <x, H y> = <H x, y>."""
    return -1.0 if ((bra_2s - ket_2s) // 2) % 2 else 1.0
