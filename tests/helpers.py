"""Shared inputs for the tests."""
import os

import numpy as np

from grand_sql_tj_amd.lattice import TJModel
from grand_sql_tj_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# the reference build is compiled for Ny = 6 and a 24-column lattice (common.f90:10)
REF_NY, REF_NX = 6, 24

# geometries (sys_len, env_len): symmetric, asymmetric, and ones where the
# cross-cut bonds are carried by sys sites or by env sites
GEOMS = [(5, 5), (6, 5), (7, 9), (11, 11), (12, 10), (13, 17)]


def model(up_bias=1, t=1.0, tp=-0.2, J=0.4, mu=0.3, ny=REF_NY, nx=REF_NX):
    return TJModel(ny=ny, nx=nx, t=t, tp=tp, J=J, mu=mu, up_bias=up_bias)


def small_superblock(m, sys_len, env_len, seed, msys=24, menv=28, median=3, nsec=8):
    return synth.superblock(m, sys_len, env_len, msys, menv, seed=seed, median=median, nsec=nsec)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def opset(length, hdr, sub, off, data):
    from grand_sql_tj_amd.blocks import OpSet
    return OpSet(int(length), np.asarray(hdr, np.int32).reshape(-1, 4), np.asarray(sub, np.int32).reshape(-1, 5),
                 np.asarray(off, np.int64), np.asarray(data, np.float64))


def random_trun(basis, seed, frac=0.7, drop=0.15):
    """A truncation operator over the basis sectors: orthonormal columns,
    about frac * dim kept per sector, some sectors dropped."""
    rng = np.random.default_rng(seed)
    sub, off, chunks, nd = [], [], [], 0
    for k in range(basis.nsec):
        u, d, n = int(basis.sec[k][0]), int(basis.sec[k][1]), int(basis.sec[k][2])
        if basis.nsec > 2 and rng.random() < drop:
            continue
        c = max(1, int(round(frac * n)))
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        sub.append((u, d, 0, n, c))
        off.append(nd)
        chunks.append(q[:, :c].T.reshape(-1))
        nd += n * c
    return opset(basis.length, [[basis.length, 0, 0, len(sub)]], sub, off, np.concatenate(chunks))


def op_blocks(ops, op):
    """{(num_up, num_down, down_dif): matrix} of operator op."""
    first = int(ops.hdr[:op, 3].sum())
    out = {}
    for s in range(first, first + int(ops.hdr[op, 3])):
        r, c, o = int(ops.sub[s, 3]), int(ops.sub[s, 4]), int(ops.off[s])
        out[(int(ops.sub[s, 0]), int(ops.sub[s, 1]), int(ops.sub[s, 2]))] = \
            np.asarray(ops.data[o:o + r * c]).reshape(c, r).T
    return out
