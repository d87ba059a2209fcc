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


def wave_transform_case(direction, seed, up_bias=1, ls=5, le=6, m=24, median=3, nsec=8, frac=0.7, drop=0.15):
    """Consistent inputs of wave_trans_LR (direction 0) / wave_trans_RL (1):
    the bond (ls | . | . | le), the next bond of the sweep and the two
    truncation operators the sweep holds there.  Returns a dict with the
    arguments in the reference routine's order plus the bases of both
    bonds ("sys_old", "env_old", "sys_new", "env_new")."""
    from grand_sql_tj_amd.blocks import get_basis, get_super_basis, site_sectors
    from grand_sql_tj_amd.synth import SplitMix64, block_sectors
    rng = SplitMix64(seed)
    site = site_sectors(up_bias)
    fill = 0.875
    n_tot = 0
    if up_bias:
        n_tot = int(round(fill * (ls + le + 2)))
        n_tot -= n_tot % 2

    def secs(length):
        return block_sectors(rng, length, up_bias, fill * length, m, median, 0.6, 600, nsec, 6)

    def kept(trun):
        return np.array([[s[0], s[1], s[4]] for s in trun.sub], np.int32).reshape(-1, 3)

    if direction == 0:
        sys_old = get_basis(ls, secs(ls), site)
        t_sys = random_trun(sys_old, seed + 1, frac, drop)
        sys_new = get_basis(ls + 1, kept(t_sys), site)
        env_new = get_basis(le - 1, secs(le - 1), site)
        t_env = random_trun(env_new, seed + 2, frac, drop)
        env_old = get_basis(le, kept(t_env), site)
        sys_bs, env_bs = sys_new, env_old
    else:
        sys_new = get_basis(ls - 1, secs(ls - 1), site)
        t_sys = random_trun(sys_new, seed + 1, frac, drop)
        sys_old = get_basis(ls, kept(t_sys), site)
        env_old = get_basis(le, secs(le), site)
        t_env = random_trun(env_old, seed + 2, frac, drop)
        env_new = get_basis(le + 1, kept(t_env), site)
        sys_bs, env_bs = sys_old, env_new
    wo = get_super_basis(sys_old, env_old, n_tot, 0)
    wn = get_super_basis(sys_new, env_new, n_tot, 0)
    d_old = int((wo[:, 4].astype(np.int64) * wo[:, 5]).sum())
    x = np.random.default_rng(seed + 3).normal(size=d_old)
    x /= np.linalg.norm(x)
    return {"direction": direction, "sys_bs": sys_bs, "env_bs": env_bs, "sys_trun": t_sys, "env_trun": t_env,
            "wsec_old": wo, "x": x, "wsec_new": wn, "sys_old": sys_old, "env_old": env_old, "sys_new": sys_new,
            "env_new": env_new, "n_tot": n_tot}


def wave_args(case):
    return (case["direction"], case["sys_bs"], case["env_bs"], case["sys_trun"], case["env_trun"], case["wsec_old"],
            case["x"], case["wsec_new"])
