"""Site expectation values on the device: <n_i> over the lattice, as the
reference's measure_operator_dia computes it (code/tjsq/measure.f90:3-57,
called by Measurement, code/tjsq/main.f90:191-332).

Every site operator is carried from its site to the block next to the bond
(Get_sys_oper_dia / Get_env_oper_dia, code/tjsq/get_sys.f90:764-891): at
each block length it is embedded into the fused basis (update_block_dia),
the new site's operator is embedded beside it (update_site_dia), and all are
projected with that length's truncation operator (update_trun_dia).  The
reference carries each site's operator separately, O(N^2) transports per
side; here all operators of a side travel together, one device program per
block length (tjsq_dia_transport_create).  The contraction with the
wavefunction (measure_sys_block / _site, measure_env_block / _site,
code/tjsq/measure.f90:1241-1470) is one more program (tjsq_measure_create)
whose outputs y_j give <O_j> = psi . y_j.  The same vectors give the
two-point functions across the bond, <n_i n_j> = y_i . y_j for i on the
system side and j on the environment side (cross_correlation).

The inputs are what a run holds, or what its checkpoint files hold: the
fused bases by length (sysbs<L> / envbs<L>), the truncation operators by
length (systruns<L> / envtruns<L>) with their data on the device, and the
wavefunction with its sector list (eigvec1).
"""
from __future__ import annotations

import math

import numpy as np

from . import capi
from .blocks import OpSet


def num_elec(up_bias: int) -> OpSet:
    """num_elec of Get_site_operator_su2_U0 (dmrg_sub.f90:370-473): the
    reduced matrix element sqrt(2) on the occupied site, 0 on the hole, as a
    one-operator set of 1x1 sub-blocks."""
    return OpSet(1, np.array([[1, 0, 0, 2]], np.int32),
                 np.array([[0, 0, 0, 1, 1], [up_bias, 1, 0, 1, 1]], np.int32),
                 np.array([0, 1], np.int64), np.array([0.0, math.sqrt(2.0)]))


def side_vectors(side, site: OpSet, bases, truns, length, wsec, psi):
    """The vectors y_j (a (length, D) device tensor) with <O_j> = psi . y_j
    for the site operator on the `length` sites of one side of the bond:
    the block's sites 1..length-1 (counted from the block's outer end), then
    the site next to the bond.  Side 0: y_j = O_j psi / sqrt(1 + 2S) on the
    system index; side 1: psi O_j / sqrt(1 + 2S) on the environment index.
    bases[L] (L = 2..length) are the fused bases, truns[L] (L = 2..length-1)
    objects with `.ops` (layout) and `.data` (device tensor) of the
    truncation that made block L."""
    import torch
    if length < 2:
        raise ValueError("a side holds at least the block site and the bond site")
    dev = psi.device
    cur = site
    cur_data = torch.from_numpy(np.ascontiguousarray(site.data, np.float64)).to(dev)
    for L in range(2, length):
        t = truns[L]
        pg = capi.Program.dia_transport(cur, site, bases[L], t.ops)
        hdr, sub, off = pg.layout()
        n = max(1, pg.sizes["ndata"])
        out = torch.empty(n, dtype=torch.float64, device=dev)
        pg.run(t.data, cur_data, out)
        cur, cur_data = OpSet(L, hdr, sub, off, np.zeros(n)), out
    mp = capi.Program.measure(side, cur, site, bases[length], wsec)
    D = psi.numel()
    y = torch.zeros(max(1, mp.sizes["ndata"]), dtype=torch.float64, device=dev)
    mp.run(cur_data, psi, y)
    return y[:mp.sizes["ndata"]].view(-1, D)


def side_values(side, site: OpSet, bases, truns, length, wsec, psi):
    """<O> of the site operator on the `length` sites of one side of the
    bond, in the order of side_vectors."""
    return (side_vectors(side, site, bases, truns, length, wsec, psi) @ psi).cpu().numpy()


def density_profile(up_bias, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, psi):
    """measure_operator_dia (measure.f90:3-57) of num_elec: <n_i> for the
    sys_len + env_len lattice sites of a wavefunction whose system part is
    sys_len sites long (block + site) and environment part env_len.
    Environment block site i is lattice site N - i + 1."""
    site = num_elec(up_bias)
    n = sys_len + env_len
    ave = np.zeros(n)
    ave[:sys_len] = side_values(0, site, sys_bases, sys_truns, sys_len, wsec, psi)
    ev = side_values(1, site, env_bases, env_truns, env_len, wsec, psi)
    for i in range(1, env_len + 1):
        ave[n - i] = ev[i - 1]
    return ave


def cross_correlation(up_bias, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, psi):
    """<n_i n_j> for every system-side site i (rows, lattice sites
    1..sys_len) and environment-side site j (columns, lattice sites
    N - j + 1 for j = 1..env_len): the pairs across the bond of
    Get_oper_cor_dia (code/tjsq/measure.f90:468-634; sys_block_env_block,
    sys_block_env_site, sys_site_env_block and sys_site_env_site_cor_dia,
    :1638-2247).  Each of those contractions is
    sum over sectors and fused sub-sectors of c_i c_j <O_i psi, psi O_j>
    with one factor per side, so the whole block is the Gram matrix of the
    two sides' measurement vectors, Y_sys Y_env^T: one GEMM."""
    site = num_elec(up_bias)
    ys = side_vectors(0, site, sys_bases, sys_truns, sys_len, wsec, psi)
    ye = side_vectors(1, site, env_bases, env_truns, env_len, wsec, psi)
    return (ys @ ye.T).cpu().numpy()


def _trivial_basis(basis):
    """The fused basis read as a block basis of one sub-sector per sector,
    for measuring operators that already live in the fused basis."""
    from .blocks import Basis
    sec = np.asarray(basis.sec)
    return Basis(basis.length, np.stack([sec[:, 0], sec[:, 1], sec[:, 2], np.ones(len(sec), np.int32)], 1)
                 .astype(np.int32),
                 np.stack([sec[:, 0], sec[:, 1], np.zeros(len(sec), np.int32), np.zeros(len(sec), np.int32),
                           np.zeros(len(sec), np.int32), sec[:, 2]], 1).astype(np.int32))


def side_pair_vectors(side, site: OpSet, bases, truns, length, wsec, psi):
    """Measurement vectors of every single-site operator n_i and every
    same-side product n_i n_j (i < j) of one side, all carried to the
    wavefunction's fused basis: returns (Y, labels), Y a (len(labels), D)
    device tensor with <O> = psi . y, labels (i,) for singles and (i, j)
    for pairs, sites counted from the side's outer end (the bond site is
    `length`).  Each product is formed when its second site joins the block,
    as a reduced operator (tjsq_dia_transport_pairs_create), then carried on
    like the singles."""
    import torch
    dev = psi.device
    cur = site
    cur_data = torch.from_numpy(np.ascontiguousarray(site.data, np.float64)).to(dev)
    singles, pairs = [(1,)], []
    for L in range(2, length + 1):
        if L < length:
            t_ops, t_data = truns[L].ops, truns[L].data
        else:                      # the fused basis of the wavefunction itself: no truncation
            tsub, toff, nd = capi.identity_truncation(bases[L])
            t_data = torch.empty(max(1, nd), dtype=torch.float64, device=dev)
            capi.identity_truncation(bases[L], t_data)
            t_ops = OpSet(L, np.array([[L, 0, 0, tsub.shape[0]]], np.int32), tsub.astype(np.int32),
                          toff.astype(np.int64), np.zeros(max(1, nd)))
        pg = capi.Program.dia_transport_pairs(cur, len(singles), site, bases[L], t_ops)
        hdr, sub, off = pg.layout()
        n = max(1, pg.sizes["ndata"])
        out = torch.empty(n, dtype=torch.float64, device=dev)
        pg.run(t_data, cur_data, out)
        pairs = pairs + [(i[0], L) for i in singles]
        singles = singles + [(L,)]
        cur, cur_data = OpSet(L, hdr, sub, off, np.zeros(n)), out
    mp = capi.Program.measure(side, cur, None, _trivial_basis(bases[length]), wsec)
    D = psi.numel()
    y = torch.zeros(max(1, mp.sizes["ndata"]), dtype=torch.float64, device=dev)
    mp.run(cur_data, psi, y)
    Y = y[:mp.sizes["ndata"]].view(-1, D)
    return Y[:-1], singles + pairs


def correlation_matrix(up_bias, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, psi):
    """<n_i n_j> for all pairs of the sys_len + env_len lattice sites (lattice
    order; environment site j from the outer end is lattice N - j + 1):
    - across the bond, the Gram matrix of the two sides' single-site vectors
      (the entries of the reference's Get_oper_cor_dia, measure.f90:468-634);
    - on one side, the carried products n_i n_j, formed with their SU(2)
      factor (where the reference's block_mul_block_dia leaves one factor
      sqrt(1 + 2S) too many, those entries here are the physical ones);
    - on the diagonal <n_i^2> = <n_i> (no double occupancy)."""
    site = num_elec(up_bias)
    n = sys_len + env_len
    cor = np.zeros((n, n))
    lat = [lambda i: i - 1, lambda j: n - j]
    vecs = []
    for side, bases, truns, length in ((0, sys_bases, sys_truns, sys_len), (1, env_bases, env_truns, env_len)):
        Y, labels = side_pair_vectors(side, site, bases, truns, length, wsec, psi)
        vals = (Y @ psi).cpu().numpy()
        f = lat[side]
        for lab, v in zip(labels, vals):
            if len(lab) == 1:
                cor[f(lab[0]), f(lab[0])] = v
            else:
                cor[f(lab[0]), f(lab[1])] = cor[f(lab[1]), f(lab[0])] = v
        vecs.append(Y[:length])
    cross = (vecs[0] @ vecs[1].T).cpu().numpy()
    for i in range(sys_len):
        for j in range(1, env_len + 1):
            cor[i, n - j] = cor[n - j, i] = cross[i, j - 1]
    return cor


# ---- correlations of rank-k site operators (Get_density_cor) ------------------------------

def spin_sd(up_bias: int) -> OpSet:
    """st_sd of Get_site_operator_su2_U0 (dmrg_sub.f90:370-473): the spin
    vector S, rank 1 (down_dif 2), reduced element sqrt(3/2) on the singly
    occupied site."""
    return OpSet(1, np.array([[1, 0, 2, 1]], np.int32), np.array([[up_bias, 1, 0, 1, 1]], np.int32),
                 np.array([0], np.int64), np.array([math.sqrt(1.5)]))


def elec_up(up_bias: int) -> OpSet:
    """st_elec_up: the creation operator c^+, rank 1/2 (down_dif 1), from the
    hole to the singly occupied site, reduced element -sqrt(2)."""
    return OpSet(1, np.array([[1, up_bias, 1, 1]], np.int32), np.array([[0, 0, 1, 1, 1]], np.int32),
                 np.array([0], np.int64), np.array([-math.sqrt(2.0)]))


def elec_down(up_bias: int) -> OpSet:
    """st_elec_down: the annihilation operator c, rank 1/2, from the singly
    occupied site to the hole, reduced element -sqrt(2)."""
    return OpSet(1, np.array([[1, -up_bias, 1, 1]], np.int32), np.array([[up_bias, 1, -1, 1, 1]], np.int32),
                 np.array([0], np.int64), np.array([-math.sqrt(2.0)]))


# the three calls of Get_density_cor in Measurement (main.f90:274-299):
# (first operator, second operator, sign flag)
COR_KINDS = {"density": (num_elec, num_elec, "B"), "spin": (spin_sd, spin_sd, "B"),
             "hopping": (elec_up, elec_down, "F")}


class DeviceRunner:
    """Runs programs on torch device tensors (the product path).  The
    correlation driver takes any object with these five methods, so that the
    same driver can execute the programs' task tables elsewhere."""

    def __init__(self, device):
        self.device = device

    def put(self, a):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(self.device)

    def run(self, pg, ops, x, clear=False):
        import torch
        n = max(1, pg.sizes["ndata"])
        y = (torch.zeros if clear else torch.empty)(n, dtype=torch.float64, device=self.device)
        pg.run(ops, x, y)
        return y

    def cat(self, a, b):
        import torch
        return torch.cat([a, b])

    def head(self, a, n):
        return a[:n]

    def values(self, y, nvec, psi):
        D = psi.numel()
        return (y[:nvec * D].view(nvec, D) @ psi).cpu().numpy()


def _extent(ops: OpSet) -> int:
    sub, off = np.asarray(ops.sub).reshape(-1, 5), np.asarray(ops.off)
    return int((off + sub[:, 3].astype(np.int64) * sub[:, 4]).max()) if off.size else 0


def _concat(a: OpSet, ad, b: OpSet, bd, runner):
    """Operator sets a and b as one (a's operators first) with their data."""
    na = _extent(a)
    hdr = np.concatenate([np.asarray(a.hdr).reshape(-1, 4), np.asarray(b.hdr).reshape(-1, 4)]).astype(np.int32)
    sub = np.concatenate([np.asarray(a.sub).reshape(-1, 5), np.asarray(b.sub).reshape(-1, 5)]).astype(np.int32)
    off = np.concatenate([np.asarray(a.off), np.asarray(b.off) + na]).astype(np.int64)
    data = runner.cat(runner.head(ad, na), runner.head(bd, _extent(b)))
    return OpSet(b.length, hdr, sub, off, np.zeros(na + _extent(b))), data


def _step(pg, length, runner, ops_data, x):
    """Run a transport-like program; returns (output OpSet layout, data)."""
    hdr, sub, off = pg.layout()
    y = runner.run(pg, ops_data, x)
    return OpSet(length, hdr, sub, off, np.zeros(max(1, pg.sizes["ndata"]))), y


def pair_correlations(kind, up_bias, sid1, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, psi,
                      runner=None):
    """The two-point functions of Get_density_cor (code/tjsq/measure.f90:
    3110-3223) between system-side site sid1 and every later site of the
    sys_len + env_len sites: returns cor (lattice order, entry sid - 1; NaN
    for sid <= sid1), each value scaled as that routine scales it
    (sqrt(1 + 2k), negated for k = 1).  kind (COR_KINDS): "density"
    <n_i n_j>, "spin" <S_i . S_j>, "hopping" the spin-summed
    <c^+_i c_j>.

    The operator of site sid1 is formed at its block length and carried to
    the block next to the bond (Get_sys_oper_ndia, Change_sys_oper_ndia,
    code/tjsq/get_sys.f90:998-1238), one ndia transport program per length.
    On the way, at every length L the product with site L's operator is
    formed (Get_sys2b, measure.f90:2785-2819: the rank-0 coupling of
    block_mul_block_ndia, projected) and carried on as a diagonal operator;
    all products travel together and are measured at the bond
    (measure_sys_block).  The environment's operators of every site travel
    together too (the reference carries each one separately), and one
    correlation program contracts the system operator with all of them, the
    environment bond site and the system bond site (sys_block_env_block_,
    sys_block_env_site_ and sys_block_site_cor_ndia)."""
    if runner is None:
        runner = DeviceRunner(psi.device)
    f1, f2, flag = COR_KINDS[kind]
    o1, o2 = f1(up_bias), f2(up_bias)
    k1, k2 = int(o1.hdr[0][2]), int(o2.hdr[0][2])
    if not 1 <= sid1 <= sys_len - 1 or env_len < 2:
        raise ValueError("sid1 must lie in the system block (1..sys_len-1) and the environment hold a block")
    n = sys_len + env_len
    scale = math.sqrt(1.0 + k1) * (-1.0 if k1 == 2 else 1.0)
    cor = np.full(n, np.nan)
    one = runner.put(np.zeros(1))
    # the system operator at block length sid1
    if sid1 == 1:
        cur, cd = o1, runner.put(o1.data)
    else:
        t = sys_truns[sid1]
        cur, cd = _step(capi.Program.ndia_transport(0, flag, None, o1, sys_bases[sid1], t.ops), sid1, runner,
                        t.data, one)
    prods, pdata, labels = None, None, []
    for L in range(sid1 + 1, sys_len):
        t = sys_truns[L]
        new, nd = _step(capi.Program.pair_ndia(0, flag, cur, o2, sys_bases[L], t.ops), L, runner, t.data, cd)
        if prods is not None:
            prods, pdata = _step(capi.Program.dia_transport(prods, None, sys_bases[L], t.ops), L, runner, t.data,
                                 pdata)
            prods, pdata = _concat(prods, pdata, new, nd, runner)
        else:
            prods, pdata = new, nd
        labels.append(L)
        cur, cd = _step(capi.Program.ndia_transport(0, flag, cur, None, sys_bases[L], t.ops), L, runner, t.data, cd)
    if prods is not None:
        mp = capi.Program.measure(0, prods, None, sys_bases[sys_len], wsec)
        vals = runner.values(runner.run(mp, pdata, psi, clear=True), len(labels), psi)
        for sid2, v in zip(labels, vals):
            cor[sid2 - 1] = scale * v
    # the environment operators of sites 1..env_len-1 (counted from the outer end)
    ecur, ed = o2, runner.put(o2.data)
    for L in range(2, env_len):
        t = env_truns[L]
        ecur, ed = _step(capi.Program.ndia_transport(1, flag, ecur, o2, env_bases[L], t.ops), L, runner, t.data, ed)
    ne = env_len - 1
    cp = capi.Program.cor_ndia(flag, cur, ecur, k2 != 0, o2, k2 == 1, o2, k2 == 1, sys_bases[sys_len],
                               env_bases[env_len], wsec)
    sys_ext = _extent(cur)
    ops = runner.cat(runner.head(cd, sys_ext), runner.head(ed, _extent(ecur)))
    vals = runner.values(runner.run(cp, ops, psi, clear=True), ne + 2, psi)
    for q in range(ne):
        cor[n - q - 1] = scale * vals[q]
    cor[sys_len] = scale * vals[ne]
    cor[sys_len - 1] = scale * vals[ne + 1]
    return cor


def density_cor(kind, up_bias, xi, xj, ny, ij, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec,
                psi, runner=None):
    """Get_density_cor's matrix oper_ndia (measure.f90:3110-3223) on the
    sys_len + env_len sites: for every sid1 in xi..xj the same-side sites
    sid1 + ny + ij, + 2 ny, ... below sys_len and the sites from sys_len on
    whose distance from sid1 + ij is a multiple of ny (the routine's
    mod(sid1 + ij - sid3, ny) >= 0 test); other entries 0."""
    n = sys_len + env_len
    out = np.zeros((n, n))
    for sid1 in range(xi, xj + 1):
        cor = pair_correlations(kind, up_bias, sid1, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns,
                                wsec, psi, runner)
        for sid2 in range(sid1 + ny + ij, sys_len, ny):
            out[sid1 - 1, sid2 - 1] = cor[sid2 - 1]
        for sid3 in range(sys_len, n + 1):
            a = sid1 + ij - sid3
            if a >= 0 or a % ny == 0:
                out[sid1 - 1, sid3 - 1] = cor[sid3 - 1]
    return out


class _DeviceTrun:
    def __init__(self, ops, data):
        self.ops, self.data = ops, data


def load_run(directory, device="cuda"):
    """What Measurement reads from a run's checkpoint files
    (code/tjsq/main.f90:213-238): eigvec1, the truncation operators
    systruns<L> / envtruns<L> (data moved to the device) and the fused bases
    sysbs<L> / envbs<L>.  Returns a dict with sys_len, env_len, wsec, psi,
    sys_bases, env_bases, sys_truns, env_truns, ready for density_profile /
    correlation_matrix."""
    import os
    import re
    import torch
    from . import checkpoint as ck
    w = ck.load_wave(os.path.join(directory, ck.wave_file(1)))
    out = {"sys_len": w.sys_len, "env_len": w.env_len, "wsec": w.wsec,
           "psi": torch.from_numpy(w.x).to(device), "sys_bases": {}, "env_bases": {}, "sys_truns": {}, "env_truns": {}}
    pat = re.compile(r"^(sys|env)(truns|bs)([0-9]+)$")
    for name in sorted(os.listdir(directory)):
        m = pat.match(name)
        if not m:
            continue
        side, L = m.group(1), int(m.group(3))
        path = os.path.join(directory, name)
        if m.group(2) == "bs":
            out[side + "_bases"][L], _ = ck.load_basis(path)
        else:
            t, _ = ck.load_trun(path)
            data = torch.from_numpy(t.data if t.data.size else np.zeros(1)).to(device)
            out[side + "_truns"][L] = _DeviceTrun(OpSet(L, t.hdr, t.sub, t.off, np.zeros(max(1, t.data.size))), data)
    return out


def measure_run(directory, up_bias=1):
    """The density profile and the whole <n_i n_j> matrix of the ground state
    stored in a checkpoint directory (the reference's or one written by
    DMRG.save_checkpoint), as Measurement computes them from the same files."""
    r = load_run(directory)
    args = (up_bias, r["sys_len"], r["env_len"], r["sys_bases"], r["env_bases"], r["sys_truns"], r["env_truns"],
            r["wsec"], r["psi"])
    return density_profile(*args), correlation_matrix(*args)
