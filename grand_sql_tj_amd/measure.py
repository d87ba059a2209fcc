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
        # the operator set grows every step (L singles, L (L - 1) / 2 pairs):
        # return the previous step's buffers and the program's scratch, so
        # that cached blocks of the smaller earlier sizes do not pile up
        # (a 70-site side at m = 2000 otherwise ran out of device memory)
        del pg, t_data
        torch.cuda.empty_cache()
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


# ---- pairing: Get_SC_Order and Get_mSCOP_Operator_Cor0 -----------------------------------

def _select(ops: OpSet, idx) -> OpSet:
    """The operators idx of ops, reading the same data array (offsets kept)."""
    hdr = np.asarray(ops.hdr).reshape(-1, 4)
    sub, off = np.asarray(ops.sub).reshape(-1, 5), np.asarray(ops.off)
    first = np.concatenate([[0], np.cumsum(hdr[:, 3])])
    rows = [np.arange(first[i], first[i + 1]) for i in idx]
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    return OpSet(ops.length, hdr[list(idx)].astype(np.int32), sub[rows].astype(np.int32),
                 off[rows].astype(np.int64), ops.data)


def _carry(side, flag, ops, od, L0, L1, bases, truns, runner):
    """Change_sys_oper_ndia / Change_env_oper_ndia with Truns (get_sys.f90:
    1134-1238): ops at block length L0 carried to block length L1 >= L0,
    embedded into each fused basis and projected with its truncation."""
    for L in range(L0 + 1, L1 + 1):
        t = truns[L]
        ops, od = _step(capi.Program.ndia_transport(side, flag, ops, None, bases[L], t.ops), L, runner, t.data, od)
    return ops, od


def _pair_chain(side, up_bias, length, bases, truns, pairs, runner):
    """The creation operators c^+_i of the sites i = 1 .. length - 1 of one
    side and the singlet pair operators [c^+_i c^+_j]^0 of the requested
    (i, j), i < j <= length - 1, all at block length length - 1: the
    operators of Get_sys_two (measure.f90:3798-3832) and
    Get_SCOP_Operator_Env (:6211-6263).  c^+_i is the site operator at
    length 1 and otherwise formed at length i (update_site_ndia_*, projected)
    and carried with the fermion signs ('F'); the pair is formed when site j
    joins (block_mul_block_ndia with down_dif 0 of c^+_i at length j - 1 and
    the site's c^+, projected) and carried as a boson ('B').  All operators
    of one kind travel together, one program per block length.  Returns
    (cdag, cdag_data, pair_ops, pair_data, pair_labels); pair_ops is None
    when no pair was requested."""
    cd = elec_up(up_bias)
    want = {}
    for i, j in pairs:
        if not 1 <= i < j <= length - 1:
            raise ValueError(f"pair ({i}, {j}) does not lie in the block (sites 1..{length - 1})")
        want.setdefault(j, []).append(i)
    cur, cdat = cd, runner.put(cd.data)          # c^+_1 at length 1
    pr, pdat, labels = None, None, []
    for L in range(2, length):
        t = truns[L]
        if pr is not None:
            pr, pdat = _step(capi.Program.ndia_transport(side, "B", pr, None, bases[L], t.ops), L, runner, t.data,
                             pdat)
        firsts = sorted(set(want.get(L, [])))
        if firsts:
            sel = _select(cur, [i - 1 for i in firsts])
            new, nd = _step(capi.Program.product_ndia(side, "F", 0, sel, cd, bases[L], t.ops), L, runner, t.data,
                            cdat)
            if pr is None:
                pr, pdat = new, nd
            else:
                pr, pdat = _concat(pr, pdat, new, nd, runner)
            labels += [(i, L) for i in firsts]
        cur, cdat = _step(capi.Program.ndia_transport(side, "F", cur, cd, bases[L], t.ops), L, runner, t.data, cdat)
    return cur, cdat, pr, pdat, labels


def _measure_ops(side, ops, od, basis, wsec, psi, runner):
    """measure_sys_block / measure_env_block of every operator of ops."""
    mp = capi.Program.measure(side, ops, None, basis, wsec)
    return runner.values(runner.run(mp, od, psi, clear=True), ops.nops, psi)


def _check_bond(sys_len, env_len):
    if sys_len < 3 or env_len < 3:
        raise ValueError("the pairing measurements need at least two block sites on each side of the bond")


def sc_order(up_bias, neib, ny, xc1, xc2, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, psi,
             runner=None):
    """Get_SC_Order (code/tjsq/measure.f90:3226-3376): the singlet pair
    amplitude <[c^+_i c^+_j]^0> of the bonds of sites sid1 = xc1, xc1 + ny,
    ... <= xc2 with their neighbours neib[sid1, 0] (direction 1, along y) and
    neib[sid1, 2] (direction 3, along x), on a chain of n = sys_len +
    env_len sites (environment site e is lattice site n - e + 1).  Returns
    rows (sid1, neib[sid1, 2], value of direction 1, value of direction 3),
    the first four columns the routine writes to Sc_order.dat (its last two
    columns are never assigned there).

    Each bond takes the routine's branch for where its sites lie: both in
    the system block, the pair operator (Get_sys_two) measured with
    measure_sys_block; both in the environment block, Get_SCOP_Operator_Env
    with measure_env_block; across the bond, the contraction of c^+ with the
    transposed c^+ (or with st_elec_down at a bond site) of the other side
    (sys_site_env_site_, sys_site_env_block_, env_block_site_ (negated),
    sys_block_site_, sys_block_env_site_ and sys_block_env_block_cor_ndia).
    The amplitude connects sectors of different particle number, which a
    wavefunction has only without the particle label: up_bias must be 0, as
    the reference configures it (common.f90:11)."""
    if up_bias != 0:
        raise ValueError("the pair amplitude needs the SU(2)-only basis (up_bias 0)")
    _check_bond(sys_len, env_len)
    if runner is None:
        runner = DeviceRunner(psi.device)
    n = sys_len + env_len
    neib = np.asarray(neib)
    sids = list(range(xc1, xc2 + 1, ny))
    bonds = []
    for sid1 in sids:
        for ij in (1, 3):
            sid2 = int(neib[sid1, ij - 1])
            if not 1 <= sid1 <= n or not 1 <= sid2 <= n:
                raise ValueError(f"site {sid1} has no neighbour in direction {ij} on the {n}-site chain")
            bonds.append((sid1, ij, min(sid1, sid2), max(sid1, sid2)))
    sys_pairs = {(a, b) for _, _, a, b in bonds if b < sys_len}
    env_pairs = {(n - b + 1, n - a + 1) for _, _, a, b in bonds if a > sys_len + 1 and n - a + 1 <= env_len - 1}
    scd, scdat, sp, spdat, slab = _pair_chain(0, up_bias, sys_len, sys_bases, sys_truns, sys_pairs, runner)
    ecd, ecdat, ep, epdat, elab = _pair_chain(1, up_bias, env_len, env_bases, env_truns, env_pairs, runner)
    sb, eb = sys_bases[sys_len], env_bases[env_len]
    val = {}
    if sp is not None:
        for lab, v in zip(slab, _measure_ops(0, sp, spdat, sb, wsec, psi, runner)):
            val[("s",) + lab] = v
    if ep is not None:
        for lab, v in zip(elab, _measure_ops(1, ep, epdat, eb, wsec, psi, runner)):
            val[("e",) + lab] = v
    up, dn = elec_up(up_bias), elec_down(up_bias)
    one = runner.put(np.zeros(1))
    # across the bond: environment c^+ of site e sits at index e - 1 of ecd
    cross_e = sorted({n - b + 1 for _, _, a, b in bonds if a >= sys_len and b > sys_len + 1} |
                     {n - b + 1 for _, _, a, b in bonds if a < sys_len and b > sys_len + 1})
    e_ops = _select(ecd, [e - 1 for e in cross_e]) if cross_e else None
    epos = {e: q for q, e in enumerate(cross_e)}
    for sid1, ij, a, b in bonds:
        key = (sid1, ij)
        if b < sys_len:
            val[key] = val[("s", a, b)]
        elif a > sys_len + 1:
            val[key] = val[("e", n - b + 1, n - a + 1)]
        elif a == sys_len and b == sys_len + 1:
            pg = capi.Program.site_cor(0, "F", up, 0, None, 0, dn, 0, sb, eb, wsec)
            val[key] = runner.values(runner.run(pg, one, psi, clear=True), 1, psi)[0]
        elif a == sys_len:
            pg = capi.Program.site_cor(1, "F", up, 0, _select(e_ops, [epos[n - b + 1]]), 1, None, 0, sb, eb, wsec)
            val[key] = runner.values(runner.run(pg, ecdat, psi, clear=True), 1, psi)[0]
        elif a == sys_len + 1:
            pg = capi.Program.site_cor(2, "F", None, 0, _select(e_ops, [epos[n - b + 1]]), 1, up, 0, sb, eb, wsec)
            val[key] = -runner.values(runner.run(pg, ecdat, psi, clear=True), 1, psi)[0]
        else:                                  # a < sys_len <= b
            sop = _select(scd, [a - 1])
            if b == sys_len:
                pg = capi.Program.cor_ndia("F", sop, None, 0, None, 0, dn, 0, sb, eb, wsec)
                ops = scdat
            elif b == sys_len + 1:
                pg = capi.Program.cor_ndia("F", sop, None, 0, dn, 0, None, 0, sb, eb, wsec)
                ops = scdat
            else:
                eo = _select(e_ops, [epos[n - b + 1]])
                pg = capi.Program.cor_ndia("F", sop, eo, 1, None, 0, None, 0, sb, eb, wsec)
                ops = runner.cat(runner.head(scdat, _extent(sop)), runner.head(ecdat, _extent(eo)))
            val[key] = runner.values(runner.run(pg, ops, psi, clear=True), 1, psi)[0]
    return [(sid1, int(neib[sid1, 2]), val[(sid1, 1)], val[(sid1, 3)]) for sid1 in sids]


def sc_pair_cor0(up_bias, neib, ny, xc1, ij, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec,
                 psi, runner=None):
    """Get_mSCOP_Operator_Cor0 (code/tjsq/measure.f90:3377-3507) with
    xc1 = xc2: the correlations of the singlet pair operator of bond
    (sid1, neib[sid1, ij - 1]), sid1 = xc1, with the pair operators of the
    bonds (sid3, neib[sid3, i - 1]), i = 1, 3, 5, further along the chain
    (n = sys_len + env_len sites).  Returns the rows the routine writes to its
    file: (sid1, sid3, mod(sid3 - sid1 + n, ny), (sid3 - 1)//ny -
    (sid1 - 1)//ny, v1, v3, v5) (its last three columns come from a table
    the reference never fills and are 0):
    - sid3 = sid1 + ny, ... <= sys_len - 3 on the system side: the
      four-fermion operator of Get_sys_two1 (:3835-3889: the pair carried to
      min(sid3, sid4), coupled with st_elec_down to 2S = 1, carried, coupled
      with st_elec_down to 2S = 0, carried as a scalar) measured with
      measure_sys_block; bonds reaching back to the first pair or past
      sys_len - 1 stay 0;
    - when ny >= 6, the bond of sid1 + 4 with sid3 = sys_len + 2 ..
      sys_len + ny + 1 in line with it, and then the first bond with
      sid3 = sys_len + 2 .. n - ny in line with it, on the environment side:
      sys_block_env_block_cor_ndia of the two pair operators
      (Get_SCOP_Operator_Env), bonds that leave the environment block stay 0.
    Valid for both particle labels."""
    _check_bond(sys_len, env_len)
    if runner is None:
        runner = DeviceRunner(psi.device)
    n = sys_len + env_len
    neib = np.asarray(neib)
    dn = elec_down(up_bias)
    sb, eb = sys_bases[sys_len], env_bases[env_len]

    def row(s1, s3, v):
        return (s1, s3, (s3 - s1 + n) % ny, (s3 - 1) // ny - (s1 - 1) // ny, v[0], v[1], v[2])

    def sys_pair(s1):
        s2 = int(neib[s1, ij - 1])
        if s2 == 0 or not s1 < s2 <= sys_len - 1:
            raise ValueError(f"the bond ({s1}, {s2}) must point forward inside the system block")
        return s1, s2

    s1, s2 = sys_pair(xc1)
    _, _, pr, pdat, _ = _pair_chain(0, up_bias, s2 + 1, sys_bases, sys_truns, [(s1, s2)], runner)
    rows = []
    # (a) both pairs on the system side
    for s3 in range(s1 + ny, sys_len - 2, ny):
        v = [0.0, 0.0, 0.0]
        for k, i in enumerate((1, 3, 5)):
            s4 = int(neib[s3, i - 1])
            if min(s3, s4) <= max(s1, s2) or max(s3, s4) > sys_len - 1:
                continue
            i3, i4 = min(s3, s4), max(s3, s4)
            q, qd = _carry(0, "B", pr, pdat, s2, i3 - 1, sys_bases, sys_truns, runner)
            t = sys_truns[i3]
            q, qd = _step(capi.Program.product_ndia(0, "F", 1, q, dn, sys_bases[i3], t.ops), i3, runner, t.data, qd)
            q, qd = _carry(0, "F", q, qd, i3, i4 - 1, sys_bases, sys_truns, runner)
            t = sys_truns[i4]
            q, qd = _step(capi.Program.product_ndia(0, "F", 0, q, dn, sys_bases[i4], t.ops), i4, runner, t.data, qd)
            for L in range(i4 + 1, sys_len):
                t = sys_truns[L]
                q, qd = _step(capi.Program.dia_transport(q, None, sys_bases[L], t.ops), L, runner, t.data, qd)
            v[k] = _measure_ops(0, q, qd, sb, wsec, psi, runner)[0]
        rows.append(row(s1, s3, v))
    # (b), (c) the second pair on the environment side
    parts = []
    if ny >= 6:
        parts.append((sys_pair(xc1 + 4), range(sys_len + 2, sys_len + ny + 2)))
    parts.append(((s1, s2), range(sys_len + 2, n - ny + 1)))
    work = []
    for (a1, a2), s3s in parts:
        bl = []
        for s3 in s3s:
            if (s3 - a1) % ny != 0:
                continue
            if s3 > n:
                raise ValueError(f"site {s3} lies past the end of the {n}-site chain")
            e1 = n - s3 + 1
            es = []
            for i in (1, 3, 5):
                s4 = int(neib[s3, i - 1])
                e2 = n - s4 + 1
                es.append((min(e1, e2), max(e1, e2)) if s4 != 0 and max(e1, e2) <= env_len - 1 else None)
            bl.append((s3, es))
        work.append(((a1, a2), bl))
    env_pairs = {p for _, bl in work for _, es in bl for p in es if p is not None}
    if env_pairs:
        _, _, ep, epdat, elab = _pair_chain(1, up_bias, env_len, env_bases, env_truns, env_pairs, runner)
    for (a1, a2), bl in work:
        _, _, sp, spdat, _ = _pair_chain(0, up_bias, a2 + 1, sys_bases, sys_truns, [(a1, a2)], runner)
        sp, spdat = _carry(0, "B", sp, spdat, a2, sys_len - 1, sys_bases, sys_truns, runner)
        vals = {}
        if env_pairs:
            pg = capi.Program.cor_ndia("B", sp, ep, 0, None, 0, None, 0, sb, eb, wsec)
            ops = runner.cat(runner.head(spdat, _extent(sp)), runner.head(epdat, _extent(ep)))
            vals = dict(zip(elab, runner.values(runner.run(pg, ops, psi, clear=True), ep.nops, psi)))
        for s3, es in bl:
            rows.append(row(a1, s3, [vals[p] if p is not None else 0.0 for p in es]))
    return rows


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
