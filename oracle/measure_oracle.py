"""numpy restatement of the reference's density measurement (TEST
INFRASTRUCTURE ONLY: the checker of grand_sql_tj_amd.measure; only tests/
import it).

measure_operator_dia (code/tjsq/measure.f90:3-57) with the site operator
num_elec: every site operator n_i carried from its site to the block next to
the measured bond (Get_sys_oper_dia / Get_env_oper_dia,
code/tjsq/get_sys.f90:764-891, built from update_site_dia, update_block_dia
and update_trun_dia, code/tjsq/block.f90:382-498 and 1311-1358), then
contracted with the wavefunction (measure_sys_block, measure_sys_site,
measure_env_block, measure_env_site, code/tjsq/measure.f90:1241-1470).

Blocks are lists of (num_up, num_down, matrix) in the reference's sub-block
order; bases and truncation operators come in the library's flat layouts
(Basis: sec (nsec, 4), sub (nsub, 6); truncation OpSet with one operator).
Pure-Python loops over sectors: small cases only.
"""
from __future__ import annotations

import math

import numpy as np


def _basis_secs(bs):
    """[(new_up, new_down, dim, [(bl_up, bl_down, st_up, st_down, spos, sdim), ...]), ...]"""
    out, q = [], 0
    for up, down, dim, num in (tuple(int(v) for v in r) for r in np.asarray(bs.sec)):
        subs = [tuple(int(v) for v in r) for r in np.asarray(bs.sub)[q:q + num]]
        q += num
        out.append((up, down, dim, subs))
    return out


def _trun_blocks(trun):
    out = []
    for s in range(int(trun.hdr[0, 3])):
        u, d, _, r, c = (int(v) for v in trun.sub[s])
        o = int(trun.off[s])
        out.append((u, d, np.asarray(trun.data[o:o + r * c]).reshape(c, r).T))
    return out


def site_num_elec(up_bias):
    """num_elec of Get_site_operator_su2_U0 (dmrg_sub.f90:370-473): the
    reduced matrix element sqrt(2) on the occupied site, 0 on the hole."""
    return [(0, 0, np.zeros((1, 1))), (up_bias, 1, np.full((1, 1), math.sqrt(2.0)))]


def update_site_dia(site, bs):
    """update_site_dia (block.f90:437-498)."""
    out = []
    for up, down, dim, subs in _basis_secs(bs):
        m = np.zeros((dim, dim))
        for bu, bd, su, sd, spos, sdim in subs:
            for u, d, mat in site:
                if u == su and d == sd:
                    c = math.sqrt((1.0 + down) / (1.0 + d))
                    m[spos:spos + sdim, spos:spos + sdim] += np.eye(sdim) * mat[0, 0] * c
                    break
        out.append((up, down, m))
    return out


def update_block_dia(block, bs):
    """update_block_dia (block.f90:382-432)."""
    out = []
    for up, down, dim, subs in _basis_secs(bs):
        m = np.zeros((dim, dim))
        for bu, bd, su, sd, spos, sdim in subs:
            for u, d, mat in block:
                if u == bu and d == bd:
                    c = math.sqrt((1.0 + down) / (1.0 + d))
                    m[spos:spos + sdim, spos:spos + sdim] += mat[:sdim, :sdim] * c
                    break
        out.append((up, down, m))
    return out


def update_trun_dia(op, trun):
    """update_trun_dia (block.f90:1311-1358)."""
    out = []
    for u, d, t in _trun_blocks(trun):
        m = np.zeros((t.shape[1], t.shape[1]))
        for ou, od, mat in op:
            if ou == u and od == d:
                m = t.T @ (mat @ t)
                break
        out.append((u, d, m))
    return out


def transport(site, i, bases, truns, new_idx):
    """Get_sys_oper_dia / Get_env_oper_dia (get_sys.f90:764-891) with
    trun_idx = 1 and Truns = .true.: the operator of site i of the block in
    the basis of the block of length new_idx (>= i >= 1; new_idx >= 2).
    bases[L] / truns[L] are sys_bsm(L) / systruns(L) (or the env ones).
    At new_idx = 1 the block is the site itself and is returned as it is
    (the reference would project it with systruns(1), which Measurement
    never reads from disk, main.f90:220-228)."""
    if new_idx == 1:
        return site
    if i == new_idx:
        op = update_site_dia(site, bases[i]) if i > 1 else site
        return update_trun_dia(op, truns[new_idx])
    if i <= 1:
        op = site
    else:
        op = update_trun_dia(update_site_dia(site, bases[i]), truns[i])
    idx = i
    while idx + 1 <= new_idx - 1:
        op = update_trun_dia(update_block_dia(op, bases[idx + 1]), truns[idx + 1])
        idx += 1
    op = update_block_dia(op, bases[idx + 1])
    return update_trun_dia(op, truns[new_idx])


def _wave_blocks(wsec, x):
    out, o = [], 0
    for r in np.asarray(wsec):
        su, sd, eu, ed, ns, ne = (int(v) for v in r)
        out.append((su, sd, eu, ed, np.asarray(x[o:o + ns * ne]).reshape(ne, ns).T))
        o += ns * ne
    return out


def _secs_by_qn(bs):
    return {(up, down): subs for up, down, _, subs in _basis_secs(bs)}


def measure_block(side, op, bs, wsec, x):
    """measure_sys_block (side 0, measure.f90:1241-1295) / measure_env_block
    (side 1, :1359-1414)."""
    secs = _secs_by_qn(bs)
    val = 0.0
    for su, sd, eu, ed, v in _wave_blocks(wsec, x):
        subs = secs.get((su, sd) if side == 0 else (eu, ed), [])
        for bu, bd, _, _, spos, sdim in subs:
            for u, d, mat in op:
                if u == bu and d == bd:
                    c = 1.0 / math.sqrt(1.0 + bd)
                    if side == 0:
                        vx = v[spos:spos + sdim, :]
                        val += float(np.sum(vx * (c * mat @ vx)))
                    else:
                        vx = v[:, spos:spos + sdim]
                        val += float(np.sum((vx @ mat) * vx)) * c
                    break
    return val


def measure_site(side, site, bs, wsec, x):
    """measure_sys_site (side 0, measure.f90:1299-1355) / measure_env_site
    (side 1, :1418-1470)."""
    secs = _secs_by_qn(bs)
    val = 0.0
    for su, sd, eu, ed, v in _wave_blocks(wsec, x):
        subs = secs.get((su, sd) if side == 0 else (eu, ed), [])
        for _, _, tu, td, spos, sdim in subs:
            for u, d, mat in site:
                if u == tu and d == td:
                    vx = v[spos:spos + sdim, :] if side == 0 else v[:, spos:spos + sdim]
                    val += float(np.sum(vx * vx)) * mat[0, 0] / math.sqrt(1.0 + td)
                    break
    return val


def density_profile(up_bias, nsite, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x):
    """measure_operator_dia (measure.f90:3-57) with num_elec: <n_i> for the
    nsite lattice sites.  sys_len / env_len are the wavefunction's lengths
    (block + site), bases keyed by length 2..sys_len (env: 2..env_len),
    truncation operators by length 2..sys_len - 1 (env_len - 1)."""
    site = site_num_elec(up_bias)
    ave = np.zeros(nsite)
    for i in range(1, sys_len):
        op = transport(site, i, sys_bases, sys_truns, sys_len - 1)
        ave[i - 1] = measure_block(0, op, sys_bases[sys_len], wsec, x)
    ave[sys_len - 1] = measure_site(0, site, sys_bases[sys_len], wsec, x)
    for i in range(1, env_len):
        op = transport(site, i, env_bases, env_truns, env_len - 1)
        ave[nsite - i] = measure_block(1, op, env_bases[env_len], wsec, x)
    ave[nsite - env_len] = measure_site(1, site, env_bases[env_len], wsec, x)
    return ave


def _cross_pair(sys_op, sys_is_site, env_op, env_is_site, sys_bs, env_bs, wsec, x):
    """sys_block_env_block_cor_dia (measure.f90:1638-1735) and its site
    variants sys_block_env_site (:1736-1834), sys_site_env_block
    (:1835-1932), sys_site_env_site (:2155-2247): per wavefunction sector,
    system sub-sector and environment sub-sector, the first matching
    sub-block of each side's operator (block part, or site part for a site
    operator), coefficient 1 / sqrt((1 + 2S_sys)(1 + 2S_env)) times the
    site operators' values."""
    ssecs, esecs = _secs_by_qn(sys_bs), _secs_by_qn(env_bs)
    val = 0.0
    for su, sd, eu, ed, v in _wave_blocks(wsec, x):
        for bu, bd, tu, td, spos, sdim in ssecs.get((su, sd), []):
            qs = (tu, td) if sys_is_site else (bu, bd)
            ms = next((m for u, d, m in sys_op if (u, d) == qs), None)
            if ms is None:
                continue
            for ebu, ebd, etu, etd, epos, edim in esecs.get((eu, ed), []):
                qe = (etu, etd) if env_is_site else (ebu, ebd)
                me = next((m for u, d, m in env_op if (u, d) == qe), None)
                if me is None:
                    continue
                c = 1.0 / math.sqrt((1.0 + qs[1]) * (1.0 + qe[1]))
                vb = v[spos:spos + sdim, epos:epos + edim]
                left = ms[0, 0] * vb if sys_is_site else ms @ vb
                right = me[0, 0] * vb if env_is_site else vb @ me
                val += float(np.sum(left * right)) * c
    return val


def cross_correlation(up_bias, sys_len, env_len, sys_bases, env_bases, sys_truns, env_truns, wsec, x):
    """The across-the-bond block of Get_oper_cor_dia (measure.f90:468-634)
    for num_elec: <n_i n_j> for system-side sites i = 1..sys_len (rows) and
    environment-side sites j = 1..env_len (columns; j = env_len is the
    environment's bond site, lattice N - env_len + 1)."""
    site = site_num_elec(up_bias)
    sops = [transport(site, i, sys_bases, sys_truns, sys_len - 1) for i in range(1, sys_len)]
    eops = [transport(site, j, env_bases, env_truns, env_len - 1) for j in range(1, env_len)]
    sb, eb = sys_bases[sys_len], env_bases[env_len]
    out = np.zeros((sys_len, env_len))
    for i in range(sys_len):
        so, ss = (sops[i], False) if i < sys_len - 1 else (site, True)
        for j in range(env_len):
            eo, es = (eops[j], False) if j < env_len - 1 else (site, True)
            out[i, j] = _cross_pair(so, ss, eo, es, sb, eb, wsec, x)
    return out
