"""Exact diagonalisation of the t-J model on the reference's square cylinder
(test helper): the physics pin of the device DMRG.

H = sum_<ij> [ -t_ij sum_s (c+_is c_js + h.c.) + J_ij (S_i.S_j - n_i n_j / 4) ]
    - mu sum_i n_i,
nearest-neighbour bonds (t, J) and next-nearest (t', J' = J (t'/t)^2) taken
from the neighbour table of dmrg_setup (code/tjsq/dmrg_sub.f90:87-143: y
periodic, x open, no doubled rung on a two-leg ladder), no double occupancy,
fixed N and S_z = 0.
"""
import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from grand_sql_tj_amd.lattice import square_neighbours


def bonds(model):
    n = model.nsite
    neib = square_neighbours(model.ny, model.nx, n)
    out = []
    for i in range(1, n + 1):
        for k, (t, J) in ((0, (model.t, model.J)), (2, (model.t, model.J)), (4, (model.tp, model.jprime)),
                          (6, (model.tp, model.jprime))):
            j = int(neib[i, k])
            if j and (t != 0.0 or J != 0.0):
                out.append((i - 1, j - 1, t, J))
    return out


def _states(n, nup, ndn):
    ups = [sum(1 << b for b in c) for c in itertools.combinations(range(n), nup)]
    out = []
    for u in ups:
        free = [b for b in range(n) if not (u >> b) & 1]
        for c in itertools.combinations(free, ndn):
            out.append((u << n) | sum(1 << b for b in c))
    return np.array(sorted(out), dtype=np.int64)


def _popcount_between(m, a, b):
    """number of set bits of m strictly between positions a and b"""
    lo, hi = min(a, b), max(a, b)
    mask = ((1 << hi) - 1) & ~((1 << (lo + 1)) - 1)
    x = m & mask
    c = np.zeros_like(x)
    while True:
        nz = x != 0
        if not nz.any():
            return c
        c += nz
        x = x & (x - 1)


def hamiltonian(model, n_elec):
    """Sparse H on the (N = n_elec, S_z = 0) sector; returns (H, states)."""
    n = model.nsite
    st = _states(n, n_elec // 2, n_elec // 2)
    full = (1 << n) - 1
    up, dn = st >> n, st & full
    rows, cols, vals = [], [], []
    diag = np.full(st.size, -model.mu * n_elec, dtype=np.float64)
    for i, j, t, J in bonds(model):
        bi, bj = 1 << i, 1 << j
        ui, uj = (up & bi) != 0, (up & bj) != 0
        di, dj = (dn & bi) != 0, (dn & bj) != 0
        # exchange, diagonal: J (Sz Sz - n n / 4): antiparallel -J/2, parallel 0
        anti = (ui & dj) | (di & uj)
        diag += np.where(anti, -0.5 * J, 0.0)
        # spin flip J/2 (S+_i S-_j + h.c.): (i up, j dn) <-> (i dn, j up).
        # S-_i S+_j = c+_id c_iu c+_ju c_jd = -(c+_id c_jd)(c+_ju c_iu): minus
        # the product of the two hops, each with its string sign
        for sel, nu, nd in (((ui & dj), up ^ bi ^ bj, dn ^ bi ^ bj), ((di & uj), up ^ bi ^ bj, dn ^ bi ^ bj)):
            idx = np.nonzero(sel)[0]
            if idx.size == 0 or J == 0.0:
                continue
            sign = (-1.0) ** (_popcount_between(up[idx], i, j) + _popcount_between(dn[idx], i, j))
            tgt = np.searchsorted(st, (nu[idx] << n) | nd[idx])
            rows.append(tgt); cols.append(idx); vals.append(-0.5 * J * sign)
        # hopping -t (c+_is c_js + h.c.) for s = up, down
        if t != 0.0:
            for spin in (0, 1):
                m = up if spin == 0 else dn
                occ_i, occ_j = (m & bi) != 0, (m & bj) != 0
                empty_i = ((up | dn) & bi) == 0
                empty_j = ((up | dn) & bj) == 0
                for src, dst, ok in ((j, i, occ_j & empty_i), (i, j, occ_i & empty_j)):
                    idx = np.nonzero(ok)[0]
                    if idx.size == 0:
                        continue
                    mm = m[idx] ^ (1 << src) ^ (1 << dst)
                    sign = (-1.0) ** _popcount_between(m[idx], src, dst)
                    new = (mm << n) | dn[idx] if spin == 0 else (up[idx] << n) | mm
                    tgt = np.searchsorted(st, new)
                    rows.append(tgt); cols.append(idx); vals.append(-t * sign)
    rows.append(np.arange(st.size)); cols.append(np.arange(st.size)); vals.append(diag)
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(st.size, st.size))
    return H, st


def ground_energy(model, n_elec):
    H, _ = hamiltonian(model, n_elec)
    assert abs(H - H.T).max() < 1e-12
    if H.shape[0] < 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spla.eigsh(H, k=1, which="SA", tol=1e-13)[0][0])


def ground_state(model, n_elec):
    """(E0, g, states) of the (N = n_elec, S_z = 0) sector."""
    H, st = hamiltonian(model, n_elec)
    if H.shape[0] <= 2000:
        w, v = np.linalg.eigh(H.toarray())
    else:
        w, v = spla.eigsh(H, k=1, which="SA", tol=1e-13)
    return float(w[0]), v[:, 0], st


def spin_correlation(model, g, st):
    """<S_i . S_j> (n x n) in the state g over the basis st: S^z S^z on the
    diagonal of the basis, (S^+_i S^-_j + S^-_i S^+_j) / 2 with the string
    signs of the exchange term of hamiltonian()."""
    n = model.nsite
    full = (1 << n) - 1
    up, dn = st >> n, st & full
    w = g * g
    out = np.zeros((n, n))
    sz = np.array([0.5 * (((up >> b) & 1) - ((dn >> b) & 1)) for b in range(n)], np.float64)
    for i in range(n):
        for j in range(n):
            if i == j:
                out[i, i] = 0.75 * float((((up | dn) >> i) & 1) @ w)
                continue
            v = float((sz[i] * sz[j]) @ w)
            bi, bj = 1 << i, 1 << j
            ui, uj, di, dj = (up & bi) != 0, (up & bj) != 0, (dn & bi) != 0, (dn & bj) != 0
            for sel in ((ui & dj), (di & uj)):
                idx = np.nonzero(sel)[0]
                if idx.size == 0:
                    continue
                sign = (-1.0) ** (_popcount_between(up[idx], i, j) + _popcount_between(dn[idx], i, j))
                tgt = np.searchsorted(st, ((up[idx] ^ bi ^ bj) << n) | (dn[idx] ^ bi ^ bj))
                v += float((-0.5 * sign * g[tgt]) @ g[idx])
            out[i, j] = v
    return out


def hopping_correlation(model, g, st):
    """sum over s of <c^+_is c_js> (n x n, i != j) in the state g, with the
    string signs of the hopping term of hamiltonian()."""
    n = model.nsite
    full = (1 << n) - 1
    up, dn = st >> n, st & full
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            bi, bj = 1 << i, 1 << j
            v = 0.0
            for spin in (0, 1):
                m = up if spin == 0 else dn
                ok = ((m & bj) != 0) & (((up | dn) & bi) == 0)
                idx = np.nonzero(ok)[0]
                if idx.size == 0:
                    continue
                mm = m[idx] ^ bj ^ bi
                sign = (-1.0) ** _popcount_between(m[idx], i, j)
                new = (mm << n) | dn[idx] if spin == 0 else (up[idx] << n) | mm
                v += float((sign * g[np.searchsorted(st, new)]) @ g[idx])
            out[i, j] = v
    return out


def density_correlation(model, g, st):
    n = model.nsite
    full = (1 << n) - 1
    up, dn = st >> n, st & full
    occ = np.array([((up >> b) & 1) + ((dn >> b) & 1) for b in range(n)], dtype=np.float64)
    return (occ * (g * g)) @ occ.T


def _pair_annihilate(n, st, g, k, l, target):
    """Delta_kl |g> = (c_l,dn c_k,up - c_l,up c_k,dn) / sqrt(2) |g> on the
    states `target`, modes ordered all up spins (by site), then all down
    spins: the ordering whose string signs hamiltonian() uses."""
    full = (1 << n) - 1
    up, dn = st >> n, st & full
    out = np.zeros(target.size)

    def pc(x):
        c = np.zeros_like(x)
        while True:
            nz = x != 0
            if not nz.any():
                return c
            c += nz
            x = x & (x - 1)

    for first_up, sgn in ((True, 1.0), (False, -1.0)):
        if first_up:       # c_k,up then c_l,dn
            ok = ((up >> k) & 1 == 1)
            u1 = up ^ (1 << k)
            s1 = pc(up & ((1 << k) - 1))
            ok &= ((dn >> l) & 1 == 1)
            d1 = dn ^ (1 << l)
            s2 = pc(u1) + pc(dn & ((1 << l) - 1))
            nu, nd = u1, d1
        else:              # c_k,dn then c_l,up
            ok = ((dn >> k) & 1 == 1)
            d1 = dn ^ (1 << k)
            s1 = pc(up) + pc(dn & ((1 << k) - 1))
            ok &= ((up >> l) & 1 == 1)
            u1 = up ^ (1 << l)
            s2 = pc(up & ((1 << l) - 1))
            nu, nd = u1, d1
        idx = np.nonzero(ok)[0]
        if idx.size == 0:
            continue
        tgt = np.searchsorted(target, (nu[idx] << n) | nd[idx])
        np.add.at(out, tgt, sgn * (-1.0) ** (s1[idx] + s2[idx]) * g[idx] / np.sqrt(2.0))
    return out


def pair_correlation(model, g, st, n_elec, i, j, k, l):
    """<Delta^+_ij Delta_kl> in the state g of the (N = n_elec, S_z = 0)
    sector over st, with the singlet pair Delta^+_ij = (c^+_i,up c^+_j,dn -
    c^+_i,dn c^+_j,up) / sqrt(2) (sites 0-based)."""
    n = model.nsite
    target = _states(n, n_elec // 2 - 1, n_elec // 2 - 1)
    a = _pair_annihilate(n, st, g, i, j, target)
    b = _pair_annihilate(n, st, g, k, l, target)
    return float(a @ b)


def pair_amplitude(model, g_lo, st_lo, g_hi, st_hi, i, j):
    """<g_hi| Delta^+_ij |g_lo> between a state of N electrons (g_lo over
    st_lo) and one of N + 2 (g_hi over st_hi), both S_z = 0 (sites 0-based)."""
    return float(_pair_annihilate(model.nsite, st_hi, g_hi, i, j, st_lo) @ g_lo)
