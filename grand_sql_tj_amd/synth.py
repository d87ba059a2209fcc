"""Synthetic, symmetric environments for H_eff and Lanczos workloads.

A block of length L carries SU(2) multiplet sectors (N, 2S) with multiplet
dimensions drawn log-normally, and random reduced operators obeying the
reference's selection rules:

* Hamiltonian and density n: scalar, one symmetric square block per sector;
* c^+ (elec_up): rank 1/2, (N, 2S) -> (N + up_bias, 2S +/- 1);
* c   (elec_down): rank 1/2, (N, 2S) -> (N - up_bias, 2S +/- 1);
* S   (spin_sd): rank 1, (N, 2S) -> (N, 2S + {-2, 0, 2}), triangle allowed,
  with the hermiticity relation between (a -> b) and (b -> a) blocks that
  makes the spin-spin terms of H_eff symmetric.

Fermion hopping terms are symmetric by construction (the apply adds the
hermitian-conjugate product itself), so c^+ and c need no relation.  With
symmetric scalar blocks and hermitian S the whole H_eff is symmetric, so
Lanczos on it is well posed.

Randomness comes from SplitMix64 (Steele, Lea, Flood 2014) seeded explicitly,
so a workload is the same on every machine and numpy version.
"""
from __future__ import annotations

import math

import numpy as np

from .blocks import (Basis, OpSetBuilder, Superblock, get_basis, get_super_basis,
                     site_sectors, spin_hermitian_factor)
from .lattice import TJModel, active_sites

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator with vectorised uniform/normal draws."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self, n: int) -> np.ndarray:
        idx = np.arange(1, n + 1, dtype=np.uint64)
        gamma = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + idx * gamma
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Uniform doubles in (0, 1) from the top 53 bits."""
        return ((self.next_u64(n) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Standard normals by Box-Muller."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


class FastNormals(SplitMix64):
    """SplitMix64 for the structure draws (sector dimensions), numpy's PCG64
    for the operator entries: the same block structure as SplitMix64 with
    the same seed, different (and ~10x cheaper) entries.  For tests at the
    benchmark sizes, where only the structure and the symmetries matter."""

    def __init__(self, seed: int):
        super().__init__(seed)
        self.fast = np.random.default_rng(seed & _MASK)

    def entries(self, n: int) -> np.ndarray:
        return self.fast.standard_normal(n)


def block_sectors(rng: SplitMix64, length: int, up_bias: int, n_center: float, m: int,
                  median: float, sigma: float, maxdim: int, nsec: int, smax: int):
    """Sector list [(N, 2S, dim)] of a block, ordered like the reference's
    enlarged bases (2S descending, then N descending; block.f90:212-239)."""
    cands = []
    if up_bias:
        n0 = int(round(n_center))
        for dn in range(-length, length + 1):
            N = n0 + dn
            if N < 0 or N > length:
                continue
            for s2 in range(N % 2, min(N, smax) + 1, 2):
                cands.append((abs(dn) + 0.35 * s2, N, s2))
    else:
        for s2 in range(0, min(length, smax) + 1):
            cands.append((0.5 * s2, 0, s2))
    cands.sort()
    chosen = [(N, s2) for _, N, s2 in cands[:max(1, nsec)]]
    z = rng.normal(len(chosen))
    dims = np.exp(np.log(median) + sigma * z)
    dims = dims * (m / dims.sum())
    dims = np.clip(np.rint(dims), 1, maxdim).astype(np.int64)
    secs = [(N, s2, int(d)) for (N, s2), d in zip(chosen, dims)]
    secs.sort(key=lambda t: (-t[1], -t[0]))
    return np.array(secs, dtype=np.int32).reshape(-1, 3)


def _rand(rng: SplitMix64, r: int, c: int, scale: float) -> np.ndarray:
    draw = rng.entries if isinstance(rng, FastNormals) else rng.normal
    return (draw(r * c) * scale).reshape(c, r).T.copy()


def _sym(rng: SplitMix64, d: int, scale: float) -> np.ndarray:
    a = _rand(rng, d, d, scale)
    return 0.5 * (a + a.T)


def random_model(rng: SplitMix64, length: int, secs: np.ndarray, nact: int, up_bias: int,
                 scale: float = 1.0, ham_shift: float = 0.0):
    """Random block model (Hamiltonian + nact active-site operator sets)."""
    index = {(int(u), int(d)): int(k) for u, d, k in secs}
    b = OpSetBuilder(length)
    ham = []
    for u, d, k in secs:
        h = _sym(rng, int(k), scale / math.sqrt(max(1, int(k))))
        h += ham_shift * np.eye(int(k))
        ham.append((int(u), int(d), 0, h))
    b.add(0, 0, ham)
    for _site in range(nact):
        for which in range(4):
            blocks = []
            if which in (0, 1):
                du = up_bias if which == 0 else -up_bias
                for u, d, k in secs:
                    for dd in (-1, 1):
                        bra = (int(u) + du, int(d) + dd)
                        if bra[1] < 0 or bra not in index:
                            continue
                        blocks.append((int(u), int(d), dd,
                                       _rand(rng, index[bra], int(k), scale / math.sqrt(index[bra] + int(k)))))
                b.add(du, 1, blocks)
            elif which == 2:
                mats = {}
                for u, d, k in secs:
                    for dd in (-2, 0, 2):
                        bd = int(d) + dd
                        if bd < 0 or (int(u), bd) not in index:
                            continue
                        if int(d) == 0 and bd == 0:
                            continue
                        key = (int(u), int(d), bd)
                        rev = (int(u), bd, int(d))
                        if rev in mats:
                            mats[key] = spin_hermitian_factor(bd, int(d)) * mats[rev].T
                        elif dd == 0:
                            a = _sym(rng, int(k), scale / math.sqrt(int(k)))
                            mats[key] = a
                        else:
                            mats[key] = _rand(rng, index[(int(u), bd)], int(k),
                                              scale / math.sqrt(index[(int(u), bd)] + int(k)))
                        blocks.append((int(u), int(d), dd, mats[key]))
                b.add(0, 2, blocks)
            else:
                blocks = [(int(u), int(d), 0, _sym(rng, int(k), scale / math.sqrt(int(k))))
                          for u, d, k in secs]
                b.add(0, 0, blocks)
    return b.build()


def superblock(model: TJModel, sys_len: int, env_len: int, m_sys: int, m_env: int, seed: int,
               median: float = 4.0, sigma: float = 0.6, maxdim: int = 600, nsec: int = 8,
               smax: int = 6, filling: float | None = None, total_up: int | None = None,
               scale: float = 1.0, nact: int | None = None, fast: bool = False) -> Superblock:
    """A random but symmetric two-site superblock [sys . site . site . env]
    (fast: operator entries from FastNormals)."""
    rng = FastNormals(seed) if fast else SplitMix64(seed)
    ub = model.up_bias
    if filling is None:
        filling = 0.875
    _, nleg11 = active_sites(model)
    na_sys = nact or max(1, int(nleg11[sys_len]))
    na_env = nact or max(1, int(nleg11[env_len]))
    s_secs = block_sectors(rng, sys_len, ub, filling * sys_len, m_sys, median, sigma, maxdim, nsec, smax)
    e_secs = block_sectors(rng, env_len, ub, filling * env_len, m_env, median, sigma, maxdim, nsec, smax)
    sys = random_model(rng, sys_len, s_secs, na_sys, ub, scale)
    env = random_model(rng, env_len, e_secs, na_env, ub, scale)
    site = site_sectors(ub)
    sys_bs = get_basis(sys_len, s_secs, site)
    env_bs = get_basis(env_len, e_secs, site)
    if total_up is None:
        if ub:
            n_tot = int(round(filling * (sys_len + env_len + 2)))
            n_tot -= n_tot % 2
            total_up = n_tot
        else:
            total_up = 0
    wsec = get_super_basis(sys_bs, env_bs, total_up, 0)
    return Superblock(sys_len, env_len, sys, env, sys_bs, env_bs, wsec,
                      meta={"seed": seed, "total_up": total_up, "model": model})


def block(model: TJModel, length: int, m: int, seed: int, median: float = 4.0, sigma: float = 0.6,
          maxdim: int = 600, nsec: int = 8, smax: int = 6, filling: float = 0.875, scale: float = 1.0,
          nact: int | None = None, fast: bool = False):
    """One random symmetric block model of `length` sites and about m
    multiplets (the sys side of superblock() with the same arguments)."""
    rng = FastNormals(seed) if fast else SplitMix64(seed)
    _, nleg11 = active_sites(model)
    secs = block_sectors(rng, length, model.up_bias, filling * length, m, median, sigma, maxdim, nsec, smax)
    return random_model(rng, length, secs, nact or max(1, int(nleg11[length])), model.up_bias, scale)


def random_vector(sb: Superblock, seed: int) -> np.ndarray:
    rng = SplitMix64(seed ^ 0x5DEECE66D)
    x = rng.normal(sb.dim)
    return x / np.linalg.norm(x)


def as_basis(sec, sub, length) -> Basis:
    return Basis(length, np.asarray(sec, dtype=np.int32), np.asarray(sub, dtype=np.int32))
