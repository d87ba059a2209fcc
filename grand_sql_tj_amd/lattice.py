"""Square-cylinder lattice description and the couplings of the t-J model.

The coupling constants follow the reference's conventions
(code/tjsq/main.f90:35-47, code/tjsq/readme.md:28-34):

* hopping      jt = -sqrt(2) t   on nearest-neighbour bonds (directions 1-4),
               jt = -sqrt(2) t'  on next-nearest bonds (directions 5-8);
* exchange     jd = -sqrt(3) J,  J' = J (t'/t)^2 on next-nearest bonds;
* density      jn = -J/4 (and -J'/4);
* on-site      esite = -mu.

The neighbour table restates dmrg_setup (code/tjsq/dmrg_sub.f90:31-143, square
lattice, y periodic, x open) and the active-site table dmrg_setup0
(dmrg_sub.f90:1-29).  The C++ plan builder holds the same construction; this
module serves the Python-side generators.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

NEIBT = 12


@dataclass(frozen=True)
class TJModel:
    ny: int                 # cylinder circumference (nleg)
    nx: int                 # cylinder length in columns
    t: float = 1.0
    tp: float = 0.0         # t'
    J: float = 0.4
    mu: float = 0.0
    up_bias: int = 1        # 1: conserve (N, 2S); 0: SU(2) only (grand canonical)
    nx_tables: int = 0      # columns used for the active-site tables (0: nx)

    @property
    def nsite(self) -> int:
        return self.ny * self.nx

    @property
    def jt(self):
        return (-math.sqrt(2.0) * self.t, -math.sqrt(2.0) * self.tp)

    @property
    def jprime(self) -> float:
        return self.J * (self.tp / self.t) ** 2 if self.t != 0 else 0.0

    @property
    def jd(self):
        return (-math.sqrt(3.0) * self.J, -math.sqrt(3.0) * self.jprime)

    @property
    def jn(self):
        return (-self.J / 4.0, -self.jprime / 4.0)

    @property
    def esite(self) -> float:
        return -self.mu

    @property
    def table_columns(self) -> int:
        return self.nx_tables or self.nx


def square_neighbours(nleg: int, nxc: int, nsite: int) -> np.ndarray:
    """neib[i, k-1] for sites 1..nleg*nxc (dmrg_sub.f90:87-143, open = 1)."""
    neib = np.zeros((nsite + 1, NEIBT), dtype=np.int32)
    for ix in range(1, nxc + 1):
        for iy in range(1, nleg + 1):
            ii = iy + (ix - 1) * nleg
            jy = ii - iy + (iy % nleg) + 1
            if iy != nleg or nleg != 2:
                neib[ii, 0] = jy
                neib[jy, 1] = ii
            jx = ii + nleg
            if jx > nleg * nxc:
                jx = 0
            if jx:
                neib[ii, 2] = jx
                neib[jx, 3] = ii
                jy = jx - 1 + (nleg if iy == 1 else 0)
                if jy:
                    neib[ii, 4] = jy
                    neib[jy, 5] = ii
                jy = jx + 1 - (nleg if iy == nleg else 0)
                if jy:
                    neib[ii, 6] = jy
                    neib[jy, 7] = ii
    return neib


def active_sites(model: TJModel):
    """(inb1, nleg11) of dmrg_setup0 over the full lattice
    (dmrg_sub.f90:1-29 after dmrg_setup(N/2-1, N/2-1))."""
    n = model.ny * model.table_columns
    neib = square_neighbours(model.ny, model.table_columns, n)
    inb1 = np.zeros((n + 1, n + 1), dtype=np.int32)
    nleg11 = np.zeros(n + 1, dtype=np.int32)
    for ns in range(1, n - 1):
        k0 = 0
        for i1 in range(1, ns + 1):
            inb1[i1, ns] = 10000
            if np.count_nonzero(neib[i1] >= ns):
                k0 += 1
                inb1[i1, ns] = k0
        nleg11[ns] = k0
    return inb1, nleg11
