"""Physics pins of the oracle (CPU): the spectrum of H_eff built by the
oracle's hmerge + model_wave on small superblocks whose blocks hold their
full Hilbert space (site models grown with identity truncations through the
oracle's Get_Hamiltonian) lies inside the exact-diagonalisation spectrum of
the t-J model on the same cylinder (tests/ed.py), and its ground state is
the exact one.  This checks the 14 kernels' reduced matrix elements, 6j
coefficients and fermion signs against the physics rather than against the
reference."""
import numpy as np
import pytest

import ed
import oracle
from helpers import opset
from grand_sql_tj_amd import sweep
from grand_sql_tj_amd.blocks import Superblock, get_basis, get_super_basis, site_sectors
from grand_sql_tj_amd.lattice import TJModel


def _identity(basis, length):
    sub, off, data, nd = [], [], [], 0
    for u, d, n, _ in basis.sec:
        sub.append((u, d, 0, n, n))
        off.append(nd)
        data.append(np.eye(int(n)).reshape(-1))
        nd += int(n) * int(n)
    return opset(length, [[length, 0, 0, len(sub)]], sub, off, np.concatenate(data))


def _heff(m, sb):
    d = sb.dim
    H = np.zeros((d, d))
    for k in range(d):
        e = np.zeros(d)
        e[k] = 1.0
        H[:, k] = oracle.model_wave(m, sb, e)
    return H


def _blocks(m, length):
    """Exact block models of length 1..length (identity truncations)."""
    site = site_sectors(m.up_bias)
    blk = {1: (sweep.site_model(m), sweep.site_model(m))}
    for L in range(1, length):
        s, e = blk[L]
        sbs, ebs = get_basis(L, s.sectors(), site), get_basis(L, e.sectors(), site)
        ns = oracle.get_hamiltonian(m, 0, L, L, s, sbs, _identity(sbs, L + 1))
        ne = oracle.get_hamiltonian(m, 1, L, L, e, ebs, _identity(ebs, L + 1))
        blk[L + 1] = tuple(opset(L + 1, r.hdr, r.sub, r.off, r.data) for r in (ns, ne))
    return blk


def test_ed_closed_forms():
    ring = TJModel(ny=2, nx=2, t=1.0, J=0.4, up_bias=1)
    assert abs(ed.ground_energy(ring, 4) - (-2 * 0.4 - 0.4)) < 1e-12     # Heisenberg 4-ring: -2J - 4 J/4
    free = TJModel(ny=2, nx=2, t=1.0, J=0.0, up_bias=1)
    assert abs(ed.ground_energy(free, 2) + 2 * np.sqrt(2)) < 1e-12      # two hard-core electrons on the ring
    # the singlet feels J: the ground state of the stoquastic pair is symmetric
    assert ed.ground_energy(ring, 2) < ed.ground_energy(free, 2) - 0.1


# t' only on cylinders with ny >= 4: on a two-leg ladder the neighbour
# table (dmrg_setup, dmrg_sub.f90:128-141) lists each diagonal bond twice and
# the reference's Hamiltonian is not the ladder's t-t' model
@pytest.mark.parametrize("ny,nx,L,N,tp,mu", [(2, 2, 1, 2, 0.0, 0.0), (2, 3, 2, 4, 0.0, 0.0), (2, 3, 2, 4, 0.0, 0.2),
                                             (2, 4, 3, 6, 0.0, 0.0), (4, 2, 3, 6, -0.2, 0.0), (4, 2, 3, 4, 0.3, 0.1)])
def test_heff_spectrum_inside_ed(ny, nx, L, N, tp, mu):
    m = TJModel(ny=ny, nx=nx, t=1.0, tp=tp, J=0.4, mu=mu, up_bias=1)
    assert 2 * L + 2 == m.nsite
    s, e = _blocks(m, L)[L]
    site = site_sectors(1)
    sbs, ebs = get_basis(L, s.sectors(), site), get_basis(L, e.sectors(), site)
    w = get_super_basis(sbs, ebs, N, 0)
    sb = Superblock(L, L, s, e, sbs, ebs, w, meta={})
    H = _heff(m, sb)
    assert np.abs(H - H.T).max() < 1e-12
    ev = np.linalg.eigvalsh(H)
    Hf, _ = ed.hamiltonian(m, N)
    full = np.linalg.eigvalsh(Hf.toarray()) if Hf.shape[0] <= 3000 else None
    e0 = ed.ground_energy(m, N)
    assert abs(ev[0] - e0) < 1e-10 * max(1.0, abs(e0))
    if full is not None:      # every S = 0 level is a level of the full problem
        for v in ev:
            assert np.abs(full - v).min() < 1e-9
