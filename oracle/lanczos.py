"""numpy restatement of lanczos_memo_new (code/tjsq/eigen.f90:2-175), levels=1.

TEST INFRASTRUCTURE ONLY (the checker).  The H_eff apply is passed in as a
callable (the C oracle's model_wave in the tests).  Steps follow the
reference: normalise the start vector (:49-50), three-term recurrence
(:63-72), breakdown exit when beta <= lanzero (:74-77), classical
Gram-Schmidt against every previous vector with the <v_j, v_j> division of
Gram_Schmidt_process (block.f90:2515-2551), eigenvalues of the tridiagonal
matrix each step (dstev 'N', :93-96), stop when the lowest moved by at most
lanzero or at loop_in vectors (:98-105), ground state from the eigenvector of
the final tridiagonal matrix (:138-154).  DSTEV is replaced by
scipy.linalg.eigh_tridiagonal (LAPACK stemr/stebz), a different LAPACK driver
for the same symmetric tridiagonal eigenproblem.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh_tridiagonal


def lanczos(apply, x0: np.ndarray, lanzero: float = 1e-10, loop_in: int = 110):
    v = [x0 / np.linalg.norm(x0)]
    alpha, beta = [], []
    w = apply(v[0])
    alpha.append(float(v[0] @ w))
    ritz = alpha[0]
    breakdown = False
    while len(v) < loop_in:
        last = len(v)
        if last < 2:
            w = w - alpha[-1] * v[-1]
        else:
            w = w - alpha[-1] * v[-1] - beta[-1] * v[-2]
        b = float(np.sqrt(w @ w))
        beta.append(b)
        if abs(b) <= lanzero:
            breakdown = True
            break
        w = w / b
        uv = np.array([vj @ w for vj in v])
        uu = np.array([vj @ vj for vj in v])
        corr = np.zeros_like(w)
        for j, vj in enumerate(v):
            corr = corr + (uv[j] / uu[j]) * vj
        w = w - corr
        w = w / np.sqrt(w @ w)
        v.append(w)
        w = apply(v[-1])
        alpha.append(float(v[-1] @ w))
        k = len(v)
        ev = eigh_tridiagonal(np.array(alpha), np.array(beta[:k - 1]), eigvals_only=True)
        done = abs(ev[0] - ritz) <= lanzero or k == loop_in
        ritz = ev[0]
        if done:
            break
    k = len(v)
    if k == 1:
        return alpha[0], v[0], k, breakdown
    ev, z = eigh_tridiagonal(np.array(alpha[:k]), np.array(beta[:k - 1]))
    psi = sum(z[i, 0] * v[i] for i in range(k))
    psi = psi / np.linalg.norm(psi)
    return float(ev[0]), psi, k, breakdown
