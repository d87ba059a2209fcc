"""Writes ed_energies.json: exact ground-state energies of the t-J model on
small cylinders (S_z = 0 sector, which holds the S = 0 ground state here),
from tests/ed.py.  Run from the repository root: python tests/golden/make_ed.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import ed  # noqa: E402
from grand_sql_tj_amd.lattice import TJModel  # noqa: E402

CASES = [  # (ny, nx, N, t, tp, J, mu)
    (2, 6, 10, 1.0, 0.0, 0.4, 0.0),
    (4, 3, 10, 1.0, 0.0, 0.4, 0.0),
    (2, 6, 12, 1.0, 0.0, 0.4, 0.0),
    (2, 8, 14, 1.0, 0.0, 0.4, 0.0),
    (4, 4, 14, 1.0, 0.0, 0.4, 0.0),
    (4, 3, 10, 1.0, -0.2, 0.4, 0.0),     # next-nearest hopping t' and J' = J (t'/t)^2
    (2, 6, 10, 1.0, 0.0, 0.4, 0.3),      # an on-site term esite = -mu
]

out = []
for ny, nx, n, t, tp, J, mu in CASES:
    m = TJModel(ny=ny, nx=nx, t=t, tp=tp, J=J, mu=mu, up_bias=1)
    e = ed.ground_energy(m, n)
    out.append({"ny": ny, "nx": nx, "N": n, "t": t, "tp": tp, "J": J, "mu": mu, "E0": e})
    print(out[-1], flush=True)
with open(os.path.join(HERE, "ed_energies.json"), "w") as f:
    json.dump(out, f, indent=1)
