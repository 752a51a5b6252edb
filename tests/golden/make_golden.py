"""Writes the golden vectors of the NLS operators from the CPU oracle.

Run from the repository root: python tests/golden/make_golden.py
Cases: the 64x64 right-triangle square (mu = 0.1) and the 4^3 Kuhn cube
(mu = 0.4), V = -1, g = 1, thickness 1, psi seed 0, phi seed 1. Stored: the
edge list, cell->edge map, edge lengths/covolumes, control volumes, F(psi),
K phi, J(psi) phi, dF/dg and dF/dmu (deriv_parameter "mu"), and the
keo_regularized product P phi.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from nosh_amd import meshgen as mg  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {"square64": (lambda: mg.square_tri(64), 0.1),
         "kuhn4": (lambda: mg.kuhn_cube(4), 0.4)}


def make(name):
    gen, mu = CASES[name]
    coords, cells = gen()
    m = O.Mesh(coords, cells)
    md = O.Model(m, mg.magnetic_potential(coords))
    md.set_params(1.0, mu)
    psi = mg.random_state(m.n, 0)
    phi = mg.random_state(m.n, 1)
    md.jac_rebuild(psi)
    md.prec_rebuild(psi)
    ln, cov = m.edge_data()
    return dict(mu=np.array(mu), edges=m.edges(), cell_edges=m.cell_edges(), length=ln,
                covolume=cov, cv=m.control_volumes(), F=md.f(psi), Kphi=md.keo(phi),
                Jphi=md.jac(phi), dFdg=md.dfdp(psi, "g"), dFdmu=md.dfdp(psi, "mu"),
                Pphi=md.prec_op(phi))


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **make(name))
        print("wrote", name)
