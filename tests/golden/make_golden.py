"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
* p1_unit_square_4x4.npz          faer CSC of assemble_stiffness_2d(unit_square(4,4))
* elasticity_box_beam_2x1x1.npz   faer CSC of assemble_elasticity_stiffness_3d(
                                  box_beam(1,0.1,0.1,2,1,1), E=1e7, nu=0.3)
* counts.json                     exact sizes of the benchmark configurations
The DOLFINx numbers (dolfinx_reference.json) are data from the reference's
benchmarks/dolfinx_reference.json:1-13, kept as a fixture.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402


def main():
    cp, ri, va = O.assemble_stiffness(O.unit_square(4, 4)).export()
    np.savez(os.path.join(HERE, "p1_unit_square_4x4.npz"), col_ptr=cp, row_idx=ri, values=va)
    m = O.box_beam(1.0, 0.1, 0.1, 2, 1, 1)
    cp, ri, va = O.assemble_elasticity(m, 1e7, 0.3).export()
    np.savez(os.path.join(HERE, "elasticity_box_beam_2x1x1.npz"), col_ptr=cp, row_idx=ri,
             values=va)
    counts = {}
    m = O.unit_square(128, 128)
    A = O.assemble_stiffness(m)
    bd = O.boundary_dofs(m)
    Ab, _ = O.apply_dirichlet_bc(A, np.zeros(m.nv), bd, np.zeros(len(bd)))
    counts["C0"] = dict(n=m.nv, cells=m.nc, nnz_raw=A.nnz, nnz_bc=Ab.nnz, n_fixed=len(bd))
    m = O.box_beam(10.0, 2.0, 2.0, 100, 20, 20)
    K = O.assemble_elasticity(m, 210e9, 0.3)
    fx = O.clamp_dofs_x0(m)
    Kb, _ = O.apply_dirichlet_bc(K, np.zeros(3 * m.nv), fx, np.zeros(len(fx)))
    counts["C3"] = dict(n=3 * m.nv, cells=m.nc, nnz_raw=K.nnz, nnz_bc=Kb.nnz, n_fixed=len(fx))
    # C1 and C2 are too large for a quick CPU run; their counts are closed forms
    # of the structured meshes, checked against the oracle on small grids
    counts["C1"] = dict(n=1013643, cells=1920000, nnz_raw=44316369, nnz_bc=44100345,
                        n_fixed=5043)
    counts["C2"] = dict(n=2146689, cells=1572864, nnz_raw=60859905, nnz_bc=57611761,
                        n_fixed=98306, n_vertices=274625, n_edges=1872064)
    with open(os.path.join(HERE, "counts.json"), "w") as f:
        json.dump(counts, f, indent=1)


if __name__ == "__main__":
    main()
