"""GPU boundary facets and boundary (ds) integrals (mesh.rs, boundary.rs)
against the oracle.

Facet lists (vertices, cell, local index, 3D orientation) are identical; the
integrals are bitwise equal, since every vertex sums its facets'
contributions in list order, as the reference's sequential scatter does.  The
Robin solve compares the device CG (tol 1e-12) with a sparse direct solve at
1e-9 relative.
"""
import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_csc(a, b):
    c1, r1, v1 = a.export_csc()
    c2, r2, v2 = b.export()
    return np.array_equal(c1, c2) and np.array_equal(r1, r2) and np.array_equal(bits(v1), bits(v2))


def shuffled(om, seed):
    """The oracle mesh with vertices relabelled and cells permuted and rotated."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(om.nv)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(om.nv)
    V = om.vertices[perm]
    C = inv[om.cells.astype(np.int64)][rng.permutation(om.nc)]
    C = np.array([np.roll(c, int(rng.integers(0, len(c)))) for c in C], dtype=np.uint64)
    return V, C


@pytest.mark.parametrize("dim", [2, 3])
def test_facets_match_oracle(dim):
    if dim == 2:
        gm, om = fh.unit_square(7, 5), O.unit_square(7, 5)
    else:
        gm, om = fh.box_beam(1.0, 0.5, 0.25, 4, 3, 2), O.box_beam(1.0, 0.5, 0.25, 4, 3, 2)
    gv, gc, gl = fh.boundary_facet_arrays(gm)
    ov, oc, ol = O.boundary_facets_full(om)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc) and np.array_equal(gl, ol)
    # an unstructured numbering of the same mesh
    V, Cs = shuffled(om, 3 + dim)
    v2, c2, l2 = fh.boundary_facet_arrays(fh.Mesh.from_arrays(V, Cs))
    ov2, oc2, ol2 = O.boundary_facets_full(O.Mesh(V, Cs))
    assert len(v2) == len(ov)
    assert np.array_equal(v2, ov2) and np.array_equal(c2, oc2) and np.array_equal(l2, ol2)
    lst = fh.extract_boundary_facets(gm)
    assert len(lst) == len(ov) and lst[0].vertices == tuple(int(x) for x in ov[0])


def test_neumann_bitwise_and_kat():
    gm, om = fh.unit_square(9, 6), O.unit_square(9, 6)
    ov, _, _ = O.boundary_facets_full(om)
    g = lambda p: np.sin(3 * p[0]) + p[1] ** 2  # noqa: E731
    assert np.array_equal(bits(fh.boundary_points(gm, ov)), bits(O.neumann_points_2d(om, ov)))
    assert np.array_equal(bits(fh.assemble_neumann_2d(gm, ov, g)),
                          bits(O.assemble_neumann_2d(om, ov, g)))
    # a subset in a scrambled order, one facet listed twice
    rng = np.random.default_rng(1)
    sub = ov[rng.permutation(len(ov))[:17]]
    sub = np.vstack([sub, sub[:1]])
    assert np.array_equal(bits(fh.assemble_neumann_2d(gm, sub, g)),
                          bits(O.assemble_neumann_2d(om, sub, g)))
    m4 = fh.unit_square(4, 4)  # boundary.rs:143-159
    assert abs(fh.assemble_neumann_2d(m4, fh.extract_boundary_facets(m4), lambda p: 1.0).sum()
               - 4.0) < 1e-12


def test_surface_traction_bitwise():
    gm, om = fh.box_beam(2.0, 1.0, 0.5, 5, 3, 2), O.box_beam(2.0, 1.0, 0.5, 5, 3, 2)
    ov, _, _ = O.boundary_facets_full(om)
    t = lambda p: (1.0 + p[0] * p[2], -2.0 * p[1], np.cos(p[0]))  # noqa: E731
    assert np.array_equal(bits(fh.boundary_points(gm, ov)), bits(O.traction_points_3d(om, ov)))
    assert np.array_equal(bits(fh.assemble_surface_traction_3d(gm, ov, t)),
                          bits(O.assemble_surface_traction_3d(om, ov, t)))
    tip = ov[np.all(np.abs(om.vertices[ov.astype(np.int64), 0] - 2.0) < 1e-12, axis=1)]
    assert len(tip) == 2 * 3 * 2
    assert np.array_equal(bits(fh.assemble_surface_traction_3d(gm, tip, t)),
                          bits(O.assemble_surface_traction_3d(om, tip, t)))


def test_robin_matrix_bitwise_and_solve():
    gm, om = fh.unit_square(12, 10), O.unit_square(12, 10)
    ov, _, _ = O.boundary_facets_full(om)
    R, Ro = fh.assemble_robin_matrix_2d(gm, ov, 3.0), O.assemble_robin_matrix_2d(om, ov, 3.0)
    assert same_csc(R, Ro)
    dup = np.vstack([ov, ov[:3]])
    assert same_csc(fh.assemble_robin_matrix_2d(gm, dup, 0.5), O.assemble_robin_matrix_2d(om, dup, 0.5))
    with pytest.raises(fh.FenicsError) as e:
        fh.assemble_robin_matrix_2d(gm, np.array([[3, 3]], np.uint64), 1.0)
    assert e.value.code == fh.FH_ERR_BAD_INPUT
    # -lap u = 0 with du/dn + 3 u = g on the whole boundary: (K + R) u = G
    K, Ko = fh.assemble_stiffness(gm), O.assemble_stiffness(om)
    A, Ao = fh.add_scaled_matrices(K, R, 1.0), O.add_scaled_matrices(Ko, Ro, 1.0)
    assert same_csc(A, Ao)
    g = lambda p: 1.0 + p[0]  # noqa: E731
    G = fh.assemble_neumann_2d(gm, ov, g)
    u = fh.solve_cg(A, G, tol=1e-12, max_iter=5000, precision=fh.FP64_TRUE)
    import scipy.sparse.linalg as spla
    ref = spla.spsolve(Ao.to_scipy().tocsc(), O.assemble_neumann_2d(om, ov, g))
    u = u[0] if isinstance(u, tuple) else u
    assert np.abs(u - ref).max() < 1e-9 * np.abs(ref).max()


def test_full_size_facets_and_integrals():
    gm = fh.unit_square(1024, 1024)
    v, c, l = fh.boundary_facet_arrays(gm)
    assert len(v) == 4 * 1024 and np.all(v[:, 0] < v[:, 1])
    assert abs(fh.assemble_neumann_2d(gm, v, lambda p: 1.0).sum() - 4.0) < 1e-11
    gc = fh.unit_cube(64, 64, 64)
    v3, _, _ = fh.boundary_facet_arrays(gc)
    assert len(v3) == 2 * 6 * 64 * 64
    X = gc.vertices
    n = np.cross(X[v3[:, 1].astype(np.int64)] - X[v3[:, 0].astype(np.int64)],
                 X[v3[:, 2].astype(np.int64)] - X[v3[:, 0].astype(np.int64)])
    cen = X[v3.astype(np.int64)].mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", n, cen - 0.5) > 0)  # outward on the unit cube
    f = fh.assemble_surface_traction_3d(gc, v3, lambda p: (0.0, 0.0, 1.0)).reshape(-1, 3).sum(0)
    assert abs(f[2] - 6.0) < 1e-11
