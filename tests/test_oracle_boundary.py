"""Pins the oracle's boundary facets and boundary (ds) integrals (mesh.rs,
boundary.rs) to the reference's known answer and to closed forms.  CPU only.

Reference KAT: boundary.rs:143-159 (a unit flux integrates to the perimeter
4 of the unit square).  Independent pins: the facet list against the plain
boundary-facet listing; outward orientation (each normal points away from its
cell, the oriented areas of a closed surface cancel); the 2-point Gauss rule
is exact for g * phi up to cubic along an edge (checked against a 10-point
Gauss-Legendre integral); the midpoint traction rule gives each face vertex a
third of the face's exact integral of a linear traction; the Robin matrix is
gamma times the assembled 1D P1 mass of the boundary edges.
"""
import numpy as np

import oracle as O


def test_facets_full_2d_match_plain_listing():
    m = O.unit_square(5, 3)
    v, c, l = O.boundary_facets_full(m)
    assert np.array_equal(v, O.boundary_facets(m))
    for f, cell, loc in zip(v, c, l):
        t = m.cells[cell]
        edge = sorted((t[(loc + 1) % 3], t[(loc + 2) % 3]))
        assert list(f) == edge


def test_facets_full_3d_orientation():
    m = O.box_beam(1.0, 0.5, 0.25, 4, 3, 2)
    v, c, l = O.boundary_facets_full(m)
    assert np.array_equal(np.sort(v, axis=1), O.boundary_facets(m))
    X = m.vertices
    total = np.zeros(3)
    for f, cell, loc in zip(v, c, l):
        a, b, d = X[f[0]], X[f[1]], X[f[2]]
        n = np.cross(b - a, d - a)
        opp = X[m.cells[cell][loc]]
        assert n @ (a - opp) > 0  # away from the cell's opposite vertex
        assert set(int(x) for x in f) == set(int(x) for x in np.delete(m.cells[cell], loc))
        total += 0.5 * n
    assert np.abs(total).max() < 1e-14  # closed surface


def test_neumann_unit_flux_is_perimeter():  # boundary.rs:143-159
    m = O.unit_square(4, 4)
    v, _, _ = O.boundary_facets_full(m)
    assert abs(O.assemble_neumann_2d(m, v, lambda p: 1.0).sum() - 4.0) < 1e-12


def test_neumann_exact_for_quadratic_flux():
    m = O.unit_square(3, 5)
    v, _, _ = O.boundary_facets_full(m)
    g = lambda p: p[0] * p[0] - 2.0 * p[1] + 0.5 * p[0] * p[1]  # noqa: E731
    got = O.assemble_neumann_2d(m, v, g)
    xs, ws = np.polynomial.legendre.leggauss(10)
    s, w = 0.5 * (xs + 1), 0.5 * ws
    ref = np.zeros(m.nv)
    for a, b in v:
        pa, pb = m.vertices[a], m.vertices[b]
        L = np.linalg.norm(pb - pa)
        gq = np.array([g((1 - t) * pa + t * pb) for t in s])
        ref[a] += L * np.sum(w * gq * (1 - s))
        ref[b] += L * np.sum(w * gq * s)
    assert np.abs(got - ref).max() < 1e-13
    pts = O.neumann_points_2d(m, v)
    assert pts.shape == (len(v), 2, 2)


def test_surface_traction_linear_is_a_third_of_the_face_integral():
    m = O.box_beam(2.0, 1.0, 0.5, 3, 2, 2)
    v, _, _ = O.boundary_facets_full(m)
    t = lambda p: (1.0 + p[0], -2.0 * p[1], 0.5 * p[2] - p[0])  # noqa: E731
    got = O.assemble_surface_traction_3d(m, v, t)
    ref = np.zeros(3 * m.nv)
    X = m.vertices
    for f in v:
        a, b, d = X[f[0]], X[f[1]], X[f[2]]
        area = 0.5 * np.linalg.norm(np.cross(b - a, d - a))
        integral = area * np.array(t((a + b + d) / 3.0))  # exact for linear t
        for k in f:
            ref[3 * k:3 * k + 3] += integral / 3.0
    assert np.abs(got - ref).max() < 1e-13
    # constant traction: the total force is t times the boundary area
    tot = O.assemble_surface_traction_3d(m, v, lambda p: (0.0, 0.0, 1.0)).reshape(-1, 3).sum(0)
    area = 2 * (2.0 * 1.0 + 2.0 * 0.5 + 1.0 * 0.5)
    assert abs(tot[2] - area) < 1e-12 and abs(tot[0]) == 0.0


def test_robin_matrix_is_gamma_times_edge_mass():
    m = O.unit_square(4, 3)
    v, _, _ = O.boundary_facets_full(m)
    gamma = 2.5
    A = O.assemble_robin_matrix_2d(m, v, gamma).to_scipy().toarray()
    ref = np.zeros((m.nv, m.nv))
    for a, b in v:
        L = np.linalg.norm(m.vertices[b] - m.vertices[a])
        ref[np.ix_([a, b], [a, b])] += gamma * L / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
    assert np.abs(A - ref).max() < 1e-15
    one = np.ones(m.nv)
    assert abs(one @ A @ one - gamma * 4.0) < 1e-13  # gamma * perimeter
    # a subset of the boundary (the edge x = 1)
    right = np.array([f for f in v if np.all(np.abs(m.vertices[f, 0] - 1.0) < 1e-12)])
    Ar = O.assemble_robin_matrix_2d(m, right, 1.0).to_scipy()
    assert abs(one @ (Ar @ one) - 1.0) < 1e-14
