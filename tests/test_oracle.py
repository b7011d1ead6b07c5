"""Pins the CPU oracle (the checker) to the reference's own known answers and
to independent mathematics.  CPU only.

Reference KATs restated (file:line in fenics-rs):
  generators.rs:133-157, jacobian.rs:107-147, quadrature.rs:226-273,
  fenics-form/src/lib.rs:54-70, cg.rs:156-169, poisson_tests.rs:6-73,
  elasticity_tests.rs:6-126, ab_benchmarks.rs:13-100,
  benchmarks/dolfinx_reference.json.
Independent pins: scipy/sympy closed forms, rigid-body modes, a B^T D B
element matrix written in Voigt notation, direct sparse solves, linear
displacement fields for the stresses, finite differences for the adjoint.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PI = np.pi


def src2d(x):
    return 2 * PI * PI * np.sin(PI * x[..., 0]) * np.sin(PI * x[..., 1])


def exact2d(x):
    return np.sin(PI * x[..., 0]) * np.sin(PI * x[..., 1])


def solve_poisson(m, src, exact, tol, max_iter):
    A = O.assemble_stiffness(m)
    fq = src(O.quad_points(m, 2))
    b = O.assemble_rhs(m, fq)
    bd = O.boundary_dofs(m)
    g = exact(m.vertices[bd.astype(np.int64)])
    Ab, bb = O.apply_dirichlet_bc(A, b, bd, g)
    st, u, it, res = O.solve_cg(Ab, bb, None, tol, max_iter)
    assert st == 0
    return u, it, res


# ---------------------------------------------------------------- mesh KATs
def test_unit_square_properties():  # generators.rs:133-144
    m = O.unit_square(4, 4)
    assert (m.nv, m.nc) == (25, 32)
    assert abs(O.total_volume(m) - 1.0) < 1e-12
    assert abs(O.boundary_measure(m) - 4.0) < 1e-12
    assert len(O.boundary_facets(m)) == 16


def test_unit_cube_properties():  # generators.rs:146-157
    m = O.unit_cube(3, 3, 3)
    assert (m.nv, m.nc) == (64, 162)
    assert abs(O.total_volume(m) - 1.0) < 1e-12
    assert abs(O.boundary_measure(m) - 6.0) < 1e-12
    assert len(O.boundary_facets(m)) == 108


def test_box_beam_volume_and_kuhn_orientation():
    m = O.box_beam(10.0, 2.0, 2.0, 5, 2, 2)
    assert abs(O.total_volume(m) - 40.0) < 1e-10
    # all six Kuhn tetrahedra share the p0-p7 diagonal and have the same |det|
    v = m.vertices[m.cells.astype(np.int64)]
    det = np.linalg.det(v[:, 1:] - v[:, :1])
    assert np.allclose(np.abs(det), np.abs(det[0]))


# ---------------------------------------------------------------- element KATs
def test_scaled_triangle_mapping():  # jacobian.rs:107-127
    st, J, invT, det = O.affine_map([[1.0, 1.0], [3.0, 1.0], [1.0, 4.0]])
    assert st == 0 and abs(det - 6.0) < 1e-12
    assert np.allclose(invT @ np.array([6.0, 15.0]), [3.0, 5.0], atol=1e-12)
    assert np.allclose(np.array([1.0, 1.0]) + J @ np.array([0.5, 0.5]), [2.0, 2.5])


def test_scaled_tetrahedron_mapping():  # jacobian.rs:129-147
    st, J, invT, det = O.affine_map([[0, 0, 0], [2.0, 0, 0], [0, 3.0, 0], [0, 0, 4.0]])
    assert st == 0 and abs(det - 24.0) < 1e-12
    assert np.allclose(invT @ np.array([2.0, 6.0, 12.0]), [1, 2, 3], atol=1e-12)


def test_degenerate_simplex_rejected():  # jacobian.rs:29-31
    st, *_ = O.affine_map([[0, 0], [1, 1], [2, 2]])
    assert st != 0


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_quadrature_volumes(deg):  # quadrature.rs:226-245
    assert abs(O.quadrature(1, deg)[1].sum() - 1.0) < 1e-12
    assert abs(O.quadrature(2, deg)[1].sum() - 0.5) < 1e-12
    assert abs(O.quadrature(3, deg)[1].sum() - 1.0 / 6.0) < 1e-12


def test_polynomial_integration():  # quadrature.rs:247-273
    p, w = O.quadrature(2, 2)
    assert abs(((p[:, 0] ** 2 + 2 * p[:, 1]) * w).sum() - 5.0 / 12.0) < 1e-12
    p, w = O.quadrature(3, 2)
    assert abs(((p[:, 0] ** 2) * w).sum() - 1.0 / 60.0) < 1e-12


def test_p2_basis_kronecker_and_partition():  # lagrange.rs:480-559 style
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0],
                      [0, .5, 0], [0, 0, .5], [.5, 0, .5], [0, .5, .5]])
    B = np.array([O.p2_tet_basis(x) for x in nodes])
    assert np.allclose(B, np.eye(10), atol=1e-14)
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.dirichlet(np.ones(4))[1:]
        assert abs(O.p2_tet_basis(x).sum() - 1.0) < 1e-13
        assert np.allclose(O.p2_tet_grads(x).sum(axis=0), 0.0, atol=1e-13)


def test_p1_reference_triangle_local_matrices():  # fenics-form/src/lib.rs:54-70
    one = O.Mesh(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]]))
    K = O.assemble_stiffness(one).to_scipy().toarray()
    assert np.allclose(K, [[1, -.5, -.5], [-.5, .5, 0], [-.5, 0, .5]], atol=1e-12)
    M = O.assemble_mass_2d(one).to_scipy().toarray()
    assert np.allclose(np.diag(M), 1 / 12) and abs(M[0, 1] - 1 / 24) < 1e-12


# ---------------------------------------------------------------- CSC container
def test_csc_matches_scipy_canonical_form():
    rng = np.random.default_rng(1)
    n, nt = 40, 600
    r, c = rng.integers(0, n, nt), rng.integers(0, n, nt)
    v = rng.standard_normal(nt)
    v[::7] = 0.0  # explicit zeros must be kept
    A = O.csc_from_triplets(n, r, c, v)
    cp, ri, va = A.export()
    S = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(cp, S.indptr) and np.array_equal(ri, S.indices)
    assert np.allclose(va, S.data, rtol=1e-14, atol=1e-15)
    for j in range(n):  # rows sorted within each column
        assert np.all(np.diff(ri[cp[j]:cp[j + 1]].astype(np.int64)) > 0)


def test_csc_rejects_out_of_bounds():  # faer try_new_from_triplets Err
    with pytest.raises(ValueError):
        O.csc_from_triplets(3, [0, 3], [0, 0], [1.0, 1.0])


# ---------------------------------------------------------------- CG
def test_small_spd_system():  # cg.rs:156-169
    A = O.csc_from_triplets(2, [0, 0, 1, 1], [0, 1, 0, 1], [4.0, -1.0, -1.0, 4.0])
    st, x, it, res = O.solve_cg(A, [3.0, 3.0], None, 1e-10, 100)
    assert st == 0 and it <= 2 and res < 1e-10 and np.allclose(x, 1.0, atol=1e-8)


def test_cg_matches_direct_solve_and_status_codes():
    m = O.unit_square(12, 12)
    A = O.assemble_stiffness(m)
    bd = O.boundary_dofs(m)
    b = O.assemble_rhs(m, np.ones((m.nc, 3)))
    Ab, bb = O.apply_dirichlet_bc(A, b, bd, np.zeros(len(bd)))
    st, x, it, res = O.solve_cg(Ab, bb, None, 1e-12, 1000)
    xd = spla.spsolve(Ab.to_scipy().tocsr(), bb)
    assert st == 0 and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    assert O.solve_cg(Ab, bb, None, 1e-12, 2)[0] == 2  # not converged
    st, x0, it, res = O.solve_cg(Ab, np.zeros(Ab.n), None, 1e-12, 10)
    assert st == 0 and it == 0 and not x0.any()  # ||b|| = 0 -> x = 0
    Z = O.csc_from_triplets(2, [0, 1], [0, 1], [0.0, 0.0])
    assert O.solve_cg(Z, [1.0, 1.0], None, 1e-10, 10)[0] == 1  # breakdown


def test_csc_to_csr_inverse_diagonal():  # gpu.rs:1552-1583
    A = O.assemble_stiffness(O.unit_square(5, 5))
    rp, ci, va, inv = A.to_csr()
    S = A.to_scipy().tocsr()
    S.sort_indices()
    assert np.array_equal(rp, S.indptr) and np.array_equal(ci, S.indices)
    assert np.array_equal(va, S.data)
    assert np.allclose(inv, 1.0 / S.diagonal())


# ---------------------------------------------------------------- Poisson (C0 and KATs)
def test_c0_counts_and_iterations():
    m = O.unit_square(128, 128)
    A = O.assemble_stiffness(m)
    assert (m.nv, m.nc, A.nnz) == (16641, 32768, 115457)
    bd = O.boundary_dofs(m)
    assert len(bd) == 512
    b = O.assemble_rhs(m, src2d(O.quad_points(m, 2)))
    Ab, bb = O.apply_dirichlet_bc(A, b, bd, np.zeros(len(bd)))
    assert Ab.nnz == 112401
    st, u, it, res = O.solve_cg(Ab, bb, None, 1e-12, 5000)
    assert st == 0 and it == 223 and abs(u.max() - 0.99995) < 1e-4


def test_stiffness_kernel_and_symmetry():
    m = O.unit_cube(4, 4, 4)
    K = O.assemble_stiffness(m).to_scipy()
    assert np.abs(K @ np.ones(m.nv)).max() < 1e-13  # constants in the kernel
    assert abs(K - K.T).max() < 1e-15


def test_rhs_integrates_source():
    m = O.unit_square(7, 5)
    b = O.assemble_rhs(m, np.full((m.nc, 3), 3.0))
    assert abs(b.sum() - 3.0) < 1e-12  # sum_i b_i = int f
    m3 = O.unit_cube(3, 2, 2)
    b3 = O.assemble_rhs(m3, np.full((m3.nc, 4), 2.0))
    assert abs(b3.sum() - 2.0) < 1e-12


def test_poisson_2d_manufactured_rate():  # poisson_tests.rs:6-52
    errs = []
    for n in (8, 16, 32):
        m = O.unit_square(n, n)
        u, it, res = solve_poisson(m, src2d, exact2d, 1e-10, 500)
        errs.append(O.l2_error(m, u, exact2d(O.quad_points(m, 3))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.8 < r1 < 4.2 and 3.8 < r2 < 4.2


def test_poisson_3d_manufactured():  # poisson_tests.rs:55-73
    ex = lambda x: np.sin(PI * x[..., 0]) * np.sin(PI * x[..., 1]) * np.sin(PI * x[..., 2])
    src = lambda x: 3 * PI * PI * ex(x)
    m = O.unit_cube(8, 8, 8)
    u, it, res = solve_poisson(m, src, ex, 1e-10, 1000)
    assert O.l2_error(m, u, ex(O.quad_points(m, 3))) < 0.05


def test_ab_poisson_2d_vs_dolfinx():  # ab_benchmarks.rs:13-48
    ref = json.load(open(os.path.join(GOLDEN, "dolfinx_reference.json")))["poisson_2d"]
    m = O.unit_square(16, 16)
    u, it, res = solve_poisson(m, src2d, exact2d, 1e-12, 1000)
    l2 = O.l2_error(m, u, exact2d(O.quad_points(m, 3)))
    assert abs(l2 - ref["l2_error"]) / ref["l2_error"] < 0.05
    assert abs(u.max() - ref["max_value"]) / ref["max_value"] < 1e-3


# ---------------------------------------------------------------- elasticity
def cantilever(m, E, nu, f, tol, max_iter):
    K = O.assemble_elasticity(m, E, nu)
    b = O.assemble_body_force(m, f)
    fx = O.clamp_dofs_x0(m)
    Kb, bb = O.apply_dirichlet_bc(K, b, fx, np.zeros(len(fx)))
    st, u, it, res = O.solve_cg(Kb, bb, None, tol, max_iter)
    assert st == 0
    return u


def test_ab_cantilever_vs_dolfinx():  # ab_benchmarks.rs:51-100
    ref = json.load(open(os.path.join(GOLDEN, "dolfinx_reference.json")))["elasticity_3d"]
    m = O.box_beam(1.0, 0.1, 0.1, 10, 2, 2)
    u = cantilever(m, 1e7, 0.3, [0.0, -1000.0, 0.0], 1e-8, 1000)
    tip = np.abs(m.vertices[:, 0] - 1.0) < 1e-6
    uy = u.reshape(-1, 3)[tip, 1].mean()
    assert abs(uy - ref["tip_uy"]) / abs(ref["tip_uy"]) < 0.01


def test_axial_bar_extension():  # elasticity_tests.rs:6-49
    m = O.box_beam(1.0, 0.1, 0.1, 10, 2, 2)
    u = cantilever(m, 1.0e6, 0.0, [1000.0, 0.0, 0.0], 1e-10, 1000)
    tip = np.abs(m.vertices[:, 0] - 1.0) < 1e-6
    assert abs(u.reshape(-1, 3)[tip, 0].mean() - 5e-4) / 5e-4 < 0.03


def test_cantilever_beam_bending():  # elasticity_tests.rs:52-126
    m = O.box_beam(1.0, 0.1, 0.1, 20, 3, 3)
    u = cantilever(m, 1.0e6, 0.3, [0.0, 0.0, -1000.0], 1e-9, 2000)
    tip = np.abs(m.vertices[:, 0] - 1.0) < 1e-6
    uz = u.reshape(-1, 3)[tip, 2].mean()
    assert abs(uz - (-0.0890)) / 0.0890 < 0.06 and uz < 0
    s = O.element_stresses(m, u, 1.0e6, 0.3)
    assert O.vertex_von_mises(m, s[:, 18]).max() > 1e4


def voigt_element(x, E, nu):
    """Independent 12x12 element matrix: vol * B^T D B in Voigt notation."""
    J = (x[1:] - x[0]).T
    g = np.vstack([-np.linalg.inv(J).sum(axis=0), np.linalg.inv(J)])  # rows: grad phi_i
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    B = np.zeros((6, 12))
    for i in range(4):
        gx, gy, gz = g[i]
        B[:, 3 * i:3 * i + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz],
                                 [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
    return abs(np.linalg.det(J)) / 6 * B.T @ D @ B


def test_elasticity_element_matches_voigt_bdb():
    rng = np.random.default_rng(3)
    x = rng.random((4, 3))
    m = O.Mesh(x, np.array([[0, 1, 2, 3]]))
    K = O.assemble_elasticity(m, 2.5e3, 0.27).to_scipy().toarray()
    Kv = voigt_element(x, 2.5e3, 0.27)
    assert np.abs(K - Kv).max() < 1e-10 * np.abs(Kv).max()


def test_elasticity_rigid_modes_and_homogeneity():
    m = O.box_beam(2.0, 0.5, 0.5, 4, 2, 2)
    K = O.assemble_elasticity(m, 7.0, 0.3).to_scipy()
    X = m.vertices
    modes = []
    for a in range(3):
        t = np.zeros_like(X)
        t[:, a] = 1
        modes.append(t.reshape(-1))
    for a, b in [(0, 1), (1, 2), (2, 0)]:
        r = np.zeros_like(X)
        r[:, a], r[:, b] = -X[:, b], X[:, a]
        modes.append(r.reshape(-1))
    scale = abs(K).max()
    for r in modes:
        assert np.abs(K @ r).max() < 1e-11 * scale * np.abs(r).max() * 10
    E = np.full(m.nc, 7.0)
    Kc = O.assemble_elasticity(m, E, 0.3).to_scipy()
    assert np.array_equal(Kc.data, K.data)  # per-cell E == uniform E bitwise
    K2 = O.assemble_elasticity(m, 14.0, 0.3).to_scipy()
    assert np.allclose(K2.data, 2 * K.data, rtol=1e-14)


def test_body_force_total():
    m = O.box_beam(10.0, 2.0, 2.0, 6, 2, 2)
    b = O.assemble_body_force(m, [0.0, 0.0, -7.85e4]).reshape(-1, 3)
    assert abs(b[:, 2].sum() - (-7.85e4 * 40.0)) < 1e-6 * 7.85e4 * 40
    assert not b[:, :2].any()


def test_dirichlet_condensation_matches_definition():
    m = O.unit_square(6, 6)
    A = O.assemble_stiffness(m)
    b = np.random.default_rng(4).standard_normal(m.nv)
    bd = O.boundary_dofs(m)
    g = np.linspace(1.0, 2.0, len(bd))
    Ab, bb = O.apply_dirichlet_bc(A, b, bd, g)
    S = A.to_scipy().toarray()
    fixed = np.zeros(m.nv, bool)
    fixed[bd.astype(np.int64)] = True
    ref = S.copy()
    ref[fixed, :] = 0
    ref[:, fixed] = 0
    ref[fixed, fixed] = 1
    gb = np.zeros(m.nv)
    gb[bd.astype(np.int64)] = g
    rb = b - S[:, fixed] @ gb[fixed]
    rb[fixed] = g
    assert np.allclose(Ab.to_scipy().toarray(), ref)
    assert np.allclose(bb, rb, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- P2
def test_p2_counts_and_closed_form_stiffness():
    import sympy as S
    one = O.Mesh(np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]), np.array([[0, 1, 2, 3]]))
    p2 = O.P2(one)
    assert (p2.ndofs, p2.nedges) == (10, 6)
    K = p2.stiffness().to_scipy().toarray()
    x, y, z = S.symbols("x y z")
    l = [1 - x - y - z, x, y, z]
    phi = [li * (2 * li - 1) for li in l] + [4 * l[a] * l[b] for a, b in
                                             [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]]
    cd = p2.cell_dofs()[0].astype(int)
    Kex = np.zeros((10, 10))
    for i in range(10):
        for j in range(10):
            e = sum(S.diff(phi[i], v) * S.diff(phi[j], v) for v in (x, y, z))
            val = S.integrate(S.integrate(S.integrate(e, (z, 0, 1 - x - y)), (y, 0, 1 - x)), (x, 0, 1))
            Kex[cd[i], cd[j]] = float(val)
    assert np.abs(K - Kex).max() < 1e-14


def test_p2_dofmap_counts_unit_cube():
    m = O.unit_cube(4, 4, 4)
    p2 = O.P2(m)
    assert p2.ndofs == 9 ** 3  # P2 on a 4^3 Kuhn mesh = vertices of the 8^3 grid
    bd = p2.boundary_dofs()
    assert len(bd) == 9 ** 3 - 7 ** 3
    K = p2.stiffness().to_scipy()
    assert np.abs(K @ np.ones(p2.ndofs)).max() < 1e-12


def test_p2_edge_numbering_convention():
    """P2 dofs: vertices keep their numbers; edge dofs follow ordered by the
    number of cells around the edge (most first), then by max - min, then by
    min; every edge's midpoint is its dof coordinate."""
    m = O.unit_cube(3, 4, 2)
    p2 = O.P2(m)
    cd = p2.cell_dofs().reshape(-1, 10).astype(np.int64)
    C = m.cells.astype(np.int64)
    assert np.array_equal(cd[:, :4], C)
    pairs = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    ne = p2.ndofs - m.nv
    key = np.zeros((ne, 2), dtype=np.int64)
    for k, (i, j) in enumerate(pairs):
        e = cd[:, 4 + k] - m.nv
        key[e, 0] = np.minimum(C[:, i], C[:, j])
        key[e, 1] = np.maximum(C[:, i], C[:, j])
    around = np.bincount((cd[:, 4:] - m.nv).ravel(), minlength=ne)
    order = sorted(range(ne), key=lambda e: (-around[e], key[e, 1] - key[e, 0], key[e, 0]))
    assert order == list(range(ne))
    assert len(np.unique(key, axis=0)) == ne
    X = p2.dof_coords()
    V = m.vertices
    assert np.array_equal(X[m.nv:], (V[key[:, 0]] + V[key[:, 1]]) * 0.5)


def test_p2_quadratic_patch():
    """A quadratic u with -Lap u = 6 (u = x^2+y^2+z^2) is reproduced exactly."""
    m = O.unit_cube(3, 3, 3)
    p2 = O.P2(m)
    X = p2.dof_coords()
    ue = (X ** 2).sum(axis=1)
    K = p2.stiffness()
    b = p2.rhs_cellconst(np.full(m.nc, -6.0))
    bd = p2.boundary_dofs()
    Kb, bb = O.apply_dirichlet_bc(K, b, bd, ue[bd.astype(np.int64)])
    st, u, it, res = O.solve_cg(Kb, bb, None, 1e-13, 2000)
    assert st == 0 and np.abs(u - ue).max() < 1e-9


# ---------------------------------------------------------------- adjoint / stresses
def test_adjoint_gradient_fd_and_homogeneity():
    m = O.box_beam(1.0, 0.2, 0.2, 4, 2, 2)
    rng = np.random.default_rng(5)
    E = 1e3 * np.exp(0.3 * rng.standard_normal(m.nc))
    nu = 0.3
    fb = [0.0, 0.0, -10.0]
    fx = O.clamp_dofs_x0(m)

    def solve(Ec):
        K = O.assemble_elasticity(m, Ec, nu)
        b = O.assemble_body_force(m, fb)
        Kb, bb = O.apply_dirichlet_bc(K, b, fx, np.zeros(len(fx)))
        u = spla.spsolve(Kb.to_scipy().tocsc(), bb)
        return u, bb

    u, f = solve(E)
    J = f @ u
    g = O.cell_bilinear_k0(m, nu, u, u)
    assert abs((E * g).sum() + J) < 1e-10 * abs(J)  # sum_c E_c dJ/dE_c = -J
    for c in rng.choice(m.nc, 6, replace=False):
        h = 1e-3 * E[c]
        Ep, Em = E.copy(), E.copy()
        Ep[c] += h
        Em[c] -= h
        fd = (f @ solve(Ep)[0] - f @ solve(Em)[0]) / (2 * h)
        assert abs(fd - g[c]) < 1e-5 * abs(g[c])


def test_stresses_of_linear_displacement():
    m = O.box_beam(1.0, 1.0, 1.0, 2, 2, 2)
    G = np.array([[1e-3, 2e-4, 0.0], [0.0, -5e-4, 1e-4], [3e-4, 0.0, 2e-4]])
    u = (m.vertices @ G.T).reshape(-1)
    E, nu = 210e9, 0.3
    s = O.element_stresses(m, u, E, nu)
    eps = 0.5 * (G + G.T)
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    assert np.allclose(s[:, :9], sig.reshape(-1), rtol=1e-9)
    assert np.allclose(s[:, 9:18], eps.reshape(-1), rtol=1e-9)
    d = sig
    vm = np.sqrt(0.5 * ((d[0, 0] - d[1, 1]) ** 2 + (d[1, 1] - d[2, 2]) ** 2 + (d[2, 2] - d[0, 0]) ** 2
                        + 6 * (d[0, 1] ** 2 + d[1, 2] ** 2 + d[2, 0] ** 2)))
    assert np.allclose(s[:, 18], vm, rtol=1e-9)
    assert np.allclose(O.vertex_von_mises(m, s[:, 18]), vm, rtol=1e-9)


# ---------------------------------------------------------------- golden fixtures
def test_golden_fixtures_match_oracle():
    g = np.load(os.path.join(GOLDEN, "p1_unit_square_4x4.npz"))
    m = O.unit_square(4, 4)
    cp, ri, va = O.assemble_stiffness(m).export()
    assert np.array_equal(cp, g["col_ptr"]) and np.array_equal(ri, g["row_idx"])
    assert np.array_equal(va.view(np.uint64), g["values"].view(np.uint64))
    e = np.load(os.path.join(GOLDEN, "elasticity_box_beam_2x1x1.npz"))
    mb = O.box_beam(1.0, 0.1, 0.1, 2, 1, 1)
    K = O.assemble_elasticity(mb, 1e7, 0.3)
    cp, ri, va = K.export()
    assert np.array_equal(cp, e["col_ptr"]) and np.array_equal(ri, e["row_idx"])
    assert np.array_equal(va.view(np.uint64), e["values"].view(np.uint64))
    # the same matrix against the independent Voigt form, element by element
    Kd = np.zeros((3 * mb.nv, 3 * mb.nv))
    for cell in mb.cells.astype(np.int64):
        Ke = voigt_element(mb.vertices[cell], 1e7, 0.3)
        idx = (3 * cell[:, None] + np.arange(3)).reshape(-1)
        Kd[np.ix_(idx, idx)] += Ke
    assert np.abs(K.to_scipy().toarray() - Kd).max() < 1e-9 * np.abs(Kd).max()


# ---------------------------------------------------------------- modal analysis
def _cantilever_km(nx, ny, nz, rho=1000.0):
    m = O.box_beam(1.0, 0.1, 0.1, nx, ny, nz)
    return m, O.assemble_elasticity(m, 1.0e7, 0.3), O.assemble_elasticity_mass_3d(m, rho)


def test_elasticity_mass_3d_invariants():
    """mass.rs:86-126: sum of all entries = 3 rho V (three components), each
    component block equals the scalar P1 mass (vol/10 diagonal, vol/20 off),
    symmetric, no coupling between components."""
    m, _, M = _cantilever_km(4, 2, 2, rho=7.5)
    S = M.to_scipy().tocoo()
    assert abs(S.sum() - 3 * 7.5 * 0.01) < 1e-15 * 3 * 7.5
    assert np.all(S.row % 3 == S.col % 3)
    D = M.to_scipy().toarray()
    assert np.array_equal(D, D.T)
    # one tetrahedron by hand
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    t = O.Mesh(V, np.array([[0, 1, 2, 3]], np.uint64))
    Mt = O.assemble_elasticity_mass_3d(t, 2.0).to_scipy().toarray()
    vol = 1.0 / 6.0
    assert abs(Mt[0, 0] - 2.0 * vol / 10.0) < 1e-17 and abs(Mt[0, 3] - 2.0 * vol / 20.0) < 1e-17
    assert Mt[0, 1] == 0.0 and Mt[0, 4] == 0.0


def test_fold_dirichlet_definition():
    """eigen.rs:51-78: constrained rows/columns dropped, unit diagonal when asked."""
    m, K, M = _cantilever_km(3, 1, 1)
    fx = O.clamp_dofs_x0(m)
    Kd = O.fold_dirichlet(K, fx, True).to_scipy().toarray()
    Md = O.fold_dirichlet(M, fx, False).to_scipy().toarray()
    Kf, Mf = K.to_scipy().toarray(), M.to_scipy().toarray()
    free = np.setdiff1d(np.arange(K.n), fx)
    assert np.array_equal(Kd[np.ix_(free, free)], Kf[np.ix_(free, free)])
    assert np.array_equal(Kd[np.ix_(fx, fx)], np.eye(len(fx)))
    assert not Kd[np.ix_(fx, free)].any() and not Md[fx].any() and not Md[:, fx].any()


def test_cantilever_natural_frequencies():  # modal_tests.rs:9-86
    m, K, M = _cantilever_km(16, 2, 2)
    st, f, om, ev, _ = O.vibration_modes(K, M, O.clamp_dofs_x0(m), 3, 30, 1e-5)
    assert st == 0
    assert 1.0 < f[0] < 3.0 and 1.5 < f[1] < 3.5 and 8.0 < f[2] < 20.0
    assert np.allclose(om, 2 * np.pi * f) and np.allclose(ev, om ** 2)


def test_frequencies_scale_as_one_over_sqrt_density():  # modal_tests.rs:88-113
    m, K, M1 = _cantilever_km(8, 1, 1)
    M2 = O.assemble_elasticity_mass_3d(m, 1000.0 * 10000.0)
    fx = O.clamp_dofs_x0(m)
    a = O.vibration_modes(K, M1, fx, 2, 60, 1e-8)[1]
    b = O.vibration_modes(K, M2, fx, 2, 60, 1e-8)[1]
    assert np.allclose(b / a, 0.01, rtol=0, atol=1e-4)


def test_vibration_modes_vs_shift_invert_lanczos():
    """Independent pin: scipy's shift-invert eigsh on the free-dof generalized
    problem K v = lambda M v gives the same lowest eigenvalues (1e-9)."""
    import scipy.sparse.linalg as sla
    m, K, M = _cantilever_km(12, 2, 2)
    fx = O.clamp_dofs_x0(m)
    st, f, om, ev, sh = O.vibration_modes(K, M, fx, 3, 300, 1e-13, inner_tol=1e-12,
                                          inner_max_iter=5000)
    free = np.setdiff1d(np.arange(K.n), fx)
    Kf = K.to_scipy().tocsc()[free][:, free]
    Mf = M.to_scipy().tocsc()[free][:, free]
    lam = np.sort(sla.eigsh(Kf, k=3, M=Mf, sigma=0, which="LM")[0])
    assert st == 0 and np.allclose(ev, lam, rtol=1e-9, atol=0)
    # modes are M-orthogonal with one common norm (normalised in the rescaled
    # mass s M, eigen.rs:172-181, so v^T M v = 1/s) and vanish on the clamped dofs
    Ms = M.to_scipy()
    G = sh @ (Ms @ sh.T)
    assert np.allclose(G / G[0, 0], np.eye(3), atol=1e-6) and not sh[:, fx].any()


def test_p2_mass_products_closed_forms():  # mass.rs:270-289
    P = O.p2_mass_products()
    for (a, b), v in {(0, 0): 1 / 420, (4, 4): 4 / 315, (0, 1): 1 / 2520, (0, 4): -1 / 630,
                      (4, 6): 2 / 315, (4, 9): 1 / 315}.items():
        assert abs(P[a, b] - v) < 1e-14
    rows = P.sum(axis=1)
    assert np.allclose(rows[:4], -1 / 120, rtol=0, atol=1e-14)
    assert np.allclose(rows[4:], 1 / 30, rtol=0, atol=1e-14)
    assert abs(P.sum() - 1 / 6) < 1e-13 and np.array_equal(P, P.T)


def test_p2_vector_mass_one_tet_kat():  # mass.rs:296-330
    nodes = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0],
             [0, 0, 0.5], [0.5, 0, 0.5], [0, 0.5, 0.5]]
    M = O.assemble_elasticity_mass_p2_3d(nodes, [list(range(10))], 1.0).to_scipy()
    one = np.ones(30)
    assert abs(one @ (M @ one) - 0.5) < 1e-13
    assert abs(M[0, 0] - 1 / 420) < 1e-13
    assert M.nnz == 10 * 10 * 3  # equal components only


def test_p2_vector_mass_integrates_quadratics():
    """sum_i (M u)_i = rho int u for the P2 interpolant of a quadratic field
    (sum_i phi_i = 1), component by component, on a box of 2 x 1 x 0.5."""
    m = O.box_beam(2.0, 1.0, 0.5, 3, 2, 2)
    p2 = O.P2(m)
    X = p2.dof_coords()
    M = O.assemble_elasticity_mass_p2_3d(X, p2.cell_dofs(), 7.0).to_scipy()
    u = np.zeros((len(X), 3))
    u[:, 0] = X[:, 0] ** 2
    u[:, 1] = X[:, 1] * X[:, 2]
    u[:, 2] = 1.0 + X[:, 2] ** 2
    tot = (M @ u.reshape(-1)).reshape(-1, 3).sum(axis=0)
    vol = 2.0 * 1.0 * 0.5
    exact = 7.0 * np.array([vol * 4 / 3, vol * 0.5 * 0.25, vol * (1 + 0.25 / 3)])
    assert np.allclose(tot, exact, rtol=1e-13, atol=0)


def test_exact_residual_checker_vs_fractions():
    """The exact-residual checker of the GPU residual tests agrees with a
    rational-arithmetic evaluation on a small cantilever."""
    from fractions import Fraction
    from exactsum import exact_residual
    om = O.box_beam(1.0, 0.2, 0.2, 4, 1, 1)
    K = O.assemble_elasticity(om, 210e9, 0.3)
    b = O.assemble_body_force(om, [0, 0, -7.85e4])
    fx = O.clamp_dofs_x0(om)
    Kb, bb = O.apply_dirichlet_bc(K, b, fx, np.zeros(len(fx)))
    st, uo, it, res = O.solve_cg(Kb, bb, None, 1e-13, 2000)
    cp, ri, va = Kb.export()
    for x in (uo, np.random.default_rng(1).standard_normal(len(uo))):
        r = [Fraction(v) for v in bb]
        for c in range(len(cp) - 1):
            for e in range(int(cp[c]), int(cp[c + 1])):
                r[int(ri[e])] -= Fraction(va[e]) * Fraction(x[c])
        rf = np.array([float(v) for v in r]).astype(np.longdouble)
        ref = float(np.sqrt((rf ** 2).sum()) / np.sqrt((bb.astype(np.longdouble) ** 2).sum()))
        assert exact_residual(Kb, bb, x) == ref


def test_adjoint_contraction_vs_exact_rationals():
    """-w_c^T K0_c u_c (the adjoint kernel's quantity) against exact rational
    arithmetic: the affine map, gradients, Lame constants and the
    elasticity.rs:133-136 element matrix contracted with w and u, all in
    Fractions from the FP64 inputs.  Fields dominated by a rigid motion (the
    situation near a cantilever's free end) leave a strain energy many orders
    below |u|^2 |K0|; the oracle's strain-form evaluation loses only about
    eps * |grad u| / |strain| of it (the rotation left in the gradient), where
    contracting the FP64 element matrix itself loses about
    eps * |u|^2 |K0| / (u^T K0 u), the square of that ratio times more."""
    from fractions import Fraction as Fr
    m = O.box_beam(1.0, 0.4, 0.3, 2, 2, 1)
    nu = 0.3
    rng = np.random.default_rng(8)
    R = np.array([[0.0, -1.0, 0.5], [1.0, 0.0, -0.2], [-0.5, 0.2, 0.0]])  # skew: a rotation
    X = m.vertices
    for scale in (1.0, 1e-6, 1e-9):
        u = (np.array([3.0, -2.0, 1.0]) + X @ R.T) + scale * rng.standard_normal(X.shape)
        u = u.reshape(-1)
        w = u.copy()
        got = O.cell_bilinear_k0(m, nu, w, u)
        nuq = Fr(nu)
        lam = nuq / ((1 + nuq) * (1 - 2 * nuq))
        mu = Fr(1) / (2 * (1 + nuq))
        for c in range(m.nc):
            t = [int(v) for v in m.cells[c]]
            P = [[Fr(float(X[v][k])) for k in range(3)] for v in t]
            J = [[P[j + 1][i] - P[0][i] for j in range(3)] for i in range(3)]
            det = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2])
                   - J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2])
                   + J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]))
            cof = [[(J[(i + 1) % 3][(j + 1) % 3] * J[(i + 2) % 3][(j + 2) % 3]
                     - J[(i + 1) % 3][(j + 2) % 3] * J[(i + 2) % 3][(j + 1) % 3])
                    for j in range(3)] for i in range(3)]
            invT = [[cof[i][j] / det for j in range(3)] for i in range(3)]  # J^{-T}
            G = [[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
            g = [[sum(invT[i][k] * G[a][k] for k in range(3)) for i in range(3)] for a in range(4)]
            vol = abs(det) / 6
            acc = Fr(0)
            for i in range(4):
                for j in range(4):
                    gg = sum(g[i][k] * g[j][k] for k in range(3))
                    for a in range(3):
                        for b in range(3):
                            K = vol * (lam * g[j][b] * g[i][a] + mu * gg * (1 if a == b else 0)
                                       + mu * g[j][a] * g[i][b])
                            acc += Fr(float(w[3 * t[i] + a])) * K * Fr(float(u[3 * t[j] + b]))
            ref = -float(acc)
            # the strain form's own bound: rounding of G (entries of the size
            # of the rotation) relative to its symmetric part (the strain)
            tol = 8 * 2.2e-16 * (1.0 + 1.0 / scale)
            assert abs(got[c] - ref) <= tol * abs(ref), (scale, c, got[c], ref)
