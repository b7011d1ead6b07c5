"""GPU parity of the assembly path against the CPU oracle.

Integer data (meshes, boundary dofs, CSC patterns, P2 dofmaps) must be
identical; FP64 values must be bit-identical (the GPU sums every slot's
duplicates in ascending cell order with the oracle's operation order and no
FMA).  Full-size configurations are checked through exact counts and
size-independent invariants."""
import json
import os

import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_csc(gpu_mat, ora_mat):
    cg, rg, vg = gpu_mat.export_csc()
    co, ro, vo = ora_mat.export()
    assert np.array_equal(cg, co), "col_ptr differs"
    assert np.array_equal(rg, ro), "row_idx differs"
    assert np.array_equal(bits(vg), bits(vo)), \
        f"values differ (max |d| {np.abs(vg - vo).max():.3e})"


def gpu_mesh_equals_oracle(gm, om):
    assert np.array_equal(bits(gm.vertices), bits(om.vertices))
    assert np.array_equal(gm.cells, om.cells)


@pytest.mark.parametrize("nx,ny", [(1, 1), (4, 4), (7, 3), (32, 32)])
def test_unit_square_bitwise(nx, ny):
    gm, om = fh.unit_square(nx, ny), O.unit_square(nx, ny)
    gpu_mesh_equals_oracle(gm, om)
    assert np.array_equal(gm.boundary_dofs(), O.boundary_dofs(om))
    same_csc(fh.assemble_stiffness_2d(gm), O.assemble_stiffness(om))
    same_csc(fh.assemble_mass(gm), O.assemble_mass_2d(om))


@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0, 3, 3, 3), (10.0, 2.0, 2.0, 8, 3, 2)])
def test_box_beam_bitwise(dims):
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    gpu_mesh_equals_oracle(gm, om)
    assert np.array_equal(gm.boundary_dofs(), O.boundary_dofs(om))
    same_csc(fh.assemble_stiffness_3d(gm), O.assemble_stiffness(om))


def test_golden_p1_fixture():
    g = np.load(os.path.join(GOLDEN, "p1_unit_square_4x4.npz"))
    cp, ri, va = fh.assemble_stiffness(fh.unit_square(4, 4)).export_csc()
    assert np.array_equal(cp, g["col_ptr"]) and np.array_equal(ri, g["row_idx"])
    assert np.array_equal(bits(va), bits(g["values"]))


def test_golden_elasticity_fixture():
    g = np.load(os.path.join(GOLDEN, "elasticity_box_beam_2x1x1.npz"))
    m = fh.box_beam(1.0, 0.1, 0.1, 2, 1, 1)
    cp, ri, va = fh.assemble_elasticity_stiffness_3d(m, fh.ElasticMaterial(1e7, 0.3)).export_csc()
    assert np.array_equal(cp, g["col_ptr"]) and np.array_equal(ri, g["row_idx"])
    assert np.array_equal(bits(va), bits(g["values"]))


def test_mesh_from_arrays_roundtrip():
    om = O.box_beam(2.0, 1.0, 1.0, 3, 2, 2)
    gm = fh.Mesh.from_arrays(om.vertices, om.cells)
    gpu_mesh_equals_oracle(gm, om)
    same_csc(fh.assemble_stiffness(gm), O.assemble_stiffness(om))


def test_quadrature_points_and_rhs_bitwise():
    for gm, om in [(fh.unit_square(9, 5), O.unit_square(9, 5)),
                   (fh.unit_cube(3, 2, 4), O.unit_cube(3, 2, 4))]:
        xg, xo = gm.quad_points(2), O.quad_points(om, 2)
        assert np.array_equal(bits(xg), bits(xo))
        fq = np.sin(3 * xo[..., 0]) + xo[..., 1] ** 2
        assert np.array_equal(bits(fh.assemble_rhs(gm, fq)), bits(O.assemble_rhs(om, fq)))


@pytest.mark.parametrize("E", [210e9, "cells"])
def test_elasticity_bitwise(E):
    dims = (1.0, 0.1, 0.1, 6, 2, 2)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    if E == "cells":
        Ec = 210e9 * np.exp(0.3 * np.random.default_rng(7).standard_normal(om.nc))
        g = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(210e9, 0.3), E_cell=Ec)
        o = O.assemble_elasticity(om, Ec, 0.3)
    else:
        g = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(E, 0.3))
        o = O.assemble_elasticity(om, E, 0.3)
    same_csc(g, o)
    f = [0.3, -1.0, -7.85e4]
    assert np.array_equal(bits(fh.assemble_elasticity_body_force_3d(gm, f)),
                          bits(O.assemble_body_force(om, f)))


@pytest.mark.parametrize("E", [210e9, "cells"])
def test_elasticity_mfma_matches_fp64(E):
    """FP32 element matrices B^T (vol D) B on the MFMA units (method 1),
    summed per row in FP32, against the FP64 assembly: the same pattern, every
    value within 2e-6 of the largest entry (FP32 rounding of the gradients,
    vol, lambda and mu, of each entry's six-term MFMA sum, and of up to 24
    additions per stored value; one cell's contribution is about max|K| / 24,
    so a dropped, doubled or misplaced cell fails it), rigid translations in
    the kernel to FP32 rounding.  Method 2 keeps the FP64 values bitwise the
    oracle's and carries the MFMA values as its FP32 copy."""
    dims = (1.0, 0.3, 0.2, 9, 4, 3)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    mat = fh.ElasticMaterial(2e11, 0.3)
    ec = None if E != "cells" else np.random.default_rng(5).uniform(1e10, 3e11, om.nc)
    c1, r1, v1 = fh.assemble_elasticity_stiffness_3d(gm, mat, ec).export_csc()
    K32 = fh.assemble_elasticity_stiffness_3d(gm, mat, ec, method=1)
    c2, r2, v2 = K32.export_csc()
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    assert np.abs(v2 - v1).max() <= 2e-6 * np.abs(v1).max()
    assert np.array_equal(v2, v2.astype(np.float32).astype(np.float64))
    for a in range(3):
        t = np.zeros((gm.nv, 3))
        t[:, a] = 1.0
        assert np.abs(K32.spmv(t.reshape(-1))).max() <= 1e-5 * np.abs(v1).max()
    K2 = fh.assemble_elasticity_stiffness_3d(gm, mat, ec, method=2)
    same_csc(K2, O.assemble_elasticity(om, ec if ec is not None else 2e11, 0.3))


def test_elasticity_mfma_one_cell_and_errors():
    """One tetrahedron: a single MFMA group with three of its four cells
    masked and a one-slice gather, against the oracle's FP64 values; a
    skipped (degenerate) cell adds nothing; unknown methods and a Poisson
    ratio outside (-1, 0.5) are rejected, as elasticity.rs:14-48 rejects them."""
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    C = np.array([[0, 1, 2, 3]], dtype=np.uint64)
    mat = fh.ElasticMaterial(2e11, 0.3)
    gm, om = fh.Mesh.from_arrays(V, C), O.Mesh(V, C)
    _, _, v64 = O.assemble_elasticity(om, 2e11, 0.3).export()
    _, _, v32 = fh.assemble_elasticity_stiffness_3d(gm, mat, method=1).export_csc()
    assert np.abs(v32 - v64).max() <= 2e-6 * np.abs(v64).max()
    # the same cell plus a flat copy of it (det J = 0): the values do not change
    C2 = np.array([[0, 1, 2, 3], [0, 1, 2, 0]], dtype=np.uint64)
    g2 = fh.Mesh.from_arrays(V, C2)
    K2 = fh.assemble_elasticity_stiffness_3d(g2, mat, method=1)
    assert K2.skipped_cells() == 1
    _, _, w32 = K2.export_csc()
    assert np.array_equal(w32, v32)
    for bad in (-1, 3):
        with pytest.raises(fh.FenicsError, match="method"):
            fh.assemble_elasticity_stiffness_3d(gm, mat, method=bad)
    with pytest.raises(fh.FenicsError, match="Poisson"):
        fh._new_matrix(fh.lib().fh_assemble_elasticity_method, gm.h, 2e11, 0.5, None, 1)


def test_elasticity_mfma_copy_drives_the_mixed_solve():
    """Method 2 (FP64 values, MFMA FP32 copy): the mixed solve of a clamped
    cantilever reaches the FP64 solve's true residual and solution."""
    gm = fh.box_beam(2.0, 0.4, 0.4, 20, 4, 4)
    mat = fh.ElasticMaterial(210e9, 0.3)
    b = fh.assemble_elasticity_body_force_3d(gm, [0, 0, -7.85e4])
    bc = fh.clamp_dofs_x0(gm.vertices)
    out = {}
    for method in (0, 2):
        K = fh.assemble_elasticity_stiffness_3d(gm, mat, method=method)
        Kb, bb = fh.apply_vector_dirichlet_bc(K, b, bc)
        prec = fh.FP64_TRUE if method == 0 else fh.MIXED
        out[method] = Kb.solve(bb, tol=1e-10, max_iter=20000, precision=prec)
    (x0, s0), (x2, s2) = out[0], out[2]
    assert s2.true_rel_residual <= max(1e-10, 2.5 * s0.true_rel_residual)
    assert np.linalg.norm(x2 - x0) / np.linalg.norm(x0) < 1e-8


def test_dirichlet_scalar_bitwise():
    gm, om = fh.unit_square(10, 8), O.unit_square(10, 8)
    b = np.random.default_rng(2).standard_normal(om.nv)
    bd = O.boundary_dofs(om)
    g = np.linspace(-1.0, 2.0, len(bd))
    A1, b1 = fh.apply_dirichlet_bc(fh.assemble_stiffness(gm), b, fh.DirichletBC(bd, g))
    A2, b2 = O.apply_dirichlet_bc(O.assemble_stiffness(om), b, bd, g)
    same_csc(A1, A2)
    assert np.array_equal(bits(b1), bits(b2))


def test_dirichlet_vector_partial_and_duplicate_bitwise():
    dims = (1.0, 0.2, 0.2, 4, 2, 2)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    K1 = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(1e6, 0.25))
    K2 = O.assemble_elasticity(om, 1e6, 0.25)
    b = fh.assemble_elasticity_body_force_3d(gm, [0, 0, -1.0])
    # x-clamp plus a few single-component constraints with nonzero values and
    # one dof listed twice (reference: last value wins, diagonal 1.0 + 1.0)
    pres = [(d, 0.0) for d in O.clamp_dofs_x0(om)] + [(31, 1e-3), (47, -2e-3), (31, 5e-4)]
    A1, b1 = fh.apply_vector_dirichlet_bc(K1, b, fh.VectorDirichletBC(pres))
    A2, b2 = O.apply_dirichlet_bc(K2, b, [d for d, _ in pres], [v for _, v in pres])
    same_csc(A1, A2)
    assert np.array_equal(bits(b1), bits(b2))


def test_p2_dofmap_stiffness_rhs_bitwise():
    gm, om = fh.unit_cube(3, 3, 2), O.unit_cube(3, 3, 2)
    p2 = O.P2(om)
    assert gm.p2_info() == (p2.ndofs, p2.nedges)
    assert np.array_equal(gm.p2_cell_dofs(), p2.cell_dofs())
    assert np.array_equal(gm.p2_boundary_dofs(), p2.boundary_dofs())
    same_csc(fh.assemble_p2_stiffness(gm, 0), p2.stiffness())
    fc = np.random.default_rng(9).random(om.nc)
    assert np.array_equal(bits(fh.assemble_p2_rhs_cellconst(gm, fc)), bits(p2.rhs_cellconst(fc)))


def test_p2_stiffness_bitwise_many_slices():
    """P2 stiffness over 77 SELL slices (unit_cube(8, 8, 8), 4913 dofs, vertex
    and edge rows of different widths) is the oracle's, bit for bit."""
    gm, om = fh.unit_cube(8, 8, 8), O.unit_cube(8, 8, 8)
    same_csc(fh.assemble_p2_stiffness(gm, 0), O.P2(om).stiffness())


# ---------------------------------------------------------------- full-size configurations
COUNTS = json.load(open(os.path.join(GOLDEN, "counts.json")))


def test_c0_full_size_bitwise():
    gm, om = fh.unit_square(128, 128), O.unit_square(128, 128)
    A = fh.assemble_stiffness(gm)
    same_csc(A, O.assemble_stiffness(om))
    bd = gm.boundary_dofs()
    Ab, _ = fh.apply_dirichlet_bc(A, np.zeros(gm.nv), fh.DirichletBC(bd, np.zeros(len(bd))))
    assert Ab.nnz() == COUNTS["C0"]["nnz_bc"]


def test_c3_full_size_bitwise():
    dims = (10.0, 2.0, 2.0, 100, 20, 20)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    K = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(210e9, 0.3))
    same_csc(K, O.assemble_elasticity(om, 210e9, 0.3))


def test_c1_full_size_counts_and_invariants():
    gm = fh.box_beam(10.0, 2.0, 2.0, 200, 40, 40)
    c = COUNTS["C1"]
    assert 3 * gm.nv == c["n"] and gm.nc == c["cells"]
    K = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(210e9, 0.3))
    assert K.nnz() == c["nnz_raw"]
    # rigid translations are in the kernel before the boundary conditions
    for a in range(3):
        t = np.zeros((gm.nv, 3))
        t[:, a] = 1.0
        r = K.spmv(t.reshape(-1))
        assert np.abs(r).max() < 1e-6 * 210e9 * 0.05
    bc = fh.clamp_dofs_x0(gm.vertices)
    assert len(bc.prescribed_dofs) == c["n_fixed"]
    b = fh.assemble_elasticity_body_force_3d(gm, [0, 0, -7.85e4])
    assert abs(b.sum() - (-7.85e4 * 40.0)) < 1e-9 * 7.85e4 * 40.0
    Kb, bb = fh.apply_vector_dirichlet_bc(K, b, bc)
    assert Kb.nnz() == c["nnz_bc"]


def test_c1_elasticity_mfma_matches_fp64():
    """The FP32 MFMA elasticity assembly at full C1 size: the pattern of the
    FP64 assembly, every value within 2e-6 of the largest entry (one cell's
    contribution is about max|K| / 24, so a dropped or doubled cell fails it),
    and every value an FP32 number."""
    gm = fh.box_beam(10.0, 2.0, 2.0, 200, 40, 40)
    mat = fh.ElasticMaterial(210e9, 0.3)
    c1, r1, v1 = fh.assemble_elasticity_stiffness_3d(gm, mat).export_csc()
    c2, r2, v2 = fh.assemble_elasticity_stiffness_3d(gm, mat, method=1).export_csc()
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    err = np.abs(v2 - v1).max() / np.abs(v1).max()
    print(f"C1 MFMA elasticity: max |K32 - K64| / max |K64| = {err:.3e}")
    assert err <= 2e-6
    assert np.array_equal(v2, v2.astype(np.float32).astype(np.float64))


def test_c2_full_size_p2_counts():
    gm = fh.unit_cube(64, 64, 64)
    c = COUNTS["C2"]
    nd, ne = gm.p2_info()
    assert (nd, ne) == (c["n"], c["n_edges"])
    assert len(gm.p2_boundary_dofs()) == c["n_fixed"]
    K = fh.assemble_p2_stiffness(gm, 0)
    assert K.nnz() == c["nnz_raw"]
    r = K.spmv(np.ones(nd))
    assert np.abs(r).max() < 1e-12


def unstructured(om, seed, jitter):
    """The oracle mesh with interior vertices moved by up to jitter * h,
    vertices renumbered and cells permuted and rotated (orientation kept)."""
    rng = np.random.default_rng(seed)
    V = om.vertices.copy()
    lo, hi = V.min(0), V.max(0)
    interior = np.all((V > lo + 1e-9) & (V < hi - 1e-9), axis=1)
    h = (hi - lo).min() / 8.0
    V[interior] += jitter * h * rng.uniform(-1, 1, (int(interior.sum()), V.shape[1]))
    perm = rng.permutation(om.nv)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(om.nv)
    C = inv[om.cells.astype(np.int64)][rng.permutation(om.nc)]
    k = C.shape[1]
    C = np.array([np.roll(c, 2 * int(rng.integers(0, k // 2))) for c in C], dtype=np.uint64)
    return O.Mesh(V[perm], C)


@pytest.mark.parametrize("dim", [2, 3])
def test_unstructured_mesh_bitwise(dim):
    """Assembly, loads, Dirichlet condensation and boundary dofs on a jittered,
    renumbered mesh given through the C ABI: bitwise the oracle's."""
    base = O.unit_square(9, 7) if dim == 2 else O.box_beam(1.0, 0.5, 0.25, 6, 4, 3)
    om = unstructured(base, 7 + dim, 0.25)
    gm = fh.Mesh.from_arrays(om.vertices, om.cells)
    gpu_mesh_equals_oracle(gm, om)
    same_csc(fh.assemble_stiffness(gm), O.assemble_stiffness(om))
    if dim == 2:
        same_csc(fh.assemble_mass(gm), O.assemble_mass_2d(om))
    xo = O.quad_points(om, 2)
    assert np.array_equal(bits(gm.quad_points(2)), bits(xo))
    fq = np.cos(2 * xo[..., 0]) + xo[..., 1]
    assert np.array_equal(bits(fh.assemble_rhs(gm, fq)), bits(O.assemble_rhs(om, fq)))
    bd = O.boundary_dofs(om)
    assert np.array_equal(gm.boundary_dofs(), bd)
    if dim == 3:
        K, Ko = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(7e10, 0.33)), \
            O.assemble_elasticity(om, 7e10, 0.33)
        same_csc(K, Ko)
        b, bo = fh.assemble_elasticity_body_force_3d(gm, [1.0, -2.0, 3.0]), \
            O.assemble_body_force(om, [1.0, -2.0, 3.0])
        assert np.array_equal(bits(b), bits(bo))
        fixed = [(3 * int(i) + a, 0.01 * a) for i in bd[:9] for a in range(3)]
        Kc, bc = fh.apply_vector_dirichlet_bc(K, b, fh.VectorDirichletBC(fixed))
        Kco, bco = O.apply_dirichlet_bc(Ko, bo, [d for d, _ in fixed], [v for _, v in fixed])
        same_csc(Kc, Kco)
        assert np.array_equal(bits(bc), bits(bco))
        # the P2 numbering (cells around each edge, then max - min, then min:
        # on a renumbered mesh max - min is no direction, only the tie rule)
        # and the P2 stiffness, bitwise; the MFMA elasticity within its bound
        p2 = O.P2(om)
        assert gm.p2_info() == (p2.ndofs, p2.nedges)
        assert np.array_equal(gm.p2_cell_dofs(), p2.cell_dofs())
        same_csc(fh.assemble_p2_stiffness(gm, 0), p2.stiffness())
        _, _, v64 = K.export_csc()
        _, _, v32 = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(7e10, 0.33),
                                                        method=1).export_csc()
        assert np.abs(v32 - v64).max() <= 2e-6 * np.abs(v64).max()


def test_high_valence_vertex_bitwise():
    """A vertex shared by every cell: the tetrahedra (centre, hull triangle)
    of 180 points on a sphere around it.  The centre's row holds 181 blocks
    (P1) and 895 (P2): more than the 255 a u8 slot map could index, and more
    than the 48 (elasticity) and 512 (P2 gather) blocks the row kernels sum in
    LDS at once, so both kernels cover it in passes.  The assembled matrices
    stay bitwise the oracle's."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(11)
    p = rng.normal(size=(180, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    tri = ConvexHull(p).simplices
    V = np.vstack([np.zeros((1, 3)), p]).astype(np.float64)
    C = np.hstack([np.zeros((len(tri), 1), dtype=np.uint64), tri.astype(np.uint64) + 1])
    om = O.Mesh(V, C)
    gm = fh.Mesh.from_arrays(V, C)
    same_csc(fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(7e10, 0.3)),
             O.assemble_elasticity(om, 7e10, 0.3))
    ec = rng.uniform(1e9, 5e9, len(C))
    same_csc(fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(7e10, 0.3), ec),
             O.assemble_elasticity(om, ec, 0.3))
    same_csc(fh.assemble_stiffness(gm), O.assemble_stiffness(om))
    p2 = O.P2(om)
    assert gm.p2_info() == (p2.ndofs, p2.nedges)
    assert np.array_equal(gm.p2_cell_dofs(), p2.cell_dofs())
    same_csc(fh.assemble_p2_stiffness(gm, 0), p2.stiffness())
    # the FP32 MFMA element path through the same wide-row gather.  The
    # centre's entries sum up to 356 FP32 element contributions, so the bound
    # is the recursive-summation one, (k + 16) u max|K| with k the most cells
    # around one dof and u = 2^-24 (the 16 covers the element matrix itself);
    # dropping or doubling one cell's contribution (about max|K| / k) fails it
    _, _, v64 = fh.assemble_p2_stiffness(gm, 0).export_csc()
    _, _, v32 = fh.assemble_p2_stiffness(gm, 1).export_csc()
    k = np.bincount(gm.p2_cell_dofs().ravel().astype(np.int64)).max()
    assert k == len(C)
    assert np.abs(v32 - v64).max() <= (k + 16) * 2.0 ** -24 * np.abs(v64).max()
    # the FP32 MFMA elasticity path through the multi-pass row gather (the
    # centre's row of 181 blocks is summed in four passes of 48), under the
    # same recursive-summation bound with the centre's 356 cells
    mat = fh.ElasticMaterial(7e10, 0.3)
    _, _, e64 = fh.assemble_elasticity_stiffness_3d(gm, mat).export_csc()
    _, _, e32 = fh.assemble_elasticity_stiffness_3d(gm, mat, method=1).export_csc()
    assert np.abs(e32 - e64).max() <= (k + 16) * 2.0 ** -24 * np.abs(e64).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_degenerate_cells_are_skipped_and_counted(dim):
    """jacobian.rs:29-31 / poisson.rs:42-45: a cell with |det J| < 1e-14 adds
    nothing.  Two cells with a repeated vertex (det J = 0 exactly) appended to
    a mesh leave every matrix and load bitwise the oracle's, and the assembly
    reports them as skipped.  The repeated vertices couple only vertices that
    other cells couple already, so the pattern is the reference's."""
    base = O.unit_square(6, 5) if dim == 2 else O.box_beam(1.0, 0.5, 0.25, 5, 3, 2)
    C = base.cells.astype(np.uint64)
    k = C.shape[1]
    extra = []
    for c in (C[3], C[len(C) // 2]):
        d = c.copy()
        d[k - 1] = d[0]
        extra.append(d)
    om = O.Mesh(base.vertices.copy(), np.concatenate([C, np.array(extra, dtype=np.uint64)]))
    gm = fh.Mesh.from_arrays(om.vertices, om.cells)
    K = fh.assemble_stiffness(gm)
    assert K.skipped_cells() == 2
    same_csc(K, O.assemble_stiffness(om))
    xo = O.quad_points(om, 2)
    fq = np.cos(2 * xo[..., 0]) + xo[..., 1]
    assert np.array_equal(bits(fh.assemble_rhs(gm, fq)), bits(O.assemble_rhs(om, fq)))
    if dim == 2:
        M = fh.assemble_mass(gm)
        assert M.skipped_cells() == 2
        same_csc(M, O.assemble_mass_2d(om))
    else:
        Ke = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(7e10, 0.3))
        assert Ke.skipped_cells() == 2
        same_csc(Ke, O.assemble_elasticity(om, 7e10, 0.3))
        b = fh.assemble_elasticity_body_force_3d(gm, [1.0, -2.0, 3.0])
        assert np.array_equal(bits(b), bits(O.assemble_body_force(om, [1.0, -2.0, 3.0])))
        bc = fh.clamp_dofs_x0(gm.vertices)
        Kc, _ = fh.apply_vector_dirichlet_bc(Ke, b, bc)
        assert Kc.skipped_cells() == 2
    # a regular mesh skips nothing
    assert fh.assemble_stiffness(fh.Mesh.from_arrays(base.vertices, base.cells)).skipped_cells() == 0
