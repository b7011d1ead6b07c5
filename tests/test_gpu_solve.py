"""GPU parity of the solvers against the CPU oracle and the reference's own
tests (cg.rs:156-169, poisson_tests.rs, elasticity_tests.rs,
ab_benchmarks.rs, gpu.rs:1609-1769).

Tolerances: FP64 CG differs from the oracle's only in the summation order of
its dot products.  That changes the iteration count by a few percent on
ill-conditioned elasticity systems (the recursive residual drifts from the
true one differently), while solutions agree to ~1e-8 relative at tol 1e-12.
The reference semantics stop on the recursive residual, so their true FP64
residual may exceed tol (checked within 10x).  FP64_TRUE and the mixed modes
stop on the true FP64 residual, so that one is checked against tol itself, and
mixed solutions agree with FP64_TRUE's to 1e-9 relative (2.4e-13 measured on a
CPU model of the same iteration)."""
import json
import os

import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PI = np.pi


def src2d(x):
    return 2 * PI * PI * np.sin(PI * x[0]) * np.sin(PI * x[1])


def exact2d(x):
    return np.sin(PI * x[0]) * np.sin(PI * x[1])


def test_small_spd_system():  # cg.rs:156-169
    res = fh.solve_cg_gpu_or_cpu([0, 2, 4], [0, 1, 0, 1], [4.0, -1.0, -1.0, 4.0], [3.0, 3.0],
                                 None, 1e-10, 100)
    assert res.iterations <= 2 and res.stats.final_residual < 1e-10
    assert np.allclose(res.solution, 1.0, atol=1e-8)


def c0_system():
    om = O.unit_square(128, 128)
    A = O.assemble_stiffness(om)
    b = O.assemble_rhs(om, src2d(np.moveaxis(O.quad_points(om, 2), -1, 0)))
    bd = O.boundary_dofs(om)
    return O.apply_dirichlet_bc(A, b, bd, np.zeros(len(bd)))


def test_c0_csc_drop_in_vs_oracle():
    Ab, bb = c0_system()
    st, xo, it_o, res_o = O.solve_cg(Ab, bb, None, 1e-12, 5000)
    cp, ri, va = Ab.export()
    r = fh.solve_cg_gpu_or_cpu(cp, ri, va, bb, None, 1e-12, 5000)
    assert abs(r.iterations - it_o) <= 2
    assert np.linalg.norm(r.solution - xo) / np.linalg.norm(xo) < 1e-8
    assert r.stats.true_rel_residual < 1e-11


def test_ordered_spmv_is_bitwise_reference_spmv():
    Ab, _ = c0_system()
    cp, ri, va = Ab.export()
    x = np.random.default_rng(0).standard_normal(Ab.n)
    x[::5] = 0.0
    gm = fh.unit_square(128, 128)
    A = fh.assemble_stiffness(gm)
    bd = gm.boundary_dofs()
    Ag, _ = fh.apply_dirichlet_bc(A, np.zeros(gm.nv), fh.DirichletBC(bd, np.zeros(len(bd))))
    y = Ag.spmv(x, fh.ORDER_ORDERED)
    assert np.array_equal(y.view(np.uint64), Ab.spmv(x).view(np.uint64))
    om = O.box_beam(1.0, 0.2, 0.2, 5, 2, 2)
    K = fh.assemble_elasticity_stiffness_3d(fh.box_beam(1.0, 0.2, 0.2, 5, 2, 2),
                                            fh.ElasticMaterial(3e3, 0.3))
    xe = np.random.default_rng(1).standard_normal(3 * om.nv)
    ye = O.assemble_elasticity(om, 3e3, 0.3).spmv(xe)
    assert np.array_equal(K.spmv(xe, fh.ORDER_ORDERED).view(np.uint64), ye.view(np.uint64))


def test_poisson_2d_manufactured_rate():  # poisson_tests.rs:6-52
    errs = []
    for n in (8, 16, 32):
        m = fh.unit_square(n, n)
        sol = fh.solve_poisson_2d(m, src2d, exact2d, 1e-10, 500)
        om = O.unit_square(n, n)
        e = fh.compute_l2_error_2d(m, sol.u, exact2d)
        eo = O.l2_error(om, sol.u, exact2d(np.moveaxis(O.quad_points(om, 3), -1, 0)))
        assert abs(e - eo) < 1e-13 * eo  # device partial sums vs the sequential loop
        errs.append(e)
    assert 3.8 < errs[0] / errs[1] < 4.2 and 3.8 < errs[1] / errs[2] < 4.2


def test_poisson_3d_manufactured():  # poisson_tests.rs:55-73
    ex = lambda x: np.sin(PI * x[0]) * np.sin(PI * x[1]) * np.sin(PI * x[2])
    src = lambda x: 3 * PI * PI * ex(x)
    m = fh.unit_cube(8, 8, 8)
    sol = fh.solve_poisson_3d(m, src, ex, 1e-10, 1000)
    om = O.unit_cube(8, 8, 8)
    e = fh.compute_l2_error_3d(m, sol.u, ex)
    assert abs(e - O.l2_error(om, sol.u, ex(np.moveaxis(O.quad_points(om, 3), -1, 0)))) < 1e-13 * e
    assert e < 0.05


def test_ab_poisson_2d_vs_dolfinx():  # ab_benchmarks.rs:13-48
    ref = json.load(open(os.path.join(GOLDEN, "dolfinx_reference.json")))["poisson_2d"]
    m = fh.unit_square(16, 16)
    sol = fh.solve_poisson_2d(m, src2d, exact2d, 1e-12, 1000)
    l2 = fh.compute_l2_error_2d(m, sol.u, exact2d)
    assert abs(l2 - ref["l2_error"]) / ref["l2_error"] < 0.05
    assert abs(sol.u.max() - ref["max_value"]) / ref["max_value"] < 1e-3


def tip_mean(m, u, comp, L=1.0):
    tip = np.abs(m.vertices[:, 0] - L) < 1e-6
    return u.reshape(-1, 3)[tip, comp].mean()


def test_ab_cantilever_vs_dolfinx():  # ab_benchmarks.rs:51-100
    ref = json.load(open(os.path.join(GOLDEN, "dolfinx_reference.json")))["elasticity_3d"]
    m = fh.box_beam(1.0, 0.1, 0.1, 10, 2, 2)
    sol = fh.solve_elasticity_3d(m, fh.ElasticMaterial(1e7, 0.3), [0, -1000.0, 0],
                                 fh.clamp_dofs_x0(m.vertices), 1e-8, 1000)
    assert abs(tip_mean(m, sol.displacements, 1) - ref["tip_uy"]) / abs(ref["tip_uy"]) < 0.01


def test_axial_bar_extension():  # elasticity_tests.rs:6-49
    m = fh.box_beam(1.0, 0.1, 0.1, 10, 2, 2)
    sol = fh.solve_elasticity_3d(m, fh.ElasticMaterial(1.0e6, 0.0), [1000.0, 0, 0],
                                 fh.clamp_dofs_x0(m.vertices), 1e-10, 1000)
    assert abs(tip_mean(m, sol.displacements, 0) - 5e-4) / 5e-4 < 0.03


def test_cantilever_beam_bending_vs_oracle():  # elasticity_tests.rs:52-126
    m = fh.box_beam(1.0, 0.1, 0.1, 20, 3, 3)
    sol = fh.solve_elasticity_3d(m, fh.ElasticMaterial(1.0e6, 0.3), [0, 0, -1000.0],
                                 fh.clamp_dofs_x0(m.vertices), 1e-9, 2000)
    uz = tip_mean(m, sol.displacements, 2)
    assert abs(uz - (-0.0890)) / 0.0890 < 0.06 and uz < 0
    om = O.box_beam(1.0, 0.1, 0.1, 20, 3, 3)
    K = O.assemble_elasticity(om, 1.0e6, 0.3)
    b = O.assemble_body_force(om, [0, 0, -1000.0])
    fx = O.clamp_dofs_x0(om)
    Kb, bb = O.apply_dirichlet_bc(K, b, fx, np.zeros(len(fx)))
    st, uo, it, res = O.solve_cg(Kb, bb, None, 1e-9, 2000)
    assert abs(sol.iterations - it) <= max(3, 0.05 * it)
    assert np.linalg.norm(sol.displacements - uo) / np.linalg.norm(uo) < 1e-6


@pytest.mark.parametrize("prec", [fh.MIXED, fh.MIXED_IR])
def test_mixed_precision_reaches_fp64_accuracy(prec):
    m = fh.box_beam(10.0, 2.0, 2.0, 50, 10, 10)
    mat = fh.ElasticMaterial(210e9, 0.3)
    bc = fh.clamp_dofs_x0(m.vertices)
    f = [0, 0, -7.85e4]
    kw = dict(compute_stresses=False)
    sref = fh.solve_elasticity_3d(m, mat, f, bc, 1e-10, 20000, **kw)
    stru = fh.solve_elasticity_3d(m, mat, f, bc, 1e-10, 20000, precision=fh.FP64_TRUE, **kw)
    smx = fh.solve_elasticity_3d(m, mat, f, bc, 1e-10, 40000, precision=prec, **kw)
    # recursive-residual stop: the plain FP64 product's rounding floor shows
    assert sref.stats.rel_residual < 1e-10 and sref.stats.true_rel_residual < 1e-9
    assert stru.stats.true_rel_residual < 1e-10 and smx.stats.true_rel_residual < 1e-10
    assert stru.stats.replacements >= 1 and smx.stats.replacements >= 1
    ut, umx = stru.displacements, smx.displacements
    assert np.linalg.norm(umx - ut) / np.linalg.norm(ut) < 1e-9
    assert np.linalg.norm(sref.displacements - ut) / np.linalg.norm(ut) < 1e-7
    if prec == fh.MIXED:  # relaxation keeps the FP64 Krylov convergence
        assert smx.iterations < 1.1 * stru.iterations
        assert 0 < smx.stats.iterations_low < smx.iterations
    else:  # FP32 values from the first iteration (FP64 again only after a stall)
        assert smx.stats.iterations_low > 0


def test_fp64_true_is_reference_cg_until_tol():
    """FP64_TRUE is cg.rs's iteration with the product evaluated in difference
    form (same operator, different rounding) and a true-residual stop; on a
    well-conditioned system it stops within an iteration or two of cg.rs."""
    Ab, bb = c0_system()
    cp, ri, va = Ab.export()
    st, xo, it_o, res_o = O.solve_cg(Ab, bb, None, 1e-10, 5000)
    m = fh.unit_square(128, 128)
    A = fh.assemble_stiffness(m)
    bd = m.boundary_dofs()
    Ag, bg = fh.apply_dirichlet_bc(A, bb, fh.DirichletBC(bd, np.zeros(len(bd))))
    x, s = Ag.solve(bb, tol=1e-10, max_iter=5000, precision=fh.FP64_TRUE)
    assert abs(int(s.iterations) - it_o) <= 2 and s.true_rel_residual < 1e-10
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_error_paths():
    Ab, bb = c0_system()
    cp, ri, va = Ab.export()
    with pytest.raises(fh.FenicsError) as e:
        fh.solve_cg_gpu_or_cpu(cp, ri, va, bb, None, 1e-12, 3)
    assert e.value.code == fh.FH_ERR_NOT_CONVERGED
    r = fh.solve_cg_gpu_or_cpu(cp, ri, va, np.zeros(len(bb)), None, 1e-12, 10)
    assert r.iterations == 0 and not r.solution.any()
    with pytest.raises(fh.FenicsError) as e:
        fh.solve_cg_gpu_or_cpu([0, 1, 2], [0, 1], [0.0, 0.0], [1.0, 1.0], None, 1e-10, 10)
    assert e.value.code == fh.FH_ERR_BREAKDOWN
    with pytest.raises(fh.FenicsError) as e:
        fh.solve_cg_gpu_or_cpu([0, 1, 2], [0, 5], [1.0, 1.0], [1.0, 1.0], None, 1e-10, 10)
    assert e.value.code == fh.FH_ERR_BAD_INPUT
    with pytest.raises(ValueError):
        fh.ElasticMaterial(1.0, 0.5)


def test_edge_cases():
    # x0 already the solution: the initial test of cg.rs:80-84 returns at once
    Ab, bb = c0_system()
    cp, ri, va = Ab.export()
    _, xo, _, _ = O.solve_cg(Ab, bb, None, 1e-12, 5000)
    assert fh.solve_cg_gpu_or_cpu(cp, ri, va, bb, xo, 1e-6, 10).iterations == 0
    # a 1x1 system
    r = fh.solve_cg_gpu_or_cpu([0, 1], [0], [4.0], [2.0], None, 1e-12, 5)
    assert abs(r.solution[0] - 0.5) < 1e-15 and r.iterations <= 1
    # the smallest mesh: every vertex is on the boundary, u is the boundary data
    m = fh.unit_square(1, 1)
    sol = fh.solve_poisson_2d(m, src2d, lambda v: 1.0 + v[0], 1e-10, 10)
    assert np.allclose(sol.u, 1.0 + m.vertices[:, 0], rtol=0, atol=1e-14)
    # non-homogeneous boundary data on a larger mesh, against the oracle pipeline
    gm, om = fh.unit_square(24, 17), O.unit_square(24, 17)
    g = lambda v: np.sin(v[0]) + v[1] * v[1]  # noqa: E731
    sol = fh.solve_poisson_2d(gm, src2d, g, 1e-12, 2000)
    A = O.assemble_stiffness(om)
    b = O.assemble_rhs(om, src2d(np.moveaxis(O.quad_points(om, 2), -1, 0)))
    bd = O.boundary_dofs(om)
    Ab2, bb2 = O.apply_dirichlet_bc(A, b, bd, np.array([g(om.vertices[int(i)]) for i in bd]))
    _, xo2, _, _ = O.solve_cg(Ab2, bb2, None, 1e-12, 2000)
    assert np.linalg.norm(sol.u - xo2) < 1e-9 * np.linalg.norm(xo2)


def laplace7(nx):
    """gpu.rs:1614-1643: 3D 7-point finite-difference Laplacian."""
    n = nx ** 3
    rows, cols, vals = [], [], []
    for z in range(nx):
        for y in range(nx):
            for x in range(nx):
                i = (z * nx + y) * nx + x
                rows.append(i), cols.append(i), vals.append(6.0)
                for d, ok in ((-1, x > 0), (1, x + 1 < nx), (-nx, y > 0), (nx, y + 1 < nx),
                              (-nx * nx, z > 0), (nx * nx, z + 1 < nx)):
                    if ok:
                        rows.append(i), cols.append(i + d), vals.append(-1.0)
    return O.csc_from_triplets(n, rows, cols, vals)


def test_csr_cg_vs_cpu():  # gpu.rs:1609-1663
    A = laplace7(20)
    b = np.ones(A.n)
    st, xo, it, res = O.solve_cg(A, b, None, 1e-9, 1000)
    rp, ci, va, inv = A.to_csr()
    r = fh.solve_csr_cg_gpu(rp, ci, va, inv, b, None, 1e-9, 1000)
    for i in range(0, A.n, 137):
        assert abs(r.solution[i] - xo[i]) <= 1e-5 * max(1.0, abs(xo[i]))
    cp, ri, vv = A.export()
    r2 = fh.solve_cg_gpu_or_cpu(cp, ri, vv, b, None, 1e-9, 1000)
    assert np.allclose(r2.solution, xo, rtol=1e-6, atol=1e-8)
    # x0 is honoured when forming r0 (gpu.rs:1194 ignores it)
    r3 = fh.solve_csr_cg_gpu(rp, ci, va, inv, b, xo, 1e-9, 1000)
    assert r3.iterations <= 1


def test_solve_without_stored_residual_matches_kept_residual():
    """Scalar Jacobi with the library's own inverse diagonal keeps no residual
    vector: the update recovers r_i as z_i / D^-1_i from the (z, p) pairs.  A
    caller-supplied inverse diagonal (the CSR entry point) keeps r.  On the
    same system, with the same stopping rule, the two take the same number of
    iterations up to 2, reach the same solution to 1e-8 and true residuals
    within 2x of tol; a residual recovered wrongly (off by a factor, or
    stale) fails both."""
    A = laplace7(24)
    b = np.random.default_rng(21).uniform(-1.0, 1.0, A.n)
    rp, ci, va, inv = A.to_csr()
    kept = fh.solve_csr_cg_gpu(rp, ci, va, inv, b, None, 1e-10, 5000)
    cp, ri, vv = A.export()
    zr = fh.solve_cg_gpu_or_cpu(cp, ri, vv, b, None, 1e-10, 5000)
    assert abs(kept.iterations - zr.iterations) <= 2
    assert np.linalg.norm(kept.solution - zr.solution) <= 1e-8 * np.linalg.norm(kept.solution)
    S = A.to_scipy()
    for r in (kept, zr):
        assert np.linalg.norm(b - S @ r.solution) <= 2e-10 * np.linalg.norm(b)


def test_deferred_x_after_each_iteration_count():
    """The update kernel writes x every second iteration and a flush completes
    it before it is read.  Stopped after k = 1..6 iterations (both parities),
    the GPU iterate equals the oracle's cg.rs iterate after k iterations to
    rounding; a missing flush (x short of alpha_k p_k) or a doubled one fails
    by far more."""
    Ab, bb = c0_system()
    K = fh.matrix_from_csc(*Ab.export())
    for k in range(1, 7):
        st, xo, it, _ = O.solve_cg(Ab, bb, None, 1e-30, k)
        assert st == 2 and it == k
        x, s = K.solve(bb, tol=1e-30, max_iter=k, precision=fh.FP64, allow_fail=True)
        assert s.iterations == k
        assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo), k


def bsr3_synthetic(nx=15):
    """gpu.rs:1671-1741: synthetic SPD 3x3 block-CSR on a 15^3 grid."""
    n_nodes = nx ** 3
    diag = np.array([18.0, 1.0, 0.5, 1.0, 18.0, 0.5, 0.5, 0.5, 18.0])
    off = np.array([-2.0, 0, 0, 0, -2.0, 0, 0, 0, -2.0])
    inv = np.linalg.inv(diag.reshape(3, 3)).reshape(-1)
    row_ptr, cols, vals = [0], [], []
    for z in range(nx):
        for y in range(nx):
            for x in range(nx):
                i = (z * nx + y) * nx + x
                nb = [i] + [i + d for d, ok in ((-1, x > 0), (1, x + 1 < nx), (-nx, y > 0),
                                                (nx, y + 1 < nx), (-nx * nx, z > 0),
                                                (nx * nx, z + 1 < nx)) if ok]
                for j in sorted(nb):
                    cols.append(j)
                    vals.append(diag if j == i else off)
                row_ptr.append(len(cols))
    return n_nodes, np.array(row_ptr), np.array(cols), np.concatenate(vals), np.tile(inv, n_nodes)


def test_bsr3_context_multi_and_pcg():  # gpu.rs:1665-1768
    n_nodes, rp, cc, va, dinv = bsr3_synthetic()
    x = (np.arange(3 * n_nodes) % 17) * 0.1
    ctx = fh.GpuBsr3Context(n_nodes, rp, cc, va, dinv)
    y1 = ctx.spmv(x)
    y2, used = fh.spmv_bsr3_multi_gpu(n_nodes, rp, cc, va, x, 1)  # one GPU only
    assert len(used) >= 1 and np.allclose(y1, y2, atol=1e-12, rtol=0)
    # reference product on the host
    yr = np.zeros(3 * n_nodes)
    for r in range(n_nodes):
        for p in range(rp[r], rp[r + 1]):
            yr[3 * r:3 * r + 3] += va[9 * p:9 * p + 9].reshape(3, 3) @ x[3 * cc[p]:3 * cc[p] + 3]
    assert np.allclose(y1, yr, rtol=1e-14, atol=1e-12)
    assert ctx.nnz_blocks() == len(cc) and ctx.device_name()
    res = fh.solve_bsr3_pcg_gpu(n_nodes, rp, cc, va, dinv, np.ones(3 * n_nodes), None, 1e-9, 500)
    assert res.stats.rel_residual < 1e-9 and res.stats.true_rel_residual < 1e-8
    # block-Jacobi smoothing reduces the residual of A z = r
    r = np.ones(3 * n_nodes)
    z = ctx.jacobi_smooth(r, np.zeros(3 * n_nodes), 6, 0.8, True)
    assert np.linalg.norm(r - ctx.spmv(z)) < 0.5 * np.linalg.norm(r)
    z1 = ctx.jacobi_smooth(r, np.zeros(3 * n_nodes), 1, 0.8, True)
    assert np.allclose(z1, (dinv.reshape(-1, 3, 3) @ r.reshape(-1, 3, 1)).reshape(-1))


def test_solver_determinism():
    m = fh.unit_square(64, 64)
    a = fh.solve_poisson_2d(m, src2d, exact2d, 1e-12, 2000)
    b = fh.solve_poisson_2d(m, src2d, exact2d, 1e-12, 2000)
    assert a.iterations == b.iterations and np.array_equal(a.u, b.u)


def test_export_bsr3_feeds_the_block_solver():  # gpu.rs:453-458, 775-786 arrays from any mesh
    om = O.box_beam(1.0, 0.3, 0.2, 9, 3, 3)
    rng = np.random.default_rng(5)
    perm = rng.permutation(om.nv)  # renumbered vertices: not a generated layout
    inv = np.empty_like(perm)
    inv[perm] = np.arange(om.nv)
    verts = om.vertices[perm]
    cells = inv[om.cells][rng.permutation(len(om.cells))]
    m = fh.Mesh.from_arrays(verts, cells)
    # moduli near 1: the block solver's breakdown test is an absolute
    # p.Ap <= 1e-30 (gpu.rs:1023), which a 1e-12 solve of a stiff, weakly loaded
    # beam would reach
    E_cell = 1.0 + rng.random(len(cells))
    K = fh.assemble_elasticity_stiffness_3d(m, fh.ElasticMaterial(1.0, 0.3), E_cell)
    b = fh.assemble_elasticity_body_force_3d(m, [0.0, 0.0, -100.0])
    Kc, bc = fh.apply_vector_dirichlet_bc(K, b, fh.clamp_dofs_x0(verts))
    n_nodes, rp, cc, va, di = Kc.export_bsr3()
    assert n_nodes == om.nv and rp[0] == 0 and rp[-1] == len(cc)
    for r in range(n_nodes):
        row = cc[rp[r]:rp[r + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and r in row
    # the blocks hold exactly the exported CSC operator (itself bitwise the oracle's)
    dense = np.zeros((3 * n_nodes, 3 * n_nodes))
    for r in range(n_nodes):
        for p in range(rp[r], rp[r + 1]):
            c = int(cc[p])
            dense[3 * r:3 * r + 3, 3 * c:3 * c + 3] = va[9 * p:9 * p + 9].reshape(3, 3)
    assert np.array_equal(dense, Kc.to_scipy().toarray())
    # dinv is the inverse of each diagonal block
    D = np.stack([dense[3 * r:3 * r + 3, 3 * r:3 * r + 3] for r in range(n_nodes)])
    Dinv = di.reshape(-1, 3, 3)
    assert np.abs(Dinv @ D - np.eye(3)).max() < 1e-12
    assert np.allclose(Dinv, np.linalg.inv(D), rtol=1e-12, atol=0)
    # the exported arrays drive the block solver to the device CG's solution
    res = fh.solve_bsr3_pcg_gpu(n_nodes, rp, cc, va, di, bc, None, 1e-12, 5000)
    x, st = Kc.solve(bc, None, 1e-12, 5000)
    assert np.linalg.norm(res.solution - x) / np.linalg.norm(x) < 1e-6  # cond(K) ~ 5e4
    ctx = fh.GpuBsr3Context(n_nodes, rp, cc, va, di)
    assert np.allclose(ctx.spmv(x), Kc.spmv(x), rtol=1e-13, atol=1e-9 * np.abs(bc).max())
    # the scalar operator has no 3x3 blocks
    with pytest.raises(fh.FenicsError, match="3x3"):
        fh.assemble_stiffness_2d(fh.unit_square(3, 3)).export_bsr3()


def test_calls_from_two_threads_match_sequential_calls():
    """Calls from several threads share each device's stream and CG state, so
    the library runs them one after the other: two threads assembling and
    solving different systems get bitwise the results of a sequential run."""
    import threading

    def work(N):
        m = fh.unit_square(N, N)
        K = fh.assemble_stiffness(m)
        bd = m.boundary_dofs()
        b = np.ones(m.num_vertices()) / N ** 2
        Kb, bb = fh.apply_dirichlet_bc(K, b, fh.DirichletBC(bd, np.zeros(len(bd))))
        return [Kb.solve(bb, tol=1e-10, max_iter=5000, precision=prec)[0]
                for prec in (fh.FP64, fh.FP64_TRUE, fh.MIXED)]

    sizes = (40, 56)
    seq = [work(N) for N in sizes]
    out, errs = [None, None], []

    def run(i):
        try:
            for _ in range(3):
                out[i] = work(sizes[i])
        except BaseException as e:  # reported after join
            errs.append(e)

    ts = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for a, b in zip(seq, out):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_repeated_solves_reuse_the_kept_workspace_bitwise():
    """A solve keeps its workspace and iteration graphs on the matrix; later
    solves of the same operator in the same configuration replay them.  Every
    repeat (with solves in other precisions in between) returns the bits of a
    solve of a freshly assembled copy of the operator, with the same
    iteration counts."""
    import torch
    m = fh.box_beam(10.0, 2.0, 2.0, 40, 8, 8)
    mat = fh.ElasticMaterial(210e9, 0.3)
    bc = fh.clamp_dofs_x0(m.vertices)

    def system():
        K = fh.assemble_elasticity_stiffness_3d(m, mat)
        b = fh.assemble_elasticity_body_force_3d(m, [0.0, 0.0, -7.85e4])
        return fh.apply_vector_dirichlet_bc(K, b, bc)

    Kb, bb = system()
    bt = torch.tensor(bb, device="cuda:0")
    runs = {}
    for prec in (fh.FP64_TRUE, fh.MIXED, fh.FP64_TRUE, fh.FP64, fh.MIXED, fh.FP64_TRUE):
        x = torch.empty_like(bt)
        st = Kb.solve_dev(bt, x, tol=1e-10, max_iter=20000, precision=prec)
        runs.setdefault(prec, []).append((x.cpu().numpy().view(np.uint64), st.iterations))
    for prec, got in runs.items():
        K2, b2 = system()
        x2 = torch.empty_like(bt)
        st2 = K2.solve_dev(torch.tensor(b2, device="cuda:0"), x2, tol=1e-10, max_iter=20000,
                           precision=prec)
        ref = x2.cpu().numpy().view(np.uint64)
        for bits, its in got:
            assert its == st2.iterations
            assert np.array_equal(bits, ref), prec
