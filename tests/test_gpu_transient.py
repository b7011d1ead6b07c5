"""GPU transient heat conduction (transient.rs, heat_solver.rs) against the
oracle and the reference's known answers.

Tolerances: add_scaled_matrices and the CSC round trip are bitwise (the same
a + alpha * b per stored position, the same union).  A time step replaces the
reference's dense LU by the device CG, so states agree with the oracle's LU
solution to the CG tolerance times the step matrix's condition number: one
step at tol 1e-13 to 1e-10 of max|u|, a 15-step march (errors are damped by
the stable schemes, not amplified) to 1e-9 of max|u|.  The full-size march
(unit_square(1024, 1024), 1,050,625 dofs) is checked step by step against
the oracle's right-hand side and a sparse direct solve on its first step,
and FP32-assisted (MIXED) against FP64 marching to 1e-8 of max|u|.
"""
import time

import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu

BE, CN = fh.TimeSteppingScheme.BackwardEuler, fh.TimeSteppingScheme.CrankNicolson


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_csc(a, b):
    c1, r1, v1 = a.export_csc()
    c2, r2, v2 = b.export()
    return np.array_equal(c1, c2) and np.array_equal(r1, r2) and np.array_equal(bits(v1), bits(v2))


def walls(m):
    x = m.vertices[:, 0]
    left = np.nonzero(np.abs(x) < 1e-6)[0]
    right = np.nonzero(np.abs(x - 1.0) < 1e-6)[0]
    return left, right


def test_add_scaled_same_pattern_bitwise():
    gm, om = fh.unit_square(8, 8), O.unit_square(8, 8)
    M, K = fh.assemble_mass(gm), fh.assemble_stiffness(gm)
    Mo, Ko = O.assemble_mass_2d(om), O.assemble_stiffness(om)
    for alpha in (0.05, -0.025, 0.0, 3.5):
        assert same_csc(fh.add_scaled_matrices(M, K, alpha), O.add_scaled_matrices(Mo, Ko, alpha))
    # 3x3 blocks: the elasticity mass (equal components only) plus the stiffness
    gb, ob = fh.box_beam(1.0, 0.2, 0.2, 4, 2, 2), O.box_beam(1.0, 0.2, 0.2, 4, 2, 2)
    Me, Ke = fh.assemble_elasticity_mass_3d(gb, 7.8), fh.assemble_elasticity_stiffness_3d(
        gb, fh.ElasticMaterial(2.0e5, 0.3))
    Meo, Keo = O.assemble_elasticity_mass_3d(ob, 7.8), O.assemble_elasticity(ob, 2.0e5, 0.3)
    assert same_csc(fh.add_scaled_matrices(Me, Ke, 1e-3), O.add_scaled_matrices(Meo, Keo, 1e-3))
    assert same_csc(fh.add_scaled_matrices(Ke, Me, -2.0), O.add_scaled_matrices(Keo, Meo, -2.0))
    assert same_csc(fh.add_scaled_matrices(Me, Me, 0.5), O.add_scaled_matrices(Meo, Meo, 0.5))


def test_csc_round_trip_and_general_union_bitwise():
    rng = np.random.default_rng(11)
    n = 70

    def rand_csc(nnz):
        r, c = rng.integers(0, n, nnz), rng.integers(0, n, nnz)
        return O.csc_from_triplets(n, r, c, rng.standard_normal(nnz))

    A, B = rand_csc(300), rand_csc(260)
    gA, gB = fh.matrix_from_csc(*A.export()), fh.matrix_from_csc(*B.export())
    assert same_csc(gA, A) and same_csc(gB, B)
    for alpha in (0.3, -1.75):
        assert same_csc(fh.add_scaled_matrices(gA, gB, alpha), O.add_scaled_matrices(A, B, alpha))
    # a condensed operand goes through the exported (condensed) parts
    gm, om = fh.unit_square(6, 6), O.unit_square(6, 6)
    bnd = O.boundary_dofs(om)
    Kc, _ = fh.apply_dirichlet_bc(fh.assemble_stiffness(gm), np.zeros(gm.num_vertices()),
                                  fh.DirichletBC(list(bnd), [0.0] * len(bnd)))
    Kco, _ = O.apply_dirichlet_bc(O.assemble_stiffness(om), np.zeros(om.nv), bnd,
                                  np.zeros(len(bnd)))
    M, Mo = fh.assemble_mass(gm), O.assemble_mass_2d(om)
    assert same_csc(fh.add_scaled_matrices(M, Kc, 0.1), O.add_scaled_matrices(Mo, Kco, 0.1))
    # folded operands (eigen.rs:51-78 drops their exact zeros) share the stored
    # pattern but not the exported one; entries zero in both must not appear
    Kf, Mf = fh.fold_dirichlet(fh.assemble_stiffness(gm), bnd, True), fh.fold_dirichlet(M, bnd, False)
    Kfo, Mfo = O.fold_dirichlet(O.assemble_stiffness(om), bnd, True), O.fold_dirichlet(Mo, bnd, False)
    for alpha in (0.25, -1.0):
        assert same_csc(fh.add_scaled_matrices(Mf, Kf, alpha),
                        O.add_scaled_matrices(Mfo, Kfo, alpha))
    assert same_csc(fh.add_scaled_matrices(M, Kf, 0.5), O.add_scaled_matrices(Mo, Kfo, 0.5))


def test_step_exponential_decay_kat():  # transient.rs:104-126
    one = lambda v: fh.matrix_from_csc([0, 1], [0], [v])  # noqa: E731
    M, K = one(1.0), one(2.0)
    solver = fh.TransientHeatSolver(CN, 0.01)
    u, f = np.array([1.0]), np.zeros(1)
    for _ in range(100):
        u = solver.step(M, K, u, f, f)
    assert abs(u[0] - np.exp(-2.0)) < 1e-4
    assert abs(u[0] - ((1.0 - 0.01) / (1.0 + 0.01)) ** 100) < 1e-13
    be = fh.TransientHeatSolver(BE, 0.5)
    assert abs(be.step(one(1.0), one(0.0), np.zeros(1), np.zeros(1), np.ones(1))[0] - 0.5) < 1e-15


def test_step_matches_oracle():
    gm, om = fh.unit_square(16, 16), O.unit_square(16, 16)
    M, K = fh.assemble_mass(gm), fh.assemble_stiffness(gm)
    Mo, Ko = O.assemble_mass_2d(om), O.assemble_stiffness(om)
    x, y = om.vertices[:, 0], om.vertices[:, 1]
    u = np.sin(np.pi * x) * np.cos(2 * y) + x
    f0, f1 = 3.0 + x * y, 3.5 + x * y
    for scheme in (BE, CN):
        s = fh.TransientHeatSolver(scheme, 0.01, tol=1e-13)
        got = s.step(M, K, u, f0, f1)
        ref = O.transient_step(Mo, Ko, scheme, 0.01, u, f0, f1)
        assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()
        assert s.last_stats.iterations > 0


def test_march_matches_oracle_with_bcs_and_source():
    gm, om = fh.unit_square(16, 16), O.unit_square(16, 16)
    left, right = walls(om)
    # a repeated dof: the later value wins (heat_solver.rs:82-86 assigns in order)
    bcs = [fh.DirichletBC(list(left), [80.0] * len(left)),
           fh.DirichletBC(list(right), [5.0] * len(right)),
           fh.DirichletBC([int(left[0])], [90.0])]
    obcs = [(b.dofs, b.values) for b in bcs]
    src = lambda t, p: 20.0 * np.sin(np.pi * p[1]) * (1.0 + t)  # noqa: E731
    x = om.vertices[:, 0]
    u0 = 10.0 * x * (1 - x)
    for scheme in (BE, CN):
        sol = fh.solve_transient_heat_2d(gm, 0.8, u0, bcs, 0.02, 15, scheme, src, tol=1e-13)
        times, hist = O.transient_heat_2d(om, 0.8, u0, obcs, 0.02, 15, scheme, src)
        assert sol.time_steps == times
        got = np.array(sol.temperature_history)
        assert np.array_equal(got[0], u0)
        assert np.abs(got - hist).max() < 1e-9 * np.abs(hist).max()
        assert np.array_equal(sol.final_temperature, got[-1])
        assert sol.final_temperature[left[0]] == 90.0 and np.all(sol.final_temperature[right] == 5.0)
        assert sol.stats["steps"] == 15 and sol.stats["cg_iterations"] >= 15


def test_steady_state_kat():  # heat_solver.rs:116-171
    m = fh.unit_square(8, 8)
    left, right = walls(m)
    bcs = [fh.DirichletBC(list(left), [100.0] * len(left)),
           fh.DirichletBC(list(right), [0.0] * len(right))]
    sol = fh.solve_transient_heat_2d(m, 1.0, np.zeros(m.num_vertices()), bcs, 0.05, 40, CN)
    mid = np.abs(m.vertices[:, 0] - 0.5) < 1e-6
    assert mid.sum() == 9
    assert np.all(np.abs(sol.final_temperature[mid] - 50.0) <= 1.5)
    assert len(sol.time_steps) == 41


def test_constant_source_kat_and_errors():  # ab_benchmarks.rs:136-151
    m = fh.unit_square(6, 6)
    sol = fh.solve_transient_heat_2d(m, 1.0, np.zeros(m.num_vertices()), [], 0.01, 10, CN,
                                     lambda t, p: 10.0)
    assert len(sol.time_steps) == 11 and sol.final_temperature[0] > 0.0
    times, hist = O.transient_heat_2d(O.unit_square(6, 6), 1.0, np.zeros(49), [], 0.01, 10, CN,
                                      lambda t, p: 10.0)
    assert np.abs(sol.final_temperature - hist[-1]).max() < 1e-9 * np.abs(hist[-1]).max()
    with pytest.raises(fh.FenicsError, match="Initial temperature size 3 != mesh vertices 49"):
        fh.solve_transient_heat_2d(m, 1.0, np.zeros(3), [], 0.01, 1, CN)
    with pytest.raises(fh.FenicsError) as e:
        fh.solve_transient_heat_2d(m, 1.0, np.zeros(49), [fh.DirichletBC([49], [1.0])], 0.01, 1, CN)
    assert e.value.code == fh.FH_ERR_BAD_INPUT
    with pytest.raises(ZeroDivisionError):  # an exception in the source reaches the caller
        fh.solve_transient_heat_2d(m, 1.0, np.zeros(49), [], 0.01, 1, CN, lambda t, p: 1 / 0)


def full_size_bcs(m):
    left, right = walls(m)
    return [fh.DirichletBC(list(left), [100.0] * len(left)),
            fh.DirichletBC(list(right), [0.0] * len(right))]


def test_full_size_first_step_vs_sparse_direct():
    """unit_square(512, 512) (263,169 dofs): the first CN step against the
    oracle's right-hand side solved by a sparse direct factorization."""
    import scipy.sparse.linalg as spla
    N = 512
    gm, om = fh.unit_square(N, N), O.unit_square(N, N)
    bcs = full_size_bcs(gm)
    dt = 1e-4
    sol = fh.solve_transient_heat_2d(gm, 1.0, np.zeros(gm.num_vertices()), bcs, dt, 1, CN,
                                     tol=1e-13)
    st = O.TransientStepper.__new__(O.TransientStepper)
    Mo, Ko = O.assemble_mass_2d(om), O.assemble_stiffness(om)
    st.cn, st.h = True, 0.5 * dt
    st.lhs = O.add_scaled_matrices(Mo, Ko, st.h)
    st.rmat = O.add_scaled_matrices(Mo, Ko, -st.h)
    u = np.zeros(om.nv)
    for b in bcs:
        u[np.asarray(b.dofs)] = b.values
    rhs = st.rhs(u, np.zeros(om.nv), np.zeros(om.nv))
    ref = spla.spsolve(st.lhs.to_scipy().tocsc(), rhs)
    for b in bcs:
        ref[np.asarray(b.dofs)] = b.values
    assert np.abs(sol.final_temperature - ref).max() < 1e-9 * 100.0


def test_full_size_march_mixed_vs_fp64():
    """unit_square(1024, 1024), 1,050,625 dofs, CN, 10 steps of dt = 1e-4."""
    N = 1024
    gm = fh.unit_square(N, N)
    bcs = full_size_bcs(gm)
    n = gm.num_vertices()
    out = {}
    for prec in (fh.FP64_TRUE, fh.MIXED):
        t0 = time.perf_counter()
        sol = fh.solve_transient_heat_2d(gm, 1.0, np.zeros(n), bcs, 1e-4, 10, CN, tol=1e-10,
                                         precision=prec, keep_history=False)
        wall = time.perf_counter() - t0
        out[prec] = sol.final_temperature
        st = sol.stats
        print(f"\nheat {N}x{N} prec={prec}: {wall * 1e3:.1f} ms, solve {st['solve_ms']:.1f} ms, "
              f"cg {st['cg_iterations']} (low {st['cg_iterations_low']}, max {st['cg_iterations_max']})",
              flush=True)
        u = sol.final_temperature
        assert np.all(np.isfinite(u)) and u.max() <= 100.0 + 1e-6 and u.min() > -1.0
    assert np.abs(out[fh.MIXED] - out[fh.FP64_TRUE]).max() < 1e-8 * 100.0
