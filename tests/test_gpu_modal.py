"""GPU modal analysis (the CG's consumer in eigen.rs) against the oracle and
the reference's modal tests (modal_tests.rs).

Tolerances: the elasticity mass and the Dirichlet folds are assembled with
the oracle's arithmetic, so their CSC exports are bitwise equal.  The power
iteration runs the reference's sequence on the device; dot products are tree
sums and the inner CG differs from cg.rs in its dot order, so eigenvalues from
tightly converged runs (tol 1e-12, inner 1e-12) agree with the oracle's to
1e-9 relative and mode shapes to 1e-6 (up to sign); loosely converged runs
(tol 1e-5, the reference's test) are compared through the reference's own
frequency ranges and to the oracle at 1e-4."""
import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def beam(nx, ny, nz, rho=1000.0):
    gm, om = fh.box_beam(1.0, 0.1, 0.1, nx, ny, nz), O.box_beam(1.0, 0.1, 0.1, nx, ny, nz)
    K = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(1.0e7, 0.3))
    M = fh.assemble_elasticity_mass_3d(gm, rho)
    Ko = O.assemble_elasticity(om, 1.0e7, 0.3)
    Mo = O.assemble_elasticity_mass_3d(om, rho)
    fixed = np.array([d for d, _ in fh.clamp_dofs_x0(gm.vertices).prescribed_dofs], np.uint64)
    return gm, om, K, M, Ko, Mo, fixed


def same_csc(a, b):
    c1, r1, v1 = a.export_csc()
    c2, r2, v2 = b.export()
    return np.array_equal(c1, c2) and np.array_equal(r1, r2) and np.array_equal(bits(v1), bits(v2))


def test_elasticity_mass_and_folds_bitwise():
    gm, om, K, M, Ko, Mo, fixed = beam(4, 2, 2, rho=7.5)
    assert same_csc(M, Mo)
    assert M.nnz() == len(Mo.export()[2])  # the mass couples equal components only
    assert same_csc(fh.fold_dirichlet(K, fixed, True), O.fold_dirichlet(Ko, fixed, True))
    assert same_csc(fh.fold_dirichlet(M, fixed, False), O.fold_dirichlet(Mo, fixed, False))
    # duplicates and out-of-range indices (eigen.rs:57-62) are harmless
    dup = np.concatenate([fixed, fixed[:5], [10 ** 9]]).astype(np.uint64)
    assert same_csc(fh.fold_dirichlet(K, dup, True), O.fold_dirichlet(Ko, fixed, True))


def test_p2_vector_mass_bitwise():
    om = O.box_beam(2.0, 1.0, 0.5, 5, 3, 2)
    p2 = O.P2(om)
    X, cd = p2.dof_coords(), p2.cell_dofs()
    M, Mo = fh.assemble_elasticity_mass_p2_3d(X, cd, 7.5), O.assemble_elasticity_mass_p2_3d(X, cd, 7.5)
    assert same_csc(M, Mo)
    assert M.nnz() == len(Mo.export()[2])
    nodes = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0],
             [0, 0, 0.5], [0.5, 0, 0.5], [0, 0.5, 0.5]]  # mass.rs:296-330
    M1 = fh.assemble_elasticity_mass_p2_3d(nodes, [list(range(10))], 1.0)
    one = np.ones(30)
    assert abs(one @ M1.spmv(one) - 0.5) < 1e-13


def test_cantilever_natural_frequencies():  # modal_tests.rs:9-86
    gm, om, K, M, Ko, Mo, fixed = beam(16, 2, 2)
    modes = fh.solve_vibration_modes(K, M, fixed, 3, 30, 1e-5)
    f = np.array([md.frequency_hz for md in modes])
    assert 1.0 < f[0] < 3.0 and 1.5 < f[1] < 3.5 and 8.0 < f[2] < 20.0
    st, fo, _, _, _ = O.vibration_modes(Ko, Mo, fixed, 3, 30, 1e-5)
    assert st == 0 and np.allclose(f, fo, rtol=1e-4, atol=0)
    for md in modes:
        assert abs(md.omega_rad_s - 2 * np.pi * md.frequency_hz) < 1e-12 * md.omega_rad_s
        assert not md.mode_shape[fixed].any()


def test_frequencies_scale_as_one_over_sqrt_density():  # modal_tests.rs:88-113
    gm, om, K, M1, Ko, Mo, fixed = beam(8, 1, 1)
    M2 = fh.assemble_elasticity_mass_3d(gm, 1000.0 * 10000.0)
    a = fh.solve_vibration_modes(K, M1, fixed, 2, 60, 1e-8)
    b = fh.solve_vibration_modes(K, M2, fixed, 2, 60, 1e-8)
    for hi, lo in zip(b, a):
        assert abs(hi.frequency_hz / lo.frequency_hz - 0.01) < 1e-4


def test_tight_modes_match_oracle():
    """The power iteration stops when the Rayleigh quotient moves by < 1e-13:
    the quotient is accurate to about that, the vectors only to about its
    square root over the spectral gap, so shapes are compared at 1e-5 and each
    GPU shape is checked as an eigenvector of the folded pencil directly."""
    gm, om, K, M, Ko, Mo, fixed = beam(12, 2, 2)
    modes = fh.solve_vibration_modes(K, M, fixed, 3, 300, 1e-13, inner_tol=1e-12,
                                     inner_max_iter=5000)
    st, fo, omo, evo, sho = O.vibration_modes(Ko, Mo, fixed, 3, 300, 1e-13, inner_tol=1e-12,
                                              inner_max_iter=5000)
    assert st == 0
    ev = np.array([md.eigenvalue for md in modes])
    assert np.allclose(ev, evo, rtol=1e-9, atol=0)
    Kd = O.fold_dirichlet(Ko, fixed, True).to_scipy()
    Md = O.fold_dirichlet(Mo, fixed, False).to_scipy()
    for md, so in zip(modes, sho):
        sgn = np.sign(md.mode_shape @ so)
        assert np.linalg.norm(sgn * md.mode_shape - so) < 1e-5 * np.linalg.norm(so)
        kphi = Kd @ md.mode_shape
        res = kphi - md.eigenvalue * (Md @ md.mode_shape)
        assert np.linalg.norm(res) < 1e-6 * np.linalg.norm(kphi)  # oracle's own: 2e-10..2e-7
    assert modes.inner_iterations > 0


def test_inner_solve_failure_is_reported():
    """eigen.rs:190 propagates the inner solve's Err: an iteration cap that CG
    cannot meet makes the whole call fail with the CG's status."""
    gm, om, K, M, Ko, Mo, fixed = beam(16, 2, 2)
    with pytest.raises(fh.FenicsError) as e:
        fh.solve_vibration_modes(K, M, fixed, 1, 5, 1e-8, inner_tol=1e-14, inner_max_iter=3)
    assert e.value.code == fh.FH_ERR_NOT_CONVERGED


def test_with_inner_device_cg_callback_is_bitwise_the_default():
    """eigen.rs:282-300: solve_vibration_modes is solve_vibration_modes_with_inner
    with solve_cg on the folded K; the device CG passed back as the caller's
    inner solve (b and x0 through host memory) gives bitwise the same modes."""
    gm, om, K, M, Ko, Mo, fixed = beam(8, 2, 2)
    Kd = fh.fold_dirichlet(K, fixed, True)
    calls = []

    def inner(b, x0):
        calls.append(len(b))
        return Kd.solve(b, x0, 1e-9, 1000)[0]

    a = fh.solve_vibration_modes(K, M, fixed, 2, 100, 1e-8)
    b = fh.solve_vibration_modes_with_inner(K, M, fixed, 2, 100, 1e-8, inner)
    assert calls and all(c == K.n for c in calls)
    for ma, mb in zip(a, b):
        assert bits(ma.eigenvalue) == bits(mb.eigenvalue)
        assert np.array_equal(bits(ma.mode_shape), bits(mb.mode_shape))


def test_with_inner_direct_solve_matches_oracle():
    """A caller's sparse direct inner solve (eigen.rs:89-93: any solver of the
    folded K) reaches the oracle's tightly converged eigenvalues."""
    import scipy.sparse.linalg as spla
    gm, om, K, M, Ko, Mo, fixed = beam(12, 2, 2)
    lu = spla.splu(O.fold_dirichlet(Ko, fixed, True).to_scipy().tocsc())
    modes = fh.solve_vibration_modes_with_inner(K, M, fixed, 3, 300, 1e-13,
                                                lambda b, x0: lu.solve(b))
    st, fo, omo, evo, sho = O.vibration_modes(Ko, Mo, fixed, 3, 300, 1e-13, inner_tol=1e-12,
                                              inner_max_iter=5000)
    assert st == 0
    assert np.allclose([md.eigenvalue for md in modes], evo, rtol=1e-9, atol=0)
    assert np.allclose([md.frequency_hz for md in modes], fo, rtol=1e-9, atol=0)
    for md in modes:  # constrained entries stay at exactly zero
        assert np.all(md.mode_shape[fixed.astype(np.int64)] == 0.0)


def test_with_inner_errors_propagate():
    """eigen.rs:189 propagates the inner closure's Err."""
    gm, om, K, M, Ko, Mo, fixed = beam(4, 2, 2)

    def boom(b, x0):
        raise ValueError("inner solver exploded")

    with pytest.raises(ValueError, match="exploded"):
        fh.solve_vibration_modes_with_inner(K, M, fixed, 1, 10, 1e-8, boom)
    with pytest.raises(fh.FenicsError, match="shape"):
        fh.solve_vibration_modes_with_inner(K, M, fixed, 1, 10, 1e-8, lambda b, x0: b[:-1])
    # the library stays usable afterwards
    assert fh.solve_vibration_modes(K, M, fixed, 1, 50, 1e-6)[0].frequency_hz > 0
