"""Pins the oracle's transient heat restatement (transient.rs, heat_solver.rs)
to the reference's known answers and to independent mathematics.  CPU only.

Reference KATs restated: transient.rs:104-126 (Crank-Nicolson exponential
decay), heat_solver.rs:116-171 (steady state at x = 0.5), ab_benchmarks.rs:
136-151 (Crank-Nicolson with a constant source).  Independent pins: the
schemes' closed-form amplification factors (scalar and, through the
generalized eigenvectors of (K, M), on a mesh), heat conservation
1^T M u_{n+1} = 1^T M u_n + theta dt 1^T (f terms) (1^T K = 0 for the pure
Neumann stiffness), and add_scaled_matrices against dense arithmetic on two
different patterns.
"""
import numpy as np
import scipy.linalg as sl

import oracle as O


def scalar(v):
    return O.csc_from_triplets(1, [0], [0], [v])


def test_exponential_decay_crank_nicolson():  # transient.rs:104-126
    M, K = scalar(1.0), scalar(2.0)
    st = O.TransientStepper(M, K, O.CRANK_NICOLSON, 0.01)
    u = np.array([1.0])
    f = np.zeros(1)
    for _ in range(100):
        u = st.step(u, f, f)
    assert abs(u[0] - np.exp(-2.0)) < 1e-4
    g = (1.0 - 0.01) / (1.0 + 0.01)  # CN amplification (1 - lam dt/2) / (1 + lam dt/2)
    assert abs(u[0] - g ** 100) < 1e-14


def test_exponential_decay_backward_euler():
    u = np.array([1.0])
    f = np.zeros(1)
    for _ in range(100):
        u = O.transient_step(scalar(1.0), scalar(2.0), O.BACKWARD_EULER, 0.01, u, f, f)
    assert abs(u[0] - (1.0 / 1.02) ** 100) < 1e-14
    # forcing enters BE at t_{n+1} only: du/dt = 1 from u = 0 gives u = dt after a step
    one = np.ones(1)
    assert O.transient_step(scalar(1.0), scalar(0.0), O.BACKWARD_EULER, 0.5, np.zeros(1),
                            np.zeros(1), one)[0] == 0.5
    assert O.transient_step(scalar(1.0), scalar(0.0), O.CRANK_NICOLSON, 0.5, np.zeros(1),
                            np.zeros(1), one)[0] == 0.25


def test_add_scaled_matrices_union_of_patterns():
    rng = np.random.default_rng(5)
    n = 40

    def rand_csc(nnz):
        r, c = rng.integers(0, n, nnz), rng.integers(0, n, nnz)
        return O.csc_from_triplets(n, r, c, rng.standard_normal(nnz))

    A, B = rand_csc(150), rand_csc(120)
    for alpha in (0.3, -1.75, 0.0):
        S = O.add_scaled_matrices(A, B, alpha)
        Ad, Bd = A.to_scipy().toarray(), B.to_scipy().toarray()
        assert np.array_equal(S.to_scipy().toarray(), Ad + alpha * Bd)
        pa, pb = Ad != 0, Bd != 0
        assert S.nnz == np.count_nonzero(pa | pb)  # the union keeps every stored position


def spectral_march(m, kappa, u0, dt, steps, scheme):
    M = O.assemble_mass_2d(m).to_scipy().toarray()
    K = kappa * O.assemble_stiffness(m).to_scipy().toarray()
    mu, phi = sl.eigh(K, M)
    c = phi.T @ (M @ u0)
    if scheme == O.CRANK_NICOLSON:
        g = (1 - 0.5 * dt * mu) / (1 + 0.5 * dt * mu)
    else:
        g = 1 / (1 + dt * mu)
    return phi @ (g ** steps * c)


def test_march_matches_generalized_eigen_closed_form():
    m = O.unit_square(5, 4)
    x, y = m.vertices[:, 0], m.vertices[:, 1]
    u0 = x * x + np.sin(3 * y)
    for scheme in (O.BACKWARD_EULER, O.CRANK_NICOLSON):
        _, hist = O.transient_heat_2d(m, 0.7, u0, [], 0.03, 12, scheme)
        ref = spectral_march(m, 0.7, u0, 0.03, 12, scheme)
        assert np.abs(hist[-1] - ref).max() < 1e-11 * np.abs(ref).max()
        assert np.array_equal(hist[0], u0)


def test_heat_conservation_with_time_dependent_source():
    m = O.unit_square(6, 5)
    M = O.assemble_mass_2d(m).to_scipy()
    src = lambda t, p: 10.0 + 5.0 * t + p[0] * p[1]  # noqa: E731
    u0 = np.cos(m.vertices[:, 0])
    dt, steps = 0.02, 9
    for scheme in (O.BACKWARD_EULER, O.CRANK_NICOLSON):
        times, hist = O.transient_heat_2d(m, 1.3, u0, [], dt, steps, scheme, src)
        total = M @ u0
        heat = total.sum()
        for k in range(1, steps + 1):
            f0 = np.array([src(times[k - 1], p) for p in m.vertices])
            f1 = np.array([src(times[k], p) for p in m.vertices])
            heat += (0.5 * dt * (f0 + f1)).sum() if scheme == O.CRANK_NICOLSON else (dt * f1).sum()
            assert abs((M @ hist[k]).sum() - heat) < 1e-12 * abs(heat)


def test_steady_state_kat():  # heat_solver.rs:116-171
    m = O.unit_square(8, 8)
    x = m.vertices[:, 0]
    left = np.nonzero(np.abs(x) < 1e-6)[0]
    right = np.nonzero(np.abs(x - 1.0) < 1e-6)[0]
    bcs = [(left, [100.0] * len(left)), (right, [0.0] * len(right))]
    times, hist = O.transient_heat_2d(m, 1.0, np.zeros(m.nv), bcs, 0.05, 40, O.CRANK_NICOLSON)
    mid = np.abs(x - 0.5) < 1e-6
    assert mid.sum() == 9
    assert np.all(np.abs(hist[-1][mid] - 50.0) <= 1.5)
    assert len(times) == 41 and times[-1] == 40 * 0.05
    assert np.all(hist[-1][left] == 100.0) and np.all(hist[-1][right] == 0.0)


def test_constant_source_kat():  # ab_benchmarks.rs:136-151
    m = O.unit_square(6, 6)
    times, hist = O.transient_heat_2d(m, 1.0, np.zeros(m.nv), [], 0.01, 10, O.CRANK_NICOLSON,
                                      lambda t, p: 10.0)
    assert len(times) == 11
    assert hist[-1][0] > 0.0
    # the nodal source enters as a load vector (not M f): the total heat after
    # t = 0.1 is t * sum_i f_i while the profile is not uniform
    heat = (O.assemble_mass_2d(m).to_scipy() @ hist[-1]).sum()
    assert abs(heat - 0.1 * 10.0 * m.nv) < 1e-12 * heat
    assert hist[-1].max() > 1.1 * hist[-1].min()
