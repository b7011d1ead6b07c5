"""The headline solve at full size: the 1,013,643-DOF cantilever
box_beam(10, 2, 2, 200, 40, 40), E = 210e9, nu = 0.3, clamped at x = 0, body
force (0, 0, -7.85e4) (elasticity_solver.rs:31-72 on the C1 mesh).

What pins it (no oracle CG at this size: the single-thread cg.rs loop would
take minutes):

* a double-double solution x_dd = x_hi + x_lo whose double-double residual is
  below 1e-14 (iterative refinement through fh_solve_refined_dev), so x_dd is
  the solution of the assembled system to about kappa * 1e-14;
* the residual kernel that certifies it is checked against an exact
  evaluation (products split by Dekker's method, rows summed by math.fsum) of
  the oracle's CSC on a smaller cantilever;
* x_hi is x_dd rounded to FP64: its true residual is the measured floor that
  no FP64 vector near the solution goes below by much;
* the tip deflection against the Euler-Bernoulli closed form q L^4 / (8 E I).

Tolerances: mixed and FP64_TRUE solutions agree with x_dd to 1e-9 relative
(kappa ~ 1e8 times their residuals of a few 1e-10 bounds the error by ~1e-2
in the worst case; the measured errors are ~1e-12, so 1e-9 leaves a wide
margin while catching any solve that stopped early); their true residuals are
below tol or at most 2.5x the measured floor (1.92e-10 on this system, so
1e-10 is out of reach for any FP64 vector near the solution; the solves stop
at about 2.2x it); the reference-semantics solve (recursive-residual stop)
agrees with x_dd to 1e-6.
"""
import numpy as np
import pytest

import fenics_hip as fh
import oracle as O
from exactsum import exact_residual

pytestmark = pytest.mark.gpu

E, NU = 210e9, 0.3
F = [0.0, 0.0, -7.85e4]
TOL = 1e-10


def test_true_residual_kernel_vs_exact():
    import torch
    dims = (10.0, 2.0, 2.0, 20, 4, 4)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    K = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(E, NU))
    b = fh.assemble_elasticity_body_force_3d(gm, F)
    bc = fh.clamp_dofs_x0(gm.vertices)
    Kb, bb = fh.apply_vector_dirichlet_bc(K, b, bc)
    Ko = O.assemble_elasticity(om, E, NU)
    bo = O.assemble_body_force(om, F)
    fx = [d for d, _ in bc.prescribed_dofs]
    Kbo, bbo = O.apply_dirichlet_bc(Ko, bo, fx, np.zeros(len(fx)))
    st, uo, it, res = O.solve_cg(Kbo, bbo, None, 1e-12, 20000)
    rng = np.random.default_rng(5)
    dev = "cuda:0"
    for x in (uo, uo * (1 + 1e-7 * rng.standard_normal(len(uo))), rng.standard_normal(len(uo))):
        ref = exact_residual(Kbo, bbo, x)
        got = Kb.true_residual_dev(torch.tensor(bb, device=dev), torch.tensor(x, device=dev))
        assert abs(got - ref) <= 1e-12 * ref, (got, ref)
    # x = hi + lo, with lo far below the last bit of hi
    lo = 1e-20 * rng.standard_normal(len(uo)) * np.abs(uo)
    hi = uo
    got = Kb.true_residual_dev(torch.tensor(bb, device=dev), torch.tensor(hi, device=dev),
                               torch.tensor(lo, device=dev))
    ref = exact_residual(Kbo, bbo, hi, lo)
    assert abs(got - ref) <= 1e-12 * ref, (got, ref)


def test_c1_full_size_solve():
    import torch
    gm = fh.box_beam(10.0, 2.0, 2.0, 200, 40, 40)
    n = 3 * gm.nv
    assert n == 1013643
    K = fh.assemble_elasticity_stiffness_3d(gm, fh.ElasticMaterial(E, NU))
    b = fh.assemble_elasticity_body_force_3d(gm, F)
    bc = fh.clamp_dofs_x0(gm.vertices)
    Kb, bb = fh.apply_vector_dirichlet_bc(K, b, bc)
    del K
    dev = "cuda:0"
    bt = torch.tensor(bb, device=dev)

    # the double-double solution: certified residual, then rounded to FP64
    xh = torch.empty(n, dtype=torch.float64, device=dev)
    xl = torch.empty_like(xh)
    rs = Kb.solve_refined_dev(bt, xh, xl, target=1e-14, max_outer=8, inner_tol=1e-6,
                              max_iter=20000)
    floor = rs.rel_residual_rounded
    print(f"\nC1 double-double solution: {rs.outer_steps} correction solves, "
          f"{rs.cg_iterations} CG iterations, residual {rs.rel_residual_pair:.3e}; "
          f"rounded to FP64: {floor:.3e}", flush=True)
    assert rs.rel_residual_pair < 1e-14
    assert floor > 0.0
    xdd = xh.double() + xl
    nx = torch.linalg.norm(xdd)

    sols = {}
    for prec in (fh.FP64_TRUE, fh.MIXED):
        u = torch.empty_like(xh)
        st = Kb.solve_dev(bt, u, tol=TOL, max_iter=50000, precision=prec)
        err = float(torch.linalg.norm(u - xdd) / nx)
        print(f"precision {prec}: {st.iterations} it ({st.iterations_low} FP32), true rel res "
              f"{st.true_rel_residual:.3e}, estimated floor {st.residual_floor:.3e}, "
              f"error vs x_dd {err:.2e}", flush=True)
        assert st.true_rel_residual < TOL or st.true_rel_residual <= 2.5 * floor
        # the library's own evaluation agrees with the stats
        again = Kb.true_residual_dev(bt, u)
        assert abs(again - st.true_rel_residual) <= 1e-6 * again
        assert err < 1e-9
        sols[prec] = u
    assert float(torch.linalg.norm(sols[fh.MIXED] - sols[fh.FP64_TRUE]) / nx) < 1e-9

    # the reference's recursive-residual stop on the same system
    u = torch.empty_like(xh)
    st = Kb.solve_dev(bt, u, tol=TOL, max_iter=50000, precision=fh.FP64)
    err = float(torch.linalg.norm(u - xdd) / nx)
    print(f"reference semantics: {st.iterations} it, recursive {st.rel_residual:.3e}, true "
          f"{st.true_rel_residual:.3e}, error vs x_dd {err:.2e}", flush=True)
    assert st.rel_residual < TOL and err < 1e-6

    # tip deflection on the centre line at x = L against Euler-Bernoulli
    # q L^4 / (8 E I), q = rho g W H, I = W H^3 / 12 (1.402e-3 m); shear and
    # the 3D clamp shift it by a few per cent
    V = gm.vertices
    tip = np.flatnonzero((np.abs(V[:, 0] - 10.0) < 1e-9) & (np.abs(V[:, 1] - 1.0) < 1e-9)
                         & (np.abs(V[:, 2] - 1.0) < 1e-9))
    assert len(tip) == 1
    uz = float(xdd[3 * int(tip[0]) + 2])
    q, L, W, H = 7.85e4 * 2.0 * 2.0, 10.0, 2.0, 2.0
    eb = q * L ** 4 / (8.0 * E * (W * H ** 3 / 12.0))
    print(f"tip u_z {uz:.5e} m, Euler-Bernoulli {-eb:.5e} m", flush=True)
    assert abs(-uz / 1.42e-3 - 1.0) < 0.05
