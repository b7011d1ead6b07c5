"""GPU checks of the paths beyond the reference: MFMA P2 element matrices,
compliance adjoint (torch autograd op), stress recovery and the batched
multi-RHS solver.

Tolerances: FP64 kernels that mirror the oracle are bitwise; the FP32 MFMA
element matrices agree with the FP64 VALU ones to 2e-6 of the largest entry
(FP32 rounding of the geometry and of 6-term sums); adjoint gradients agree
with central finite differences to 1e-6 relative (north_star) and satisfy
sum_c E_c dJ/dE_c = -J to 1e-9; batched FP64 columns agree with the oracle's
CG to 1e-8 relative, mixed-precision ones to 1e-7."""
import numpy as np
import pytest

import fenics_hip as fh
import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_p2_mfma_element_matrices_match_fp64():
    gm = fh.unit_cube(6, 5, 4)
    A64 = fh.assemble_p2_stiffness(gm, 0)
    A32 = fh.assemble_p2_stiffness(gm, 1)
    c1, r1, v1 = A64.export_csc()
    c2, r2, v2 = A32.export_csc()
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    assert np.abs(v1 - v2).max() < 2e-6 * np.abs(v1).max()
    # constants stay in the kernel of the FP32 matrix up to FP32 rounding
    nd, _ = gm.p2_info()
    assert np.abs(A32.spmv(np.ones(nd))).max() < 1e-5 * np.abs(v1).max()


def test_p2_poisson_solve_mixed_with_mfma_matrix():
    gm = fh.unit_cube(8, 8, 8)
    nd, _ = gm.p2_info()
    fc = np.random.default_rng(11).random(gm.nc)
    b = fh.assemble_p2_rhs_cellconst(gm, fc)
    bd = gm.p2_boundary_dofs()
    K = fh.assemble_p2_stiffness(gm, 2)  # FP64 VALU values + MFMA FP32 copy
    Kb, bb = fh.apply_dirichlet_bc(K, b, fh.DirichletBC(bd, np.zeros(len(bd))))
    x64, s64 = Kb.solve(bb, tol=1e-10, max_iter=20000, precision=fh.FP64)
    xmx, smx = Kb.solve(bb, tol=1e-10, max_iter=40000, precision=fh.MIXED)
    assert s64.true_rel_residual < 1.5e-10 and smx.true_rel_residual < 1e-10
    assert np.linalg.norm(xmx - x64) / np.linalg.norm(x64) < 1e-7
    # against the oracle's FP64 P2 system
    om = O.unit_cube(8, 8, 8)
    p2 = O.P2(om)
    Ko, bo = O.apply_dirichlet_bc(p2.stiffness(), p2.rhs_cellconst(fc), p2.boundary_dofs(),
                                  np.zeros(len(bd)))
    st, xo, it, res = O.solve_cg(Ko, bo, None, 1e-10, 20000)
    assert np.linalg.norm(x64 - xo) / np.linalg.norm(xo) < 1e-8


def test_stresses_bitwise():
    dims = (1.0, 0.1, 0.1, 8, 2, 2)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    u = np.random.default_rng(3).standard_normal(3 * om.nv) * 1e-3
    sg = fh.element_stresses(gm, u, fh.ElasticMaterial(2e11, 0.3))
    so = O.element_stresses(om, u, 2e11, 0.3)
    assert np.array_equal(bits(sg), bits(so))
    assert np.array_equal(bits(fh.vertex_von_mises(gm, sg[:, 18])),
                          bits(O.vertex_von_mises(om, so[:, 18])))


def test_cell_bilinear_bitwise():
    dims = (1.0, 0.2, 0.2, 6, 2, 2)
    gm, om = fh.box_beam(*dims), O.box_beam(*dims)
    rng = np.random.default_rng(4)
    w, u = rng.standard_normal(3 * om.nv), rng.standard_normal(3 * om.nv)
    assert np.array_equal(bits(fh.cell_bilinear_k0(gm, 0.3, w, u)),
                          bits(O.cell_bilinear_k0(om, 0.3, w, u)))


def test_compliance_adjoint_torch_vs_fd():
    torch = pytest.importorskip("torch")
    from fenics_hip.torch_ops import ElasticityProblem
    m = fh.box_beam(10.0, 2.0, 2.0, 20, 4, 4)
    rng = np.random.default_rng(12)
    E = torch.tensor(210e9 * np.exp(0.3 * rng.standard_normal(m.nc)), dtype=torch.float64,
                     device="cuda:0", requires_grad=True)
    prob = ElasticityProblem(m, 0.3, [0, 0, -7.85e4], fh.clamp_dofs_x0(m.vertices), tol=1e-13)
    J = prob.compliance(E)
    J.backward()
    g = E.grad.detach().clone()
    # homogeneity of degree -1: sum_c E_c dJ/dE_c = -J
    assert abs(float((E.detach() * g).sum()) + float(J)) < 1e-9 * abs(float(J))
    # central differences h = 1e-6 E_c, cancellation-free:
    # J(E+) - J(E-) = -h (u+ + u-)_c^T K0_c u_c  (K' u' = K u = f)
    with torch.no_grad():
        u, _ = prob.solve(E.detach())
        for c in rng.choice(m.nc, 8, replace=False):
            h = 1e-6 * float(E[c])
            Ep, Em = E.detach().clone(), E.detach().clone()
            Ep[c] += h
            Em[c] -= h
            up, _ = prob.solve(Ep)
            um, _ = prob.solve(Em)
            fd = float(prob.cell_gradient(up + um, u)[c]) / 2.0
            assert abs(fd - float(g[c])) < 1e-6 * abs(float(g[c])), (c, fd, float(g[c]))
    # and against the oracle's FP64 adjoint on the same displacement
    om = O.box_beam(10.0, 2.0, 2.0, 20, 4, 4)
    go = O.cell_bilinear_k0(om, 0.3, u.cpu().numpy(), u.cpu().numpy())
    assert np.allclose(g.cpu().numpy(), go, rtol=1e-12, atol=0)


def oracle_batched_column(N, f):
    om = O.unit_square(N, N)
    K = O.assemble_stiffness(om)
    M = O.assemble_mass_2d(om)
    b = M.spmv(f)
    bd = O.boundary_dofs(om)
    Kb, bb = O.apply_dirichlet_bc(K, b, bd, np.zeros(len(bd)))
    st, x, it, res = O.solve_cg(Kb, bb, None, 1e-10, 20000)
    return x, it


@pytest.mark.parametrize("prec", [fh.FP64, fh.MIXED])
def test_batched_multi_rhs_vs_oracle(prec):
    rng = np.random.default_rng(21)
    sizes = [3, 8, 17, 30, 41]
    fs = [rng.uniform(-1, 1, ((N + 1) ** 2, 4)) for N in sizes]
    us, its, st = fh.batched_poisson_block_cg(sizes, fs, tol=1e-10, precision=prec)
    for i, N in enumerate(sizes):
        for k in range(4):
            xo, ito = oracle_batched_column(N, fs[i][:, k])
            rel = np.linalg.norm(us[i][:, k] - xo) / max(np.linalg.norm(xo), 1e-300)
            assert rel < (1e-8 if prec == fh.FP64 else 1e-7), (N, k, rel)
            if prec == fh.FP64:
                assert abs(int(its[i, k]) - ito) <= 2


def test_batched_single_rhs_block_equals_cg():
    """A batch of one instance reproduces the single-system solver column by column."""
    N = 24
    rng = np.random.default_rng(5)
    f = rng.uniform(-1, 1, ((N + 1) ** 2, 4))
    us, its, st = fh.batched_poisson_block_cg([N], [f], tol=1e-12, precision=fh.FP64)
    m = fh.unit_square(N, N)
    K = fh.assemble_stiffness(m)
    M = fh.assemble_mass(m)
    bd = m.boundary_dofs()
    for k in range(4):
        Kb, bb = fh.apply_dirichlet_bc(K, M.spmv(f[:, k]), fh.DirichletBC(bd, np.zeros(len(bd))))
        x, s1 = Kb.solve(bb, tol=1e-12, max_iter=5000)
        assert np.linalg.norm(us[0][:, k] - x) / np.linalg.norm(x) < 1e-10
        assert abs(int(its[0, k]) - int(s1.iterations)) <= 1
