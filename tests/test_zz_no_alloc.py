"""The allocation-free solve (fh_solve_dev_noalloc) against fh_solve_dev:
the same kernels on the same inputs, so solutions and iteration counts are
bitwise the ordinary entry point's.  Two calls in a row per precision, the
second on a smaller system, so the shared scratch holds the first call's
state when the second starts."""
import numpy as np
import pytest

import fenics_hip as fh

pytestmark = pytest.mark.gpu


def poisson(nx):
    m = fh.unit_square(nx, nx)
    A = fh.assemble_stiffness(m)
    b = fh.assemble_rhs(m, lambda x: 2 * np.pi ** 2 * np.sin(np.pi * x[0]) * np.sin(np.pi * x[1]))
    bd = m.boundary_dofs()
    return fh.apply_dirichlet_bc(A, b, fh.DirichletBC(bd, np.zeros(len(bd))))


def cantilever(nx):
    m = fh.box_beam(10.0, 2.0, 2.0, nx, nx // 5, nx // 5)
    K = fh.assemble_elasticity_stiffness_3d(m, fh.ElasticMaterial(210e9, 0.3))
    b = fh.assemble_elasticity_body_force_3d(m, [0.0, 0.0, -7.85e4])
    return fh.apply_vector_dirichlet_bc(K, b, fh.clamp_dofs_x0(m.vertices))


def bits(t):
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("prec", [fh.FP64, fh.FP64_TRUE, fh.MIXED])
@pytest.mark.parametrize("family", ["poisson", "cantilever"])
def test_noalloc_matches_solve_dev(family, prec):
    import torch
    make, sizes = (poisson, (96, 40)) if family == "poisson" else (cantilever, (40, 20))
    tol = 1e-10
    systems = [make(nx) for nx in sizes]  # larger first, then smaller
    refs = []
    for A, b in systems:
        bt = torch.tensor(b, device="cuda:0")
        x = torch.empty_like(bt)
        st = A.solve_dev(bt, x, tol=tol, max_iter=20000, precision=prec)
        refs.append((bt, x, st))
    for (A, _), (bt, xr, sr) in zip(systems, refs):
        x = torch.full_like(bt, float("nan"))
        st = A.solve_dev_noalloc(bt, x, tol=tol, max_iter=20000, precision=prec)
        assert st.iterations == sr.iterations and st.iterations_low == sr.iterations_low
        assert st.replacements == sr.replacements
        assert np.array_equal(bits(x), bits(xr))
        assert st.true_rel_residual == sr.true_rel_residual
    # a warm start goes through the same path
    A, _ = systems[1]
    bt, xr, _ = refs[1]
    x0 = xr * 0.5
    y1, y2 = torch.empty_like(bt), torch.empty_like(bt)
    s1 = A.solve_dev(bt, y1, x0_t=x0, tol=tol, max_iter=20000, precision=prec)
    s2 = A.solve_dev_noalloc(bt, y2, x0_t=x0, tol=tol, max_iter=20000, precision=prec)
    assert s1.iterations == s2.iterations and np.array_equal(bits(y1), bits(y2))


def test_noalloc_needs_a_prepared_matrix_and_bounds_the_size():
    import torch
    A, b = poisson(24)
    bt = torch.tensor(b, device="cuda:0")
    x = torch.empty_like(bt)
    with pytest.raises(fh.FenicsError) as e:
        A.solve_dev_noalloc(bt, x, tol=1e-10, max_iter=1000, precision=fh.FP64)
    assert e.value.code == fh.FH_ERR_BAD_INPUT and "prepared" in str(e.value)
    # an FP64 solve does not build the FP32 copy a mixed solve reads
    A.solve_dev(bt, x, tol=1e-10, max_iter=1000, precision=fh.FP64)
    with pytest.raises(fh.FenicsError) as e:
        A.solve_dev_noalloc(bt, x, tol=1e-10, max_iter=1000, precision=fh.MIXED)
    assert e.value.code == fh.FH_ERR_BAD_INPUT
    # 1401^2 = 1.96M unknowns, nine vectors in a true-residual solve: beyond
    # the scratch
    B, c = poisson(1400)
    ct = torch.tensor(c, device="cuda:0")
    y = torch.empty_like(ct)
    B.solve_dev(ct, y, tol=1e-10, max_iter=1, precision=fh.FP64_TRUE, allow_fail=True)
    with pytest.raises(fh.FenicsError) as e:
        B.solve_dev_noalloc(ct, y, tol=1e-10, max_iter=1, precision=fh.FP64_TRUE)
    assert e.value.code == fh.FH_ERR_BAD_INPUT and "MiB" in str(e.value)
