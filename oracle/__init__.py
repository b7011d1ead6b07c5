"""CPU FP64 oracle for the fenics-rs assembly + CG hot path.

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
it, and never as the thing measured or shipped.  The HIP product library
does not link or call it.

It wraps ``oracle/_build/liboracle.so`` (built by ``oracle/Makefile`` from
``oracle_*.cpp``), a C++ restatement of the reference Rust code; each C
function cites the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "or_unit_square": (C.c_int, [C.c_uint64, C.c_uint64, f64p, u64p]),
            "or_box_beam": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, C.c_uint64, f64p, u64p]),
            "or_boundary_facets_2d": (C.c_int, [C.c_uint64, u64p, vp, C.POINTER(C.c_uint64)]),
            "or_boundary_facets_3d": (C.c_int, [C.c_uint64, u64p, vp, C.POINTER(C.c_uint64)]),
            "or_boundary_dofs_2d": (C.c_int, [C.c_uint64, u64p, vp, C.POINTER(C.c_uint64)]),
            "or_boundary_dofs_3d": (C.c_int, [C.c_uint64, u64p, vp, C.POINTER(C.c_uint64)]),
            "or_total_volume_2d": (C.c_double, [f64p, C.c_uint64, u64p]),
            "or_total_volume_3d": (C.c_double, [f64p, C.c_uint64, u64p]),
            "or_boundary_measure_2d": (C.c_double, [f64p, C.c_uint64, u64p]),
            "or_boundary_measure_3d": (C.c_double, [f64p, C.c_uint64, u64p]),
            "or_affine_map_2d": (C.c_int, [f64p, f64p, f64p, C.POINTER(C.c_double)]),
            "or_affine_map_3d": (C.c_int, [f64p, f64p, f64p, C.POINTER(C.c_double)]),
            "or_quadrature": (C.c_int, [C.c_int, C.c_int, f64p, f64p, C.POINTER(C.c_int)]),
            "or_p2_tet_basis": (None, [f64p, f64p]),
            "or_p2_tet_grads": (None, [f64p, f64p]),
            "or_csc_from_triplets": (vp, [C.c_uint64, C.c_uint64, u64p, u64p, f64p]),
            "or_csc_from_parts": (vp, [C.c_uint64, u64p, u64p, f64p]),
            "or_csc_n": (C.c_uint64, [vp]),
            "or_csc_nnz": (C.c_uint64, [vp]),
            "or_csc_export": (None, [vp, u64p, u64p, f64p]),
            "or_csc_free": (None, [vp]),
            "or_assemble_stiffness_2d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p, C.c_int]),
            "or_assemble_stiffness_3d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p, C.c_int]),
            "or_assemble_mass_2d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p, C.c_int]),
            "or_assemble_elasticity_3d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p, C.c_double,
                                               C.c_double, C.c_int]),
            "or_assemble_elasticity_3d_cellE": (vp, [C.c_uint64, f64p, C.c_uint64, u64p, f64p,
                                                     C.c_double, C.c_int]),
            "or_assemble_elasticity_mass_3d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p,
                                                    C.c_double, C.c_int]),
            "or_fold_dirichlet": (vp, [vp, C.c_uint64, u64p, C.c_int]),
            "or_vibration_modes": (C.c_int, [vp, vp, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                             C.c_double, C.c_double, C.c_uint64, f64p, f64p,
                                             f64p, vp]),
            "or_quad_points_2d": (C.c_int, [f64p, C.c_uint64, u64p, C.c_int, f64p]),
            "or_quad_points_3d": (C.c_int, [f64p, C.c_uint64, u64p, C.c_int, f64p]),
            "or_assemble_rhs_2d": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_assemble_rhs_3d": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_assemble_body_force_3d": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p, f64p,
                                                    f64p]),
            "or_apply_dirichlet_bc": (vp, [vp, f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_l2_error_2d": (C.c_double, [f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_l2_error_3d": (C.c_double, [f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_spmv": (None, [vp, f64p, f64p]),
            "or_solve_cg": (C.c_int, [vp, f64p, vp, C.c_double, C.c_uint64, f64p,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
            "or_csc_to_csr": (None, [vp, u32p, u32p, f64p, f64p]),
            "or_time_cg_iterations": (C.c_double, [vp, f64p, C.c_uint64]),
            "or_p2_build": (vp, [C.c_uint64, f64p, C.c_uint64, u64p]),
            "or_p2_ndofs": (C.c_uint64, [vp]),
            "or_p2_nedges": (C.c_uint64, [vp]),
            "or_p2_cell_dofs": (None, [vp, u64p]),
            "or_p2_dof_coords": (None, [vp, f64p]),
            "or_p2_boundary_dofs": (C.c_int, [vp, vp, C.POINTER(C.c_uint64)]),
            "or_assemble_p2_stiffness_3d": (vp, [vp, C.c_int]),
            "or_assemble_p2_rhs_cellconst": (C.c_int, [vp, f64p, f64p]),
            "or_p2_free": (None, [vp]),
            "or_cell_bilinear_k0": (C.c_int, [f64p, C.c_uint64, u64p, C.c_double, f64p, f64p,
                                              f64p]),
            "or_element_stresses": (C.c_int, [f64p, C.c_uint64, u64p, f64p, C.c_double,
                                              C.c_double, f64p]),
            "or_vertex_von_mises": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p, f64p, f64p]),
            "or_p2_mass_products": (None, [f64p]),
            "or_assemble_elasticity_mass_p2_3d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p,
                                                       C.c_double, C.c_int]),
            "or_boundary_facets_full": (C.c_int, [C.c_int, f64p, C.c_uint64, u64p, vp, vp, vp,
                                                  C.POINTER(C.c_uint64)]),
            "or_neumann_points_2d": (C.c_int, [f64p, C.c_uint64, u64p, f64p]),
            "or_assemble_neumann_2d": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p, f64p,
                                                 f64p]),
            "or_traction_points_3d": (C.c_int, [f64p, C.c_uint64, u64p, f64p]),
            "or_assemble_surface_traction_3d": (C.c_int, [C.c_uint64, f64p, C.c_uint64, u64p,
                                                          f64p, f64p]),
            "or_assemble_robin_matrix_2d": (vp, [C.c_uint64, f64p, C.c_uint64, u64p,
                                                 C.c_double]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------- meshes
class Mesh:
    """SimplicialMesh (mesh.rs:10-20): vertices [nv, dim] f64, cells [nc, k] u64."""

    def __init__(self, vertices, cells):
        self.vertices = _f64(vertices)
        self.cells = _u64(cells)
        self.dim = self.vertices.shape[1]

    @property
    def nv(self):
        return self.vertices.shape[0]

    @property
    def nc(self):
        return self.cells.shape[0]


def unit_square(nx, ny) -> Mesh:
    v = np.empty(((nx + 1) * (ny + 1), 2))
    c = np.empty((2 * nx * ny, 3), dtype=np.uint64)
    assert lib().or_unit_square(nx, ny, v, c) == 0
    return Mesh(v, c)


def box_beam(L, W, H, nx, ny, nz) -> Mesh:
    v = np.empty(((nx + 1) * (ny + 1) * (nz + 1), 3))
    c = np.empty((6 * nx * ny * nz, 4), dtype=np.uint64)
    assert lib().or_box_beam(L, W, H, nx, ny, nz, v, c) == 0
    return Mesh(v, c)


def unit_cube(nx, ny, nz) -> Mesh:
    return box_beam(1.0, 1.0, 1.0, nx, ny, nz)


def boundary_facets(m: Mesh) -> np.ndarray:
    n = C.c_uint64()
    fn = lib().or_boundary_facets_2d if m.dim == 2 else lib().or_boundary_facets_3d
    fn(m.nc, m.cells, None, C.byref(n))
    out = np.empty((n.value, m.dim), dtype=np.uint64)
    fn(m.nc, m.cells, out.ctypes.data, C.byref(n))
    return out


def boundary_dofs(m: Mesh) -> np.ndarray:
    n = C.c_uint64()
    fn = lib().or_boundary_dofs_2d if m.dim == 2 else lib().or_boundary_dofs_3d
    fn(m.nc, m.cells, None, C.byref(n))
    out = np.empty(n.value, dtype=np.uint64)
    fn(m.nc, m.cells, out.ctypes.data, C.byref(n))
    return out


def total_volume(m: Mesh) -> float:
    fn = lib().or_total_volume_2d if m.dim == 2 else lib().or_total_volume_3d
    return fn(m.vertices, m.nc, m.cells)


def boundary_measure(m: Mesh) -> float:
    fn = lib().or_boundary_measure_2d if m.dim == 2 else lib().or_boundary_measure_3d
    return fn(m.vertices, m.nc, m.cells)


def affine_map(verts):
    verts = _f64(verts)
    d = verts.shape[1]
    J = np.empty(d * d)
    invT = np.empty(d * d)
    det = C.c_double()
    fn = lib().or_affine_map_2d if d == 2 else lib().or_affine_map_3d
    st = fn(verts.reshape(-1), J, invT, C.byref(det))
    return st, J.reshape(d, d), invT.reshape(d, d), det.value


def quadrature(dim, degree):
    pts = np.empty(16)
    w = np.empty(8)
    n = C.c_int()
    lib().or_quadrature(dim, degree, pts, w, C.byref(n))
    return pts[: n.value * dim].reshape(n.value, dim), w[: n.value].copy()


def p2_tet_basis(xi):
    out = np.empty(10)
    lib().or_p2_tet_basis(_f64(xi), out)
    return out


def p2_tet_grads(xi):
    out = np.empty(30)
    lib().or_p2_tet_grads(_f64(xi), out)
    return out.reshape(10, 3)


# ---------------------------------------------------------------- CSC
class Csc:
    """faer SparseColMat<usize, f64> raw parts (col_ptr, row_idx, values)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle returned no matrix (out-of-bounds triplet or bad input)")
        self.h = C.c_void_p(handle)

    def __del__(self):
        try:
            if self.h:
                lib().or_csc_free(self.h)
        except Exception:
            pass

    @property
    def n(self):
        return int(lib().or_csc_n(self.h))

    @property
    def nnz(self):
        return int(lib().or_csc_nnz(self.h))

    def export(self):
        n, nnz = self.n, self.nnz
        cp = np.empty(n + 1, np.uint64)
        ri = np.empty(nnz, np.uint64)
        va = np.empty(nnz)
        lib().or_csc_export(self.h, cp, ri, va)
        return cp, ri, va

    def to_scipy(self):
        import scipy.sparse as sp
        cp, ri, va = self.export()
        return sp.csc_matrix((va, ri.astype(np.int64), cp.astype(np.int64)), shape=(self.n, self.n))

    def spmv(self, x):
        y = np.empty(self.n)
        lib().or_spmv(self.h, _f64(x), y)
        return y

    def to_csr(self):
        n, nnz = self.n, self.nnz
        rp = np.empty(n + 1, np.uint32)
        ci = np.empty(nnz, np.uint32)
        va = np.empty(nnz)
        inv = np.empty(n)
        lib().or_csc_to_csr(self.h, rp, ci, va, inv)
        return rp, ci, va, inv


def csc_from_triplets(n, rows, cols, vals) -> Csc:
    rows, cols, vals = _u64(rows), _u64(cols), _f64(vals)
    return Csc(lib().or_csc_from_triplets(n, len(vals), rows, cols, vals))


def csc_from_parts(col_ptr, row_idx, vals) -> Csc:
    cp, ri, va = _u64(col_ptr), _u64(row_idx), _f64(vals)
    return Csc(lib().or_csc_from_parts(len(cp) - 1, cp, ri, va))


def assemble_stiffness(m: Mesh, nthreads=0) -> Csc:
    fn = lib().or_assemble_stiffness_2d if m.dim == 2 else lib().or_assemble_stiffness_3d
    return Csc(fn(m.nv, m.vertices, m.nc, m.cells, nthreads))


def assemble_mass_2d(m: Mesh, nthreads=0) -> Csc:
    return Csc(lib().or_assemble_mass_2d(m.nv, m.vertices, m.nc, m.cells, nthreads))


def assemble_elasticity(m: Mesh, E, nu, nthreads=0) -> Csc:
    if np.ndim(E) == 0:
        return Csc(lib().or_assemble_elasticity_3d(m.nv, m.vertices, m.nc, m.cells, float(E),
                                                   float(nu), nthreads))
    return Csc(lib().or_assemble_elasticity_3d_cellE(m.nv, m.vertices, m.nc, m.cells, _f64(E),
                                                     float(nu), nthreads))


def assemble_elasticity_mass_3d(m: Mesh, density, nthreads=0) -> Csc:
    """mass.rs:86-126."""
    return Csc(lib().or_assemble_elasticity_mass_3d(m.nv, m.vertices, m.nc, m.cells,
                                                    float(density), nthreads))


def fold_dirichlet(A: Csc, fixed, unit_diag: bool) -> Csc:
    """eigen.rs:51-78."""
    return Csc(lib().or_fold_dirichlet(A.h, len(fixed), _u64(fixed), 1 if unit_diag else 0))


def vibration_modes(K: Csc, M: Csc, fixed, num_modes, max_power_iters, tol, inner_tol=1e-9,
                    inner_max_iter=1000, shapes=True):
    """eigen.rs:94-300 with the Jacobi-CG inner solve.  Returns (status,
    freq_hz, omega, eigenvalue, shapes [num_modes, n] or None)."""
    f, om, ev = np.zeros(num_modes), np.zeros(num_modes), np.zeros(num_modes)
    sh = np.zeros((num_modes, K.n)) if shapes else None
    st = lib().or_vibration_modes(K.h, M.h, len(fixed), _u64(fixed), num_modes, max_power_iters,
                                  tol, inner_tol, inner_max_iter, f, om, ev,
                                  sh.ctypes.data if shapes else None)
    return st, f, om, ev, sh


def quad_points(m: Mesh, degree):
    nq = len(quadrature(m.dim, degree)[1])
    xq = np.empty((m.nc, nq, m.dim))
    fn = lib().or_quad_points_2d if m.dim == 2 else lib().or_quad_points_3d
    fn(m.vertices, m.nc, m.cells, degree, xq.reshape(-1))
    return xq


def assemble_rhs(m: Mesh, fq) -> np.ndarray:
    b = np.empty(m.nv)
    fn = lib().or_assemble_rhs_2d if m.dim == 2 else lib().or_assemble_rhs_3d
    fn(m.nv, m.vertices, m.nc, m.cells, _f64(fq).reshape(-1), b)
    return b


def assemble_body_force(m: Mesh, f3) -> np.ndarray:
    b = np.empty(3 * m.nv)
    lib().or_assemble_body_force_3d(m.nv, m.vertices, m.nc, m.cells, _f64(f3), b)
    return b


def apply_dirichlet_bc(A: Csc, rhs, dofs, vals):
    out = np.empty(A.n)
    h = lib().or_apply_dirichlet_bc(A.h, _f64(rhs), len(dofs), _u64(dofs), _f64(vals), out)
    return Csc(h), out


def solve_cg(A: Csc, b, x0=None, tol=1e-10, max_iter=1000):
    """cg.rs:40-149.  Returns (status, x, iterations, absolute residual)."""
    x = np.empty(A.n)
    it = C.c_uint64()
    res = C.c_double()
    x0p = _f64(x0).ctypes.data if x0 is not None else None
    keep = _f64(x0) if x0 is not None else None
    if keep is not None:
        x0p = keep.ctypes.data
    st = lib().or_solve_cg(A.h, _f64(b), x0p, tol, max_iter, x, C.byref(it), C.byref(res))
    return st, x, int(it.value), res.value


def time_cg_iterations(A: Csc, b, iters) -> float:
    return lib().or_time_cg_iterations(A.h, _f64(b), iters)


def l2_error(m: Mesh, u, uq) -> float:
    fn = lib().or_l2_error_2d if m.dim == 2 else lib().or_l2_error_3d
    return fn(m.vertices, m.nc, m.cells, _f64(u), _f64(uq).reshape(-1))


# ---------------------------------------------------------------- P2, adjoint, stress
class P2:
    def __init__(self, m: Mesh):
        self.mesh = m
        self.h = C.c_void_p(lib().or_p2_build(m.nv, m.vertices, m.nc, m.cells))

    def __del__(self):
        try:
            lib().or_p2_free(self.h)
        except Exception:
            pass

    @property
    def ndofs(self):
        return int(lib().or_p2_ndofs(self.h))

    @property
    def nedges(self):
        return int(lib().or_p2_nedges(self.h))

    def cell_dofs(self):
        out = np.empty((self.mesh.nc, 10), np.uint64)
        lib().or_p2_cell_dofs(self.h, out.reshape(-1))
        return out

    def dof_coords(self):
        out = np.empty((self.ndofs, 3))
        lib().or_p2_dof_coords(self.h, out.reshape(-1))
        return out

    def boundary_dofs(self):
        n = C.c_uint64()
        lib().or_p2_boundary_dofs(self.h, None, C.byref(n))
        out = np.empty(n.value, np.uint64)
        lib().or_p2_boundary_dofs(self.h, out.ctypes.data, C.byref(n))
        return out

    def stiffness(self, nthreads=0) -> Csc:
        return Csc(lib().or_assemble_p2_stiffness_3d(self.h, nthreads))

    def rhs_cellconst(self, f_cell):
        b = np.empty(self.ndofs)
        lib().or_assemble_p2_rhs_cellconst(self.h, _f64(f_cell), b)
        return b


def cell_bilinear_k0(m: Mesh, nu, w, u):
    out = np.empty(m.nc)
    lib().or_cell_bilinear_k0(m.vertices, m.nc, m.cells, float(nu), _f64(w), _f64(u), out)
    return out


def element_stresses(m: Mesh, u, E, nu):
    out = np.empty((m.nc, 20))
    lib().or_element_stresses(m.vertices, m.nc, m.cells, _f64(u), float(E), float(nu),
                              out.reshape(-1))
    return out


def vertex_von_mises(m: Mesh, vm_cell):
    out = np.empty(m.nv)
    lib().or_vertex_von_mises(m.nv, m.vertices, m.nc, m.cells, _f64(vm_cell), out)
    return out


def clamp_dofs_x0(m: Mesh, tol=1e-6):
    """elasticity_tests.rs:94-101: all 3 dofs of vertices with x < 1e-6."""
    idx = np.nonzero(m.vertices[:, 0] < tol)[0].astype(np.uint64)
    return np.sort(np.concatenate([3 * idx, 3 * idx + 1, 3 * idx + 2])).astype(np.uint64)


# ------------------------------------------------------- transient heat
BACKWARD_EULER, CRANK_NICOLSON = 0, 1


def add_scaled_matrices(a: Csc, b: Csc, alpha) -> Csc:
    """transient.rs:72-97: column by column, the triplets of a and then those
    of alpha * b; try_new_from_triplets sums the duplicates in that order."""
    ca, ra, va = a.export()
    cb, rb, vb = b.export()
    n = a.n
    cols = np.concatenate([np.repeat(np.arange(n, dtype=np.uint64), np.diff(ca).astype(np.int64)),
                           np.repeat(np.arange(n, dtype=np.uint64), np.diff(cb).astype(np.int64))])
    rows = np.concatenate([ra, rb])
    vals = np.concatenate([va, alpha * vb])
    order = np.argsort(cols, kind="stable")  # a's entries of a column precede b's
    return csc_from_triplets(n, rows[order], cols[order], vals[order])


class DenseLU:
    """direct.rs:26-53 solve_direct: the sparse matrix densified and factored
    by partial-pivot LU (LAPACK getrf / getrs).  The reference refactors the
    unchanged step matrix every step; factoring it once gives the same
    factors."""

    def __init__(self, A: Csc):
        import scipy.linalg as sl
        self.lu = sl.lu_factor(A.to_scipy().toarray(), check_finite=False)

    def solve(self, b):
        import scipy.linalg as sl
        return sl.lu_solve(self.lu, b, check_finite=False)


class TransientStepper:
    """transient.rs:20-69 TransientHeatSolver for fixed (M, K): BE solves
    (M + dt K) u' = M u + dt f', CN (M + dt/2 K) u' = (M - dt/2 K) u + dt/2 (f + f')."""

    def __init__(self, M: Csc, K: Csc, scheme, dt):
        self.cn = scheme == CRANK_NICOLSON
        self.h = 0.5 * dt if self.cn else dt
        self.lhs = add_scaled_matrices(M, K, self.h)
        self.rmat = add_scaled_matrices(M, K, -self.h) if self.cn else M
        self.lu = DenseLU(self.lhs)

    def rhs(self, u_n, f_n, f_next):
        y = self.rmat.spmv(_f64(u_n))
        if self.cn:
            return y + self.h * (_f64(f_n) + _f64(f_next))
        return y + self.h * _f64(f_next)

    def step(self, u_n, f_n, f_next):
        return self.lu.solve(self.rhs(u_n, f_n, f_next))


def transient_step(M: Csc, K: Csc, scheme, dt, u_n, f_n, f_next):
    """transient.rs:32-68."""
    return TransientStepper(M, K, scheme, dt).step(u_n, f_n, f_next)


def transient_heat_2d(m: Mesh, diffusivity, u_initial, bcs, dt, num_steps, scheme,
                      heat_source=None):
    """heat_solver.rs:23-108.  bcs: sequence of (dofs, values); heat_source(t,
    (x, y)) or None (zero).  Returns (time_steps, history [(steps+1), nv])."""
    n = m.nv
    M = assemble_mass_2d(m)
    cp, ri, va = assemble_stiffness(m).export()
    K = csc_from_parts(cp, ri, diffusivity * va)  # heat_solver.rs:49-58
    stepper = TransientStepper(M, K, scheme, dt)
    u = _f64(u_initial).copy()
    hist = [u.copy()]
    times = [0.0]
    fc, fn = np.zeros(n), np.zeros(n)
    for step in range(1, num_steps + 1):
        t_curr, t_next = (step - 1) * dt, step * dt
        if heat_source is not None:
            for i in range(n):
                pt = (m.vertices[i, 0], m.vertices[i, 1])
                fc[i] = heat_source(t_curr, pt)
                fn[i] = heat_source(t_next, pt)
        for dofs, vals in bcs:
            for d, v in zip(dofs, vals):
                u[d] = v
        u = stepper.step(u, fc, fn)
        for dofs, vals in bcs:
            for d, v in zip(dofs, vals):
                u[d] = v
        times.append(t_next)
        hist.append(u.copy())
    return times, np.array(hist)


# ------------------------------------------------- boundary (ds) integrals
def boundary_facets_full(m: Mesh):
    """mesh.rs:79-113 / 145-210: (vertices [nf, dim] (3D: outward order),
    cell [nf], local facet [nf]), sorted by the sorted vertex tuple."""
    n = C.c_uint64()
    lib().or_boundary_facets_full(m.dim, m.vertices, m.nc, m.cells, None, None, None,
                                  C.byref(n))
    v = np.empty((n.value, m.dim), np.uint64)
    c = np.empty(n.value, np.uint64)
    l = np.empty(n.value, np.uint64)  # noqa: E741
    lib().or_boundary_facets_full(m.dim, m.vertices, m.nc, m.cells, v.ctypes.data, c.ctypes.data,
                                  l.ctypes.data, C.byref(n))
    return v, c, l


def neumann_points_2d(m: Mesh, facets):
    f = _u64(facets).reshape(-1, 2)
    out = np.empty((len(f), 2, 2))
    lib().or_neumann_points_2d(m.vertices, len(f), f, out.reshape(-1))
    return out


def assemble_neumann_2d(m: Mesh, facets, flux_fn):
    """boundary.rs:8-46: flux_fn((x, y)) called per facet, per Gauss point."""
    f = _u64(facets).reshape(-1, 2)
    pts = neumann_points_2d(m, f)
    g = np.array([[flux_fn((p[0], p[1])) for p in fp] for fp in pts], dtype=np.float64)
    out = np.empty(m.nv)
    lib().or_assemble_neumann_2d(m.nv, m.vertices, len(f), f, g.reshape(-1), out)
    return out


def traction_points_3d(m: Mesh, facets):
    f = _u64(facets).reshape(-1, 3)
    out = np.empty((len(f), 3, 3))
    lib().or_traction_points_3d(m.vertices, len(f), f, out.reshape(-1))
    return out


def assemble_surface_traction_3d(m: Mesh, facets, traction_fn):
    """boundary.rs:49-108: traction_fn((x, y, z)) -> (tx, ty, tz) per facet, per midpoint."""
    f = _u64(facets).reshape(-1, 3)
    mids = traction_points_3d(m, f)
    t = np.array([[traction_fn(tuple(p)) for p in fm] for fm in mids], dtype=np.float64)
    out = np.empty(3 * m.nv)
    lib().or_assemble_surface_traction_3d(m.nv, m.vertices, len(f), f, t.reshape(-1), out)
    return out


def assemble_robin_matrix_2d(m: Mesh, facets, gamma) -> Csc:
    """boundary.rs:111-139."""
    f = _u64(facets).reshape(-1, 2)
    return Csc(lib().or_assemble_robin_matrix_2d(m.nv, m.vertices, len(f), f, float(gamma)))


def p2_mass_products():
    """mass.rs:190-250: [10, 10] reference products int phi_a phi_b."""
    out = np.empty(100)
    lib().or_p2_mass_products(out)
    return out.reshape(10, 10)


def assemble_elasticity_mass_p2_3d(nodes, cells, density, nthreads=0) -> Csc:
    """mass.rs:136-180."""
    v, c = _f64(nodes), _u64(cells)
    return Csc(lib().or_assemble_elasticity_mass_p2_3d(len(v), v, len(c), c, float(density),
                                                        nthreads))
