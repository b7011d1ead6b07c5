"""fenics_hip — Python mirror of fenics-rs's assemble/solve API over the
libfenics_hip C ABI (include/fenics_hip.h), running on MI355X (gfx950).

Names, argument meaning and error behaviour follow the reference crates:

* meshes: ``unit_square``, ``box_beam``, ``unit_cube``
  (crates/fenics-mesh/src/generators.rs:24-117) built directly on the GPU;
* assembly: ``assemble_stiffness_2d/3d``, ``assemble_rhs_2d/3d``,
  ``apply_dirichlet_bc``, ``boundary_dofs_2d/3d`` (fenics-assembly/src/poisson.rs),
  ``assemble_elasticity_stiffness_3d``, ``assemble_elasticity_body_force_3d``,
  ``apply_vector_dirichlet_bc`` (fenics-assembly/src/elasticity.rs);
* solvers: ``solve_cg`` (fenics-solver/src/cg.rs:40), ``solve_cg_gpu_or_cpu``,
  ``solve_csr_cg_gpu``, ``solve_bsr3_pcg_gpu``, ``GpuBsr3Context``,
  ``spmv_bsr3_multi_gpu``, ``discover_gpus``, ``best_gpu_index``
  (fenics-solver/src/gpu.rs) and the drivers ``solve_poisson_2d/3d``,
  ``solve_elasticity_3d``.

Errors that the reference returns as ``Err(String)`` raise :class:`FenicsError`
(non-convergence and CG breakdown included).  There is no CPU fallback: when
the shared library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Callable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FENICS_HIP_LIB selects another build of the library (e.g. an experimental one)
LIB_PATH = os.environ.get("FENICS_HIP_LIB") or os.path.join(os.path.dirname(_HERE),
                                                           "libfenics_hip.so")

FH_OK = 0
FH_ERR_NO_DEVICE = 1
FH_ERR_NOT_CONVERGED = 2
FH_ERR_BREAKDOWN = 3
FH_ERR_BAD_INPUT = 4
FH_ERR_HIP = 5
FH_ERR_UNSUPPORTED = 6

FP64, MIXED, MIXED_IR, FP64_TRUE = 0, 1, 2, 3
ORDER_FAST, ORDER_ORDERED = 0, 1


class FenicsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class DeviceInfo(C.Structure):
    _fields_ = [("index", C.c_int), ("name", C.c_char * 128), ("arch", C.c_char * 64),
                ("total_memory_mib", C.c_uint64), ("free_memory_mib", C.c_uint64),
                ("compute_units", C.c_int), ("clock_mhz", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("iterations_low", C.c_uint64),
                ("replacements", C.c_uint64),
                ("final_residual", C.c_double), ("rel_residual", C.c_double),
                ("true_rel_residual", C.c_double), ("residual_floor", C.c_double),
                ("wall_ms", C.c_double),
                ("kernel_ms", C.c_double), ("gflops", C.c_double),
                ("effective_bw_gbps", C.c_double), ("spmv_bytes", C.c_double),
                ("iter_bytes", C.c_double), ("spmv_avg_ms", C.c_double),
                ("spmv_timed", C.c_uint64), ("spmv32_bytes", C.c_double),
                ("spmv32_avg_ms", C.c_double), ("spmv32_timed", C.c_uint64),
                ("device", C.c_int), ("precision", C.c_int),
                ("status", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RefineStats(C.Structure):
    _fields_ = [("outer_steps", C.c_uint64), ("cg_iterations", C.c_uint64),
                ("rel_residual_pair", C.c_double), ("rel_residual_rounded", C.c_double)]


_lib = None
vp = C.c_void_p
u64 = C.c_uint64
dbl = C.c_double
PP = C.POINTER(C.c_void_p)


def lib():
    """Load libfenics_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FenicsError(FH_ERR_UNSUPPORTED,
                          f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    # PyTorch wheels bundle their own HIP runtime under the same soname.  When
    # torch is importable, load it first so that the library binds to that one
    # runtime; two HIP runtimes in one process make the second see no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    sig = {
        "fh_last_error": (C.c_char_p, []),
        "fh_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "fh_get_device_info": (C.c_int, [C.c_int, C.POINTER(DeviceInfo)]),
        "fh_best_device": (C.c_int, [C.POINTER(C.c_int)]),
        "fh_synchronize": (C.c_int, [C.c_int]),
        "fh_stream_bandwidth": (C.c_int, [C.c_int, u64, C.c_int, C.c_int, C.POINTER(dbl),
                                          C.POINTER(dbl)]),
        "fh_mesh_create": (C.c_int, [C.c_int, u64, vp, u64, vp, C.c_int, PP]),
        "fh_mesh_unit_square": (C.c_int, [u64, u64, C.c_int, PP]),
        "fh_mesh_box_beam": (C.c_int, [dbl, dbl, dbl, u64, u64, u64, C.c_int, PP]),
        "fh_mesh_sizes": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(u64)]),
        "fh_mesh_download": (C.c_int, [vp, vp, vp]),
        "fh_mesh_boundary_dofs": (C.c_int, [vp, vp, C.POINTER(u64)]),
        "fh_quad_points": (C.c_int, [vp, C.c_int, vp]),
        "fh_mesh_destroy": (C.c_int, [vp]),
        "fh_assemble_stiffness": (C.c_int, [vp, PP]),
        "fh_assemble_mass": (C.c_int, [vp, PP]),
        "fh_assemble_elasticity": (C.c_int, [vp, dbl, dbl, vp, PP]),
        "fh_assemble_elasticity_dev": (C.c_int, [vp, dbl, dbl, vp, PP]),
        "fh_assemble_elasticity_method": (C.c_int, [vp, dbl, dbl, vp, C.c_int, PP]),
        "fh_assemble_p2_stiffness": (C.c_int, [vp, C.c_int, PP]),
        "fh_p2_info": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "fh_p2_cell_dofs": (C.c_int, [vp, vp]),
        "fh_p2_boundary_dofs": (C.c_int, [vp, vp, C.POINTER(u64)]),
        "fh_assemble_rhs": (C.c_int, [vp, vp, vp]),
        "fh_assemble_p2_rhs_cellconst": (C.c_int, [vp, vp, vp]),
        "fh_assemble_body_force": (C.c_int, [vp, vp, vp]),
        "fh_assemble_body_force_dev": (C.c_int, [vp, vp, vp]),
        "fh_apply_dirichlet": (C.c_int, [vp, vp, u64, vp, vp, PP, vp]),
        "fh_apply_dirichlet_dev": (C.c_int, [vp, vp, u64, vp, vp, PP, vp]),
        "fh_matrix_info": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(C.c_int)]),
        "fh_matrix_skipped_cells": (C.c_int, [vp, C.POINTER(u64)]),
        "fh_matrix_export_csc": (C.c_int, [vp, vp, vp, vp]),
        "fh_matrix_export_bsr3": (C.c_int, [vp, vp, vp, vp, vp]),
        "fh_matrix_spmv": (C.c_int, [vp, vp, vp, C.c_int]),
        "fh_matrix_spmv_dev": (C.c_int, [vp, vp, vp, C.c_int]),
        "fh_matrix_destroy": (C.c_int, [vp]),
        "fh_solve": (C.c_int, [vp, vp, vp, dbl, u64, C.c_int, vp, C.POINTER(SolveStats)]),
        "fh_solve_dev": (C.c_int, [vp, vp, vp, dbl, u64, C.c_int, vp, C.POINTER(SolveStats)]),
        "fh_true_residual_dev": (C.c_int, [vp, vp, vp, vp, C.POINTER(dbl)]),
        "fh_solve_dev_noalloc": (C.c_int, [vp, vp, vp, dbl, u64, C.c_int, vp,
                                           C.POINTER(SolveStats)]),
        "fh_solve_refined_dev": (C.c_int, [vp, vp, vp, dbl, u64, dbl, u64, vp, vp,
                                           C.POINTER(RefineStats)]),
        "fh_solve_csc_cg": (C.c_int, [u64, vp, vp, vp, vp, vp, dbl, u64, C.c_int, vp,
                                      C.POINTER(SolveStats)]),
        "fh_solve_csr_cg": (C.c_int, [u64, vp, vp, vp, vp, vp, vp, dbl, u64, C.c_int, vp,
                                      C.POINTER(SolveStats)]),
        "fh_solve_bsr3_pcg": (C.c_int, [u64, vp, vp, vp, vp, vp, vp, dbl, u64, C.c_int, vp,
                                        C.POINTER(SolveStats)]),
        "fh_bsr3_ctx_create": (C.c_int, [u64, vp, vp, vp, vp, C.c_int, PP]),
        "fh_bsr3_ctx_spmv": (C.c_int, [vp, vp, vp]),
        "fh_bsr3_ctx_jacobi_smooth": (C.c_int, [vp, vp, vp, u64, dbl, C.c_int]),
        "fh_bsr3_ctx_device_name": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
        "fh_bsr3_ctx_nnz_blocks": (C.c_int, [vp, C.POINTER(u64)]),
        "fh_bsr3_ctx_destroy": (C.c_int, [vp]),
        "fh_spmv_bsr3_multi": (C.c_int, [u64, vp, vp, vp, vp, u64, vp, C.POINTER(C.c_int)]),
        "fh_solve_poisson": (C.c_int, [vp, vp, vp, dbl, u64, C.c_int, vp, C.POINTER(SolveStats)]),
        "fh_solve_elasticity_3d": (C.c_int, [vp, dbl, dbl, vp, vp, u64, vp, vp, dbl, u64,
                                             C.c_int, vp, vp, C.POINTER(SolveStats)]),
        "fh_cell_bilinear_k0_dev": (C.c_int, [vp, dbl, vp, vp, vp]),
        "fh_cell_bilinear_k0": (C.c_int, [vp, dbl, vp, vp, vp]),
        "fh_element_stresses": (C.c_int, [vp, vp, dbl, dbl, vp]),
        "fh_vertex_von_mises": (C.c_int, [vp, vp, vp]),
        "fh_batched_poisson_block_cg": (C.c_int, [u64, vp, u64, vp, dbl, u64, C.c_int, C.c_int,
                                                  vp, vp, C.POINTER(SolveStats)]),
        "fh_assemble_elasticity_mass": (C.c_int, [vp, dbl, PP]),
        "fh_fold_dirichlet": (C.c_int, [vp, u64, vp, C.c_int, PP]),
        "fh_assemble_elasticity_mass_p2": (C.c_int, [u64, vp, u64, vp, dbl, C.c_int, PP]),
        "fh_vibration_modes": (C.c_int, [vp, vp, u64, vp, u64, u64, dbl, dbl, u64, C.c_int,
                                         vp, vp, vp, vp, C.POINTER(u64)]),
        "fh_vibration_modes_with_inner": (C.c_int, [vp, vp, u64, vp, u64, u64, dbl,
                                                    InnerSolveFn, vp, vp, vp, vp, vp]),
        "fh_matrix_from_csc": (C.c_int, [u64, vp, vp, vp, C.c_int, PP]),
        "fh_l2_error": (C.c_int, [vp, vp, vp, C.POINTER(dbl)]),
        "fh_mesh_boundary_facets": (C.c_int, [vp, vp, vp, vp, C.POINTER(u64)]),
        "fh_boundary_points": (C.c_int, [vp, u64, vp, vp]),
        "fh_assemble_neumann_2d": (C.c_int, [vp, u64, vp, vp, vp]),
        "fh_assemble_surface_traction_3d": (C.c_int, [vp, u64, vp, vp, vp]),
        "fh_assemble_robin_matrix_2d": (C.c_int, [vp, u64, vp, dbl, PP]),
        "fh_add_scaled_matrices": (C.c_int, [vp, vp, dbl, PP]),
        "fh_transient_heat_step": (C.c_int, [vp, vp, C.c_int, dbl, vp, vp, vp, dbl, u64,
                                             C.c_int, vp, C.POINTER(SolveStats)]),
        "fh_solve_transient_heat_2d": (C.c_int, [vp, dbl, vp, u64, vp, vp, dbl, u64, C.c_int,
                                                 HeatSourceFn, vp, dbl, u64, C.c_int, vp, vp,
                                                 C.POINTER(TransientStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return lib().fh_last_error().decode(errors="replace")


def _check(code: int, allow=()):
    if code != FH_OK and code not in allow:
        raise FenicsError(code, last_error())
    return code


def _dptr(t):
    """Device pointer of a contiguous torch float64 tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous() or str(t.dtype) != "torch.float64" or not t.is_cuda:
        raise FenicsError(FH_ERR_BAD_INPUT, "expected a contiguous float64 device tensor")
    return C.c_void_p(t.data_ptr())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ------------------------------------------------------------------ devices
@dataclasses.dataclass
class GpuDeviceInfo:
    """gpu.rs:12-21 GpuDeviceInfo."""
    index: int
    name: str
    arch: str
    total_memory_mib: int
    free_memory_mib: int
    multiprocessor_count: int


def discover_gpus():
    """gpu.rs:178-234.  Empty list when no HIP device is visible."""
    n = C.c_int(0)
    lib().fh_device_count(C.byref(n))
    out = []
    for i in range(n.value):
        info = DeviceInfo()
        if lib().fh_get_device_info(i, C.byref(info)) == FH_OK:
            out.append(GpuDeviceInfo(i, info.name.decode(), info.arch.decode(),
                                     info.total_memory_mib, info.free_memory_mib,
                                     info.compute_units))
    return out


def best_gpu_index() -> Optional[int]:
    """gpu.rs:237-242: the device with the most free memory."""
    i = C.c_int(-1)
    return i.value if lib().fh_best_device(C.byref(i)) == FH_OK else None


# ------------------------------------------------------------------ meshes
class Mesh:
    """Device-resident SimplicialMesh (mesh.rs:10-20)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)
        d, nv, nc = C.c_int(), u64(), u64()
        _check(lib().fh_mesh_sizes(self.h, C.byref(d), C.byref(nv), C.byref(nc)))
        self.dim, self.nv, self.nc = d.value, int(nv.value), int(nc.value)
        self._cache = {}

    def __del__(self):
        try:
            if self.h:
                lib().fh_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @staticmethod
    def from_arrays(vertices, cells, device=-1) -> "Mesh":
        v, c = _f64(vertices), _u64(cells)
        h = C.c_void_p()
        _check(lib().fh_mesh_create(v.shape[1], v.shape[0], _p(v), c.shape[0], _p(c), device,
                                    C.byref(h)))
        return Mesh(h.value)

    def num_vertices(self):
        return self.nv

    def num_cells(self):
        return self.nc

    @property
    def vertices(self) -> np.ndarray:
        if "v" not in self._cache:
            self._download()
        return self._cache["v"]

    @property
    def cells(self) -> np.ndarray:
        if "c" not in self._cache:
            self._download()
        return self._cache["c"]

    def _download(self):
        v = np.empty((self.nv, self.dim))
        c = np.empty((self.nc, self.dim + 1), np.uint64)
        _check(lib().fh_mesh_download(self.h, _p(v), _p(c)))
        self._cache["v"], self._cache["c"] = v, c

    def boundary_dofs(self) -> np.ndarray:
        """poisson.rs:266-287 boundary_dofs_2d/3d (sorted)."""
        n = u64(0)
        _check(lib().fh_mesh_boundary_dofs(self.h, None, C.byref(n)))
        out = np.empty(n.value, np.uint64)
        _check(lib().fh_mesh_boundary_dofs(self.h, _p(out), C.byref(n)))
        return out

    def quad_points(self, degree: int = 2) -> np.ndarray:
        nq = {(2, 1): 1, (2, 2): 3, (2, 3): 4, (3, 1): 1, (3, 2): 4, (3, 3): 5}[(self.dim, min(degree, 3))]
        out = np.empty((self.nc, nq, self.dim))
        _check(lib().fh_quad_points(self.h, degree, _p(out)))
        return out

    # P2 tetrahedral dofs (vertices, then edges by the number of cells around
    # them, most first, then by max - min, then by min)
    def p2_info(self):
        nd, ne = u64(), u64()
        _check(lib().fh_p2_info(self.h, C.byref(nd), C.byref(ne)))
        return int(nd.value), int(ne.value)

    def p2_cell_dofs(self):
        out = np.empty((self.nc, 10), np.uint64)
        _check(lib().fh_p2_cell_dofs(self.h, _p(out)))
        return out

    def p2_boundary_dofs(self):
        n = u64(0)
        _check(lib().fh_p2_boundary_dofs(self.h, None, C.byref(n)))
        out = np.empty(n.value, np.uint64)
        _check(lib().fh_p2_boundary_dofs(self.h, _p(out), C.byref(n)))
        return out


def unit_square(nx: int, ny: int, device: int = -1) -> Mesh:
    h = C.c_void_p()
    _check(lib().fh_mesh_unit_square(nx, ny, device, C.byref(h)))
    return Mesh(h.value)


def box_beam(length, width, height, nx, ny, nz, device: int = -1) -> Mesh:
    h = C.c_void_p()
    _check(lib().fh_mesh_box_beam(float(length), float(width), float(height), nx, ny, nz, device,
                                  C.byref(h)))
    return Mesh(h.value)


def unit_cube(nx, ny, nz, device: int = -1) -> Mesh:
    return box_beam(1.0, 1.0, 1.0, nx, ny, nz, device)


def boundary_dofs_2d(mesh: Mesh) -> np.ndarray:
    return mesh.boundary_dofs()


boundary_dofs_3d = boundary_dofs_2d


# ------------------------------------------------------------------ matrices
class SparseMatrix:
    """Device-resident assembled operator; exports faer's SparseColMat raw parts."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    def __del__(self):
        try:
            if self.h:
                lib().fh_matrix_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _info(self):
        n, nnz, bs = u64(), u64(), C.c_int()
        _check(lib().fh_matrix_info(self.h, C.byref(n), C.byref(nnz), C.byref(bs)))
        return int(n.value), int(nnz.value), bs.value

    @property
    def n(self):
        """Number of scalar rows (without counting the exported nonzeros)."""
        n = u64()
        _check(lib().fh_matrix_info(self.h, C.byref(n), None, None))
        return int(n.value)

    def nrows(self):
        return self.n

    ncols = nrows

    @property
    def block(self):
        return self._info()[2]

    def nnz(self):
        return self._info()[1]

    def skipped_cells(self) -> int:
        """Cells the assembly skipped (|det J| < 1e-14, jacobian.rs:29-31)."""
        c = C.c_uint64()
        _check(lib().fh_matrix_skipped_cells(self.h, C.byref(c)))
        return int(c.value)

    def export_csc(self):
        """(col_ptr u64[n+1], row_idx u64[nnz], values f64[nnz]) — faer's canonical CSC."""
        n = self.nrows()
        cp = np.empty(n + 1, np.uint64)
        _check(lib().fh_matrix_export_csc(self.h, _p(cp), None, None))
        nnz = int(cp[-1])
        ri = np.empty(nnz, np.uint64)
        va = np.empty(nnz)
        _check(lib().fh_matrix_export_csc(self.h, _p(cp), _p(ri), _p(va)))
        return cp, ri, va

    def export_bsr3(self, with_dinv: bool = True):
        """(n_nodes, row_ptr u32, cols u32, vals_flat f64[9 nnzb], dinv_flat f64[9 n_nodes]):
        the arguments of GpuBsr3Context / solve_bsr3_pcg_gpu (gpu.rs:453-458,
        775-786) for a 3x3-blocked operator; dinv_flat is None without with_dinv."""
        n_nodes = self.nrows() // 3
        rp = np.empty(n_nodes + 1, np.uint32)
        _check(lib().fh_matrix_export_bsr3(self.h, _p(rp), None, None, None))
        nnzb = int(rp[-1])
        cc = np.empty(nnzb, np.uint32)
        va = np.empty(9 * nnzb)
        di = np.empty(9 * n_nodes) if with_dinv else None
        _check(lib().fh_matrix_export_bsr3(self.h, _p(rp), _p(cc), _p(va), _p(di)))
        return n_nodes, rp, cc, va, di

    def to_scipy(self):
        import scipy.sparse as sp
        cp, ri, va = self.export_csc()
        n = len(cp) - 1
        return sp.csc_matrix((va, ri.astype(np.int64), cp.astype(np.int64)), shape=(n, n))

    def spmv(self, x, order: int = ORDER_FAST) -> np.ndarray:
        x = _f64(x)
        y = np.empty_like(x)
        _check(lib().fh_matrix_spmv(self.h, _p(x), _p(y), order))
        return y

    def solve(self, b, x0=None, tol=1e-10, max_iter=1000, precision=FP64, allow_fail=False):
        """Jacobi-PCG (cg.rs:40-149).  Returns (x, SolveStats)."""
        b = _f64(b)
        x = np.empty_like(b)
        x0a = None if x0 is None else _f64(x0)
        st = SolveStats()
        code = lib().fh_solve(self.h, _p(b), _p(x0a), tol, max_iter, precision, _p(x),
                              C.byref(st))
        if not allow_fail:
            _check(code)
        elif code not in (FH_OK, FH_ERR_NOT_CONVERGED, FH_ERR_BREAKDOWN):
            _check(code)
        return x, st

    def solve_dev(self, b_t, x_t, x0_t=None, tol=1e-10, max_iter=1000, precision=FP64,
                  allow_fail=False):
        """fh_solve_dev on torch float64 device tensors (x_t receives the
        solution).  Returns SolveStats."""
        st = SolveStats()
        code = lib().fh_solve_dev(self.h, _dptr(b_t), _dptr(x0_t), tol, max_iter, precision,
                                  _dptr(x_t), C.byref(st))
        if not allow_fail or code not in (FH_OK, FH_ERR_NOT_CONVERGED, FH_ERR_BREAKDOWN):
            _check(code)
        return st

    def solve_dev_noalloc(self, b_t, x_t, x0_t=None, tol=1e-10, max_iter=1000, precision=FP64,
                          allow_fail=False):
        """fh_solve_dev_noalloc: as solve_dev, with every buffer in the
        library's static per-GPU scratch (the matrix must have been solved once
        with solve_dev / solve in the same precision).  Returns SolveStats."""
        st = SolveStats()
        code = lib().fh_solve_dev_noalloc(self.h, _dptr(b_t), _dptr(x0_t), tol, max_iter,
                                          precision, _dptr(x_t), C.byref(st))
        if not allow_fail or code not in (FH_OK, FH_ERR_NOT_CONVERGED, FH_ERR_BREAKDOWN):
            _check(code)
        return st

    def true_residual_dev(self, b_t, xh_t, xl_t=None) -> float:
        """||b - A (xh + xl)|| / ||b||, evaluated in double-double."""
        out = C.c_double()
        _check(lib().fh_true_residual_dev(self.h, _dptr(b_t), _dptr(xh_t), _dptr(xl_t),
                                          C.byref(out)))
        return out.value

    def solve_refined_dev(self, b_t, xh_t, xl_t, x0_t=None, target=1e-14, max_outer=8,
                          inner_tol=1e-6, max_iter=20000, allow_fail=False):
        """Refinement with a double-double solution (xh_t + xl_t).  Returns
        RefineStats; xh_t is the FP64 vector nearest to the solution."""
        st = RefineStats()
        code = lib().fh_solve_refined_dev(self.h, _dptr(b_t), _dptr(x0_t), target, max_outer,
                                          inner_tol, max_iter, _dptr(xh_t), _dptr(xl_t),
                                          C.byref(st))
        if not allow_fail or code != FH_ERR_NOT_CONVERGED:
            _check(code)
        return st


def _new_matrix(fn, *args) -> SparseMatrix:
    h = C.c_void_p()
    _check(fn(*args, C.byref(h)))
    return SparseMatrix(h.value)


def assemble_stiffness(mesh: Mesh) -> SparseMatrix:
    """poisson.rs:23-69 / 124-165."""
    return _new_matrix(lib().fh_assemble_stiffness, mesh.h)


assemble_stiffness_2d = assemble_stiffness
assemble_stiffness_3d = assemble_stiffness


def assemble_mass(mesh: Mesh) -> SparseMatrix:
    """mass.rs:10-82."""
    return _new_matrix(lib().fh_assemble_mass, mesh.h)


def _sample(mesh: Mesh, f, degree=2):
    if callable(f):
        xq = mesh.quad_points(degree)
        vals = np.asarray(f(xq.reshape(-1, mesh.dim).T), dtype=np.float64)
        return _f64(np.broadcast_to(vals, (xq.shape[0] * xq.shape[1],)).reshape(xq.shape[:2]))
    return _f64(f)


def assemble_rhs(mesh: Mesh, f) -> np.ndarray:
    """poisson.rs:72-121 / 168-218.  `f` maps coordinates (dim x N array) to
    values, or is the array of values at mesh.quad_points(2)."""
    fq = _sample(mesh, f)
    b = np.empty(mesh.nv)
    _check(lib().fh_assemble_rhs(mesh.h, _p(fq), _p(b)))
    return b


assemble_rhs_2d = assemble_rhs
assemble_rhs_3d = assemble_rhs


@dataclasses.dataclass
class DirichletBC:
    """poisson.rs:13-20."""
    dofs: Sequence[int]
    values: Sequence[float]


@dataclasses.dataclass
class VectorDirichletBC:
    """elasticity.rs:63-69: list of (dof, value)."""
    prescribed_dofs: Sequence[tuple]


def _apply(mat: SparseMatrix, rhs, dofs, vals):
    rhs = _f64(rhs)
    dofs, vals = _u64(dofs), _f64(vals)
    out = np.empty_like(rhs)
    h = C.c_void_p()
    _check(lib().fh_apply_dirichlet(mat.h, _p(rhs), len(dofs), _p(dofs), _p(vals), C.byref(h),
                                    _p(out)))
    return SparseMatrix(h.value), out


def apply_dirichlet_bc(mat: SparseMatrix, rhs, bc: DirichletBC):
    """poisson.rs:222-263."""
    return _apply(mat, rhs, bc.dofs, bc.values)


def apply_vector_dirichlet_bc(mat: SparseMatrix, rhs, bc: VectorDirichletBC):
    """elasticity.rs:188-227."""
    pd = list(bc.prescribed_dofs)
    dofs = [d for d, _ in pd]
    vals = [v for _, v in pd]
    return _apply(mat, rhs, dofs, vals)


@dataclasses.dataclass
class ElasticMaterial:
    """elasticity.rs:14-48."""
    youngs_modulus: float
    poissons_ratio: float

    def __post_init__(self):
        if not (-1.0 < self.poissons_ratio < 0.5):
            raise ValueError("Poisson's ratio must be in (-1.0, 0.5) for physical stability")

    def lame_lambda(self):
        e, nu = self.youngs_modulus, self.poissons_ratio
        return (e * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))

    def lame_mu(self):
        e, nu = self.youngs_modulus, self.poissons_ratio
        return e / (2.0 * (1.0 + nu))


def assemble_elasticity_stiffness_3d(mesh: Mesh, material: ElasticMaterial,
                                     E_cell=None, method: int = 0) -> SparseMatrix:
    """elasticity.rs:73-149 (E_cell: optional per-cell Young's moduli).

    method 0: FP64 VALU (the reference's values, bitwise); 1: FP32 element
    matrices B^T (vol D) B on the MFMA units, summed per row in FP32; 2: the
    FP64 values of method 0 with the MFMA FP32 values as the FP32 copy."""
    ec = None if E_cell is None else _f64(E_cell)
    if method == 0:
        return _new_matrix(lib().fh_assemble_elasticity, mesh.h, material.youngs_modulus,
                           material.poissons_ratio, _p(ec))
    return _new_matrix(lib().fh_assemble_elasticity_method, mesh.h, material.youngs_modulus,
                       material.poissons_ratio, _p(ec), int(method))


def assemble_elasticity_body_force_3d(mesh: Mesh, f_body) -> np.ndarray:
    """elasticity.rs:153-185."""
    f = _f64(f_body)
    b = np.empty(3 * mesh.nv)
    _check(lib().fh_assemble_body_force(mesh.h, _p(f), _p(b)))
    return b


def assemble_p2_stiffness(mesh: Mesh, method: int = 0) -> SparseMatrix:
    """P2 tetrahedral Poisson stiffness (method 0: FP64 VALU, 1: FP32 MFMA)."""
    return _new_matrix(lib().fh_assemble_p2_stiffness, mesh.h, method)


def assemble_p2_rhs_cellconst(mesh: Mesh, f_cell) -> np.ndarray:
    f = _f64(f_cell)
    nd, _ = mesh.p2_info()
    b = np.empty(nd)
    _check(lib().fh_assemble_p2_rhs_cellconst(mesh.h, _p(f), _p(b)))
    return b


# ------------------------------------------------------------------ solvers
def solve_cg(mat: SparseMatrix, b, x0=None, tol=1e-10, max_iter=1000, precision=FP64):
    """cg.rs:40-149: returns (x, iterations, final absolute residual);
    raises FenicsError on breakdown / non-convergence like the reference's Err."""
    x, st = mat.solve(b, x0, tol, max_iter, precision)
    return x, int(st.iterations), st.final_residual


@dataclasses.dataclass
class GpuSolveResult:
    """gpu.rs:24-34."""
    solution: np.ndarray
    iterations: int
    final_residual: float
    wall_ms: float
    kernel_ms: float
    gflops: float
    effective_bw_gbps: float
    devices_used: list
    stats: SolveStats = None


def _result(x, st: SolveStats):
    infos = discover_gpus()
    name = infos[st.device].name if st.device < len(infos) else f"device {st.device}"
    return GpuSolveResult(x, int(st.iterations), st.rel_residual, st.wall_ms, st.kernel_ms,
                          st.gflops, st.effective_bw_gbps, [name], st)


def solve_cg_gpu_or_cpu(col_ptr, row_idx, values, b, x0=None, tol=1e-10, max_iter=1000,
                        device=-1) -> GpuSolveResult:
    """gpu.rs:1542-1602 on faer's raw CSC parts (no CPU fallback here)."""
    cp, ri, va, b = _u64(col_ptr), _u64(row_idx), _f64(values), _f64(b)
    x = np.empty_like(b)
    x0a = None if x0 is None else _f64(x0)
    st = SolveStats()
    _check(lib().fh_solve_csc_cg(len(cp) - 1, _p(cp), _p(ri), _p(va), _p(b), _p(x0a), tol,
                                 max_iter, device, _p(x), C.byref(st)))
    return _result(x, st)


def solve_csr_cg_gpu(row_ptr, col_idx, vals, inv_diag, b, x0=None, rtol=1e-10, max_iter=1000,
                     device_index=None) -> GpuSolveResult:
    """gpu.rs:1145-1442."""
    rp, ci, va, inv, b = _u32(row_ptr), _u32(col_idx), _f64(vals), _f64(inv_diag), _f64(b)
    x = np.empty_like(b)
    x0a = None if x0 is None else _f64(x0)
    st = SolveStats()
    dev = -1 if device_index is None else device_index
    _check(lib().fh_solve_csr_cg(len(rp) - 1, _p(rp), _p(ci), _p(va), _p(inv), _p(b), _p(x0a),
                                 rtol, max_iter, dev, _p(x), C.byref(st)))
    return _result(x, st)


def solve_bsr3_pcg_gpu(n_nodes, row_ptr, cols, vals_flat, dinv_flat, b, x0=None, rtol=1e-10,
                       max_iter=1000, device_index=None) -> GpuSolveResult:
    """gpu.rs:775-1141 (3x3 block CSR, 3x3 block-Jacobi)."""
    rp, cc, va, di, b = _u32(row_ptr), _u32(cols), _f64(vals_flat), _f64(dinv_flat), _f64(b)
    x = np.empty_like(b)
    x0a = None if x0 is None else _f64(x0)
    st = SolveStats()
    dev = -1 if device_index is None else device_index
    _check(lib().fh_solve_bsr3_pcg(n_nodes, _p(rp), _p(cc), _p(va), _p(di), _p(b), _p(x0a),
                                   rtol, max_iter, dev, _p(x), C.byref(st)))
    return _result(x, st)


class GpuBsr3Context:
    """gpu.rs:432-658: persistent device-resident 3x3 BSR operator."""

    def __init__(self, n_nodes, row_ptr, cols, vals_flat, dinv_flat, device_index=None):
        self._keep = (_u32(row_ptr), _u32(cols), _f64(vals_flat), _f64(dinv_flat))
        rp, cc, va, di = self._keep
        self.n = 3 * n_nodes
        h = C.c_void_p()
        dev = -1 if device_index is None else device_index
        _check(lib().fh_bsr3_ctx_create(n_nodes, _p(rp), _p(cc), _p(va), _p(di), dev, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            lib().fh_bsr3_ctx_destroy(self.h)
        except Exception:
            pass

    def spmv(self, x) -> np.ndarray:
        x = _f64(x)
        y = np.empty(self.n)
        _check(lib().fh_bsr3_ctx_spmv(self.h, _p(x), _p(y)))
        return y

    def jacobi_smooth(self, r, z, sweeps, omega, zero_initial_z) -> np.ndarray:
        r = _f64(r)
        z = np.array(z, dtype=np.float64, copy=True)
        _check(lib().fh_bsr3_ctx_jacobi_smooth(self.h, _p(r), _p(z), sweeps, omega,
                                               int(bool(zero_initial_z))))
        return z

    def device_name(self) -> str:
        buf = C.create_string_buffer(128)
        _check(lib().fh_bsr3_ctx_device_name(self.h, buf, 128))
        return buf.value.decode()

    def nnz_blocks(self) -> int:
        v = u64()
        _check(lib().fh_bsr3_ctx_nnz_blocks(self.h, C.byref(v)))
        return int(v.value)


def spmv_bsr3_multi_gpu(n_nodes, row_ptr, cols, vals_flat, x, max_gpus):
    """gpu.rs:662-771: returns (y, device names used)."""
    rp, cc, va, x = _u32(row_ptr), _u32(cols), _f64(vals_flat), _f64(x)
    y = np.empty(3 * n_nodes)
    used = C.c_int()
    _check(lib().fh_spmv_bsr3_multi(n_nodes, _p(rp), _p(cc), _p(va), _p(x), max_gpus, _p(y),
                                    C.byref(used)))
    names = [g.name for g in discover_gpus()][: used.value]
    return y, names


# ------------------------------------------------------------------ drivers
@dataclasses.dataclass
class PoissonSolution:
    """poisson_solver.rs:16-23."""
    u: np.ndarray
    iterations: int
    residual: float
    stats: SolveStats = None


def _solve_poisson(mesh: Mesh, source, boundary_value, tol, max_iter, precision):
    fq = _sample(mesh, source)
    bnd = mesh.boundary_dofs()
    verts = mesh.vertices
    if boundary_value is None:
        g = np.zeros(len(bnd))
    else:
        g = _f64([boundary_value(verts[int(i)]) for i in bnd])
    u = np.empty(mesh.nv)
    st = SolveStats()
    _check(lib().fh_solve_poisson(mesh.h, _p(fq), _p(g), tol, max_iter, precision, _p(u),
                                  C.byref(st)))
    return PoissonSolution(u, int(st.iterations), st.final_residual, st)


def solve_poisson_2d(mesh, source, boundary_value, tol, max_iter, precision=FP64):
    """poisson_solver.rs:26-59: source(x: dim x N array) -> values,
    boundary_value(vertex) -> value."""
    return _solve_poisson(mesh, source, boundary_value, tol, max_iter, precision)


solve_poisson_3d = solve_poisson_2d


def compute_l2_error_2d(mesh: Mesh, u_h, u_exact) -> float:
    """poisson_solver.rs:62-96 (and compute_l2_error_3d, :135-172): the L2
    norm of u_h - u_exact with the degree-3 rule; u_exact(x: dim x N array)
    -> values, or its values at mesh.quad_points(3)."""
    uq = _sample(mesh, u_exact, degree=3)
    u = _f64(u_h)
    out = dbl()
    _check(lib().fh_l2_error(mesh.h, _p(u), _p(uq), C.byref(out)))
    return out.value


compute_l2_error_3d = compute_l2_error_2d


@dataclasses.dataclass
class ElasticitySolution:
    """elasticity_solver.rs:12-28 (stresses filled by compute_stresses=True)."""
    displacements: np.ndarray
    iterations: int
    residual: float
    deformed_vertices: np.ndarray = None
    element_stresses: np.ndarray = None
    vertex_von_mises: np.ndarray = None
    max_von_mises: float = 0.0
    stats: SolveStats = None


def solve_elasticity_3d(mesh: Mesh, material: ElasticMaterial, body_force,
                        bc: VectorDirichletBC, tol, max_iter, precision=FP64, E_cell=None,
                        compute_stresses=True) -> ElasticitySolution:
    """elasticity_solver.rs:31-72."""
    pd = list(bc.prescribed_dofs)
    dofs = _u64([d for d, _ in pd])
    vals = _f64([v for _, v in pd])
    f = _f64(body_force)
    ec = None if E_cell is None else _f64(E_cell)
    u = np.empty(3 * mesh.nv)
    st = SolveStats()
    _check(lib().fh_solve_elasticity_3d(mesh.h, material.youngs_modulus, material.poissons_ratio,
                                        _p(ec), _p(f), len(dofs), _p(dofs), _p(vals), tol,
                                        max_iter, precision, _p(u), None, C.byref(st)))
    sol = ElasticitySolution(u, int(st.iterations), st.final_residual, stats=st)
    sol.deformed_vertices = mesh.vertices + u.reshape(-1, 3)
    if compute_stresses and hasattr(lib(), "fh_element_stresses"):
        sol.element_stresses = element_stresses(mesh, u, material)
        sol.vertex_von_mises = vertex_von_mises(mesh, sol.element_stresses[:, 18])
        sol.max_von_mises = float(max(0.0, sol.vertex_von_mises.max(initial=0.0)))
    return sol


def element_stresses(mesh: Mesh, u, material: ElasticMaterial) -> np.ndarray:
    """elasticity.rs:230-320: per cell stress[9], strain[9], von_mises, trace."""
    u = _f64(u)
    out = np.empty((mesh.nc, 20))
    _check(lib().fh_element_stresses(mesh.h, _p(u), material.youngs_modulus,
                                     material.poissons_ratio, _p(out)))
    return out


def vertex_von_mises(mesh: Mesh, vm_cell) -> np.ndarray:
    """elasticity.rs:323-360."""
    vm = _f64(vm_cell)
    out = np.empty(mesh.nv)
    _check(lib().fh_vertex_von_mises(mesh.h, _p(vm), _p(out)))
    return out


def cell_bilinear_k0(mesh: Mesh, nu, w, u) -> np.ndarray:
    """out[c] = -w_c^T K0_c u_c (E = 1 element stiffness); dJ/dE_c for w = u."""
    w, u = _f64(w), _f64(u)
    out = np.empty(mesh.nc)
    _check(lib().fh_cell_bilinear_k0(mesh.h, float(nu), _p(w), _p(u), _p(out)))
    return out


def batched_poisson_block_cg(sizes, f_nodal, tol=1e-10, max_iter=20000, precision=FP64,
                             device=-1):
    """Many unit_square(N_i, N_i) Poisson problems solved by block CG in one
    persistent launch, with nrhs = 1..4 right-hand sides each (b = M f from
    nodal loads f: list of [(N_i+1)^2, nrhs] arrays, or [(N_i+1)^2] for one).
    Returns (list of solutions of f's shape, iterations [n_inst, nrhs] (the
    block iteration at which each column converged), SolveStats)."""
    sizes = _u64(sizes)
    arrs = [np.asarray(a, dtype=np.float64) for a in f_nodal]
    squeeze = arrs[0].ndim == 1
    nrhs = 1 if squeeze else arrs[0].shape[1]
    f = _f64(np.concatenate([a.reshape(-1, nrhs) for a in arrs]))
    u = np.empty_like(f)
    its = np.empty((len(sizes), nrhs), np.uint64)
    st = SolveStats()
    _check(lib().fh_batched_poisson_block_cg(len(sizes), _p(sizes), nrhs, _p(f), tol, max_iter,
                                             precision, device, _p(u), _p(its), C.byref(st)))
    out, o = [], 0
    for N in sizes:
        k = (int(N) + 1) ** 2
        out.append(u[o:o + k, 0].copy() if squeeze else u[o:o + k])
        o += k
    return out, its, st


def assemble_elasticity_mass_3d(mesh: Mesh, density: float) -> SparseMatrix:
    """mass.rs:86-126: the 3D vector mass rho * int phi_i phi_j delta_ab."""
    return _new_matrix(lib().fh_assemble_elasticity_mass, mesh.h, float(density))


def assemble_elasticity_mass_p2_3d(nodes, cells, density: float, device: int = -1) -> SparseMatrix:
    """mass.rs:136-180: P2 nodes [n, 3], cells [nc, 10] (corners, then edges
    01, 12, 02, 03, 13, 23)."""
    v, c = _f64(nodes).reshape(-1, 3), _u64(cells).reshape(-1, 10)
    return _new_matrix(lib().fh_assemble_elasticity_mass_p2, len(v), _p(v), len(c), _p(c),
                       float(density), device)


def fold_dirichlet(mat: SparseMatrix, fixed_dofs, unit_diag: bool) -> SparseMatrix:
    """eigen.rs:51-78."""
    d = _u64(fixed_dofs)
    return _new_matrix(lib().fh_fold_dirichlet, mat.h, len(d), _p(d), 1 if unit_diag else 0)


class VibrationMode:
    """eigen.rs:9-19."""

    def __init__(self, frequency_hz, omega_rad_s, eigenvalue, mode_shape):
        self.frequency_hz = frequency_hz
        self.omega_rad_s = omega_rad_s
        self.eigenvalue = eigenvalue
        self.mode_shape = mode_shape


def solve_vibration_modes(k: SparseMatrix, m: SparseMatrix, fixed_dofs, num_modes: int,
                          max_power_iters: int, tol: float, inner_tol: float = 1e-9,
                          inner_max_iter: int = 1000, precision: int = FP64):
    """eigen.rs:288-300 (and solve_vibration_modes_with_inner, eigen.rs:94-266,
    with the device Jacobi-CG as the inner solve).  Returns the list of
    VibrationMode; the inner-solve iteration total is in .inner_iterations of
    the returned list."""
    d = _u64(fixed_dofs)
    n = k.n
    f, om, ev = np.zeros(num_modes), np.zeros(num_modes), np.zeros(num_modes)
    shapes = np.zeros((num_modes, n))
    it = u64(0)
    _check(lib().fh_vibration_modes(k.h, m.h, len(d), _p(d), num_modes, max_power_iters, tol,
                                    inner_tol, inner_max_iter, precision, _p(f), _p(om),
                                    _p(ev), _p(shapes), C.byref(it)))
    out = _ModeList(VibrationMode(float(f[i]), float(om[i]), float(ev[i]), shapes[i])
                    for i in range(num_modes))
    out.inner_iterations = int(it.value)
    return out


class _ModeList(list):
    inner_iterations = 0


InnerSolveFn = C.CFUNCTYPE(C.c_int, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.c_void_p)


def solve_vibration_modes_with_inner(k: SparseMatrix, m: SparseMatrix, fixed_dofs,
                                     num_modes: int, max_power_iters: int, tol: float, inner):
    """eigen.rs:94-266 with the caller's inner solve: inner(b, x0) returns x
    solving K x = b in the Dirichlet frame (K with constrained rows and columns
    zero and a unit diagonal on them), x0 the previous iterate.  The power
    iteration runs on the device; b, x0 and x cross to the host per call.  An
    exception raised by inner is re-raised after the call."""
    d = _u64(fixed_dofs)
    n = k.n
    f, om, ev = np.zeros(num_modes), np.zeros(num_modes), np.zeros(num_modes)
    shapes = np.zeros((num_modes, n))
    errors = []

    def _inner(nn, bp, x0p, xp, _ctx):
        try:
            b = np.ctypeslib.as_array(bp, (nn,)).copy()
            x0 = np.ctypeslib.as_array(x0p, (nn,)).copy()
            x = np.asarray(inner(b, x0), dtype=np.float64)
            if x.shape != (nn,):
                raise FenicsError(FH_ERR_BAD_INPUT,
                                  f"inner solve returned shape {x.shape}, expected ({nn},)")
            np.ctypeslib.as_array(xp, (nn,))[:] = x
            return 0
        except BaseException as e:  # re-raised after the call
            errors.append(e)
            return 1

    cb = InnerSolveFn(_inner)
    code = lib().fh_vibration_modes_with_inner(k.h, m.h, len(d), _p(d), num_modes,
                                               max_power_iters, tol, cb, None, _p(f), _p(om),
                                               _p(ev), _p(shapes))
    if errors:
        raise errors[0]
    _check(code)
    return [VibrationMode(float(f[i]), float(om[i]), float(ev[i]), shapes[i])
            for i in range(num_modes)]


# ---------------------------------------------- boundary (ds) integrals
@dataclasses.dataclass
class BoundaryFacet:
    """mesh.rs BoundaryFacet: vertices (2D the sorted edge, 3D outward
    order), the cell holding it and its local facet index."""
    vertices: tuple
    cell_idx: int
    local_facet_idx: int


def boundary_facet_arrays(mesh: Mesh):
    """extract_boundary_facets as arrays: (vertices [nf, dim], cell [nf],
    local [nf]), sorted by the sorted vertex tuple."""
    n = u64(0)
    _check(lib().fh_mesh_boundary_facets(mesh.h, None, None, None, C.byref(n)))
    nf = int(n.value)
    dim = mesh.vertices.shape[1]
    v = np.empty((nf, dim), np.uint64)
    c = np.empty(nf, np.uint64)
    l = np.empty(nf, np.uint64)  # noqa: E741
    if nf:
        _check(lib().fh_mesh_boundary_facets(mesh.h, _p(v), _p(c), _p(l), C.byref(n)))
    return v, c, l


def extract_boundary_facets(mesh: Mesh):
    """mesh.rs:79-113 / 145-210."""
    v, c, l = boundary_facet_arrays(mesh)
    return [BoundaryFacet(tuple(int(a) for a in v[i]), int(c[i]), int(l[i])) for i in range(len(c))]


def _facet_array(facets, dim):
    if isinstance(facets, np.ndarray):
        return _u64(facets).reshape(-1, dim)
    return _u64([f.vertices if isinstance(f, BoundaryFacet) else f for f in facets]).reshape(-1, dim)


def boundary_points(mesh: Mesh, facets) -> np.ndarray:
    """Where boundary.rs samples its closures: [nf, 2, 2] Gauss points (2D),
    [nf, 3, 3] edge midpoints (3D)."""
    dim = mesh.vertices.shape[1]
    f = _facet_array(facets, dim)
    out = np.empty((len(f), 2, 2) if dim == 2 else (len(f), 3, 3))
    _check(lib().fh_boundary_points(mesh.h, len(f), _p(f), _p(out)))
    return out


def assemble_neumann_2d(mesh: Mesh, boundary_facets, flux_fn: Callable) -> np.ndarray:
    """boundary.rs:8-46: flux_fn((x, y)) -> float, sampled facet by facet."""
    f = _facet_array(boundary_facets, 2)
    pts = boundary_points(mesh, f)
    g = np.array([[flux_fn((p[0], p[1])) for p in fp] for fp in pts], np.float64).reshape(len(f), 2)
    out = np.empty(mesh.num_vertices())
    _check(lib().fh_assemble_neumann_2d(mesh.h, len(f), _p(f), _p(g), _p(out)))
    return out


def assemble_surface_traction_3d(mesh: Mesh, boundary_facets, traction_fn: Callable) -> np.ndarray:
    """boundary.rs:49-108: traction_fn((x, y, z)) -> (tx, ty, tz) at each
    facet's edge midpoints."""
    f = _facet_array(boundary_facets, 3)
    mids = boundary_points(mesh, f)
    t = np.array([[traction_fn(tuple(p)) for p in fm] for fm in mids],
                 np.float64).reshape(len(f), 3, 3)
    out = np.empty(3 * mesh.num_vertices())
    _check(lib().fh_assemble_surface_traction_3d(mesh.h, len(f), _p(f), _p(t), _p(out)))
    return out


def assemble_robin_matrix_2d(mesh: Mesh, boundary_facets, gamma: float) -> SparseMatrix:
    """boundary.rs:111-139."""
    f = _facet_array(boundary_facets, 2)
    return _new_matrix(lib().fh_assemble_robin_matrix_2d, mesh.h, len(f), _p(f), float(gamma))


# ------------------------------------------------------- transient heat
class TimeSteppingScheme:
    """transient.rs:12-18."""
    BackwardEuler = 0
    CrankNicolson = 1


class TransientStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("cg_iterations", C.c_uint64),
                ("cg_iterations_low", C.c_uint64), ("cg_iterations_max", C.c_uint64),
                ("solve_ms", C.c_double), ("wall_ms", C.c_double), ("device", C.c_int),
                ("precision", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


HeatSourceFn = C.CFUNCTYPE(C.c_double, C.c_double, C.POINTER(C.c_double), C.c_void_p)


def matrix_from_csc(col_ptr, row_idx, values, device: int = -1) -> SparseMatrix:
    """A faer SparseColMat's raw parts (column pointers, row indices, values)
    as a device operator."""
    cp, ri, va = _u64(col_ptr), _u64(row_idx), _f64(values)
    return _new_matrix(lib().fh_matrix_from_csc, len(cp) - 1, _p(cp), _p(ri), _p(va), device)


def add_scaled_matrices(a: SparseMatrix, b: SparseMatrix, alpha: float) -> SparseMatrix:
    """transient.rs:72-97: a + alpha * b."""
    return _new_matrix(lib().fh_add_scaled_matrices, a.h, b.h, float(alpha))


class TransientHeatSolver:
    """transient.rs:20-69.  The implicit system is solved by the device CG
    (from x0 = u_n, to ||r||/||b|| < tol) in place of the reference's dense LU."""

    def __init__(self, scheme: int, dt: float, tol: float = 1e-12, max_iter: int = 10000,
                 precision: int = FP64_TRUE):
        self.scheme, self.dt = scheme, dt
        self.tol, self.max_iter, self.precision = tol, max_iter, precision
        self.last_stats = None

    def step(self, mass_mat: SparseMatrix, stiff_mat: SparseMatrix, u_n, f_n, f_next):
        n = mass_mat.n
        u, fa, fb = _f64(u_n), _f64(f_n), _f64(f_next)
        if not (len(u) == len(fa) == len(fb) == n):
            raise FenicsError(FH_ERR_BAD_INPUT, f"vector sizes {len(u)}/{len(fa)}/{len(fb)} "
                                                f"do not match the matrix dimension {n}")
        out = np.empty(n)
        st = SolveStats()
        _check(lib().fh_transient_heat_step(mass_mat.h, stiff_mat.h, self.scheme, self.dt,
                                            _p(u), _p(fa), _p(fb), self.tol, self.max_iter,
                                            self.precision, _p(out), C.byref(st)))
        self.last_stats = st
        return out


@dataclasses.dataclass
class TransientHeatSolution:
    """heat_solver.rs:13-19 (plus the device statistics)."""
    time_steps: list
    temperature_history: list
    final_temperature: np.ndarray
    stats: dict


def solve_transient_heat_2d(mesh: Mesh, diffusivity: float, u_initial, bcs: Sequence[DirichletBC],
                            dt: float, num_steps: int, scheme: int,
                            heat_source: Optional[Callable] = None, tol: float = 1e-12,
                            max_iter: int = 10000, precision: int = FP64_TRUE,
                            keep_history: bool = True) -> TransientHeatSolution:
    """heat_solver.rs:23-108.  heat_source(t, point) -> float is sampled at the
    vertices at t_n and t_{n+1} of every step (None: no source)."""
    n = mesh.num_vertices()
    u0 = _f64(u_initial)
    if len(u0) != n:
        raise FenicsError(FH_ERR_BAD_INPUT,
                          f"Initial temperature size {len(u0)} != mesh vertices {n}")
    dofs = [int(d) for bc in bcs for d in bc.dofs]
    vals = [float(v) for bc in bcs for v in bc.values]
    if any(len(bc.dofs) != len(bc.values) for bc in bcs):
        raise FenicsError(FH_ERR_BAD_INPUT, "DirichletBC dofs and values differ in length")
    d, v = _u64(dofs), _f64(vals)
    errors = []
    if heat_source is None:
        cb = C.cast(None, HeatSourceFn)
    else:
        def _src(t, pt, _ctx):
            try:
                return float(heat_source(t, (pt[0], pt[1])))
            except BaseException as e:  # re-raised after the call
                errors.append(e)
                return 0.0
        cb = HeatSourceFn(_src)
    hist = np.empty((num_steps + 1, n)) if keep_history else None
    final = np.empty(n)
    st = TransientStats()
    code = lib().fh_solve_transient_heat_2d(mesh.h, float(diffusivity), _p(u0), len(d), _p(d),
                                            _p(v), float(dt), int(num_steps), int(scheme), cb,
                                            None, tol, max_iter, precision, _p(hist), _p(final),
                                            C.byref(st))
    if errors:
        raise errors[0]
    _check(code)
    return TransientHeatSolution([k * dt for k in range(num_steps + 1)],
                                 list(hist) if keep_history else [], final, st.as_dict())


def clamp_dofs_x0(vertices, tol=1e-6):
    """elasticity_tests.rs:94-101: all three dofs of the vertices with x < 1e-6."""
    idx = np.nonzero(np.asarray(vertices)[:, 0] < tol)[0]
    return VectorDirichletBC([(3 * int(i) + a, 0.0) for i in idx for a in range(3)])


def stream_bandwidth(nbytes: int, kind: str = "read", reps: int = 5, device: int = 0):
    """Measured bandwidth ceiling of the device (fh_stream_bandwidth):
    kind 'read', 'copy' (read + write bytes) or 'write'.  Returns (GB/s, ms)."""
    k = {"read": 0, "copy": 1, "write": 2}[kind]
    g, ms = C.c_double(), C.c_double()
    _check(lib().fh_stream_bandwidth(device, int(nbytes), k, reps, C.byref(g), C.byref(ms)))
    return g.value, ms.value
