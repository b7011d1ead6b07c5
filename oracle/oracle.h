/*
 * oracle.h — CPU FP64 restatement of the fenics-rs assembly + CG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (libfenics_hip.so) never links, loads or
 * calls it.
 *
 * Every function restates the reference algorithm it names (file:line into
 * the fenics-rs sources).  Arithmetic is written operation by operation so
 * that it is reproducible: the library is compiled with -ffp-contract=off
 * (no fused multiply-add), matching Rust's default of never contracting
 * a*b+c.  Where the reference delegates to faer 0.20.2 or nalgebra 0.33.3
 * (third-party crates not present in the reference tree) the textbook
 * algorithm is used and the rounding of the last ulp is "parity unpinned"
 * (see the per-function notes).
 *
 * All indices are uint64_t (Rust usize); all arrays are caller-allocated.
 * Status codes: 0 = Ok, 1 = CG breakdown, 2 = CG not converged,
 *               3 = bad input.
 */
#ifndef FENICS_ORACLE_H
#define FENICS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_csc or_csc;   /* canonical CSC, faer SparseColMat semantics */
typedef struct or_p2 or_p2;     /* P2 tetrahedral dofmap                        */

/* ---------------- mesh: generators.rs, mesh.rs ---------------- */
int or_unit_square(uint64_t nx, uint64_t ny, double* verts, uint64_t* cells);
int or_box_beam(double L, double W, double H, uint64_t nx, uint64_t ny, uint64_t nz,
                double* verts, uint64_t* cells);
/* boundary facets (count == 1), returned sorted by their sorted vertex tuple */
int or_boundary_facets_2d(uint64_t nc, const uint64_t* cells, uint64_t* facets_out,
                          uint64_t* n_out);
int or_boundary_facets_3d(uint64_t nc, const uint64_t* cells, uint64_t* facets_out,
                          uint64_t* n_out);
/* with cell, local index and (3D) outward vertex order, same sort */
int or_boundary_facets_full(int dim, const double* verts, uint64_t nc, const uint64_t* cells,
                            uint64_t* verts_out, uint64_t* cell_out, uint64_t* local_out,
                            uint64_t* n_out);
int or_boundary_dofs_2d(uint64_t nc, const uint64_t* cells, uint64_t* dofs_out, uint64_t* n_out);
int or_boundary_dofs_3d(uint64_t nc, const uint64_t* cells, uint64_t* dofs_out, uint64_t* n_out);
double or_total_volume_2d(const double* verts, uint64_t nc, const uint64_t* cells);
double or_total_volume_3d(const double* verts, uint64_t nc, const uint64_t* cells);
double or_boundary_measure_2d(const double* verts, uint64_t nc, const uint64_t* cells);
double or_boundary_measure_3d(const double* verts, uint64_t nc, const uint64_t* cells);

/* ---------------- element: jacobian.rs, lagrange.rs, quadrature.rs ---------------- */
int or_affine_map_2d(const double* v /*3x2*/, double* J /*2x2 row-major*/, double* invT,
                     double* det);
int or_affine_map_3d(const double* v /*4x3*/, double* J /*3x3*/, double* invT, double* det);
int or_quadrature(int dim, int degree, double* points, double* weights, int* npts);
void or_p2_tet_basis(const double* xi, double* out10);
void or_p2_tet_grads(const double* xi, double* out30);

/* ---------------- CSC container (faer try_new_from_triplets semantics) ---------------- */
or_csc* or_csc_from_triplets(uint64_t n, uint64_t nt, const uint64_t* rows, const uint64_t* cols,
                             const double* vals);
/* adopt raw CSC parts that are already canonical (sorted, unique) */
or_csc* or_csc_from_parts(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                          const double* vals);
uint64_t or_csc_n(const or_csc* a);
uint64_t or_csc_nnz(const or_csc* a);
void or_csc_export(const or_csc* a, uint64_t* col_ptr, uint64_t* row_idx, double* vals);
void or_csc_free(or_csc* a);

/* ---------------- assembly: poisson.rs, elasticity.rs, mass.rs ---------------- */
or_csc* or_assemble_stiffness_2d(uint64_t nv, const double* verts, uint64_t nc,
                                 const uint64_t* cells, int nthreads);
or_csc* or_assemble_stiffness_3d(uint64_t nv, const double* verts, uint64_t nc,
                                 const uint64_t* cells, int nthreads);
or_csc* or_assemble_mass_2d(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells,
                            int nthreads);
or_csc* or_assemble_elasticity_3d(uint64_t nv, const double* verts, uint64_t nc,
                                  const uint64_t* cells, double E, double nu, int nthreads);
/* per-cell Young's modulus (E_cell[nc]); E_cell == E everywhere reproduces the above */
or_csc* or_assemble_elasticity_3d_cellE(uint64_t nv, const double* verts, uint64_t nc,
                                        const uint64_t* cells, const double* E_cell, double nu,
                                        int nthreads);
/* physical quadrature points of the degree-`degree` rule, [nc][nq][dim] */
int or_quad_points_2d(const double* verts, uint64_t nc, const uint64_t* cells, int degree,
                      double* xq);
int or_quad_points_3d(const double* verts, uint64_t nc, const uint64_t* cells, int degree,
                      double* xq);
/* load vectors with f sampled at the degree-2 quadrature points, fq[nc][nq] */
int or_assemble_rhs_2d(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells,
                       const double* fq, double* b);
int or_assemble_rhs_3d(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells,
                       const double* fq, double* b);
int or_assemble_body_force_3d(uint64_t nv, const double* verts, uint64_t nc,
                              const uint64_t* cells, const double* f3, double* b);
/* Dirichlet condensation; returns the new matrix, writes the new rhs */
or_csc* or_apply_dirichlet_bc(const or_csc* a, const double* rhs, uint64_t nfix,
                              const uint64_t* dofs, const double* vals, double* rhs_out);
/* L2 error with u_exact sampled at the degree-3 quadrature points, uq[nc][nq] */
double or_l2_error_2d(const double* verts, uint64_t nc, const uint64_t* cells, const double* u,
                      const double* uq);
double or_l2_error_3d(const double* verts, uint64_t nc, const uint64_t* cells, const double* u,
                      const double* uq);

/* mass.rs:86-126 (3D vector mass for elasticity, 3*nv dofs) */
or_csc* or_assemble_elasticity_mass_3d(uint64_t nv, const double* verts, uint64_t nc,
                                       const uint64_t* cells, double density, int nthreads);

/* ---------------- modal analysis: eigen.rs ---------------- */
/* eigen.rs:51-78: drop entries on constrained rows/columns and exact zeros;
   unit_diag: constrained diagonal = 1 */
or_csc* or_fold_dirichlet(const or_csc* a, uint64_t nfix, const uint64_t* fixed, int unit_diag);
/* eigen.rs:94-266 with the inner solve solve_cg(kd, b, x0, inner_tol,
   inner_max_iter) (eigen.rs:288-300 uses 1e-9, 1000).  shapes: [num_modes][n]
   or NULL.  Returns 0, or the inner CG's status on its failure (the Err of
   eigen.rs), or 4 when no dof is free. */
int or_vibration_modes(const or_csc* k, const or_csc* m, uint64_t nfix, const uint64_t* fixed,
                       uint64_t num_modes, uint64_t max_power_iters, double tol,
                       double inner_tol, uint64_t inner_max_iter, double* freq_hz,
                       double* omega, double* eigenvalue, double* shapes);

/* ---------------- solver: cg.rs, gpu.rs CSC->CSR ---------------- */
void or_spmv(const or_csc* a, const double* x, double* y);
int or_solve_cg(const or_csc* a, const double* b, const double* x0, double tol,
                uint64_t max_iter, double* x, uint64_t* iters, double* res);
void or_csc_to_csr(const or_csc* a, uint32_t* row_ptr, uint32_t* col_idx, double* vals,
                   double* inv_diag);
/* single-thread CG restricted to `iters` iterations, for timing; returns seconds */
double or_time_cg_iterations(const or_csc* a, const double* b, uint64_t iters);

/* ---------------- P2 tetrahedra (new, not in the reference) ---------------- */
or_p2* or_p2_build(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells);
uint64_t or_p2_ndofs(const or_p2* p);
uint64_t or_p2_nedges(const or_p2* p);
void or_p2_cell_dofs(const or_p2* p, uint64_t* out /* nc*10 */);
void or_p2_dof_coords(const or_p2* p, double* out /* ndofs*3 */);
int or_p2_boundary_dofs(const or_p2* p, uint64_t* out, uint64_t* n_out);
or_csc* or_assemble_p2_stiffness_3d(const or_p2* p, int nthreads);
int or_assemble_p2_rhs_cellconst(const or_p2* p, const double* f_cell, double* b);
void or_p2_free(or_p2* p);

/* ---------------- adjoint / stresses (new; elasticity.rs:230-360) ---------------- */
/* out[c] = -(w_c)^T K0_c u_c with K0_c the E = 1 element stiffness;
 * w = u gives dJ/dE_c for compliance J = f^T u */
int or_cell_bilinear_k0(const double* verts, uint64_t nc, const uint64_t* cells, double nu,
                        const double* w, const double* u, double* out);
/* per cell: stress[9], strain[9], von_mises, trace (20 doubles) */
int or_element_stresses(const double* verts, uint64_t nc, const uint64_t* cells, const double* u,
                        double E, double nu, double* out);
int or_vertex_von_mises(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells,
                        const double* vm_cell, double* out);

/* mass.rs:190-250 reference products int phi_a phi_b, out[a * 10 + b] */
void or_p2_mass_products(double* out);
/* mass.rs:136-180 on nodes [n][3], cells [nc][10] */
or_csc* or_assemble_elasticity_mass_p2_3d(uint64_t n_nodes, const double* verts, uint64_t nc,
                                          const uint64_t* cells, double density, int nthreads);

/* ---- boundary (ds) integrals, boundary.rs ---- */
/* Gauss points [nf][2][2] of each facet edge; the flux sampled there is g[nf][2] */
int or_neumann_points_2d(const double* verts, uint64_t nf, const uint64_t* facets, double* xq);
int or_assemble_neumann_2d(uint64_t nv, const double* verts, uint64_t nf, const uint64_t* facets,
                           const double* g, double* rhs);
/* edge midpoints [nf][3][3] of each facet triangle; the traction there is t[nf][3][3] */
int or_traction_points_3d(const double* verts, uint64_t nf, const uint64_t* facets, double* mids);
int or_assemble_surface_traction_3d(uint64_t nv, const double* verts, uint64_t nf,
                                    const uint64_t* facets, const double* t, double* rhs);
or_csc* or_assemble_robin_matrix_2d(uint64_t nv, const double* verts, uint64_t nf,
                                    const uint64_t* facets, double gamma);

#ifdef __cplusplus
}
#endif
#endif
