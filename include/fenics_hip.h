/*
 * fenics_hip.h — C ABI of libfenics_hip, the MI355X (gfx950) implementation
 * of the fenics-rs assembly + conjugate-gradient hot path.
 *
 * Every entry point is plain C: pointers and sizes, no C++ or torch types.
 * Conventions (mirroring fenics-rs, crates/fenics-solver/src/gpu.rs):
 *   - every call returns an int status (FH_OK == 0) and never aborts across
 *     the FFI; fh_last_error() returns a thread-local message for the last
 *     failure (the Rust side maps it to Err(String), as gpu.rs:170-176 does
 *     for CUDA errors);
 *   - host arrays are borrowed for the duration of the call; outputs go to
 *     caller-allocated buffers sized by a query-then-fill pattern;
 *   - objects (fh_mesh, fh_matrix, fh_bsr3_ctx) own their device memory and
 *     are used by one thread at a time (Send, not Sync — gpu.rs:450);
 *   - a CPU fallback stays a caller decision (gpu.rs:1584-1601): this library
 *     never computes on the host.
 * Entry points whose names end in _dev take device pointers that must live
 * on the object's device.  Indices at the CSC boundary are uint64_t
 * (faer's usize); device-side indices are uint32_t.
 */
#ifndef FENICS_HIP_H
#define FENICS_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define FH_OK 0
#define FH_ERR_NO_DEVICE 1
#define FH_ERR_NOT_CONVERGED 2 /* cg.rs:139-148 "CG failed to converge"        */
#define FH_ERR_BREAKDOWN 3     /* cg.rs:102-107 |p^T A p| < 1e-30                */
#define FH_ERR_BAD_INPUT 4     /* shapes, indices out of range, bad material     */
#define FH_ERR_HIP 5           /* HIP runtime error, message in fh_last_error()  */
#define FH_ERR_UNSUPPORTED 6

/* arithmetic of a solve */
#define FH_PREC_FP64 0      /* the reference's CG: FP64 SpMV, stop on the recursive residual */
#define FH_PREC_MIXED 1     /* as FH_PREC_FP64_TRUE, with FP32 matrix values (difference-form
                               SpMV, FP64 vectors) once ||r||/||b|| < 1e-3                    */
#define FH_PREC_MIXED_IR 2  /* FP32 matrix values from the start, restarted at each FP64
                               replacement (iterative refinement)                            */
#define FH_PREC_FP64_TRUE 3 /* FP64 SpMV with periodic replacement by the
                               double-double true residual; stops when the true residual is
                               below tol or at the FP64 rounding floor (residual_floor)      */

/* SpMV accumulation order */
#define FH_ORDER_FAST 0    /* fused multiply-add, any order within a row           */
#define FH_ORDER_ORDERED 1 /* cg.rs:13-21 order: column-ascending, no FMA (bitwise) */

/* replaces GpuDeviceInfo (gpu.rs:12-21) */
typedef struct fh_device_info {
    int index;
    char name[128];
    char arch[64]; /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    uint64_t total_memory_mib;
    uint64_t free_memory_mib;
    int compute_units; /* multiprocessor_count */
    int clock_mhz;
} fh_device_info;

/* replaces GpuSolveResult (gpu.rs:24-34) minus the solution, which goes to x_out */
typedef struct fh_solve_stats {
    uint64_t iterations;          /* CG iterations (all precisions)                      */
    uint64_t iterations_low;      /* of which on the FP32 matrix values                  */
    uint64_t replacements;        /* FP64 residual replacements / refinement steps       */
    double final_residual;        /* cg.rs semantics: absolute recursive ||r||           */
    double rel_residual;          /* recursive ||r|| / ||b|| (gpu.rs final_residual)     */
    double true_rel_residual;     /* ||b - A x|| / ||b||, A x summed in double-double     */
    double residual_floor;        /* true-residual modes: relative residual that rounding
                                     x to FP64 alone leaves, 2 eps ||(sum_j (a_ij x_j)^2)^.5||
                                     / ||b||; a solve also converges when its true
                                     residual reaches it (0 for FH_PREC_FP64)          */
    double wall_ms;
    double kernel_ms;
    double gflops;
    double effective_bw_gbps;     /* algorithmic bytes per iteration / time per iteration */
    double spmv_bytes;            /* algorithmic bytes of one SpMV of this matrix        */
    double iter_bytes;            /* algorithmic bytes of one CG iteration               */
    double spmv_avg_ms;           /* mean duration of the CG SpMV kernel (HIP events), 0 if
                                     not timed (set FH_TIME_SPMV=1 to time it)           */
    uint64_t spmv_timed;          /* SpMV launches that average covers                   */
    double spmv32_bytes;          /* algorithmic bytes of one FP32-value SpMV            */
    double spmv32_avg_ms;         /* mean duration of the FP32-value SpMV kernel         */
    uint64_t spmv32_timed;
    int device;
    int precision;
    int status;
} fh_solve_stats;

typedef struct fh_mesh fh_mesh;         /* device-resident simplicial mesh           */
typedef struct fh_matrix fh_matrix;     /* device-resident assembled sparse operator */
typedef struct fh_bsr3_ctx fh_bsr3_ctx; /* replaces GpuBsr3Context (gpu.rs:432-658)   */

/* ------------------------------------------------------- devices, errors */
const char* fh_last_error(void);
int fh_device_count(int* count);
int fh_get_device_info(int index, fh_device_info* out);      /* gpu.rs:178-234 discover_gpus */
int fh_best_device(int* index);                           /* gpu.rs:237-242 best_gpu_index */
int fh_synchronize(int device);
/* Measurement utility (new; the reference prints per-solve GB/s only,
 * gpu.rs:1119-1129): the best of `reps` event-timed passes of a read stream
 * (kind 0, 16-byte non-temporal loads), a copy (kind 1; read + write bytes
 * counted) or a write stream (kind 2) over `bytes` of device memory.       */
int fh_stream_bandwidth(int device, uint64_t bytes, int kind, int reps, double* gbps,
                        double* best_ms);

/* ------------------------------------------------------------------ mesh
 * replaces SimplicialMesh<DIM,NODES> (crates/fenics-mesh/src/mesh.rs:10-20);
 * vertices AoS [nv][dim] f64, cells [nc][dim+1] u64 as in the reference.   */
int fh_mesh_create(int dim, uint64_t nv, const double* vertices, uint64_t nc,
                   const uint64_t* cells, int device, fh_mesh** out);
/* generators.rs:24-55 unit_square, built on the device */
int fh_mesh_unit_square(uint64_t nx, uint64_t ny, int device, fh_mesh** out);
/* generators.rs:58-111 box_beam (unit_cube = box_beam(1,1,1,..)), on the device */
int fh_mesh_box_beam(double L, double W, double H, uint64_t nx, uint64_t ny, uint64_t nz,
                     int device, fh_mesh** out);
int fh_mesh_sizes(const fh_mesh* m, int* dim, uint64_t* nv, uint64_t* nc);
int fh_mesh_download(const fh_mesh* m, double* vertices, uint64_t* cells);
/* poisson.rs:266-287 boundary_dofs_2d/3d: sorted vertex ids on boundary
 * facets (mesh.rs:79-113,145-210).  out may be NULL to query the count.   */
int fh_mesh_boundary_dofs(fh_mesh* m, uint64_t* out, uint64_t* count);
/* physical points of the degree-`degree` rule, [nc][nq][dim]
 * (map_to_physical, jacobian.rs:44-48,88-92); nq = 3/4 (2D/3D, degree 2)   */
int fh_quad_points(fh_mesh* m, int degree, double* xq_out);
int fh_mesh_destroy(fh_mesh* m);

/* -------------------------------------------------------------- assembly
 * Matrices keep faer's canonical CSC pattern of the connectivity (sorted
 * rows per column, duplicates summed in ascending cell order, structural
 * zeros kept); stored on the device in 64-row sliced ELL (SELL-64) with
 * 1x1 or 3x3 blocks, FP64 values plus an FP32 copy for mixed solves.       */
/* poisson.rs:23-69 / 124-165 assemble_stiffness_2d/3d (by mesh dim) */
int fh_assemble_stiffness(fh_mesh* m, fh_matrix** out);
/* mass.rs:10-44 / 47-82 assemble_mass_2d/3d */
int fh_assemble_mass(fh_mesh* m, fh_matrix** out);
/* elasticity.rs:73-149 assemble_elasticity_stiffness_3d; E_cell may be NULL
 * (uniform E) or a host array [nc] of per-cell Young's moduli              */
int fh_assemble_elasticity(fh_mesh* m, double E, double nu, const double* E_cell,
                           fh_matrix** out);
int fh_assemble_elasticity_dev(fh_mesh* m, double E, double nu, const double* E_cell_dev,
                               fh_matrix** out);
/* elasticity.rs:73-149 as above, by method: 0 = VALU FP64 (the values of
 * fh_assemble_elasticity, bitwise the reference's order); 1 = MFMA FP32:
 * element matrices B^T (vol D) B on the matrix cores (B the 6x12 Voigt
 * strain-displacement matrix, D the 6x6 isotropic stiffness), summed per row
 * in FP32, the FP64 values their widening; 2 = both: the FP64 values of
 * method 0 and the MFMA FP32 values as the matrix's FP32 copy             */
int fh_assemble_elasticity_method(fh_mesh* m, double E, double nu, const double* E_cell,
                                  int method, fh_matrix** out);
/* P2 tetrahedral Poisson stiffness (new): dofs = vertices (their own
 * numbers), then edges ordered by the number of cells around them, most
 * first, then by max - min, then by min; method 0 = VALU FP64 (exact
 * order of the oracle), 1 = MFMA FP32 element matrices (FP32 accumulate)  */
int fh_assemble_p2_stiffness(fh_mesh* m, int method, fh_matrix** out);
int fh_p2_info(fh_mesh* m, uint64_t* ndofs, uint64_t* nedges);
int fh_p2_cell_dofs(fh_mesh* m, uint64_t* out /* [nc][10] */);
int fh_p2_boundary_dofs(fh_mesh* m, uint64_t* out, uint64_t* count);
/* poisson.rs:72-121 / 168-218 assemble_rhs_2d/3d with the source sampled at
 * fh_quad_points(m, 2, ...): fq [nc][nq]                                   */
int fh_assemble_rhs(fh_mesh* m, const double* fq, double* b_out);
/* P2 load with a per-cell constant source f_cell [nc] (same rule)         */
int fh_assemble_p2_rhs_cellconst(fh_mesh* m, const double* f_cell, double* b_out);
/* elasticity.rs:153-185 assemble_elasticity_body_force_3d */
int fh_assemble_body_force(fh_mesh* m, const double f[3], double* b_out);
int fh_assemble_body_force_dev(fh_mesh* m, const double f[3], double* b_dev);
/* poisson.rs:222-263 apply_dirichlet_bc / elasticity.rs:188-227
 * apply_vector_dirichlet_bc: returns a new condensed matrix; b_in -> b_out
 * (may alias).  Fixed rows become identity rows, couplings to fixed dofs are
 * dropped (kept as explicit zeros on the device, dropped in the export).    */
int fh_apply_dirichlet(const fh_matrix* a, const double* b_in, uint64_t nfix,
                       const uint64_t* dofs, const double* vals, fh_matrix** out,
                       double* b_out);
int fh_apply_dirichlet_dev(const fh_matrix* a, const double* b_in_dev, uint64_t nfix,
                           const uint64_t* dofs, const double* vals, fh_matrix** out,
                           double* b_out_dev);
/* cells the assembly of `a` skipped because |det J| < 1e-14 (the silent
 * skip of jacobian.rs:29-31 / poisson.rs:42-45, counted; new)              */
int fh_matrix_skipped_cells(const fh_matrix* a, uint64_t* count);
/* matrix shape, scalar nnz of the canonical CSC, block size */
int fh_matrix_info(const fh_matrix* a, uint64_t* n, uint64_t* nnz, int* block);
/* faer SparseColMat raw parts: col_ptr[n+1], row_idx[nnz], values[nnz]     */
int fh_matrix_export_csc(const fh_matrix* a, uint64_t* col_ptr, uint64_t* row_idx, double* vals);
/* The arrays GpuBsr3Context::new / solve_bsr3_pcg_gpu take (gpu.rs:453-458,
 * 775-786) for a 3x3-blocked operator on any mesh: row_ptr[n_nodes+1],
 * cols[nnzb] (ascending per row), vals_flat[9 nnzb] (row-major blocks,
 * condensed couplings as explicit zeros) and dinv_flat[9 n_nodes], the
 * inverse of each diagonal block (adjugate over determinant; a singular block
 * is an error).  row_ptr is required; cols, vals_flat, dinv_flat may be null,
 * so a first call can size the arrays from row_ptr[n_nodes].  The reference
 * has no exporter: its callers build these arrays by hand.                  */
int fh_matrix_export_bsr3(const fh_matrix* a, uint32_t* row_ptr, uint32_t* cols,
                          double* vals_flat, double* dinv_flat);
int fh_matrix_spmv(const fh_matrix* a, const double* x, double* y, int order);
int fh_matrix_spmv_dev(const fh_matrix* a, const double* x_dev, double* y_dev, int order);
int fh_matrix_destroy(fh_matrix* a);

/* ---------------------------------------------------------------- solve
 * cg.rs:40-149 solve_cg with Jacobi preconditioning.  FH_PREC_FP64 keeps the
 * reference's stopping rule (||r_k||/||b|| < tol on the recursive residual).
 * The true-residual modes (FH_PREC_FP64_TRUE, FH_PREC_MIXED, FH_PREC_MIXED_IR)
 * stop on the double-double true residual ||b - A x||/||b||: when it is below
 * tol, when it is at or below residual_floor (the residual that rounding x to
 * FP64 alone leaves), or when replacement cycles stop halving it while it is
 * within 8x of residual_floor; stats->true_rel_residual says which residual
 * was reached.  Stagnation further above the floor returns
 * FH_ERR_NOT_CONVERGED.  x0 may be NULL.  Returns FH_ERR_NOT_CONVERGED /
 * FH_ERR_BREAKDOWN like cg.rs (x_out is still written).                      */
int fh_solve(const fh_matrix* a, const double* b, const double* x0, double tol,
             uint64_t max_iter, int precision, double* x_out, fh_solve_stats* stats);
int fh_solve_dev(const fh_matrix* a, const double* b_dev, const double* x0_dev, double tol,
                 uint64_t max_iter, int precision, double* x_dev, fh_solve_stats* stats);

/* New (no counterpart in the reference), for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc):
 * fh_solve_dev without any allocation.  Its buffers come from a static
 * scratch array of FH_NOALLOC_SCRATCH_MIB MiB per GPU (held until the
 * process ends), which bounds the system to about 1.8 million unknowns
 * (FH_ERR_BAD_INPUT beyond; the message gives the size needed).  The matrix
 * must already carry what its precision reads (FP32 copy, row sums, block
 * tail records, XCD map): solve it once with fh_solve_dev or fh_solve in the
 * same precision first, else FH_ERR_BAD_INPUT.  No HIP graphs, events or
 * pinned memory: kernels are launched one by one, and kernel_ms reports the
 * wall time.  Results are those of fh_solve_dev.  Every call on a GPU shares
 * that GPU's scratch array, and each call leaves it holding its own state:
 * the library's call lock serialises calls within a process, and work that
 * a caller enqueues on the library's stream meanwhile must not touch it.   */
#define FH_NOALLOC_SCRATCH_MIB 128
int fh_solve_dev_noalloc(const fh_matrix* a, const double* b_dev, const double* x0_dev,
                         double tol, uint64_t max_iter, int precision, double* x_dev,
                         fh_solve_stats* stats);
/* New (no counterpart in the reference, whose CG reports only its recursive
 * residual, cg.rs:139-148): the true relative residual ||b - A x||/||b|| of a
 * device vector x = x_hi + x_lo (x_lo may be NULL), evaluated in
 * double-double.                                                           */
int fh_true_residual_dev(const fh_matrix* a, const double* b_dev, const double* x_hi_dev,
                         const double* x_lo_dev, double* rel_out);
typedef struct fh_refine_stats {
    uint64_t outer_steps;         /* correction solves                               */
    uint64_t cg_iterations;       /* CG iterations over all correction solves        */
    double rel_residual_pair;     /* ||b - A (x_hi + x_lo)|| / ||b||                  */
    double rel_residual_rounded;  /* ||b - A x_hi|| / ||b||: x_hi is x_hi + x_lo
                                     rounded to FP64, so this is the residual the
                                     nearest FP64 vector to the solution leaves      */
} fh_refine_stats;
/* New (no counterpart in the reference): the solution in double-double.
 * Iterative refinement from x0 (NULL: zero): r = b - A (x_hi + x_lo) in
 * double-double, A d = r by the FH_PREC_FP64_TRUE CG to inner_tol (at most
 * max_iter iterations), (x_hi, x_lo) += d, until rel_residual_pair < target
 * or max_outer correction solves.  Returns FH_ERR_NOT_CONVERGED when the
 * target is not reached (outputs still written).                           */
int fh_solve_refined_dev(const fh_matrix* a, const double* b_dev, const double* x0_dev,
                         double target, uint64_t max_outer, double inner_tol, uint64_t max_iter,
                         double* x_hi_dev, double* x_lo_dev, fh_refine_stats* stats);

/* ---------------------------------------------- drop-ins for gpu.rs entries */
/* solve_cg (cg.rs:40-46) / solve_cg_gpu_or_cpu (gpu.rs:1542-1548) on faer's
 * raw CSC parts; device -1 = best device                                  */
int fh_solve_csc_cg(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                    const double* vals, const double* b, const double* x0, double tol,
                    uint64_t max_iter, int device, double* x_out, fh_solve_stats* stats);
/* solve_csr_cg_gpu (gpu.rs:1145-1155): caller-provided inverse diagonal; x0
 * is honoured when forming r0 (the reference ignores it, gpu.rs:1194)      */
int fh_solve_csr_cg(uint64_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                    const double* vals, const double* inv_diag, const double* b,
                    const double* x0, double rtol, uint64_t max_iter, int device,
                    double* x_out, fh_solve_stats* stats);
/* solve_bsr3_pcg_gpu (gpu.rs:775-786): 3x3 block CSR, 9 row-major values per
 * block, 3x3 block-Jacobi with caller-provided dinv_flat [n_nodes][9]      */
int fh_solve_bsr3_pcg(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                      const double* vals_flat, const double* dinv_flat, const double* b,
                      const double* x0, double rtol, uint64_t max_iter, int device,
                      double* x_out, fh_solve_stats* stats);
/* GpuBsr3Context::new / spmv / jacobi_smooth / device_name / nnz_blocks */
int fh_bsr3_ctx_create(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                       const double* vals_flat, const double* dinv_flat, int device,
                       fh_bsr3_ctx** out);
int fh_bsr3_ctx_spmv(fh_bsr3_ctx* c, const double* x, double* y);
int fh_bsr3_ctx_jacobi_smooth(fh_bsr3_ctx* c, const double* r, double* z, uint64_t sweeps,
                              double omega, int zero_initial_z);
int fh_bsr3_ctx_device_name(const fh_bsr3_ctx* c, char* buf, size_t len);
int fh_bsr3_ctx_nnz_blocks(const fh_bsr3_ctx* c, uint64_t* nnzb);
int fh_bsr3_ctx_destroy(fh_bsr3_ctx* c);
/* spmv_bsr3_multi_gpu (gpu.rs:662-669): contiguous block-row ranges over up
 * to max_gpus devices; *n_used receives the number of devices used         */
int fh_spmv_bsr3_multi(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                       const double* vals_flat, const double* x, uint64_t max_gpus, double* y,
                       int* n_used);

/* -------------------------------------------------------------- drivers */
/* poisson_solver.rs:26-59 / 99-132 solve_poisson_2d/3d: source at the
 * degree-2 points fq [nc][nq], boundary values bvals at the boundary dofs
 * (ascending, as fh_mesh_boundary_dofs returns them)                        */
int fh_solve_poisson(fh_mesh* m, const double* fq, const double* bvals, double tol,
                     uint64_t max_iter, int precision, double* u_out, fh_solve_stats* stats);
/* elasticity_solver.rs:31-44 solve_elasticity_3d (assembly, body force,
 * Dirichlet, CG); E_cell NULL or per-cell moduli on the host.  u_dev (device,
 * 3*nv) receives the solution when non-NULL, u_out (host) otherwise.        */
int fh_solve_elasticity_3d(fh_mesh* m, double E, double nu, const double* E_cell,
                           const double f[3], uint64_t nfix, const uint64_t* dofs,
                           const double* vals, double tol, uint64_t max_iter, int precision,
                           double* u_out, double* u_dev, fh_solve_stats* stats);

/* ------------------------------------------------- adjoint and stresses */
/* out[c] = -w_c^T K0_c u_c with K0_c the E = 1 elasticity element matrix
 * (elasticity.rs:133-136); with w = u this is dJ/dE_c of the compliance
 * J = f^T u for K = sum_c E_c K0_c (new; no counterpart in the reference)   */
int fh_cell_bilinear_k0_dev(fh_mesh* m, double nu, const double* w_dev, const double* u_dev,
                            double* out_dev);
int fh_cell_bilinear_k0(fh_mesh* m, double nu, const double* w, const double* u, double* out);
/* replaces compute_l2_error_2d / compute_l2_error_3d (poisson_solver.rs:62-96,
   135-172): u [nv] nodal values, u_exact_q [nc][nq] the exact solution at the
   degree-3 points (fh_quad_points(m, 3, .)); a degenerate cell is an error */
int fh_l2_error(fh_mesh* m, const double* u, const double* u_exact_q, double* out);
/* elasticity.rs:230-320 compute_element_stresses: per cell stress[9],
 * strain[9], von_mises, trace (20 doubles)                                  */
int fh_element_stresses(fh_mesh* m, const double* u, double E, double nu, double* out);
/* elasticity.rs:323-360 compute_vertex_von_mises */
int fh_vertex_von_mises(fh_mesh* m, const double* vm_cell, double* out);

/* ------------------------------------------------ batched multi-RHS solve
 * New (the reference solves one system per solve_poisson_2d call,
 * poisson_solver.rs:26): n_inst independent unit_square(N_i, N_i) Poisson P1
 * problems (generators.rs:24-55, u = 0 on the boundary) with nrhs (1..4)
 * right-hand sides each, b = M f (M the P1 mass matrix, mass.rs:10-44; f the
 * nodal loads f_nodal[sum_i (N_i+1)^2][nrhs]).  Every instance is solved by
 * Jacobi-preconditioned block CG (O'Leary; nrhs = 1 is cg.rs:40-149) in one
 * persistent launch over a device-side work queue, each column to
 * ||r||/||b|| < tol on its recursive residual; converged columns leave the
 * block.  FH_PREC_FP64 reads FP64 matrix values, the mixed modes an FP32 copy
 * (exact for this operator); vectors and reductions are FP64.  u_out has
 * f_nodal's layout, iters_out[n_inst][nrhs] the block iteration at which each
 * column converged.  stats: iterations = the largest count, replacements =
 * block iterations summed over instances, true_rel_residual = the largest
 * true FP64 residual of any column, gflops = unknowns x columns x iterations
 * per second / 1e9, spmv_bytes / iter_bytes = algorithmic bytes of one
 * product pass / block iteration over the whole batch.                       */
int fh_batched_poisson_block_cg(uint64_t n_inst, const uint64_t* sizes, uint64_t nrhs,
                                const double* f_nodal, double tol, uint64_t max_iter,
                                int precision, int device, double* u_out,
                                uint64_t* iters_out, fh_solve_stats* stats);

/* ---------------- modal analysis (the CG's consumer in eigen.rs) ---------------- */
/* replaces assemble_elasticity_mass_3d (mass.rs:86-126): 3x3-block vector mass
   M_{(i,a),(j,a)} = density * vol * (1/10 diagonal, 1/20 off) per tetrahedron */
int fh_assemble_elasticity_mass(fh_mesh* m, double density, fh_matrix** out);
/* replaces assemble_elasticity_mass_p2_3d (mass.rs:136-180) on a P2 mesh given
   as node positions [n_nodes][3] and cells [nc][10] (4 corners, then edges
   01, 12, 02, 03, 13, 23): M_{(i,a),(j,a)} = density |det J| int phi_i phi_j */
int fh_assemble_elasticity_mass_p2(uint64_t n_nodes, const double* nodes, uint64_t nc,
                                   const uint64_t* cells, double density, int device,
                                   fh_matrix** out);
/* replaces fold_dirichlet (eigen.rs:51-78): constrained rows and columns
   removed, constrained diagonal 1 (unit_diag) or 0; the CSC export drops exact
   zeros as the reference does */
int fh_fold_dirichlet(const fh_matrix* a, uint64_t nfix, const uint64_t* dofs, int unit_diag,
                      fh_matrix** out);
/* replaces solve_vibration_modes (eigen.rs:288-300) and, with the device CG
   as the inner solve, solve_vibration_modes_with_inner (eigen.rs:94-266):
   inverse power iteration with M-deflation; the inner solve K w = s M v runs
   from x0 = v to inner_tol in at most inner_max_iter iterations (the
   reference: 1e-9, 1000) with the given FH_PREC_* precision; shapes
   [num_modes][n] may be NULL.  Errors of the inner solve are returned. */
int fh_vibration_modes(const fh_matrix* K, const fh_matrix* M, uint64_t nfix,
                       const uint64_t* fixed, uint64_t num_modes, uint64_t max_power_iters,
                       double tol, double inner_tol, uint64_t inner_max_iter, int precision,
                       double* freq_hz, double* omega, double* eigenvalue, double* shapes,
                       uint64_t* inner_iterations);
/* The caller's inner solve: x[n] solving K x = b[n] in the Dirichlet frame
   (constrained rows and columns of K zero, unit diagonal on them), starting
   from x0[n], the previous iterate.  Returns 0 on success; any other value
   stops the iteration and the library returns FH_ERR_BAD_INPUT naming it. */
typedef int (*fh_inner_solve_fn)(uint64_t n, const double* b, const double* x0, double* x,
                                 void* ctx);
/* replaces solve_vibration_modes_with_inner (eigen.rs:94-266) with the inner
   closure (eigen.rs:101) as a function pointer plus context: the power
   iteration, deflation and Rayleigh quotients run on the device; b and x0 are
   copied to host memory for each call of `inner` and x back.  The library
   calls `inner` on the calling thread, once per power iteration. */
int fh_vibration_modes_with_inner(const fh_matrix* K, const fh_matrix* M, uint64_t nfix,
                                  const uint64_t* fixed, uint64_t num_modes,
                                  uint64_t max_power_iters, double tol, fh_inner_solve_fn inner,
                                  void* inner_ctx, double* freq_hz, double* omega,
                                  double* eigenvalue, double* shapes);

/* ------------------------------- boundary facets and boundary (ds) integrals ------------------------------- */
/* replaces extract_boundary_facets (mesh.rs:79-113 triangles, 145-210 tetrahedra):
   per facet its vertices ([count][dim]; 2D the sorted edge, 3D the face with
   an outward normal (v1 - v0) x (v2 - v0)), cell_idx and local_facet_idx, in
   the order of the sorted vertex tuples.  Query-then-fill: with all three
   outputs NULL only *count is set; otherwise *count is the capacity (0: do not
   check) and receives the facet count. */
int fh_mesh_boundary_facets(fh_mesh* m, uint64_t* vertices, uint64_t* cell_idx,
                            uint64_t* local_idx, uint64_t* count);
/* the points where boundary.rs samples its closures, facet by facet: on a
   triangle mesh the 2-point Gauss points (1 - s) v0 + s v1 of each edge
   ([nf][2][2], boundary.rs:29-36), on a tetrahedral mesh the edge midpoints
   v0v1, v1v2, v2v0 of each face ([nf][3][3], boundary.rs:77-93) */
int fh_boundary_points(fh_mesh* m, uint64_t nf, const uint64_t* facets, double* points);
/* replaces assemble_neumann_2d (boundary.rs:8-46): g[nf][2] the flux at the
   fh_boundary_points of each facet; rhs_out [nv] */
int fh_assemble_neumann_2d(fh_mesh* m, uint64_t nf, const uint64_t* facets, const double* g,
                           double* rhs_out);
/* replaces assemble_surface_traction_3d (boundary.rs:49-108): t[nf][3][3] the
   traction at the face's edge midpoints; rhs_out [3 nv] */
int fh_assemble_surface_traction_3d(fh_mesh* m, uint64_t nf, const uint64_t* facets,
                                    const double* t, double* rhs_out);
/* replaces assemble_robin_matrix_2d (boundary.rs:111-139); a facet that
   repeats a vertex is rejected (FH_ERR_BAD_INPUT) */
int fh_assemble_robin_matrix_2d(fh_mesh* m, uint64_t nf, const uint64_t* facets, double gamma,
                                fh_matrix** out);

/* ------------- transient heat (the CG's consumer in transient.rs / heat_solver.rs) ------------- */
/* replaces TimeSteppingScheme (transient.rs:12-18) */
#define FH_SCHEME_BACKWARD_EULER 0
#define FH_SCHEME_CRANK_NICOLSON 1

/* replaces the heat_source closure FnMut(f64, &[f64; 2]) -> f64 (heat_solver.rs:32-34):
   called with (t, vertex coordinates, ctx) */
typedef double (*fh_heat_source_fn)(double t, const double* point, void* ctx);

typedef struct fh_transient_stats {
    uint64_t steps;
    uint64_t cg_iterations;     /* over all steps                               */
    uint64_t cg_iterations_low; /* of which on FP32 matrix values (FH_PREC_MIXED) */
    uint64_t cg_iterations_max; /* largest count of one step                     */
    double solve_ms;            /* wall time inside the per-step CG solves       */
    double wall_ms;             /* whole call, assembly included                 */
    int device;
    int precision;
} fh_transient_stats;

/* replaces faer's SparseColMat at this boundary: CSC raw parts (column
   pointers, row indices, values; gpu.rs:1552-1583 reads the same parts) ->
   device operator, for the entries below that take fh_matrix operands */
int fh_matrix_from_csc(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                       const double* vals, int device, fh_matrix** out);
/* replaces add_scaled_matrices (transient.rs:72-97): a + alpha b on the union
   pattern; operands on one pattern combine on the device, others through the
   host union of their CSC parts (a's entry, then alpha times b's, summed) */
int fh_add_scaled_matrices(const fh_matrix* a, const fh_matrix* b, double alpha,
                           fh_matrix** out);
/* replaces TransientHeatSolver::step (transient.rs:32-68): BE solves
   (M + dt K) u_next = M u_n + dt f_next, CN (M + dt/2 K) u_next =
   (M - dt/2 K) u_n + dt/2 (f_n + f_next).  solve_direct (direct.rs:26-53, a
   dense LU) is replaced by the device CG from x0 = u_n to ||r||/||b|| < tol,
   so M + theta dt K must be symmetric positive definite; a CG failure is
   returned (FH_ERR_NOT_CONVERGED / FH_ERR_BREAKDOWN).  Host vectors of n. */
int fh_transient_heat_step(const fh_matrix* M, const fh_matrix* K, int scheme, double dt,
                           const double* u_n, const double* f_n, const double* f_next,
                           double tol, uint64_t max_iter, int precision, double* u_next,
                           fh_solve_stats* stats);
/* replaces solve_transient_heat_2d (heat_solver.rs:23-108) on a triangle
   mesh: P1 mass and diffusivity-scaled stiffness, the Dirichlet list (the
   DirichletBC entries concatenated in order) written into u before and after
   every step, source sampled at the vertices at t_n and t_{n+1} (NULL: zero).
   history [(num_steps+1)][nv] may be NULL; final_out [nv]. */
int fh_solve_transient_heat_2d(fh_mesh* m, double diffusivity, const double* u_initial,
                               uint64_t nbc, const uint64_t* bc_dofs, const double* bc_vals,
                               double dt, uint64_t num_steps, int scheme, fh_heat_source_fn source,
                               void* source_ctx, double tol, uint64_t max_iter, int precision,
                               double* history, double* final_out, fh_transient_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
