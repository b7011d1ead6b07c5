// fh_internal.hpp — device-resident objects behind the opaque C handles.
#pragma once
#include <map>
#include <memory>

#include "fh_common.hpp"

namespace fh {

constexpr int SLICE = 64;  // rows per SELL slice == one wavefront

// Position of component j (row-major within the bs x bs block) of the block
// stored at SELL entry e0 + lane, where e0 = slice_off[slice] + k * SLICE is
// the entry of lane 0 at step k.  Scalar matrices: e0 + lane.  3x3 blocks: the
// nine components of a step are 576 consecutive values; components 0-7 are
// stored as lane pairs (2p, 2p+1) at 128 p + 2 lane, component 8 at 512 + lane,
// so a wave reads a step as four 16-byte and one 8-byte load per lane, each
// over a contiguous range.
__host__ __device__ inline uint64_t vpos(uint64_t e0, int j, uint32_t lane, int bs2) {
    if (bs2 == 9)
        return e0 * 9 + (j < 8 ? (uint64_t)(j >> 1) * 128 + 2 * (uint64_t)lane + (j & 1)
                               : 512 + (uint64_t)lane);
    return e0 * bs2 + (uint64_t)j * SLICE + lane;
}

// A symbolic sparsity pattern over "block rows" (a vertex or P2 dof), with the
// CSR arrays (for export) and the SELL-64 slice layout (for the kernels).
struct Pattern {
    int dev = 0;
    uint32_t nrows = 0;     // block rows
    uint64_t nnzb = 0;      // stored blocks (CSR)
    uint32_t nslices = 0;
    uint64_t sell_entries = 0;  // padded entries (64 * sum of slice widths)
    DBuf<uint32_t> row_ptr;     // [nrows+1] CSR
    DBuf<uint32_t> cols;        // [nnzb] CSR, ascending per row
    DBuf<uint32_t> slice_off;   // [nslices+1] entry offset of each slice
    DBuf<uint32_t> sell_cols;   // [sell_entries]
    DBuf<uint32_t> row_len;     // [nrows]
    uint32_t max_width = 0;
    // slice ranges of the 8 XCDs for the CG product, computed at the first
    // solve on this pattern (fh_cg.hip xcd_map)
    bool xcd_ready = false;
    uint32_t xcd_b[9] = {};
    uint32_t xcd_interleave = 0;
    uint32_t xcd_blocks = 0;
};

// Dof -> cell adjacency of a cell-dof table (K dofs per cell), every dof's
// cells in ascending order, plus for each (dof, cell) pair the position of
// every cell dof inside that dof's pattern row ("slot map").
struct Adjacency {
    uint32_t ndofs = 0;
    int K = 0;
    DBuf<uint32_t> ptr;    // [ndofs+1]
    DBuf<uint32_t> ent;    // cell * K + local
    DBuf<uint8_t> slots;     // [entries][K] position in the pattern row (rows of <= 255 blocks)
    DBuf<uint16_t> slots16;  // the same as u16, instead, for rows of 256-65535 blocks
};

// A solve's workspace and its captured iteration graphs (FP64 and FP32
// operator), kept on the matrix for the next solve with the same kernel
// configuration: repeated solves of one operator skip the allocation and the
// graph capture.  Defined in fh_cg.hip.
struct CgCache {
    DBuf<uint8_t> ws;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    CgCache() = default;
    CgCache(const CgCache&) = delete;
    CgCache& operator=(const CgCache&) = delete;
    ~CgCache();
};

struct MeshP2 {
    uint32_t nedges = 0;
    DBuf<uint32_t> cell_dofs;  // [nc][10]
    std::unique_ptr<Adjacency> adj;
    std::shared_ptr<Pattern> pat;
};

}  // namespace fh

struct fh_mesh {
    int dev = 0;
    int dim = 2;  // 2 (triangles) or 3 (tetrahedra)
    int k = 3;    // vertices per cell
    uint32_t nv = 0, nc = 0;
    fh::DBuf<double> verts;   // [nv][dim]
    fh::DBuf<uint32_t> cells; // [nc][k]
    // lazily built
    std::unique_ptr<fh::Adjacency> adj;
    std::shared_ptr<fh::Pattern> pat;   // vertex graph pattern
    std::unique_ptr<fh::MeshP2> p2;
    int64_t n_degenerate = -1;  // cells with |det J| < 1e-14 (counted once, lazily)
};

struct fh_matrix {
    int dev = 0;
    int bs = 1;            // block size (1 scalar, 3 elasticity)
    uint32_t n = 0;        // scalar rows
    std::shared_ptr<fh::Pattern> pat;
    fh::DBuf<double> vals;   // SELL layout, [entries][bs*bs] interleaved by lane
    fh::DBuf<float> vals32;  // FP32 copy (lazily made)
    // 3x3 blocks: per SELL entry, the block's last value and its column in one
    // record (the FP64 CG product loads both with one instruction) [entries][3 words]
    fh::DBuf<uint32_t> tail64;
    // 3x3 blocks, FP32 product of the CG: per SELL step of 64 entries, 640 words:
    // values 0-3 of each lane as a 16-byte group (words 4 lane ..), values 4-7
    // likewise (256 + 4 lane ..), then value 8 and the column (512 + 2 lane ..)
    fh::DBuf<uint32_t> quad32;
    fh::DBuf<double> rowsum64;  // block row sums, lane-interleaved [slice][bs*bs][64] (difference form)
    fh::DBuf<float> rowsum32;   // FP32 copy of rowsum64
    fh::DBuf<uint8_t> fixed; // [n] Dirichlet mask after condensation (may be empty)
    uint64_t nnz_scalar = 0; // canonical CSC nnz (export)
    bool condensed = false;
    bool folded = false;     // Dirichlet fold: the export drops every exact zero
    bool comp_diag = false;  // blocks couple equal components only (elasticity mass):
                             // the export leaves out the off-diagonal components
    fh::DBuf<uint32_t> fixed_count;  // duplicates in the prescribed list (diag value)
    uint64_t skipped_cells = 0;      // cells the assembly skipped (|det J| < 1e-14)
    // solve workspaces and graphs by kernel configuration (fh_cg.hip run_cg)
    std::map<uint64_t, std::unique_ptr<fh::CgCache>> cg_cache;
};

namespace fh {

// quadrature rules of quadrature.rs:83-219 (points in reference coordinates)
struct QuadTab {
    double pts[5][3];
    double w[5];
    int n;
};
QuadTab quad_table(int dim, int degree);
std::vector<uint64_t> mask_to_sorted(int dev, const uint8_t* mask, size_t n, hipStream_t s);

// mesh / pattern (fh_mesh.hip, fh_pattern.hip)
Adjacency& mesh_adjacency(fh_mesh* m);
std::shared_ptr<Pattern> mesh_pattern(fh_mesh* m);
MeshP2& mesh_p2(fh_mesh* m);
std::unique_ptr<Adjacency> build_adjacency(int dev, const uint32_t* cell_dofs, uint32_t nc,
                                           int K, uint32_t ndofs, hipStream_t s);
std::shared_ptr<Pattern> build_pattern(int dev, const uint32_t* cell_dofs, int K,
                                       Adjacency& adj, hipStream_t s);
void build_slots(const uint32_t* cell_dofs, int K, Adjacency& adj, const Pattern& pat,
                 hipStream_t s);
void boundary_vertex_mask(fh_mesh* m, uint8_t* mask_dev, hipStream_t s,
                          const uint32_t* cd10, uint8_t* mask2);

// matrices (fh_assemble.hip)
fh_matrix* assemble_p1(fh_mesh* m, int kind);  // 0 stiffness, 1 mass
fh_matrix* assemble_elasticity(fh_mesh* m, double E, double nu, const double* Ecell_dev);
void assemble_elasticity_into(fh_mesh* m, double E, double nu, const double* Ecell_dev,
                              fh_matrix* a, hipStream_t s);
void rhs_p1_dev(fh_mesh* m, const double* fq, double* b, hipStream_t s);
void body_force_dev(fh_mesh* m, const double f[3], double* b, hipStream_t s);
// fold_diag >= 0: the Dirichlet fold of eigen.rs:51-78 (constrained diagonal =
// fold_diag, exact zeros dropped on export) instead of the condensation
fh_matrix* apply_dirichlet(const fh_matrix* a, const double* b_in, uint64_t nfix,
                           const uint64_t* dofs, const double* vals, double* b_out,
                           hipStream_t s, int fold_diag = -1);
fh_matrix* assemble_elasticity_mass(fh_mesh* m, double density);
fh_matrix* matrix_from_csr(int dev, uint32_t nrows, int bs, const uint32_t* row_ptr,
                           const uint32_t* cols, const double* vals);
fh_matrix* matrix_from_csc(int dev, uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                           const double* vals);
void ensure_vals32(fh_matrix* a);
// cells of the mesh that the assembly skips (jacobian.rs:29-31: |det J| < 1e-14)
uint64_t mesh_degenerate_cells(fh_mesh* m);
// scalar CSR of the operator the matrix exports (condensed couplings, folded
// zeros and the off-diagonal components of a component-diagonal mass left out)
struct HostCsr {
    std::vector<uint64_t> rp;
    std::vector<uint32_t> ci;
    std::vector<double> v;
};
HostCsr export_scalar_csr(const fh_matrix* a);
void assemble_p2_mfma(fh_mesh* m, fh_matrix* a, bool fill64, hipStream_t s);
// P1 elasticity on the MFMA units (fh_element_mfma.hip): FP32 element
// matrices, summed per row into a->vals32 (and, with fill64, their FP64
// widening into a->vals)
void assemble_elasticity_mfma(fh_mesh* m, double E, double nu, const double* Ecell_dev,
                              fh_matrix* a, bool fill64, hipStream_t s);
// per-cell gradient records [nc][16] (fh_assemble.hip k_elas_cell_geom)
void elas_cell_records(fh_mesh* m, double nu, const double* Ecell_dev, double* G, hipStream_t s);
// sum FP32 elasticity element blocks (Ke8 [nc][4][4][8], Ke9 [nc][4][4]) into
// the SELL values of pattern p (vals64: also as doubles, when not null)
void elas_gather_f32(fh_mesh* m, const Pattern& p, const float* Ke8, const float* Ke9,
                     float* vals32, double* vals64, hipStream_t s);
// sum P2 element matrices Ke [nc][10][10] (FP64, or FP32 when !f64) into SELL
// values; FP32 sums are also written as doubles to vals64 when it is not null
void p2_gather(fh_mesh* m, const void* Ke, bool f64, void* vals, double* vals64, hipStream_t s);
uint64_t sell_value_count(const fh_matrix* a);  // entries * bs*bs

// SpMV (fh_cg.hip)
void spmv_f64(const fh_matrix* a, const double* x, double* y, int order, hipStream_t s);
void inv_diag_f64(const fh_matrix* a, double* out, hipStream_t s);
void ensure_rowsum(fh_matrix* a, bool f32);
void ensure_tail(fh_matrix* a);
void ensure_quad32(fh_matrix* a);

// Solve (fh_cg.hip)
struct SolveArgs {
    const double* b;     // device
    const double* x0;    // device or null
    double* x;           // device out
    double tol;
    uint64_t max_iter;
    int precision;
    int order;           // FH_ORDER_*
    const double* inv_diag_override = nullptr;  // device, scalar Jacobi
    const double* block_dinv = nullptr;         // device, [nrows][9] block Jacobi
};
// ws: a caller-provided workspace of solve_workspace_bytes() (null: the
// solve allocates its own)
int solve_device(const fh_matrix* a, const SolveArgs& args, fh_solve_stats* st,
                 uint8_t* ws = nullptr);
size_t solve_workspace_bytes(const fh_matrix* a, const SolveArgs& args);

double spmv_algorithmic_bytes(const fh_matrix* a, int vbytes, int mbytes);

}  // namespace fh
