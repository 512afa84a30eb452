// Launch interfaces of the HIP kernels (host-callable, defined in *.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace dfa {

// Which element tensors one pass computes: the bilinear form (A), the linear form (b) and
// whether Dirichlet rows/columns are eliminated per cell (SystemAssembler::apply_bc).
struct P1Pass {
  int gdim = 3;
  int a_form = -1;            // DFA_FORM_P1_POISSON / DFA_FORM_P1_MASS / -1 (no matrix)
  int L_form = -1;            // DFA_FORM_P1_LOAD / -1 (no vector)
  bool want_A = false, want_b = false, bc = false, add = false;
  const double* x = nullptr;
  const double* kappa = nullptr;   // a_form coefficient (P1 nodal)
  const double* f = nullptr;       // L_form coefficient (P1 nodal)
  const uint8_t* bc_mask = nullptr;
  const double* bc_value = nullptr;
  double* values = nullptr;
  double* b = nullptr;
};

struct P1AtomicTopo {
  int64_t num_cells = 0;
  const int32_t* cells = nullptr;      // [nc x nvc] vertex ids
  const int32_t* cell_dofs = nullptr;  // [nc x nvc]
  const int32_t* coeff_dofs = nullptr; // [nc x nvc]
  const int32_t* row_ptr = nullptr;
  const int32_t* col = nullptr;
};

// Graph-coloured scatter: cells stored in colour order (cells of one colour share no dof),
// each with its vertex ids, dofs, coefficient dofs and the rank of column d[b] in row d[a]
// (one byte per pair). One launch per colour over [color_ptr[k], color_ptr[k+1]).
struct P1ColorTopo {
  int64_t begin = 0, end = 0;          // cell range of the colour being launched
  const int32_t* cells = nullptr;      // [nc x nvc], colour order
  const int32_t* cell_dofs = nullptr;  // [nc x nvc], colour order
  const int32_t* coeff_dofs = nullptr; // [nc x nvc], colour order (may alias cell_dofs)
  const uint8_t* rank = nullptr;       // [nc x nvc x nvc]
  const int32_t* row_ptr = nullptr;
};

struct P1BlockTopo {
  int32_t num_blocks = 0;
  int32_t max_nodes = 0, max_rows = 0, max_nnz = 0, max_row_len = 0;
  int32_t threads = 256;               // workgroup size of the blocked kernels (256 or 512)
  bool wide_nodes = true;              // 48-byte LDS node records (kappa, f, Dirichlet value inline)
  int32_t max_cells = 0;               // largest cell-appearance list of a block
  // edge chains (tetrahedra): kChainWords uint32 per chain, chain_ptr[nb + 1]
  int32_t max_chains = 0;
  int32_t chain_threads = 256;
  const int32_t* chain_ptr = nullptr;
  const uint32_t* chain_rec = nullptr;
  // Fixed-stride per-block slots of the edge-chain kernels, padded to the largest block, so
  // that a block's node ids and owned-row words are addressed from the block index alone
  // (no dependent load of the block's offsets before them):
  int32_t slot_nodes = 0, slot_rows = 0;
  const int4* blk_hdr = nullptr;         // [nb] first chain, chains, nodes, owned rows
  const int32_t* slot_vertex = nullptr;  // [nb x slot_nodes] (padding repeats the block's node 0)
  const int32_t* slot_kdof = nullptr;    // [nb x slot_nodes]; NULL when kdof_is_vertex
  const int2* slot_row = nullptr;        // [nb x slot_rows] CSR start, length | diagonal << 8
  const int32_t* slot_rdof = nullptr;    // [nb x slot_rows] dof of each owned row
  bool kdof_is_vertex = false;
  const int32_t* node_ptr = nullptr;
  const int32_t* node_vertex = nullptr;
  const int32_t* node_kdof = nullptr;
  const int32_t* node_dof = nullptr;
  const int32_t* cell_ptr = nullptr;
  // per cell appearance, interleaved: [2k] 4 x 16-bit local node ids, [2k+1] 16 x 4-bit ranks
  const uint64_t* cell_rec = nullptr;
  const uint64_t* cell_rank_wide = nullptr;  // 8-bit ranks, 2 words per appearance (rows > 16)
  const int32_t* row_ptr_b = nullptr;
  const int32_t* row_dof = nullptr;
  const uint32_t* row_info = nullptr;
  const uint8_t* row_diag = nullptr;
  const int32_t* blk_nnz = nullptr;
  const int32_t* run_ptr = nullptr;
  const uint64_t* runs = nullptr;
  const int32_t* csr_row_ptr = nullptr;
};

hipError_t launch_p1_atomic(const P1AtomicTopo& t, const P1Pass& p, hipStream_t s);
hipError_t launch_p1_colored(const P1ColorTopo& t, const P1Pass& p, hipStream_t s);
// Colour-order copies of a handle's cell arrays and the column ranks of the coloured scatter:
// for i in [0, nc): cell perm[i]'s vertex ids, dofs and coefficient dofs go to position i.
// Fails (hipErrorInvalidValue) if a row holds more than 256 entries.
hipError_t launch_p1_color_gather(int64_t num_cells, int32_t nvc, const int32_t* perm,
                                  const int32_t* cells, const int32_t* cell_dofs,
                                  const int32_t* coeff_dofs, const int32_t* row_ptr,
                                  const int32_t* col, int32_t* out_cells, int32_t* out_dofs,
                                  int32_t* out_kdofs, uint8_t* out_rank, hipStream_t s);
hipError_t launch_p1_blocked(const P1BlockTopo& t, const P1Pass& p, hipStream_t s);
size_t p1_blocked_lds_bytes(const P1BlockTopo& t, const P1Pass& p);
// Name and template arguments of the kernel launch_p1_blocked runs for (t, p).
// True when launch_p1_blocked runs the edge-chain kernel for this pass.
bool p1_chains_apply(const P1BlockTopo& t, const P1Pass& p);
std::string p1_blocked_kernel_name(const P1BlockTopo& t, const P1Pass& p);
hipError_t launch_fill_zero(double* p, int64_t n, hipStream_t s);

// ---- P1 exterior facets -----------------------------------------------------------------
struct FacetTopo {
  int64_t num_facets = 0;
  const int32_t* facet_cell = nullptr;
  const int8_t* facet_local = nullptr;  // facet opposite this local vertex
  const int32_t* cells = nullptr;
  const int32_t* cell_dofs = nullptr;
  const int32_t* coeff_dofs = nullptr;
  const int32_t* row_ptr = nullptr;
  const int32_t* col = nullptr;
};

struct FacetPass {
  int gdim = 3;
  bool has_a = false;              // alpha*u*v*ds is part of the bilinear form
  bool has_b = false;              // g*v*ds is part of the linear form
  double alpha = 0.0;
  const double* x = nullptr;
  const double* g = nullptr;
  const uint8_t* mask = nullptr;   // per exterior facet, NULL = all
  const uint8_t* bc_mask = nullptr;
  const double* bc_value = nullptr;
  double* values = nullptr;        // NULL: matrix not written (its Dirichlet lifting still is)
  double* b = nullptr;
  bool vp2_traction = false;       // vector-P2 inner(t, v)*ds, g = the 3 traction components
};
hipError_t launch_facets(const FacetTopo& t, const FacetPass& p, hipStream_t s);

// ---- device CSR consumer side ------------------------------------------------------------
// Streamed SpMV: a workgroup takes one row block (<= kSpmvBudget nonzeros, <= kSpmvMaxRows
// rows, from build_spmv_blocks) and stages its products in LDS.
constexpr int kSpmvBudget = 4096;
constexpr int kSpmvMaxRows = 256;
struct CsrView {
  int64_t nrows = 0, nnz = 0;
  const int32_t* row_ptr = nullptr;
  const int32_t* col = nullptr;
  const double* values = nullptr;
  const int32_t* blocks = nullptr;  // nblocks + 1 pairs (first row, row_ptr of that row)
  int64_t nblocks = 0;
  // narrow columns: per row block the smallest column (or -1: the block reads col) and the
  // columns as 16-bit offsets from it
  const int32_t* blk_col0 = nullptr;
  const uint16_t* col16 = nullptr;
  // vector P2 expanded from the node graph (node_layout 1 = component-major, 2 = node-major):
  // every node entry (n, m) is a 3 x 3 block of the dof rows; 0 = a general CSR
  int32_t node_layout = 0;
  int64_t node_nnz = 0;
  const int32_t* node_rp = nullptr;      // nrows / 3 + 1
  const int32_t* node_col = nullptr;     // node_nnz
  const int32_t* node_chunks = nullptr;  // node_nchunks + 1 pairs (first node, first node entry)
  int64_t node_nchunks = 0;
};
// Node-blocked SpMV: a workgroup takes a chunk of consecutive nodes with at most kNodeBudget
// node entries and kNodeMaxNodes nodes (from build_node_chunks).
constexpr int kNodeBudget = 1024;
constexpr int kNodeMaxNodes = 128;
hipError_t launch_spmv(const CsrView& A, double alpha, const double* x, double beta, double* y,
                       hipStream_t s);
hipError_t launch_diagonal(const CsrView& A, double* d, hipStream_t s);
// Dirichlet conditions applied to an assembled square system as SystemAssembler's per-cell
// elimination sums up (cnt: cells per dof, from launch_dof_cell_count); b may be NULL.
hipError_t launch_dof_cell_count(int64_t nk, const int32_t* cell_dofs, int32_t* cnt, hipStream_t s);
hipError_t launch_dirichlet_rows(const CsrView& A, const int32_t* tmap, const uint8_t* mask, const double* g,
                                 const int32_t* cnt, double* values, double* b, hipStream_t s);
// tmap may be NULL (A_ji found by binary search). tmap[k] = the slot of A_ji for the entry
// k = A_ij (-1 when the pattern has none).
hipError_t launch_transpose_map(const CsrView& A, int32_t* tmap, hipStream_t s);
// The same for a vector-P2 structure expanded from the node graph (A.node_layout != 0): the
// Dirichlet node rows, then the free rows next to them pull their Dirichlet columns; scratch:
// two bytes per node.
hipError_t launch_dirichlet_nodes(const CsrView& A, const uint8_t* mask, const double* g, const int32_t* cnt,
                                  uint8_t* scratch, double* values, double* b, hipStream_t s);
// Conjugate gradients on the device: scalars of one solve (device memory); the partials
// buffer holds 3 * kCgMaxGroups doubles.
struct CgState {
  double rzs[2];  // r.z at the start of iteration it in rzs[it & 1]
  double rr, bb, alpha, beta, tol2;
  int64_t done;   // iteration at which |r|^2 <= tol2 (-1: not yet)
  int32_t broke;  // p.Ap = 0 with r != 0
  int32_t pad;
};
constexpr int kCgMaxGroups = 512;
hipError_t launch_cg_init(int64_t n, const double* b, const double* r, const double* dinv, double* p,
                          double* part, CgState* st, hipStream_t s);
hipError_t launch_cg_iteration(const CsrView& A, int64_t it, double* x, double* p, double* r, double* ap,
                               const double* dinv, double* part, CgState* st, hipStream_t s);
hipError_t launch_cg_jacobi(int64_t n, double* d, hipStream_t s);
// The narrow-column arrays of the streamed SpMV (blk_col0 [nblocks], col16 [nnz]).
hipError_t launch_narrow_columns(const CsrView& A, int32_t* blk_col0, uint16_t* col16, hipStream_t s);
// Host-side block insertion (GenericMatrix::add_local): slot[i*n+j] = position of
// (rows[i], cols[j]) in the CSR (-1 if the pattern lacks it), then values[slot] += block.
hipError_t launch_block_slots(int32_t m, const int32_t* rows, int32_t n, const int32_t* cols,
                              const int32_t* row_ptr, const int32_t* col, int32_t* slot,
                              hipStream_t s);
hipError_t launch_block_add(int64_t nk, const int32_t* slot, const double* block, double* values,
                            hipStream_t s);
// out[i] = mask[i] ? x0[i] - g[i] : 0 (the boundary values of an incremental solve).
hipError_t launch_dirichlet_x0(int64_t n, const uint8_t* mask, const double* g, const double* x0,
                               double* out, hipStream_t s);

// ---- vector P2 elasticity --------------------------------------------------------------
struct VP2Topo {
  int64_t num_cells = 0;
  const int32_t* cells = nullptr;      // [nc x 4]
  const int32_t* cell_dofs = nullptr;  // [nc x 30] component-major local order
  const int32_t* coeff_dofs = nullptr; // [nc x 4] P1 dofs of mu, lmbda
  const int32_t* row_ptr = nullptr;
  const uint8_t* rank = nullptr;       // [nc x 10 x 10] rank of node b among node a's neighbours
  bool component_major = true;         // CSR column order of a row: component-major or node-major
};

struct VP2Pass {
  bool want_A = false, want_b = false, bc = false;
  bool add = false;  // add to the tensors instead of overwriting them (blocked scatter)
  const double* x = nullptr;
  const double* mu = nullptr;
  const double* lmbda = nullptr;
  double f3[3] = {0, 0, 0};
  const uint8_t* bc_mask = nullptr;
  const double* bc_value = nullptr;
  double* values = nullptr;
  double* b = nullptr;
};

hipError_t launch_vp2_atomic(const VP2Topo& t, const VP2Pass& p, hipStream_t s);

struct VP2BlockTopo {
  int32_t num_blocks = 0;
  int32_t max_verts = 0, max_owned = 0, max_vals = 0;
  bool component_major = true;
  const int32_t* vert_ptr = nullptr;
  const int32_t* vert_id = nullptr;
  const int32_t* vert_kdof = nullptr;
  const int32_t* own_ptr = nullptr;
  const int32_t* own_rows = nullptr;
  const uint32_t* own_info = nullptr;
  const int32_t* task_ptr = nullptr;
  const uint64_t* task = nullptr;
  const int32_t* task_cell = nullptr;
  const int32_t* task_nvert = nullptr;  // [nb] vertex-node work items (first in each block)
  const int32_t* blk_vals = nullptr;
  const int32_t* row_ptr = nullptr;
  const int32_t* cell_dofs = nullptr;
};
hipError_t launch_vp2_blocked(const VP2BlockTopo& t, const VP2Pass& p, hipStream_t s);
size_t vp2_blocked_lds_bytes(const VP2BlockTopo& t, bool want_b);
// Name and template arguments of the kernel launch_vp2_blocked / launch_vp2_atomic runs.
std::string vp2_kernel_name(bool blocked, const VP2Pass& p);

// ---- adjoint -------------------------------------------------------------------------
struct AdjointP1 {
  int gdim = 3;
  int64_t num_cells = 0;
  const int32_t* cells = nullptr;
  const int32_t* cell_dofs = nullptr;
  const int32_t* coeff_dofs = nullptr;
  const double* x = nullptr;
  const double* kappa = nullptr;
  const double* f = nullptr;
  const double* u = nullptr;
  double* J = nullptr;
  double* dJdk = nullptr;
  double* dJdx = nullptr;
};
hipError_t launch_adjoint_p1(const AdjointP1& a, hipStream_t s);
hipError_t launch_p1_adjoint_blocked(const P1BlockTopo& t, const AdjointP1& a, hipStream_t s);

struct AdjointVP2 {
  int64_t num_cells = 0;
  const int32_t* cells = nullptr;
  const int32_t* cell_dofs = nullptr;
  const int32_t* coeff_dofs = nullptr;
  const double* x = nullptr;
  const double* mu = nullptr;
  const double* lmbda = nullptr;
  double f3[3] = {0, 0, 0};
  const double* u = nullptr;
  double* J = nullptr;
  double* dJdmu = nullptr;
  double* dJdlmbda = nullptr;
  double* dJdx = nullptr;
};
hipError_t launch_adjoint_vp2(const AdjointVP2& a, hipStream_t s);

}  // namespace dfa
