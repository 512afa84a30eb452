// Internal declarations shared by the host structure builder, the kernels and the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/dolfin_amd.h"

namespace dfa {

// Thread-local error channel behind dfa_last_error().
void set_error(const std::string& msg);
const char* last_error();

// Allocator whose resize() leaves new elements uninitialised: large host arrays that are
// filled in parallel right after allocation are first touched by the filling threads.
template <class T>
struct UninitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};

struct HostCSR {
  std::vector<int32_t> row_ptr;
  std::vector<int32_t, UninitAlloc<int32_t>> col;
  // vector P2 built from its node graph: the node-level CSR (node n = dof of component 0 for
  // the component-major layout, that dof / 3 for the node-major one) and the layout
  // (0: not built that way, 1: component-major, 2: node-major)
  int node_layout = 0;
  std::vector<int32_t> node_row_ptr;
  std::vector<int32_t, UninitAlloc<int32_t>> node_col;
};

// Sorted unique column set per row from cell_dofs x cell_dofs (+ optional diagonal):
// the structure of SparsityPatternBuilder::build + diagonal_pattern(sorted).
void build_csr(int64_t num_cells, int32_t nd, const int32_t* cell_dofs, int64_t num_rows,
               bool keep_diagonal, HostCSR& out);

// Work partition of the blocked (owner-computes) scatter for P1 elements. Rows (dofs) are
// grouped into blocks by recursive coordinate bisection; a workgroup owns one block,
// accumulates its rows in LDS from every cell touching them and writes each CSR value
// once. Cells are stored per block in block-local node numbering (uint16).
struct P1BlockPlan {
  int32_t nvc = 0;                      // vertices per cell (gdim + 1)
  int32_t num_blocks = 0;
  int32_t max_nodes = 0, max_rows = 0, max_nnz = 0, max_cells = 0, max_row_len = 0;
  bool kdof_is_vertex = false;          // coefficient dof == vertex id everywhere
  bool kdof_injective = false;          // distinct vertices carry distinct coefficient dofs
  // Local nodes 0..(owned rows - 1) are the owned rows in order; ghosts follow.
  std::vector<int32_t> node_ptr;        // [nb+1]
  std::vector<int32_t> node_vertex;     // global vertex of each local node
  std::vector<int32_t> node_kdof;       // coefficient dof of each local node
  std::vector<int32_t> node_dof;        // argument dof of each local node (BC lookups)
  std::vector<int32_t> cell_ptr;        // [nb+1]
  std::vector<uint64_t> cell_ids;       // 4 x uint16 local node ids per cell appearance
  std::vector<uint64_t> cell_rank;      // 16 x 4-bit column ranks (row a, col b) per appearance
  // when some CSR row has more than kMaxRank entries: 16 x 8-bit ranks per appearance, two
  // words (byte a * 4 + b), and cell_rank is zero
  std::vector<uint64_t> cell_rank_wide;
  std::vector<int32_t> row_ptr;         // [nb+1] owned rows per block
  std::vector<int32_t> row_dof;         // global row (dof) of each owned row
  std::vector<uint32_t> row_info;       // (local node id) | (LDS value offset << 16)
  std::vector<uint8_t> row_diag;        // position of the diagonal entry in each owned row
  std::vector<int32_t> blk_nnz;         // [nb] values per block
  // Write-out runs: maximal sequences of owned rows with consecutive dofs are one contiguous
  // CSR range. Packed as csr_start | local_offset << 32 | length << 48.
  std::vector<int32_t> run_ptr;         // [nb+1]
  std::vector<uint64_t> runs;
};

// Edge chains of a tetrahedral block plan. The cell appearances of each block are grouped
// into chains around an edge (p, q): cell k of a chain is (p, q, r_k, r_{k+1}), so
// consecutive cells share the face (p, q, r_{k+1}) and a chain of m cells touches m + 3
// nodes. One lane walks one chain: it reads every node once and sums the contributions to
// (p, q), (p, r_k), (q, r_k) and the load vector in registers before adding them to LDS.
// Record (12 x uint32): [0] p | q << 16, [1..3] r_0 .. r_5 as 16-bit pairs, [4] r_6 | m << 16,
// [5..10] one byte per emitted pair, rank(u -> v) | rank(v -> u) << 4, in the order
// (p, r_k), (q, r_k), (r_k, r_{k+1}) for k = 0..5, then (p, r_m), (q, r_m), (p, q) at
// bytes 18, 19, 20; [11] unused. Chains are sorted by length (longest first) per block.
constexpr int kChainLen = 6;
// Most nodes (owned rows and ghosts) a block of the edge-chain kernels may hold: two node
// slots per thread of a 512-thread workgroup.
constexpr int kMaxChainBlockNodes = 1024;
constexpr int kChainWords = 12;
struct P1ChainPlan {
  int32_t max_chains = 0;
  std::vector<int32_t> chain_ptr;   // [nb+1]
  std::vector<uint32_t> chain_rec;  // kChainWords per chain
};
void build_p1_chain_plan(const P1BlockPlan& plan, P1ChainPlan& chains);

// Exterior facets (facets of exactly one cell), sorted by cell then local facet index; local
// facet k is the facet opposite local vertex k. Facet keys are grouped by their smallest
// vertex (a counting sort), so the work is linear in the number of cell facets.
void build_exterior_facets(int32_t gdim, int64_t num_cells, const int32_t* cells,
                           std::vector<int32_t>& facet_cell, std::vector<int8_t>& facet_local);

// Reverse Cuthill-McKee numbering of a graph given as sorted adjacency lists (no self
// loops): old_to_new[v] = new index of v. Components in order of their lowest vertex, each
// started at a pseudo-peripheral vertex (George & Liu), breadth-first with each vertex's
// unvisited neighbours appended in ascending degree, the whole sequence reversed -- the
// ordering BoostGraphOrdering::compute_cuthill_mckee(graph, true) returns.
void rcm_ordering(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int32_t* old_to_new);

// Greedy colouring of cells in index order: cell c takes the smallest colour held by none of
// its neighbours of lower index, two cells being neighbours when they share at least
// `min_shared` of their `nvc` entity ids. Returns the number of colours.
int32_t greedy_cell_coloring(int64_t num_cells, int32_t nvc, const int32_t* cell_entities,
                             int64_t num_entities, int32_t min_shared, int32_t* colors);
// Same from a start vertex given by the caller (one component).
void cuthill_mckee_from(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int32_t start,
                        std::vector<int32_t>& order);

// Row blocks of the streamed CSR SpMV: maximal runs of consecutive rows holding at most
// `budget` nonzeros and at most `max_rows` rows; a row longer than `budget` is a block of its
// own. Returns the first row of each block followed by the row count.
std::vector<int32_t> build_spmv_blocks(const std::vector<int32_t>& row_ptr, int32_t budget,
                                       int32_t max_rows);
// Node chunks of the node-blocked SpMV (vector P2 expanded from the node graph, node_layout 1
// or 2): (first node, first node entry) pairs of consecutive nodes with at most `budget` node
// entries and `max_nodes` nodes each, closed by (nodes, node entries); empty when one node
// alone has more than `budget` entries.
std::vector<int32_t> build_node_chunks(const std::vector<int32_t>& row_ptr, int node_layout, int32_t budget,
                                       int32_t max_nodes);

// Throws std::runtime_error on inconsistent input or exceeded limits.
void build_p1_block_plan(int32_t gdim, int64_t num_vertices, int64_t num_cells,
                         const int32_t* cells, const int32_t* cell_dofs,
                         const int32_t* coeff_dofs, const double* x, const HostCSR& csr,
                         int32_t target_rows, P1BlockPlan& plan);

// Blocked-scatter partition for vector P2 (3 dofs per node, node-blocked CSR). P2 nodes
// (vertices and edges) are grouped into blocks whose rows fit in LDS; a work item is a pair
// (cell, owned node a) and adds the 10 3x3 blocks K[(.,a),(.,b)] of that cell to node a's rows.
struct VP2BlockPlan {
  int32_t num_blocks = 0;
  int32_t max_verts = 0, max_owned = 0, max_vals = 0, max_tasks = 0;
  bool component_major = true;
  std::vector<int32_t> vert_ptr;   // [nb+1] local vertices per block
  std::vector<int32_t> vert_id;    // global vertex of each local vertex
  std::vector<int32_t> vert_kdof;  // P1 coefficient dof of each local vertex
  std::vector<int32_t> own_ptr;    // [nb+1] owned nodes per block
  std::vector<int32_t> own_rows;   // 3 global rows (component dofs) per owned node
  std::vector<uint32_t> own_info;  // LDS value offset | (neighbour count << 16)
  std::vector<int32_t> task_ptr;   // [nb+1]
  std::vector<uint64_t, UninitAlloc<uint64_t>> task;  // 3 words per work item (see build_vp2_block_plan)
  std::vector<int32_t, UninitAlloc<int32_t>> task_cell;  // cell of each work item (Dirichlet path only)
  std::vector<int32_t> task_nvert; // [nb] work items of vertex nodes (they come first)
  std::vector<int32_t> blk_vals;   // [nb] LDS values per block
};

// Rank of node b among node a's column nodes for every (cell, a, b) (100 bytes per cell);
// verifies that the CSR is node-blocked and returns true for component-major column order
// (UFC numbering), false for node-major (block-interleaved numbering).
bool build_vp2_ranks(int64_t num_cells, const int32_t* cell_dofs, int64_t ndofs, const HostCSR& csr,
                     std::vector<uint8_t>& rank);

void build_vp2_block_plan(int64_t num_cells, const int32_t* cells, const int32_t* cell_dofs,
                          const int32_t* coeff_dofs, const double* x, const HostCSR& csr,
                          const std::vector<uint8_t>& rank, bool component_major,
                          int32_t max_block_vals, VP2BlockPlan& plan);

// Limits of the blocked kernel's LDS-resident block.
constexpr int kMaxBlockNodes = 2048;
constexpr int kMaxBlockNnz = 8192;
constexpr int kMaxRank = 16;  // 4-bit column rank
// LDS row stride (doubles) of the strided variant: rows of <= 16 entries, odd stride so that
// equal ranks of different rows fall into different LDS banks.
constexpr int kRowStride = 17;

}  // namespace dfa
