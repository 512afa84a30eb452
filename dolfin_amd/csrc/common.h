// Internal declarations shared by the host structure builder, the kernels and the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dolfin_amd.h"

namespace dfa {

// Thread-local error channel behind dfa_last_error().
void set_error(const std::string& msg);
const char* last_error();

struct HostCSR {
  std::vector<int32_t> row_ptr;
  std::vector<int32_t> col;
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
  std::vector<int32_t> row_ptr;         // [nb+1] owned rows per block
  std::vector<int32_t> row_dof;         // global row (dof) of each owned row
  std::vector<uint32_t> row_info;       // (local node id) | (LDS value offset << 16)
  std::vector<int32_t> blk_nnz;         // [nb] values per block
  // Write-out runs: maximal sequences of owned rows with consecutive dofs are one contiguous
  // CSR range. Packed as csr_start | local_offset << 32 | length << 48.
  std::vector<int32_t> run_ptr;         // [nb+1]
  std::vector<uint64_t> runs;
};

// Throws std::runtime_error on inconsistent input or exceeded limits.
void build_p1_block_plan(int32_t gdim, int64_t num_vertices, int64_t num_cells,
                         const int32_t* cells, const int32_t* cell_dofs,
                         const int32_t* coeff_dofs, const double* x, const HostCSR& csr,
                         int32_t target_rows, P1BlockPlan& plan);

// Limits of the blocked kernel's LDS-resident block.
constexpr int kMaxBlockNodes = 2048;
constexpr int kMaxBlockNnz = 8192;
constexpr int kMaxRank = 16;  // 4-bit column rank
// LDS row stride (doubles) of the strided variant: rows of <= 16 entries, odd stride so that
// equal ranks of different rows fall into different LDS banks.
constexpr int kRowStride = 17;

}  // namespace dfa
