// Internal state of a dfa_assembler handle (shared by the C ABI and the batched launcher).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace dfa {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceArena {
  std::vector<void*> ptrs;
  int64_t bytes = 0;
  template <class T, class A>
  T* upload(const std::vector<T, A>& v, hipStream_t s) {
    return upload(v.data(), v.size(), s);
  }
  template <class T>
  T* upload(const T* h, size_t n, hipStream_t s) {
    void* d = nullptr;
    const size_t nb = std::max<size_t>(n * sizeof(T), 16);
    hip_check(hipMalloc(&d, nb), "hipMalloc");
    ptrs.push_back(d);
    bytes += static_cast<int64_t>(n * sizeof(T));
    if (n) hip_check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    return static_cast<T*>(d);
  }
  // take ownership of a device allocation made elsewhere
  template <class T>
  T* adopt(T* d, int64_t nbytes) {
    ptrs.push_back(d);
    bytes += nbytes;
    return d;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
  }
};

// Reverse Cuthill-McKee numbering of a device graph (rcm_device.hip); false when the device
// build does not apply and the host's rcm_ordering is to be used.
bool rcm_ordering_device(int64_t n, const int32_t* d_ap, const int32_t* d_adj, int32_t max_degree,
                         int32_t* old_to_new, hipStream_t s);
// Colouring of the cells for the graph-coloured scatter (cells sharing a dof differ), on the
// device (color_device.hip): the cells in colour order and the colour offsets; false when it
// does not apply (the host greedy colouring is then used).
bool color_cells_device(int64_t nc, int32_t nvc, const int32_t* d_cdofs, int64_t ndofs,
                        std::vector<int32_t>& perm, std::vector<int64_t>& color_ptr, hipStream_t s);
// The node graph (adjacency without self loops) of a device CSR structure.
void node_graph_device(int64_t n, const int32_t* rp, const int32_t* col, int32_t** ap, int32_t** adj,
                       int32_t& max_degree, hipStream_t s);

// The CSR structure of build_csr, built on the device from device cell_dofs (csr_build.hip):
// the arrays stay on the device (returned through d_row_ptr / d_col; the caller frees them)
// and are also copied into `out`. Returns false, with nothing allocated, when device memory
// is short. For a vector-P2 CSR expanded from the node graph (out.node_layout != 0), the node graph is
// returned through d_node_rp / d_node_col when both are given (the caller frees them; the host
// copy in `out` is then left empty), else copied into out.node_row_ptr / node_col.
bool build_csr_device(int64_t num_cells, int32_t nd, const int32_t* d_cell_dofs,
                      int64_t num_rows, bool keep_diagonal, HostCSR& out, int32_t** d_row_ptr,
                      int32_t** d_col, hipStream_t stream, bool download_col = true,
                      int32_t** d_node_rp = nullptr, int32_t** d_node_col = nullptr);

// Device build of the P1 block plan and edge chains (plan_device.hip): the same arrays as
// build_p1_block_plan + build_p1_chain_plan + the slot arrays of the edge-chain kernels,
// allocated in `plan` and described by `t`. Tetrahedra only.
struct DevicePlanInput {
  int32_t gdim = 3;
  int64_t nv = 0, nc = 0, ndofs = 0, nkdofs = 0;
  const int32_t* d_cells = nullptr;
  const int32_t* d_cdofs = nullptr;
  const int32_t* d_kdofs = nullptr;
  const double* h_x = nullptr;       // host coordinates (uploaded for the partition)
  const int32_t* d_row_ptr = nullptr;
  const int32_t* d_col = nullptr;
  int32_t max_row_len = 0;
};
struct DevicePlanStats {
  int64_t appearances = 0;
  bool kdof_injective = false;
  bool chains_built = false;
  int32_t lds_bytes = 0;
};
constexpr int kDevicePlanOk = 0;
constexpr int kDevicePlanError = 1;        // the input has no blocked plan (message in err)
constexpr int kDevicePlanUnsupported = 2;  // outside the device build's limits: build on the host
int build_p1_plan_device(const DevicePlanInput& in, int32_t target_rows, bool want_chains,
                         int32_t block_threads, int32_t chain_threads, DeviceArena& plan,
                         P1BlockTopo& t, DevicePlanStats& stats, std::string& err, hipStream_t s);

// Vector P2 on the device (plan_vp2_device.hip): the rank table of build_vp2_ranks (false when
// a column is missing or a row has more than 256 nodes) and the arrays of build_vp2_block_plan,
// for CSR structures expanded from the node graph (node_layout 1 = component-major,
// 2 = node-major).
bool vp2_ranks_device(int64_t nc, const int32_t* d_cdofs, const int32_t* d_row_ptr, const int32_t* d_col,
                      bool node_major, uint8_t* d_rank, hipStream_t s);
struct DeviceVP2Input {
  int64_t nv = 0, nc = 0, ndofs = 0;
  int node_layout = 0;
  const int32_t* d_cells = nullptr;
  const int32_t* d_cdofs = nullptr;
  const int32_t* d_kdofs = nullptr;
  const double* h_x = nullptr;  // host coordinates (uploaded for the partition)
  const int32_t* d_row_ptr = nullptr;
  const uint8_t* d_rank = nullptr;
};
int build_vp2_plan_device(const DeviceVP2Input& in, int32_t max_block_vals, DeviceArena& plan, VP2BlockTopo& t,
                          int64_t& appearances, std::string& err, hipStream_t s);

}  // namespace dfa

struct dfa_assembler_s {
  int32_t gdim = 0, element = 0, nd = 0, nvc = 0;
  int64_t nv = 0, nc = 0, ndofs = 0, nkdofs = 0, nnz = 0;
  bool rows_covered = true;   // every CSR row belongs to a dof of some cell
  int32_t* d_cells = nullptr;
  int32_t* d_cdofs = nullptr;
  int32_t* d_kdofs = nullptr;
  int32_t* d_row_ptr = nullptr;
  int32_t* d_col = nullptr;
  int32_t* d_spmv_blocks = nullptr;  // SpMV row blocks: (first row, its row_ptr) pairs
  int64_t spmv_nblocks = 0;
  dfa::DeviceArena base;
  // blocked P1 plan
  bool has_plan = false;
  bool plan_on_device = false;   // the P1 block plan was built by the device planner
  std::string plan_error;
  std::string plan_verify_error; // DFA_PLAN_VERIFY=1: how the device plan differs from the host's
  dfa::P1BlockTopo btopo;
  dfa::DeviceArena plan;
  int64_t appearances = 0;
  bool kdof_injective = false;
  // vector P2: per cell, rank of node b among the neighbours of node a (100 bytes) and the
  // CSR strides (node stride s1, component stride s2 per row)
  uint8_t* d_vp2_rank = nullptr;
  // vector P2 expanded from the node graph: the node CSR (nodes = component-0 dofs, or dof / 3
  // when node-major), layout 1 = component-major, 2 = node-major; 0 = no node structure
  int32_t node_layout = 0;
  int64_t node_nnz = 0;
  int32_t* d_node_rp = nullptr;
  int32_t* d_node_col = nullptr;
  int32_t* d_node_chunks = nullptr;  // node-blocked SpMV chunks (none: the streamed CSR SpMV)
  int64_t node_nchunks = 0;
  // streamed CSR SpMV: per row block the smallest column (-1: 4-byte columns) and the 16-bit
  // column offsets
  int32_t* d_blk_col0 = nullptr;
  uint16_t* d_col16 = nullptr;
  int32_t vp2_s1 = 0;
  bool vp2_component_major = true;
  bool has_vp2_plan = false;
  dfa::VP2BlockTopo vtopo;
  // graph-coloured scatter (P1), built on first use: cells in colour order
  bool has_colors = false;
  std::vector<int64_t> color_ptr;      // [ncolors + 1]
  int32_t* d_ccells = nullptr;
  int32_t* d_cdofs_c = nullptr;
  int32_t* d_kdofs_c = nullptr;        // aliases d_cdofs_c when the dofmaps are equal
  uint8_t* d_crank = nullptr;
  dfa::DeviceArena colors;
  // exterior facets (P1), built on first use
  bool has_facets = false;
  std::vector<int32_t> facet_cell;
  std::vector<int8_t> facet_local;
  int32_t* d_facet_cell = nullptr;
  int8_t* d_facet_local = nullptr;
  dfa::DeviceArena facets;
  // Dirichlet values shifted by x0 (dfa_assemble_system_x0), allocated on first use
  double* d_bc_x0 = nullptr;
  dfa::DeviceArena x0buf;
  // Dirichlet conditions on the assembled rows: cells per dof (the diagonal of a Dirichlet row),
  // for P1 the transpose map of the CSR (the slot of A_ji for each A_ij)
  int32_t* d_dof_cells = nullptr;
  int32_t* d_tmap = nullptr;
  uint8_t* d_node_bits = nullptr;      // vector P2 on the node graph: Dirichlet bits, touched flags
  // conjugate-gradient work vectors and scalars (dfa_cg_solve), allocated on first use
  double* d_cg_r = nullptr;
  double* d_cg_p = nullptr;
  double* d_cg_ap = nullptr;
  double* d_cg_dinv = nullptr;
  double* d_cg_part = nullptr;
  dfa::CgState* d_cg_state = nullptr;
  dfa::DeviceArena cg;
};
