// C ABI of the assembly engine (include/dolfin_amd.h): handle management, structure phase
// and dispatch of the cell kernels. Counterpart of dolfin::Assembler / SystemAssembler
// (dolfin/fem/Assembler.cpp:52-103, SystemAssembler.cpp:138-301) for one (mesh, space).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "assembler.h"
#include "kernels.h"

namespace dfa {

namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }
}  // namespace dfa

namespace {

using dfa::hip_check;

int fail(const std::string& m, int code) {
  dfa::set_error(m);
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    dfa::set_error("");
    return f();
  } catch (const dfa::HipError& e) {
    return fail(e.what(), DFA_ERR_HIP);
  } catch (const std::exception& e) {
    return fail(e.what(), DFA_ERR_ARG);
  }
}

// Plan parameters for assemblers created afterwards by this thread (0 = built-in default);
// set by dfa_set_plan_defaults, e.g. from an autotuner.
thread_local int tl_p1_rows = 0, tl_p1_threads = 0, tl_p2_vals = 0;

int default_block_threads() {
  if (tl_p1_threads) return tl_p1_threads;
  if (const char* e = std::getenv("DFA_BLOCK_THREADS")) {
    const int v = std::atoi(e);
    if (v == 256 || v == 512) return v;
  }
  return 512;
}

int default_p2_block_vals() {
  if (tl_p2_vals) return tl_p2_vals;
  if (const char* e = std::getenv("DFA_P2_BLOCK_VALS")) {
    const int v = std::atoi(e);
    if (v >= 1024 && v <= 16384) return v;
  }
  return 4608;
}

int default_chain_threads() {
  if (const char* e = std::getenv("DFA_CHAIN_THREADS")) {
    const int v = std::atoi(e);
    if (v == 256 || v == 512) return v;
  }
  return 512;
}

int default_block_rows(int gdim) {
  if (tl_p1_rows) return tl_p1_rows;
  if (const char* e = std::getenv("DFA_BLOCK_ROWS")) {
    const int v = std::atoi(e);
    if (v >= 8 && v <= 4096) return v;
  }
  return gdim == 3 ? 160 : 512;
}

// Reverse Cuthill-McKee on the device unless DFA_HOST_RCM=1 or no GPU is present.
bool rcm_on_device() {
  if (const char* e = std::getenv("DFA_HOST_RCM"))
    if (std::atoi(e) != 0) return false;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

// dfa_node_reordering on the device: cells -> device CSR (sort of the cell pairs) -> node graph
// -> reverse Cuthill-McKee. False when the device build does not apply (the host's then runs).
bool rcm_nodes_device(int64_t nc, int32_t npc, const int32_t* cells, int64_t n, int32_t* old_to_new) {
  hipStream_t s = nullptr;
  dfa::DeviceArena a;
  int32_t *rp = nullptr, *col = nullptr, *ap = nullptr, *adj = nullptr;
  auto cleanup = [&]() {
    for (int32_t* p : {rp, col, ap, adj})
      if (p) (void)hipFree(p);
    a.release();
  };
  try {
    const int32_t* dcells = a.upload(cells, static_cast<size_t>(nc) * npc, s);
    dfa::HostCSR csr;
    if (!dfa::build_csr_device(nc, npc, dcells, n, false, csr, &rp, &col, s, false)) {
      cleanup();
      return false;
    }
    int32_t maxd = 0;
    dfa::node_graph_device(n, rp, col, &ap, &adj, maxd, s);
    bool ok = false;
    int32_t m = 0;
    hip_check(hipMemcpy(&m, ap + n, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    if (m == 0) {  // no edges: the identity, as the host build gives
      for (int64_t v = 0; v < n; ++v) old_to_new[v] = static_cast<int32_t>(v);
      ok = true;
    } else {
      ok = dfa::rcm_ordering_device(n, ap, adj, maxd, old_to_new, s);
    }
    cleanup();
    return ok;
  } catch (...) {
    cleanup();
    throw;
  }
}

}  // namespace

extern "C" {

const char* dfa_last_error(void) { return dfa::last_error(); }
const char* dfa_version(void) { return "dolfin_amd 0.1 (gfx950)"; }

int dfa_set_plan_defaults(int32_t p1_block_rows, int32_t p1_threads, int32_t p2_block_vals) {
  if ((p1_block_rows && (p1_block_rows < 8 || p1_block_rows > 4096)) ||
      (p1_threads && p1_threads != 256 && p1_threads != 512) ||
      (p2_block_vals && (p2_block_vals < 1024 || p2_block_vals > 16384)))
    return fail("plan defaults out of range (rows 8..4096, threads 256/512, P2 values 1024..16384)",
                DFA_ERR_ARG);
  tl_p1_rows = p1_block_rows;
  tl_p1_threads = p1_threads;
  tl_p2_vals = p2_block_vals;
  return DFA_OK;
}

extern "C++" {  // internal helpers keep C++ linkage inside the C block
namespace {
// Phase times of dfa_assembler_create on stderr when DFA_TIMINGS=1 (the role of DOLFIN's
// "Build sparsity" / "Init tensor" timers, dolfin/common/Timer.h).
struct PhaseTimer {
  bool on;
  std::chrono::steady_clock::time_point t;
  PhaseTimer() {
    const char* e = std::getenv("DFA_TIMINGS");
    on = e && e[0] == '1';
    t = std::chrono::steady_clock::now();
  }
  void mark(const char* what) {
    if (!on) return;
    const auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[dfa] %-28s %8.3f s\n", what, std::chrono::duration<double>(n - t).count());
    t = n;
  }
};

bool env_flag(const char* name) {
  const char* e = std::getenv(name);
  return e && e[0] == '1';
}

// The host build of the P1 block plan, its edge chains and the chain kernels' fixed-stride
// slots (the arrays the device build produces as well).
struct HostP1Plan {
  dfa::P1BlockPlan P;
  dfa::P1ChainPlan C;
  bool chains = false;
  int32_t SN = 0, SR = 0;
  std::vector<int4> hdr;
  std::vector<int32_t> sv, sk, sd;
  std::vector<int2> sr;
};

bool build_host_p1_plan(int32_t gdim, int64_t nv, int64_t nc, const int32_t* cells, const int32_t* cd,
                        const int32_t* kd, const double* x, const dfa::HostCSR& csr, bool want_chains,
                        HostP1Plan& hp, std::string& err) {
  try {
    auto& P = hp.P;
    dfa::build_p1_block_plan(gdim, nv, nc, cells, cd, kd, x, csr, default_block_rows(gdim), P);
    if (!want_chains) return true;
    dfa::build_p1_chain_plan(P, hp.C);
    hp.chains = true;
    const auto& C = hp.C;
    hp.SN = P.max_nodes;
    hp.SR = std::max<int32_t>(P.max_rows, 1);
    const int32_t SN = hp.SN, SR = hp.SR;
    const int64_t nb = P.num_blocks;
    hp.hdr.assign(static_cast<size_t>(nb), make_int4(0, 0, 0, 0));
    hp.sv.assign(static_cast<size_t>(nb * SN), 0);
    hp.sr.assign(static_cast<size_t>(nb * SR), make_int2(0, 0));
    hp.sd.assign(static_cast<size_t>(nb * SR), 0);
    if (!P.kdof_is_vertex) hp.sk.assign(static_cast<size_t>(nb * SN), 0);
    for (int64_t b = 0; b < nb; ++b) {
      const int32_t n0 = P.node_ptr[b], nn = P.node_ptr[b + 1] - n0;
      const int32_t r0 = P.row_ptr[b], nr = P.row_ptr[b + 1] - r0;
      hp.hdr[b] = make_int4(C.chain_ptr[b], C.chain_ptr[b + 1] - C.chain_ptr[b], nn, nr);
      for (int32_t i = 0; i < SN; ++i) {  // padding repeats the block's node 0
        const int32_t j = n0 + (i < nn ? i : 0);
        hp.sv[b * SN + i] = P.node_vertex[j];
        if (!hp.sk.empty()) hp.sk[b * SN + i] = P.node_kdof[j];
      }
      for (int32_t i = 0; i < nr; ++i) {
        const int32_t d = P.row_dof[r0 + i];
        const int32_t c0 = csr.row_ptr[d], len = csr.row_ptr[d + 1] - c0;
        hp.sr[b * SR + i] = make_int2(c0, len | (static_cast<int32_t>(P.row_diag[r0 + i]) << 8));
        hp.sd[b * SR + i] = d;
      }
    }
    return true;
  } catch (const std::exception& e) {
    err = e.what();
    return false;
  }
}

void upload_host_p1_plan(dfa_assembler_s* h, const HostP1Plan& hp, hipStream_t s) {
  const auto& P = hp.P;
  auto& t = h->btopo;
  t.num_blocks = P.num_blocks;
  t.max_nodes = P.max_nodes;
  t.max_rows = P.max_rows;
  t.max_nnz = P.max_nnz;
  t.max_row_len = P.max_row_len;
  t.threads = default_block_threads();
  if (const char* e = std::getenv("DFA_NODE_RECORD")) t.wide_nodes = std::atoi(e) != 32;
  t.max_cells = P.max_cells;
  t.kdof_is_vertex = P.kdof_is_vertex;
  t.node_ptr = h->plan.upload(P.node_ptr, s);
  t.node_vertex = h->plan.upload(P.node_vertex, s);
  t.node_kdof = h->plan.upload(P.node_kdof, s);
  t.node_dof = h->plan.upload(P.node_dof, s);
  t.cell_ptr = h->plan.upload(P.cell_ptr, s);
  std::vector<uint64_t> rec(2 * P.cell_ids.size());
  for (size_t k = 0; k < P.cell_ids.size(); ++k) {
    rec[2 * k] = P.cell_ids[k];
    rec[2 * k + 1] = P.cell_rank[k];
  }
  t.cell_rec = h->plan.upload(rec, s);
  if (!P.cell_rank_wide.empty()) t.cell_rank_wide = h->plan.upload(P.cell_rank_wide, s);
  if (hp.chains) {
    t.max_chains = hp.C.max_chains;
    t.chain_threads = default_chain_threads();
    t.chain_ptr = h->plan.upload(hp.C.chain_ptr, s);
    t.chain_rec = h->plan.upload(hp.C.chain_rec, s);
    t.slot_nodes = hp.SN;
    t.slot_rows = hp.SR;
    t.blk_hdr = h->plan.upload(hp.hdr, s);
    t.slot_vertex = h->plan.upload(hp.sv, s);
    if (!hp.sk.empty()) t.slot_kdof = h->plan.upload(hp.sk, s);
    t.slot_row = h->plan.upload(hp.sr, s);
    t.slot_rdof = h->plan.upload(hp.sd, s);
  }
  t.row_ptr_b = h->plan.upload(P.row_ptr, s);
  t.row_dof = h->plan.upload(P.row_dof, s);
  t.row_info = h->plan.upload(P.row_info, s);
  t.row_diag = h->plan.upload(P.row_diag, s);
  t.blk_nnz = h->plan.upload(P.blk_nnz, s);
  t.run_ptr = h->plan.upload(P.run_ptr, s);
  t.runs = h->plan.upload(P.runs, s);
  h->kdof_injective = P.kdof_injective;
  t.csr_row_ptr = h->d_row_ptr;
  h->appearances = static_cast<int64_t>(P.cell_ids.size());
  h->has_plan = true;
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors die after this
}

// A device array against its host build ("" = equal, else the first differing element).
template <class T, class V>
std::string same_array(const char* name, const T* d, const V& hv) {
  const size_t n = hv.size();
  if (n && !d) return std::string(name) + ": missing on the device";
  std::vector<std::remove_cv_t<T>> dv(n);
  if (n) hip_check(hipMemcpy(dv.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  if (n && std::memcmp(dv.data(), hv.data(), n * sizeof(T)) != 0) {
    size_t i = 0;
    while (std::memcmp(&dv[i], &hv[i], sizeof(T)) == 0) ++i;
    return std::string(name) + ": first difference at element " + std::to_string(i) + " of " +
           std::to_string(n);
  }
  return "";
}

// The SpMV view of a handle's structure with the values d_values: the node-blocked kernel for
// vector P2 expanded from the node graph, 16-bit column offsets otherwise (DFA_SPMV_CSR=1: the
// plain 4-byte columns).
dfa::CsrView spmv_view(dfa_assembler_s* h, const double* d_values) {
  dfa::CsrView A;
  A.nrows = h->ndofs;
  A.nnz = h->nnz;
  A.row_ptr = h->d_row_ptr;
  A.col = h->d_col;
  A.values = d_values;
  A.blocks = h->d_spmv_blocks;
  A.nblocks = h->spmv_nblocks;
  if (!std::getenv("DFA_SPMV_CSR")) {
    A.blk_col0 = h->d_blk_col0;
    A.col16 = h->d_col16;
    if (h->d_node_chunks) {
      A.node_layout = h->node_layout;
      A.node_nnz = h->node_nnz;
      A.node_rp = h->d_node_rp;
      A.node_col = h->d_node_col;
      A.node_chunks = h->d_node_chunks;
      A.node_nchunks = h->node_nchunks;
    }
  }
  return A;
}

// Outcome of two plan builds when at least one failed ("" = both failed alike).
std::string compare_plan_failures(bool host_ok, const std::string& herr, bool dev_failed,
                                  const std::string& derr) {
  if (host_ok != !dev_failed)
    return std::string("device plan ") + (dev_failed ? "failed (" + derr + ")" : "built") +
           ", host plan " + (host_ok ? "built" : "failed (" + herr + ")");
  return herr == derr ? "" : "different errors: device '" + derr + "', host '" + herr + "'";
}

// Element-by-element comparison of the device-built plan with the host build ("" = equal).
std::string compare_p1_plans(dfa_assembler_s* h, const HostP1Plan& hp, bool host_ok,
                             const std::string& herr, bool dev_failed, const std::string& derr,
                             hipStream_t s) {
  if (!host_ok || dev_failed) return compare_plan_failures(host_ok, herr, dev_failed, derr);
  const auto& P = hp.P;
  const auto& t = h->btopo;
  auto same = [](const char* name, const auto* d, const auto& hv) { return same_array(name, d, hv); };
  (void)s;
  std::string r;
  auto chk = [&](std::string x) {
    if (r.empty() && !x.empty()) r = x;
  };
  auto scalar = [&](const char* name, int64_t a, int64_t b) {
    if (a != b) chk(std::string(name) + ": device " + std::to_string(a) + ", host " + std::to_string(b));
  };
  scalar("num_blocks", t.num_blocks, P.num_blocks);
  scalar("max_nodes", t.max_nodes, P.max_nodes);
  scalar("max_rows", t.max_rows, P.max_rows);
  scalar("max_nnz", t.max_nnz, P.max_nnz);
  scalar("max_cells", t.max_cells, P.max_cells);
  scalar("max_row_len", t.max_row_len, P.max_row_len);
  scalar("kdof_is_vertex", t.kdof_is_vertex, P.kdof_is_vertex);
  scalar("kdof_injective", h->kdof_injective, P.kdof_injective);
  scalar("appearances", h->appearances, static_cast<int64_t>(P.cell_ids.size()));
  if (!r.empty()) return r;
  std::vector<uint64_t> rec(2 * P.cell_ids.size());
  for (size_t k = 0; k < P.cell_ids.size(); ++k) {
    rec[2 * k] = P.cell_ids[k];
    rec[2 * k + 1] = P.cell_rank[k];
  }
  chk(same("node_ptr", t.node_ptr, P.node_ptr));
  chk(same("node_vertex", t.node_vertex, P.node_vertex));
  chk(same("node_kdof", t.node_kdof, P.node_kdof));
  chk(same("node_dof", t.node_dof, P.node_dof));
  chk(same("cell_ptr", t.cell_ptr, P.cell_ptr));
  chk(same("cell_rec", t.cell_rec, rec));
  chk(same("cell_rank_wide", t.cell_rank_wide, P.cell_rank_wide));
  chk(same("row_ptr", t.row_ptr_b, P.row_ptr));
  chk(same("row_dof", t.row_dof, P.row_dof));
  chk(same("row_info", t.row_info, P.row_info));
  chk(same("row_diag", t.row_diag, P.row_diag));
  chk(same("blk_nnz", t.blk_nnz, P.blk_nnz));
  chk(same("run_ptr", t.run_ptr, P.run_ptr));
  chk(same("runs", t.runs, P.runs));
  if (hp.chains) {
    scalar("max_chains", t.max_chains, hp.C.max_chains);
    scalar("slot_nodes", t.slot_nodes, hp.SN);
    scalar("slot_rows", t.slot_rows, hp.SR);
    if (!r.empty()) return r;
    chk(same("chain_ptr", t.chain_ptr, hp.C.chain_ptr));
    chk(same("chain_rec", t.chain_rec, hp.C.chain_rec));
    chk(same("blk_hdr", t.blk_hdr, hp.hdr));
    chk(same("slot_vertex", t.slot_vertex, hp.sv));
    chk(same("slot_kdof", t.slot_kdof, hp.sk));
    chk(same("slot_row", t.slot_row, hp.sr));
    chk(same("slot_rdof", t.slot_rdof, hp.sd));
  } else if (t.chain_ptr) {
    chk("chains: built on the device only");
  }
  return r;
}

// The same for the vector-P2 plan and rank table.
std::string compare_vp2_plans(dfa_assembler_s* h, const dfa::VP2BlockPlan& P, const std::vector<uint8_t>& rank,
                              bool host_ok, const std::string& herr, bool dev_failed, const std::string& derr) {
  std::string r = same_array("rank", h->d_vp2_rank, rank);
  if (!r.empty()) return r;
  if (!host_ok || dev_failed) return compare_plan_failures(host_ok, herr, dev_failed, derr);
  const auto& t = h->vtopo;
  auto chk = [&](std::string x) {
    if (r.empty() && !x.empty()) r = x;
  };
  auto scalar = [&](const char* name, int64_t a, int64_t b) {
    if (a != b) chk(std::string(name) + ": device " + std::to_string(a) + ", host " + std::to_string(b));
  };
  scalar("num_blocks", t.num_blocks, P.num_blocks);
  scalar("max_verts", t.max_verts, P.max_verts);
  scalar("max_owned", t.max_owned, P.max_owned);
  scalar("max_vals", t.max_vals, P.max_vals);
  scalar("component_major", t.component_major, P.component_major);
  scalar("appearances", h->appearances, static_cast<int64_t>(P.task_cell.size()));
  if (!r.empty()) return r;
  chk(same_array("vert_ptr", t.vert_ptr, P.vert_ptr));
  chk(same_array("vert_id", t.vert_id, P.vert_id));
  chk(same_array("vert_kdof", t.vert_kdof, P.vert_kdof));
  chk(same_array("own_ptr", t.own_ptr, P.own_ptr));
  chk(same_array("own_rows", t.own_rows, P.own_rows));
  chk(same_array("own_info", t.own_info, P.own_info));
  chk(same_array("task_ptr", t.task_ptr, P.task_ptr));
  chk(same_array("task", t.task, P.task));
  chk(same_array("task_cell", t.task_cell, P.task_cell));
  chk(same_array("task_nvert", t.task_nvert, P.task_nvert));
  chk(same_array("blk_vals", t.blk_vals, P.blk_vals));
  return r;
}
}  // namespace
}  // extern "C++"

int dfa_assembler_create(dfa_assembler* out, int32_t gdim, int64_t nv, int64_t nc,
                         const int32_t* h_cells, int32_t element, int32_t nd,
                         const int32_t* h_cell_dofs, int64_t ndofs, const int32_t* h_coeff_dofs,
                         int64_t nkdofs, const double* h_x, int32_t keep_diagonal, void* stream) {
  return guarded([&]() -> int {
    if (!out) return fail("out is NULL", DFA_ERR_ARG);
    *out = nullptr;
    if (gdim < 1 || gdim > 3) return fail("gdim must be 1, 2 or 3", DFA_ERR_ARG);
    if (nv <= 0 || nc < 0 || ndofs <= 0 || !h_cells || !h_cell_dofs || !h_coeff_dofs)
      return fail("invalid mesh/dofmap arguments", DFA_ERR_ARG);
    if (element == DFA_ELEMENT_P1) {
      if (nd != gdim + 1) return fail("P1 needs gdim+1 dofs per cell", DFA_ERR_ARG);
    } else if (element == DFA_ELEMENT_VECTOR_P2) {
      if (gdim != 3 || nd != 30) return fail("vector P2 needs tetrahedra and 30 dofs per cell", DFA_ERR_ARG);
    } else {
      return fail("unknown element", DFA_ERR_ARG);
    }
    if (ndofs > INT32_MAX || nv > INT32_MAX || nc > INT32_MAX)
      return fail("sizes exceed int32 (la_index)", DFA_ERR_ARG);
    const int nvc = gdim + 1;
    for (int64_t i = 0; i < nc * nvc; ++i)
      if (h_cells[i] < 0 || h_cells[i] >= nv || h_coeff_dofs[i] < 0 || h_coeff_dofs[i] >= nkdofs)
        return fail("cell vertex or coefficient dof out of range", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    PhaseTimer timer;
    auto* h = new dfa_assembler_s();
    try {
      h->gdim = gdim;
      h->element = element;
      h->nd = nd;
      h->nvc = nvc;
      h->nv = nv;
      h->nc = nc;
      h->ndofs = ndofs;
      h->nkdofs = nkdofs;
      {
        std::vector<uint8_t> seen(ndofs, 0);
        for (int64_t i = 0; i < nc * nd; ++i) {
          const int32_t d = h_cell_dofs[i];
          if (d < 0 || d >= ndofs) throw std::runtime_error("cell_dofs entry out of range");
          seen[d] = 1;
        }
        for (int64_t r = 0; r < ndofs; ++r)
          if (!seen[r]) h->rows_covered = false;
      }
      h->d_cells = h->base.upload(h_cells, static_cast<size_t>(nc * nvc), s);
      h->d_cdofs = h->base.upload(h_cell_dofs, static_cast<size_t>(nc * nd), s);
      h->d_kdofs = h->base.upload(h_coeff_dofs, static_cast<size_t>(nc * nvc), s);
      timer.mark("check + upload mesh");
      // CSR structure: built on the device (sort of the cell pairs) unless device memory is
      // short or DFA_HOST_CSR=1, then on the host; both give the same arrays. The column
      // indices come back to the host only when a host plan builder needs them.
      const bool verify_plan = env_flag("DFA_PLAN_VERIFY");
      const bool device_plan = element == DFA_ELEMENT_P1 && gdim == 3 && h_x &&
                               !env_flag("DFA_HOST_PLAN");
      dfa::HostCSR csr;
      int32_t* drp = nullptr;
      int32_t* dcol = nullptr;
      int32_t* dnrp = nullptr;
      int32_t* dncol = nullptr;
      const bool vp2 = element == DFA_ELEMENT_VECTOR_P2;
      bool col_on_host = true;
      // vector P2 needs the columns on the host only when its CSR was not expanded from the
      // node graph (the ranks are then read from the node CSR)
      const bool want_col = element == DFA_ELEMENT_P1 ? (!device_plan || verify_plan) : false;
      if (!env_flag("DFA_HOST_CSR") &&
          dfa::build_csr_device(nc, nd, h->d_cdofs, ndofs, keep_diagonal != 0, csr, &drp, &dcol, s,
                                want_col, vp2 ? &dnrp : nullptr, vp2 ? &dncol : nullptr)) {
        h->nnz = csr.row_ptr.back();
        h->d_row_ptr = h->base.adopt(drp, static_cast<int64_t>(csr.row_ptr.size() * sizeof(int32_t)));
        h->d_col = h->base.adopt(dcol, h->nnz * static_cast<int64_t>(sizeof(int32_t)));
        col_on_host = want_col;
        if (csr.node_layout) {  // vector P2 expanded from the node graph, kept on the device
          const int64_t nn = ndofs / 3;
          h->node_layout = csr.node_layout;
          h->node_nnz = h->nnz / 9;
          h->d_node_rp = h->base.adopt(dnrp, (nn + 1) * static_cast<int64_t>(sizeof(int32_t)));
          h->d_node_col = h->base.adopt(dncol, h->node_nnz * static_cast<int64_t>(sizeof(int32_t)));
        }
      } else {
        dfa::build_csr(nc, nd, h_cell_dofs, ndofs, keep_diagonal != 0, csr);
        h->nnz = static_cast<int64_t>(csr.col.size());
        h->d_row_ptr = h->base.upload(csr.row_ptr, s);
        h->d_col = h->base.upload(csr.col, s);
      }
      timer.mark("CSR structure");
      auto ensure_host_col = [&]() {
        if (col_on_host) return;
        csr.col.resize(static_cast<size_t>(h->nnz));
        if (h->nnz)
          hip_check(hipMemcpyAsync(csr.col.data(), h->d_col, h->nnz * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        col_on_host = true;
      };
      {
        const std::vector<int32_t> sb = dfa::build_spmv_blocks(csr.row_ptr, dfa::kSpmvBudget,
                                                               dfa::kSpmvMaxRows);
        std::vector<int32_t> pairs(2 * sb.size());
        for (size_t i = 0; i < sb.size(); ++i) {
          pairs[2 * i] = sb[i];
          pairs[2 * i + 1] = csr.row_ptr[sb[i]];
        }
        h->spmv_nblocks = static_cast<int64_t>(sb.size()) - 1;
        h->d_spmv_blocks = h->base.upload(pairs, s);
        if (h->node_layout) {
          const std::vector<int32_t> nc2 =
              dfa::build_node_chunks(csr.row_ptr, h->node_layout, dfa::kNodeBudget, dfa::kNodeMaxNodes);
          if (!nc2.empty()) {
            h->node_nchunks = static_cast<int64_t>(nc2.size() / 2) - 1;
            h->d_node_chunks = h->base.upload(nc2, s);
          }
        }
        if (!h->d_node_chunks && h->spmv_nblocks > 0) {  // 16-bit column offsets per row block
          void* p0 = nullptr;
          void* p1 = nullptr;
          hip_check(hipMalloc(&p0, static_cast<size_t>(h->spmv_nblocks) * sizeof(int32_t)), "hipMalloc");
          h->d_blk_col0 = h->base.adopt(static_cast<int32_t*>(p0), h->spmv_nblocks * 4);
          hip_check(hipMalloc(&p1, std::max<size_t>(static_cast<size_t>(h->nnz) * sizeof(uint16_t), 16)),
                    "hipMalloc");
          h->d_col16 = h->base.adopt(static_cast<uint16_t*>(p1), h->nnz * 2);
          dfa::CsrView A;
          A.nrows = ndofs;
          A.nnz = h->nnz;
          A.col = h->d_col;
          A.blocks = h->d_spmv_blocks;
          A.nblocks = h->spmv_nblocks;
          hip_check(dfa::launch_narrow_columns(A, h->d_blk_col0, h->d_col16, s), "narrow_columns_kernel");
        }
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors die here
      }
      {  // Dirichlet conditions after assembly: cells per dof, and for P1 the transpose map
        void* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(ndofs) * sizeof(int32_t), 16)), "hipMalloc");
        h->d_dof_cells = h->base.adopt(static_cast<int32_t*>(d), ndofs * 4);
        hip_check(hipMemsetAsync(d, 0, static_cast<size_t>(ndofs) * sizeof(int32_t), s), "hipMemsetAsync");
        hip_check(dfa::launch_dof_cell_count(nc * nd, h->d_cdofs, h->d_dof_cells, s), "dof_cell_count_kernel");
        if (h->node_layout) {  // vector P2 on the node graph: Dirichlet bits and touched flags per node
          void* nb = nullptr;
          hip_check(hipMalloc(&nb, std::max<size_t>(static_cast<size_t>(2 * (ndofs / 3)), 16)), "hipMalloc");
          h->d_node_bits = h->base.adopt(static_cast<uint8_t*>(nb), 2 * (ndofs / 3));
        }
      }
      if (element == DFA_ELEMENT_P1) {
        void* t = nullptr;
        hip_check(hipMalloc(&t, std::max<size_t>(static_cast<size_t>(h->nnz) * sizeof(int32_t), 16)), "hipMalloc");
        h->d_tmap = h->base.adopt(static_cast<int32_t*>(t), h->nnz * 4);
        dfa::CsrView A;
        A.nrows = ndofs;
        A.nnz = h->nnz;
        A.row_ptr = h->d_row_ptr;
        A.col = h->d_col;
        hip_check(dfa::launch_transpose_map(A, h->d_tmap, s), "transpose_map_kernel");
      }
      if (element == DFA_ELEMENT_P1) {
        if (h_x) {
          const char* ce = std::getenv("DFA_P1_CHAINS");
          const bool want_chains = gdim == 3 && !(ce && std::atoi(ce) == 0);
          int rc = dfa::kDevicePlanUnsupported;
          std::string derr;
          if (device_plan) {
            dfa::DevicePlanInput in;
            in.gdim = gdim;
            in.nv = nv;
            in.nc = nc;
            in.ndofs = ndofs;
            in.nkdofs = nkdofs;
            in.d_cells = h->d_cells;
            in.d_cdofs = h->d_cdofs;
            in.d_kdofs = h->d_kdofs;
            in.h_x = h_x;
            in.d_row_ptr = h->d_row_ptr;
            in.d_col = h->d_col;
            for (size_t r = 0; r + 1 < csr.row_ptr.size(); ++r)
              in.max_row_len = std::max(in.max_row_len, csr.row_ptr[r + 1] - csr.row_ptr[r]);
            dfa::DevicePlanStats st;
            dfa::P1BlockTopo t;
            rc = dfa::build_p1_plan_device(in, default_block_rows(gdim), want_chains,
                                           default_block_threads(), default_chain_threads(), h->plan,
                                           t, st, derr, s);
            if (rc == dfa::kDevicePlanOk) {
              if (const char* e = std::getenv("DFA_NODE_RECORD")) t.wide_nodes = std::atoi(e) != 32;
              t.csr_row_ptr = h->d_row_ptr;
              h->btopo = t;
              h->appearances = st.appearances;
              h->kdof_injective = st.kdof_injective;
              h->has_plan = true;
              h->plan_on_device = true;
              timer.mark("P1 block + chain plan (device)");
            } else if (rc == dfa::kDevicePlanError) {
              h->plan.release();
              h->plan_error = derr;
            } else {
              h->plan.release();
            }
          }
          if (rc == dfa::kDevicePlanUnsupported || verify_plan) {
            ensure_host_col();
            HostP1Plan hp;
            std::string herr;
            const bool ok = build_host_p1_plan(gdim, nv, nc, h_cells, h_cell_dofs, h_coeff_dofs, h_x, csr,
                                               want_chains, hp, herr);
            timer.mark("P1 block + chain plan (host)");
            if (rc == dfa::kDevicePlanUnsupported) {
              if (ok) upload_host_p1_plan(h, hp, s);
              else h->plan_error = herr;
            } else {
              h->plan_verify_error = compare_p1_plans(h, hp, ok, herr, rc == dfa::kDevicePlanError, derr, s);
              if (!h->plan_verify_error.empty())
                std::fprintf(stderr, "[dfa] plan verification: %s\n", h->plan_verify_error.c_str());
            }
          }
        } else {
          h->plan_error = "no coordinates given at creation";
        }
      } else {
        // rank table and block plan on the device when the CSR was expanded from the node
        // graph, else (or with DFA_HOST_PLAN=1) on the host; DFA_PLAN_VERIFY=1 builds both
        std::vector<uint8_t> rank;
        bool ranks_on_device = false;
        if (csr.node_layout && !env_flag("DFA_HOST_PLAN")) {
          void* dr = nullptr;
          hip_check(hipMalloc(&dr, std::max<size_t>(static_cast<size_t>(nc) * 100, 16)), "hipMalloc");
          h->d_vp2_rank = h->base.adopt(static_cast<uint8_t*>(dr), nc * 100);
          ranks_on_device = dfa::vp2_ranks_device(nc, h->d_cdofs, h->d_row_ptr, h->d_col, csr.node_layout == 2,
                                                  h->d_vp2_rank, s);
          if (ranks_on_device) {
            h->vp2_component_major = csr.node_layout == 1;
            timer.mark("vector-P2 ranks (device)");
          }
        }
        auto host_ranks = [&]() {
          if (!csr.node_layout) ensure_host_col();
          if (csr.node_layout && csr.node_row_ptr.empty()) {  // the node graph, copied on demand
            const int64_t nn = ndofs / 3;
            csr.node_row_ptr.resize(static_cast<size_t>(nn) + 1);
            csr.node_col.resize(static_cast<size_t>(h->node_nnz));
            hip_check(hipMemcpyAsync(csr.node_row_ptr.data(), h->d_node_rp, (nn + 1) * sizeof(int32_t),
                                     hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
            if (h->node_nnz)
              hip_check(hipMemcpyAsync(csr.node_col.data(), h->d_node_col, h->node_nnz * sizeof(int32_t),
                                       hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
          }
          const bool cm = dfa::build_vp2_ranks(nc, h_cell_dofs, ndofs, csr, rank);
          timer.mark("vector-P2 ranks");
          return cm;
        };
        if (!ranks_on_device) {
          h->vp2_component_major = host_ranks();
          if (h->d_vp2_rank) {
            hip_check(hipMemcpyAsync(h->d_vp2_rank, rank.data(), rank.size(), hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
          } else {
            h->d_vp2_rank = h->base.upload(rank, s);
          }
        }
        if (h_x) {
          int rc = dfa::kDevicePlanUnsupported;
          std::string derr;
          if (ranks_on_device) {
            dfa::DeviceVP2Input in;
            in.nv = nv;
            in.nc = nc;
            in.ndofs = ndofs;
            in.node_layout = csr.node_layout;
            in.d_cells = h->d_cells;
            in.d_cdofs = h->d_cdofs;
            in.d_kdofs = h->d_kdofs;
            in.h_x = h_x;
            in.d_row_ptr = h->d_row_ptr;
            in.d_rank = h->d_vp2_rank;
            dfa::VP2BlockTopo t;
            int64_t app = 0;
            rc = dfa::build_vp2_plan_device(in, default_p2_block_vals(), h->plan, t, app, derr, s);
            if (rc == dfa::kDevicePlanOk) {
              t.row_ptr = h->d_row_ptr;
              t.cell_dofs = h->d_cdofs;
              h->vtopo = t;
              h->appearances = app;
              h->has_vp2_plan = true;
              h->plan_on_device = true;
              timer.mark("vector-P2 block plan (device)");
            } else {
              h->plan.release();
              if (rc == dfa::kDevicePlanError) h->plan_error = derr;
            }
          }
          if (rc == dfa::kDevicePlanUnsupported || verify_plan) {
            if (rank.empty()) {
              const bool cm = host_ranks();
              if (cm != h->vp2_component_major) throw std::runtime_error("vector P2 layout differs");
            }
            dfa::VP2BlockPlan P;
            bool ok = true;
            std::string herr;
            try {
              dfa::build_vp2_block_plan(nc, h_cells, h_cell_dofs, h_coeff_dofs, h_x, csr, rank,
                                        h->vp2_component_major, default_p2_block_vals(), P);
            } catch (const dfa::HipError&) {
              throw;
            } catch (const std::exception& e) {
              ok = false;
              herr = e.what();
            }
            timer.mark("vector-P2 block plan");
            if (rc != dfa::kDevicePlanUnsupported) {
              h->plan_verify_error = compare_vp2_plans(h, P, rank, ok, herr, rc == dfa::kDevicePlanError, derr);
              if (!h->plan_verify_error.empty())
                std::fprintf(stderr, "[dfa] plan verification: %s\n", h->plan_verify_error.c_str());
            } else if (!ok) {
              h->plan_error = herr;
            } else {
              auto& t = h->vtopo;
              t.num_blocks = P.num_blocks;
              t.max_verts = P.max_verts;
              t.max_owned = P.max_owned;
              t.max_vals = P.max_vals;
              t.component_major = P.component_major;
              t.vert_ptr = h->plan.upload(P.vert_ptr, s);
              t.vert_id = h->plan.upload(P.vert_id, s);
              t.vert_kdof = h->plan.upload(P.vert_kdof, s);
              t.own_ptr = h->plan.upload(P.own_ptr, s);
              t.own_rows = h->plan.upload(P.own_rows, s);
              t.own_info = h->plan.upload(P.own_info, s);
              t.task_ptr = h->plan.upload(P.task_ptr, s);
              t.task = h->plan.upload(P.task, s);
              t.task_nvert = h->plan.upload(P.task_nvert, s);
              t.task_cell = h->plan.upload(P.task_cell, s);
              t.blk_vals = h->plan.upload(P.blk_vals, s);
              t.row_ptr = h->d_row_ptr;
              t.cell_dofs = h->d_cdofs;
              h->appearances = static_cast<int64_t>(P.task_cell.size());
              h->has_vp2_plan = true;
              hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors die after this
            }
          }
        }
      }
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      timer.mark("uploads");
    } catch (...) {
      h->plan.release();
      h->base.release();
      delete h;
      throw;
    }
    *out = h;
    return DFA_OK;
  });
}

int dfa_assembler_destroy(dfa_assembler h) {
  if (!h) return DFA_OK;
  h->plan.release();
  h->base.release();
  h->colors.release();
  h->facets.release();
  h->x0buf.release();
  h->cg.release();
  delete h;
  return DFA_OK;
}

int dfa_csr_dims(dfa_assembler h, int64_t* num_rows, int64_t* nnz) {
  if (!h) return fail("NULL handle", DFA_ERR_ARG);
  if (num_rows) *num_rows = h->ndofs;
  if (nnz) *nnz = h->nnz;
  return DFA_OK;
}

int dfa_csr_structure(dfa_assembler h, int32_t* d_row_ptr, int32_t* d_col, void* stream) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d_row_ptr)
      hip_check(hipMemcpyAsync(d_row_ptr, h->d_row_ptr, (h->ndofs + 1) * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    if (d_col && h->nnz)
      hip_check(hipMemcpyAsync(d_col, h->d_col, h->nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s),
                "hipMemcpyAsync");
    return DFA_OK;
  });
}

static int resolve_scatter(dfa_assembler h, int32_t scatter) {
  if (h->element == DFA_ELEMENT_P1) {
    if (scatter == DFA_SCATTER_AUTO) return h->has_plan ? DFA_SCATTER_BLOCKED : DFA_SCATTER_ATOMIC;
    if (scatter == DFA_SCATTER_BLOCKED && !h->has_plan)
      throw std::runtime_error("blocked scatter unavailable for this mesh: " + h->plan_error);
    if (scatter != DFA_SCATTER_ATOMIC && scatter != DFA_SCATTER_BLOCKED &&
        scatter != DFA_SCATTER_COLORED)
      throw std::runtime_error("unknown scatter strategy");
    return scatter;
  }
  if (scatter == DFA_SCATTER_COLORED)
    throw std::runtime_error("the coloured scatter is implemented for P1 elements");
  if (scatter == DFA_SCATTER_AUTO) return h->has_vp2_plan ? DFA_SCATTER_BLOCKED : DFA_SCATTER_ATOMIC;
  if (scatter == DFA_SCATTER_BLOCKED && !h->has_vp2_plan)
    throw std::runtime_error("blocked scatter unavailable for this mesh: " + h->plan_error);
  if (scatter != DFA_SCATTER_ATOMIC && scatter != DFA_SCATTER_BLOCKED)
    throw std::runtime_error("unknown scatter strategy");
  return scatter;
}

// Colour-ordered cell arrays of the coloured scatter (P1), built on first use: a greedy
// colouring of the cells in which two cells sharing a dof differ (for the vertex dofmap this
// is MeshColoring::color_cells(mesh, "vertex"), dolfin/mesh/MeshColoring.cpp:45-56), cells
// stably sorted by colour, column ranks found once on the device.
static void ensure_colors(dfa_assembler h, hipStream_t s) {
  if (h->has_colors) return;
  const int32_t nvc = h->nvc;
  const size_t n = static_cast<size_t>(h->nc) * nvc;
  std::vector<int32_t> cd(n), kd(n), rp(static_cast<size_t>(h->ndofs) + 1);
  hip_check(hipMemcpyAsync(cd.data(), h->d_cdofs, n * sizeof(int32_t), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(kd.data(), h->d_kdofs, n * sizeof(int32_t), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(rp.data(), h->d_row_ptr, rp.size() * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  for (int64_t r = 0; r < h->ndofs; ++r)
    if (rp[r + 1] - rp[r] > 256)
      throw std::runtime_error("coloured scatter: a CSR row holds more than 256 entries");
  // colours on the device (Jones-Plassmann), else DOLFIN's sequential greedy on the host
  std::vector<int64_t> ptr;
  std::vector<int32_t> perm;
  if (env_flag("DFA_HOST_COLOR") || !dfa::color_cells_device(h->nc, nvc, h->d_cdofs, h->ndofs, perm, ptr, s)) {
    std::vector<int32_t> color(static_cast<size_t>(h->nc));
    const int32_t ncol = dfa::greedy_cell_coloring(h->nc, nvc, cd.data(), h->ndofs, 1, color.data());
    ptr.assign(static_cast<size_t>(ncol) + 1, 0);
    for (int32_t c : color) ++ptr[c + 1];
    for (int32_t k = 0; k < ncol; ++k) ptr[k + 1] += ptr[k];
    perm.assign(static_cast<size_t>(h->nc), 0);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < h->nc; ++c) perm[fill[color[c]]++] = static_cast<int32_t>(c);
  }
  const bool same = cd == kd;
  dfa::DeviceArena tmp;
  try {
    const int32_t* d_perm = tmp.upload(perm, s);
    void* p = nullptr;
    hip_check(hipMalloc(&p, n * sizeof(int32_t)), "hipMalloc");
    h->d_ccells = h->colors.adopt(static_cast<int32_t*>(p), n * sizeof(int32_t));
    hip_check(hipMalloc(&p, n * sizeof(int32_t)), "hipMalloc");
    h->d_cdofs_c = h->colors.adopt(static_cast<int32_t*>(p), n * sizeof(int32_t));
    if (same) {
      h->d_kdofs_c = h->d_cdofs_c;
    } else {
      hip_check(hipMalloc(&p, n * sizeof(int32_t)), "hipMalloc");
      h->d_kdofs_c = h->colors.adopt(static_cast<int32_t*>(p), n * sizeof(int32_t));
    }
    hip_check(hipMalloc(&p, n * nvc), "hipMalloc");
    h->d_crank = h->colors.adopt(static_cast<uint8_t*>(p), n * nvc);
    hip_check(dfa::launch_p1_color_gather(h->nc, nvc, d_perm, h->d_cells, h->d_cdofs,
                                          same ? nullptr : h->d_kdofs, h->d_row_ptr, h->d_col,
                                          h->d_ccells, h->d_cdofs_c, same ? nullptr : h->d_kdofs_c,
                                          h->d_crank, s),
              "p1_color_gather_kernel");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  } catch (...) {
    tmp.release();
    h->colors.release();
    throw;
  }
  tmp.release();
  h->color_ptr = std::move(ptr);
  h->has_colors = true;
}

// The Dirichlet rows of an assembled system: dirichlet_node_rows_kernel and its pull pass on a
// vector-P2 structure expanded from the node graph, else dirichlet_rows_kernel (the cells-per-dof
// counts and the per-node flags are allocated at creation, so this allocates nothing).
static bool node_dirichlet(dfa_assembler h) {
  return h->node_layout && h->d_node_bits && !env_flag("DFA_BC_ROWS");
}

static void apply_dirichlet_rows(dfa_assembler h, const uint8_t* bcm, const double* bcv, double* values,
                                 double* b, hipStream_t s) {
  dfa::CsrView A;
  A.nrows = h->ndofs;
  A.nnz = h->nnz;
  A.row_ptr = h->d_row_ptr;
  A.col = h->d_col;
  if (node_dirichlet(h)) {
    A.node_layout = h->node_layout;
    A.node_nnz = h->node_nnz;
    A.node_rp = h->d_node_rp;
    A.node_col = h->d_node_col;
    hip_check(dfa::launch_dirichlet_nodes(A, bcm, bcv, h->d_dof_cells, h->d_node_bits, values, b, s),
              "dirichlet_node_rows_kernel");
    return;
  }
  hip_check(dfa::launch_dirichlet_rows(A, h->d_tmap, bcm, bcv, h->d_dof_cells, values, b, s),
            "dirichlet_rows_kernel");
}

// One fused pass over the cells (Assembler::assemble_cells / SystemAssembler::cell_wise_assembly).
// With name_only non-NULL nothing is launched: the name of the kernel the pass would run is
// written there instead (pointer arguments are then only tested for NULL).

static int run_pass(dfa_assembler h, int32_t a_form, int32_t L_form, const double* x,
                    const double* a_c0, const double* a_c1, const double* L_c0, const double* f3,
                    const uint8_t* bcm, const double* bcv, double* values, double* b,
                    int32_t scatter, int32_t flags, hipStream_t s,
                    std::string* name_only = nullptr) {
  if (!x) throw std::runtime_error("coordinates are NULL");
  const bool add = (flags & DFA_ADD_VALUES) != 0;
  const bool bc = bcm != nullptr;
  if (bc && !bcv) throw std::runtime_error("bc_mask given without bc_value");
  if (h->element == DFA_ELEMENT_P1) {
    if (a_form >= 0 && a_form != DFA_FORM_P1_POISSON && a_form != DFA_FORM_P1_MASS)
      throw std::runtime_error("form does not match the P1 element");
    if (L_form >= 0 && L_form != DFA_FORM_P1_LOAD)
      throw std::runtime_error("linear form does not match the P1 element");
    if (a_form == DFA_FORM_P1_POISSON && !a_c0) throw std::runtime_error("kappa is NULL");
    if (L_form == DFA_FORM_P1_LOAD && !L_c0) throw std::runtime_error("f is NULL");
    if (bc && a_form < 0) throw std::runtime_error("boundary conditions need the bilinear form");
    dfa::P1Pass p;
    p.gdim = h->gdim;
    p.a_form = a_form;
    p.L_form = L_form;
    p.want_A = values != nullptr;
    p.want_b = b != nullptr;
    p.bc = bc;
    p.add = add;
    p.x = x;
    p.kappa = a_c0;
    p.f = L_c0;
    p.bc_mask = bcm;
    p.bc_value = bcv;
    p.values = values;
    p.b = b;
    const int sc = resolve_scatter(h, scatter);
    // Dirichlet conditions after assembly: the edge-chain kernel assembles the system without
    // them, then dirichlet_rows_kernel applies what the per-cell elimination sums up to (not
    // for add_values, whose earlier values must stay untouched; DFA_BC_INLINE=1 keeps the
    // per-cell elimination inside the blocked kernel)
    dfa::P1Pass q = p;
    q.bc = false;
    q.bc_mask = nullptr;
    q.bc_value = nullptr;
    const bool bc_after = bc && !add && values && (!b || L_form >= 0) && sc == DFA_SCATTER_BLOCKED &&
                          h->has_plan && h->d_tmap &&
                          dfa::p1_chains_apply(h->btopo, q) && !env_flag("DFA_BC_INLINE");
    if (name_only) {
      if (sc == DFA_SCATTER_ATOMIC) *name_only = "p1_atomic_kernel";
      else if (sc == DFA_SCATTER_COLORED) *name_only = "p1_colored_kernel";
      else if (bc_after) *name_only = dfa::p1_blocked_kernel_name(h->btopo, q) + " + dirichlet_rows_kernel";
      else *name_only = dfa::p1_blocked_kernel_name(h->btopo, p);
      return DFA_OK;
    }
    if (bc_after) {
      if (!h->rows_covered) {
        hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
        if (b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
      }
      hip_check(dfa::launch_p1_blocked(h->btopo, q, s), "p1_chain_kernel");
      apply_dirichlet_rows(h, bcm, bcv, values, b, s);
      return DFA_OK;
    }
    if (sc == DFA_SCATTER_ATOMIC) {
      if (!add && values) hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
      if (!add && b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
      dfa::P1AtomicTopo t;
      t.num_cells = h->nc;
      t.cells = h->d_cells;
      t.cell_dofs = h->d_cdofs;
      t.coeff_dofs = h->d_kdofs;
      t.row_ptr = h->d_row_ptr;
      t.col = h->d_col;
      hip_check(dfa::launch_p1_atomic(t, p, s), "p1_atomic_kernel");
    } else if (sc == DFA_SCATTER_COLORED) {
      ensure_colors(h, s);
      if (!add && values) hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
      if (!add && b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
      dfa::P1ColorTopo t;
      t.cells = h->d_ccells;
      t.cell_dofs = h->d_cdofs_c;
      t.coeff_dofs = h->d_kdofs_c;
      t.rank = h->d_crank;
      t.row_ptr = h->d_row_ptr;
      for (size_t k = 0; k + 1 < h->color_ptr.size(); ++k) {
        t.begin = h->color_ptr[k];
        t.end = h->color_ptr[k + 1];
        hip_check(dfa::launch_p1_colored(t, p, s), "p1_colored_kernel");
      }
    } else {
      if (!add && !h->rows_covered) {
        if (values) hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
        if (b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
      }
      const size_t lds = dfa::p1_blocked_lds_bytes(h->btopo, p);
      if (lds > 160 * 1024) throw std::runtime_error("blocked plan needs more LDS than a CU has");
      hip_check(dfa::launch_p1_blocked(h->btopo, p, s), "p1_blocked_kernel");
    }
    return DFA_OK;
  }
  // vector P2
  if (a_form >= 0 && a_form != DFA_FORM_VP2_ELASTICITY)
    throw std::runtime_error("form does not match the vector P2 element");
  if (L_form >= 0 && L_form != DFA_FORM_VP2_BODY_FORCE)
    throw std::runtime_error("linear form does not match the vector P2 element");
  if (a_form >= 0 && (!a_c0 || !a_c1)) throw std::runtime_error("mu/lmbda are NULL");
  if (bc && a_form < 0) throw std::runtime_error("boundary conditions need the bilinear form");
  const int sc = resolve_scatter(h, scatter);
  dfa::VP2Pass p;
  p.want_A = values != nullptr && a_form >= 0;
  p.want_b = b != nullptr && L_form >= 0;
  p.bc = bc;
  p.x = x;
  p.mu = a_c0;
  p.lmbda = a_c1;
  if (f3) for (int i = 0; i < 3; ++i) p.f3[i] = f3[i];
  p.bc_mask = bcm;
  p.bc_value = bcv;
  p.values = values;
  p.b = b;
  p.add = add;
  // Dirichlet conditions after the blocked pass, on the assembled rows (what the per-cell
  // elimination sums up to), as for P1
  dfa::VP2Pass q = p;
  q.bc = false;
  q.bc_mask = nullptr;
  q.bc_value = nullptr;
  const bool bc_after = bc && !add && p.want_A && (!b || p.want_b) && sc == DFA_SCATTER_BLOCKED &&
                        h->has_vp2_plan && h->d_dof_cells && !env_flag("DFA_BC_INLINE");
  if (name_only) {
    *name_only = bc_after ? dfa::vp2_kernel_name(true, q) +
                                (node_dirichlet(h) ? " + dirichlet_node_rows_kernel" : " + dirichlet_rows_kernel")
                          : dfa::vp2_kernel_name(sc == DFA_SCATTER_BLOCKED, p);
    return DFA_OK;
  }
  if (bc_after) {
    if (!h->rows_covered) {
      hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
      if (b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
    }
    if (dfa::vp2_blocked_lds_bytes(h->vtopo, q.want_b) > 160 * 1024)
      throw std::runtime_error("vector P2 blocked plan needs more LDS than a CU has");
    hip_check(dfa::launch_vp2_blocked(h->vtopo, q, s), "vp2_blocked_kernel");
    apply_dirichlet_rows(h, bcm, bcv, values, b, s);
    return DFA_OK;
  }
  if (sc == DFA_SCATTER_BLOCKED) {
    if (!add && !h->rows_covered) {
      if (values) hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
      if (b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
    }
    if (dfa::vp2_blocked_lds_bytes(h->vtopo, p.want_b) > 160 * 1024)
      throw std::runtime_error("vector P2 blocked plan needs more LDS than a CU has");
    hip_check(dfa::launch_vp2_blocked(h->vtopo, p, s), "vp2_blocked_kernel");
    return DFA_OK;
  }
  if (!add && values) hip_check(dfa::launch_fill_zero(values, h->nnz, s), "fill_zero");
  if (!add && b) hip_check(dfa::launch_fill_zero(b, h->ndofs, s), "fill_zero");
  dfa::VP2Topo t;
  t.num_cells = h->nc;
  t.cells = h->d_cells;
  t.cell_dofs = h->d_cdofs;
  t.coeff_dofs = h->d_kdofs;
  t.row_ptr = h->d_row_ptr;
  t.rank = h->d_vp2_rank;
  t.component_major = h->vp2_component_major;
  hip_check(dfa::launch_vp2_atomic(t, p, s), "vp2_kernel");
  return DFA_OK;
}

int dfa_assemble_matrix(dfa_assembler h, int32_t form, const double* d_x, const double* d_c0,
                        const double* d_c1, double* d_values, int32_t scatter, int32_t flags,
                        void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_values) return fail("NULL handle or values", DFA_ERR_ARG);
    return run_pass(h, form, -1, d_x, d_c0, d_c1, nullptr, nullptr, nullptr, nullptr, d_values,
                    nullptr, scatter, flags, static_cast<hipStream_t>(stream));
  });
}

int dfa_assemble_vector(dfa_assembler h, int32_t form, const double* d_x, const double* d_c0,
                        const double* h_f3, double* d_b, int32_t scatter, int32_t flags,
                        void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_b) return fail("NULL handle or vector", DFA_ERR_ARG);
    if (form == DFA_FORM_VP2_BODY_FORCE && !h_f3) return fail("f3 is NULL", DFA_ERR_ARG);
    return run_pass(h, -1, form, d_x, nullptr, nullptr, d_c0, h_f3, nullptr, nullptr, nullptr,
                    d_b, scatter, flags, static_cast<hipStream_t>(stream));
  });
}

int dfa_assemble_system(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                        const double* d_a_c0, const double* d_a_c1, const double* d_L_c0,
                        const double* h_f3, const uint8_t* d_bc_mask, const double* d_bc_value,
                        double* d_values, double* d_b, int32_t scatter, int32_t flags,
                        void* stream) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    if (!d_values && !d_b) return fail("nothing to assemble (A and b are NULL)", DFA_ERR_ARG);
    if (a_form < 0 || L_form < 0) return fail("assemble_system needs a bilinear and a linear form", DFA_ERR_ARG);
    if (L_form == DFA_FORM_VP2_BODY_FORCE && !h_f3) return fail("f3 is NULL", DFA_ERR_ARG);
    return run_pass(h, a_form, L_form, d_x, d_a_c0, d_a_c1, d_L_c0, h_f3, d_bc_mask, d_bc_value,
                    d_values, d_b, scatter, flags, static_cast<hipStream_t>(stream));
  });
}

int dfa_dirichlet_x0(int64_t n, const uint8_t* d_bc_mask, const double* d_bc_value,
                     const double* d_x0, double* d_out, void* stream) {
  return guarded([&]() -> int {
    if (n < 0 || (n > 0 && (!d_bc_mask || !d_bc_value || !d_x0 || !d_out)))
      return fail("invalid arguments", DFA_ERR_ARG);
    hip_check(dfa::launch_dirichlet_x0(n, d_bc_mask, d_bc_value, d_x0, d_out,
                                       static_cast<hipStream_t>(stream)),
              "dirichlet_x0_kernel");
    return DFA_OK;
  });
}

int dfa_assemble_system_x0(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                           const double* d_a_c0, const double* d_a_c1, const double* d_L_c0,
                           const double* h_f3, const uint8_t* d_bc_mask, const double* d_bc_value,
                           const double* d_x0, double* d_values, double* d_b, int32_t scatter,
                           int32_t flags, void* stream) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    if (!d_x0) return fail("x0 is NULL (use dfa_assemble_system)", DFA_ERR_ARG);
    if (!d_values && !d_b) return fail("nothing to assemble (A and b are NULL)", DFA_ERR_ARG);
    if (a_form < 0 || L_form < 0) return fail("assemble_system needs a bilinear and a linear form", DFA_ERR_ARG);
    if (L_form == DFA_FORM_VP2_BODY_FORCE && !h_f3) return fail("f3 is NULL", DFA_ERR_ARG);
    if ((d_bc_mask == nullptr) != (d_bc_value == nullptr))
      return fail("Dirichlet mask and values go together", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double* bcv = d_bc_value;
    if (d_bc_mask) {
      // SystemAssembler.cpp:252-275: every boundary value g becomes x0[dof] - g
      if (!h->d_bc_x0) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<int64_t>(h->ndofs, 1) * sizeof(double)), "hipMalloc");
        h->d_bc_x0 = h->x0buf.adopt(static_cast<double*>(p), h->ndofs * sizeof(double));
      }
      hip_check(dfa::launch_dirichlet_x0(h->ndofs, d_bc_mask, d_bc_value, d_x0, h->d_bc_x0, s),
                "dirichlet_x0_kernel");
      bcv = h->d_bc_x0;
    }
    return run_pass(h, a_form, L_form, d_x, d_a_c0, d_a_c1, d_L_c0, h_f3, d_bc_mask, bcv,
                    d_values, d_b, scatter, flags, s);
  });
}

}  // extern "C"

// Device array of the allocation-free entry point: the shifted Dirichlet values of
// dfa_assemble_system_x0_noalloc. Each GPU holds its own copy until the process ends. In host
// code the name is only a stand-in for the symbol; its device address on each device comes
// from hipGetSymbolAddress (kept per device below) and is the only form passed to kernels.
namespace {
static __device__ __align__(256) double scratch[DFA_NOALLOC_MAX_DOFS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (resolved once per device).
double* scratch_on_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int dfa_assemble_system_x0_noalloc(dfa_assembler h, int32_t a_form, int32_t L_form,
                                   const double* d_x, const double* d_a_c0, const double* d_a_c1,
                                   const double* d_L_c0, const double* h_f3,
                                   const uint8_t* d_bc_mask, const double* d_bc_value,
                                   const double* d_x0, double* d_values, double* d_b,
                                   int32_t scatter, int32_t flags, void* stream) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    if (!d_x0) return fail("x0 is NULL (use dfa_assemble_system)", DFA_ERR_ARG);
    if (!d_values && !d_b) return fail("nothing to assemble (A and b are NULL)", DFA_ERR_ARG);
    if (a_form < 0 || L_form < 0) return fail("assemble_system needs a bilinear and a linear form", DFA_ERR_ARG);
    if (L_form == DFA_FORM_VP2_BODY_FORCE && !h_f3) return fail("f3 is NULL", DFA_ERR_ARG);
    if ((d_bc_mask == nullptr) != (d_bc_value == nullptr))
      return fail("Dirichlet mask and values go together", DFA_ERR_ARG);
    if (h->ndofs > DFA_NOALLOC_MAX_DOFS)
      return fail("dfa_assemble_system_x0_noalloc: " + std::to_string(h->ndofs) +
                      " dofs exceed its fixed device array of " +
                      std::to_string(DFA_NOALLOC_MAX_DOFS) + " (use dfa_assemble_system_x0)",
                  DFA_ERR_LIMIT);
    // the coloured scatter builds its colour plan (device memory) on first use
    if (h->element == DFA_ELEMENT_P1 && scatter == DFA_SCATTER_COLORED && !h->has_colors)
      return fail("dfa_assemble_system_x0_noalloc: the coloured scatter allocates its colour plan "
                  "on first use; run one assembly with it through an allocating entry point first",
                  DFA_ERR_STATE);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double* bcv = d_bc_value;
    if (d_bc_mask) {
      // every entry is written (x0 - g on Dirichlet dofs, 0 elsewhere): nothing of an earlier
      // call's contents is read
      double* ptr = scratch_on_current_device();
      hip_check(dfa::launch_dirichlet_x0(h->ndofs, d_bc_mask, d_bc_value, d_x0, ptr, s),
                "dirichlet_x0_kernel");
      bcv = ptr;
    }
    return run_pass(h, a_form, L_form, d_x, d_a_c0, d_a_c1, d_L_c0, h_f3, d_bc_mask, bcv,
                    d_values, d_b, scatter, flags, s);
  });
}

int dfa_plan_builder(dfa_assembler h, int32_t* on_device) {
  return guarded([&]() -> int {
    if (!h || !on_device) return fail("NULL argument", DFA_ERR_ARG);
    *on_device = (h->has_plan || h->has_vp2_plan) && h->plan_on_device ? 1 : 0;
    if (!h->plan_verify_error.empty())
      return fail("device plan differs from the host plan: " + h->plan_verify_error, DFA_ERR_STATE);
    return DFA_OK;
  });
}

int dfa_pass_kernel(dfa_assembler h, int32_t a_form, int32_t L_form, int32_t has_bc,
                    int32_t scatter, char* name, int32_t len) {
  return guarded([&]() -> int {
    if (!h || !name || len <= 0) return fail("NULL handle or name buffer", DFA_ERR_ARG);
    // stand-in non-NULL pointers: only their presence selects the kernel; nothing is read
    static const double dummy[3] = {0, 0, 0};
    static const uint8_t dmask = 0;
    double* w = const_cast<double*>(dummy);
    std::string out;
    run_pass(h, a_form, L_form, dummy, dummy, dummy, dummy, dummy, has_bc ? &dmask : nullptr,
             has_bc ? dummy : nullptr, a_form >= 0 ? w : nullptr, L_form >= 0 ? w : nullptr,
             scatter, 0, nullptr, &out);
    std::snprintf(name, static_cast<size_t>(len), "%s", out.c_str());
    return DFA_OK;
  });
}

int dfa_adjoint(dfa_assembler h, int32_t a_form, const double* d_x, const double* d_c0,
                const double* d_c1, const double* d_f, const double* h_f3, const double* d_u,
                double* d_J, double* d_dJdc0, double* d_dJdc1, double* d_dJdx, void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_x || !d_u || !d_J || !d_dJdc0 || !d_dJdx) return fail("NULL argument", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hip_check(dfa::launch_fill_zero(d_J, 1, s), "fill_zero");
    // the blocked P1 adjoint writes every owned node's gradients once; when every vertex is a
    // row of some block (one dof and one coefficient dof per vertex, each in a cell), nothing
    // is left for a fill
    const bool p1_blocked = h->element == DFA_ELEMENT_P1 && h->has_plan && h->kdof_injective;
    if (!(p1_blocked && h->rows_covered && h->ndofs == h->nv && h->nkdofs == h->nv)) {
      hip_check(dfa::launch_fill_zero(d_dJdc0, h->nkdofs, s), "fill_zero");
      hip_check(dfa::launch_fill_zero(d_dJdx, h->nv * h->gdim, s), "fill_zero");
    }
    if (h->element == DFA_ELEMENT_P1) {
      if (a_form != DFA_FORM_P1_POISSON) return fail("P1 adjoint supports the kappa-Poisson form", DFA_ERR_ARG);
      if (!d_c0 || !d_f) return fail("kappa or f is NULL", DFA_ERR_ARG);
      dfa::AdjointP1 a;
      a.gdim = h->gdim;
      a.num_cells = h->nc;
      a.cells = h->d_cells;
      a.cell_dofs = h->d_cdofs;
      a.coeff_dofs = h->d_kdofs;
      a.x = d_x;
      a.kappa = d_c0;
      a.f = d_f;
      a.u = d_u;
      a.J = d_J;
      a.dJdk = d_dJdc0;
      a.dJdx = d_dJdx;
      if (p1_blocked)
        hip_check(dfa::launch_p1_adjoint_blocked(h->btopo, a, s), "p1_adjoint_blocked_kernel");
      else
        hip_check(dfa::launch_adjoint_p1(a, s), "adjoint_p1_kernel");
      return DFA_OK;
    }
    if (a_form != DFA_FORM_VP2_ELASTICITY) return fail("vector P2 adjoint supports elasticity", DFA_ERR_ARG);
    if (!d_c0 || !d_c1 || !d_dJdc1 || !h_f3) return fail("mu, lmbda, dJ/dlmbda or f3 is NULL", DFA_ERR_ARG);
    hip_check(dfa::launch_fill_zero(d_dJdc1, h->nkdofs, s), "fill_zero");
    dfa::AdjointVP2 a;
    a.num_cells = h->nc;
    a.cells = h->d_cells;
    a.cell_dofs = h->d_cdofs;
    a.coeff_dofs = h->d_kdofs;
    a.x = d_x;
    a.mu = d_c0;
    a.lmbda = d_c1;
    for (int i = 0; i < 3; ++i) a.f3[i] = h_f3[i];
    a.u = d_u;
    a.J = d_J;
    a.dJdmu = d_dJdc0;
    a.dJdlmbda = d_dJdc1;
    a.dJdx = d_dJdx;
    hip_check(dfa::launch_adjoint_vp2(a, s), "adjoint_vp2_kernel");
    return DFA_OK;
  });
}

namespace {
void ensure_facets(dfa_assembler h, hipStream_t s) {
  if (h->has_facets) return;
  if (h->element != DFA_ELEMENT_P1 && h->element != DFA_ELEMENT_VECTOR_P2)
    throw std::runtime_error("exterior-facet forms need a P1 or vector-P2 space");
  std::vector<int32_t> cells(static_cast<size_t>(h->nc) * h->nvc);
  if (!cells.empty()) {
    hip_check(hipMemcpyAsync(cells.data(), h->d_cells, cells.size() * sizeof(int32_t),
                             hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }
  dfa::build_exterior_facets(h->gdim, h->nc, cells.data(), h->facet_cell, h->facet_local);
  h->d_facet_cell = h->facets.upload(h->facet_cell, s);
  h->d_facet_local = h->facets.upload(h->facet_local, s);
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  h->has_facets = true;
}
}  // namespace

int dfa_node_reordering(int64_t num_cells, int32_t nodes_per_cell, const int32_t* cell_nodes,
                        int64_t num_nodes, int32_t* old_to_new) {
  return guarded([&]() -> int {
    if (num_cells < 0 || nodes_per_cell <= 0 || num_nodes < 0 || (num_cells > 0 && !cell_nodes) ||
        (num_nodes > 0 && !old_to_new))
      return fail("invalid arguments", DFA_ERR_ARG);
    if (num_nodes > INT32_MAX) return fail("sizes exceed int32", DFA_ERR_ARG);
    if (num_nodes == 0) return DFA_OK;
    for (int64_t i = 0; i < num_cells * nodes_per_cell; ++i)
      if (cell_nodes[i] < 0 || cell_nodes[i] >= num_nodes) return fail("cell node out of range", DFA_ERR_ARG);
    if (rcm_on_device() && num_cells > 0 && rcm_nodes_device(num_cells, nodes_per_cell, cell_nodes, num_nodes,
                                                              old_to_new))
      return DFA_OK;
    if (env_flag("DFA_RCM_REQUIRE_DEVICE")) return fail("the device reordering did not apply", DFA_ERR_STATE);
    // node graph: the sparsity of cell_nodes x cell_nodes without the diagonal
    dfa::HostCSR g;
    dfa::build_csr(num_cells, nodes_per_cell, cell_nodes, num_nodes, false, g);
    std::vector<int32_t> ptr(static_cast<size_t>(num_nodes) + 1, 0), adj;
    adj.reserve(g.col.size());
    for (int64_t v = 0; v < num_nodes; ++v) {
      for (int32_t k = g.row_ptr[v]; k < g.row_ptr[v + 1]; ++k)
        if (g.col[k] != v) adj.push_back(g.col[k]);
      ptr[v + 1] = static_cast<int32_t>(adj.size());
    }
    dfa::rcm_ordering(num_nodes, ptr.data(), adj.data(), old_to_new);
    return DFA_OK;
  });
}

int dfa_graph_reordering(int64_t num_nodes, const int32_t* graph_ptr, const int32_t* graph_adj,
                         int32_t* old_to_new) {
  return guarded([&]() -> int {
    if (num_nodes < 0 || (num_nodes > 0 && (!graph_ptr || !old_to_new)))
      return fail("invalid arguments", DFA_ERR_ARG);
    if (num_nodes > INT32_MAX) return fail("sizes exceed int32", DFA_ERR_ARG);
    if (num_nodes == 0) return DFA_OK;
    if (graph_ptr[0] != 0) return fail("graph_ptr[0] must be 0", DFA_ERR_ARG);
    for (int64_t v = 0; v < num_nodes; ++v) {
      if (graph_ptr[v + 1] < graph_ptr[v]) return fail("graph_ptr must be non-decreasing", DFA_ERR_ARG);
      for (int32_t k = graph_ptr[v]; k < graph_ptr[v + 1]; ++k) {
        const int32_t u = graph_adj[k];
        if (u < 0 || u >= num_nodes || u == v || (k > graph_ptr[v] && graph_adj[k - 1] >= u))
          return fail("adjacency lists must be sorted, in range and without self loops", DFA_ERR_ARG);
      }
    }
    if (rcm_on_device() && graph_ptr[num_nodes] > 0) {
      int32_t maxd = 0;
      for (int64_t v = 0; v < num_nodes; ++v) maxd = std::max(maxd, graph_ptr[v + 1] - graph_ptr[v]);
      hipStream_t s = nullptr;
      dfa::DeviceArena a;
      try {
        const int32_t* dap = a.upload(graph_ptr, static_cast<size_t>(num_nodes) + 1, s);
        const int32_t* dadj = a.upload(graph_adj, static_cast<size_t>(graph_ptr[num_nodes]), s);
        const bool ok = dfa::rcm_ordering_device(num_nodes, dap, dadj, maxd, old_to_new, s);
        a.release();
        if (ok) return DFA_OK;
      } catch (...) {
        a.release();
        throw;
      }
    }
    if (env_flag("DFA_RCM_REQUIRE_DEVICE") && graph_ptr[num_nodes] > 0)
      return fail("the device reordering did not apply", DFA_ERR_STATE);
    dfa::rcm_ordering(num_nodes, graph_ptr, graph_adj, old_to_new);
    return DFA_OK;
  });
}

int dfa_exterior_facets(dfa_assembler h, int64_t* num_facets, int32_t* facet_cell,
                        int32_t* facet_local) {
  return guarded([&]() -> int {
    if (!h || !num_facets) return fail("NULL argument", DFA_ERR_ARG);
    ensure_facets(h, nullptr);
    *num_facets = static_cast<int64_t>(h->facet_cell.size());
    if (facet_cell) std::copy(h->facet_cell.begin(), h->facet_cell.end(), facet_cell);
    if (facet_local)
      for (size_t i = 0; i < h->facet_local.size(); ++i) facet_local[i] = h->facet_local[i];
    return DFA_OK;
  });
}

int dfa_assemble_facets(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                        double alpha, const double* d_g, const uint8_t* d_facet_mask,
                        const uint8_t* d_bc_mask, const double* d_bc_value, double* d_values,
                        double* d_b, void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_x) return fail("NULL handle or coordinates", DFA_ERR_ARG);
    const bool vp2 = h->element == DFA_ELEMENT_VECTOR_P2;
    if (vp2) {
      if (a_form != -1 || L_form != DFA_FORM_VP2_FACET_TRACTION)
        return fail("a vector-P2 space takes a_form -1 and L_form DFA_FORM_VP2_FACET_TRACTION", DFA_ERR_ARG);
    } else {
      if (a_form != -1 && a_form != DFA_FORM_P1_FACET_MASS)
        return fail("facet bilinear form must be DFA_FORM_P1_FACET_MASS or -1", DFA_ERR_ARG);
      if (L_form != -1 && L_form != DFA_FORM_P1_FACET_LOAD)
        return fail("facet linear form must be DFA_FORM_P1_FACET_LOAD or -1", DFA_ERR_ARG);
    }
    if (L_form >= 0 && !d_g) return fail("g is NULL", DFA_ERR_ARG);
    if ((d_bc_mask == nullptr) != (d_bc_value == nullptr))
      return fail("Dirichlet mask and values go together", DFA_ERR_ARG);
    if (!d_values && !d_b) return fail("nothing to assemble (A and b are NULL)", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ensure_facets(h, s);
    dfa::FacetTopo t;
    t.num_facets = static_cast<int64_t>(h->facet_cell.size());
    t.facet_cell = h->d_facet_cell;
    t.facet_local = h->d_facet_local;
    t.cells = h->d_cells;
    t.cell_dofs = h->d_cdofs;
    t.coeff_dofs = h->d_kdofs;
    t.row_ptr = h->d_row_ptr;
    t.col = h->d_col;
    dfa::FacetPass p;
    p.gdim = h->gdim;
    p.has_a = a_form >= 0;
    p.has_b = L_form >= 0;
    p.alpha = alpha;
    p.x = d_x;
    p.g = d_g;
    p.mask = d_facet_mask;
    p.bc_mask = d_bc_mask;
    p.bc_value = d_bc_value;
    p.values = p.has_a ? d_values : nullptr;
    p.b = d_b;
    p.vp2_traction = vp2;
    hip_check(dfa::launch_facets(t, p, s), "facet_kernel");
    return DFA_OK;
  });
}

int dfa_csr_spmv(dfa_assembler h, const double* d_values, double alpha, const double* d_x,
                 double beta, double* d_y, void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_values || !d_x || !d_y) return fail("NULL argument", DFA_ERR_ARG);
    hip_check(dfa::launch_spmv(spmv_view(h, d_values), alpha, d_x, beta, d_y, static_cast<hipStream_t>(stream)),
              "spmv_kernel");
    return DFA_OK;
  });
}

int dfa_cg_solve(dfa_assembler h, const double* d_values, double* d_x, const double* d_b, int32_t jacobi,
                 double rtol, double atol, int64_t maxiter, int32_t check_every, int64_t* iterations,
                 void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_values || !d_x || !d_b || !iterations) return fail("NULL argument", DFA_ERR_ARG);
    if (check_every < 1) return fail("check_every must be at least 1", DFA_ERR_ARG);
    if (!(rtol >= 0.0) || !(atol >= 0.0)) return fail("tolerances must be non-negative", DFA_ERR_ARG);
    *iterations = 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = h->ndofs;
    if (maxiter < 0) maxiter = 10 * n;
    if (!h->d_cg_r) {  // work vectors, kept with the handle
      auto vec = [&](int64_t m) {
        void* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(m) * sizeof(double), 16)), "hipMalloc");
        return h->cg.adopt(static_cast<double*>(d), m * 8);
      };
      h->d_cg_r = vec(n);
      h->d_cg_p = vec(n);
      h->d_cg_ap = vec(n);
      h->d_cg_dinv = vec(n);
      h->d_cg_part = vec(3 * dfa::kCgMaxGroups);
      void* d = nullptr;
      hip_check(hipMalloc(&d, sizeof(dfa::CgState)), "hipMalloc");
      h->d_cg_state = h->cg.adopt(static_cast<dfa::CgState*>(d), sizeof(dfa::CgState));
    }
    const dfa::CsrView A = spmv_view(h, d_values);
    double* r = h->d_cg_r;
    const double* dinv = jacobi ? h->d_cg_dinv : nullptr;
    // r = b - A x; the preconditioner diag(A)^-1 (1 where the diagonal is 0)
    hip_check(hipMemcpyAsync(r, d_b, n * sizeof(double), hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    hip_check(dfa::launch_spmv(A, -1.0, d_x, 1.0, r, s), "spmv_kernel");
    if (jacobi) {
      hip_check(dfa::launch_diagonal(A, h->d_cg_dinv, s), "diagonal_kernel");
      hip_check(dfa::launch_cg_jacobi(n, h->d_cg_dinv, s), "cg_jacobi_kernel");
    }
    dfa::CgState hs{};
    hs.done = -1;
    hip_check(hipMemcpyAsync(h->d_cg_state, &hs, sizeof hs, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(dfa::launch_cg_init(n, d_b, r, dinv, h->d_cg_p, h->d_cg_part, h->d_cg_state, s), "cg_init_kernel");
    hip_check(hipMemcpyAsync(&hs, h->d_cg_state, sizeof hs, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    const double tol = std::max(rtol * std::sqrt(hs.bb), atol);
    const double tol2 = tol * tol;
    if (hs.rr <= tol2) return DFA_OK;
    hip_check(hipMemcpyAsync(&h->d_cg_state->tol2, &tol2, sizeof tol2, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync");
    for (int64_t it = 1; it <= maxiter; ++it) {
      hip_check(dfa::launch_cg_iteration(A, it, d_x, h->d_cg_p, r, h->d_cg_ap, dinv, h->d_cg_part,
                                         h->d_cg_state, s),
                "cg iteration");
      if (it % check_every == 0 || it == maxiter) {
        hip_check(hipMemcpyAsync(&hs, h->d_cg_state, sizeof hs, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (hs.done >= 0) {
          *iterations = hs.done;
          return DFA_OK;
        }
        if (hs.broke) {
          *iterations = it;
          return fail("conjugate gradients broke down (p.Ap = 0) at iteration <= " + std::to_string(it),
                      DFA_ERR_STATE);
        }
      }
    }
    *iterations = maxiter;
    return fail("conjugate gradients did not converge in " + std::to_string(maxiter) + " iterations",
                DFA_ERR_STATE);
  });
}

int dfa_csr_add_local(dfa_assembler h, double* d_values, int32_t m, const int32_t* h_rows,
                      int32_t n, const int32_t* h_cols, const double* h_block, void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_values || m < 0 || n < 0) return fail("invalid arguments", DFA_ERR_ARG);
    const int64_t nk = static_cast<int64_t>(m) * n;
    if (nk == 0) return DFA_OK;
    if (!h_rows || !h_cols || !h_block) return fail("NULL block or index array", DFA_ERR_ARG);
    for (int32_t i = 0; i < m; ++i)
      if (h_rows[i] < 0 || h_rows[i] >= h->ndofs) return fail("row index out of range", DFA_ERR_ARG);
    for (int32_t j = 0; j < n; ++j)
      if (h_cols[j] < 0 || h_cols[j] >= h->ndofs) return fail("column index out of range", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    dfa::DeviceArena tmp;
    try {
      const int32_t* d_rows = tmp.upload(h_rows, static_cast<size_t>(m), s);
      const int32_t* d_cols = tmp.upload(h_cols, static_cast<size_t>(n), s);
      const double* d_block = tmp.upload(h_block, static_cast<size_t>(nk), s);
      std::vector<int32_t> slot(static_cast<size_t>(nk));
      int32_t* d_slot = tmp.upload(slot, s);
      hip_check(dfa::launch_block_slots(m, d_rows, n, d_cols, h->d_row_ptr, h->d_col, d_slot, s),
                "block_slots_kernel");
      hip_check(hipMemcpyAsync(slot.data(), d_slot, nk * sizeof(int32_t), hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      for (int64_t k = 0; k < nk; ++k)
        if (slot[k] < 0) {
          tmp.release();
          return fail("entry (" + std::to_string(h_rows[k / n]) + ", " + std::to_string(h_cols[k % n]) +
                          ") is not in the sparsity pattern",
                      DFA_ERR_ARG);
        }
      hip_check(dfa::launch_block_add(nk, d_slot, d_block, d_values, s), "block_add_kernel");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    } catch (...) {
      tmp.release();
      throw;
    }
    tmp.release();
    return DFA_OK;
  });
}

int dfa_vector_add_local(double* d_b, int64_t size, int32_t m, const int32_t* h_rows,
                         const double* h_block, void* stream) {
  return guarded([&]() -> int {
    if (!d_b || size < 0 || m < 0) return fail("invalid arguments", DFA_ERR_ARG);
    if (m == 0) return DFA_OK;
    if (!h_rows || !h_block) return fail("NULL block or index array", DFA_ERR_ARG);
    for (int32_t i = 0; i < m; ++i)
      if (h_rows[i] < 0 || h_rows[i] >= size) return fail("row index out of range", DFA_ERR_ARG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    dfa::DeviceArena tmp;
    try {
      const int32_t* d_rows = tmp.upload(h_rows, static_cast<size_t>(m), s);
      const double* d_block = tmp.upload(h_block, static_cast<size_t>(m), s);
      hip_check(dfa::launch_block_add(m, d_rows, d_block, d_b, s), "block_add_kernel");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    } catch (...) {
      tmp.release();
      throw;
    }
    tmp.release();
    return DFA_OK;
  });
}

int dfa_csr_diagonal(dfa_assembler h, const double* d_values, double* d_diag, void* stream) {
  return guarded([&]() -> int {
    if (!h || !d_values || !d_diag) return fail("NULL argument", DFA_ERR_ARG);
    dfa::CsrView A;
    A.nrows = h->ndofs;
    A.nnz = h->nnz;
    A.row_ptr = h->d_row_ptr;
    A.col = h->d_col;
    A.values = d_values;
    hip_check(dfa::launch_diagonal(A, d_diag, static_cast<hipStream_t>(stream)), "diagonal_kernel");
    return DFA_OK;
  });
}

int dfa_plan_info(dfa_assembler h, int32_t scatter, int64_t* plan_bytes, int64_t* num_blocks,
                  int64_t* cell_appearances, int32_t* lds_bytes) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    const int sc = resolve_scatter(h, scatter);
    if (sc == DFA_SCATTER_BLOCKED && h->element == DFA_ELEMENT_VECTOR_P2) {
      if (plan_bytes) *plan_bytes = h->plan.bytes;
      if (num_blocks) *num_blocks = h->vtopo.num_blocks;
      if (cell_appearances) *cell_appearances = h->appearances;
      if (lds_bytes) *lds_bytes = static_cast<int32_t>(dfa::vp2_blocked_lds_bytes(h->vtopo, true));
    } else if (sc == DFA_SCATTER_BLOCKED) {
      dfa::P1Pass p;
      p.gdim = h->gdim;
      p.a_form = DFA_FORM_P1_POISSON;
      p.want_b = true;
      if (plan_bytes) *plan_bytes = h->plan.bytes;
      if (num_blocks) *num_blocks = h->btopo.num_blocks;
      if (cell_appearances) *cell_appearances = h->appearances;
      if (lds_bytes) *lds_bytes = static_cast<int32_t>(dfa::p1_blocked_lds_bytes(h->btopo, p));
    } else if (sc == DFA_SCATTER_COLORED) {
      ensure_colors(h, nullptr);
      if (plan_bytes) *plan_bytes = h->colors.bytes;
      if (num_blocks) *num_blocks = static_cast<int64_t>(h->color_ptr.size()) - 1;
      if (cell_appearances) *cell_appearances = h->nc;
      if (lds_bytes) *lds_bytes = 0;
    } else {
      if (plan_bytes) *plan_bytes = h->element == DFA_ELEMENT_VECTOR_P2 ? h->nc * 100 : 0;
      if (num_blocks) *num_blocks = 0;
      if (cell_appearances) *cell_appearances = h->nc;
      if (lds_bytes) *lds_bytes = 0;
    }
    return DFA_OK;
  });
}

int dfa_device_info(int32_t device, int64_t* info) {
  return guarded([&]() -> int {
    if (!info) return fail("info is NULL", DFA_ERR_ARG);
    const hipDeviceAttribute_t attrs[5] = {
        hipDeviceAttributeMultiprocessorCount, hipDeviceAttributeClockRate,
        hipDeviceAttributeMemoryClockRate, hipDeviceAttributeMemoryBusWidth,
        hipDeviceAttributeL2CacheSize};
    for (int i = 0; i < 5; ++i) {
      int v = 0;
      hip_check(hipDeviceGetAttribute(&v, attrs[i], device), "hipDeviceGetAttribute");
      info[i] = v;
    }
    size_t mem = 0;
    hip_check(hipDeviceTotalMem(&mem, device), "hipDeviceTotalMem");
    info[5] = static_cast<int64_t>(mem);
    return DFA_OK;
  });
}

int dfa_mesh_coloring(int64_t num_cells, int32_t nvc, const int32_t* h_cells, int64_t num_vertices,
                      int32_t entity_dim, int32_t* h_colors, int32_t* num_colors) {
  return guarded([&]() -> int {
    if (num_cells < 0 || num_vertices < 0 || nvc < 1 || entity_dim < 0 || entity_dim >= nvc)
      return fail("invalid mesh colouring arguments", DFA_ERR_ARG);
    if (num_cells > 0 && (!h_cells || !h_colors)) return fail("NULL array", DFA_ERR_ARG);
    const int32_t n = dfa::greedy_cell_coloring(num_cells, nvc, h_cells, num_vertices,
                                                entity_dim + 1, h_colors);
    if (num_colors) *num_colors = n;
    return DFA_OK;
  });
}

int dfa_color_info(dfa_assembler h, int32_t* num_colors, int64_t* h_color_ptr, void* stream) {
  return guarded([&]() -> int {
    if (!h) return fail("NULL handle", DFA_ERR_ARG);
    if (h->element != DFA_ELEMENT_P1)
      return fail("the coloured scatter is implemented for P1 elements", DFA_ERR_ARG);
    ensure_colors(h, static_cast<hipStream_t>(stream));
    if (num_colors) *num_colors = static_cast<int32_t>(h->color_ptr.size()) - 1;
    if (h_color_ptr) std::copy(h->color_ptr.begin(), h->color_ptr.end(), h_color_ptr);
    return DFA_OK;
  });
}

}  // extern "C"
