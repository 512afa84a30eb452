// Batched assembly: many (mesh, form) pairs in ONE persistent kernel, captured in a hipGraph.
//
// Work units are the row blocks of every member's blocked plan (P1 row blocks, vector-P2 node
// blocks), listed once in a unit table sorted by expected cost (vector-P2 blocks first). A
// grid of resident workgroups walks the table with a fixed stride, so the launch needs no
// queue counter and every workgroup reaches the end of the table. Each unit runs the same
// block body as the standalone kernels; values land directly in each member's CSR storage.
// The reference assembles one form per Assembler::assemble call and has no counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "assembler.h"
#include "common.h"
#include "kernels.h"
#include "p1_device.h"
#include "p2_device.h"

namespace dfa {

namespace {

constexpr int kBatchThreads = 256;
constexpr int kBatchChainsPerThread = 2;  // edge-chain units: blocks of up to 512 chains

struct BatchItem {
  int32_t kind;  // 1: P1 kappa-Poisson on tetrahedra, 2: vector-P2 elasticity, 3: as 1 by edge chains
  P1BlockTopo p1;
  P1Pass p1p;
  p1dev::LdsLayout p1L;
  VP2BlockTopo p2;
  VP2Pass p2p;
  p2dev::VP2Lds p2L;
};

struct BatchUnit {
  int32_t item;
  int32_t block;
};

__global__ __launch_bounds__(kBatchThreads) void batch_kernel(const BatchItem* __restrict__ items,
                                                              const BatchUnit* __restrict__ units,
                                                              int nunits) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const BatchUnit un = units[u];
    const BatchItem& it = items[un.item];
    if (it.kind == 3)
      p1dev::p1_chain_body<false, kBatchThreads, kBatchChainsPerThread>(it.p1, it.p1p, it.p1L, lds,
                                                                        un.block);
    else if (it.kind == 1)
      p1dev::p1_block_body<3, DFA_FORM_P1_POISSON, false, false, true, kBatchThreads, true>(
          it.p1, it.p1p, it.p1L, lds, un.block);
    else
      p2dev::vp2_block_body<true, false, false>(it.p2, it.p2p, it.p2L, lds, un.block);
    __syncthreads();  // the next unit overwrites the LDS this unit's write-out reads
  }
}

}  // namespace
}  // namespace dfa

struct dfa_batch_s {
  int32_t n = 0;
  int64_t nunits = 0, cells = 0;
  int32_t grid = 0, lds = 0;
  void* d_items = nullptr;
  void* d_units = nullptr;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int bfail(const std::string& m, int code) {
  dfa::set_error(m);
  return code;
}

void batch_release(dfa_batch_s* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_units) (void)hipFree(b->d_units);
}

}  // namespace

extern "C" {

int dfa_batch_create(dfa_batch* out, int32_t n, const dfa_assembler* hs, const int32_t* forms,
                     const double* const* d_x, const double* const* d_c0,
                     const double* const* d_c1, double* const* d_values, int32_t grid_per_cu,
                     void* stream) {
  using namespace dfa;
  try {
    set_error("");
    if (!out || n <= 0 || !hs || !forms || !d_x || !d_c0 || !d_values)
      return bfail("invalid batch arguments", DFA_ERR_ARG);
    *out = nullptr;
    std::vector<BatchItem> items(n);
    std::vector<std::pair<int64_t, BatchUnit>> units;
    int lds = 0;
    int64_t cells = 0;
    for (int i = 0; i < n; ++i) {
      const dfa_assembler h = hs[i];
      BatchItem& it = items[i];
      it = BatchItem();
      if (!h || !d_x[i] || !d_c0[i] || !d_values[i]) return bfail("NULL member in batch", DFA_ERR_ARG);
      cells += h->nc;
      if (forms[i] == DFA_FORM_P1_POISSON) {
        if (h->element != DFA_ELEMENT_P1 || h->gdim != 3 || !h->has_plan ||
            h->btopo.max_row_len > kMaxRank || !h->rows_covered)
          return bfail("batch member " + std::to_string(i) +
                           ": P1 members must be tetrahedral meshes with a blocked plan",
                       DFA_ERR_ARG);
        it.p1 = h->btopo;
        it.kind = (it.p1.chain_ptr && it.p1.chain_rec &&
                   it.p1.max_chains <= kBatchThreads * kBatchChainsPerThread) ? 3 : 1;
        it.p1p.gdim = 3;
        it.p1p.a_form = DFA_FORM_P1_POISSON;
        it.p1p.want_A = true;
        it.p1p.x = d_x[i];
        it.p1p.kappa = d_c0[i];
        it.p1p.values = d_values[i];
        it.p1L = p1dev::lds_layout(3, it.p1.max_nodes, it.p1.max_rows, it.p1.max_nnz, false, false,
                                   true, true);
        lds = std::max(lds, it.p1L.total);
        for (int b = 0; b < it.p1.num_blocks; ++b) units.push_back({1, BatchUnit{i, b}});
      } else if (forms[i] == DFA_FORM_VP2_ELASTICITY) {
        if (h->element != DFA_ELEMENT_VECTOR_P2 || !h->has_vp2_plan || !h->rows_covered)
          return bfail("batch member " + std::to_string(i) +
                           ": vector P2 members need a blocked plan",
                       DFA_ERR_ARG);
        if (!d_c1 || !d_c1[i]) return bfail("lmbda is NULL", DFA_ERR_ARG);
        it.kind = 2;
        it.p2 = h->vtopo;
        it.p2p.want_A = true;
        it.p2p.x = d_x[i];
        it.p2p.mu = d_c0[i];
        it.p2p.lmbda = d_c1[i];
        it.p2p.values = d_values[i];
        it.p2L = p2dev::vp2_lds(it.p2.max_verts, it.p2.max_owned, it.p2.max_vals, false);
        lds = std::max(lds, it.p2L.total);
        for (int b = 0; b < it.p2.num_blocks; ++b) units.push_back({8, BatchUnit{i, b}});
      } else {
        return bfail("batch supports the P1 Poisson and vector-P2 elasticity forms", DFA_ERR_ARG);
      }
    }
    if (lds > 160 * 1024) return bfail("batch needs more LDS than a CU has", DFA_ERR_LIMIT);
    // expensive units first (longest-processing-time order), stable within a cost class
    std::stable_sort(units.begin(), units.end(),
                     [](const std::pair<int64_t, BatchUnit>& a, const std::pair<int64_t, BatchUnit>& b) {
                       return a.first > b.first;
                     });
    std::vector<BatchUnit> flat(units.size());
    for (size_t k = 0; k < units.size(); ++k) flat[k] = units[k].second;

    auto* B = new dfa_batch_s();
    try {
      B->n = n;
      B->nunits = static_cast<int64_t>(flat.size());
      B->cells = cells;
      B->lds = lds;
      hipStream_t s = static_cast<hipStream_t>(stream);
      hip_check(hipMalloc(&B->d_items, sizeof(BatchItem) * items.size()), "hipMalloc");
      hip_check(hipMalloc(&B->d_units, sizeof(BatchUnit) * std::max<size_t>(flat.size(), 1)), "hipMalloc");
      hip_check(hipMemcpyAsync(B->d_items, items.data(), sizeof(BatchItem) * items.size(),
                               hipMemcpyHostToDevice, s), "hipMemcpyAsync");
      if (!flat.empty())
        hip_check(hipMemcpyAsync(B->d_units, flat.data(), sizeof(BatchUnit) * flat.size(),
                                 hipMemcpyHostToDevice, s), "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds),
                "hipFuncSetAttribute");
      int dev = 0, per_cu = 0;
      hipDeviceProp_t prop;
      hip_check(hipGetDevice(&dev), "hipGetDevice");
      hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel, kBatchThreads, lds),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor");
      if (grid_per_cu > 0) per_cu = std::min(per_cu, grid_per_cu);
      per_cu = std::max(per_cu, 1);
      B->grid = static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(per_cu) * prop.multiProcessorCount,
                                                       std::max<int64_t>(B->nunits, 1)));
      // capture the launch once; replays reuse it
      hip_check(hipStreamCreateWithFlags(&B->capture, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipStreamBeginCapture(B->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      hipLaunchKernelGGL(batch_kernel, dim3(B->grid), dim3(kBatchThreads), lds, B->capture,
                         static_cast<const BatchItem*>(B->d_items), static_cast<const BatchUnit*>(B->d_units),
                         static_cast<int>(B->nunits));
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(B->capture, &B->graph);
      hip_check(le, "batch_kernel launch");
      hip_check(ce, "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    } catch (...) {
      batch_release(B);
      delete B;
      throw;
    }
    *out = B;
    return DFA_OK;
  } catch (const dfa::HipError& e) {
    return bfail(e.what(), DFA_ERR_HIP);
  } catch (const std::exception& e) {
    return bfail(e.what(), DFA_ERR_ARG);
  }
}

int dfa_batch_launch(dfa_batch b, void* stream) {
  if (!b) return bfail("NULL batch", DFA_ERR_ARG);
  const hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return bfail(std::string("hipGraphLaunch: ") + hipGetErrorString(e), DFA_ERR_HIP);
  return DFA_OK;
}

int dfa_batch_info(dfa_batch b, int64_t* units, int32_t* grid, int32_t* lds_bytes, int64_t* cells) {
  if (!b) return bfail("NULL batch", DFA_ERR_ARG);
  if (units) *units = b->nunits;
  if (grid) *grid = b->grid;
  if (lds_bytes) *lds_bytes = b->lds;
  if (cells) *cells = b->cells;
  return DFA_OK;
}

int dfa_batch_destroy(dfa_batch b) {
  if (!b) return DFA_OK;
  batch_release(b);
  delete b;
  return DFA_OK;
}

}  // extern "C"
