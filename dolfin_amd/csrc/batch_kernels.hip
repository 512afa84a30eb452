// Batched assembly: many (mesh, form) pairs in one captured hipGraph of two kernels, one for
// the vector-P2 members and one for the P1 members, each compiled and launched with the
// register and LDS budget of its own block body (a single kernel for both would run both at
// the larger body's occupancy).
//
// Work units are the row blocks of every member's blocked plan (P1 row blocks, vector-P2 node
// blocks), listed once per kernel. By default every unit is one workgroup; with a positive
// grid_per_cu a grid of persistent workgroups walks the table with a fixed stride (no queue
// counter; every workgroup reaches the end of the table). Each unit runs the same block body as the standalone kernels; values land directly in
// each member's CSR storage. The reference assembles one form per Assembler::assemble call
// and has no counterpart.
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

constexpr int kBatchP1Threads = 512;  // the standalone edge-chain configuration
constexpr int kBatchP2Threads = p2dev::kBT;

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

// The member table is read-only for the whole launch: read through the constant address
// space, its (uniform) fields load into scalar registers even though the block bodies store
// to global memory in between.
// (The host compilation pass only parses the kernels; it sees plain const types.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) BatchItem ConstItem;
typedef const __attribute__((address_space(4))) BatchUnit ConstUnit;
#else
typedef const BatchItem ConstItem;
typedef const BatchUnit ConstUnit;
#endif

// The unit a workgroup takes: with one workgroup per unit (PERSIST false) in XCD-aware order
// (workgroup i runs on XCD i % 8; each XCD takes a contiguous range of units, i.e. adjacent
// blocks of one member, whose ghost nodes then meet in that XCD's L2); persistent workgroups
// stride through the table.
template <bool PERSIST>
__device__ __forceinline__ int first_unit(int nunits) {
  if constexpr (PERSIST) return static_cast<int>(blockIdx.x);
  const int q = nunits >> 3, r = nunits & 7, xcd = blockIdx.x & 7;
  return xcd * q + min(xcd, r) + static_cast<int>(blockIdx.x >> 3);
}

template <bool PERSIST>
__global__ __launch_bounds__(kBatchP1Threads) __attribute__((amdgpu_waves_per_eu(6)))
void batch_p1_kernel(const BatchItem* __restrict__ items, const BatchUnit* __restrict__ units,
                     int nunits) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (int u = first_unit<PERSIST>(nunits); u < nunits; u += gridDim.x) {
    const BatchUnit un = ((ConstUnit*)units)[u];
    ConstItem& it = ((ConstItem*)items)[un.item];
    const P1BlockTopo t = it.p1;
    const P1Pass p = it.p1p;
    const p1dev::LdsLayout L = it.p1L;
    if (it.kind == 3)
      p1dev::p1_chain_body<false, kBatchP1Threads, 1>(t, p, L, lds, un.block);
    else
      p1dev::p1_block_body<3, DFA_FORM_P1_POISSON, false, false, true, kBatchP1Threads, true>(
          t, p, L, lds, un.block);
    if constexpr (!PERSIST) break;
    __syncthreads();  // the next unit overwrites the LDS this unit's write-out reads
  }
}

template <bool PERSIST>
__global__ __launch_bounds__(kBatchP2Threads) __attribute__((amdgpu_waves_per_eu(4)))
void batch_p2_kernel(const BatchItem* __restrict__ items, const BatchUnit* __restrict__ units,
                     int nunits) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (int u = first_unit<PERSIST>(nunits); u < nunits; u += gridDim.x) {
    const BatchUnit un = ((ConstUnit*)units)[u];
    ConstItem& it = ((ConstItem*)items)[un.item];
    const VP2BlockTopo t = it.p2;
    const VP2Pass p = it.p2p;
    const p2dev::VP2Lds L = it.p2L;
    p2dev::vp2_block_body<true, false, false>(t, p, L, lds, un.block);
    if constexpr (!PERSIST) break;
    __syncthreads();
  }
}

}  // namespace
}  // namespace dfa

struct dfa_batch_s {
  int32_t n = 0;
  int64_t nunits = 0, cells = 0;
  int32_t grid = 0, lds = 0;
  void* d_items = nullptr;
  void* d_units = nullptr;  // vector-P2 units, then P1 units
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
    std::vector<BatchUnit> units1, units2;  // P1, vector-P2
    int lds1 = 0, lds2 = 0;
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
        it.kind = (it.p1.chain_ptr && it.p1.chain_rec && it.p1.blk_hdr &&
                   it.p1.max_chains <= kBatchP1Threads && it.p1.max_nodes <= kMaxChainBlockNodes)
                      ? 3
                      : 1;
        it.p1p.gdim = 3;
        it.p1p.a_form = DFA_FORM_P1_POISSON;
        it.p1p.want_A = true;
        it.p1p.x = d_x[i];
        it.p1p.kappa = d_c0[i];
        it.p1p.values = d_values[i];
        it.p1L = p1dev::lds_layout(3, it.p1.max_nodes, it.p1.max_rows, it.p1.max_nnz, false, false,
                                   true, true);
        lds1 = std::max(lds1, it.p1L.total);
        for (int b = 0; b < it.p1.num_blocks; ++b) units1.push_back(BatchUnit{i, b});
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
        lds2 = std::max(lds2, it.p2L.total);
        for (int b = 0; b < it.p2.num_blocks; ++b) units2.push_back(BatchUnit{i, b});
      } else {
        return bfail("batch supports the P1 Poisson and vector-P2 elasticity forms", DFA_ERR_ARG);
      }
    }
    if (std::max(lds1, lds2) > 160 * 1024) return bfail("batch needs more LDS than a CU has", DFA_ERR_LIMIT);
    std::vector<BatchUnit> flat(units2);
    flat.insert(flat.end(), units1.begin(), units1.end());
    const int64_t n2 = static_cast<int64_t>(units2.size()), n1 = static_cast<int64_t>(units1.size());
    if (n1 > INT32_MAX || n2 > INT32_MAX) return bfail("too many batch units", DFA_ERR_LIMIT);

    auto* B = new dfa_batch_s();
    try {
      B->n = n;
      B->nunits = static_cast<int64_t>(flat.size());
      B->cells = cells;
      B->lds = std::max(lds1, lds2);
      hipStream_t s = static_cast<hipStream_t>(stream);
      hip_check(hipMalloc(&B->d_items, sizeof(BatchItem) * items.size()), "hipMalloc");
      hip_check(hipMalloc(&B->d_units, sizeof(BatchUnit) * std::max<size_t>(flat.size(), 1)), "hipMalloc");
      hip_check(hipMemcpyAsync(B->d_items, items.data(), sizeof(BatchItem) * items.size(),
                               hipMemcpyHostToDevice, s), "hipMemcpyAsync");
      if (!flat.empty())
        hip_check(hipMemcpyAsync(B->d_units, flat.data(), sizeof(BatchUnit) * flat.size(),
                                 hipMemcpyHostToDevice, s), "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      int dev = 0;
      hipDeviceProp_t prop;
      hip_check(hipGetDevice(&dev), "hipGetDevice");
      hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
      // grid_per_cu > 0: persistent workgroups, that many per CU (at most what fits);
      // 0: one workgroup per unit
      const bool persist = grid_per_cu > 0;
      auto grid_of = [&](const void* kern, int threads, int lds, int64_t nunits) {
        hip_check(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds),
                  "hipFuncSetAttribute");
        if (!persist) return static_cast<int32_t>(std::max<int64_t>(nunits, 1));
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds),
                  "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu = std::max(std::min(per_cu, grid_per_cu), 1);
        return static_cast<int32_t>(std::min<int64_t>(static_cast<int64_t>(per_cu) * prop.multiProcessorCount,
                                                      std::max<int64_t>(nunits, 1)));
      };
      const void* k2 = persist ? reinterpret_cast<const void*>(batch_p2_kernel<true>)
                               : reinterpret_cast<const void*>(batch_p2_kernel<false>);
      const void* k1 = persist ? reinterpret_cast<const void*>(batch_p1_kernel<true>)
                               : reinterpret_cast<const void*>(batch_p1_kernel<false>);
      const int32_t g2 = n2 > 0 ? grid_of(k2, kBatchP2Threads, lds2, n2) : 0;
      const int32_t g1 = n1 > 0 ? grid_of(k1, kBatchP1Threads, lds1, n1) : 0;
      B->grid = g1 + g2;
      const BatchItem* di = static_cast<const BatchItem*>(B->d_items);
      const BatchUnit* du = static_cast<const BatchUnit*>(B->d_units);
      // capture the launches once; replays reuse them
      hip_check(hipStreamCreateWithFlags(&B->capture, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipStreamBeginCapture(B->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      if (n2 > 0) {
        if (persist)
          hipLaunchKernelGGL(batch_p2_kernel<true>, dim3(g2), dim3(kBatchP2Threads), lds2, B->capture,
                             di, du, static_cast<int>(n2));
        else
          hipLaunchKernelGGL(batch_p2_kernel<false>, dim3(g2), dim3(kBatchP2Threads), lds2, B->capture,
                             di, du, static_cast<int>(n2));
      }
      hipError_t le = hipGetLastError();
      if (le == hipSuccess && n1 > 0) {
        if (persist)
          hipLaunchKernelGGL(batch_p1_kernel<true>, dim3(g1), dim3(kBatchP1Threads), lds1, B->capture,
                             di, du + n2, static_cast<int>(n1));
        else
          hipLaunchKernelGGL(batch_p1_kernel<false>, dim3(g1), dim3(kBatchP1Threads), lds1, B->capture,
                             di, du + n2, static_cast<int>(n1));
        le = hipGetLastError();
      }
      hipError_t ce = hipStreamEndCapture(B->capture, &B->graph);
      hip_check(le, "batch kernel launch");
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
