// Cell colouring for the graph-coloured scatter, on the device: cells that share a dof get
// different colours, so the cells of one colour can add into their CSR rows without atomics
// (the role MeshColoring::color plays for DOLFIN's race-free assembly, MeshColoring.cpp:57-121).
//
// Jones-Plassmann: every cell gets a fixed pseudo-random priority (a hash of its index, ties
// by index); in each round, every uncoloured cell whose priority exceeds that of all its
// uncoloured neighbours takes the smallest colour none of its coloured neighbours holds. Two
// such cells are never neighbours, so each round colours an independent set; rounds read the
// previous round's colours only, so the result depends only on the mesh and dofmap. The colours are not those of DOLFIN's
// sequential greedy colouring in index order (that one is MeshColoring / dfa_mesh_coloring,
// on the host); the scatter needs any colouring with this property.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "assembler.h"

namespace dfa {
namespace {

constexpr int kCT = 256;
constexpr int kMaxColors = 128;  // two 64-bit masks per cell

unsigned cgrid(int64_t n) {
  const int64_t b = (n + kCT - 1) / kCT;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 65535)));
}

#define CGRID_LOOP(i, n)                                                                    \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

int cbits(int64_t v) {  // bits to hold values 0..v
  int b = 1;
  while (b < 63 && (int64_t{1} << b) <= v) ++b;
  return b;
}

__device__ __forceinline__ uint32_t prio_hash(uint32_t x) {  // a 32-bit integer mix (splitmix-style)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint64_t prio(int64_t c) {
  return (static_cast<uint64_t>(prio_hash(static_cast<uint32_t>(c))) << 32) | static_cast<uint64_t>(c);
}

// (dof, cell) keys of every cell entry
__global__ void col_keys(int64_t n, int32_t nvc, const int32_t* __restrict__ cd, int cb,
                         uint64_t* __restrict__ keys) {
  CGRID_LOOP(i, n) {
    const int64_t c = i / nvc;
    keys[i] = (static_cast<uint64_t>(cd[i]) << cb) | static_cast<uint64_t>(c);
  }
}

// dof -> first position among the sorted keys (dptr[ndofs] = n), and the cell of each key
__global__ void col_ptr(int64_t n, int64_t ndofs, int cb, const uint64_t* __restrict__ keys,
                        int64_t* __restrict__ dptr, int32_t* __restrict__ cells) {
  const uint64_t mask = (uint64_t{1} << cb) - 1;
  CGRID_LOOP(k, n + 1) {
    const int64_t prev = k == 0 ? -1 : static_cast<int64_t>(keys[k - 1] >> cb);
    const int64_t cur = k == n ? ndofs : static_cast<int64_t>(keys[k] >> cb);
    for (int64_t d = prev + 1; d <= cur; ++d) dptr[d] = k;
    if (k < n) cells[k] = static_cast<int32_t>(keys[k] & mask);
  }
}

// One Jones-Plassmann round from the colours of the previous round (cur) into next, so the
// result does not depend on the order in which threads run; flags[0] = cells left
// uncoloured, flags[1] = a cell found no colour below kMaxColors.
__global__ void col_round(int64_t nc, int32_t nvc, const int32_t* __restrict__ cd,
                          const int64_t* __restrict__ dptr, const int32_t* __restrict__ dcells,
                          const int32_t* __restrict__ cur, int32_t* __restrict__ next,
                          int32_t* __restrict__ flags) {
  CGRID_LOOP(c, nc) {
    if (cur[c] >= 0) {
      next[c] = cur[c];
      continue;
    }
    const uint64_t pc = prio(c);
    uint64_t m0 = 0, m1 = 0;
    bool top = true;
    for (int a = 0; a < nvc && top; ++a) {
      const int32_t d = cd[c * nvc + a];
      for (int64_t k = dptr[d]; k < dptr[d + 1]; ++k) {
        const int32_t n = dcells[k];
        if (n == c) continue;
        const int32_t cn = cur[n];
        if (cn < 0) {
          if (prio(n) > pc) {
            top = false;
            break;
          }
        } else if (cn < 64) {
          m0 |= uint64_t{1} << cn;
        } else {
          m1 |= uint64_t{1} << (cn - 64);
        }
      }
    }
    int k = -1;
    if (top) {
      if (~m0) k = __builtin_ctzll(~m0);
      else if (~m1) k = 64 + __builtin_ctzll(~m1);
      else atomicOr(&flags[1], 1);
    }
    if (k < 0) atomicAdd(&flags[0], 1);
    next[c] = k;
  }
}

__global__ void col_sort_keys(int64_t nc, int cb, const int32_t* __restrict__ color, uint64_t* __restrict__ keys) {
  CGRID_LOOP(c, nc) keys[c] = (static_cast<uint64_t>(color[c]) << cb) | static_cast<uint64_t>(c);
}

__global__ void col_perm(int64_t nc, int cb, const uint64_t* __restrict__ keys, int32_t* __restrict__ perm,
                         int64_t* __restrict__ hist) {
  const uint64_t mask = (uint64_t{1} << cb) - 1;
  CGRID_LOOP(i, nc) {
    perm[i] = static_cast<int32_t>(keys[i] & mask);
    atomicAdd(reinterpret_cast<unsigned long long*>(&hist[keys[i] >> cb]), 1ull);
  }
}

struct CBuf {
  void* p = nullptr;
  ~CBuf() {
    if (p) (void)hipFree(p);
  }
};

// radix sort of n keys (bits significant bits) in a / b; returns the sorted buffer
uint64_t* sort_keys(uint64_t* a, uint64_t* b, int64_t n, int bits, hipStream_t s, CBuf& tmp) {
  rocprim::double_buffer<uint64_t> keys(a, b);
  size_t need = 0;
  hip_check(rocprim::radix_sort_keys(nullptr, need, keys, static_cast<size_t>(n), 0, bits, s), "radix_sort_keys (size)");
  if (tmp.p) {
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    (void)hipFree(tmp.p);
    tmp.p = nullptr;
  }
  hip_check(hipMalloc(&tmp.p, std::max<size_t>(need, 16)), "hipMalloc");
  hip_check(rocprim::radix_sort_keys(tmp.p, need, keys, static_cast<size_t>(n), 0, bits, s), "radix_sort_keys");
  return keys.current();
}

}  // namespace

// Colours the cells of (nc, nvc, d_cdofs) on the device and puts them in colour order: perm
// (host, nc: the cells of colour 0 in ascending index, then colour 1, ...) and color_ptr
// (host, colours + 1 offsets). False when more than kMaxColors colours would be needed (the
// caller then colours on the host).
bool color_cells_device(int64_t nc, int32_t nvc, const int32_t* d_cdofs, int64_t ndofs, std::vector<int32_t>& perm,
                        std::vector<int64_t>& color_ptr, hipStream_t s) {
  if (nc <= 0 || nc >= INT32_MAX || ndofs <= 0) return false;
  const int64_t n = nc * nvc;
  const int cbc = cbits(nc), cbd = cbits(ndofs);
  if (cbc + cbd > 64) return false;
  CBuf ka, kb, tmp, dptr, dcells, color, color2, flags, permd, hist;
  hip_check(hipMalloc(&ka.p, static_cast<size_t>(n) * sizeof(uint64_t)), "hipMalloc");
  hip_check(hipMalloc(&kb.p, static_cast<size_t>(n) * sizeof(uint64_t)), "hipMalloc");
  hip_check(hipMalloc(&dptr.p, static_cast<size_t>(ndofs + 1) * sizeof(int64_t)), "hipMalloc");
  hip_check(hipMalloc(&dcells.p, static_cast<size_t>(n) * sizeof(int32_t)), "hipMalloc");
  hip_check(hipMalloc(&color.p, static_cast<size_t>(nc) * sizeof(int32_t)), "hipMalloc");
  hip_check(hipMalloc(&color2.p, static_cast<size_t>(nc) * sizeof(int32_t)), "hipMalloc");
  hip_check(hipMalloc(&flags.p, 2 * sizeof(int32_t)), "hipMalloc");
  auto* a = static_cast<uint64_t*>(ka.p);
  auto* b = static_cast<uint64_t*>(kb.p);
  // dof -> cells
  hipLaunchKernelGGL(col_keys, dim3(cgrid(n)), dim3(kCT), 0, s, n, nvc, d_cdofs, cbc, a);
  hip_check(hipGetLastError(), "col_keys");
  const uint64_t* sorted = sort_keys(a, b, n, cbc + cbd, s, tmp);
  hipLaunchKernelGGL(col_ptr, dim3(cgrid(n + 1)), dim3(kCT), 0, s, n, ndofs, cbc, sorted,
                     static_cast<int64_t*>(dptr.p), static_cast<int32_t*>(dcells.p));
  hip_check(hipGetLastError(), "col_ptr");
  // Jones-Plassmann rounds
  const std::vector<int32_t> minus(static_cast<size_t>(nc), -1);
  hip_check(hipMemcpyAsync(color.p, minus.data(), minus.size() * sizeof(int32_t), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
  for (;;) {
    int32_t fl[2] = {0, 0};
    hip_check(hipMemcpyAsync(flags.p, fl, sizeof(fl), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hipLaunchKernelGGL(col_round, dim3(cgrid(nc)), dim3(kCT), 0, s, nc, nvc, d_cdofs,
                       static_cast<const int64_t*>(dptr.p), static_cast<const int32_t*>(dcells.p),
                       static_cast<const int32_t*>(color.p), static_cast<int32_t*>(color2.p),
                       static_cast<int32_t*>(flags.p));
    hip_check(hipGetLastError(), "col_round");
    std::swap(color.p, color2.p);
    hip_check(hipMemcpyAsync(fl, flags.p, sizeof(fl), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (fl[1]) return false;
    if (fl[0] == 0) break;
  }
  // colour order: sort (colour, cell)
  const int cbk = cbits(kMaxColors);
  hipLaunchKernelGGL(col_sort_keys, dim3(cgrid(nc)), dim3(kCT), 0, s, nc, cbc, static_cast<const int32_t*>(color.p), a);
  const uint64_t* ck = sort_keys(a, b, nc, cbc + cbk, s, tmp);
  hip_check(hipMalloc(&permd.p, static_cast<size_t>(nc) * sizeof(int32_t)), "hipMalloc");
  hip_check(hipMalloc(&hist.p, (kMaxColors + 1) * sizeof(int64_t)), "hipMalloc");
  hip_check(hipMemsetAsync(hist.p, 0, (kMaxColors + 1) * sizeof(int64_t), s), "hipMemsetAsync");
  hipLaunchKernelGGL(col_perm, dim3(cgrid(nc)), dim3(kCT), 0, s, nc, cbc, ck, static_cast<int32_t*>(permd.p),
                     static_cast<int64_t*>(hist.p));
  hip_check(hipGetLastError(), "col_perm");
  std::vector<int64_t> h(kMaxColors + 1);
  perm.resize(static_cast<size_t>(nc));
  hip_check(hipMemcpyAsync(h.data(), hist.p, h.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(perm.data(), permd.p, perm.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  int ncol = 0;
  for (int k = 0; k < kMaxColors; ++k)
    if (h[k]) ncol = k + 1;
  color_ptr.assign(static_cast<size_t>(ncol) + 1, 0);
  for (int k = 0; k < ncol; ++k) color_ptr[k + 1] = color_ptr[k] + h[k];
  return true;
}

}  // namespace dfa
