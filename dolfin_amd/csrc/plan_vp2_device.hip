// Device build of the vector-P2 blocked-scatter plan: the arrays build_vp2_block_plan and
// build_vp2_ranks (structure.cpp) produce, computed on the GPU from the device mesh, dofmaps
// and CSR structure, element for element equal to the host build.
//
//  0. rank table: per cell, the rank of node b among node a's neighbour nodes (binary search
//     in the component-0 columns of node a's CSR row);
//  1. node table: every P2 node (named by the dof of its component 0) with its position as
//     float (vertex, or edge midpoint), the node ids in ascending order;
//  2. recursive coordinate bisection, one level at a time: per segment the summed LDS values
//     (3 per CSR entry of the node's row) and the bounding box are reduced on the device;
//     the host decides each segment as the recursion does (leaf when its values fit the
//     block, else split at n * (leaves / 2) / leaves along its longest extent), and one
//     stable radix sort by (coordinate on the segment's axis, node) and one by segment put
//     every segment in the order the split needs. The split takes the same node sets as the
//     recursion's nth_element (the order is total), and leaves in position order are its
//     depth-first leaves;
//  3. per block: owned nodes ascending, the work items (cell, owned node) in cell order, the
//     cells' vertices sorted and unique, the LDS offset of every owned node's rows, and the
//     items interleaved over the owned nodes (vertex nodes first) by one sort of
//     (block, kind, round, owned node);
//  4. task words: local vertex ids (binary search in the block's vertices) and the ranks.
// Block boundaries come from binary searches in the sorted keys, not from counters.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "assembler.h"

namespace dfa {
namespace {

constexpr int kVT = 256;
constexpr int kRC = 8;  // positions per thread in the segment reduction

unsigned vgrid(int64_t n) {
  const int64_t b = (n + kVT - 1) / kVT;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 65535)));
}

#define VGRID_LOOP(i, n)                                                                    \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

int vbits(int64_t v) {  // bits to hold values 0..v
  int b = 1;
  while (b < 63 && (int64_t{1} << b) <= v) ++b;
  return b;
}

// order-preserving map of a float onto uint32 (+0 and -0 equal, as the float comparison has them)
__device__ __forceinline__ uint32_t fkey(float v) {
  if (v == 0.0f) v = 0.0f;
  const uint32_t u = __float_as_uint(v);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}

float unkey(uint32_t k) {
  const uint32_t u = (k >> 31) ? (k & 0x7fffffffu) : ~k;
  float f;
  std::memcpy(&f, &u, sizeof f);
  return f;
}

// cell-local edges of the P2 edge nodes 4..9 (UFC ordering: edge e joins the two vertices
// other than those of the opposite edge)
__constant__ int kVEdge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

enum { G_COEFF = 0, G_RANK, G_NFLAGS };

__global__ __launch_bounds__(kVT) void vp_rank_table(int64_t nc, const int32_t* __restrict__ cd,
                                                     const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col, bool node_major,
                                                     uint8_t* __restrict__ rank, int32_t* __restrict__ bad) {
  VGRID_LOOP(k, nc * 10) {
    const int64_t c = k / 10;
    const int a = static_cast<int>(k - c * 10);
    const int32_t na = cd[c * 30 + a];
    const int32_t r0 = rp[na], deg = (rp[na + 1] - r0) / 3;
    const int st = node_major ? 3 : 1;  // column (component 0, q-th neighbour) at r0 + st * q
    int fail = 0;
    for (int b = 0; b < 10; ++b) {
      const int32_t nb = cd[c * 30 + b];
      int32_t lo = 0, hi = deg;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (col[r0 + st * mid] < nb) lo = mid + 1;
        else hi = mid;
      }
      if (lo >= deg || col[r0 + st * lo] != nb || lo > 255) fail = 1;
      rank[c * 100 + a * 10 + b] = static_cast<uint8_t>(lo & 0xff);
    }
    if (fail) atomicOr(bad, 1);
  }
}

__global__ __launch_bounds__(kVT) void vp_node_table(int64_t nc, const int32_t* __restrict__ cells,
                                                     const int32_t* __restrict__ cd,
                                                     const double* __restrict__ x, int32_t* __restrict__ isnode,
                                                     float* __restrict__ nx) {
  VGRID_LOOP(k, nc * 10) {
    const int64_t c = k / 10;
    const int a = static_cast<int>(k - c * 10);
    const int32_t n = cd[c * 30 + a];
    isnode[n] = 1;
    for (int q = 0; q < 3; ++q) {
      double v;
      if (a < 4) {
        v = x[static_cast<int64_t>(cells[c * 4 + a]) * 3 + q];
      } else {
        v = 0.5 * (x[static_cast<int64_t>(cells[c * 4 + kVEdge[a - 4][0]]) * 3 + q] +
                   x[static_cast<int64_t>(cells[c * 4 + kVEdge[a - 4][1]]) * 3 + q]);
      }
      nx[3 * static_cast<int64_t>(n) + q] = static_cast<float>(v);
    }
  }
}

__global__ __launch_bounds__(kVT) void vp_compact(int64_t nd, const int32_t* __restrict__ flag,
                                                  const int32_t* __restrict__ pos, int32_t* __restrict__ ids) {
  VGRID_LOOP(d, nd) if (flag[d]) ids[pos[d]] = static_cast<int32_t>(d);
}

// segment of every position: the last segment start <= p (starts ascending, lo[0] = 0)
__global__ __launch_bounds__(kVT) void vp_segid(int64_t m, int32_t nseg, const int32_t* __restrict__ lo,
                                                int32_t* __restrict__ seg) {
  VGRID_LOOP(p, m) {
    int32_t a = 0, b = nseg - 1;
    while (a < b) {
      const int32_t mid = (a + b + 1) >> 1;
      if (lo[mid] <= p) a = mid;
      else b = mid - 1;
    }
    seg[p] = a;
  }
}

__device__ __forceinline__ void vp_flush(int32_t s, unsigned long long acc, const uint32_t* mn, const uint32_t* mx,
                                         unsigned long long* tot, uint32_t* bmin, uint32_t* bmax) {
  atomicAdd(&tot[s], acc);
  for (int q = 0; q < 3; ++q) {
    atomicMin(&bmin[3 * s + q], mn[q]);
    atomicMax(&bmax[3 * s + q], mx[q]);
  }
}

// per segment: summed values (3 per entry of the node's CSR row) and bounding box (as
// orderable keys). Each thread folds kRC consecutive positions; a wavefront whose positions
// lie in one segment folds its lanes with shuffles and adds once.
__global__ __launch_bounds__(kVT) void vp_reduce(int64_t m, const int32_t* __restrict__ perm,
                                                 const int32_t* __restrict__ seg, const int32_t* __restrict__ rp,
                                                 const float* __restrict__ nx, unsigned long long* __restrict__ tot,
                                                 uint32_t* __restrict__ bmin, uint32_t* __restrict__ bmax) {
  const int64_t nthr = (m + kRC - 1) / kRC;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t wb = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u); wb < nthr;
       wb += stride) {  // wb is the same on every lane: the loop is uniform per wavefront
    const int64_t t = wb + lane;
    const int64_t wp0 = wb * kRC, wp1 = std::min<int64_t>(m, (wb + 64) * kRC);
    const bool uniform = seg[wp0] == seg[wp1 - 1];
    int32_t cs = -1;
    unsigned long long acc = 0;
    uint32_t mn[3] = {~0u, ~0u, ~0u}, mx[3] = {0u, 0u, 0u};
    if (t < nthr) {
      const int64_t p0 = t * kRC, p1 = std::min<int64_t>(m, p0 + kRC);
      for (int64_t p = p0; p < p1; ++p) {
        const int32_t s = seg[p];
        if (s != cs) {
          if (cs >= 0) {
            vp_flush(cs, acc, mn, mx, tot, bmin, bmax);
            acc = 0;
            for (int q = 0; q < 3; ++q) {
              mn[q] = ~0u;
              mx[q] = 0u;
            }
          }
          cs = s;
        }
        const int32_t n = perm[p];
        acc += static_cast<unsigned long long>(3 * (rp[n + 1] - rp[n]));
        for (int q = 0; q < 3; ++q) {
          const uint32_t k = fkey(nx[3 * static_cast<int64_t>(n) + q]);
          mn[q] = min(mn[q], k);
          mx[q] = max(mx[q], k);
        }
      }
    }
    if (uniform) {
      for (int o = 32; o >= 1; o >>= 1) {
        acc += __shfl_xor(acc, o);
        for (int q = 0; q < 3; ++q) {
          mn[q] = min(mn[q], static_cast<uint32_t>(__shfl_xor(mn[q], o)));
          mx[q] = max(mx[q], static_cast<uint32_t>(__shfl_xor(mx[q], o)));
        }
      }
      if (lane == 0) vp_flush(seg[wp0], acc, mn, mx, tot, bmin, bmax);
    } else if (cs >= 0) {
      vp_flush(cs, acc, mn, mx, tot, bmin, bmax);
    }
  }
}

// keys for one level: splitting segments by (coordinate on their axis, node), others by node
__global__ __launch_bounds__(kVT) void vp_level_keys(int64_t m, const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ seg,
                                                     const int8_t* __restrict__ axis, const float* __restrict__ nx,
                                                     int nb, uint64_t* __restrict__ keys,
                                                     int32_t* __restrict__ vals) {
  VGRID_LOOP(p, m) {
    const int32_t n = perm[p];
    const int ax = axis[seg[p]];
    uint64_t k = static_cast<uint64_t>(n);
    if (ax >= 0) k |= static_cast<uint64_t>(fkey(nx[3 * static_cast<int64_t>(n) + ax])) << nb;
    keys[p] = k;
    vals[p] = static_cast<int32_t>(p);
  }
}

__global__ __launch_bounds__(kVT) void vp_seg_keys(int64_t m, const int32_t* __restrict__ order,
                                                   const int32_t* __restrict__ seg, uint64_t* __restrict__ keys) {
  VGRID_LOOP(i, m) keys[i] = static_cast<uint64_t>(seg[order[i]]);
}

__global__ __launch_bounds__(kVT) void vp_gather(int64_t m, const int32_t* __restrict__ order,
                                                 const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
  VGRID_LOOP(i, m) dst[i] = src[order[i]];
}

__global__ __launch_bounds__(kVT) void vp_leaf_keys(int64_t m, const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ seg, int nb,
                                                    uint64_t* __restrict__ keys) {
  VGRID_LOOP(p, m) keys[p] = (static_cast<uint64_t>(seg[p]) << nb) | static_cast<uint64_t>(perm[p]);
}

// owner block and index among the block's owned nodes; LDS values of each position
__global__ __launch_bounds__(kVT) void vp_owner(int64_t m, const uint64_t* __restrict__ keys, int nb,
                                                const int32_t* __restrict__ lo, const int32_t* __restrict__ rp,
                                                int32_t* __restrict__ perm, int32_t* __restrict__ owner,
                                                int32_t* __restrict__ own_index, int64_t* __restrict__ vals) {
  const uint64_t mask = (uint64_t{1} << nb) - 1;
  VGRID_LOOP(p, m) {
    const int32_t n = static_cast<int32_t>(keys[p] & mask);
    const int32_t b = static_cast<int32_t>(keys[p] >> nb);
    perm[p] = n;
    owner[n] = b;
    own_index[n] = static_cast<int32_t>(p - lo[b]);
    vals[p] = 3 * static_cast<int64_t>(rp[n + 1] - rp[n]);
  }
}

// own_info (LDS offset | neighbour nodes << 16), own_rows (the node's three dof rows), blk_vals
__global__ __launch_bounds__(kVT) void vp_own(int64_t m, const int32_t* __restrict__ perm,
                                              const int32_t* __restrict__ owner, const int32_t* __restrict__ lo,
                                              const int64_t* __restrict__ voff, const int64_t* __restrict__ vals,
                                              int64_t comp_stride, uint32_t* __restrict__ info,
                                              int32_t* __restrict__ rows, int32_t* __restrict__ blk_vals) {
  VGRID_LOOP(p, m) {
    const int32_t n = perm[p], b = owner[n];
    const int64_t off = voff[p] - voff[lo[b]];
    const int64_t len = vals[p] / 3;  // CSR row length
    info[p] = static_cast<uint32_t>(off) | (static_cast<uint32_t>(len / 3) << 16);
    for (int i = 0; i < 3; ++i) rows[3 * p + i] = static_cast<int32_t>(n + i * comp_stride);
    if (p + 1 == lo[b + 1]) blk_vals[b] = static_cast<int32_t>(off + vals[p]);
  }
}

// (block, item) keys of the work items (cell c, node a), item = 10 c + a
__global__ __launch_bounds__(kVT) void vp_item_keys(int64_t nc, const int32_t* __restrict__ cd,
                                                    const int32_t* __restrict__ owner, int ib,
                                                    uint64_t* __restrict__ keys) {
  VGRID_LOOP(k, nc * 10) {
    const int64_t c = k / 10;
    const int a = static_cast<int>(k - c * 10);
    keys[k] = (static_cast<uint64_t>(owner[cd[c * 30 + a]]) << ib) | static_cast<uint64_t>(k);
  }
}

// out[q] = first position i in keys[0, n) with (keys[i] >> shift) >= q * mul + add, mapped
// through map (map[i], or map_end at n) when given, minus sub[q] when given
__global__ __launch_bounds__(kVT) void vp_bounds(int64_t nq, const uint64_t* __restrict__ keys, int64_t n,
                                                 int shift, int64_t mul, int64_t add,
                                                 const int32_t* __restrict__ map, int32_t map_end,
                                                 const int32_t* __restrict__ sub, int32_t* __restrict__ out) {
  VGRID_LOOP(q, nq) {
    const uint64_t target = static_cast<uint64_t>(q * mul + add);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((keys[mid] >> shift) < target) lo = mid + 1;
      else hi = mid;
    }
    int64_t v = map ? (lo < n ? map[lo] : map_end) : lo;
    if (sub) v -= sub[q];
    out[q] = static_cast<int32_t>(v);
  }
}

// (block, vertex) for the four vertices of every item's cell, items in block order
__global__ __launch_bounds__(kVT) void vp_vert_keys(int64_t nt, const uint64_t* __restrict__ items, int ib,
                                                    const int32_t* __restrict__ cells, int vb,
                                                    uint64_t* __restrict__ keys) {
  const uint64_t mask = (uint64_t{1} << ib) - 1;
  VGRID_LOOP(t, nt) {
    const uint64_t b = items[t] >> ib;
    const int64_t c = static_cast<int64_t>(items[t] & mask) / 10;
    for (int v = 0; v < 4; ++v) keys[4 * t + v] = (b << vb) | static_cast<uint64_t>(cells[c * 4 + v]);
  }
}

__global__ __launch_bounds__(kVT) void vp_first(int64_t n, const uint64_t* __restrict__ keys,
                                                int32_t* __restrict__ flag) {
  VGRID_LOOP(i, n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(kVT) void vp_verts(int64_t n, const uint64_t* __restrict__ keys,
                                                const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                                int vb, const int32_t* __restrict__ vk, int32_t* __restrict__ vid,
                                                int32_t* __restrict__ vkd) {
  const uint64_t mask = (uint64_t{1} << vb) - 1;
  VGRID_LOOP(i, n) {
    if (!flag[i]) continue;
    const int32_t v = static_cast<int32_t>(keys[i] & mask);
    vid[pos[i]] = v;
    vkd[pos[i]] = vk[v];
  }
}

// vertex -> coefficient dof; a second pass finds vertices with two different dofs
__global__ __launch_bounds__(kVT) void vp_vk(int64_t n4, const int32_t* __restrict__ cells,
                                             const int32_t* __restrict__ kd, int32_t* __restrict__ vk) {
  VGRID_LOOP(i, n4) vk[cells[i]] = kd[i];
}

__global__ __launch_bounds__(kVT) void vp_vk_check(int64_t n4, const int32_t* __restrict__ cells,
                                                   const int32_t* __restrict__ kd, const int32_t* __restrict__ vk,
                                                   int32_t* __restrict__ flags) {
  VGRID_LOOP(i, n4) if (vk[cells[i]] != kd[i]) atomicOr(&flags[G_COEFF], 1);
}

// (block, owned node, item) keys
__global__ __launch_bounds__(kVT) void vp_round_keys(int64_t nt, const uint64_t* __restrict__ items, int ib,
                                                     const int32_t* __restrict__ cd,
                                                     const int32_t* __restrict__ own_index, int ob,
                                                     uint64_t* __restrict__ keys) {
  const uint64_t mask = (uint64_t{1} << ib) - 1;
  VGRID_LOOP(t, nt) {
    const uint64_t b = items[t] >> ib, it = items[t] & mask;
    const int64_t c = static_cast<int64_t>(it / 10);
    const int a = static_cast<int>(it - static_cast<uint64_t>(c) * 10);
    const uint64_t o = static_cast<uint64_t>(own_index[cd[c * 30 + a]]);
    keys[t] = (((b << ob) | o) << ib) | it;
  }
}

// position of each (block, owned node) run's head, 0 elsewhere (a maximum scan spreads it)
__global__ __launch_bounds__(kVT) void vp_heads(int64_t nt, const uint64_t* __restrict__ keys, int ib,
                                                int32_t* __restrict__ head) {
  VGRID_LOOP(t, nt) head[t] = (t == 0 || (keys[t] >> ib) != (keys[t - 1] >> ib)) ? static_cast<int32_t>(t) : 0;
}

// order keys (block, edge node, round, owned node): round = the item's index among the items
// of its owned node in cell order; value = the item
__global__ __launch_bounds__(kVT) void vp_final_keys(int64_t nt, const uint64_t* __restrict__ keys,
                                                     const int32_t* __restrict__ start, int ib, int ob, int rb,
                                                     uint64_t* __restrict__ fk, int32_t* __restrict__ fv,
                                                     int32_t* __restrict__ flags) {
  const uint64_t imask = (uint64_t{1} << ib) - 1, omask = (uint64_t{1} << ob) - 1;
  VGRID_LOOP(t, nt) {
    const uint64_t it = keys[t] & imask;
    const uint64_t bo = keys[t] >> ib;
    const uint64_t o = bo & omask, b = bo >> ob;
    const uint64_t seen = static_cast<uint64_t>(t - start[t]);
    if (seen >> rb) atomicOr(&flags[G_RANK], 1);
    const uint64_t edge = (it % 10) >= 4 ? 1 : 0;
    fk[t] = (((((b << 1) | edge) << rb) | seen) << ob) | o;
    fv[t] = static_cast<int32_t>(it);
  }
}

// task words of the items in final order:
// w0 = 4 x 16-bit local vertex ids; w1 = owned index (16) | a (8) << 16 | ranks b = 0..4
// (8 bits each) << 24; w2 = ranks b = 5..9 (8 bits each)
__global__ __launch_bounds__(kVT) void vp_tasks(int64_t nt, const int32_t* __restrict__ item,
                                                const uint64_t* __restrict__ fk, int bshift,
                                                const int32_t* __restrict__ cells, const int32_t* __restrict__ cd,
                                                const int32_t* __restrict__ own_index,
                                                const int32_t* __restrict__ vptr, const int32_t* __restrict__ vid,
                                                const uint8_t* __restrict__ rank, uint64_t* __restrict__ task,
                                                int32_t* __restrict__ task_cell) {
  VGRID_LOOP(t, nt) {
    const int64_t c = item[t] / 10;
    const int a = item[t] - static_cast<int>(c) * 10;
    const int32_t b = static_cast<int32_t>(fk[t] >> bshift);
    const int32_t v0 = vptr[b], v1 = vptr[b + 1];
    uint64_t w0 = 0;
    for (int v = 0; v < 4; ++v) {
      const int32_t x = cells[c * 4 + v];
      int32_t lo = v0, hi = v1;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (vid[mid] < x) lo = mid + 1;
        else hi = mid;
      }
      w0 |= static_cast<uint64_t>(lo - v0) << (16 * v);
    }
    uint64_t w1 = static_cast<uint64_t>(own_index[cd[c * 30 + a]]) | (static_cast<uint64_t>(a) << 16);
    uint64_t w2 = 0;
    for (int bb = 0; bb < 10; ++bb) {
      const uint64_t r = rank[c * 100 + a * 10 + bb];
      if (bb < 5) w1 |= r << (24 + 8 * bb);
      else w2 |= r << (8 * (bb - 5));
    }
    task[3 * t] = w0;
    task[3 * t + 1] = w1;
    task[3 * t + 2] = w2;
    task_cell[t] = static_cast<int32_t>(c);
  }
}

// temporaries of one build, freed once the stream has finished with them
struct VTemp {
  hipStream_t s;
  std::vector<void*> p;
  explicit VTemp(hipStream_t st) : s(st) {}
  ~VTemp() {
    (void)hipStreamSynchronize(s);
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  T* alloc(int64_t n) {
    void* d = nullptr;
    hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(n) * sizeof(T), 16)), "hipMalloc (vp2 plan)");
    p.push_back(d);
    return static_cast<T*>(d);
  }
};

template <class T>
T* varena(DeviceArena& a, int64_t n) {
  void* d = nullptr;
  hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(n) * sizeof(T), 16)), "hipMalloc (vp2 plan)");
  a.ptrs.push_back(d);
  a.bytes += n * static_cast<int64_t>(sizeof(T));
  return static_cast<T*>(d);
}

template <class T>
std::vector<T> vdown(const T* d, int64_t n, hipStream_t s) {
  std::vector<T> h(static_cast<size_t>(n));
  if (n) hip_check(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return h;
}

template <class T>
void vup(T* d, const std::vector<T>& h, hipStream_t s) {
  if (!h.empty())
    hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
}

template <class T>
void vscan(VTemp& tmp, const T* in, T* out, int64_t n, hipStream_t s) {
  size_t bytes = 0;
  hip_check(rocprim::exclusive_scan(nullptr, bytes, in, out, T{0}, static_cast<size_t>(n), rocprim::plus<T>(), s),
            "exclusive_scan (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::exclusive_scan(t, bytes, in, out, T{0}, static_cast<size_t>(n), rocprim::plus<T>(), s),
            "exclusive_scan");
}

template <class V>
void vsort_pairs(VTemp& tmp, uint64_t*& k, uint64_t*& ka, V*& v, V*& va, int64_t n, int bits, hipStream_t s) {
  rocprim::double_buffer<uint64_t> kk(k, ka);
  rocprim::double_buffer<V> vv(v, va);
  size_t bytes = 0;
  hip_check(rocprim::radix_sort_pairs(nullptr, bytes, kk, vv, static_cast<size_t>(n), 0,
                                      static_cast<unsigned>(bits), s),
            "radix_sort_pairs (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::radix_sort_pairs(t, bytes, kk, vv, static_cast<size_t>(n), 0, static_cast<unsigned>(bits), s),
            "radix_sort_pairs");
  k = kk.current();
  ka = kk.alternate();
  v = vv.current();
  va = vv.alternate();
}

void vsort_keys(VTemp& tmp, uint64_t*& k, uint64_t*& ka, int64_t n, int bits, hipStream_t s) {
  rocprim::double_buffer<uint64_t> kk(k, ka);
  size_t bytes = 0;
  hip_check(rocprim::radix_sort_keys(nullptr, bytes, kk, static_cast<size_t>(n), 0, static_cast<unsigned>(bits), s),
            "radix_sort_keys (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::radix_sort_keys(t, bytes, kk, static_cast<size_t>(n), 0, static_cast<unsigned>(bits), s),
            "radix_sort_keys");
  k = kk.current();
  ka = kk.alternate();
}

int64_t count_flags(const int32_t* flag, const int32_t* pos, int64_t n, hipStream_t s) {
  int32_t lp = 0, lf = 0;
  hip_check(hipMemcpyAsync(&lp, pos + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(&lf, flag + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return static_cast<int64_t>(lp) + lf;
}

}  // namespace

bool vp2_ranks_device(int64_t nc, const int32_t* d_cdofs, const int32_t* d_row_ptr, const int32_t* d_col,
                      bool node_major, uint8_t* d_rank, hipStream_t s) {
  if (nc <= 0) return true;
  VTemp tmp(s);
  int32_t* bad = tmp.alloc<int32_t>(1);
  hip_check(hipMemsetAsync(bad, 0, sizeof(int32_t), s), "hipMemsetAsync");
  hipLaunchKernelGGL(vp_rank_table, dim3(vgrid(nc * 10)), dim3(kVT), 0, s, nc, d_cdofs, d_row_ptr, d_col,
                     node_major, d_rank, bad);
  hip_check(hipGetLastError(), "vp_rank_table");
  return vdown(bad, 1, s)[0] == 0;
}

int build_vp2_plan_device(const DeviceVP2Input& in, int32_t max_block_vals, DeviceArena& plan, VP2BlockTopo& t,
                          int64_t& appearances, std::string& err, hipStream_t s) {
  const int64_t nc = in.nc, nd = in.ndofs;
  if (nc <= 0 || nd <= 0 || !in.node_layout || !in.d_rank || max_block_vals < 1) return kDevicePlanUnsupported;
  const bool node_major = in.node_layout == 2;
  VTemp tmp(s);
  // 1. node table
  double* dx = tmp.alloc<double>(in.nv * 3);
  hip_check(hipMemcpyAsync(dx, in.h_x, static_cast<size_t>(in.nv) * 3 * sizeof(double), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
  int32_t* isnode = tmp.alloc<int32_t>(nd);
  float* nx = tmp.alloc<float>(3 * nd);
  hip_check(hipMemsetAsync(isnode, 0, static_cast<size_t>(nd) * sizeof(int32_t), s), "hipMemsetAsync");
  hipLaunchKernelGGL(vp_node_table, dim3(vgrid(nc * 10)), dim3(kVT), 0, s, nc, in.d_cells, in.d_cdofs, dx, isnode,
                     nx);
  hip_check(hipGetLastError(), "vp_node_table");
  int32_t* pos = tmp.alloc<int32_t>(nd);
  vscan(tmp, isnode, pos, nd, s);
  const int64_t m = count_flags(isnode, pos, nd, s);  // nodes
  int32_t* perm = tmp.alloc<int32_t>(m);
  hipLaunchKernelGGL(vp_compact, dim3(vgrid(nd)), dim3(kVT), 0, s, nd, isnode, pos, perm);
  const int nbits = vbits(nd);
  // 2. recursive coordinate bisection, level by level; segments [lo, hi) on the host
  std::vector<int32_t> lo{0}, hi{static_cast<int32_t>(m)};
  std::vector<char> leaf{0};
  int32_t* seg = tmp.alloc<int32_t>(m);
  uint64_t* ka = tmp.alloc<uint64_t>(m);
  uint64_t* kb = tmp.alloc<uint64_t>(m);
  int32_t* va = tmp.alloc<int32_t>(m);
  int32_t* vb = tmp.alloc<int32_t>(m);
  int32_t* perm2 = tmp.alloc<int32_t>(m);
  for (;;) {
    const int32_t ns = static_cast<int32_t>(lo.size());
    VTemp lt(s);  // this level's buffers
    int32_t* dlo = lt.alloc<int32_t>(ns);
    vup(dlo, lo, s);
    hipLaunchKernelGGL(vp_segid, dim3(vgrid(m)), dim3(kVT), 0, s, m, ns, dlo, seg);
    auto* tot = lt.alloc<unsigned long long>(ns);
    auto* bmin = lt.alloc<uint32_t>(3 * static_cast<int64_t>(ns));
    auto* bmax = lt.alloc<uint32_t>(3 * static_cast<int64_t>(ns));
    hip_check(hipMemsetAsync(tot, 0, static_cast<size_t>(ns) * sizeof(unsigned long long), s), "hipMemsetAsync");
    hip_check(hipMemsetAsync(bmin, 0xff, 3 * static_cast<size_t>(ns) * sizeof(uint32_t), s), "hipMemsetAsync");
    hip_check(hipMemsetAsync(bmax, 0, 3 * static_cast<size_t>(ns) * sizeof(uint32_t), s), "hipMemsetAsync");
    hipLaunchKernelGGL(vp_reduce, dim3(vgrid((m + kRC - 1) / kRC)), dim3(kVT), 0, s, m, perm, seg, in.d_row_ptr,
                       nx, tot, bmin, bmax);
    hip_check(hipGetLastError(), "vp_reduce");
    const auto htot = vdown(tot, ns, s);
    const auto hmin = vdown(bmin, 3 * static_cast<int64_t>(ns), s);
    const auto hmax = vdown(bmax, 3 * static_cast<int64_t>(ns), s);
    // decide every open segment as the recursion does
    std::vector<int8_t> axis(static_cast<size_t>(ns), -1);
    std::vector<int32_t> mid(static_cast<size_t>(ns), 0);
    bool any = false;
    for (int32_t g = 0; g < ns; ++g) {
      if (leaf[g]) continue;
      const int64_t total = static_cast<int64_t>(htot[g]);
      if (total <= max_block_vals || hi[g] - lo[g] == 1) {
        if (total > max_block_vals) {
          err = "vector P2 node rows exceed the LDS block";
          return kDevicePlanError;
        }
        leaf[g] = 1;
        continue;
      }
      float ext[3];
      for (int q = 0; q < 3; ++q) ext[q] = unkey(hmax[3 * g + q]) - unkey(hmin[3 * g + q]);
      int ax = 0;
      for (int q = 1; q < 3; ++q)
        if (ext[q] > ext[ax]) ax = q;
      const int64_t nleaf = (total + max_block_vals - 1) / max_block_vals;
      int64_t md = lo[g] + static_cast<int64_t>(hi[g] - lo[g]) * (nleaf / 2) / nleaf;
      if (md <= lo[g]) md = lo[g] + 1;
      axis[g] = static_cast<int8_t>(ax);
      mid[g] = static_cast<int32_t>(md);
      any = true;
    }
    if (!any) break;
    // every splitting segment in (coordinate on its axis, node) order, segments in place
    int8_t* dax = lt.alloc<int8_t>(ns);
    vup(dax, axis, s);
    hipLaunchKernelGGL(vp_level_keys, dim3(vgrid(m)), dim3(kVT), 0, s, m, perm, seg, dax, nx, nbits, ka, va);
    hip_check(hipGetLastError(), "vp_level_keys");
    uint64_t *k1 = ka, *k2 = kb;
    int32_t *v1 = va, *v2 = vb;
    vsort_pairs(lt, k1, k2, v1, v2, m, 32 + nbits, s);  // v1: positions in (coordinate, node) order
    hipLaunchKernelGGL(vp_seg_keys, dim3(vgrid(m)), dim3(kVT), 0, s, m, v1, seg, k2);
    vsort_pairs(lt, k2, k1, v1, v2, m, vbits(ns), s);  // stable: (coordinate, node) order kept
    hipLaunchKernelGGL(vp_gather, dim3(vgrid(m)), dim3(kVT), 0, s, m, v1, perm, perm2);
    hip_check(hipGetLastError(), "vp_gather");
    std::swap(perm, perm2);
    std::vector<int32_t> nlo, nhi;
    std::vector<char> nleafv;
    for (int32_t g = 0; g < ns; ++g) {
      if (axis[g] < 0) {
        nlo.push_back(lo[g]);
        nhi.push_back(hi[g]);
        nleafv.push_back(leaf[g]);
      } else {
        nlo.push_back(lo[g]);
        nhi.push_back(mid[g]);
        nleafv.push_back(0);
        nlo.push_back(mid[g]);
        nhi.push_back(hi[g]);
        nleafv.push_back(0);
      }
    }
    lo.swap(nlo);
    hi.swap(nhi);
    leaf.swap(nleafv);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // before the level's buffers go
  }
  const int32_t nb = static_cast<int32_t>(lo.size());
  int32_t max_owned = 0;
  for (int32_t b = 0; b < nb; ++b) max_owned = std::max(max_owned, hi[b] - lo[b]);
  // 3. blocks: owned nodes ascending
  int32_t* dlo = tmp.alloc<int32_t>(nb + 1);
  {
    std::vector<int32_t> l2(lo);
    l2.push_back(static_cast<int32_t>(m));
    vup(dlo, l2, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // l2 dies here
  }
  hipLaunchKernelGGL(vp_segid, dim3(vgrid(m)), dim3(kVT), 0, s, m, nb, dlo, seg);
  hipLaunchKernelGGL(vp_leaf_keys, dim3(vgrid(m)), dim3(kVT), 0, s, m, perm, seg, nbits, ka);
  uint64_t *lk = ka, *lk2 = kb;
  vsort_keys(tmp, lk, lk2, m, vbits(nb) + nbits, s);
  int32_t* owner = tmp.alloc<int32_t>(nd);
  int32_t* own_index = tmp.alloc<int32_t>(nd);
  int64_t* vals = tmp.alloc<int64_t>(m);
  int64_t* voff = tmp.alloc<int64_t>(m);
  hipLaunchKernelGGL(vp_owner, dim3(vgrid(m)), dim3(kVT), 0, s, m, lk, nbits, dlo, in.d_row_ptr, perm, owner,
                     own_index, vals);
  hip_check(hipGetLastError(), "vp_owner");
  vscan(tmp, vals, voff, m, s);
  t = VP2BlockTopo();
  t.num_blocks = nb;
  t.component_major = !node_major;
  int32_t* own_ptr = varena<int32_t>(plan, nb + 1);
  hip_check(hipMemcpyAsync(own_ptr, dlo, static_cast<size_t>(nb + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s),
            "hipMemcpyAsync");
  uint32_t* info = varena<uint32_t>(plan, m);
  int32_t* rows = varena<int32_t>(plan, 3 * m);
  int32_t* blk_vals = varena<int32_t>(plan, nb);
  hipLaunchKernelGGL(vp_own, dim3(vgrid(m)), dim3(kVT), 0, s, m, perm, owner, dlo, voff, vals,
                     node_major ? int64_t{1} : nd / 3, info, rows, blk_vals);
  hip_check(hipGetLastError(), "vp_own");
  t.own_ptr = own_ptr;
  t.own_info = info;
  t.own_rows = rows;
  t.blk_vals = blk_vals;
  // work items per block, in cell order
  const int64_t nt = nc * 10;
  const int ib = vbits(nt), bb = vbits(nb), vb2 = vbits(in.nv), ob = vbits(max_owned);
  const int rb = std::min(24, 64 - bb - 1 - ob);
  if (bb + ib > 64 || bb + vb2 > 64 || bb + ob + ib > 64 || rb < 8) return kDevicePlanUnsupported;
  uint64_t* items = tmp.alloc<uint64_t>(nt);
  uint64_t* items2 = tmp.alloc<uint64_t>(nt);
  hipLaunchKernelGGL(vp_item_keys, dim3(vgrid(nt)), dim3(kVT), 0, s, nc, in.d_cdofs, owner, ib, items);
  hip_check(hipGetLastError(), "vp_item_keys");
  vsort_keys(tmp, items, items2, nt, bb + ib, s);
  int32_t* task_ptr = varena<int32_t>(plan, nb + 1);
  hipLaunchKernelGGL(vp_bounds, dim3(vgrid(nb + 1)), dim3(kVT), 0, s, int64_t{nb} + 1, items, nt, ib, int64_t{1},
                     int64_t{0}, nullptr, 0, nullptr, task_ptr);
  t.task_ptr = task_ptr;
  // block vertices: sorted unique (block, vertex)
  uint64_t* vka = tmp.alloc<uint64_t>(4 * nt);
  uint64_t* vkb = tmp.alloc<uint64_t>(4 * nt);
  hipLaunchKernelGGL(vp_vert_keys, dim3(vgrid(nt)), dim3(kVT), 0, s, nt, items, ib, in.d_cells, vb2, vka);
  hip_check(hipGetLastError(), "vp_vert_keys");
  vsort_keys(tmp, vka, vkb, 4 * nt, bb + vb2, s);
  int32_t* vflag = tmp.alloc<int32_t>(4 * nt);
  int32_t* vpos = tmp.alloc<int32_t>(4 * nt);
  hipLaunchKernelGGL(vp_first, dim3(vgrid(4 * nt)), dim3(kVT), 0, s, 4 * nt, vka, vflag);
  vscan(tmp, vflag, vpos, 4 * nt, s);
  const int64_t nverts = count_flags(vflag, vpos, 4 * nt, s);
  int32_t* flags = tmp.alloc<int32_t>(G_NFLAGS);
  hip_check(hipMemsetAsync(flags, 0, G_NFLAGS * sizeof(int32_t), s), "hipMemsetAsync");
  int32_t* vkmap = tmp.alloc<int32_t>(in.nv);
  hipLaunchKernelGGL(vp_vk, dim3(vgrid(nc * 4)), dim3(kVT), 0, s, nc * 4, in.d_cells, in.d_kdofs, vkmap);
  hipLaunchKernelGGL(vp_vk_check, dim3(vgrid(nc * 4)), dim3(kVT), 0, s, nc * 4, in.d_cells, in.d_kdofs, vkmap,
                     flags);
  int32_t* vid = varena<int32_t>(plan, nverts);
  int32_t* vkd = varena<int32_t>(plan, nverts);
  hipLaunchKernelGGL(vp_verts, dim3(vgrid(4 * nt)), dim3(kVT), 0, s, 4 * nt, vka, vflag, vpos, vb2, vkmap, vid, vkd);
  int32_t* vptr = varena<int32_t>(plan, nb + 1);
  hipLaunchKernelGGL(vp_bounds, dim3(vgrid(nb + 1)), dim3(kVT), 0, s, int64_t{nb} + 1, vka, 4 * nt, vb2,
                     int64_t{1}, int64_t{0}, vpos, static_cast<int32_t>(nverts), nullptr, vptr);
  hip_check(hipGetLastError(), "vp_verts");
  t.vert_ptr = vptr;
  t.vert_id = vid;
  t.vert_kdof = vkd;
  // item order: rounds over the owned nodes, vertex nodes first
  hipLaunchKernelGGL(vp_round_keys, dim3(vgrid(nt)), dim3(kVT), 0, s, nt, items, ib, in.d_cdofs, own_index, ob,
                     items2);
  hip_check(hipGetLastError(), "vp_round_keys");
  vsort_keys(tmp, items2, items, nt, bb + ob + ib, s);
  int32_t* head = tmp.alloc<int32_t>(nt);
  int32_t* start = tmp.alloc<int32_t>(nt);
  hipLaunchKernelGGL(vp_heads, dim3(vgrid(nt)), dim3(kVT), 0, s, nt, items2, ib, head);
  {
    size_t bytes = 0;
    hip_check(rocprim::inclusive_scan(nullptr, bytes, head, start, static_cast<size_t>(nt),
                                      rocprim::maximum<int32_t>(), s),
              "inclusive_scan (size)");
    void* tt = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
    hip_check(rocprim::inclusive_scan(tt, bytes, head, start, static_cast<size_t>(nt), rocprim::maximum<int32_t>(), s),
              "inclusive_scan");
  }
  uint64_t* fk = tmp.alloc<uint64_t>(nt);
  uint64_t* fk2 = tmp.alloc<uint64_t>(nt);
  int32_t* fv = tmp.alloc<int32_t>(nt);
  int32_t* fv2 = tmp.alloc<int32_t>(nt);
  hipLaunchKernelGGL(vp_final_keys, dim3(vgrid(nt)), dim3(kVT), 0, s, nt, items2, start, ib, ob, rb, fk, fv, flags);
  hip_check(hipGetLastError(), "vp_final_keys");
  vsort_pairs(tmp, fk, fk2, fv, fv2, nt, bb + 1 + rb + ob, s);
  // task words; vertex-node items lead each block
  uint64_t* task = varena<uint64_t>(plan, 3 * nt);
  int32_t* task_cell = varena<int32_t>(plan, nt);
  int32_t* nvert = varena<int32_t>(plan, nb);
  hipLaunchKernelGGL(vp_tasks, dim3(vgrid(nt)), dim3(kVT), 0, s, nt, fv, fk, 1 + rb + ob, in.d_cells, in.d_cdofs,
                     own_index, vptr, vid, in.d_rank, task, task_cell);
  hipLaunchKernelGGL(vp_bounds, dim3(vgrid(nb)), dim3(kVT), 0, s, int64_t{nb}, fk, nt, rb + ob, int64_t{2},
                     int64_t{1}, nullptr, 0, task_ptr, nvert);
  hip_check(hipGetLastError(), "vp_tasks");
  t.task = task;
  t.task_cell = task_cell;
  t.task_nvert = nvert;
  // checks (in the host build's order) and maxima
  const auto hf = vdown(flags, G_NFLAGS, s);
  if (hf[G_RANK]) return kDevicePlanUnsupported;  // a node in more than 2^rb items of a block
  const auto hvp = vdown(vptr, nb + 1, s);
  const auto hbv = vdown(blk_vals, nb, s);
  t.max_verts = t.max_vals = 0;
  t.max_owned = max_owned;
  for (int32_t b = 0; b < nb; ++b) {
    t.max_verts = std::max(t.max_verts, hvp[b + 1] - hvp[b]);
    t.max_vals = std::max(t.max_vals, hbv[b]);
  }
  if (t.max_verts > 65535) {
    err = "vector P2 plan: too many vertices in a block";
    return kDevicePlanError;
  }
  if (hf[G_COEFF]) {
    err = "coefficient dofmap is not a function of the vertex";
    return kDevicePlanError;
  }
  appearances = nt;
  return kDevicePlanOk;
}

}  // namespace dfa
