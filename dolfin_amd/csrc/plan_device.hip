// Device build of the P1 blocked-scatter plan and its edge chains for tetrahedral meshes:
// the arrays build_p1_block_plan and build_p1_chain_plan (structure.cpp) produce, computed
// on the GPU from the device-resident mesh, dofmaps and CSR structure, element for element
// equal to the host build. The partition is DOLFIN's locality ordering problem (the role of
// the graph partitioning in DofMapBuilder.cpp:1473-1667); the per-block cell order is the
// greedy colouring of MeshColoring.cpp:57-121 restricted to one block's rows.
//
//  1. vertex <-> dof maps and their consistency (a P1 dofmap must be a function of the vertex);
//  2. recursive coordinate bisection, level by level: every dof gets its rank along each axis
//     once (a stable radix sort of the coordinate bits over the dofs in ascending order, so
//     ties go to the smaller dof, as the host comparison does); each level then splits every
//     segment larger than the target at (n * (leaves / 2) / leaves) along its longest extent
//     by one radix sort of (segment, rank on that segment's axis). The leaves, in position
//     order, are the host recursion's depth-first leaves;
//  3. cell appearances: the (block, cell) pairs of every cell and each distinct block owning
//     one of its dofs, radix sorted, i.e. per block in ascending cell order;
//  4. one workgroup per block: owned dofs sorted, ghost vertices collected in an LDS hash set
//     and sorted, local node ids, owned-row offsets, diagonals and write-out runs, the greedy
//     "shares an owned row" colouring in cell order (one lane: it is sequential by
//     definition) and the stable (colour, smallest local node) order, the 4-bit column ranks
//     by binary search in the CSR rows, and the edge chains: face neighbours through per-node
//     cell lists, then the greedy chain growth of the host in seed order, the six candidate
//     edges of a seed walked by six lanes at once (the host keeps the first longest);
//  5. chain records in (length descending, seed order) and the fixed-stride slot arrays.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstring>

#include "assembler.h"

namespace dfa {
namespace {

constexpr int kT = 256;            // threads of the helper kernels
constexpr int kPT = 256;           // threads of the per-block planning kernel
constexpr int kGhostCap = 4096;    // LDS hash slots for a block's ghost vertices
constexpr int kDevMaxRows = 1024;  // largest leaf the device build takes

// flag words
enum { F_DOFMAP = 0, F_COEFF, F_KNOTID, F_KNOTINJ, F_NODES, F_NNZ, F_RANK, F_CHAIN, F_NFLAGS };

unsigned grid_for(int64_t n, int t = kT) {
  const int64_t b = (n + t - 1) / t;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 65535)));
}

#define GRID_LOOP(i, n)                                                                     \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

int bitlen(uint64_t v) {  // bits to hold v (at least 1)
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

// order-preserving map of a double onto uint64 (+0 and -0 equal)
__device__ inline uint64_t dkey(double v) {
  if (v == 0.0) v = 0.0;
  const uint64_t u = static_cast<uint64_t>(__double_as_longlong(v));
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
double dkey_inv(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  std::memcpy(&d, &u, sizeof d);
  return d;
}

__global__ void fill_i32(int64_t n, int32_t* p, int32_t v) {
  GRID_LOOP(i, n) p[i] = v;
}

__global__ void maps_kernel(int64_t n4, const int32_t* __restrict__ cells, const int32_t* __restrict__ cd,
                            const int32_t* __restrict__ kd, int32_t* vdof, int32_t* dofv, int32_t* vk) {
  GRID_LOOP(i, n4) {
    const int32_t v = cells[i], d = cd[i];
    vdof[v] = d;
    dofv[d] = v;
    vk[v] = kd[i];
  }
}

__global__ void maps_check_kernel(int64_t n4, const int32_t* __restrict__ cells,
                                  const int32_t* __restrict__ cd, const int32_t* __restrict__ kd,
                                  const int32_t* __restrict__ vdof, const int32_t* __restrict__ dofv,
                                  const int32_t* __restrict__ vk, int32_t* flags) {
  GRID_LOOP(i, n4) {
    const int32_t v = cells[i], d = cd[i];
    if (vdof[v] != d || dofv[d] != v) flags[F_DOFMAP] = 1;
    if (vk[v] != kd[i]) flags[F_COEFF] = 1;
  }
}

__global__ void kdof_kernel(int64_t nv, const int32_t* __restrict__ vk, int32_t* kcnt, int32_t* flags) {
  GRID_LOOP(v, nv) {
    const int32_t k = vk[v];
    if (k < 0) continue;
    if (k != v) flags[F_KNOTID] = 1;
    if (atomicAdd(&kcnt[k], 1) > 0) flags[F_KNOTINJ] = 1;
  }
}

__global__ void present_kernel(int64_t nd, const int32_t* __restrict__ dofv, int32_t* flag) {
  GRID_LOOP(d, nd) flag[d] = dofv[d] >= 0 ? 1 : 0;
}

// ids[pos[d]] = d for present dofs (pos = exclusive scan of the flags)
__global__ void compact_ids_kernel(int64_t nd, const int32_t* __restrict__ dofv,
                                   const int32_t* __restrict__ pos, int32_t* ids) {
  GRID_LOOP(d, nd) if (dofv[d] >= 0) ids[pos[d]] = static_cast<int32_t>(d);
}

__global__ void xs_kernel(int64_t n, int32_t gdim, const int32_t* __restrict__ ids,
                          const int32_t* __restrict__ dofv, const double* __restrict__ x, double* xs) {
  GRID_LOOP(p, n) {
    const int64_t v = dofv[ids[p]];
    for (int k = 0; k < gdim; ++k) xs[p * gdim + k] = x[v * gdim + k];
  }
}

__global__ void axis_keys_kernel(int64_t n, int32_t gdim, int32_t ax, const double* __restrict__ xs,
                                 uint64_t* keys, int32_t* vals) {
  GRID_LOOP(p, n) {
    keys[p] = dkey(xs[p * gdim + ax]);
    vals[p] = static_cast<int32_t>(p);
  }
}

__global__ void scatter_rank_kernel(int64_t n, const int32_t* __restrict__ sorted, int32_t* rank) {
  GRID_LOOP(j, n) rank[sorted[j]] = static_cast<int32_t>(j);
}

__global__ void iota_kernel(int64_t n, int32_t* p) {
  GRID_LOOP(i, n) p[i] = static_cast<int32_t>(i);
}

// per-segment coordinate extents of the segments still to split (wave-reduced when the 64
// lanes of a wave lie in one segment, which is the common case)
__global__ __launch_bounds__(kT) void bbox_kernel(int64_t n, int32_t gdim, const int32_t* __restrict__ sid,
                                                  const int32_t* __restrict__ perm,
                                                  const double* __restrict__ xs,
                                                  const int8_t* __restrict__ seg_axis,
                                                  unsigned long long* bmin, unsigned long long* bmax) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t base0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  for (int64_t b = base0; b < n; b += stride) {  // uniform per wave: all lanes iterate together
    const int64_t i = b + threadIdx.x;
    const bool in = i < n;
    const int32_t s = in ? sid[i] : -1;
    const bool act = in && seg_axis[s] >= 0;
    uint64_t mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
    if (act) {
      const int64_t p = perm[i];
      for (int k = 0; k < gdim; ++k) mn[k] = mx[k] = dkey(xs[p * gdim + k]);
    }
    const int32_t s0 = __shfl(s, 0);
    const bool uniform = __all(s == s0);
    if (uniform) {
      if (!__any(act)) continue;
      for (int k = 0; k < gdim; ++k)
        for (int off = 32; off > 0; off >>= 1) {
          const uint64_t a = __shfl_xor(mn[k], off), c = __shfl_xor(mx[k], off);
          mn[k] = a < mn[k] ? a : mn[k];
          mx[k] = c > mx[k] ? c : mx[k];
        }
      if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < gdim; ++k) {
          atomicMin(&bmin[s0 * 3 + k], static_cast<unsigned long long>(mn[k]));
          atomicMax(&bmax[s0 * 3 + k], static_cast<unsigned long long>(mx[k]));
        }
    } else if (act) {
      for (int k = 0; k < gdim; ++k) {
        atomicMin(&bmin[s * 3 + k], static_cast<unsigned long long>(mn[k]));
        atomicMax(&bmax[s * 3 + k], static_cast<unsigned long long>(mx[k]));
      }
    }
  }
}

__global__ void rcb_keys_kernel(int64_t n, int rb, const int32_t* __restrict__ sid,
                                const int32_t* __restrict__ perm, const int32_t* __restrict__ seg_lo,
                                const int8_t* __restrict__ seg_axis, const int32_t* __restrict__ r0,
                                const int32_t* __restrict__ r1, const int32_t* __restrict__ r2,
                                uint64_t* keys) {
  GRID_LOOP(i, n) {
    const int32_t s = sid[i];
    const int ax = seg_axis[s];
    uint64_t r;
    if (ax < 0) r = static_cast<uint64_t>(i - seg_lo[s]);
    else {
      const int32_t p = perm[i];
      r = static_cast<uint64_t>(ax == 0 ? r0[p] : ax == 1 ? r1[p] : r2[p]);
    }
    keys[i] = static_cast<uint64_t>(s) << rb | r;
  }
}

__global__ void rcb_sid_kernel(int64_t n, int32_t* sid, const int8_t* __restrict__ seg_axis,
                               const int32_t* __restrict__ seg_mid, const int32_t* __restrict__ newbase) {
  GRID_LOOP(i, n) {
    const int32_t s = sid[i];
    sid[i] = newbase[s] + (seg_axis[s] >= 0 && i >= seg_mid[s] ? 1 : 0);
  }
}

__global__ void owner_kernel(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ ids,
                             const int32_t* __restrict__ sid, int32_t* owner, int32_t* rowdof_u) {
  GRID_LOOP(i, n) {
    const int32_t d = ids[perm[i]];
    owner[d] = sid[i];
    rowdof_u[i] = d;
  }
}

__device__ inline int distinct_owners(const int32_t* __restrict__ cd, const int32_t* __restrict__ owner,
                                      int64_t c, int32_t* bl) {
  int n = 0;
  for (int a = 0; a < 4; ++a) {
    const int32_t b = owner[cd[c * 4 + a]];
    bool seen = false;
    for (int k = 0; k < n; ++k) seen |= bl[k] == b;
    if (!seen) bl[n++] = b;
  }
  return n;
}

__global__ void app_count_kernel(int64_t nc, const int32_t* __restrict__ cd,
                                 const int32_t* __restrict__ owner, int32_t* cnt) {
  GRID_LOOP(c, nc) {
    int32_t bl[4];
    cnt[c] = distinct_owners(cd, owner, c, bl);
  }
}

__global__ void app_fill_kernel(int64_t nc, const int32_t* __restrict__ cd, const int32_t* __restrict__ owner,
                                const int32_t* __restrict__ off, int cbits, uint64_t* keys) {
  GRID_LOOP(c, nc) {
    int32_t bl[4];
    const int n = distinct_owners(cd, owner, c, bl);
    for (int k = 0; k < n; ++k)
      keys[off[c] + k] = static_cast<uint64_t>(static_cast<uint32_t>(bl[k])) << cbits | static_cast<uint64_t>(c);
  }
}

// cell_ptr[b] = first appearance of block b; app[k] = cell of appearance k
__global__ void app_bounds_kernel(int64_t napp, int64_t nb, int cbits, const uint64_t* __restrict__ keys,
                                  int32_t* cell_ptr, int32_t* app) {
  const uint64_t mask = (uint64_t{1} << cbits) - 1;
  GRID_LOOP(k, napp + 1) {
    const int64_t prev = k == 0 ? -1 : static_cast<int64_t>(keys[k - 1] >> cbits);
    const int64_t cur = k == napp ? nb : static_cast<int64_t>(keys[k] >> cbits);
    for (int64_t r = prev + 1; r <= cur; ++r) cell_ptr[r] = static_cast<int32_t>(k);
    if (k < napp) app[k] = static_cast<int32_t>(keys[k] & mask);
  }
}

// ---- per-block planning kernel --------------------------------------------------------------

struct BlockArgs {
  int32_t nb;
  const int32_t* cells;
  const int32_t* vdof;
  const int32_t* dofv;
  const int32_t* owner;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* seg_lo;    // [nb + 1] leaf boundaries (the owned-row pointer)
  const int32_t* rowdof_u;  // dofs of each leaf, in RCB order
  const int32_t* cell_ptr;  // [nb + 1]
  const int32_t* app;       // cells of each block, ascending
  int32_t R;                // power of two >= the largest leaf
  int32_t C;                // power of two >= the largest cell list
  int32_t runs_stride;      // >= the largest leaf
  int32_t wide;             // 8-bit column ranks
  int32_t chains;           // build the edge chains
  // LDS byte offsets
  int32_t o_sdof, o_nodes, o_hash, o_gtmp, o_lens, o_st, o_cl, o_keys, o_used;
  int32_t o_incp, o_usedc, o_nbr, o_inc;
  // outputs
  int32_t* row_dof;
  uint32_t* row_info;
  uint8_t* row_diag;
  int32_t* blk_nnz;
  int32_t* node_v;   // [nb x kMaxBlockNodes]
  int32_t* nnodes;   // [nb]
  uint64_t* runs_s;  // [nb x runs_stride]
  int32_t* nruns;    // [nb]
  uint64_t* cell_rec;
  uint64_t* cell_rank_wide;
  uint16_t* desc;    // 16 per chain, at the block's first appearance
  int32_t* nchains;  // [nb]
  int32_t* classoff; // [nb x 8]: first slot of chains of length m (longest first)
  int32_t* flags;
};

template <class T>
__device__ void lds_bitonic(T* a, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const T x = a[i], y = a[l];
          if ((x > y) == ((i & k) == 0)) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

// exclusive prefix sum of a[0..n) in place; returns the total (all threads)
__device__ int32_t lds_scan(int32_t* a, int n, int32_t* part) {
  const int nt = blockDim.x, tid = threadIdx.x;
  const int per = (n + nt - 1) / nt;
  const int b0 = min(n, tid * per), b1 = min(n, b0 + per);
  int32_t s = 0;
  for (int i = b0; i < b1; ++i) s += a[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    const int32_t v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int32_t run = part[tid] - s;
  const int32_t total = part[nt - 1];
  for (int i = b0; i < b1; ++i) {
    const int32_t x = a[i];
    a[i] = run;
    run += x;
  }
  __syncthreads();
  return total;
}

__device__ inline int lower_bound_i32(const int32_t* a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__device__ inline uint32_t lid(uint64_t w, int a) { return static_cast<uint32_t>(w >> (16 * a)) & 0xffffu; }

__global__ __launch_bounds__(kPT) void block_plan_kernel(BlockArgs A) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int32_t part[kPT];
  __shared__ int32_t misc[16];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lo = A.seg_lo[b], nr = A.seg_lo[b + 1] - lo;
  const int c0 = A.cell_ptr[b], ncb = A.cell_ptr[b + 1] - c0;
  int32_t* sdof = reinterpret_cast<int32_t*>(lds + A.o_sdof);
  int32_t* nodes = reinterpret_cast<int32_t*>(lds + A.o_nodes);
  int32_t* hash = reinterpret_cast<int32_t*>(lds + A.o_hash);
  int32_t* gtmp = reinterpret_cast<int32_t*>(lds + A.o_gtmp);

  // ---- owned rows, sorted by dof
  for (int i = tid; i < A.R; i += nt) sdof[i] = i < nr ? A.rowdof_u[lo + i] : INT_MAX;
  for (int i = tid; i < kGhostCap; i += nt) hash[i] = -1;
  if (tid < 16) misc[tid] = 0;
  __syncthreads();
  lds_bitonic(sdof, A.R);
  for (int i = tid; i < nr; i += nt) nodes[i] = A.dofv[sdof[i]];
  // ---- ghost vertices: every vertex of the block's cells not owned by it
  for (int j = tid; j < 4 * ncb; j += nt) {
    const int32_t v = A.cells[static_cast<int64_t>(A.app[c0 + (j >> 2)]) * 4 + (j & 3)];
    if (A.owner[A.vdof[v]] == b) continue;
    uint32_t h = (static_cast<uint32_t>(v) * 2654435761u) >> 20;  // 12 bits
    bool placed = false;
    for (int probe = 0; probe < kGhostCap; ++probe) {
      const int32_t old = atomicCAS(&hash[h], -1, v);
      if (old == -1) {
        atomicAdd(&misc[0], 1);
        placed = true;
        break;
      }
      if (old == v) {
        placed = true;
        break;
      }
      h = (h + 1) & (kGhostCap - 1);
    }
    if (!placed) misc[1] = 1;
  }
  __syncthreads();
  const int ng = misc[0];
  if (misc[1] || nr + ng > kMaxBlockNodes) {
    if (tid == 0) A.flags[F_NODES] = 1;
    return;  // uniform: every thread read the same LDS words after the barrier
  }
  for (int i = tid; i < kGhostCap; i += nt) {
    const int32_t v = hash[i];
    if (v != -1) gtmp[atomicAdd(&misc[2], 1)] = v;
  }
  int G = 1;
  while (G < ng) G <<= 1;
  __syncthreads();
  for (int i = ng + tid; i < G; i += nt) gtmp[i] = INT_MAX;
  __syncthreads();
  lds_bitonic(gtmp, G);
  const int nn = nr + ng;
  for (int i = tid; i < ng; i += nt) nodes[nr + i] = gtmp[i];
  __syncthreads();
  for (int i = tid; i < nn; i += nt) A.node_v[static_cast<int64_t>(b) * kMaxBlockNodes + i] = nodes[i];
  if (tid == 0) A.nnodes[b] = nn;

  // ---- owned rows: CSR lengths, diagonals, LDS offsets, write-out runs
  int32_t* lens = reinterpret_cast<int32_t*>(lds + A.o_lens);
  int32_t* st = reinterpret_cast<int32_t*>(lds + A.o_st);
  for (int i = tid; i < nr; i += nt) {
    const int32_t d = sdof[i];
    const int32_t rs = A.row_ptr[d], len = A.row_ptr[d + 1] - rs;
    lens[i] = len;
    const int pos = lower_bound_i32(A.col + rs, len, d);
    A.row_diag[lo + i] = static_cast<uint8_t>(min(pos, 255));
    A.row_dof[lo + i] = d;
    st[i] = (i == 0 || sdof[i] != sdof[i - 1] + 1) ? 1 : 0;
  }
  __syncthreads();
  const int32_t total = lds_scan(lens, nr, part);  // lens[i] = LDS offset of owned row i
  if (total > kMaxBlockNnz) {
    if (tid == 0) A.flags[F_NNZ] = 1;
    return;
  }
  for (int i = tid; i < nr; i += nt) A.row_info[lo + i] = static_cast<uint32_t>(i) | (static_cast<uint32_t>(lens[i]) << 16);
  if (tid == 0) A.blk_nnz[b] = total;
  // write-out runs: maximal sequences of consecutive dofs; run index = starts before it
  const int32_t nrun = lds_scan(st, nr, part);
  for (int i = tid; i < nr; i += nt) {
    const bool start_here = (i == 0 || sdof[i] != sdof[i - 1] + 1);
    if (!start_here) continue;
    int j = i + 1;
    while (j < nr && sdof[j] == sdof[j - 1] + 1) ++j;
    const uint64_t loff = static_cast<uint32_t>(lens[i]);
    const uint64_t len = static_cast<uint32_t>((j < nr ? lens[j] : total) - lens[i]);
    const uint64_t start = static_cast<uint32_t>(A.row_ptr[sdof[i]]);
    A.runs_s[static_cast<int64_t>(b) * A.runs_stride + st[i]] = start | (loff << 32) | (len << 48);
  }
  if (tid == 0) A.nruns[b] = nrun;
  __syncthreads();

  // ---- local node ids of the cells (ascending cell order)
  uint64_t* cl = reinterpret_cast<uint64_t*>(lds + A.o_cl);
  uint64_t* keys = reinterpret_cast<uint64_t*>(lds + A.o_keys);
  uint64_t* used = reinterpret_cast<uint64_t*>(lds + A.o_used);
  for (int j = tid; j < ncb; j += nt) {
    const int64_t c = A.app[c0 + j];
    uint64_t w = 0;
    for (int a = 0; a < 4; ++a) {
      const int32_t v = A.cells[c * 4 + a];
      const int32_t d = A.vdof[v];
      const int l = A.owner[d] == b ? lower_bound_i32(sdof, nr, d) : nr + lower_bound_i32(nodes + nr, ng, v);
      w |= static_cast<uint64_t>(l) << (16 * a);
    }
    cl[j] = w;
  }
  for (int i = tid; i < nr; i += nt) used[i] = 0;
  for (int j = ncb + tid; j < A.C; j += nt) keys[j] = ~0ull;
  __syncthreads();
  // greedy colouring in cell order, one lane
  if (tid == 0) {
    for (int j = 0; j < ncb; ++j) {
      const uint64_t w = cl[j];
      uint64_t busy = 0;
      uint32_t lmin = 0xffffu;
      for (int a = 0; a < 4; ++a) {
        const uint32_t l = lid(w, a);
        lmin = min(lmin, l);
        if (static_cast<int>(l) < nr) busy |= used[l];
      }
      const uint64_t freeb = ~busy & 0x7fffffffffffffffull;
      const int colour = freeb ? __ffsll(static_cast<long long>(freeb)) - 1 : 63;
      for (int a = 0; a < 4; ++a) {
        const uint32_t l = lid(w, a);
        if (static_cast<int>(l) < nr) used[l] |= 1ull << colour;
      }
      keys[j] = (static_cast<uint64_t>(colour) * 65536u + lmin) << 32 | static_cast<uint32_t>(j);
    }
  }
  __syncthreads();
  lds_bitonic(keys, A.C);
  // column ranks, in the sorted order
  for (int t = tid; t < ncb; t += nt) {
    const int j = static_cast<int>(keys[t] & 0xffffffffu);
    const uint64_t w = cl[j];
    uint64_t rk = 0, wr[2] = {0, 0};
    for (int a = 0; a < 4; ++a) {
      const int la = static_cast<int>(lid(w, a));
      if (la >= nr) continue;
      const int32_t da = sdof[la];
      const int32_t rs = A.row_ptr[da], len = A.row_ptr[da + 1] - rs;
      for (int bb = 0; bb < 4; ++bb) {
        const int lb = static_cast<int>(lid(w, bb));
        const int32_t db = lb < nr ? sdof[lb] : A.vdof[nodes[lb]];
        const uint32_t r = static_cast<uint32_t>(lower_bound_i32(A.col + rs, len, db));
        if (r >= (A.wide ? 256u : 16u)) {
          A.flags[F_RANK] = 1;
          continue;
        }
        if (A.wide) wr[a >> 1] |= static_cast<uint64_t>(r) << (8 * ((a & 1) * 4 + bb));
        else rk |= static_cast<uint64_t>(r) << (4 * (a * 4 + bb));
      }
    }
    A.cell_rec[2 * static_cast<int64_t>(c0 + t)] = w;
    A.cell_rec[2 * static_cast<int64_t>(c0 + t) + 1] = rk;
    if (A.wide) {
      A.cell_rank_wide[2 * static_cast<int64_t>(c0 + t)] = wr[0];
      A.cell_rank_wide[2 * static_cast<int64_t>(c0 + t) + 1] = wr[1];
    }
  }
  __syncthreads();
  if (!A.chains) return;

  // ---- edge chains over the sorted cells
  // keys[t] becomes the local ids of sorted cell t (each thread rewrites its own entries)
  for (int t = tid; t < ncb; t += nt) keys[t] = cl[static_cast<int>(keys[t] & 0xffffffffu)];
  __syncthreads();
  const uint64_t* cs = keys;
  int32_t* incp = reinterpret_cast<int32_t*>(lds + A.o_incp);  // [nn + 1]
  uint8_t* usedc = lds + A.o_usedc;                             // [ncb]
  int16_t* nbr = reinterpret_cast<int16_t*>(lds + A.o_nbr);     // [4 ncb]
  uint16_t* inc = reinterpret_cast<uint16_t*>(lds + A.o_inc);   // [4 ncb]
  for (int i = tid; i <= nn; i += nt) incp[i] = 0;
  for (int k = tid; k < ncb; k += nt) usedc[k] = 0;
  __syncthreads();
  for (int j = tid; j < 4 * ncb; j += nt) atomicAdd(&incp[lid(cs[j >> 2], j & 3)], 1);
  __syncthreads();
  lds_scan(incp, nn + 1, part);  // incp[u] = first slot of node u
  for (int j = tid; j < 4 * ncb; j += nt) {
    const int pos = atomicAdd(&incp[lid(cs[j >> 2], j & 3)], 1);
    inc[pos] = static_cast<uint16_t>(j >> 2);
  }
  __syncthreads();
  // now incp[u] = end of node u's list, its start is incp[u - 1] (0 for u = 0)
  for (int j = tid; j < 4 * ncb; j += nt) {
    const int k = j >> 2, o = j & 3;
    const uint64_t w = cs[k];
    uint32_t f[3];
    int n3 = 0;
    for (int a = 0; a < 4; ++a)
      if (a != o) f[n3++] = lid(w, a);
    const int u = static_cast<int>(f[0]);
    const int s0 = u == 0 ? 0 : incp[u - 1], s1 = incp[u];
    int found = -1;
    for (int s = s0; s < s1; ++s) {
      const int c = inc[s];
      if (c == k) continue;
      const uint64_t wc = cs[c];
      bool h1 = false, h2 = false;
      for (int a = 0; a < 4; ++a) {
        const uint32_t v = lid(wc, a);
        h1 |= v == f[1];
        h2 |= v == f[2];
      }
      if (h1 && h2) {
        found = c;
        break;
      }
    }
    nbr[j] = static_cast<int16_t>(found);
  }
  __syncthreads();
  if (tid < 64) {
    constexpr int K = kChainLen;
    const int e = tid;  // lanes 0..5 walk edge e of the seed
    const int ea = e < 6 ? (0x211000 >> (4 * e)) & 0xf : 0;
    const int eb = e < 6 ? (0x332321 >> (4 * e)) & 0xf : 1;
    int cnt_m[K + 1];
#pragma unroll
    for (int m = 0; m <= K; ++m) cnt_m[m] = 0;
    int nch = 0;
    for (int k0 = 0; k0 < ncb; ++k0) {
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
      if (usedc[k0]) continue;
      const uint64_t w0 = cs[k0];
      const uint32_t p = lid(w0, ea), q = lid(w0, eb);
      const int rest = 0xf & ~(1 << ea) & ~(1 << eb);
      const int a0 = __ffs(rest) - 1, a1 = __ffs(rest & (rest - 1)) - 1;
      const uint32_t rs0 = lid(w0, a0), rs1 = lid(w0, a1);
      int32_t fc[K - 1], gc[K - 1];
      uint32_t fr[K - 1], gr[K - 1];
      int nf = 0, ngc = 0;
      // the cell across face (p, q, r) of cell k (the face opposite its vertex not in it)
      auto across = [&](int32_t k, uint32_t r) -> int32_t {
        const uint64_t w = cs[k];
        int o = 0;
        for (int a = 3; a >= 0; --a) {
          const uint32_t v = lid(w, a);
          if (v != p && v != q && v != r) o = a;
        }
        const int32_t c = nbr[4 * k + o];
        return (c >= 0 && !usedc[c]) ? c : -1;
      };
      auto fourth = [&](int32_t k, uint32_t r) -> uint32_t {
        const uint64_t w = cs[k];
        uint32_t res = 0xffffffffu;
        for (int a = 3; a >= 0; --a) {
          const uint32_t v = lid(w, a);
          if (v != p && v != q && v != r) res = v;
        }
        return res;
      };
      if (e < 6) {
        int32_t cur = k0;
        uint32_t rl = rs1;
        bool go = true;
#pragma unroll
        for (int s = 0; s < K - 1; ++s) {
          if (go) {
            const int32_t c = across(cur, rl);
            bool dup = c == k0;
#pragma unroll
            for (int t = 0; t < s; ++t) dup |= c == fc[t];
            if (c < 0 || dup) {
              go = false;
            } else {
              const uint32_t nx = fourth(c, rl);
              fc[s] = c;
              fr[s] = nx;
              nf = s + 1;
              cur = c;
              rl = nx;
            }
          }
        }
        cur = k0;
        uint32_t rf = rs0;
        go = true;
#pragma unroll
        for (int s = 0; s < K - 1; ++s) {
          if (go && 1 + nf + s < K) {
            const int32_t c = across(cur, rf);
            bool dup = c == k0;
#pragma unroll
            for (int t = 0; t < K - 1; ++t) dup |= (t < nf && c == fc[t]);
#pragma unroll
            for (int t = 0; t < s; ++t) dup |= c == gc[t];
            if (c < 0 || dup) {
              go = false;
            } else {
              const uint32_t nx = fourth(c, rf);
              gc[s] = c;
              gr[s] = nx;
              ngc = s + 1;
              cur = c;
              rf = nx;
            }
          }
        }
      }
      const int m = e < 6 ? 1 + nf + ngc : 0;
      int key = e < 6 ? m * 8 + (7 - e) : 0;
#pragma unroll
      for (int off = 1; off < 8; off <<= 1) key = max(key, __shfl_xor(key, off));
      const int best_e = 7 - (key & 7), best_m = key >> 3;
      int idx = 0;
#pragma unroll
      for (int mm = 1; mm <= K; ++mm)
        if (mm == best_m) idx = cnt_m[mm];
      if (e == best_e) {
        uint16_t* dsc = A.desc + 16 * static_cast<int64_t>(c0 + nch);
        dsc[0] = static_cast<uint16_t>(p);
        dsc[1] = static_cast<uint16_t>(q);
        // cells g[ng-1] .. g[0], k0, f[0] .. f[nf-1]; nodes gr[ng-1] .. gr[0], rs0, rs1, fr[..]
#pragma unroll
        for (int t = 0; t < K - 1; ++t)
          if (t < ngc) {
            dsc[9 + (ngc - 1 - t)] = static_cast<uint16_t>(gc[t]);
            dsc[2 + (ngc - 1 - t)] = static_cast<uint16_t>(gr[t]);
            usedc[gc[t]] = 1;
          }
        dsc[9 + ngc] = static_cast<uint16_t>(k0);
        dsc[2 + ngc] = static_cast<uint16_t>(rs0);
        dsc[2 + ngc + 1] = static_cast<uint16_t>(rs1);
        usedc[k0] = 1;
#pragma unroll
        for (int t = 0; t < K - 1; ++t)
          if (t < nf) {
            dsc[9 + ngc + 1 + t] = static_cast<uint16_t>(fc[t]);
            dsc[2 + ngc + 2 + t] = static_cast<uint16_t>(fr[t]);
            usedc[fc[t]] = 1;
          }
        dsc[15] = static_cast<uint16_t>(best_m | (idx << 3));
      }
#pragma unroll
      for (int mm = 1; mm <= K; ++mm)
        if (mm == best_m) cnt_m[mm] += 1;
      ++nch;
    }
    if (tid == 0) {
      A.nchains[b] = nch;
      int off = 0;
      for (int mm = K; mm >= 1; --mm) {
        A.classoff[b * 8 + mm] = off;
        off += cnt_m[mm];
      }
    }
  }
}

// chain records at their final positions (longest first, seed order within a length)
__global__ __launch_bounds__(kT) void chain_records_kernel(
    int32_t nb, const int32_t* __restrict__ cell_ptr, const int32_t* __restrict__ chain_ptr,
    const int32_t* __restrict__ classoff, const uint16_t* __restrict__ desc,
    const uint64_t* __restrict__ cell_rec, uint32_t* chain_rec, int32_t* flags) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const int c0 = cell_ptr[b];
  const int n = chain_ptr[b + 1] - chain_ptr[b];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint16_t* d = desc + 16 * static_cast<int64_t>(c0 + i);
    uint16_t h[16];
    for (int k = 0; k < 16; ++k) h[k] = d[k];
    const int m = h[15] & 7, idx = h[15] >> 3;
    const uint32_t p = h[0], q = h[1];
    uint32_t ids[9];
    ids[0] = p;
    ids[1] = q;
    for (int k = 0; k < 7; ++k) ids[2 + k] = k <= m ? h[2 + k] : p;
    uint32_t* w = chain_rec + static_cast<int64_t>(chain_ptr[b] + classoff[b * 8 + m] + idx) * kChainWords;
    w[0] = ids[0] | ids[1] << 16;
    w[1] = ids[2] | ids[3] << 16;
    w[2] = ids[4] | ids[5] << 16;
    w[3] = ids[6] | ids[7] << 16;
    w[4] = ids[8] | static_cast<uint32_t>(m) << 16;
    uint32_t bytes[6] = {0, 0, 0, 0, 0, 0};
    bool bad = false;
    auto rank_in = [&](int k, uint32_t u, uint32_t v) -> uint32_t {
      const uint64_t cw = cell_rec[2 * static_cast<int64_t>(c0 + k)];
      const uint64_t rk = cell_rec[2 * static_cast<int64_t>(c0 + k) + 1];
      int au = -1, av = -1;
      for (int a = 0; a < 4; ++a) {
        if (lid(cw, a) == u) au = a;
        if (lid(cw, a) == v) av = a;
      }
      if (au < 0 || av < 0) return 0xffu;
      return static_cast<uint32_t>((rk >> (4 * (au * 4 + av))) & 15u);
    };
    auto pair = [&](int j, int k, uint32_t u, uint32_t v) {
      const uint32_t a = rank_in(k, u, v), c = rank_in(k, v, u);
      if (a > 15 || c > 15) bad = true;
      bytes[j >> 2] |= ((a & 15u) | (c & 15u) << 4) << (8 * (j & 3));
    };
    for (int k = 0; k < m; ++k) {
      const int cell = h[9 + k];
      pair(3 * k, cell, p, h[2 + k]);
      pair(3 * k + 1, cell, q, h[2 + k]);
      pair(3 * k + 2, cell, h[2 + k], h[2 + k + 1]);
    }
    pair(18, h[9 + m - 1], p, h[2 + m]);
    pair(19, h[9 + m - 1], q, h[2 + m]);
    pair(20, h[9], p, q);
    for (int k = 0; k < 6; ++k) w[5 + k] = bytes[k];
    w[11] = 0;
    if (bad) flags[F_CHAIN] = 1;
  }
}

__global__ __launch_bounds__(kT) void nodes_out_kernel(
    int32_t nb, int32_t SN, const int32_t* __restrict__ node_v, const int32_t* __restrict__ node_ptr,
    const int32_t* __restrict__ vk, const int32_t* __restrict__ vdof, int32_t* node_vertex,
    int32_t* node_kdof, int32_t* node_dof, int32_t* slot_vertex, int32_t* slot_kdof) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const int n0 = node_ptr[b], nn = node_ptr[b + 1] - n0;
  const int32_t* src = node_v + static_cast<int64_t>(b) * kMaxBlockNodes;
  for (int i = threadIdx.x; i < nn; i += blockDim.x) {
    const int32_t v = src[i];
    node_vertex[n0 + i] = v;
    node_kdof[n0 + i] = vk[v];
    node_dof[n0 + i] = vdof[v];
  }
  if (slot_vertex)
    for (int i = threadIdx.x; i < SN; i += blockDim.x) {
      const int32_t v = src[i < nn ? i : 0];
      slot_vertex[static_cast<int64_t>(b) * SN + i] = v;
      if (slot_kdof) slot_kdof[static_cast<int64_t>(b) * SN + i] = vk[v];
    }
}

__global__ __launch_bounds__(kT) void rows_out_kernel(
    int32_t nb, int32_t SR, int32_t runs_stride, const int32_t* __restrict__ seg_lo,
    const int32_t* __restrict__ row_dof, const uint8_t* __restrict__ row_diag,
    const int32_t* __restrict__ row_ptr, const uint64_t* __restrict__ runs_s,
    const int32_t* __restrict__ run_ptr, const int32_t* __restrict__ chain_ptr,
    const int32_t* __restrict__ node_ptr, uint64_t* runs, int4* blk_hdr, int2* slot_row,
    int32_t* slot_rdof) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const int r0 = run_ptr[b], nrun = run_ptr[b + 1] - r0;
  for (int i = threadIdx.x; i < nrun; i += blockDim.x)
    runs[r0 + i] = runs_s[static_cast<int64_t>(b) * runs_stride + i];
  if (!blk_hdr) return;
  const int lo = seg_lo[b], nr = seg_lo[b + 1] - lo;
  if (threadIdx.x == 0)
    blk_hdr[b] = make_int4(chain_ptr[b], chain_ptr[b + 1] - chain_ptr[b], node_ptr[b + 1] - node_ptr[b], nr);
  for (int i = threadIdx.x; i < SR; i += blockDim.x) {
    int2 sr = make_int2(0, 0);
    int32_t sd = 0;
    if (i < nr) {
      const int32_t d = row_dof[lo + i];
      const int32_t c0 = row_ptr[d], len = row_ptr[d + 1] - c0;
      sr = make_int2(c0, len | (static_cast<int32_t>(row_diag[lo + i]) << 8));
      sd = d;
    }
    slot_row[static_cast<int64_t>(b) * SR + i] = sr;
    slot_rdof[static_cast<int64_t>(b) * SR + i] = sd;
  }
}

// ---- host side -------------------------------------------------------------------------------

struct Temp {
  std::vector<void*> p;
  ~Temp() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  T* alloc(int64_t n) {
    void* d = nullptr;
    hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(n) * sizeof(T), 16)), "hipMalloc (plan)");
    p.push_back(d);
    return static_cast<T*>(d);
  }
};

template <class T>
T* arena_alloc(DeviceArena& a, int64_t n) {
  void* d = nullptr;
  hip_check(hipMalloc(&d, std::max<size_t>(static_cast<size_t>(n) * sizeof(T), 16)), "hipMalloc (plan)");
  a.ptrs.push_back(d);
  a.bytes += n * static_cast<int64_t>(sizeof(T));
  return static_cast<T*>(d);
}

template <class T>
std::vector<T> download(const T* d, int64_t n, hipStream_t s) {
  std::vector<T> h(static_cast<size_t>(n));
  if (n) hip_check(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return h;
}

template <class T>
void upload_to(T* d, const std::vector<T>& h, hipStream_t s) {
  if (!h.empty())
    hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
}

void exclusive_scan_i32(Temp& tmp, const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  size_t bytes = 0;
  hip_check(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t{0}, static_cast<size_t>(n),
                                    rocprim::plus<int32_t>(), s), "exclusive_scan (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::exclusive_scan(t, bytes, in, out, int32_t{0}, static_cast<size_t>(n),
                                    rocprim::plus<int32_t>(), s), "exclusive_scan");
}

template <class V>
void sort_pairs(Temp& tmp, uint64_t*& keys, uint64_t*& keys_alt, V*& vals, V*& vals_alt, int64_t n,
                int end_bit, hipStream_t s) {
  rocprim::double_buffer<uint64_t> k(keys, keys_alt);
  rocprim::double_buffer<V> v(vals, vals_alt);
  size_t bytes = 0;
  hip_check(rocprim::radix_sort_pairs(nullptr, bytes, k, v, static_cast<size_t>(n), 0,
                                      static_cast<unsigned>(end_bit), s), "radix_sort_pairs (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::radix_sort_pairs(t, bytes, k, v, static_cast<size_t>(n), 0,
                                      static_cast<unsigned>(end_bit), s), "radix_sort_pairs");
  keys = k.current();
  keys_alt = k.alternate();
  vals = v.current();
  vals_alt = v.alternate();
}

void sort_keys(Temp& tmp, uint64_t*& keys, uint64_t*& keys_alt, int64_t n, int end_bit, hipStream_t s) {
  rocprim::double_buffer<uint64_t> k(keys, keys_alt);
  size_t bytes = 0;
  hip_check(rocprim::radix_sort_keys(nullptr, bytes, k, static_cast<size_t>(n), 0,
                                     static_cast<unsigned>(end_bit), s), "radix_sort_keys (size)");
  void* t = tmp.alloc<unsigned char>(static_cast<int64_t>(bytes));
  hip_check(rocprim::radix_sort_keys(t, bytes, k, static_cast<size_t>(n), 0,
                                     static_cast<unsigned>(end_bit), s), "radix_sort_keys");
  keys = k.current();
  keys_alt = k.alternate();
}

int32_t align16(int64_t v) { return static_cast<int32_t>((v + 15) & ~int64_t{15}); }

}  // namespace

int build_p1_plan_device(const DevicePlanInput& in, int32_t target_rows, bool want_chains,
                         int32_t block_threads, int32_t chain_threads, DeviceArena& plan,
                         P1BlockTopo& t, DevicePlanStats& stats, std::string& err, hipStream_t s) {
  if (in.gdim != 3 || target_rows < 1 || target_rows > kDevMaxRows) return kDevicePlanUnsupported;
  const int64_t nv = in.nv, nc = in.nc, ndofs = in.ndofs, gdim = in.gdim;
  Temp tmp;
  int32_t* flags = tmp.alloc<int32_t>(F_NFLAGS);
  hip_check(hipMemsetAsync(flags, 0, F_NFLAGS * sizeof(int32_t), s), "hipMemsetAsync");
  // 1. maps
  int32_t* vdof = tmp.alloc<int32_t>(nv);
  int32_t* vk = tmp.alloc<int32_t>(nv);
  int32_t* dofv = tmp.alloc<int32_t>(ndofs);
  hipLaunchKernelGGL(fill_i32, dim3(grid_for(nv)), dim3(kT), 0, s, nv, vdof, -1);
  hipLaunchKernelGGL(fill_i32, dim3(grid_for(nv)), dim3(kT), 0, s, nv, vk, -1);
  hipLaunchKernelGGL(fill_i32, dim3(grid_for(ndofs)), dim3(kT), 0, s, ndofs, dofv, -1);
  hipLaunchKernelGGL(maps_kernel, dim3(grid_for(4 * nc)), dim3(kT), 0, s, 4 * nc, in.d_cells, in.d_cdofs,
                     in.d_kdofs, vdof, dofv, vk);
  hipLaunchKernelGGL(maps_check_kernel, dim3(grid_for(4 * nc)), dim3(kT), 0, s, 4 * nc, in.d_cells,
                     in.d_cdofs, in.d_kdofs, vdof, dofv, vk, flags);
  int32_t* kcnt = tmp.alloc<int32_t>(std::max<int64_t>(in.nkdofs, 1));
  hip_check(hipMemsetAsync(kcnt, 0, std::max<int64_t>(in.nkdofs, 1) * sizeof(int32_t), s), "hipMemsetAsync");
  hipLaunchKernelGGL(kdof_kernel, dim3(grid_for(nv)), dim3(kT), 0, s, nv, vk, kcnt, flags);
  hip_check(hipGetLastError(), "plan maps");
  {
    const std::vector<int32_t> f = download(flags, F_NFLAGS, s);
    if (f[F_DOFMAP]) {
      err = "P1 dofmap is not a function of the vertex";
      return kDevicePlanError;
    }
    if (f[F_COEFF]) {
      err = "P1 coefficient dofmap is not a function of the vertex";
      return kDevicePlanError;
    }
    t.kdof_is_vertex = !f[F_KNOTID];
    stats.kdof_injective = !f[F_KNOTINJ];
  }
  // 2. RCB over the present dofs
  int32_t* pflag = tmp.alloc<int32_t>(ndofs);
  int32_t* ppos = tmp.alloc<int32_t>(ndofs);
  hipLaunchKernelGGL(present_kernel, dim3(grid_for(ndofs)), dim3(kT), 0, s, ndofs, dofv, pflag);
  exclusive_scan_i32(tmp, pflag, ppos, ndofs, s);
  int64_t n = 0;
  {
    int32_t last_pos = 0, last_flag = 0;
    hip_check(hipMemcpyAsync(&last_pos, ppos + ndofs - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&last_flag, pflag + ndofs - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    n = static_cast<int64_t>(last_pos) + last_flag;
  }
  if (n == 0) return kDevicePlanUnsupported;
  int32_t* ids = tmp.alloc<int32_t>(n);
  hipLaunchKernelGGL(compact_ids_kernel, dim3(grid_for(ndofs)), dim3(kT), 0, s, ndofs, dofv, ppos, ids);
  double* dx = tmp.alloc<double>(nv * gdim);
  hip_check(hipMemcpyAsync(dx, in.h_x, nv * gdim * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  double* xs = tmp.alloc<double>(n * gdim);
  hipLaunchKernelGGL(xs_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, static_cast<int32_t>(gdim), ids, dofv,
                     dx, xs);
  uint64_t* ka = tmp.alloc<uint64_t>(std::max<int64_t>(n, 1));
  uint64_t* kb = tmp.alloc<uint64_t>(std::max<int64_t>(n, 1));
  int32_t* va = tmp.alloc<int32_t>(n);
  int32_t* vb = tmp.alloc<int32_t>(n);
  int32_t* rank[3] = {nullptr, nullptr, nullptr};
  for (int ax = 0; ax < gdim; ++ax) {
    rank[ax] = tmp.alloc<int32_t>(n);
    hipLaunchKernelGGL(axis_keys_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, static_cast<int32_t>(gdim), ax,
                       xs, ka, va);
    sort_pairs(tmp, ka, kb, va, vb, n, 64, s);
    hipLaunchKernelGGL(scatter_rank_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, va, rank[ax]);
  }
  int32_t* perm = va;  // the current buffer of the pair sort
  int32_t* perm_alt = vb;
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, perm);
  int32_t* sid = tmp.alloc<int32_t>(n);
  hip_check(hipMemsetAsync(sid, 0, n * sizeof(int32_t), s), "hipMemsetAsync");
  std::vector<int32_t> seg_lo = {0, static_cast<int32_t>(n)};
  const int rb = bitlen(static_cast<uint64_t>(n - 1));
  const int64_t max_segs = 4 * ((n + target_rows - 1) / target_rows) + 16;
  int8_t* d_axis = tmp.alloc<int8_t>(max_segs);
  int32_t* d_lo = tmp.alloc<int32_t>(max_segs + 1);
  int32_t* d_mid = tmp.alloc<int32_t>(max_segs);
  int32_t* d_newbase = tmp.alloc<int32_t>(max_segs);
  unsigned long long* d_bmin = tmp.alloc<unsigned long long>(3 * max_segs);
  unsigned long long* d_bmax = tmp.alloc<unsigned long long>(3 * max_segs);
  for (;;) {
    const int64_t nseg = static_cast<int64_t>(seg_lo.size()) - 1;
    std::vector<int8_t> axis(nseg, -1);
    bool any = false;
    for (int64_t q = 0; q < nseg; ++q)
      if (seg_lo[q + 1] - seg_lo[q] > target_rows) {
        axis[q] = 0;
        any = true;
      }
    if (!any) break;
    if (nseg > max_segs) throw std::runtime_error("device plan: segment table overflow");
    upload_to(d_axis, axis, s);
    hip_check(hipMemsetAsync(d_bmin, 0xff, 3 * nseg * sizeof(unsigned long long), s), "hipMemsetAsync");
    hip_check(hipMemsetAsync(d_bmax, 0, 3 * nseg * sizeof(unsigned long long), s), "hipMemsetAsync");
    hipLaunchKernelGGL(bbox_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, static_cast<int32_t>(gdim), sid, perm,
                       xs, d_axis, d_bmin, d_bmax);
    hip_check(hipGetLastError(), "bbox_kernel");
    const std::vector<unsigned long long> bmin = download(d_bmin, 3 * nseg, s);
    const std::vector<unsigned long long> bmax = download(d_bmax, 3 * nseg, s);
    std::vector<int32_t> mid(nseg, 0), newbase(nseg, 0), nlo;
    nlo.reserve(2 * nseg + 1);
    for (int64_t q = 0; q < nseg; ++q) {
      const int64_t lo = seg_lo[q], hi = seg_lo[q + 1];
      newbase[q] = static_cast<int32_t>(nlo.size());
      nlo.push_back(static_cast<int32_t>(lo));
      if (axis[q] < 0) continue;
      int ax = 0;
      double ext[3];
      for (int k = 0; k < gdim; ++k) ext[k] = dkey_inv(bmax[3 * q + k]) - dkey_inv(bmin[3 * q + k]);
      for (int k = 1; k < gdim; ++k)
        if (ext[k] > ext[ax]) ax = k;
      axis[q] = static_cast<int8_t>(ax);
      const int64_t nleaf = (hi - lo + target_rows - 1) / target_rows;
      const int64_t m = lo + (hi - lo) * (nleaf / 2) / nleaf;
      mid[q] = static_cast<int32_t>(m);
      nlo.push_back(static_cast<int32_t>(m));
    }
    nlo.push_back(static_cast<int32_t>(n));
    upload_to(d_axis, axis, s);
    std::vector<int32_t> lo32(seg_lo.begin(), seg_lo.end());
    upload_to(d_lo, lo32, s);
    upload_to(d_mid, mid, s);
    upload_to(d_newbase, newbase, s);
    const int sb = bitlen(static_cast<uint64_t>(std::max<int64_t>(nseg - 1, 1)));
    if (sb + rb > 64) throw std::runtime_error("device plan: key overflow");
    hipLaunchKernelGGL(rcb_keys_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, rb, sid, perm, d_lo, d_axis,
                       rank[0], rank[1], gdim > 2 ? rank[2] : rank[0], ka);
    sort_pairs(tmp, ka, kb, perm, perm_alt, n, sb + rb, s);
    hipLaunchKernelGGL(rcb_sid_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, sid, d_axis, d_mid, d_newbase);
    hip_check(hipGetLastError(), "rcb level");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // the host tables die here
    seg_lo.swap(nlo);
  }
  const int64_t nb = static_cast<int64_t>(seg_lo.size()) - 1;
  if (nb > 65535 * 256) return kDevicePlanUnsupported;
  int32_t* owner = tmp.alloc<int32_t>(ndofs);
  hipLaunchKernelGGL(fill_i32, dim3(grid_for(ndofs)), dim3(kT), 0, s, ndofs, owner, -1);
  int32_t* rowdof_u = tmp.alloc<int32_t>(n);
  hipLaunchKernelGGL(owner_kernel, dim3(grid_for(n)), dim3(kT), 0, s, n, perm, ids, sid, owner, rowdof_u);
  int32_t* row_ptr_b = arena_alloc<int32_t>(plan, nb + 1);
  upload_to(row_ptr_b, seg_lo, s);
  int32_t max_rows = 0;
  for (int64_t q = 0; q < nb; ++q) max_rows = std::max(max_rows, seg_lo[q + 1] - seg_lo[q]);
  // 3. cell appearances
  int32_t* ccnt = tmp.alloc<int32_t>(nc);
  int32_t* coff = tmp.alloc<int32_t>(nc);
  hipLaunchKernelGGL(app_count_kernel, dim3(grid_for(nc)), dim3(kT), 0, s, nc, in.d_cdofs, owner, ccnt);
  exclusive_scan_i32(tmp, ccnt, coff, nc, s);
  int64_t napp = 0;
  {
    int32_t lp = 0, lc = 0;
    hip_check(hipMemcpyAsync(&lp, coff + nc - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&lc, ccnt + nc - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    napp = static_cast<int64_t>(lp) + lc;
  }
  const int cbits = bitlen(static_cast<uint64_t>(nc - 1));
  const int bbits = bitlen(static_cast<uint64_t>(nb - 1));
  if (cbits + bbits > 64) return kDevicePlanUnsupported;
  uint64_t* aka = tmp.alloc<uint64_t>(napp);
  uint64_t* akb = tmp.alloc<uint64_t>(napp);
  hipLaunchKernelGGL(app_fill_kernel, dim3(grid_for(nc)), dim3(kT), 0, s, nc, in.d_cdofs, owner, coff, cbits, aka);
  sort_keys(tmp, aka, akb, napp, cbits + bbits, s);
  int32_t* cell_ptr = arena_alloc<int32_t>(plan, nb + 1);
  int32_t* app = tmp.alloc<int32_t>(napp);
  hipLaunchKernelGGL(app_bounds_kernel, dim3(grid_for(napp + 1)), dim3(kT), 0, s, napp, nb, cbits, aka, cell_ptr, app);
  hip_check(hipGetLastError(), "plan appearances");
  const std::vector<int32_t> h_cell_ptr = download(cell_ptr, nb + 1, s);
  int32_t max_cells = 0;
  for (int64_t q = 0; q < nb; ++q) max_cells = std::max(max_cells, h_cell_ptr[q + 1] - h_cell_ptr[q]);
  // 4. per-block tables
  int R = 1, C = 1;
  while (R < max_rows) R <<= 1;
  while (C < max_cells) C <<= 1;
  BlockArgs A{};
  A.nb = static_cast<int32_t>(nb);
  A.cells = in.d_cells;
  A.vdof = vdof;
  A.dofv = dofv;
  A.owner = owner;
  A.row_ptr = in.d_row_ptr;
  A.col = in.d_col;
  A.seg_lo = row_ptr_b;
  A.rowdof_u = rowdof_u;
  A.cell_ptr = cell_ptr;
  A.app = app;
  A.R = R;
  A.C = C;
  A.runs_stride = std::max(max_rows, 1);
  A.wide = in.max_row_len > kMaxRank;
  A.chains = want_chains;
  // LDS: phase 1 [sdof R][nodes][hash][gtmp][lens R][st R]; phase 2 [sdof][nodes][cl C][keys C]
  // [used R]; phase 3 reuses everything below keys (incp, usedc, nbr) and appends inc
  int32_t o = 0;
  A.o_sdof = o;
  o = align16(o + 4 * R);
  A.o_nodes = o;
  o = align16(o + 4 * kMaxBlockNodes);
  const int32_t o_after_nodes = o;
  A.o_hash = o;
  A.o_gtmp = align16(A.o_hash + 4 * kGhostCap);
  A.o_lens = align16(A.o_gtmp + 4 * kMaxBlockNodes);
  A.o_st = align16(A.o_lens + 4 * R);
  const int32_t end1 = align16(A.o_st + 4 * R);
  A.o_cl = o_after_nodes;
  A.o_keys = align16(A.o_cl + 8 * C);
  A.o_used = align16(A.o_keys + 8 * C);
  const int32_t end2 = align16(A.o_used + 8 * R);
  A.o_incp = 0;
  A.o_usedc = align16(A.o_incp + 4 * (kMaxBlockNodes + 1));
  A.o_nbr = align16(A.o_usedc + max_cells);
  int32_t end3 = align16(A.o_nbr + 8 * max_cells);
  if (end3 > A.o_keys) {  // does not fit below keys: everything of phase 3 after keys
    A.o_incp = align16(A.o_keys + 8 * C);
    A.o_usedc = align16(A.o_incp + 4 * (kMaxBlockNodes + 1));
    A.o_nbr = align16(A.o_usedc + max_cells);
    end3 = align16(A.o_nbr + 8 * max_cells);
    A.o_inc = end3;
  } else {
    A.o_inc = align16(A.o_keys + 8 * C);
  }
  end3 = std::max(end3, align16(A.o_inc + 8 * max_cells));
  const int32_t lds_bytes = std::max(end1, std::max(end2, A.chains ? end3 : 0));
  int dev = 0, lds_max = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
  const int static_lds = static_cast<int>(sizeof(int32_t) * (kPT + 16));
  if (lds_bytes + static_lds > lds_max || max_cells > 32767 || C > (1 << 16)) return kDevicePlanUnsupported;
  int32_t* blk_nnz = arena_alloc<int32_t>(plan, nb);
  int32_t* row_dof = arena_alloc<int32_t>(plan, n);
  uint32_t* row_info = arena_alloc<uint32_t>(plan, n);
  uint8_t* row_diag = arena_alloc<uint8_t>(plan, n);
  uint64_t* cell_rec = arena_alloc<uint64_t>(plan, 2 * napp);
  uint64_t* cell_rank_wide = A.wide ? arena_alloc<uint64_t>(plan, 2 * napp) : nullptr;
  A.row_dof = row_dof;
  A.row_info = row_info;
  A.row_diag = row_diag;
  A.blk_nnz = blk_nnz;
  A.node_v = tmp.alloc<int32_t>(nb * kMaxBlockNodes);
  A.nnodes = tmp.alloc<int32_t>(nb);
  A.runs_s = tmp.alloc<uint64_t>(nb * A.runs_stride);
  A.nruns = tmp.alloc<int32_t>(nb);
  A.cell_rec = cell_rec;
  A.cell_rank_wide = cell_rank_wide;
  A.desc = A.chains ? tmp.alloc<uint16_t>(16 * napp) : nullptr;
  A.nchains = tmp.alloc<int32_t>(nb);
  A.classoff = tmp.alloc<int32_t>(8 * nb);
  A.flags = flags;
  hip_check(hipMemsetAsync(A.nchains, 0, nb * sizeof(int32_t), s), "hipMemsetAsync");
  hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(block_plan_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes),
            "hipFuncSetAttribute (block plan)");
  hipLaunchKernelGGL(block_plan_kernel, dim3(static_cast<unsigned>(nb)), dim3(kPT), lds_bytes, s, A);
  hip_check(hipGetLastError(), "block_plan_kernel");
  {
    const std::vector<int32_t> f = download(flags, F_NFLAGS, s);
    if (f[F_NODES]) {
      err = "blocked plan: too many nodes in a block";
      return kDevicePlanError;
    }
    if (f[F_NNZ]) {
      err = "blocked plan: too many nonzeros in a block";
      return kDevicePlanError;
    }
    if (f[F_RANK]) {
      err = "blocked plan: a CSR row has more than 256 entries";
      return kDevicePlanError;
    }
  }
  // 5. pointers, compaction and slots
  const std::vector<int32_t> nnodes = download(A.nnodes, nb, s);
  const std::vector<int32_t> nruns = download(A.nruns, nb, s);
  const std::vector<int32_t> h_nnz = download(blk_nnz, nb, s);
  std::vector<int32_t> node_ptr(nb + 1, 0), run_ptr(nb + 1, 0);
  int32_t max_nodes = 0, max_nnz = 0;
  for (int64_t q = 0; q < nb; ++q) {
    node_ptr[q + 1] = node_ptr[q] + nnodes[q];
    run_ptr[q + 1] = run_ptr[q] + nruns[q];
    max_nodes = std::max(max_nodes, nnodes[q]);
    max_nnz = std::max(max_nnz, h_nnz[q]);
  }
  int32_t* d_node_ptr = arena_alloc<int32_t>(plan, nb + 1);
  int32_t* d_run_ptr = arena_alloc<int32_t>(plan, nb + 1);
  upload_to(d_node_ptr, node_ptr, s);
  upload_to(d_run_ptr, run_ptr, s);
  const int64_t tot_nodes = node_ptr[nb];
  int32_t* node_vertex = arena_alloc<int32_t>(plan, tot_nodes);
  int32_t* node_kdof = arena_alloc<int32_t>(plan, tot_nodes);
  int32_t* node_dof = arena_alloc<int32_t>(plan, tot_nodes);
  uint64_t* runs = arena_alloc<uint64_t>(plan, run_ptr[nb]);
  int32_t* d_chain_ptr = nullptr;
  uint32_t* chain_rec = nullptr;
  int32_t max_chains = 0;
  std::vector<int32_t> chain_ptr;  // alive until the final synchronisation
  if (A.chains) {
    const std::vector<int32_t> nch = download(A.nchains, nb, s);
    chain_ptr.assign(nb + 1, 0);
    for (int64_t q = 0; q < nb; ++q) {
      chain_ptr[q + 1] = chain_ptr[q] + nch[q];
      max_chains = std::max(max_chains, nch[q]);
    }
    d_chain_ptr = arena_alloc<int32_t>(plan, nb + 1);
    upload_to(d_chain_ptr, chain_ptr, s);
    chain_rec = arena_alloc<uint32_t>(plan, static_cast<int64_t>(chain_ptr[nb]) * kChainWords);
    hipLaunchKernelGGL(chain_records_kernel, dim3(static_cast<unsigned>(nb)), dim3(kT), 0, s,
                       static_cast<int32_t>(nb), cell_ptr, d_chain_ptr, A.classoff, A.desc, cell_rec,
                       chain_rec, flags);
    hip_check(hipGetLastError(), "chain_records_kernel");
  }
  const int32_t SN = max_nodes, SR = std::max<int32_t>(max_rows, 1);
  int4* blk_hdr = nullptr;
  int32_t *slot_vertex = nullptr, *slot_kdof = nullptr, *slot_rdof = nullptr;
  int2* slot_row = nullptr;
  if (A.chains) {
    blk_hdr = arena_alloc<int4>(plan, nb);
    slot_vertex = arena_alloc<int32_t>(plan, nb * SN);
    if (!t.kdof_is_vertex) slot_kdof = arena_alloc<int32_t>(plan, nb * SN);
    slot_row = arena_alloc<int2>(plan, nb * SR);
    slot_rdof = arena_alloc<int32_t>(plan, nb * SR);
  }
  hipLaunchKernelGGL(nodes_out_kernel, dim3(static_cast<unsigned>(nb)), dim3(kT), 0, s, static_cast<int32_t>(nb),
                     SN, A.node_v, d_node_ptr, vk, vdof, node_vertex, node_kdof, node_dof, slot_vertex, slot_kdof);
  hipLaunchKernelGGL(rows_out_kernel, dim3(static_cast<unsigned>(nb)), dim3(kT), 0, s, static_cast<int32_t>(nb),
                     SR, A.runs_stride, row_ptr_b, row_dof, row_diag, in.d_row_ptr, A.runs_s, d_run_ptr,
                     d_chain_ptr, d_node_ptr, runs, blk_hdr, slot_row, slot_rdof);
  hip_check(hipGetLastError(), "plan output");
  {
    const std::vector<int32_t> f = download(flags, F_NFLAGS, s);  // also waits for the kernels
    if (f[F_CHAIN]) {
      err = "edge chains: inconsistent cell records";
      return kDevicePlanError;
    }
  }
  t.num_blocks = static_cast<int32_t>(nb);
  t.max_nodes = max_nodes;
  t.max_rows = max_rows;
  t.max_nnz = max_nnz;
  t.max_row_len = in.max_row_len;
  t.threads = block_threads;
  t.max_cells = max_cells;
  t.node_ptr = d_node_ptr;
  t.node_vertex = node_vertex;
  t.node_kdof = node_kdof;
  t.node_dof = node_dof;
  t.cell_ptr = cell_ptr;
  t.cell_rec = cell_rec;
  t.cell_rank_wide = cell_rank_wide;
  t.row_ptr_b = row_ptr_b;
  t.row_dof = row_dof;
  t.row_info = row_info;
  t.row_diag = row_diag;
  t.blk_nnz = blk_nnz;
  t.run_ptr = d_run_ptr;
  t.runs = runs;
  if (A.chains) {
    t.max_chains = max_chains;
    t.chain_threads = chain_threads;
    t.chain_ptr = d_chain_ptr;
    t.chain_rec = chain_rec;
    t.slot_nodes = SN;
    t.slot_rows = SR;
    t.blk_hdr = blk_hdr;
    t.slot_vertex = slot_vertex;
    t.slot_kdof = slot_kdof;
    t.slot_row = slot_row;
    t.slot_rdof = slot_rdof;
  }
  stats.appearances = napp;
  stats.lds_bytes = lds_bytes;
  stats.chains_built = A.chains;
  return kDevicePlanOk;
}

}  // namespace dfa
