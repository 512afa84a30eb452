// Reverse Cuthill-McKee numbering on the device: the order rcm_ordering (structure.cpp) and
// BoostGraphOrdering::compute_cuthill_mckee(graph, true) (dolfin/graph/BoostGraphOrdering.cpp:34-75)
// give, computed one breadth-first level at a time.
//
// A breadth-first queue visits the children of a vertex u in the order u appends them: its
// unvisited neighbours in adjacency order (ascending index), sorted by degree in the
// Cuthill-McKee sweep (std::sort over at most 16 items is an insertion sort, so equal degrees
// keep adjacency order). A vertex w first discovered from level L is appended by its
// neighbour of smallest queue position in level L. So level L + 1 in queue order is level
// L + 1 sorted by (smallest queue position of a level-L neighbour, [degree,] index) -- one
// radix sort per level. When some vertex appends more than 16 children at once, the host's
// std::sort is an introsort whose order of equal degrees is not restated here: the device
// build reports it and the caller takes the host build.
//
// Per component, as the host does: the pseudo-peripheral start of George and Liu (breadth-first
// sweeps from r, then from the first vertex of least degree in the last level, while the
// eccentricity grows), then the Cuthill-McKee sweep from it; components in order of their
// smallest vertex; the whole sequence reversed.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "assembler.h"

namespace dfa {
namespace {

constexpr int kRT = 256;
constexpr int kMaxDeviceComponents = 16;

unsigned rgrid(int64_t n) {
  const int64_t b = (n + kRT - 1) / kRT;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 65535)));
}

#define RGRID_LOOP(i, n)                                                                    \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

int rbits(int64_t v) {  // bits to hold values 0..v
  int b = 1;
  while (b < 63 && (int64_t{1} << b) <= v) ++b;
  return b;
}

__global__ void rcm_fill(int64_t n, int32_t* p, int32_t v) {
  RGRID_LOOP(i, n) p[i] = v;
}

// Expand level L (frontier f[0..nf), queue positions pos0 + i): claim unvisited neighbours for
// level L + 1 and keep, per claimed vertex, the smallest queue position of a parent.
__global__ void rcm_expand(int32_t nf, const int32_t* __restrict__ f, int32_t pos0,
                           const int32_t* __restrict__ ap, const int32_t* __restrict__ adj,
                           int32_t next_level, int32_t* __restrict__ lvl, int32_t* __restrict__ par,
                           int32_t* __restrict__ nxt, int32_t* __restrict__ cnt) {
  RGRID_LOOP(i, nf) {
    const int32_t u = f[i];
    const int32_t pu = pos0 + static_cast<int32_t>(i);
    for (int32_t k = ap[u]; k < ap[u + 1]; ++k) {
      const int32_t w = adj[k];
      const int32_t prev = atomicCAS(&lvl[w], -1, next_level);
      if (prev == -1) nxt[atomicAdd(cnt, 1)] = w;
      if (prev == -1 || prev == next_level) atomicMin(&par[w], pu);
    }
  }
}

// keys (parent position, [degree,] vertex) of the claimed vertices
__global__ void rcm_keys(int32_t m, const int32_t* __restrict__ nxt, const int32_t* __restrict__ par,
                         const int32_t* __restrict__ ap, int with_degree, int bw, int bd,
                         uint64_t* __restrict__ keys) {
  RGRID_LOOP(i, m) {
    const int32_t w = nxt[i];
    uint64_t k = static_cast<uint64_t>(par[w]);
    if (with_degree) k = (k << bd) | static_cast<uint64_t>(ap[w + 1] - ap[w]);
    keys[i] = (k << bw) | static_cast<uint64_t>(w);
  }
}

// Sorted keys -> the level in queue order (vertices at positions pos0 + i); flag[0] = 1 when
// a parent has more than 16 children (the position 16 places back has the same parent).
__global__ void rcm_place(int32_t m, const uint64_t* __restrict__ keys, int bw, int pshift,
                          int32_t pos0, int32_t* __restrict__ f, int32_t* __restrict__ order,
                          int32_t* __restrict__ flag) {
  const uint64_t wmask = (uint64_t{1} << bw) - 1;
  RGRID_LOOP(i, m) {
    const int32_t w = static_cast<int32_t>(keys[i] & wmask);
    f[i] = w;
    if (order) order[pos0 + i] = w;
    if (i >= 16 && (keys[i] >> pshift) == (keys[i - 16] >> pshift)) *flag = 1;
  }
}

// The first vertex of least degree in f[0..m) (queue order): atomicMin over (degree, index).
__global__ void rcm_last_min(int32_t m, const int32_t* __restrict__ f, const int32_t* __restrict__ ap,
                             unsigned long long* __restrict__ best) {
  RGRID_LOOP(i, m) {
    const int32_t w = f[i];
    const unsigned long long k =
        (static_cast<unsigned long long>(ap[w + 1] - ap[w]) << 32) | static_cast<unsigned long long>(i);
    atomicMin(best, k);
  }
}

// smallest vertex not yet numbered (INT_MAX if none)
__global__ void rcm_first_open(int64_t n, const int32_t* __restrict__ done, int32_t* __restrict__ out) {
  RGRID_LOOP(v, n) {
    if (!done[v]) atomicMin(out, static_cast<int32_t>(v));
  }
}

__global__ void rcm_mark_done(int32_t m, const int32_t* __restrict__ order, int32_t pos0,
                              int32_t* __restrict__ done) {
  RGRID_LOOP(i, m) done[order[pos0 + i]] = 1;
}

__global__ void rcm_reverse(int64_t n, const int32_t* __restrict__ order, int32_t* __restrict__ old_to_new) {
  RGRID_LOOP(i, n) old_to_new[order[i]] = static_cast<int32_t>(n - 1 - i);
}

struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* get() const {
    return static_cast<T*>(p);
  }
};

struct Sweep {
  int64_t n;
  const int32_t* ap;
  const int32_t* adj;
  int32_t *lvl, *par, *fa, *fb, *cnt, *flag;
  uint64_t *ka, *kb;
  void* tmp;
  size_t tmp_bytes;
  unsigned long long* best;
  int bw, bd;
  hipStream_t s;

  // One breadth-first sweep from `start`; with `order`, the vertices are written there from
  // position pos0 on in Cuthill-McKee order (degree in the key). Returns the number of
  // vertices reached and sets ecc and the first vertex of least degree in the last level.
  // Returns -1 when a vertex appends more than 16 children in the Cuthill-McKee sweep.
  int64_t run(int32_t start, int32_t* order, int32_t pos0, int& ecc, int32_t& last_min) {
    hipLaunchKernelGGL(rcm_fill, dim3(rgrid(n)), dim3(kRT), 0, s, n, lvl, -1);
    hipLaunchKernelGGL(rcm_fill, dim3(rgrid(n)), dim3(kRT), 0, s, n, par, INT_MAX);
    hip_check(hipMemcpyAsync(fa, &start, sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    const int32_t zero = 0;
    hip_check(hipMemcpyAsync(lvl + start, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(flag, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    if (order) hip_check(hipMemcpyAsync(order + pos0, &start, sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    int32_t nf = 1, pos = pos0, level = 0;
    int64_t reached = 1;
    int32_t* f = fa;
    int32_t* g = fb;
    const int with_degree = order ? 1 : 0;
    const int pshift = bw + (with_degree ? bd : 0);
    const int kbits = pshift + bw;  // parent positions < n as well
    for (;;) {
      hip_check(hipMemsetAsync(cnt, 0, sizeof(int32_t), s), "hipMemsetAsync");
      hipLaunchKernelGGL(rcm_expand, dim3(rgrid(nf)), dim3(kRT), 0, s, nf, f, pos, ap, adj, level + 1, lvl,
                         par, g, cnt);
      hip_check(hipGetLastError(), "rcm_expand");
      int32_t m = 0;
      hip_check(hipMemcpyAsync(&m, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      if (m == 0) break;
      hipLaunchKernelGGL(rcm_keys, dim3(rgrid(m)), dim3(kRT), 0, s, m, g, par, ap, with_degree, bw, bd, ka);
      rocprim::double_buffer<uint64_t> keys(ka, kb);
      size_t need = 0;
      hip_check(rocprim::radix_sort_keys(nullptr, need, keys, static_cast<size_t>(m), 0, kbits, s),
                "radix_sort_keys (size)");
      if (need > tmp_bytes) {  // rocPRIM picks its algorithm by size: grow the scratch
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (tmp) hip_check(hipFree(tmp), "hipFree");
        tmp = nullptr;
        hip_check(hipMalloc(&tmp, need), "hipMalloc");
        tmp_bytes = need;
      }
      hip_check(rocprim::radix_sort_keys(tmp, need, keys, static_cast<size_t>(m), 0, kbits, s), "radix_sort_keys");
      pos += nf;
      hipLaunchKernelGGL(rcm_place, dim3(rgrid(m)), dim3(kRT), 0, s, m, keys.current(), bw, pshift, pos, g,
                         order, flag);
      hip_check(hipGetLastError(), "rcm_place");
      std::swap(f, g);
      nf = m;
      reached += m;
      ++level;
    }
    ecc = level;
    const unsigned long long init = ~0ull;
    hip_check(hipMemcpyAsync(best, &init, sizeof(init), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hipLaunchKernelGGL(rcm_last_min, dim3(rgrid(nf)), dim3(kRT), 0, s, nf, f, ap, best);
    unsigned long long b = 0;
    int32_t fl = 0;
    hip_check(hipMemcpyAsync(&b, best, sizeof(b), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&fl, flag, sizeof(fl), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    int32_t lm = 0;
    hip_check(hipMemcpyAsync(&lm, f + static_cast<int32_t>(b & 0xffffffffull), sizeof(int32_t),
                             hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    last_min = lm;
    if (order && fl) return -1;
    return reached;
  }
};

__global__ void rcm_diag_flags(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                               int32_t* __restrict__ has) {
  RGRID_LOOP(v, n) {
    int32_t h = 0;
    for (int32_t k = rp[v]; k < rp[v + 1]; ++k)
      if (col[k] == v) h = 1;
    has[v] = h;
  }
}

// adjacency without the diagonal: row v starts at rp[v] - before[v] (before = exclusive scan
// of the diagonal flags)
__global__ void rcm_drop_diag(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                              const int32_t* __restrict__ before, int32_t* __restrict__ ap,
                              int32_t* __restrict__ adj) {
  RGRID_LOOP(v, n) {
    int32_t o = rp[v] - before[v];
    ap[v] = o;
    for (int32_t k = rp[v]; k < rp[v + 1]; ++k)
      if (col[k] != v) adj[o++] = col[k];
    if (v == n - 1) ap[n] = o;
  }
}

}  // namespace

// The node graph of a cell -> node map (nodes sharing a cell, no self loops) from the device
// CSR structure (rp, col: n rows, sorted): device arrays ap [n + 1] and adj, owned by the
// caller (hipFree); max_degree on the host.
void node_graph_device(int64_t n, const int32_t* rp, const int32_t* col, int32_t** ap, int32_t** adj,
                       int32_t& max_degree, hipStream_t s) {
  Buf has, before, tmp;
  const size_t ni = static_cast<size_t>(n) * sizeof(int32_t);
  hip_check(hipMalloc(&has.p, ni), "hipMalloc");
  hip_check(hipMalloc(&before.p, ni), "hipMalloc");
  hipLaunchKernelGGL(rcm_diag_flags, dim3(rgrid(n)), dim3(kRT), 0, s, n, rp, col, has.get<int32_t>());
  size_t bytes = 0;
  hip_check(rocprim::exclusive_scan(nullptr, bytes, has.get<int32_t>(), before.get<int32_t>(), int32_t{0},
                                    static_cast<size_t>(n), rocprim::plus<int32_t>(), s), "exclusive_scan (size)");
  hip_check(hipMalloc(&tmp.p, std::max<size_t>(bytes, 16)), "hipMalloc");
  hip_check(rocprim::exclusive_scan(tmp.p, bytes, has.get<int32_t>(), before.get<int32_t>(), int32_t{0},
                                    static_cast<size_t>(n), rocprim::plus<int32_t>(), s), "exclusive_scan");
  int32_t nnz = 0, lastb = 0, lasth = 0;
  hip_check(hipMemcpyAsync(&nnz, rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(&lastb, before.get<int32_t>() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(&lasth, has.get<int32_t>() + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  const int64_t m = static_cast<int64_t>(nnz) - lastb - lasth;
  hip_check(hipMalloc(reinterpret_cast<void**>(ap), ni + sizeof(int32_t)), "hipMalloc");
  if (hipMalloc(reinterpret_cast<void**>(adj), std::max<size_t>(static_cast<size_t>(m) * sizeof(int32_t), 16)) != hipSuccess) {
    (void)hipFree(*ap);
    hip_check(hipErrorOutOfMemory, "hipMalloc");
  }
  hipLaunchKernelGGL(rcm_drop_diag, dim3(rgrid(n)), dim3(kRT), 0, s, n, rp, col, before.get<int32_t>(), *ap, *adj);
  hip_check(hipGetLastError(), "rcm_drop_diag");
  std::vector<int32_t> h(static_cast<size_t>(n) + 1);
  hip_check(hipMemcpyAsync(h.data(), *ap, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  max_degree = 0;
  for (int64_t v = 0; v < n; ++v) max_degree = std::max(max_degree, h[v + 1] - h[v]);
}

// Device reverse Cuthill-McKee of the graph (d_ap, d_adj: sorted adjacency lists without self
// loops, on the device); old_to_new on the host. Returns false (nothing written) when the
// device build does not apply: a vertex appends more than 16 children at once, more than 16
// components, or keys wider than 64 bits.
bool rcm_ordering_device(int64_t n, const int32_t* d_ap, const int32_t* d_adj, int32_t max_degree,
                         int32_t* old_to_new, hipStream_t s) {
  if (n <= 0 || n >= INT32_MAX) return false;
  const int bw = rbits(n), bd = rbits(max_degree);
  if (2 * bw + bd > 64) return false;
  Buf lvl, par, fa, fb, cnt, flag, ka, kb, best, order, done, first, o2n;
  const size_t ni = static_cast<size_t>(n) * sizeof(int32_t);
  hip_check(hipMalloc(&lvl.p, ni), "hipMalloc");
  hip_check(hipMalloc(&par.p, ni), "hipMalloc");
  hip_check(hipMalloc(&fa.p, ni), "hipMalloc");
  hip_check(hipMalloc(&fb.p, ni), "hipMalloc");
  hip_check(hipMalloc(&order.p, ni), "hipMalloc");
  hip_check(hipMalloc(&done.p, ni), "hipMalloc");
  hip_check(hipMalloc(&o2n.p, ni), "hipMalloc");
  hip_check(hipMalloc(&ka.p, static_cast<size_t>(n) * sizeof(uint64_t)), "hipMalloc");
  hip_check(hipMalloc(&kb.p, static_cast<size_t>(n) * sizeof(uint64_t)), "hipMalloc");
  hip_check(hipMalloc(&cnt.p, 16), "hipMalloc");
  hip_check(hipMalloc(&flag.p, 16), "hipMalloc");
  hip_check(hipMalloc(&best.p, 16), "hipMalloc");
  hip_check(hipMalloc(&first.p, 16), "hipMalloc");
  Sweep sw{};
  sw.n = n;
  sw.ap = d_ap;
  sw.adj = d_adj;
  sw.lvl = lvl.get<int32_t>();
  sw.par = par.get<int32_t>();
  sw.fa = fa.get<int32_t>();
  sw.fb = fb.get<int32_t>();
  sw.cnt = cnt.get<int32_t>();
  sw.flag = flag.get<int32_t>();
  sw.ka = ka.get<uint64_t>();
  sw.kb = kb.get<uint64_t>();
  sw.best = best.get<unsigned long long>();
  sw.bw = bw;
  sw.bd = bd;
  sw.s = s;
  {
    // scratch for the largest sort of either key width (the size rocPRIM asks for depends on both)
    rocprim::double_buffer<uint64_t> keys(sw.ka, sw.kb);
    size_t need = 0;
    for (int bits : {2 * bw, 2 * bw + bd}) {
      size_t b = 0;
      hip_check(rocprim::radix_sort_keys(nullptr, b, keys, static_cast<size_t>(n), 0, bits, s),
                "radix_sort_keys (size)");
      need = std::max(need, b);
    }
    hip_check(hipMalloc(&sw.tmp, std::max<size_t>(need, 16)), "hipMalloc");
    sw.tmp_bytes = std::max<size_t>(need, 16);
  }
  struct FreeScratch {
    Sweep& w;
    ~FreeScratch() {
      if (w.tmp) (void)hipFree(w.tmp);
    }
  } free_scratch{sw};
  hipLaunchKernelGGL(rcm_fill, dim3(rgrid(n)), dim3(kRT), 0, s, n, done.get<int32_t>(), 0);
  int64_t numbered = 0;
  for (int comp = 0; numbered < n; ++comp) {
    if (comp == kMaxDeviceComponents) return false;
    const int32_t big = INT_MAX;
    hip_check(hipMemcpyAsync(first.p, &big, sizeof(int32_t), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hipLaunchKernelGGL(rcm_first_open, dim3(rgrid(n)), dim3(kRT), 0, s, n, done.get<int32_t>(), first.get<int32_t>());
    int32_t r0 = 0;
    hip_check(hipMemcpyAsync(&r0, first.p, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    // pseudo-peripheral start (George and Liu), as the host's loop
    int ecc_r = 0, ecc_x = 0;
    int32_t x = 0, y = 0;
    if (sw.run(r0, nullptr, 0, ecc_r, x) < 0 || sw.run(x, nullptr, 0, ecc_x, y) < 0) return false;
    int32_t r = r0;
    while (ecc_x > ecc_r) {
      r = x;
      ecc_r = ecc_x;
      x = y;
      if (sw.run(x, nullptr, 0, ecc_x, y) < 0) return false;
    }
    (void)r;
    int e = 0;
    int32_t lm = 0;
    const int64_t got = sw.run(x, order.get<int32_t>(), static_cast<int32_t>(numbered), e, lm);
    if (got < 0) return false;
    hipLaunchKernelGGL(rcm_mark_done, dim3(rgrid(got)), dim3(kRT), 0, s, static_cast<int32_t>(got),
                       order.get<int32_t>(), static_cast<int32_t>(numbered), done.get<int32_t>());
    numbered += got;
  }
  hipLaunchKernelGGL(rcm_reverse, dim3(rgrid(n)), dim3(kRT), 0, s, n, order.get<int32_t>(), o2n.get<int32_t>());
  hip_check(hipMemcpyAsync(old_to_new, o2n.p, ni, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return true;
}

}  // namespace dfa
