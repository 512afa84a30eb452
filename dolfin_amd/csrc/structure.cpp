// Host-side structure phase: CSR sparsity and the blocked-scatter work partition.
// Runs once per (mesh, dofmap), like the first call of AssemblerBase::init_global_tensor
// (dolfin/fem/AssemblerBase.cpp:137-285).
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <exception>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "common.h"

namespace dfa {

namespace {

int num_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return static_cast<int>(std::min(n, 16u));
}

// Run f(begin, end) over [0, n) split into contiguous chunks, one per thread.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t, int)>& f) {
  const int T = static_cast<int>(std::min<int64_t>(num_threads(), std::max<int64_t>(n, 1)));
  if (T <= 1) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t) {
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    th.emplace_back(f, b, e, t);
  }
  for (auto& x : th) x.join();
}

}  // namespace

// Row r's columns are the sorted unique dofs sharing a cell with r
// (SparsityPatternBuilder.cpp:95-111 -> SparsityPattern::insert_entries, sorted export in
// SparsityPattern::diagonal_pattern, SparsityPattern.cpp:511-542), plus (r, r) when the
// diagonal is kept (SparsityPatternBuilder.cpp:245-264). Built row-parallel from a
// dof -> cell adjacency with a per-thread marker array instead of dolfin::Set inserts.
void build_csr(int64_t num_cells, int32_t nd, const int32_t* cd, int64_t num_rows,
               bool keep_diagonal, HostCSR& out) {
  std::vector<int64_t> aptr(num_rows + 1, 0);
  for (int64_t i = 0; i < num_cells * nd; ++i) {
    const int32_t d = cd[i];
    if (d < 0 || d >= num_rows) throw std::runtime_error("cell_dofs entry out of range");
    aptr[d + 1]++;
  }
  for (int64_t r = 0; r < num_rows; ++r) aptr[r + 1] += aptr[r];
  std::vector<int32_t> adj(aptr[num_rows]);
  {
    std::vector<int64_t> pos(aptr.begin(), aptr.end() - 1);
    for (int64_t c = 0; c < num_cells; ++c)
      for (int j = 0; j < nd; ++j) adj[pos[cd[c * nd + j]]++] = static_cast<int32_t>(c);
  }
  // Rows with the same (non-empty) cell list have the same columns, the union of those
  // cells' dofs (the dofs of one vector-element node, say): rep[r] is the smallest row of r's
  // first cell with r's cell list, and only representatives are built; the rest are copies.
  const int32_t* adjp = adj.data();
  std::vector<int32_t> rep(num_rows);
  parallel_for(num_rows, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      rep[r] = static_cast<int32_t>(r);
      const int64_t n = aptr[r + 1] - aptr[r];
      if (n == 0) continue;
      const int32_t c0 = adjp[aptr[r]];
      for (int j = 0; j < nd; ++j) {
        const int32_t q = cd[static_cast<int64_t>(c0) * nd + j];
        if (q >= rep[r] || aptr[q + 1] - aptr[q] != n) continue;
        if (std::equal(adjp + aptr[q], adjp + aptr[q + 1], adjp + aptr[r])) rep[r] = q;
      }
    }
  });
  std::vector<int32_t> len(num_rows);
  const int T = num_threads();
  std::vector<std::vector<int32_t>> marks(T);
  parallel_for(num_rows, [&](int64_t b, int64_t e, int t) {
    auto& mark = marks[t];
    mark.assign(num_rows, -1);
    for (int64_t r = b; r < e; ++r) {
      if (rep[r] != r) continue;
      int32_t n = 0;
      for (int64_t k = aptr[r]; k < aptr[r + 1]; ++k) {
        const int32_t* d = cd + static_cast<int64_t>(adj[k]) * nd;
        for (int j = 0; j < nd; ++j)
          if (mark[d[j]] != r) {
            mark[d[j]] = static_cast<int32_t>(r);
            ++n;
          }
      }
      if (keep_diagonal && mark[r] != r) ++n;
      len[r] = n;
    }
  });
  parallel_for(num_rows, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) len[r] = len[rep[r]];
  });
  out.row_ptr.assign(num_rows + 1, 0);
  int64_t acc = 0;
  for (int64_t r = 0; r < num_rows; ++r) {
    acc += len[r];
    if (acc > INT32_MAX) throw std::runtime_error("nnz exceeds int32 (la_index) range");
    out.row_ptr[r + 1] = static_cast<int32_t>(acc);
  }
  out.col.resize(acc);
  parallel_for(num_rows, [&](int64_t b, int64_t e, int t) {
    auto& mark = marks[t];
    std::fill(mark.begin(), mark.end(), -1);
    for (int64_t r = b; r < e; ++r) {
      if (rep[r] != r) continue;
      int32_t* dst = out.col.data() + out.row_ptr[r];
      int32_t n = 0;
      for (int64_t k = aptr[r]; k < aptr[r + 1]; ++k) {
        const int32_t* d = cd + static_cast<int64_t>(adj[k]) * nd;
        for (int j = 0; j < nd; ++j)
          if (mark[d[j]] != r) {
            mark[d[j]] = static_cast<int32_t>(r);
            dst[n++] = d[j];
          }
      }
      if (keep_diagonal && mark[r] != r) dst[n++] = static_cast<int32_t>(r);
      std::sort(dst, dst + n);
    }
    std::vector<int32_t>().swap(mark);
  });
  parallel_for(num_rows, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r)
      if (rep[r] != r)
        std::copy(out.col.data() + out.row_ptr[rep[r]], out.col.data() + out.row_ptr[rep[r] + 1],
                  out.col.data() + out.row_ptr[r]);
  });
}

// Blocked-scatter partition for P1 (one dof per vertex).
//  1. vertex <-> dof maps from cell_dofs (must be consistent: P1 dofs live on vertices);
//  2. recursive coordinate bisection of the dofs into blocks of <= target_rows rows;
//  3. every cell is listed once in each distinct block owning one of its dofs;
//  4. per block: local nodes (all vertices of its cells, ascending vertex id), owned rows
//     (ascending dof) with their LDS value offsets, and per cell appearance the local node
//     ids plus the 4-bit rank of column b inside CSR row a for every owned a.
void build_p1_block_plan(int32_t gdim, int64_t nv, int64_t nc, const int32_t* cells,
                         const int32_t* cd, const int32_t* kd, const double* x,
                         const HostCSR& csr, int32_t target_rows, P1BlockPlan& P) {
  const int nvc = gdim + 1;
  const int64_t ndofs = static_cast<int64_t>(csr.row_ptr.size()) - 1;
  P = P1BlockPlan();
  P.nvc = nvc;
  std::vector<int32_t> vdof(nv, -1), dofv(ndofs, -1), vk(nv, -1);
  for (int64_t c = 0; c < nc; ++c)
    for (int a = 0; a < nvc; ++a) {
      const int32_t v = cells[c * nvc + a], d = cd[c * nvc + a], k = kd[c * nvc + a];
      if (v < 0 || v >= nv) throw std::runtime_error("cell vertex out of range");
      if ((vdof[v] >= 0 && vdof[v] != d) || (dofv[d] >= 0 && dofv[d] != v))
        throw std::runtime_error("P1 dofmap is not a function of the vertex");
      if (vk[v] >= 0 && vk[v] != k)
        throw std::runtime_error("P1 coefficient dofmap is not a function of the vertex");
      vdof[v] = d;
      dofv[d] = v;
      vk[v] = k;
    }
  P.kdof_is_vertex = true;
  for (int64_t v = 0; v < nv; ++v)
    if (vk[v] >= 0 && vk[v] != v) P.kdof_is_vertex = false;
  {
    P.kdof_injective = true;
    std::vector<int32_t> kv;
    kv.reserve(nv);
    for (int64_t v = 0; v < nv; ++v)
      if (vk[v] >= 0) kv.push_back(vk[v]);
    std::sort(kv.begin(), kv.end());
    for (size_t i = 1; i < kv.size(); ++i)
      if (kv[i] == kv[i - 1]) P.kdof_injective = false;
  }

  // 2. RCB over the dofs that appear in cells
  std::vector<int32_t> ids;
  ids.reserve(ndofs);
  for (int64_t d = 0; d < ndofs; ++d)
    if (dofv[d] >= 0) ids.push_back(static_cast<int32_t>(d));
  using Leaves = std::vector<std::pair<int64_t, int64_t>>;
  // The two halves of the upper levels are split on threads of their own; each subtree's
  // leaves are appended in depth-first order, so the result equals the serial recursion.
  std::function<void(int64_t, int64_t, int, Leaves&)> rcb = [&](int64_t lo, int64_t hi, int depth,
                                                               Leaves& out) {
    if (hi - lo <= target_rows) {
      if (hi > lo) out.emplace_back(lo, hi);
      return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = lo; i < hi; ++i) {
      const double* p = x + static_cast<int64_t>(dofv[ids[i]]) * gdim;
      for (int k = 0; k < gdim; ++k) {
        mn[k] = std::min(mn[k], p[k]);
        mx[k] = std::max(mx[k], p[k]);
      }
    }
    int ax = 0;
    for (int k = 1; k < gdim; ++k)
      if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
    // split so that the leaves end up as equal as possible with <= target_rows each
    const int64_t nleaf = (hi - lo + target_rows - 1) / target_rows;
    const int64_t mid = lo + (hi - lo) * (nleaf / 2) / nleaf;
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                     [&](int32_t a, int32_t b) {
                       const double pa = x[static_cast<int64_t>(dofv[a]) * gdim + ax];
                       const double pb = x[static_cast<int64_t>(dofv[b]) * gdim + ax];
                       return pa < pb || (pa == pb && a < b);
                     });
    if (depth < 4 && hi - lo > 65536) {
      Leaves left, right;
      std::thread th([&]() { rcb(lo, mid, depth + 1, left); });
      rcb(mid, hi, depth + 1, right);
      th.join();
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
    } else {
      rcb(lo, mid, depth + 1, out);
      rcb(mid, hi, depth + 1, out);
    }
  };
  Leaves leaves;
  rcb(0, static_cast<int64_t>(ids.size()), 0, leaves);
  const int32_t nb = static_cast<int32_t>(leaves.size());
  P.num_blocks = nb;
  std::vector<int32_t> owner(ndofs, -1);
  for (int32_t b = 0; b < nb; ++b)
    for (int64_t i = leaves[b].first; i < leaves[b].second; ++i) owner[ids[i]] = b;

  // 3. cell appearances
  P.cell_ptr.assign(nb + 1, 0);
  auto distinct_blocks = [&](int64_t c, int32_t* bl) {
    int n = 0;
    for (int a = 0; a < nvc; ++a) {
      const int32_t b = owner[cd[c * nvc + a]];
      bool seen = false;
      for (int k = 0; k < n; ++k) seen |= (bl[k] == b);
      if (!seen) bl[n++] = b;
    }
    return n;
  };
  // counted and filled over contiguous cell ranges, one per thread; thread t's cells of a
  // block follow thread t-1's, so each block lists its cells in ascending order
  const int nthr = static_cast<int>(std::min<int64_t>(num_threads(), std::max<int64_t>(nc / 65536, 1)));
  std::vector<std::vector<int32_t>> tcount(nthr, std::vector<int32_t>(nb, 0));
  auto cell_range = [&](int t, int64_t& c0, int64_t& c1) {
    c0 = nc * t / nthr;
    c1 = nc * (t + 1) / nthr;
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nthr; ++t)
      th.emplace_back([&, t]() {
        int64_t c0, c1;
        cell_range(t, c0, c1);
        for (int64_t c = c0; c < c1; ++c) {
          int32_t bl[4];
          const int n = distinct_blocks(c, bl);
          for (int k = 0; k < n; ++k) tcount[t][bl[k]]++;
        }
      });
    for (auto& x : th) x.join();
  }
  for (int32_t b = 0; b < nb; ++b) {
    int32_t s0 = 0;
    for (int t = 0; t < nthr; ++t) s0 += tcount[t][b];
    P.cell_ptr[b + 1] = P.cell_ptr[b] + s0;
  }
  std::vector<int32_t> app(P.cell_ptr[nb]);
  {
    for (int32_t b = 0; b < nb; ++b) {  // tcount becomes each thread's first position per block
      int32_t pos = P.cell_ptr[b];
      for (int t = 0; t < nthr; ++t) {
        const int32_t n = tcount[t][b];
        tcount[t][b] = pos;
        pos += n;
      }
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthr; ++t)
      th.emplace_back([&, t]() {
        int64_t c0, c1;
        cell_range(t, c0, c1);
        auto& pos = tcount[t];
        for (int64_t c = c0; c < c1; ++c) {
          int32_t bl[4];
          const int n = distinct_blocks(c, bl);
          for (int k = 0; k < n; ++k) app[pos[bl[k]]++] = static_cast<int32_t>(c);
        }
      });
    for (auto& x : th) x.join();
  }

  // 4. per-block tables, built in parallel into per-block vectors then concatenated
  std::vector<std::vector<int32_t>> bnodes(nb);
  std::vector<std::vector<int32_t>> brows(nb);
  std::vector<std::vector<uint32_t>> binfo(nb);
  std::vector<std::vector<uint64_t>> bruns(nb);
  std::vector<std::vector<uint8_t>> bdiag(nb);
  P.cell_ids.resize(app.size());
  P.cell_rank.resize(app.size());
  bool wide_ranks = false;  // a row longer than the 4-bit ranks reach
  for (int64_t r = 0; r < ndofs && !wide_ranks; ++r)
    wide_ranks = csr.row_ptr[r + 1] - csr.row_ptr[r] > kMaxRank;
  if (wide_ranks) P.cell_rank_wide.assign(2 * app.size(), 0);
  P.blk_nnz.assign(nb, 0);
  std::atomic<int> failed{0};
  std::string fail_msg;
  parallel_for(nb, [&](int64_t b0, int64_t b1, int) {
    for (int64_t b = b0; b < b1; ++b) {
      // Local node numbering: the owned rows first, in ascending dof order (local node i is
      // owned row i), then the ghost vertices in ascending vertex order.
      std::vector<int32_t> blk_dofs(ids.begin() + leaves[b].first, ids.begin() + leaves[b].second);
      std::sort(blk_dofs.begin(), blk_dofs.end());
      auto& nodes = bnodes[b];
      // vertex -> local node: a per-thread dense map, cleared again when the block is done
      thread_local std::vector<int32_t> vloc;
      if (static_cast<int64_t>(vloc.size()) < nv) vloc.assign(static_cast<size_t>(nv), -1);
      struct Clear {
        std::vector<int32_t>& m;
        const std::vector<int32_t>& n;
        ~Clear() {
          for (int32_t v : n) m[v] = -1;
        }
      } clear_on_exit{vloc, nodes};
      for (int32_t d : blk_dofs) {
        vloc[dofv[d]] = static_cast<int32_t>(nodes.size());
        nodes.push_back(dofv[d]);
      }
      const size_t nown = nodes.size();
      for (int32_t k = P.cell_ptr[b]; k < P.cell_ptr[b + 1]; ++k)
        for (int a = 0; a < nvc; ++a) {
          const int32_t v = cells[static_cast<int64_t>(app[k]) * nvc + a];
          if (vloc[v] == -1) {  // a ghost, first seen
            vloc[v] = -2;
            nodes.push_back(v);
          }
        }
      std::sort(nodes.begin() + static_cast<std::ptrdiff_t>(nown), nodes.end());
      for (size_t i = nown; i < nodes.size(); ++i) vloc[nodes[i]] = static_cast<int32_t>(i);
      auto local = [&](int32_t v) { return static_cast<uint32_t>(vloc[v]); };
      if (nodes.size() > static_cast<size_t>(kMaxBlockNodes)) {
        failed = 1;
        continue;
      }
      // owned rows
      std::vector<int32_t> row_off_of_node(nodes.size(), -1);
      uint32_t off = 0;
      for (int32_t d : blk_dofs) {
        const uint32_t ln = local(dofv[d]);
        {
          const int32_t* rb = csr.col.data() + csr.row_ptr[d];
          const int32_t* re = csr.col.data() + csr.row_ptr[d + 1];
          bdiag[b].push_back(static_cast<uint8_t>(std::min<int64_t>(std::lower_bound(rb, re, d) - rb, 255)));
        }
        brows[b].push_back(d);
        binfo[b].push_back(ln | (off << 16));
        row_off_of_node[ln] = static_cast<int32_t>(off);
        off += static_cast<uint32_t>(csr.row_ptr[d + 1] - csr.row_ptr[d]);
      }
      if (off > static_cast<uint32_t>(kMaxBlockNnz)) {
        failed = 2;
        continue;
      }
      {
        const auto& bd = blk_dofs;
        uint32_t loff = 0;
        size_t i = 0;
        while (i < bd.size()) {
          size_t j = i + 1;
          while (j < bd.size() && bd[j] == bd[j - 1] + 1) ++j;
          const int32_t start = csr.row_ptr[bd[i]];
          const uint32_t len = static_cast<uint32_t>(csr.row_ptr[bd[j - 1] + 1] - start);
          bruns[b].push_back(static_cast<uint64_t>(static_cast<uint32_t>(start)) |
                             (static_cast<uint64_t>(loff) << 32) | (static_cast<uint64_t>(len) << 48));
          loff += len;
          i = j;
        }
      }
      P.blk_nnz[b] = static_cast<int32_t>(off);
      // Order the block's cells by a greedy colouring of "shares an owned row": cells of one
      // colour never add to the same LDS entry, so a wavefront working through one colour
      // issues conflict-free LDS atomics (the order does not change any sum's terms).
      const int32_t kb0 = P.cell_ptr[b], kb1 = P.cell_ptr[b + 1];
      {
        std::vector<uint64_t> used(nodes.size(), 0);
        std::vector<std::pair<int, int32_t>> col(kb1 - kb0);
        for (int32_t k = kb0; k < kb1; ++k) {
          const int64_t c = app[k];
          uint64_t busy = 0;
          uint32_t ln[4];
          for (int a = 0; a < nvc; ++a) {
            ln[a] = local(cells[c * nvc + a]);
            if (row_off_of_node[ln[a]] >= 0) busy |= used[ln[a]];
          }
          int colour = 63;
          for (int q = 0; q < 63; ++q)
            if (!(busy >> q & 1)) { colour = q; break; }
          for (int a = 0; a < nvc; ++a)
            if (row_off_of_node[ln[a]] >= 0) used[ln[a]] |= (1ull << colour);
          // secondary key: the cell's smallest local node, so that the lanes of a wavefront
          // read nearby node records (different LDS banks) within a colour
          uint32_t lmin = ln[0];
          for (int a = 1; a < nvc; ++a) lmin = std::min(lmin, ln[a]);
          col[k - kb0] = {colour * 65536 + static_cast<int>(lmin), static_cast<int32_t>(c)};
        }
        // stable order by (colour, smallest local node): two stable counting sorts
        std::vector<std::pair<int, int32_t>> tmp(col.size());
        auto counting = [](const std::vector<std::pair<int, int32_t>>& in,
                           std::vector<std::pair<int, int32_t>>& out, int range, auto digit) {
          std::vector<int32_t> cnt(static_cast<size_t>(range) + 1, 0);
          for (const auto& e : in) cnt[digit(e) + 1]++;
          for (int i = 0; i < range; ++i) cnt[i + 1] += cnt[i];
          for (const auto& e : in) out[cnt[digit(e)]++] = e;
        };
        counting(col, tmp, static_cast<int>(nodes.size()), [](const std::pair<int, int32_t>& e) { return e.first & 0xffff; });
        counting(tmp, col, 64, [](const std::pair<int, int32_t>& e) { return e.first >> 16; });
        for (int32_t k = kb0; k < kb1; ++k) app[k] = col[k - kb0].second;
      }
      // column ranks: rankof[i * nn + j] = position of local node j's dof in owned row i (every
      // column of an owned row is a vertex of one of the block's cells, so it has a local node)
      const size_t nn = nodes.size();
      thread_local std::vector<uint16_t> rankof;
      if (rankof.size() < nown * nn) rankof.resize(nown * nn);
      for (size_t i = 0; i < nown; ++i) {
        const int32_t d = blk_dofs[i];
        const int32_t r0 = csr.row_ptr[d];
        for (int32_t r = r0; r < csr.row_ptr[d + 1]; ++r)
          rankof[i * nn + static_cast<size_t>(vloc[dofv[csr.col[r]]])] =
              static_cast<uint16_t>(std::min<int32_t>(r - r0, 65535));
      }
      for (int32_t k = P.cell_ptr[b]; k < P.cell_ptr[b + 1]; ++k) {
        const int64_t c = app[k];
        uint64_t idw = 0, rk = 0;
        uint32_t ln[4];
        for (int a = 0; a < nvc; ++a) {
          ln[a] = local(cells[c * nvc + a]);
          idw |= static_cast<uint64_t>(ln[a]) << (16 * a);
        }
        uint64_t wr[2] = {0, 0};
        for (int a = 0; a < nvc; ++a) {
          if (ln[a] >= nown) continue;  // not an owned row
          for (int bb = 0; bb < nvc; ++bb) {
            const uint32_t r = rankof[ln[a] * nn + ln[bb]];
            if (r >= (wide_ranks ? 256u : static_cast<uint32_t>(kMaxRank))) {
              failed = 3;
              continue;
            }
            if (wide_ranks) wr[a >> 1] |= static_cast<uint64_t>(r) << (8 * ((a & 1) * 4 + bb));
            else rk |= static_cast<uint64_t>(r) << (4 * (a * 4 + bb));
          }
        }
        P.cell_ids[k] = idw;
        P.cell_rank[k] = rk;
        if (wide_ranks) {
          P.cell_rank_wide[2 * static_cast<size_t>(k)] = wr[0];
          P.cell_rank_wide[2 * static_cast<size_t>(k) + 1] = wr[1];
        }
      }
    }
  });
  if (failed == 1) throw std::runtime_error("blocked plan: too many nodes in a block");
  if (failed == 2) throw std::runtime_error("blocked plan: too many nonzeros in a block");
  if (failed == 3) throw std::runtime_error("blocked plan: a CSR row has more than 256 entries");

  P.node_ptr.assign(nb + 1, 0);
  P.row_ptr.assign(nb + 1, 0);
  P.run_ptr.assign(nb + 1, 0);
  for (int32_t b = 0; b < nb; ++b) P.run_ptr[b + 1] = P.run_ptr[b] + static_cast<int32_t>(bruns[b].size());
  P.runs.resize(P.run_ptr[nb]);
  for (int32_t b = 0; b < nb; ++b) std::copy(bruns[b].begin(), bruns[b].end(), P.runs.begin() + P.run_ptr[b]);
  for (int32_t b = 0; b < nb; ++b) {
    P.node_ptr[b + 1] = P.node_ptr[b] + static_cast<int32_t>(bnodes[b].size());
    P.row_ptr[b + 1] = P.row_ptr[b] + static_cast<int32_t>(brows[b].size());
    P.max_nodes = std::max<int32_t>(P.max_nodes, static_cast<int32_t>(bnodes[b].size()));
    P.max_rows = std::max<int32_t>(P.max_rows, static_cast<int32_t>(brows[b].size()));
    P.max_nnz = std::max(P.max_nnz, P.blk_nnz[b]);
    P.max_cells = std::max(P.max_cells, P.cell_ptr[b + 1] - P.cell_ptr[b]);
  }
  for (size_t r = 0; r + 1 < csr.row_ptr.size(); ++r)
    P.max_row_len = std::max(P.max_row_len, csr.row_ptr[r + 1] - csr.row_ptr[r]);
  P.node_vertex.resize(P.node_ptr[nb]);
  P.node_kdof.resize(P.node_ptr[nb]);
  P.node_dof.resize(P.node_ptr[nb]);
  P.row_dof.resize(P.row_ptr[nb]);
  P.row_info.resize(P.row_ptr[nb]);
  P.row_diag.resize(P.row_ptr[nb]);
  parallel_for(nb, [&](int64_t b0, int64_t b1, int) {
    for (int64_t b = b0; b < b1; ++b) {
      int32_t o = P.node_ptr[b];
      for (int32_t v : bnodes[b]) {
        P.node_vertex[o] = v;
        P.node_kdof[o] = vk[v];
        P.node_dof[o] = vdof[v];
        ++o;
      }
      std::copy(brows[b].begin(), brows[b].end(), P.row_dof.begin() + P.row_ptr[b]);
      std::copy(binfo[b].begin(), binfo[b].end(), P.row_info.begin() + P.row_ptr[b]);
      std::copy(bdiag[b].begin(), bdiag[b].end(), P.row_diag.begin() + P.row_ptr[b]);
    }
  });
}

namespace {

// Breadth-first sweep from s over the white vertices of its component; returns the number
// of levels - 1 (the eccentricity of s) and in `last_min` the first vertex of smallest
// degree in the last level (the rcm_queue bookkeeping of the Boost implementation).
int bfs_levels(const int32_t* ap, const int32_t* adj, int32_t s, std::vector<int32_t>& mark,
               int32_t stamp, std::vector<int32_t>& queue, int32_t& last_min) {
  auto deg = [&](int32_t v) { return ap[v + 1] - ap[v]; };
  queue.clear();
  queue.push_back(s);
  mark[s] = stamp;
  size_t head = 0;
  int ecc = -1;
  while (head < queue.size()) {
    const size_t level_end = queue.size();
    last_min = queue[head];
    for (; head < level_end; ++head) {
      const int32_t u = queue[head];
      if (deg(u) < deg(last_min)) last_min = u;
      for (int32_t k = ap[u]; k < ap[u + 1]; ++k)
        if (mark[adj[k]] != stamp) {
          mark[adj[k]] = stamp;
          queue.push_back(adj[k]);
        }
    }
    ++ecc;
  }
  return ecc;
}

}  // namespace

void cuthill_mckee_from(int64_t n, const int32_t* ap, const int32_t* adj, int32_t start,
                        std::vector<int32_t>& order) {
  std::vector<char> seen(static_cast<size_t>(n), 0);
  auto deg = [&](int32_t v) { return ap[v + 1] - ap[v]; };
  const size_t first = order.size();
  order.push_back(start);
  seen[start] = 1;
  for (size_t head = first; head < order.size(); ++head) {
    const int32_t u = order[head];
    const size_t begin = order.size();
    for (int32_t k = ap[u]; k < ap[u + 1]; ++k)
      if (!seen[adj[k]]) {
        seen[adj[k]] = 1;
        order.push_back(adj[k]);
      }
    // the same comparison sort the Boost visitor applies to the newly queued vertices
    std::sort(order.begin() + static_cast<std::ptrdiff_t>(begin), order.end(),
              [&](int32_t a, int32_t b) { return deg(a) < deg(b); });
  }
}

void rcm_ordering(int64_t n, const int32_t* ap, const int32_t* adj, int32_t* old_to_new) {
  if (n <= 0) return;
  if (ap[n] == 0) {  // no edges: the identity (BoostGraphOrdering.cpp:50-55)
    for (int64_t v = 0; v < n; ++v) old_to_new[v] = static_cast<int32_t>(v);
    return;
  }
  std::vector<int32_t> mark(static_cast<size_t>(n), -1), queue;
  std::vector<int32_t> comp_of(static_cast<size_t>(n), -1);
  std::vector<int32_t> order;
  order.reserve(static_cast<size_t>(n));
  std::vector<char> done(static_cast<size_t>(n), 0);
  auto deg = [&](int32_t v) { return ap[v + 1] - ap[v]; };
  int32_t stamp = 0;
  for (int64_t r0 = 0; r0 < n; ++r0) {
    if (done[r0]) continue;
    // pseudo-peripheral start of r0's component
    int32_t r = static_cast<int32_t>(r0), x = 0, y = 0;
    int ecc_r = bfs_levels(ap, adj, r, mark, stamp++, queue, x);
    int ecc_x = bfs_levels(ap, adj, x, mark, stamp++, queue, y);
    while (ecc_x > ecc_r) {
      r = x;
      ecc_r = ecc_x;
      x = y;
      ecc_x = bfs_levels(ap, adj, x, mark, stamp++, queue, y);
    }
    // Cuthill-McKee sweep of the component from x
    const size_t first = order.size();
    order.push_back(x);
    done[x] = 1;
    for (size_t head = first; head < order.size(); ++head) {
      const int32_t u = order[head];
      const size_t begin = order.size();
      for (int32_t k = ap[u]; k < ap[u + 1]; ++k)
        if (!done[adj[k]]) {
          done[adj[k]] = 1;
          order.push_back(adj[k]);
        }
      std::sort(order.begin() + static_cast<std::ptrdiff_t>(begin), order.end(),
                [&](int32_t a, int32_t b) { return deg(a) < deg(b); });
    }
  }
  for (int64_t i = 0; i < n; ++i) old_to_new[order[i]] = static_cast<int32_t>(n - 1 - i);
}

std::vector<int32_t> build_spmv_blocks(const std::vector<int32_t>& row_ptr, int32_t budget,
                                       int32_t max_rows) {
  const int64_t n = static_cast<int64_t>(row_ptr.size()) - 1;
  std::vector<int32_t> blocks;
  blocks.reserve(static_cast<size_t>(n / 16 + 2));
  int64_t r = 0;
  while (r < n) {
    blocks.push_back(static_cast<int32_t>(r));
    int64_t e = r + 1;  // the first row always joins, however long
    while (e < n && e - r < max_rows && row_ptr[e + 1] - row_ptr[r] <= budget) ++e;
    r = e;
  }
  blocks.push_back(static_cast<int32_t>(n));
  return blocks;
}

std::vector<int32_t> build_node_chunks(const std::vector<int32_t>& row_ptr, int node_layout, int32_t budget,
                                       int32_t max_nodes) {
  const int64_t nn = (static_cast<int64_t>(row_ptr.size()) - 1) / 3;
  // node entries before node n: component-0 row n holds 3 per entry (component-major), the
  // three rows of node n start at 9 per entry (node-major)
  auto nrp = [&](int64_t n) -> int64_t {
    if (n == nn) return static_cast<int64_t>(row_ptr.back()) / 9;
    return node_layout == 2 ? row_ptr[3 * n] / 9 : row_ptr[n] / 3;
  };
  std::vector<int32_t> chunks;
  int64_t n = 0;
  while (n < nn) {
    if (nrp(n + 1) - nrp(n) > budget) return {};  // a node longer than a chunk: no node path
    chunks.push_back(static_cast<int32_t>(n));
    chunks.push_back(static_cast<int32_t>(nrp(n)));
    int64_t e = n + 1;
    while (e < nn && e - n < max_nodes && nrp(e + 1) - nrp(n) <= budget) ++e;
    n = e;
  }
  chunks.push_back(static_cast<int32_t>(nn));
  chunks.push_back(static_cast<int32_t>(nrp(nn)));
  return chunks;
}

void build_exterior_facets(int32_t gdim, int64_t nc, const int32_t* cells,
                           std::vector<int32_t>& facet_cell, std::vector<int8_t>& facet_local) {
  const int nv = gdim + 1;
  facet_cell.clear();
  facet_local.clear();
  if (nc == 0) return;
  int32_t nvert = 0;
  for (int64_t i = 0; i < nc * nv; ++i) nvert = std::max(nvert, cells[i] + 1);
  // bucket every cell facet by its smallest vertex
  auto facet_key = [&](int64_t c, int f, int32_t (&k)[3]) {
    int n = 0;
    for (int a = 0; a < nv; ++a)
      if (a != f) k[n++] = cells[c * nv + a];
    for (int i = n; i < 3; ++i) k[i] = -1;
    std::sort(k, k + n);
  };
  std::vector<int64_t> start(static_cast<size_t>(nvert) + 1, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < nv; ++f) {
      int32_t k[3];
      facet_key(c, f, k);
      start[k[0] + 1]++;
    }
  for (int32_t v = 0; v < nvert; ++v) start[v + 1] += start[v];
  std::vector<int64_t> item(static_cast<size_t>(nc) * nv);  // c * nv + f
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (int64_t c = 0; c < nc; ++c)
      for (int f = 0; f < nv; ++f) {
        int32_t k[3];
        facet_key(c, f, k);
        item[pos[k[0]]++] = c * nv + f;
      }
  }
  // within a bucket, facets with equal keys are adjacent after sorting by the other vertices
  std::vector<char> exterior(item.size(), 0);
  parallel_for(nvert, [&](int64_t v0, int64_t v1, int) {
    std::vector<std::pair<std::pair<int32_t, int32_t>, int64_t>> bucket;
    for (int64_t v = v0; v < v1; ++v) {
      bucket.clear();
      for (int64_t i = start[v]; i < start[v + 1]; ++i) {
        int32_t k[3];
        facet_key(item[i] / nv, static_cast<int>(item[i] % nv), k);
        bucket.push_back({{k[1], k[2]}, i});
      }
      std::sort(bucket.begin(), bucket.end());
      for (size_t i = 0; i < bucket.size();) {
        size_t j = i + 1;
        while (j < bucket.size() && bucket[j].first == bucket[i].first) ++j;
        if (j - i == 1) exterior[bucket[i].second] = 1;
        i = j;
      }
    }
  });
  std::vector<int64_t> ext;
  for (size_t i = 0; i < item.size(); ++i)
    if (exterior[i]) ext.push_back(item[i]);
  std::sort(ext.begin(), ext.end());
  facet_cell.resize(ext.size());
  facet_local.resize(ext.size());
  for (size_t i = 0; i < ext.size(); ++i) {
    facet_cell[i] = static_cast<int32_t>(ext[i] / nv);
    facet_local[i] = static_cast<int8_t>(ext[i] % nv);
  }
}

// Chains are grown greedily: an unassigned cell seeds one chain per edge, walking across the
// faces that contain the edge in both directions through unassigned cells of the block, and
// the longest (capped at kChainLen) is kept. Seeds follow the block's cell order.
void build_p1_chain_plan(const P1BlockPlan& P, P1ChainPlan& C) {
  if (P.nvc != 4) throw std::runtime_error("edge chains need tetrahedra");
  const int32_t nb = P.num_blocks;
  std::vector<std::vector<uint32_t>> brec(nb);
  std::atomic<int> failed{0};
  parallel_for(nb, [&](int64_t b0, int64_t b1, int) {
    for (int64_t b = b0; b < b1; ++b) {
      const int32_t c0 = P.cell_ptr[b], ncb = P.cell_ptr[b + 1] - c0;
      std::vector<std::array<uint32_t, 4>> cl(ncb);
      for (int32_t k = 0; k < ncb; ++k)
        for (int a = 0; a < 4; ++a) cl[k][a] = static_cast<uint32_t>((P.cell_ids[c0 + k] >> (16 * a)) & 0xffffu);
      // face key = the sorted local ids of its three vertices
      auto fkey = [](uint32_t a, uint32_t b, uint32_t c) {
        if (a > b) std::swap(a, b);
        if (b > c) std::swap(b, c);
        if (a > b) std::swap(a, b);
        return static_cast<uint64_t>(a) << 32 | static_cast<uint64_t>(b) << 16 | c;
      };
      // face neighbours: nbr[4k + o] = the other cell of the block across the face of cell k
      // opposite its vertex o (a face has at most two cells), -1 on the block's boundary,
      // matched through an open-addressing table of face keys (linear in the block's faces)
      std::vector<int32_t> nbr(4 * static_cast<size_t>(ncb), -1);
      {
        size_t cap = 16;
        while (cap < 8 * static_cast<size_t>(ncb)) cap <<= 1;
        std::vector<uint64_t> hkey(cap, ~0ull);
        std::vector<int32_t> hval(cap, -1);
        for (int32_t k = 0; k < ncb; ++k)
          for (int o = 0; o < 4; ++o) {
            uint32_t v[3];
            int n = 0;
            for (int a = 0; a < 4; ++a)
              if (a != o) v[n++] = cl[k][a];
            const uint64_t key = fkey(v[0], v[1], v[2]);
            size_t h = static_cast<size_t>((key * 0x9E3779B97F4A7C15ull) >> 32) & (cap - 1);
            while (hkey[h] != ~0ull && hkey[h] != key) h = (h + 1) & (cap - 1);
            const int32_t me = 4 * k + o;
            if (hkey[h] == key) {
              const int32_t other = hval[h];
              nbr[me] = other >> 2;
              nbr[other] = k;
            } else {
              hkey[h] = key;
              hval[h] = me;
            }
          }
      }
      std::vector<char> used(ncb, 0);
      // the unassigned cell other than k across face (p, q, r) of cell k, or -1
      auto across = [&](int32_t k, uint32_t p, uint32_t q, uint32_t r) {
        for (int o = 0; o < 4; ++o) {
          const uint32_t v = cl[k][o];
          if (v != p && v != q && v != r) {
            const int32_t c = nbr[4 * static_cast<size_t>(k) + o];
            return c >= 0 && !used[c] ? c : int32_t{-1};
          }
        }
        return int32_t{-1};
      };
      auto fourth = [&](int32_t k, uint32_t p, uint32_t q, uint32_t r) {
        for (int a = 0; a < 4; ++a) {
          const uint32_t v = cl[k][a];
          if (v != p && v != q && v != r) return v;
        }
        return uint32_t{0xffffffffu};
      };
      // rank(u -> v) from any cell appearance containing both
      auto rank_in = [&](int32_t k, uint32_t u, uint32_t v) {
        int au = -1, av = -1;
        for (int a = 0; a < 4; ++a) {
          if (cl[k][a] == u) au = a;
          if (cl[k][a] == v) av = a;
        }
        if (au < 0 || av < 0) return uint32_t{0xffu};
        return static_cast<uint32_t>((P.cell_rank[c0 + k] >> (4 * (au * 4 + av))) & 15u);
      };
      // cell i of a chain is (p, q, r[i], r[i + 1]); fixed-size storage, no allocation
      struct Chain {
        uint32_t p = 0, q = 0;
        int m = 0;
        uint32_t r[kChainLen + 1];
        int32_t cells[kChainLen];
      };
      std::vector<Chain> chains;
      chains.reserve(static_cast<size_t>(ncb));
      static const int EA[6] = {0, 0, 0, 1, 1, 2}, EB[6] = {1, 2, 3, 2, 3, 3};
      constexpr int K = kChainLen;
      for (int32_t k0 = 0; k0 < ncb; ++k0) {
        if (used[k0]) continue;
        Chain best;
        for (int e = 0; e < 6; ++e) {
          const uint32_t p = cl[k0][EA[e]], q = cl[k0][EB[e]];
          uint32_t rs[2];
          int n = 0;
          for (int a = 0; a < 4; ++a)
            if (a != EA[e] && a != EB[e]) rs[n++] = cl[k0][a];
          // cells cb[lo, hi), nodes rb[lo, hi]: room for K - 1 cells on either side of k0
          int32_t cb[2 * K];
          uint32_t rb[2 * K + 1];
          int lo = K - 1, hi = K;
          cb[lo] = k0;
          rb[lo] = rs[0];
          rb[hi] = rs[1];
          used[k0] = 1;
          int32_t cur = k0;
          while (hi - lo < K) {  // forward across (p, q, r_last)
            const int32_t c = across(cur, p, q, rb[hi]);
            if (c < 0) break;
            rb[hi + 1] = fourth(c, p, q, rb[hi]);
            cb[hi++] = c;
            used[c] = 1;
            cur = c;
          }
          cur = k0;
          while (hi - lo < K) {  // backward across (p, q, r_first)
            const int32_t c = across(cur, p, q, rb[lo]);
            if (c < 0) break;
            const uint32_t nx = fourth(c, p, q, rb[lo]);
            cb[--lo] = c;
            rb[lo] = nx;
            used[c] = 1;
            cur = c;
          }
          for (int i = lo; i < hi; ++i) used[cb[i]] = 0;
          if (hi - lo > best.m) {
            best.p = p;
            best.q = q;
            best.m = hi - lo;
            for (int i = lo; i < hi; ++i) best.cells[i - lo] = cb[i];
            for (int i = lo; i <= hi; ++i) best.r[i - lo] = rb[i];
          }
          if (best.m == K) break;
        }
        for (int i = 0; i < best.m; ++i) used[best.cells[i]] = 1;
        chains.push_back(best);
      }
      std::stable_sort(chains.begin(), chains.end(),
                       [](const Chain& a, const Chain& b) { return a.m > b.m; });
      auto& out = brec[b];
      out.assign(chains.size() * kChainWords, 0u);
      for (size_t i = 0; i < chains.size(); ++i) {
        const Chain& ch = chains[i];
        const int m = ch.m;
        uint32_t* w = out.data() + i * kChainWords;
        uint32_t ids[9];
        ids[0] = ch.p;
        ids[1] = ch.q;
        for (int k = 0; k < 7; ++k) ids[2 + k] = k <= m ? ch.r[k] : ch.p;
        w[0] = ids[0] | ids[1] << 16;
        w[1] = ids[2] | ids[3] << 16;
        w[2] = ids[4] | ids[5] << 16;
        w[3] = ids[6] | ids[7] << 16;
        w[4] = ids[8] | static_cast<uint32_t>(m) << 16;
        uint8_t bytes[24] = {0};
        auto pair = [&](int j, int32_t cell, uint32_t u, uint32_t v) {
          const uint32_t a = rank_in(cell, u, v), c = rank_in(cell, v, u);
          if (a > 15 || c > 15) failed = 1;
          bytes[j] = static_cast<uint8_t>((a & 15u) | (c & 15u) << 4);
        };
        for (int k = 0; k < m; ++k) {
          pair(3 * k, ch.cells[k], ch.p, ch.r[k]);
          pair(3 * k + 1, ch.cells[k], ch.q, ch.r[k]);
          pair(3 * k + 2, ch.cells[k], ch.r[k], ch.r[k + 1]);
        }
        pair(18, ch.cells[m - 1], ch.p, ch.r[m]);
        pair(19, ch.cells[m - 1], ch.q, ch.r[m]);
        pair(20, ch.cells[0], ch.p, ch.q);
        std::memcpy(w + 5, bytes, 24);
      }
    }
  });
  if (failed) throw std::runtime_error("edge chains: inconsistent cell records");
  C.chain_ptr.assign(nb + 1, 0);
  C.max_chains = 0;
  for (int32_t b = 0; b < nb; ++b) {
    const int32_t n = static_cast<int32_t>(brec[b].size() / kChainWords);
    C.chain_ptr[b + 1] = C.chain_ptr[b] + n;
    C.max_chains = std::max(C.max_chains, n);
  }
  C.chain_rec.resize(static_cast<size_t>(C.chain_ptr[nb]) * kChainWords);
  for (int32_t b = 0; b < nb; ++b)
    std::copy(brec[b].begin(), brec[b].end(), C.chain_rec.begin() + static_cast<size_t>(C.chain_ptr[b]) * kChainWords);
}

}  // namespace dfa

namespace dfa {

// Vector P2 rank table. Nodes are identified by the cell-local node index a (0..9) of the
// component-0 dof; the CSR row of dof (i, a) lists, for each column component j, the
// neighbour nodes in ascending order, either component-major (UFC numbering, stride s1 = 1,
// component stride = deg) or node-major (block-interleaved numbering, s1 = 3, stride 1).
bool build_vp2_ranks(int64_t nc, const int32_t* cd, int64_t ndofs, const HostCSR& csr,
                     std::vector<uint8_t>& rank) {
  if (csr.node_layout) {
    // the CSR was expanded from the node graph, whose layout was checked on every cell: the
    // rank of node b in node a's row is its position in the node CSR row
    const bool node_major = csr.node_layout == 2;
    rank.assign(static_cast<size_t>(nc) * 100, 0);
    std::atomic<bool> failed{false};
    parallel_for(nc, [&](int64_t c0, int64_t c1, int) {
      for (int64_t c = c0; c < c1; ++c) {
        int32_t nd[10];
        for (int a = 0; a < 10; ++a) nd[a] = node_major ? cd[c * 30 + a] / 3 : cd[c * 30 + a];
        for (int a = 0; a < 10; ++a) {
          const int32_t* b0 = csr.node_col.data() + csr.node_row_ptr[nd[a]];
          const int32_t* e0 = csr.node_col.data() + csr.node_row_ptr[nd[a] + 1];
          for (int b = 0; b < 10; ++b) {
            const int32_t* q = std::lower_bound(b0, e0, nd[b]);
            if (q == e0 || *q != nd[b] || q - b0 > 255) {
              failed = true;
              return;
            }
            rank[c * 100 + a * 10 + b] = static_cast<uint8_t>(q - b0);
          }
        }
      }
    });
    if (failed) throw std::runtime_error("vector P2 node graph: column not found or row longer than 256 nodes");
    return !node_major;
  }
  // Node id of a dof: dof of component 0 at the same cell-local node.
  std::vector<int32_t> node0(ndofs, -1), comp(ndofs, -1);
  for (int64_t c = 0; c < nc; ++c)
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < 10; ++a) {
        const int32_t d = cd[c * 30 + i * 10 + a];
        const int32_t n0 = cd[c * 30 + a];
        if ((node0[d] >= 0 && node0[d] != n0) || (comp[d] >= 0 && comp[d] != i))
          throw std::runtime_error("vector P2 dofmap is not node-blocked");
        node0[d] = n0;
        comp[d] = i;
      }
  // Determine the layout from the first row with more than one neighbour.
  int layout = 0;  // 1 = component-major, 2 = node-major
  for (int64_t r = 0; r < ndofs && !layout; ++r) {
    const int32_t b = csr.row_ptr[r], e = csr.row_ptr[r + 1];
    if (e - b < 6) continue;
    layout = (comp[csr.col[b]] == comp[csr.col[b + 1]]) ? 1 : 2;
  }
  if (!layout) layout = 1;

  rank.assign(static_cast<size_t>(nc) * 100, 0);
  std::atomic<bool> failed{false};
  parallel_for(nc, [&](int64_t c0, int64_t c1, int) {
    bool bad = false;
    for (int64_t c = c0; c < c1 && !bad; ++c)
      for (int a = 0; a < 10 && !bad; ++a) {
        const int32_t ra = cd[c * 30 + a];  // component-0 row of node a
        const int32_t rb0 = csr.row_ptr[ra], re0 = csr.row_ptr[ra + 1];
        const int32_t len = re0 - rb0;
        if (len % 3) { bad = true; break; }
        const int32_t deg = len / 3;
        for (int b = 0; b < 10; ++b) {
          // locate column (component 0, node b) in row ra
          const int32_t colb = cd[c * 30 + b];
          const int32_t* p = std::lower_bound(csr.col.data() + rb0, csr.col.data() + re0, colb);
          if (p == csr.col.data() + re0 || *p != colb) { bad = true; break; }
          const int64_t pos = p - (csr.col.data() + rb0);
          const int64_t rk = (layout == 1) ? pos : pos / 3;
          if (rk > 255 || rk >= deg) { bad = true; break; }
          rank[c * 100 + a * 10 + b] = static_cast<uint8_t>(rk);
          // verify all 9 component combinations land on the expected columns
          for (int i = 0; i < 3 && !bad; ++i) {
            const int32_t row = cd[c * 30 + i * 10 + a];
            if (csr.row_ptr[row + 1] - csr.row_ptr[row] != len) { bad = true; break; }
            for (int j = 0; j < 3; ++j) {
              const int64_t slot = csr.row_ptr[row] + ((layout == 1) ? (rk + j * deg) : (rk * 3 + j));
              if (csr.col[slot] != cd[c * 30 + j * 10 + b]) { bad = true; break; }
            }
          }
          if (bad) break;
        }
      }
    if (bad) failed = true;
  });
  const bool bad = failed;
  if (bad) throw std::runtime_error("vector P2 sparsity is not node-blocked (unsupported dofmap)");
  return layout == 1;
}


// Vector P2 blocked partition.
//  1. nodes are identified by their component-0 dof; node coordinates are vertices and edge
//     midpoints; recursive coordinate bisection groups them so that each block's rows
//     (3 rows of 3*deg entries per node) fit in max_block_vals doubles of LDS;
//  2. every (cell, local node a) pair becomes one work item of the block owning node a;
//  3. per block: the cells' vertices (local ids), the owned nodes' rows and LDS offsets, and
//     per work item the 4 local vertex ids, the owned-node index, a, and the 10 ranks of the
//     cell's nodes b among node a's neighbours (the same ranks the atomic kernel uses).
// Task words: w0 = 4 x 16-bit local vertex ids; w1 = owned index (16) | a (8) << 16 |
// ranks b=0..4 (8 bits each) << 24; w2 = ranks b=5..9 (8 bits each).
void build_vp2_block_plan(int64_t nc, const int32_t* cells, const int32_t* cd, const int32_t* kd,
                          const double* x, const HostCSR& csr, const std::vector<uint8_t>& rank,
                          bool component_major, int32_t max_block_vals, VP2BlockPlan& P) {
  static const int edge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  const int64_t ndofs = static_cast<int64_t>(csr.row_ptr.size()) - 1;
  P = VP2BlockPlan();
  P.component_major = component_major;
  // node table (indexed by component-0 dof)
  std::vector<float> nx(3 * ndofs, 0.f);
  std::vector<uint8_t> isnode(ndofs, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int a = 0; a < 10; ++a) {
      const int32_t n = cd[c * 30 + a];
      if (isnode[n]) continue;
      isnode[n] = 1;
      for (int q = 0; q < 3; ++q) {
        double v;
        if (a < 4) v = x[static_cast<int64_t>(cells[c * 4 + a]) * 3 + q];
        else v = 0.5 * (x[static_cast<int64_t>(cells[c * 4 + edge[a - 4][0]]) * 3 + q] +
                        x[static_cast<int64_t>(cells[c * 4 + edge[a - 4][1]]) * 3 + q]);
        nx[3 * n + q] = static_cast<float>(v);
      }
    }
  std::vector<int32_t> ids;
  for (int64_t n = 0; n < ndofs; ++n)
    if (isnode[n]) ids.push_back(static_cast<int32_t>(n));
  auto vals_of = [&](int32_t n) { return 3 * (csr.row_ptr[n + 1] - csr.row_ptr[n]); };
  using Leaves = std::vector<std::pair<int64_t, int64_t>>;
  // upper levels split on threads; leaves appended depth-first as in the serial recursion
  std::function<void(int64_t, int64_t, int, Leaves&)> rcb = [&](int64_t lo, int64_t hi, int depth,
                                                               Leaves& out) {
    int64_t tot = 0;
    for (int64_t i = lo; i < hi; ++i) tot += vals_of(ids[i]);
    if (tot <= max_block_vals || hi - lo == 1) {
      if (tot > max_block_vals) throw std::runtime_error("vector P2 node rows exceed the LDS block");
      out.emplace_back(lo, hi);
      return;
    }
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t i = lo; i < hi; ++i)
      for (int q = 0; q < 3; ++q) {
        mn[q] = std::min(mn[q], nx[3 * ids[i] + q]);
        mx[q] = std::max(mx[q], nx[3 * ids[i] + q]);
      }
    int ax = 0;
    for (int q = 1; q < 3; ++q)
      if (mx[q] - mn[q] > mx[ax] - mn[ax]) ax = q;
    const int64_t nleaf = (tot + max_block_vals - 1) / max_block_vals;
    int64_t mid = lo + (hi - lo) * (nleaf / 2) / nleaf;
    if (mid <= lo) mid = lo + 1;
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                     [&](int32_t a, int32_t b) {
                       const float pa = nx[3 * a + ax], pb = nx[3 * b + ax];
                       return pa < pb || (pa == pb && a < b);
                     });
    if (depth < 4 && hi - lo > 65536) {
      Leaves left, right;
      std::exception_ptr err;
      std::thread th([&]() {
        try {
          rcb(lo, mid, depth + 1, left);
        } catch (...) {
          err = std::current_exception();
        }
      });
      std::exception_ptr err2;
      try {
        rcb(mid, hi, depth + 1, right);
      } catch (...) {
        err2 = std::current_exception();
      }
      th.join();
      if (err) std::rethrow_exception(err);
      if (err2) std::rethrow_exception(err2);
      out.insert(out.end(), left.begin(), left.end());
      out.insert(out.end(), right.begin(), right.end());
    } else {
      rcb(lo, mid, depth + 1, out);
      rcb(mid, hi, depth + 1, out);
    }
  };
  Leaves leaves;
  rcb(0, static_cast<int64_t>(ids.size()), 0, leaves);
  const int32_t nb = static_cast<int32_t>(leaves.size());
  P.num_blocks = nb;
  std::vector<int32_t> owner(ndofs, -1), own_index(ndofs, -1);
  for (int32_t b = 0; b < nb; ++b) {
    std::sort(ids.begin() + leaves[b].first, ids.begin() + leaves[b].second);
    for (int64_t i = leaves[b].first; i < leaves[b].second; ++i) {
      owner[ids[i]] = b;
      own_index[ids[i]] = static_cast<int32_t>(i - leaves[b].first);
    }
  }
  // work items per block (cell order)
  P.task_ptr.assign(nb + 1, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int a = 0; a < 10; ++a) P.task_ptr[owner[cd[c * 30 + a]] + 1]++;
  for (int32_t b = 0; b < nb; ++b) P.task_ptr[b + 1] += P.task_ptr[b];
  const int64_t ntask = P.task_ptr[nb];
  std::vector<int64_t> titem(ntask);  // cell * 10 + a
  {
    std::vector<int32_t> pos(P.task_ptr.begin(), P.task_ptr.end() - 1);
    for (int64_t c = 0; c < nc; ++c)
      for (int a = 0; a < 10; ++a) titem[pos[owner[cd[c * 30 + a]]]++] = c * 10 + a;
  }
  P.task.resize(3 * ntask);
  P.task_cell.resize(ntask);
  P.task_nvert.assign(nb, 0);
  P.blk_vals.assign(nb, 0);
  std::vector<std::vector<int32_t>> bverts(nb);
  std::vector<std::vector<uint32_t>> binfo(nb);
  std::vector<std::vector<int32_t>> brows(nb);
  std::atomic<int> failed{0};
  parallel_for(nb, [&](int64_t b0, int64_t b1, int) {
    std::vector<int64_t> order;
    for (int64_t b = b0; b < b1; ++b) {
      const int32_t t0 = P.task_ptr[b], t1 = P.task_ptr[b + 1];
      auto& verts = bverts[b];
      for (int32_t k = t0; k < t1; ++k) {
        const int64_t c = titem[k] / 10;
        for (int v = 0; v < 4; ++v) verts.push_back(cells[c * 4 + v]);
      }
      std::sort(verts.begin(), verts.end());
      verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
      if (verts.size() > 65535) { failed = 1; continue; }
      // owned nodes: LDS offsets and rows
      uint32_t off = 0;
      for (int64_t i = leaves[b].first; i < leaves[b].second; ++i) {
        const int32_t n = ids[i];
        const int32_t len = csr.row_ptr[n + 1] - csr.row_ptr[n];
        binfo[b].push_back(off | (static_cast<uint32_t>(len / 3) << 16));
        off += 3 * len;
      }
      P.blk_vals[b] = static_cast<int32_t>(off);
      // rows of each owned node: taken from any cell that contains it
      brows[b].assign(3 * (leaves[b].second - leaves[b].first), -1);
      // interleave work items over owned nodes so that a wavefront's lanes add to different rows
      order.assign(titem.begin() + t0, titem.begin() + t1);
      std::vector<int32_t> seen(leaves[b].second - leaves[b].first, 0);
      std::vector<std::pair<int64_t, int64_t>> key(order.size());
      for (size_t q = 0; q < order.size(); ++q) {
        const int64_t c = order[q] / 10;
        const int a = static_cast<int>(order[q] % 10);
        const int32_t o = own_index[cd[c * 30 + a]];
        // vertex nodes first (one gradient factor), then edge nodes; within each kind the
        // items are interleaved over the owned nodes
        key[q] = {(a >= 4 ? (int64_t{1} << 40) : 0) + static_cast<int64_t>(seen[o]++) * 65536 + o, order[q]};
        for (int i = 0; i < 3; ++i) brows[b][3 * o + i] = cd[c * 30 + i * 10 + a];
      }
      std::sort(key.begin(), key.end());
      for (size_t q = 0; q < key.size(); ++q) {
        const int64_t item = key[q].second;
        const int64_t c = item / 10;
        const int a = static_cast<int>(item % 10);
        const int32_t o = own_index[cd[c * 30 + a]];
        uint64_t w0 = 0, w1 = 0, w2 = 0;
        for (int v = 0; v < 4; ++v) {
          const uint64_t lv = static_cast<uint64_t>(
              std::lower_bound(verts.begin(), verts.end(), cells[c * 4 + v]) - verts.begin());
          w0 |= lv << (16 * v);
        }
        w1 = static_cast<uint64_t>(o) | (static_cast<uint64_t>(a) << 16);
        for (int bb = 0; bb < 10; ++bb) {
          const uint64_t r = rank[c * 100 + a * 10 + bb];
          if (bb < 5) w1 |= r << (24 + 8 * bb);
          else w2 |= r << (8 * (bb - 5));
        }
        const int64_t k = t0 + static_cast<int64_t>(q);
        if (a < 4) P.task_nvert[b]++;
        P.task[3 * k] = w0;
        P.task[3 * k + 1] = w1;
        P.task[3 * k + 2] = w2;
        P.task_cell[k] = static_cast<int32_t>(c);
      }
    }
  });
  if (failed) throw std::runtime_error("vector P2 plan: too many vertices in a block");
  P.vert_ptr.assign(nb + 1, 0);
  P.own_ptr.assign(nb + 1, 0);
  for (int32_t b = 0; b < nb; ++b) {
    P.vert_ptr[b + 1] = P.vert_ptr[b] + static_cast<int32_t>(bverts[b].size());
    P.own_ptr[b + 1] = P.own_ptr[b] + static_cast<int32_t>(binfo[b].size());
    P.max_verts = std::max<int32_t>(P.max_verts, static_cast<int32_t>(bverts[b].size()));
    P.max_owned = std::max<int32_t>(P.max_owned, static_cast<int32_t>(binfo[b].size()));
    P.max_vals = std::max(P.max_vals, P.blk_vals[b]);
    P.max_tasks = std::max(P.max_tasks, P.task_ptr[b + 1] - P.task_ptr[b]);
  }
  // vertex -> coefficient dof (must be a function of the vertex)
  std::vector<int32_t> vk;
  {
    int32_t nvmax = 0;
    for (int64_t i = 0; i < nc * 4; ++i) nvmax = std::max(nvmax, cells[i] + 1);
    vk.assign(nvmax, -1);
    for (int64_t i = 0; i < nc * 4; ++i) {
      if (vk[cells[i]] >= 0 && vk[cells[i]] != kd[i])
        throw std::runtime_error("coefficient dofmap is not a function of the vertex");
      vk[cells[i]] = kd[i];
    }
  }
  P.vert_id.resize(P.vert_ptr[nb]);
  P.vert_kdof.resize(P.vert_ptr[nb]);
  P.own_info.resize(P.own_ptr[nb]);
  P.own_rows.resize(3 * static_cast<size_t>(P.own_ptr[nb]));
  for (int32_t b = 0; b < nb; ++b) {
    int32_t o = P.vert_ptr[b];
    for (int32_t v : bverts[b]) {
      P.vert_id[o] = v;
      P.vert_kdof[o] = vk[v];
      ++o;
    }
    std::copy(binfo[b].begin(), binfo[b].end(), P.own_info.begin() + P.own_ptr[b]);
    std::copy(brows[b].begin(), brows[b].end(), P.own_rows.begin() + 3 * static_cast<size_t>(P.own_ptr[b]));
  }
}

}  // namespace dfa

namespace dfa {

// MeshColoring::color for the cell graphs (tdim, dim, tdim) (dolfin/mesh/MeshColoring.cpp:
// 57-121): GraphBuilder::local_graph(mesh, tdim, dim) (dolfin/graph/GraphBuilder.cpp:124-153)
// links two cells through a shared entity of dimension dim, and Boost's
// sequential_vertex_coloring (called from dolfin/graph/BoostGraphColoring.h:87-113) visits
// the cells in index order and gives each the smallest colour not taken by an already
// coloured neighbour. In a simplicial mesh two cells share a vertex, an edge or a facet
// exactly when they share at least 1, 2 or tdim vertices, so the entity graph reduces to
// counting shared vertex ids through a vertex -> cells table. Sequential by definition.
int32_t greedy_cell_coloring(int64_t nc, int32_t nvc, const int32_t* ce, int64_t ne,
                             int32_t min_shared, int32_t* colors) {
  if (nc == 0) return 0;
  if (nvc < 1 || min_shared < 1) throw std::runtime_error("invalid colouring arguments");
  std::vector<int64_t> ptr(static_cast<size_t>(ne) + 1, 0);
  for (int64_t i = 0; i < nc * nvc; ++i) {
    const int32_t e = ce[i];
    if (e < 0 || e >= ne) throw std::runtime_error("cell entity id out of range");
    ++ptr[static_cast<size_t>(e) + 1];
  }
  for (int64_t e = 0; e < ne; ++e) ptr[e + 1] += ptr[e];
  std::vector<int32_t> adj(static_cast<size_t>(ptr[ne]));
  {
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < nc; ++c)  // cells enter each list in ascending order
      for (int32_t a = 0; a < nvc; ++a) adj[fill[ce[c * nvc + a]]++] = static_cast<int32_t>(c);
  }
  std::vector<int32_t> shared(static_cast<size_t>(nc), 0);  // count per candidate, reset after use
  std::vector<int64_t> used;                                 // colour -> stamp (c + 1)
  std::vector<int32_t> cand;
  int32_t ncolors = 0;
  for (int64_t c = 0; c < nc; ++c) {
    cand.clear();
    for (int32_t a = 0; a < nvc; ++a) {
      const int32_t e = ce[c * nvc + a];
      for (int64_t k = ptr[e]; k < ptr[e + 1]; ++k) {
        const int32_t n = adj[k];
        if (n >= c) break;  // lists are ascending: only coloured (lower) cells matter
        if (shared[n]++ == 0) cand.push_back(n);
      }
    }
    for (int32_t n : cand) {
      if (shared[n] >= min_shared) {
        const int32_t col = colors[n];
        if (static_cast<size_t>(col) >= used.size()) used.resize(col + 1, 0);
        used[col] = c + 1;
      }
      shared[n] = 0;
    }
    int32_t k = 0;
    while (static_cast<size_t>(k) < used.size() && used[k] == c + 1) ++k;
    colors[c] = k;
    ncolors = std::max(ncolors, k + 1);
  }
  return ncolors;
}

}  // namespace dfa
