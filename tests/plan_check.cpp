// CPU check of the blocked-scatter work partition (test helper, host code only).
//
// Builds the CSR and the P1 block plan with the library's own host code and walks the plan
// exactly the way the GPU kernel indexes it -- LDS node table, per-appearance local ids and
// 4-bit ranks, owned-row offsets, write-out -- asserting every index against the sizes the
// kernel's LDS layout reserves. It then accumulates a P1 mass matrix through the plan so the
// test can compare with the oracle: a wrong rank or offset shows up as a wrong value.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "../dolfin_amd/csrc/common.h"

extern "C" int plan_check_p1(int32_t gdim, int64_t nv, int64_t nc, const int32_t* cells,
                             const int32_t* cd, const int32_t* kd, const double* x,
                             int64_t ndofs, int32_t target_rows, int32_t keep_diagonal,
                             double* values, int32_t* row_ptr_out, int64_t nnz_cap,
                             int32_t* stats) {
  try {
    dfa::HostCSR csr;
    dfa::build_csr(nc, gdim + 1, cd, ndofs, keep_diagonal != 0, csr);
    if (static_cast<int64_t>(csr.col.size()) > nnz_cap) return -10;
    dfa::P1BlockPlan P;
    dfa::build_p1_block_plan(gdim, nv, nc, cells, cd, kd, x, csr, target_rows, P);
    const int nvc = gdim + 1;
    const double fact = gdim == 1 ? 1.0 : (gdim == 2 ? 2.0 : 6.0);
    std::vector<int> covered(ndofs, 0);
    int64_t total_app = 0;
    for (int b = 0; b < P.num_blocks; ++b) {
      const int n0 = P.node_ptr[b], nn = P.node_ptr[b + 1] - n0;
      if (nn <= 0 || nn > P.max_nodes) return -1;
      std::vector<int> own(nn, -1);
      const int r0 = P.row_ptr[b], nr = P.row_ptr[b + 1] - r0;
      if (nr <= 0 || nr > P.max_rows) return -2;
      const int nnzb = P.blk_nnz[b];
      if (nnzb > P.max_nnz || nnzb > dfa::kMaxBlockNnz) return -3;
      std::vector<double> vals(nnzb, 0.0);
      for (int i = 0; i < nr; ++i) {
        const uint32_t info = P.row_info[r0 + i];
        const int ln = static_cast<int>(info & 0xffffu), off = static_cast<int>(info >> 16);
        if (ln >= nn || off >= nnzb) return -4;
        if (ln != i) return -12;  // owned rows are local nodes 0..nr-1 in order
        if (P.node_dof[n0 + ln] != P.row_dof[r0 + i]) return -5;
        own[ln] = off;
        covered[P.row_dof[r0 + i]]++;
      }
      const int c0 = P.cell_ptr[b], ncb = P.cell_ptr[b + 1] - c0;
      total_app += ncb;
      for (int k = 0; k < ncb; ++k) {
        const uint64_t ids = P.cell_ids[c0 + k], rk = P.cell_rank[c0 + k];
        int l[4];
        double p[4][3];
        for (int a = 0; a < nvc; ++a) {
          l[a] = static_cast<int>((ids >> (16 * a)) & 0xffffu);
          if (l[a] >= nn) return -6;
          const int32_t v = P.node_vertex[n0 + l[a]];
          for (int q = 0; q < gdim; ++q) p[a][q] = x[static_cast<int64_t>(v) * gdim + q];
        }
        // |det J| for the mass matrix
        double det;
        if (gdim == 1) det = p[1][0] - p[0][0];
        else if (gdim == 2)
          det = (p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) - (p[1][1] - p[0][1]) * (p[2][0] - p[0][0]);
        else {
          double e[3][3];
          for (int j = 0; j < 3; ++j)
            for (int q = 0; q < 3; ++q) e[j][q] = p[j + 1][q] - p[0][q];
          det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
        }
        const double w = std::fabs(det) / (fact * (gdim + 1.0) * (gdim + 2.0));
        for (int a = 0; a < nvc; ++a) {
          if ((own[l[a]] >= 0) != (l[a] < nr)) return -13;  // strided kernel's ownership test
          if (own[l[a]] < 0) continue;
          const int32_t rdof = P.row_dof[r0 + l[a]];
          const int rlen = csr.row_ptr[rdof + 1] - csr.row_ptr[rdof];
          for (int bb = 0; bb < nvc; ++bb) {
            const bool wide = !P.cell_rank_wide.empty();
            const int rank = wide ? static_cast<int>((P.cell_rank_wide[2 * (c0 + k) + (a >> 1)] >>
                                                      (8 * ((a & 1) * 4 + bb))) & 0xffu)
                                  : static_cast<int>((rk >> (4 * (a * 4 + bb))) & 15u);
            if (rank >= rlen || (rlen <= 16 && l[a] * 17 + rank >= nr * 17)) return -14;
            const int slot = own[l[a]] + rank;
            if (slot >= nnzb) return -7;
            vals[slot] += (a == bb) ? 2.0 * w : w;
          }
        }
      }
      // write-out exactly as the kernel does it: by runs of consecutive rows
      int covered_vals = 0;
      for (int q = P.run_ptr[b]; q < P.run_ptr[b + 1]; ++q) {
        const uint64_t run = P.runs[q];
        const int64_t gstart = static_cast<int64_t>(static_cast<uint32_t>(run));
        const int loff = static_cast<int>((run >> 32) & 0xffffu);
        const int len = static_cast<int>(run >> 48);
        if (loff + len > nnzb || gstart + len > static_cast<int64_t>(csr.col.size())) return -8;
        for (int j = 0; j < len; ++j) values[gstart + j] = vals[loff + j];
        covered_vals += len;
      }
      if (covered_vals != nnzb) return -11;
    }
    for (int64_t d = 0; d < ndofs; ++d)
      if (covered[d] != 1) return -9;
    for (size_t i = 0; i < csr.row_ptr.size(); ++i) row_ptr_out[i] = csr.row_ptr[i];
    stats[0] = P.num_blocks;
    stats[1] = P.max_nodes;
    stats[2] = P.max_rows;
    stats[3] = P.max_nnz;
    stats[4] = static_cast<int32_t>(total_app);
    stats[5] = P.max_cells;
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "plan_check: %s\n", e.what());
    return -100;
  }
}

// CSR structure only (for structure parity tests without a GPU).
extern "C" int64_t csr_host(int64_t nc, int32_t nd, const int32_t* cd, int64_t ndofs,
                            int32_t keep_diagonal, int32_t* row_ptr, int32_t* col, int64_t cap) {
  dfa::HostCSR csr;
  dfa::build_csr(nc, nd, cd, ndofs, keep_diagonal != 0, csr);
  const int64_t nnz = static_cast<int64_t>(csr.col.size());
  if (nnz > cap) return -nnz;
  for (size_t i = 0; i < csr.row_ptr.size(); ++i) row_ptr[i] = csr.row_ptr[i];
  for (int64_t i = 0; i < nnz; ++i) col[i] = csr.col[i];
  return nnz;
}

// Vector P2 blocked partition walked like vp2_blocked_kernel: every index is checked against
// the block's LDS sizes and the oracle's element matrices are accumulated through the plan
// (task words, ranks, owned-row offsets, write-out rows).
extern "C" void orc_vp2_elasticity(const double* xc, const double* mu, const double* lmbda, double* A);

extern "C" int plan_check_vp2(int64_t nc, const int32_t* cells, const int32_t* cd, const int32_t* kd,
                              const double* x, int64_t ndofs, const double* mu, const double* lmbda,
                              int32_t block_vals, double* values, int64_t nnz_cap, int32_t* stats) {
  try {
    dfa::HostCSR csr;
    dfa::build_csr(nc, 30, cd, ndofs, false, csr);
    if (static_cast<int64_t>(csr.col.size()) > nnz_cap) return -10;
    std::vector<uint8_t> rank;
    const bool cm = dfa::build_vp2_ranks(nc, cd, ndofs, csr, rank);
    dfa::VP2BlockPlan P;
    dfa::build_vp2_block_plan(nc, cells, cd, kd, x, csr, rank, cm, block_vals, P);
    std::vector<int> seen(nc * 10, 0);
    std::vector<int> rowseen(ndofs, 0);
    for (int b = 0; b < P.num_blocks; ++b) {
      const int v0 = P.vert_ptr[b], nv = P.vert_ptr[b + 1] - v0;
      const int o0 = P.own_ptr[b], no = P.own_ptr[b + 1] - o0;
      const int nvals = P.blk_vals[b];
      if (nv > P.max_verts || no > P.max_owned || nvals > P.max_vals || nvals > block_vals) return -1;
      std::vector<double> vals(nvals, 0.0);
      for (int o = 0; o < no; ++o) {
        const uint32_t info = P.own_info[o0 + o];
        const int off = info & 0xffff, deg = info >> 16;
        if (off + 9 * deg > nvals) return -2;
        for (int i = 0; i < 3; ++i) {
          const int32_t row = P.own_rows[3 * (o0 + o) + i];
          if (row < 0 || row >= ndofs) return -3;
          if (csr.row_ptr[row + 1] - csr.row_ptr[row] != 3 * deg) return -4;
          rowseen[row]++;
        }
      }
      for (int k = P.task_ptr[b]; k < P.task_ptr[b + 1]; ++k) {
        const uint64_t w0 = P.task[3 * k], w1 = P.task[3 * k + 1], w2 = P.task[3 * k + 2];
        const int o = static_cast<int>(w1 & 0xffffu), a = static_cast<int>((w1 >> 16) & 0xffu);
        if (o >= no || a >= 10) return -5;
        // the kernel's segments: the block's first task_nvert items are vertex nodes, the rest edges
        if ((k - P.task_ptr[b] < P.task_nvert[b]) != (a < 4)) return -15;
        const int64_t c = P.task_cell[k];
        seen[c * 10 + a]++;
        double xc[12], m[4], l[4];
        for (int v = 0; v < 4; ++v) {
          const int lv = static_cast<int>((w0 >> (16 * v)) & 0xffffu);
          if (lv >= nv) return -6;
          const int32_t gv = P.vert_id[v0 + lv];
          if (gv != cells[c * 4 + v]) return -7;
          for (int q = 0; q < 3; ++q) xc[3 * v + q] = x[static_cast<int64_t>(gv) * 3 + q];
          m[v] = mu[P.vert_kdof[v0 + lv]];
          l[v] = lmbda[P.vert_kdof[v0 + lv]];
        }
        // the owned node's rows must be the cell's dofs at local node a
        for (int i = 0; i < 3; ++i)
          if (P.own_rows[3 * (o0 + o) + i] != cd[c * 30 + i * 10 + a]) return -8;
        double A[900];
        orc_vp2_elasticity(xc, m, l, A);
        const uint32_t info = P.own_info[o0 + o];
        const int off = info & 0xffff, deg = info >> 16;
        for (int bb = 0; bb < 10; ++bb) {
          const int rk = static_cast<int>(bb < 5 ? (w1 >> (24 + 8 * bb)) & 0xffu : (w2 >> (8 * (bb - 5))) & 0xffu);
          if (rk >= deg) return -9;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              const int slot = off + i * 3 * deg + (cm ? rk + j * deg : 3 * rk + j);
              if (slot >= nvals) return -11;
              vals[slot] += A[(i * 10 + a) * 30 + j * 10 + bb];
            }
        }
      }
      for (int o = 0; o < no; ++o) {
        const uint32_t info = P.own_info[o0 + o];
        const int off = info & 0xffff, deg = info >> 16;
        for (int i = 0; i < 3; ++i) {
          const int32_t row = P.own_rows[3 * (o0 + o) + i];
          for (int j = 0; j < 3 * deg; ++j) values[csr.row_ptr[row] + j] = vals[off + i * 3 * deg + j];
        }
      }
    }
    for (int64_t i = 0; i < nc * 10; ++i)
      if (seen[i] != 1) return -12;
    for (int64_t r = 0; r < ndofs; ++r)
      if (rowseen[r] != 1) return -13;
    stats[0] = P.num_blocks;
    stats[1] = P.max_verts;
    stats[2] = P.max_owned;
    stats[3] = P.max_vals;
    stats[4] = static_cast<int32_t>(P.task_cell.size());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "plan_check_vp2: %s\n", e.what());
    return -100;
  }
}


// Edge-chain walk of the tetrahedral kappa-Poisson kernel: every block's chains are walked as
// a lane does it (sliding window over r_k, contributions to (p, q), (p, r_k), (q, r_k) and b
// carried until complete, one LDS add per owned direction at the recorded ranks, diagonal =
// minus the row sum at write-out). Returns the assembled CSR values and load vector.
extern "C" int plan_check_p1_chains(int64_t nv, int64_t nc, const int32_t* cells,
                                    const int32_t* cd, const double* x, const double* kappa,
                                    const double* f, int64_t ndofs, int32_t target_rows,
                                    double* values, double* bvec, int64_t nnz_cap, int32_t* stats) {
  try {
    dfa::HostCSR csr;
    dfa::build_csr(nc, 4, cd, ndofs, false, csr);
    if (static_cast<int64_t>(csr.col.size()) > nnz_cap) return -10;
    dfa::P1BlockPlan P;
    dfa::build_p1_block_plan(3, nv, nc, cells, cd, cd, x, csr, target_rows, P);
    dfa::P1ChainPlan C;
    dfa::build_p1_chain_plan(P, C);
    if (static_cast<int32_t>(C.chain_ptr.size()) != P.num_blocks + 1) return -1;
    int64_t cells_walked = 0, nchains = 0;
    std::vector<int> covered(ndofs, 0);
    for (int b = 0; b < P.num_blocks; ++b) {
      const int n0 = P.node_ptr[b], nn = P.node_ptr[b + 1] - n0;
      const int r0 = P.row_ptr[b], nr = P.row_ptr[b + 1] - r0;
      if (nr > P.max_rows || nn > P.max_nodes) return -2;
      std::vector<double> sv(static_cast<size_t>(nr) * 17, 0.0), sb(nr, 0.0);
      auto node = [&](uint32_t l, double* pc, double& k, double& fv) {
        if (static_cast<int>(l) >= nn) throw std::runtime_error("chain node out of range");
        const int32_t v = P.node_vertex[n0 + l];
        for (int q = 0; q < 3; ++q) pc[q] = x[static_cast<int64_t>(v) * 3 + q];
        k = kappa[P.node_kdof[n0 + l]];
        fv = f[P.node_kdof[n0 + l]];
      };
      auto add = [&](uint32_t u, uint32_t v, uint32_t byte, double w) {
        if (static_cast<int>(u) < nr) {
          const int rlen = csr.row_ptr[P.row_dof[r0 + u] + 1] - csr.row_ptr[P.row_dof[r0 + u]];
          if (static_cast<int>(byte & 15u) >= rlen) throw std::runtime_error("rank beyond row");
          sv[u * 17 + (byte & 15u)] += w;
        }
        if (static_cast<int>(v) < nr) {
          const int rlen = csr.row_ptr[P.row_dof[r0 + v] + 1] - csr.row_ptr[P.row_dof[r0 + v]];
          if (static_cast<int>(byte >> 4) >= rlen) throw std::runtime_error("rank beyond row");
          sv[v * 17 + (byte >> 4)] += w;
        }
      };
      auto addb = [&](uint32_t u, double w) {
        if (static_cast<int>(u) < nr) sb[u] += w;
      };
      int prev_m = 1 << 30;
      for (int c = C.chain_ptr[b]; c < C.chain_ptr[b + 1]; ++c) {
        const uint32_t* w = C.chain_rec.data() + static_cast<size_t>(c) * dfa::kChainWords;
        const uint32_t lp = w[0] & 0xffffu, lq = w[0] >> 16;
        uint32_t lr[7];
        for (int k = 0; k < 7; ++k) lr[k] = (w[1 + k / 2] >> (16 * (k & 1))) & 0xffffu;
        const int m = static_cast<int>(w[4] >> 16);
        if (m < 1 || m > dfa::kChainLen) return -3;
        if (m > prev_m) return -4;  // longest first
        prev_m = m;
        auto byte = [&](int j) { return (w[5 + j / 4] >> (8 * (j % 4))) & 0xffu; };
        double P_[3], Q_[3], R_[3], N_[3], kp, kq, kr, kn, fp, fq, fr, fn;
        node(lp, P_, kp, fp);
        node(lq, Q_, kq, fq);
        node(lr[0], R_, kr, fr);
        double kpq = 0, bp = 0, bq = 0, cp = 0, cq = 0, cb = 0;
        uint32_t cur = lr[0];
        for (int k = 0; k < m; ++k) {
          node(lr[k + 1], N_, kn, fn);
          const double* pts[4] = {P_, Q_, R_, N_};
          double e[3][3], cf[4][3];
          for (int j = 0; j < 3; ++j)
            for (int q = 0; q < 3; ++q) e[j][q] = pts[j + 1][q] - pts[0][q];
          auto cross = [](const double* a, const double* b2, double* o) {
            o[0] = a[1] * b2[2] - a[2] * b2[1];
            o[1] = a[2] * b2[0] - a[0] * b2[2];
            o[2] = a[0] * b2[1] - a[1] * b2[0];
          };
          cross(e[1], e[2], cf[1]);
          cross(e[2], e[0], cf[2]);
          cross(e[0], e[1], cf[3]);
          for (int q = 0; q < 3; ++q) cf[0][q] = -(cf[1][q] + cf[2][q] + cf[3][q]);
          const double det = e[0][0] * cf[1][0] + e[0][1] * cf[1][1] + e[0][2] * cf[1][2];
          const double vol = std::fabs(det) / 6.0;
          const double kb = 0.25 * (kp + kq + kr + kn);
          double K[4][4];
          for (int a = 0; a < 4; ++a)
            for (int c2 = 0; c2 < 4; ++c2)
              K[a][c2] = kb * vol * (cf[a][0] * cf[c2][0] + cf[a][1] * cf[c2][1] + cf[a][2] * cf[c2][2]) / (det * det);
          const double F = fp + fq + fr + fn;
          const double be[4] = {vol / 20 * (F + fp), vol / 20 * (F + fq), vol / 20 * (F + fr), vol / 20 * (F + fn)};
          kpq += K[0][1];
          bp += be[0];
          bq += be[1];
          cp += K[0][2];
          cq += K[1][2];
          cb += be[2];
          add(lp, lr[k], byte(3 * k), cp);
          add(lq, lr[k], byte(3 * k + 1), cq);
          addb(lr[k], cb);
          add(lr[k], lr[k + 1], byte(3 * k + 2), K[2][3]);
          cp = K[0][3];
          cq = K[1][3];
          cb = be[3];
          for (int q = 0; q < 3; ++q) R_[q] = N_[q];
          kr = kn;
          fr = fn;
          cur = lr[k + 1];
        }
        add(lp, cur, byte(18), cp);
        add(lq, cur, byte(19), cq);
        addb(cur, cb);
        add(lp, lq, byte(20), kpq);
        addb(lp, bp);
        addb(lq, bq);
        cells_walked += m;
        ++nchains;
      }
      if (cells_walked != P.cell_ptr[b + 1]) return -5;  // every appearance in one chain
      for (int i = 0; i < nr; ++i) {
        const int32_t d = P.row_dof[r0 + i];
        covered[d]++;
        const int rb = csr.row_ptr[d], len = csr.row_ptr[d + 1] - rb;
        double rs = 0;
        for (int s2 = 0; s2 < len; ++s2) rs += sv[i * 17 + s2];
        for (int s2 = 0; s2 < len; ++s2) values[rb + s2] = (s2 == P.row_diag[r0 + i]) ? -rs : sv[i * 17 + s2];
        bvec[d] = sb[i];
      }
    }
    for (int64_t d = 0; d < ndofs; ++d)
      if (covered[d] != 1) return -9;
    stats[0] = P.num_blocks;
    stats[1] = C.max_chains;
    stats[2] = static_cast<int32_t>(nchains);
    stats[3] = static_cast<int32_t>(cells_walked);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "plan_check_p1_chains: %s\n", e.what());
    return -100;
  }
}

// The library's exterior-facet list (host code) for comparison with the oracle's.
extern "C" int64_t facets_host(int32_t gdim, int64_t nc, const int32_t* cells, int32_t* fcell,
                               int32_t* flocal) {
  std::vector<int32_t> fc;
  std::vector<int8_t> fl;
  dfa::build_exterior_facets(gdim, nc, cells, fc, fl);
  if (fcell)
    for (size_t i = 0; i < fc.size(); ++i) {
      fcell[i] = fc[i];
      flocal[i] = fl[i];
    }
  return static_cast<int64_t>(fc.size());
}

// The library's SpMV row blocks (host code); returns the number of blocks, fills blocks
// (nblocks + 1 entries) when given.
extern "C" int64_t spmv_blocks_host(const int32_t* row_ptr, int64_t nrows, int32_t budget,
                                    int32_t max_rows, int32_t* blocks) {
  std::vector<int32_t> rp(row_ptr, row_ptr + nrows + 1);
  const std::vector<int32_t> b = dfa::build_spmv_blocks(rp, budget, max_rows);
  if (blocks) std::copy(b.begin(), b.end(), blocks);
  return static_cast<int64_t>(b.size()) - 1;
}

