// C++ host program using the header-only adapter (include/dolfin_amd/HipAssembler.hpp) the way
// DOLFIN's own C++ would: build a mesh and dofmap on the host, create the Space once, then
// SystemAssembler / Assembler calls. Checks the reference's known answers:
//   P1 Poisson + DirichletBC(1) on UnitCubeMesh(4): |A|_F = 96.847818767384,
//   |b|_2 = 96.564760289080 (python/test/unit/fem/test_system_assembler.py:73-91);
//   P1 Poisson on UnitIntervalMesh(48): |A|_F = 811.75365721381274, |b| = 1.43583841167606474
//   (python/test/unit/fem/test_assembler.py:39-55);
//   the boundary measure 6.0 of the unit cube through g*v*ds (test_assembler.py:140-153).
// Exit code 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/dolfin_amd/HipAssembler.hpp"

using namespace dolfin_amd;

static void box_mesh(int n, std::vector<double>& x, std::vector<int32_t>& cells) {
  for (int iz = 0; iz <= n; ++iz)
    for (int iy = 0; iy <= n; ++iy)
      for (int ix = 0; ix <= n; ++ix) {
        x.push_back(static_cast<double>(ix) / n);
        x.push_back(static_cast<double>(iy) / n);
        x.push_back(static_cast<double>(iz) / n);
      }
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        const int v0 = iz * (n + 1) * (n + 1) + iy * (n + 1) + ix, v1 = v0 + 1, v2 = v0 + n + 1,
                  v3 = v1 + n + 1, v4 = v0 + (n + 1) * (n + 1), v5 = v1 + (n + 1) * (n + 1),
                  v6 = v2 + (n + 1) * (n + 1), v7 = v3 + (n + 1) * (n + 1);
        const int t[6][4] = {{v0, v1, v3, v7}, {v0, v1, v7, v5}, {v0, v5, v7, v4},
                             {v0, v3, v2, v7}, {v0, v6, v4, v7}, {v0, v2, v6, v7}};
        for (auto& c : t) {
          int s[4] = {c[0], c[1], c[2], c[3]};
          std::sort(s, s + 4);
          cells.insert(cells.end(), s, s + 4);
        }
      }
}

static bool close(double a, double b) { return std::fabs(a - b) < 1e-10; }

static std::vector<double> h0_of(int64_t n) { return std::vector<double>(static_cast<size_t>(n), 0.0); }

int main() {
  bool ok = true;
  {
    std::vector<double> x;
    std::vector<int32_t> cells;
    box_mesh(4, x, cells);
    const int64_t nv = static_cast<int64_t>(x.size()) / 3;
    auto V = std::make_shared<const Space>(3, nv, cells, DFA_ELEMENT_P1, 4, cells, nv, cells, nv, x);
    DeviceArray<double> dx, one, ten, g;
    DeviceArray<uint8_t> mask;
    std::vector<double> h1(nv, 1.0), h10(nv, 10.0), hg(nv, 0.0);
    std::vector<uint8_t> hm(nv, 0);
    for (int64_t v = 0; v < nv; ++v)
      for (int q = 0; q < 3; ++q)
        if (x[3 * v + q] == 0.0 || x[3 * v + q] == 1.0) {
          hm[v] = 1;
          hg[v] = 1.0;
        }
    dx.upload(x.data(), x.size());
    one.upload(h1.data(), h1.size());
    ten.upload(h10.data(), h10.size());
    g.upload(hg.data(), hg.size());
    mask.upload(hm.data(), hm.size());
    Form a{V, DFA_FORM_P1_POISSON, 2, dx.get(), one.get()};
    Form L{V, DFA_FORM_P1_LOAD, 1, dx.get(), ten.get()};
    DeviceCSRMatrix A;
    DeviceVector b;
    HipSystemAssembler sys(a, L, mask.get(), g.get());
    sys.assemble(A, b);
    const double nA = A.norm("frobenius"), nb = b.norm("l2");
    std::printf("UnitCubeMesh(4) P1 + bc: |A| = %.12f |b| = %.12f\n", nA, nb);
    ok &= close(nA, 96.847818767384) && close(nb, 96.564760289080);
    sys.assemble(A);  // re-assembly of A only
    ok &= close(A.norm("frobenius"), 96.847818767384);
    // add_values (AssemblerBase.h:53): a second pass adds instead of overwriting
    sys.add_values = true;
    sys.assemble(A, b);
    sys.add_values = false;
    std::printf("add_values: |A| = %.12f |b| = %.12f\n", A.norm("frobenius"), b.norm("l2"));
    ok &= close(A.norm("frobenius"), 2 * 96.847818767384) && close(b.norm("l2"), 2 * 96.564760289080);
    // incremental assembly (SystemAssembler.h:79,84): x0 equal to the boundary values gives
    // the system of the homogeneous condition (x0 - g = 0 on the boundary)
    DeviceVector x0, bx, bh;
    x0.init(nv);
    x0.add_local(hg.data(), [&] {
      std::vector<int32_t> r(nv);
      for (int64_t v = 0; v < nv; ++v) r[v] = static_cast<int32_t>(v);
      return r;
    }());
    DeviceCSRMatrix Ax, Ah;
    sys.assemble(Ax, bx, x0);
    DeviceArray<double> g0;
    g0.upload(h0_of(nv).data(), nv);
    HipSystemAssembler hom(a, L, mask.get(), g0.get());
    hom.assemble(Ah, bh);
    const std::vector<double> vbx = bx.get_local(), vbh = bh.get_local();
    double db = 0.0, bcmax = 0.0;
    for (int64_t v = 0; v < nv; ++v) {
      db = std::fmax(db, std::fabs(vbx[v] - vbh[v]));
      if (hm[v]) bcmax = std::fmax(bcmax, std::fabs(vbx[v]));
    }
    std::printf("x0 = g: max|b(x0) - b(g=0)| = %.3e, max|b| on the boundary = %.3e, |A| = %.12f\n",
                db, bcmax, Ax.norm("frobenius"));
    ok &= db < 1e-13 && bcmax == 0.0 && close(Ax.norm("frobenius"), 96.847818767384);
    DeviceVector bx2;
    sys.assemble(bx2, x0);  // the vector-only variant
    const std::vector<double> vbx2 = bx2.get_local();
    double db2 = 0.0;
    for (int64_t v = 0; v < nv; ++v) db2 = std::fmax(db2, std::fabs(vbx2[v] - vbx[v]));
    ok &= db2 < 1e-13;
    // GenericTensor::init(TensorLayout), local_range, add_local (host blocks)
    DeviceCSRMatrix M;
    M.init(TensorLayout(V));
    const auto lr = M.local_range(0);
    ok &= lr.first == 0 && lr.second == nv && M.nnz() == V->nnz();
    const double blk[4] = {1.0, 2.0, 3.0, 4.0};
    M.add_local(blk, {0, 1}, {0, 1});
    M.add_local(blk, {0}, {0});
    std::vector<int32_t> rp, cc;
    std::vector<double> mv;
    M.data(rp, cc, mv);
    double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0, rest = 0.0;
    for (int32_t r = 0; r < 2; ++r)
      for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
        if (r == 0 && cc[k] == 0) s00 = mv[k];
        else if (r == 0 && cc[k] == 1) s01 = mv[k];
        else if (r == 1 && cc[k] == 0) s10 = mv[k];
        else if (r == 1 && cc[k] == 1) s11 = mv[k];
        else rest += std::fabs(mv[k]);
      }
    bool refused = false;
    try {
      M.add_local(blk, {0}, {static_cast<int32_t>(nv - 1)});  // (0, last vertex): not a pair of one cell
    } catch (const std::runtime_error&) {
      refused = true;
    }
    std::printf("add_local: %.1f %.1f %.1f %.1f, entry outside the pattern refused: %s\n", s00, s01,
                s10, s11, refused ? "yes" : "no");
    ok &= s00 == 2.0 && s01 == 2.0 && s10 == 3.0 && s11 == 4.0 && rest == 0.0 && refused;
    // 0*v*dx + 1*v*ds: the entries sum to the boundary measure 6.0
    // (python/test/unit/fem/test_assembler.py:140-153)
    DeviceArray<double> zero;
    std::vector<double> h0(nv, 0.0);
    zero.upload(h0.data(), h0.size());
    FacetTerms ds;
    ds.g = one.get();
    HipAssembler fa;
    DeviceVector bs;
    fa.assemble(bs, Form{V, DFA_FORM_P1_LOAD, 1, dx.get(), zero.get()}, ds);
    double area = 0.0;
    for (double v : bs.get_local()) area += v;
    std::printf("UnitCubeMesh(4) 1*v*ds: sum = %.15f\n", area);
    ok &= close(area, 6.0);
  }
  {
    const int n = 48;
    std::vector<double> x;
    std::vector<int32_t> cells;
    for (int i = 0; i <= n; ++i) x.push_back(static_cast<double>(i) / n);
    for (int i = 0; i < n; ++i) {
      cells.push_back(i);
      cells.push_back(i + 1);
    }
    auto V = std::make_shared<const Space>(1, n + 1, cells, DFA_ELEMENT_P1, 2, cells, n + 1, cells, n + 1, x);
    DeviceArray<double> dx, one, ten;
    std::vector<double> h1(n + 1, 1.0), h10(n + 1, 10.0);
    dx.upload(x.data(), x.size());
    one.upload(h1.data(), h1.size());
    ten.upload(h10.data(), h10.size());
    HipAssembler assembler;
    DeviceCSRMatrix A;
    DeviceVector b;
    assembler.assemble(A, Form{V, DFA_FORM_P1_POISSON, 2, dx.get(), one.get()});
    assembler.assemble(b, Form{V, DFA_FORM_P1_LOAD, 1, dx.get(), ten.get()});
    std::printf("UnitIntervalMesh(48) P1: |A| = %.14f |b| = %.17f\n", A.norm("frobenius"), b.norm("l2"));
    ok &= close(A.norm("frobenius"), 811.75365721381274) && close(b.norm("l2"), 1.43583841167606474);
    // stiffness rows sum to zero, so A 1 = 0; the diagonal is 2/h inside and 1/h at the ends
    DeviceVector y, d;
    y.init(n + 1);
    d.init(n + 1);
    A.mult(one.get(), y.values());
    A.diagonal(d.values());
    double ymax = 0.0;
    for (double v : y.get_local()) ymax = std::fmax(ymax, std::fabs(v));
    const std::vector<double> hd = d.get_local();
    bool dok = close(hd[0], n) && close(hd[n], n);
    for (int i = 1; i < n; ++i) dok &= close(hd[i], 2.0 * n);
    std::printf("UnitIntervalMesh(48) P1: max|A 1| = %.3e, diagonal %s\n", ymax, dok ? "ok" : "wrong");
    ok &= ymax < 1e-10 && dok;
    // Jacobi CG on the consistent system A x = A u (A is singular: the constants are its null
    // space, so the check is the residual, not x itself)
    std::vector<double> hu(n + 1);
    for (int i = 0; i <= n; ++i) hu[i] = std::sin(3.0 * i / n);
    DeviceArray<double> du;
    du.upload(hu.data(), hu.size());
    DeviceVector rhs, sol, chk;
    rhs.init(n + 1);
    sol.init(n + 1);
    chk.init(n + 1);
    A.mult(du.get(), rhs.values());
    const int64_t its = A.solve_cg(sol.values(), rhs.values(), true, 1e-12);
    A.mult(sol.values(), chk.values());
    const std::vector<double> hr = rhs.get_local(), hc = chk.get_local();
    double rmax = 0.0, bmax = 0.0;
    for (int i = 0; i <= n; ++i) {
      rmax = std::fmax(rmax, std::fabs(hc[i] - hr[i]));
      bmax = std::fmax(bmax, std::fabs(hr[i]));
    }
    std::printf("UnitIntervalMesh(48) P1: CG %lld iterations, max|A x - b| / max|b| = %.3e\n",
                static_cast<long long>(its), rmax / bmax);
    ok &= its > 0 && its <= n + 1 && rmax / bmax < 1e-9;
  }
  std::printf(ok ? "adapter checks passed\n" : "adapter checks FAILED\n");
  return ok ? 0 : 1;
}
