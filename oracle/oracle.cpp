// CPU oracle for DOLFIN 2019.1 cell-integral assembly -- TEST INFRASTRUCTURE ONLY.
// Serial restatement of the reference algorithm; see oracle.h for what pins each part.
// All "file:line" citations are into the DOLFIN 2019.1.0 source tree.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <tuple>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------
// Geometry of an affine simplex of dimension d embedded in R^d (tdim == gdim).
// J[i][j] = x_{j+1,i} - x_{0,i}; D[k][p] = d(lambda_k)/dx_p.  UFC convention: the
// coordinate_dofs array is vertex-major, gdim-minor (dolfin/mesh/Cell.h:293-307).
// ---------------------------------------------------------------------------------------
struct Simplex {
  int d = 0;
  double det = 0.0;     // signed det J
  double vol = 0.0;     // |det J| / d!
  double D[4][3] = {};  // barycentric gradients
};

Simplex simplex_geometry(int d, const double* xc) {
  Simplex s;
  s.d = d;
  double J[3][3] = {};
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) J[i][j] = xc[(j + 1) * d + i] - xc[i];
  double K[3][3] = {};
  if (d == 1) {
    s.det = J[0][0];
    K[0][0] = 1.0 / J[0][0];
  } else if (d == 2) {
    s.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    K[0][0] = J[1][1] / s.det;
    K[0][1] = -J[0][1] / s.det;
    K[1][0] = -J[1][0] / s.det;
    K[1][1] = J[0][0] / s.det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    s.det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    // K = J^{-1} = adj(J)/det, adj(J)[j][i] = cofactor(i,j)
    K[0][0] = c00 / s.det;
    K[1][0] = c01 / s.det;
    K[2][0] = c02 / s.det;
    K[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / s.det;
    K[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / s.det;
    K[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / s.det;
    K[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / s.det;
    K[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / s.det;
    K[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / s.det;
  }
  double fact = 1.0;
  for (int i = 2; i <= d; ++i) fact *= i;
  s.vol = std::fabs(s.det) / fact;
  for (int p = 0; p < d; ++p) {
    double sum = 0.0;
    for (int k = 1; k <= d; ++k) {
      s.D[k][p] = K[k - 1][p];
      sum += K[k - 1][p];
    }
    s.D[0][p] = -sum;
  }
  return s;
}

// Exact integral of lambda_n lambda_m lambda_q over a tet divided by its volume:
// int lambda^alpha = |T| d! alpha! / (|alpha| + d)!  with d = 3, |alpha| = 3.
double bary3(int n, int m, int q) {
  if (n == m && m == q) return 6.0 * 6.0 / 720.0;           // 1/20
  if (n == m || m == q || n == q) return 6.0 * 2.0 / 720.0;  // 1/60
  return 6.0 / 720.0;                                         // 1/120
}

// UFC local edge -> local vertices of a tet (dolfin/mesh/TetrahedronCell.cpp:119-124).
const int kTetEdge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

// P2 gradient coefficients: grad phi_a = sum_m lambda_m Q[a][m][:], homogeneous in lambda
// (uses sum_m lambda_m = 1). Vertex i: (4 lambda_i - 1) grad lambda_i;
// edge (i,j): 4 lambda_j grad lambda_i + 4 lambda_i grad lambda_j.
void p2_grad_coeffs(const Simplex& s, double Q[10][4][3]) {
  std::memset(Q, 0, sizeof(double) * 10 * 4 * 3);
  for (int i = 0; i < 4; ++i)
    for (int m = 0; m < 4; ++m)
      for (int p = 0; p < 3; ++p) Q[i][m][p] = (4.0 * (m == i) - 1.0) * s.D[i][p];
  for (int e = 0; e < 6; ++e) {
    const int i = kTetEdge[e][0], j = kTetEdge[e][1];
    for (int p = 0; p < 3; ++p) {
      Q[4 + e][j][p] += 4.0 * s.D[i][p];
      Q[4 + e][i][p] += 4.0 * s.D[j][p];
    }
  }
}

// W[m][q] = sum_n c_n int(lambda_n lambda_m lambda_q)/|T|
void weight3(const double* c, double W[4][4]) {
  for (int m = 0; m < 4; ++m)
    for (int q = 0; q < 4; ++q) {
      double w = 0.0;
      for (int n = 0; n < 4; ++n) w += c[n] * bary3(n, m, q);
      W[m][q] = w;
    }
}

// ---------------------------------------------------------------------------------------
// Sparsity: row sets with dolfin::Set semantics (linear find + push_back,
// dolfin/common/Set.h:58-76), exported sorted (SparsityPattern.cpp:511-542).
// ---------------------------------------------------------------------------------------
struct Sparsity {
  std::vector<int32_t> row_ptr;
  std::vector<int32_t> col;
};

// Eigen compressed row-major coeffRef: binary search of the column in the sorted row
// (EigenMatrix.cpp:156-166). Returns the CSR slot or -1 if the entry is not in the pattern.
inline int64_t csr_find(const int32_t* row_ptr, const int32_t* col, int32_t r, int32_t c) {
  const int32_t* b = col + row_ptr[r];
  const int32_t* e = col + row_ptr[r + 1];
  const int32_t* it = std::lower_bound(b, e, c);
  if (it == e || *it != c) return -1;
  return it - col;
}

void gather_cell_coords(int gdim, const int32_t* cell, const double* x, double* xc) {
  for (int v = 0; v <= gdim; ++v)
    for (int i = 0; i < gdim; ++i) xc[v * gdim + i] = x[(int64_t)cell[v] * gdim + i];
}

// Element tensor dispatch (the ufc::cell_integral::tabulate_tensor call,
// Assembler.cpp:181-183).
int element_matrix(int kind, int gdim, const double* xc, const double* w0, const double* w1,
                   double* A) {
  switch (kind) {
    case ORC_P1_POISSON: orc_p1_stiffness(gdim, xc, w0, A); return gdim + 1;
    case ORC_P1_MASS: orc_p1_mass(gdim, xc, A); return gdim + 1;
    case ORC_VP2_ELASTICITY: orc_vp2_elasticity(xc, w0, w1, A); return 30;
    case ORC_VP1_ELASTICITY: orc_vp1_elasticity(xc, w0, w1, A); return 12;
    default: return -1;
  }
}

int element_vector(int kind, int gdim, const double* xc, const double* w0, const double* f3,
                   double* b) {
  switch (kind) {
    case ORC_P1_LOAD: orc_p1_load(gdim, xc, w0, b); return gdim + 1;
    case ORC_VP2_BODY_FORCE: orc_vp2_body_force(xc, f3, b); return 30;
    case ORC_VP1_BODY_FORCE: orc_vp1_body_force(xc, f3, b); return 12;
    default: return -1;
  }
}

int num_coeffs(int kind) {
  switch (kind) {
    case ORC_P1_POISSON: case ORC_P1_LOAD: return 1;
    case ORC_VP2_ELASTICITY: case ORC_VP1_ELASTICITY: return 2;
    default: return 0;
  }
}

// Function::restrict of a P1 coefficient (Function.cpp:449-476 -> EigenVector::get_local).
void restrict_p1(int nloc, const int32_t* cdofs, const double* c, double* w) {
  for (int k = 0; k < nloc; ++k) w[k] = c ? c[cdofs[k]] : 0.0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------
// Meshes
// ---------------------------------------------------------------------------------------
// UnitIntervalMesh(n): IntervalMesh::build, vertices x_i = a + i*(b-a)/n, cells (i, i+1).
void orc_interval_mesh(int32_t n, double* x, int32_t* cells) {
  for (int32_t i = 0; i <= n; ++i) x[i] = 0.0 + (static_cast<double>(i)) * (1.0 - 0.0) / n;
  for (int32_t i = 0; i < n; ++i) {
    cells[2 * i] = i;
    cells[2 * i + 1] = i + 1;
  }
}

// UnitSquareMesh(nx, ny, "right"): RectangleMesh.cpp:46-226 (vertices x-fastest,
// cells (v0,v1,v3), (v0,v2,v3) per square, :213-214). Already vertex-sorted, so
// mesh.order() leaves them unchanged.
void orc_rectangle_mesh_right(int32_t nx, int32_t ny, double* x, int32_t* cells) {
  int64_t v = 0;
  for (int32_t iy = 0; iy <= ny; ++iy) {
    const double yy = 0.0 + (static_cast<double>(iy)) * (1.0 - 0.0) / static_cast<double>(ny);
    for (int32_t ix = 0; ix <= nx; ++ix) {
      x[2 * v] = 0.0 + (static_cast<double>(ix)) * (1.0 - 0.0) / static_cast<double>(nx);
      x[2 * v + 1] = yy;
      ++v;
    }
  }
  int64_t c = 0;
  for (int32_t iy = 0; iy < ny; ++iy)
    for (int32_t ix = 0; ix < nx; ++ix) {
      const int32_t v0 = iy * (nx + 1) + ix, v1 = v0 + 1, v2 = v0 + (nx + 1), v3 = v1 + (nx + 1);
      const int32_t t[2][3] = {{v0, v1, v3}, {v0, v2, v3}};
      for (int k = 0; k < 2; ++k, ++c)
        for (int j = 0; j < 3; ++j) cells[3 * c + j] = t[k][j];
    }
}

// UnitCubeMesh(nx, ny, nz): BoxMesh::build_tet (BoxMesh.cpp:50-163), six Kuhn tets per
// cube (:147-152), then Mesh::order() sorts each cell's vertices ascending
// (MeshEditor.cpp:265-270, TetrahedronCell.cpp:444-449).
void orc_box_mesh(int32_t nx, int32_t ny, int32_t nz, double* x, int32_t* cells) {
  int64_t v = 0;
  for (int32_t iz = 0; iz <= nz; ++iz) {
    const double zz = 0.0 + (static_cast<double>(iz)) * (1.0 - 0.0) / static_cast<double>(nz);
    for (int32_t iy = 0; iy <= ny; ++iy) {
      const double yy = 0.0 + (static_cast<double>(iy)) * (1.0 - 0.0) / static_cast<double>(ny);
      for (int32_t ix = 0; ix <= nx; ++ix) {
        x[3 * v] = 0.0 + (static_cast<double>(ix)) * (1.0 - 0.0) / static_cast<double>(nx);
        x[3 * v + 1] = yy;
        x[3 * v + 2] = zz;
        ++v;
      }
    }
  }
  int64_t c = 0;
  for (int32_t iz = 0; iz < nz; ++iz)
    for (int32_t iy = 0; iy < ny; ++iy)
      for (int32_t ix = 0; ix < nx; ++ix) {
        const int32_t v0 = iz * (nx + 1) * (ny + 1) + iy * (nx + 1) + ix;
        const int32_t v1 = v0 + 1, v2 = v0 + (nx + 1), v3 = v1 + (nx + 1);
        const int32_t v4 = v0 + (nx + 1) * (ny + 1), v5 = v1 + (nx + 1) * (ny + 1);
        const int32_t v6 = v2 + (nx + 1) * (ny + 1), v7 = v3 + (nx + 1) * (ny + 1);
        const int32_t t[6][4] = {{v0, v1, v3, v7}, {v0, v1, v7, v5}, {v0, v5, v7, v4},
                                 {v0, v3, v2, v7}, {v0, v6, v4, v7}, {v0, v2, v6, v7}};
        for (int k = 0; k < 6; ++k, ++c) {
          int32_t s[4] = {t[k][0], t[k][1], t[k][2], t[k][3]};
          std::sort(s, s + 4);
          for (int j = 0; j < 4; ++j) cells[4 * c + j] = s[j];
        }
      }
}

// Edge numbering by sorted-vertex-key matching (TopologyComputation.cpp:152-330):
// every (cell, local edge) gets the key (vmin, vmax); keys are sorted and numbered in
// ascending order. Local edge order per TetrahedronCell::create_entities (:119-124).
int64_t orc_tet_edges(int64_t num_cells, const int32_t* cells, int32_t* cell_edges) {
  std::vector<std::tuple<int32_t, int32_t, int64_t>> keyed;
  keyed.reserve(num_cells * 6);
  for (int64_t c = 0; c < num_cells; ++c)
    for (int e = 0; e < 6; ++e) {
      int32_t a = cells[4 * c + kTetEdge[e][0]], b = cells[4 * c + kTetEdge[e][1]];
      if (a > b) std::swap(a, b);
      keyed.emplace_back(a, b, c * 6 + e);
    }
  std::sort(keyed.begin(), keyed.end());
  int64_t idx = -1;
  int32_t pa = -1, pb = -1;
  for (auto& k : keyed) {
    if (std::get<0>(k) != pa || std::get<1>(k) != pb) {
      ++idx;
      pa = std::get<0>(k);
      pb = std::get<1>(k);
    }
    cell_edges[std::get<2>(k)] = static_cast<int32_t>(idx);
  }
  return idx + 1;
}

// Vector P2 UFC dofmap with reorder_dofs_serial = False (DofMapBuilder.cpp:253-272):
// node = vertex, or nv + edge; dof = comp*(nv+ne) + node; cell-local order is
// component-major [comp][4 vertices, 6 edges] (python/test/unit/fem/test_dofmap.py:112-120).
void orc_vp2_dofmap(int64_t num_cells, const int32_t* cells, const int32_t* cell_edges,
                    int64_t nv, int64_t ne, int32_t* cell_dofs) {
  const int64_t nn = nv + ne;
  for (int64_t c = 0; c < num_cells; ++c)
    for (int comp = 0; comp < 3; ++comp) {
      for (int a = 0; a < 4; ++a)
        cell_dofs[30 * c + comp * 10 + a] = static_cast<int32_t>(comp * nn + cells[4 * c + a]);
      for (int e = 0; e < 6; ++e)
        cell_dofs[30 * c + comp * 10 + 4 + e] =
            static_cast<int32_t>(comp * nn + nv + cell_edges[6 * c + e]);
    }
}

// ---------------------------------------------------------------------------------------
// Sparsity pattern: SparsityPatternBuilder::build cell loop (SparsityPatternBuilder.cpp:95-111)
// inserting cell_dofs x cell_dofs, the optional diagonal (:245-264), sorted export.
// ---------------------------------------------------------------------------------------
void* orc_sparsity_build(int64_t num_cells, int32_t nd, const int32_t* cell_dofs,
                         int64_t num_rows, int32_t keep_diagonal) {
  std::vector<std::vector<int32_t>> rows(num_rows);
  for (int64_t c = 0; c < num_cells; ++c) {
    const int32_t* d = cell_dofs + c * nd;
    for (int i = 0; i < nd; ++i) {
      auto& r = rows[d[i]];
      for (int j = 0; j < nd; ++j)
        if (std::find(r.begin(), r.end(), d[j]) == r.end()) r.push_back(d[j]);
    }
  }
  if (keep_diagonal)
    for (int64_t i = 0; i < num_rows; ++i) {
      auto& r = rows[i];
      if (std::find(r.begin(), r.end(), (int32_t)i) == r.end()) r.push_back((int32_t)i);
    }
  auto* s = new Sparsity;
  s->row_ptr.resize(num_rows + 1);
  s->row_ptr[0] = 0;
  for (int64_t i = 0; i < num_rows; ++i) {
    std::sort(rows[i].begin(), rows[i].end());
    s->row_ptr[i + 1] = s->row_ptr[i] + (int32_t)rows[i].size();
  }
  s->col.reserve(s->row_ptr[num_rows]);
  for (auto& r : rows) s->col.insert(s->col.end(), r.begin(), r.end());
  return s;
}
int64_t orc_sparsity_nnz(void* h) { return (int64_t) static_cast<Sparsity*>(h)->col.size(); }
void orc_sparsity_copy(void* h, int32_t* row_ptr, int32_t* col) {
  auto* s = static_cast<Sparsity*>(h);
  std::memcpy(row_ptr, s->row_ptr.data(), s->row_ptr.size() * sizeof(int32_t));
  std::memcpy(col, s->col.data(), s->col.size() * sizeof(int32_t));
}
void orc_sparsity_free(void* h) { delete static_cast<Sparsity*>(h); }

// ---------------------------------------------------------------------------------------
// Element tensors. The integrands are polynomials integrated exactly, so these equal
// FFC's generated tabulate_tensor up to roundoff.
// ---------------------------------------------------------------------------------------
// inner(kappa*grad u, grad v)*dx with kappa P1: A[a][b] = mean(kappa)*vol*Dlam_a.Dlam_b.
void orc_p1_stiffness(int32_t gdim, const double* xc, const double* kappa, double* A) {
  const Simplex s = simplex_geometry(gdim, xc);
  const int n = gdim + 1;
  double kbar = 0.0;
  for (int k = 0; k < n; ++k) kbar += kappa[k];
  kbar /= n;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      double g = 0.0;
      for (int p = 0; p < gdim; ++p) g += s.D[a][p] * s.D[b][p];
      A[a * n + b] = kbar * s.vol * g;
    }
}

// u*v*dx: vol*(1 + delta_ab)/((d+1)(d+2)).
void orc_p1_mass(int32_t gdim, const double* xc, double* A) {
  const Simplex s = simplex_geometry(gdim, xc);
  const int n = gdim + 1;
  const double w = s.vol / ((gdim + 1.0) * (gdim + 2.0));
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) A[a * n + b] = w * (1.0 + (a == b));
}

// f*v*dx with f P1: b_a = vol/((d+1)(d+2)) * (f_a + sum_k f_k).
void orc_p1_load(int32_t gdim, const double* xc, const double* f, double* b) {
  const Simplex s = simplex_geometry(gdim, xc);
  const int n = gdim + 1;
  double fs = 0.0;
  for (int k = 0; k < n; ++k) fs += f[k];
  const double w = s.vol / ((gdim + 1.0) * (gdim + 2.0));
  for (int a = 0; a < n; ++a) b[a] = w * (f[a] + fs);
}

// Vector P2 elasticity, component-major local order (c*10 + a):
// K[(i,a),(j,b)] = int mu (delta_ij Dphi_a.Dphi_b + d_j phi_a d_i phi_b) + lmbda d_i phi_a d_j phi_b,
// integrated exactly with int lambda_n lambda_m lambda_q.
void orc_vp2_elasticity(const double* xc, const double* mu, const double* lmbda, double* A) {
  const Simplex s = simplex_geometry(3, xc);
  double Q[10][4][3];
  p2_grad_coeffs(s, Q);
  double Wm[4][4], Wl[4][4];
  weight3(mu, Wm);
  weight3(lmbda, Wl);
  // S[a][p][b][q] = vol * sum_{m,q'} Q[a][m][p] W[m][q'] Q[b][q'][q]
  static thread_local double Sm[10][3][10][3], Sl[10][3][10][3];
  for (int a = 0; a < 10; ++a)
    for (int p = 0; p < 3; ++p) {
      double tm[4] = {0, 0, 0, 0}, tl[4] = {0, 0, 0, 0};
      for (int m = 0; m < 4; ++m)
        for (int r = 0; r < 4; ++r) {
          tm[r] += Q[a][m][p] * Wm[m][r];
          tl[r] += Q[a][m][p] * Wl[m][r];
        }
      for (int b = 0; b < 10; ++b)
        for (int q = 0; q < 3; ++q) {
          double vm = 0.0, vl = 0.0;
          for (int r = 0; r < 4; ++r) {
            vm += tm[r] * Q[b][r][q];
            vl += tl[r] * Q[b][r][q];
          }
          Sm[a][p][b][q] = s.vol * vm;
          Sl[a][p][b][q] = s.vol * vl;
        }
    }
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 10; ++a)
      for (int j = 0; j < 3; ++j)
        for (int b = 0; b < 10; ++b) {
          double v = Sm[a][j][b][i] + Sl[a][i][b][j];
          if (i == j) v += Sm[a][0][b][0] + Sm[a][1][b][1] + Sm[a][2][b][2];
          A[(i * 10 + a) * 30 + (j * 10 + b)] = v;
        }
}

// inner(f, v)*dx, f constant: int phi_vertex = -vol/20, int phi_edge = vol/5.
void orc_vp2_body_force(const double* xc, const double* f3, double* b) {
  const Simplex s = simplex_geometry(3, xc);
  for (int i = 0; i < 3; ++i) {
    for (int a = 0; a < 4; ++a) b[i * 10 + a] = f3[i] * (-s.vol / 20.0);
    for (int e = 0; e < 6; ++e) b[i * 10 + 4 + e] = f3[i] * (s.vol / 5.0);
  }
}

// Vector P1 (or DG1) elasticity, component-major (c*4 + a); mu, lmbda P1 (means enter).
void orc_vp1_elasticity(const double* xc, const double* mu, const double* lmbda, double* A) {
  const Simplex s = simplex_geometry(3, xc);
  const double mb = 0.25 * (mu[0] + mu[1] + mu[2] + mu[3]);
  const double lb = 0.25 * (lmbda[0] + lmbda[1] + lmbda[2] + lmbda[3]);
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a)
      for (int j = 0; j < 3; ++j)
        for (int b = 0; b < 4; ++b) {
          double g = 0.0;
          for (int p = 0; p < 3; ++p) g += s.D[a][p] * s.D[b][p];
          double v = mb * (s.D[a][j] * s.D[b][i] + (i == j ? g : 0.0)) + lb * s.D[a][i] * s.D[b][j];
          A[(i * 4 + a) * 12 + (j * 4 + b)] = s.vol * v;
        }
}

void orc_vp1_body_force(const double* xc, const double* f3, double* b) {
  const Simplex s = simplex_geometry(3, xc);
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) b[i * 4 + a] = f3[i] * s.vol / 4.0;
}

// ---------------------------------------------------------------------------------------
// Assembler::assemble_cells (Assembler.cpp:105-194): per cell, gather coordinates
// (Cell.h:293-307), restrict coefficients (UFC.cpp:156-169), tabulate_tensor, add_local
// (GenericMatrix.h:106-112 -> EigenMatrix::add, EigenMatrix.cpp:156-166). Cells in
// ascending order; values accumulate in that order.
// ---------------------------------------------------------------------------------------
int32_t orc_assemble_matrix(int32_t kind, int32_t gdim, int64_t num_cells, const int32_t* cells,
                            const double* x, int32_t nd, const int32_t* cell_dofs,
                            const int32_t* coeff_dofs, const double* c0, const double* c1,
                            const int32_t* row_ptr, const int32_t* col, double* values) {
  const int nv = gdim + 1;
  std::vector<double> A(nd * nd);
  double xc[12], w0[4], w1[4];
  const int nc = num_coeffs(kind);
  for (int64_t c = 0; c < num_cells; ++c) {
    gather_cell_coords(gdim, cells + c * nv, x, xc);
    if (nc > 0) restrict_p1(nv, coeff_dofs + c * nv, c0, w0);
    if (nc > 1) restrict_p1(nv, coeff_dofs + c * nv, c1, w1);
    if (element_matrix(kind, gdim, xc, w0, w1, A.data()) != nd) return -1;
    const int32_t* d = cell_dofs + c * nd;
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) {
        const int64_t slot = csr_find(row_ptr, col, d[i], d[j]);
        if (slot < 0) return -2;
        values[slot] += A[i * nd + j];
      }
  }
  return 0;
}

// Vector assembly: EigenVector::add (EigenVector.cpp:160-165), x[rows[i]] += block[i].
int32_t orc_assemble_vector(int32_t kind, int32_t gdim, int64_t num_cells, const int32_t* cells,
                            const double* x, int32_t nd, const int32_t* cell_dofs,
                            const int32_t* coeff_dofs, const double* c0, const double* f3,
                            double* b) {
  const int nv = gdim + 1;
  std::vector<double> be(nd);
  double xc[12], w0[4];
  for (int64_t c = 0; c < num_cells; ++c) {
    gather_cell_coords(gdim, cells + c * nv, x, xc);
    if (num_coeffs(kind) > 0) restrict_p1(nv, coeff_dofs + c * nv, c0, w0);
    if (element_vector(kind, gdim, xc, w0, f3, be.data()) != nd) return -1;
    const int32_t* d = cell_dofs + c * nd;
    for (int i = 0; i < nd; ++i) b[d[i]] += be[i];
  }
  return 0;
}

// SystemAssembler::cell_wise_assembly (SystemAssembler.cpp:303-492) with apply_bc for a
// square form (:1051-1075): for each local column i carrying a Dirichlet dof: zero row i,
// b -= A(:,i)*g, zero column i, b(i) = g, A(i,i) = 1. Cells [c_begin, c_end) in ascending
// order; add(ptr, v) performs the scatter.
}  // extern "C"

template <class Add>
static int32_t system_cells(int32_t a_kind, int32_t L_kind, int32_t gdim, int64_t c_begin,
                            int64_t c_end, const int32_t* cells, const double* x, int32_t nd,
                            const int32_t* cell_dofs, const int32_t* coeff_dofs,
                            const double* a_c0, const double* a_c1, const double* L_c0,
                            const double* f3, const uint8_t* bc_mask, const double* bc_value,
                            const int32_t* row_ptr, const int32_t* col, double* values, double* b,
                            Add add) {
  const int nv = gdim + 1;
  std::vector<double> A(nd * nd), be(nd);
  double xc[12], w0[4], w1[4], l0[4];
  for (int64_t c = c_begin; c < c_end; ++c) {
    gather_cell_coords(gdim, cells + c * nv, x, xc);
    const int32_t* cd = coeff_dofs + c * nv;
    const int32_t* d = cell_dofs + c * nd;
    if (num_coeffs(a_kind) > 0) restrict_p1(nv, cd, a_c0, w0);
    if (num_coeffs(a_kind) > 1) restrict_p1(nv, cd, a_c1, w1);
    if (num_coeffs(L_kind) > 0) restrict_p1(nv, cd, L_c0, l0);
    if (element_matrix(a_kind, gdim, xc, w0, w1, A.data()) != nd) return -1;
    if (element_vector(L_kind, gdim, xc, l0, f3, be.data()) != nd) return -1;
    for (int i = 0; i < nd; ++i) {
      const int32_t ii = d[i];
      if (!bc_mask[ii]) continue;
      const double g = bc_value[ii];
      for (int j = 0; j < nd; ++j) A[i * nd + j] = 0.0;
      for (int r = 0; r < nd; ++r) be[r] -= A[r * nd + i] * g;
      for (int r = 0; r < nd; ++r) A[r * nd + i] = 0.0;
      be[i] = g;
      A[i * nd + i] = 1.0;
    }
    if (values)
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) {
          const int64_t slot = csr_find(row_ptr, col, d[i], d[j]);
          if (slot < 0) return -2;
          add(values + slot, A[i * nd + j]);
        }
    if (b)
      for (int i = 0; i < nd; ++i) add(b + d[i], be[i]);
  }
  return 0;
}

extern "C" {

int32_t orc_assemble_system(int32_t a_kind, int32_t L_kind, int32_t gdim, int64_t num_cells,
                            const int32_t* cells, const double* x, int32_t nd,
                            const int32_t* cell_dofs, const int32_t* coeff_dofs,
                            const double* a_c0, const double* a_c1, const double* L_c0,
                            const double* f3, const uint8_t* bc_mask, const double* bc_value,
                            const int32_t* row_ptr, const int32_t* col, double* values,
                            double* b) {
  return system_cells(a_kind, L_kind, gdim, 0, num_cells, cells, x, nd, cell_dofs, coeff_dofs,
                      a_c0, a_c1, L_c0, f3, bc_mask, bc_value, row_ptr, col, values, b,
                      [](double* p, double v) { *p += v; });
}

// The same loop on nthreads host threads (contiguous cell ranges), the scatter made race-free
// with atomic compare-and-swap adds: a multi-core CPU timing of the same algorithm (DOLFIN's
// assembler is serial per process). Sums equal the serial ones up to summation order.
int32_t orc_assemble_system_threads(int32_t a_kind, int32_t L_kind, int32_t gdim,
                                    int64_t num_cells, const int32_t* cells, const double* x,
                                    int32_t nd, const int32_t* cell_dofs,
                                    const int32_t* coeff_dofs, const double* a_c0,
                                    const double* a_c1, const double* L_c0, const double* f3,
                                    const uint8_t* bc_mask, const double* bc_value,
                                    const int32_t* row_ptr, const int32_t* col, double* values,
                                    double* b, int32_t nthreads) {
  if (nthreads < 1) return -3;
  auto atomic_add = [](double* p, double v) {
    uint64_t* u = reinterpret_cast<uint64_t*>(p);
    uint64_t old = __atomic_load_n(u, __ATOMIC_RELAXED);
    for (;;) {
      double cur;
      std::memcpy(&cur, &old, sizeof cur);
      const double nxt = cur + v;
      uint64_t desired;
      std::memcpy(&desired, &nxt, sizeof desired);
      if (__atomic_compare_exchange_n(u, &old, desired, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
        return;
    }
  };
  std::vector<int32_t> rc(nthreads, 0);
  std::vector<std::thread> pool;
  for (int32_t t = 0; t < nthreads; ++t)
    pool.emplace_back([&, t]() {
      const int64_t c0 = num_cells * t / nthreads, c1 = num_cells * (t + 1) / nthreads;
      rc[t] = system_cells(a_kind, L_kind, gdim, c0, c1, cells, x, nd, cell_dofs, coeff_dofs,
                           a_c0, a_c1, L_c0, f3, bc_mask, bc_value, row_ptr, col, values, b,
                           atomic_add);
    });
  for (auto& th : pool) th.join();
  for (int32_t r : rc)
    if (r != 0) return r;
  return 0;
}

// ---------------------------------------------------------------------------------------
// Adjoint: J = u^T K(kappa, x) u - F(f, x)^T u, P1. Per cell:
//   E_c = mean(kappa) vol |grad u|^2,  G_c = vol/((d+1)(d+2)) (u.f + sum u sum f)
//   dE/dkappa_k = vol |grad u|^2/(d+1)
//   dE/dx_{l,q} = mean(kappa) vol (|grad u|^2 D_lq - 2 (grad u)_q (grad u . D_l)),
//   dG/dx_{l,q} = G_c D_lq     (d vol = vol div V, d grad u = -grad V grad u)
// -- the discrete form of the shape derivative checked in
// python/test/unit/fem/test_coordinate_derivative.py:47-49.
// ---------------------------------------------------------------------------------------
// ---- exterior facets ------------------------------------------------------------------
// Facet::exterior (a facet with one incident cell), computed by sorting the vertex keys of
// all cell facets (the role of TopologyComputation's key matching, TopologyComputation.cpp
// :152-330, here only to count incidences).
int64_t orc_exterior_facets(int32_t gdim, int64_t num_cells, const int32_t* cells,
                            int32_t* facet_cell, int32_t* facet_local) {
  const int nv = gdim + 1;
  std::vector<std::tuple<std::array<int32_t, 3>, int64_t, int>> keys;
  keys.reserve(static_cast<size_t>(num_cells) * nv);
  for (int64_t c = 0; c < num_cells; ++c)
    for (int f = 0; f < nv; ++f) {
      std::array<int32_t, 3> k = {-1, -1, -1};
      int n = 0;
      for (int a = 0; a < nv; ++a)
        if (a != f) k[n++] = cells[c * nv + a];
      std::sort(k.begin(), k.begin() + n);
      keys.emplace_back(k, c, f);
    }
  std::sort(keys.begin(), keys.end());
  std::vector<std::pair<int64_t, int>> ext;
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && std::get<0>(keys[j]) == std::get<0>(keys[i])) ++j;
    if (j - i == 1) ext.emplace_back(std::get<1>(keys[i]), std::get<2>(keys[i]));
    i = j;
  }
  std::sort(ext.begin(), ext.end());
  if (facet_cell)
    for (size_t i = 0; i < ext.size(); ++i) {
      facet_cell[i] = static_cast<int32_t>(ext[i].first);
      facet_local[i] = ext[i].second;
    }
  return static_cast<int64_t>(ext.size());
}

namespace {
// measure of the facet opposite local vertex lf
double facet_measure(int gdim, const double* xc, int lf) {
  int v[3], n = 0;
  for (int a = 0; a <= gdim; ++a)
    if (a != lf) v[n++] = a;
  if (gdim == 1) return 1.0;
  if (gdim == 2) {
    const double dx = xc[v[1] * 2] - xc[v[0] * 2], dy = xc[v[1] * 2 + 1] - xc[v[0] * 2 + 1];
    return std::sqrt(dx * dx + dy * dy);
  }
  double e1[3], e2[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = xc[v[1] * 3 + i] - xc[v[0] * 3 + i];
    e2[i] = xc[v[2] * 3 + i] - xc[v[0] * 3 + i];
  }
  const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
               cz = e1[0] * e2[1] - e1[1] * e2[0];
  return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}
}  // namespace

void orc_p1_facet_mass(int32_t gdim, const double* xc, int32_t lf, double alpha, double* A) {
  const int nv = gdim + 1;
  const double w = alpha * facet_measure(gdim, xc, lf) / (gdim * (gdim + 1.0));
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nv; ++b)
      A[a * nv + b] = (a == lf || b == lf) ? 0.0 : (a == b ? 2.0 * w : w);
}

void orc_p1_facet_load(int32_t gdim, const double* xc, int32_t lf, const double* g, double* b) {
  const int nv = gdim + 1;
  const double w = facet_measure(gdim, xc, lf) / (gdim * (gdim + 1.0));
  double gs = 0.0;
  for (int a = 0; a < nv; ++a)
    if (a != lf) gs += g[a];
  for (int a = 0; a < nv; ++a) b[a] = (a == lf) ? 0.0 : w * (g[a] + gs);
}

void orc_vp2_facet_traction(const double* xc, int32_t lf, const double* t3, double* b) {
  static const int edge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  const double w = facet_measure(3, xc, lf) / 3.0;
  for (int i = 0; i < 3; ++i) {
    for (int a = 0; a < 4; ++a) b[i * 10 + a] = 0.0;
    for (int e = 0; e < 6; ++e)
      b[i * 10 + 4 + e] = (edge[e][0] != lf && edge[e][1] != lf) ? t3[i] * w : 0.0;
  }
}

int32_t orc_assemble_facets(int32_t a_kind, int32_t L_kind, int32_t gdim, const int32_t* cells,
                            const double* x, const int32_t* cell_dofs, const int32_t* coeff_dofs,
                            int64_t num_facets, const int32_t* facet_cell,
                            const int32_t* facet_local, const uint8_t* facet_mask, double alpha,
                            const double* g, const int32_t* row_ptr, const int32_t* col,
                            double* values, double* b) {
  const int nv = gdim + 1;
  double xc[12], gl[4], A[16], be[4];
  if (L_kind == ORC_VP2_FACET_TRACTION) {
    if (a_kind >= 0 || gdim != 3) return -1;
    double bt[30];
    for (int64_t k = 0; k < num_facets; ++k) {
      if (facet_mask && !facet_mask[k]) continue;
      const int64_t c = facet_cell[k];
      gather_cell_coords(3, cells + c * 4, x, xc);
      orc_vp2_facet_traction(xc, facet_local[k], g, bt);
      if (b)
        for (int j = 0; j < 30; ++j) b[cell_dofs[c * 30 + j]] += bt[j];
    }
    return 0;
  }
  for (int64_t k = 0; k < num_facets; ++k) {
    if (facet_mask && !facet_mask[k]) continue;
    const int64_t c = facet_cell[k];
    const int lf = facet_local[k];
    gather_cell_coords(gdim, cells + c * nv, x, xc);
    const int32_t* d = cell_dofs + c * nv;
    if (a_kind == ORC_P1_FACET_MASS && values) {
      orc_p1_facet_mass(gdim, xc, lf, alpha, A);
      for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) {
          const int64_t slot = csr_find(row_ptr, col, d[i], d[j]);
          if (slot < 0) return -2;
          values[slot] += A[i * nv + j];
        }
    } else if (a_kind >= 0) {
      return -1;
    }
    if (L_kind == ORC_P1_FACET_LOAD && b) {
      restrict_p1(nv, coeff_dofs + c * nv, g, gl);
      orc_p1_facet_load(gdim, xc, lf, gl, be);
      for (int i = 0; i < nv; ++i) b[d[i]] += be[i];
    } else if (L_kind >= 0) {
      return -1;
    }
  }
  return 0;
}

int32_t orc_assemble_system_facets(int32_t a_kind, int32_t L_kind, int32_t fa_kind,
                                   int32_t fL_kind, int32_t gdim, int64_t num_cells,
                                   const int32_t* cells, const double* x,
                                   const int32_t* cell_dofs, const int32_t* coeff_dofs,
                                   const double* a_c0, const double* L_c0, double alpha,
                                   const double* g, int64_t num_facets, const int32_t* facet_cell,
                                   const int32_t* facet_local, const uint8_t* facet_mask,
                                   const uint8_t* bc_mask, const double* bc_value,
                                   const int32_t* row_ptr, const int32_t* col, double* values,
                                   double* b) {
  const int nv = gdim + 1;
  // exterior facets of each cell (first index into the facet list; facets are cell-sorted)
  std::vector<int64_t> first(num_cells + 1, 0);
  for (int64_t k = 0; k < num_facets; ++k) first[facet_cell[k] + 1]++;
  for (int64_t c = 0; c < num_cells; ++c) first[c + 1] += first[c];
  double xc[12], w0[4], l0[4], gl[4], A[16], be[4], F[16], fb[4];
  for (int64_t c = 0; c < num_cells; ++c) {
    gather_cell_coords(gdim, cells + c * nv, x, xc);
    const int32_t* cd = coeff_dofs + c * nv;
    const int32_t* d = cell_dofs + c * nv;
    if (num_coeffs(a_kind) > 0) restrict_p1(nv, cd, a_c0, w0);
    if (num_coeffs(L_kind) > 0) restrict_p1(nv, cd, L_c0, l0);
    if (element_matrix(a_kind, gdim, xc, w0, nullptr, A) != nv) return -1;
    if (element_vector(L_kind, gdim, xc, l0, nullptr, be) != nv) return -1;
    for (int64_t k = first[c]; k < first[c + 1]; ++k) {
      if (facet_mask && !facet_mask[k]) continue;
      const int lf = facet_local[k];
      if (fa_kind == ORC_P1_FACET_MASS) {
        orc_p1_facet_mass(gdim, xc, lf, alpha, F);
        for (int i = 0; i < nv * nv; ++i) A[i] += F[i];
      }
      if (fL_kind == ORC_P1_FACET_LOAD) {
        restrict_p1(nv, cd, g, gl);
        orc_p1_facet_load(gdim, xc, lf, gl, fb);
        for (int i = 0; i < nv; ++i) be[i] += fb[i];
      }
    }
    for (int i = 0; i < nv; ++i) {
      const int32_t ii = d[i];
      if (!bc_mask[ii]) continue;
      const double gv = bc_value[ii];
      for (int j = 0; j < nv; ++j) A[i * nv + j] = 0.0;
      for (int r = 0; r < nv; ++r) be[r] -= A[r * nv + i] * gv;
      for (int r = 0; r < nv; ++r) A[r * nv + i] = 0.0;
      be[i] = gv;
      A[i * nv + i] = 1.0;
    }
    if (values)
      for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) {
          const int64_t slot = csr_find(row_ptr, col, d[i], d[j]);
          if (slot < 0) return -2;
          values[slot] += A[i * nv + j];
        }
    if (b)
      for (int i = 0; i < nv; ++i) b[d[i]] += be[i];
  }
  return 0;
}

double orc_adjoint_p1(int32_t gdim, int64_t num_cells, const int32_t* cells, const double* x,
                      int64_t nv, const int32_t* cell_dofs, int64_t ndofs, const double* kappa,
                      const double* f, const double* u, double* dJdkappa, double* dJdx) {
  const int n = gdim + 1;
  std::memset(dJdkappa, 0, sizeof(double) * ndofs);
  std::memset(dJdx, 0, sizeof(double) * nv * gdim);
  double J = 0.0, xc[12];
  for (int64_t c = 0; c < num_cells; ++c) {
    const int32_t* cv = cells + c * n;
    const int32_t* cd = cell_dofs + c * n;
    gather_cell_coords(gdim, cv, x, xc);
    const Simplex s = simplex_geometry(gdim, xc);
    double gu[3] = {0, 0, 0}, kb = 0, us = 0, fs = 0, uf = 0;
    for (int a = 0; a < n; ++a) {
      const double ua = u[cd[a]];
      for (int p = 0; p < gdim; ++p) gu[p] += ua * s.D[a][p];
      kb += kappa[cd[a]];
      us += ua;
      fs += f[cd[a]];
      uf += ua * f[cd[a]];
    }
    kb /= n;
    double g2 = 0.0;
    for (int p = 0; p < gdim; ++p) g2 += gu[p] * gu[p];
    const double E = kb * s.vol * g2;
    const double G = s.vol / ((gdim + 1.0) * (gdim + 2.0)) * (uf + us * fs);
    J += E - G;
    for (int a = 0; a < n; ++a) dJdkappa[cd[a]] += s.vol * g2 / n;
    for (int l = 0; l < n; ++l) {
      double gdl = 0.0;
      for (int p = 0; p < gdim; ++p) gdl += gu[p] * s.D[l][p];
      for (int q = 0; q < gdim; ++q)
        dJdx[(int64_t)cv[l] * gdim + q] +=
            kb * s.vol * (g2 * s.D[l][q] - 2.0 * gu[q] * gdl) - G * s.D[l][q];
    }
  }
  return J;
}

// Vector P2 elasticity adjoint. With grad u_h = sum_m lambda_m G_m (G_m 3x3),
// E_c = vol sum_{m,r} [Wmu'_{mr} 2 eps_m:eps_r + Wl'_{mr} tr G_m tr G_r]
// (W' = W/vol, the exact weights). dE/dmu_n, dE/dlmbda_n are the same sums with the
// weight of node n; dE/dx_{l,q} = E D_lq - sum_m P_m[i][p] G_m[i][q] D_lp with
// P_m = dE/dG_m.  The load term G_c = sum f_i u_{ia} int(phi_a) scales with vol.
double orc_adjoint_vp2(int64_t num_cells, const int32_t* cells, const double* x, int64_t nv,
                       const int32_t* cell_dofs, const int32_t* coeff_dofs, int64_t ncoeff,
                       const double* mu, const double* lmbda, const double* f3, const double* u,
                       double* dJdmu, double* dJdlmbda, double* dJdx) {
  std::memset(dJdmu, 0, sizeof(double) * ncoeff);
  std::memset(dJdlmbda, 0, sizeof(double) * ncoeff);
  std::memset(dJdx, 0, sizeof(double) * nv * 3);
  double J = 0.0, xc[12];
  for (int64_t c = 0; c < num_cells; ++c) {
    const int32_t* cv = cells + 4 * c;
    const int32_t* cd = cell_dofs + 30 * c;
    const int32_t* kd = coeff_dofs + 4 * c;
    gather_cell_coords(3, cv, x, xc);
    const Simplex s = simplex_geometry(3, xc);
    double Q[10][4][3];
    p2_grad_coeffs(s, Q);
    double G[4][3][3] = {};
    for (int m = 0; m < 4; ++m)
      for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 10; ++a) {
          const double uia = u[cd[i * 10 + a]];
          for (int p = 0; p < 3; ++p) G[m][i][p] += uia * Q[a][m][p];
        }
    double ee[4][4], tt[4][4];
    for (int m = 0; m < 4; ++m)
      for (int r = 0; r < 4; ++r) {
        double e = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int p = 0; p < 3; ++p) {
            const double em = 0.5 * (G[m][i][p] + G[m][p][i]);
            const double er = 0.5 * (G[r][i][p] + G[r][p][i]);
            e += em * er;
          }
        ee[m][r] = 2.0 * e;
        tt[m][r] = (G[m][0][0] + G[m][1][1] + G[m][2][2]) * (G[r][0][0] + G[r][1][1] + G[r][2][2]);
      }
    double mun[4], lan[4];
    for (int k = 0; k < 4; ++k) {
      mun[k] = mu[kd[k]];
      lan[k] = lmbda[kd[k]];
    }
    double Wm[4][4], Wl[4][4];
    weight3(mun, Wm);
    weight3(lan, Wl);
    double E = 0.0;
    for (int m = 0; m < 4; ++m)
      for (int r = 0; r < 4; ++r) E += s.vol * (Wm[m][r] * ee[m][r] + Wl[m][r] * tt[m][r]);
    for (int n = 0; n < 4; ++n) {
      double dm = 0.0, dl = 0.0;
      for (int m = 0; m < 4; ++m)
        for (int r = 0; r < 4; ++r) {
          dm += bary3(n, m, r) * ee[m][r];
          dl += bary3(n, m, r) * tt[m][r];
        }
      dJdmu[kd[n]] += s.vol * dm;
      dJdlmbda[kd[n]] += s.vol * dl;
    }
    // P_m[i][p] = dE/dG_m[i][p] = 2 vol sum_r (Wm_mr 2 eps_r[i][p] + Wl_mr tr G_r delta_ip)
    double P[4][3][3] = {};
    for (int m = 0; m < 4; ++m)
      for (int r = 0; r < 4; ++r) {
        const double trr = G[r][0][0] + G[r][1][1] + G[r][2][2];
        for (int i = 0; i < 3; ++i)
          for (int p = 0; p < 3; ++p) {
            const double er = 0.5 * (G[r][i][p] + G[r][p][i]);
            P[m][i][p] += 2.0 * s.vol * (Wm[m][r] * 2.0 * er + (i == p ? Wl[m][r] * trr : 0.0));
          }
      }
    double Gl = 0.0;
    for (int i = 0; i < 3; ++i) {
      for (int a = 0; a < 4; ++a) Gl += f3[i] * u[cd[i * 10 + a]] * (-s.vol / 20.0);
      for (int e = 0; e < 6; ++e) Gl += f3[i] * u[cd[i * 10 + 4 + e]] * (s.vol / 5.0);
    }
    J += E - Gl;
    for (int l = 0; l < 4; ++l)
      for (int q = 0; q < 3; ++q) {
        double t = 0.0;
        for (int m = 0; m < 4; ++m)
          for (int i = 0; i < 3; ++i)
            for (int p = 0; p < 3; ++p) t += P[m][i][p] * G[m][i][q] * s.D[l][p];
        dJdx[(int64_t)cv[l] * 3 + q] += (E - Gl) * s.D[l][q] - t;
      }
  }
  return J;
}

}  // extern "C"
