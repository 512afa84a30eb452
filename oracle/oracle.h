/*
 * CPU oracle for DOLFIN 2019.1 cell-integral assembly -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a serial C++ restatement of the reference algorithm for the cell-assembly
 * path: built-in meshes, UFC dof numbering, the sparsity pattern, the element tensors
 * that FFC generates for the forms used here, the Assembler / SystemAssembler cell loops
 * with Eigen CSR "coeffRef +=" scatter semantics, and the adjoint (derivative-form)
 * closed forms. Only the test-suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it; it is the checker, never the product path.
 *
 * Parity pins (what each function is checked against, in tests/):
 *   orc_*_mesh          reference constructors restated; pinned by vertex/cell counts,
 *                       the known-answer norms below and brute-force orientation checks.
 *   orc_tet_edges       TopologyComputation key matching; pinned by edge counts and the
 *                       P2 nnz closed form 9*(230n^3+138n^2+24n+1).
 *   orc_sparsity_*      SparsityPatternBuilder + SparsityPattern::diagonal_pattern(sorted);
 *                       pinned by brute-force dense pattern on tiny meshes and nnz closed forms.
 *   orc_p1_*            closed forms; pinned by the reference-tet K and M of
 *                       python/test/unit/fem/test_assembler.py:351-390 and the 1D/3D norms.
 *   orc_vp1_elasticity  pinned by the vector-DG1 eps:eps norms of
 *                       python/test/unit/fem/test_system_assembler.py:30-70.
 *   orc_vp2_*           exact barycentric-monomial integration; pinned by an independent
 *                       numerical quadrature (Python, tests), rigid-body null space and
 *                       the int(phi) constants.
 *   orc_assemble_*      Assembler.cpp:105-194 / SystemAssembler.cpp:303-492,1036-1112;
 *                       pinned by 96.847818767384 / 96.564760289080 and
 *                       811.75365721381274 / 1.43583841167606474.
 *   orc_exterior_facets / orc_p1_facet_* / orc_assemble_facets / orc_assemble_system_facets
 *                       Facet::exterior + the P1 facet integrals; pinned by the boundary
 *                       measures of python/test/unit/fem/test_assembler.py:140-153 (4.0 on
 *                       UnitSquareMesh(4,4), 6.0 on UnitCubeMesh(2,2,2)) and :296-322 (1.0
 *                       on the marked lower-left boundary of UnitSquareMesh(8,8)), the
 *                       exterior facet counts 4n / 12n^2, and a Monte-Carlo-free closed
 *                       form check of one facet (tests).
 *   orc_adjoint_*       closed-form derivative of J = u^T K u - F^T u; pinned by central
 *                       finite differences of orc_assemble_* (tests) -- the FD check is
 *                       independent of the derivative formula.
 *   orc_vp2_dofmap      the vector-P2 cell dofs as component sub-dofmaps concatenated (pinned
 *                       by python/test/unit/fem/test_dofmap.py:112-120); the global order of
 *                       the nodes (vertices, then edges by key) is an input convention, parity
 *                       unpinned: FFC/UFC's own node order is in neither the reference nor
 *                       this image, and the engine takes whatever dofmap DOLFIN passes.
 * The pure-Python reverse Cuthill-McKee of oracle.py is pinned by the orderings Boost's
 * documentation prints for its example graph; it refuses graphs where a vertex queues more
 * than 16 new neighbours at once (Boost's order of equal degrees there: parity unpinned).
 */
#ifndef DOLFIN_AMD_ORACLE_H
#define DOLFIN_AMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Form kinds (same numbering as the product library's). */
enum {
  ORC_P1_POISSON = 0,   /* inner(kappa*grad u, grad v)*dx, kappa P1             */
  ORC_P1_LOAD = 1,      /* f*v*dx, f P1                                          */
  ORC_P1_MASS = 2,      /* u*v*dx                                                */
  ORC_VP2_ELASTICITY = 3, /* inner(2 mu eps(u) + lmbda tr eps(u) I, eps(v))*dx   */
  ORC_VP2_BODY_FORCE = 4, /* inner(f, v)*dx, f constant vector                    */
  ORC_VP1_ELASTICITY = 5, /* vector P1 (or DG1) version of ORC_VP2_ELASTICITY     */
  ORC_VP1_BODY_FORCE = 6,
  ORC_P1_FACET_MASS = 7,  /* alpha*u*v*ds, alpha constant (Robin term)              */
  ORC_P1_FACET_LOAD = 8,  /* g*v*ds, g P1 (Neumann term)                            */
  ORC_VP2_FACET_TRACTION = 9 /* inner(t, v)*ds, t constant vector, vector P2 (traction) */
};

/* ---- meshes ---------------------------------------------------------------- */
void orc_interval_mesh(int32_t n, double* x, int32_t* cells);
void orc_rectangle_mesh_right(int32_t nx, int32_t ny, double* x, int32_t* cells);
void orc_box_mesh(int32_t nx, int32_t ny, int32_t nz, double* x, int32_t* cells);
int64_t orc_tet_edges(int64_t num_cells, const int32_t* cells, int32_t* cell_edges);

/* ---- dofmaps ----------------------------------------------------------------- */
/* vector P2 on tets, UFC (component-major) numbering: dof = comp*(nv+ne) + node */
void orc_vp2_dofmap(int64_t num_cells, const int32_t* cells, const int32_t* cell_edges,
                    int64_t nv, int64_t ne, int32_t* cell_dofs);

/* ---- sparsity ------------------------------------------------------------------ */
void* orc_sparsity_build(int64_t num_cells, int32_t ndofs_cell, const int32_t* cell_dofs,
                         int64_t num_rows, int32_t keep_diagonal);
int64_t orc_sparsity_nnz(void* h);
void orc_sparsity_copy(void* h, int32_t* row_ptr, int32_t* col);
void orc_sparsity_free(void* h);

/* ---- element tensors (ufc::cell_integral::tabulate_tensor restated) ------------ */
void orc_p1_stiffness(int32_t gdim, const double* xc, const double* kappa, double* A);
void orc_p1_mass(int32_t gdim, const double* xc, double* A);
void orc_p1_load(int32_t gdim, const double* xc, const double* f, double* b);
void orc_vp2_elasticity(const double* xc, const double* mu, const double* lmbda, double* A);
void orc_vp2_body_force(const double* xc, const double* f3, double* b);
void orc_vp1_elasticity(const double* xc, const double* mu, const double* lmbda, double* A);
void orc_vp1_body_force(const double* xc, const double* f3, double* b);

/* ---- global assembly ------------------------------------------------------------ */
/* coeffs: P1 nodal arrays addressed through coeff_dofs [num_cells x (gdim+1)];
   ncoeff = 1 (kappa / f), 2 (mu, lmbda), 0 (mass). f3: constant vector (body force). */
int32_t orc_assemble_matrix(int32_t kind, int32_t gdim, int64_t num_cells, const int32_t* cells,
                            const double* x, int32_t ndofs_cell, const int32_t* cell_dofs,
                            const int32_t* coeff_dofs, const double* c0, const double* c1,
                            const int32_t* row_ptr, const int32_t* col, double* values);
int32_t orc_assemble_vector(int32_t kind, int32_t gdim, int64_t num_cells, const int32_t* cells,
                            const double* x, int32_t ndofs_cell, const int32_t* cell_dofs,
                            const int32_t* coeff_dofs, const double* c0, const double* f3,
                            double* b);
/* SystemAssembler: (a_kind, L_kind) fused with per-cell symmetric Dirichlet lifting.
   bc_mask[dof] != 0 marks a Dirichlet dof with value bc_value[dof]. values/b may be NULL. */
int32_t orc_assemble_system(int32_t a_kind, int32_t L_kind, int32_t gdim, int64_t num_cells,
                            const int32_t* cells, const double* x, int32_t ndofs_cell,
                            const int32_t* cell_dofs, const int32_t* coeff_dofs,
                            const double* a_c0, const double* a_c1, const double* L_c0,
                            const double* f3, const uint8_t* bc_mask, const double* bc_value,
                            const int32_t* row_ptr, const int32_t* col, double* values,
                            double* b);
/* The same loop on nthreads host threads with atomic scatter (CPU timing only). */
int32_t orc_assemble_system_threads(int32_t a_kind, int32_t L_kind, int32_t gdim,
                                    int64_t num_cells, const int32_t* cells, const double* x,
                                    int32_t ndofs_cell, const int32_t* cell_dofs,
                                    const int32_t* coeff_dofs, const double* a_c0,
                                    const double* a_c1, const double* L_c0, const double* f3,
                                    const uint8_t* bc_mask, const double* bc_value,
                                    const int32_t* row_ptr, const int32_t* col, double* values,
                                    double* b, int32_t nthreads);

/* ---- exterior facets ----------------------------------------------------------------- */
/* Facets that belong to exactly one cell (Facet::exterior), listed by cell and then by local
   facet index (local facet k = the facet opposite local vertex k, UFC). Returns their number;
   facet_cell / facet_local are written when non-NULL. */
int64_t orc_exterior_facets(int32_t gdim, int64_t num_cells, const int32_t* cells,
                            int32_t* facet_cell, int32_t* facet_local);
/* P1 exterior-facet tensors in cell-local numbering ((gdim+1)^2 / gdim+1 entries, zero off
   the facet): mass alpha |F| (1 + delta_ab) / (d (d + 1)) and load
   |F| / (d (d + 1)) (g_a + sum_{b in F} g_b), d = gdim facet vertices, |F| = 1 for a point. */
void orc_p1_facet_mass(int32_t gdim, const double* xc, int32_t local_facet, double alpha,
                       double* A);
void orc_p1_facet_load(int32_t gdim, const double* xc, int32_t local_facet, const double* g,
                       double* b);
/* Vector-P2 traction inner(t, v)*ds on the facet opposite local vertex local_facet, in the
   component-major cell order (30 entries): int_F phi = 0 for the vertex nodes and |F| / 3 for
   the three edge nodes of the facet, times t_i. */
void orc_vp2_facet_traction(const double* xc, int32_t local_facet, const double* t3, double* b);
/* Assembler::assemble_exterior_facets (Assembler.cpp:196-298) for the P1 facet forms over
   the facets with facet_mask[k] != 0 (NULL: all), adding into values (a_kind) and b (L_kind,
   g through coeff_dofs). a_kind / L_kind = -1 skips that tensor. L_kind ORC_VP2_FACET_TRACTION
   (with a_kind -1) takes vector-P2 cell_dofs (30 per cell) and g = the 3 traction
   components. */
int32_t orc_assemble_facets(int32_t a_kind, int32_t L_kind, int32_t gdim, const int32_t* cells,
                            const double* x, const int32_t* cell_dofs, const int32_t* coeff_dofs,
                            int64_t num_facets, const int32_t* facet_cell,
                            const int32_t* facet_local, const uint8_t* facet_mask, double alpha,
                            const double* g, const int32_t* row_ptr, const int32_t* col,
                            double* values, double* b);
/* SystemAssembler::cell_wise_assembly with exterior-facet terms: each cell's facet tensors
   are added to its cell tensors before apply_bc (SystemAssembler.cpp:421-481). P1 only;
   fa_kind / fL_kind = -1 leaves out the facet matrix / vector. */
int32_t orc_assemble_system_facets(int32_t a_kind, int32_t L_kind, int32_t fa_kind,
                                   int32_t fL_kind, int32_t gdim, int64_t num_cells,
                                   const int32_t* cells, const double* x,
                                   const int32_t* cell_dofs, const int32_t* coeff_dofs,
                                   const double* a_c0, const double* L_c0, double alpha,
                                   const double* g, int64_t num_facets, const int32_t* facet_cell,
                                   const int32_t* facet_local, const uint8_t* facet_mask,
                                   const uint8_t* bc_mask, const double* bc_value,
                                   const int32_t* row_ptr, const int32_t* col, double* values,
                                   double* b);

/* ---- adjoint ------------------------------------------------------------------------ */
/* J = u^T K(kappa,x) u - F(f,x)^T u with K the P1 kappa-Poisson matrix and F the P1 load
   vector. Returns J; writes dJ/dkappa [P1 dofs] and dJ/dx [nv x gdim] (zeroed first). */
double orc_adjoint_p1(int32_t gdim, int64_t num_cells, const int32_t* cells, const double* x,
                      int64_t nv, const int32_t* cell_dofs, int64_t ndofs, const double* kappa,
                      const double* f, const double* u, double* dJdkappa, double* dJdx);
/* J = u^T K(mu,lmbda,x) u - F(x)^T u for vector P2 elasticity with body force f3.
   dJ/dmu, dJ/dlmbda on the P1 coefficient dofs, dJ/dx on vertices. */
double orc_adjoint_vp2(int64_t num_cells, const int32_t* cells, const double* x, int64_t nv,
                       const int32_t* cell_dofs, const int32_t* coeff_dofs, int64_t ncoeff,
                       const double* mu, const double* lmbda, const double* f3, const double* u,
                       double* dJdmu, double* dJdlmbda, double* dJdx);

#ifdef __cplusplus
}
#endif
#endif
