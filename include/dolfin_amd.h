/*
 * dolfin_amd -- MI355X (gfx950) cell-integral assembly engine, C ABI.
 *
 * Drop-in for the cell-integral hot path of DOLFIN 2019.1:
 *   dolfin::Assembler::assemble(GenericTensor&, const Form&)   (dolfin/fem/Assembler.h:66)
 *   dolfin::SystemAssembler::assemble(GenericMatrix&, GenericVector&) (SystemAssembler.h:63-84)
 * The host keeps mesh, dofmap and form objects; it hands this library plain arrays
 * (cell -> vertex connectivity, cell -> dof maps, device pointers to coordinates,
 * coefficient vectors and tensor storage) and gets back a device-resident CSR tensor whose
 * structure equals EigenMatrix::init(diagonal_pattern(sorted)) bit for bit
 * (dolfin/la/EigenMatrix.cpp:384-410, dolfin/la/SparsityPattern.cpp:511-542).
 *
 * Conventions
 *   - All indices are int32 (dolfin::la_index, dolfin/common/types.h:32-34); values FP64.
 *   - "h_" arguments are host pointers, "d_" arguments are device pointers.
 *   - stream is a hipStream_t passed as void* (NULL = the null stream).
 *   - Every function returns 0 on success and a negative code on error; the message is
 *     available from dfa_last_error() (thread-local). No exception crosses this boundary
 *     (the reference raises dolfin_error -> std::runtime_error, dolfin/log/Logger.cpp:142-185;
 *     the C++ adapter rethrows).
 *   - Calls on one handle are serialized by the caller; different handles are independent.
 */
#ifndef DOLFIN_AMD_H
#define DOLFIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define DFA_OK 0
#define DFA_ERR_ARG -1      /* invalid argument / unsupported combination */
#define DFA_ERR_HIP -2      /* HIP runtime failure */
#define DFA_ERR_LIMIT -3    /* a size limit of a kernel variant was exceeded */
#define DFA_ERR_STATE -4    /* handle in the wrong state */

/* Elements (the FunctionSpace of both form arguments). */
#define DFA_ELEMENT_P1 1          /* scalar Lagrange degree 1 on intervals/triangles/tets */
#define DFA_ELEMENT_VECTOR_P2 2   /* vector Lagrange degree 2 on tets, 30 dofs per cell   */

/* Form kinds -- the registered counterparts of FFC-generated ufc::cell_integral objects
   (the tabulate_tensor ABI called at dolfin/fem/Assembler.cpp:181-183). */
#define DFA_FORM_P1_POISSON 0       /* inner(kappa*grad(u), grad(v))*dx, kappa P1        */
#define DFA_FORM_P1_LOAD 1          /* f*v*dx, f P1                                       */
#define DFA_FORM_P1_MASS 2          /* u*v*dx                                             */
#define DFA_FORM_VP2_ELASTICITY 3   /* inner(2 mu eps(u) + lmbda tr(eps(u)) I, eps(v))*dx */
#define DFA_FORM_VP2_BODY_FORCE 4   /* inner(f, v)*dx, f a constant 3-vector              */
#define DFA_FORM_P1_FACET_MASS 7    /* alpha*u*v*ds over exterior facets, alpha constant   */
#define DFA_FORM_P1_FACET_LOAD 8    /* g*v*ds over exterior facets, g P1                   */
#define DFA_FORM_VP2_FACET_TRACTION 9 /* inner(t, v)*ds, vector P2, t a constant 3-vector  */

/* Scatter strategies (chosen per call, or by dfa_autotune). */
#define DFA_SCATTER_AUTO 0
#define DFA_SCATTER_ATOMIC 1   /* one thread per cell, FP64 atomics into CSR           */
#define DFA_SCATTER_BLOCKED 2  /* row blocks owned by workgroups, LDS accumulation,
                                  each CSR value written once, no global atomics       */
#define DFA_SCATTER_COLORED 3  /* P1 only: cells in colours that share no dof (greedy
                                  colouring as MeshColoring::color), one launch per
                                  colour, thread per cell, CSR read-add-write, no atomics */

/* Flags of AssemblerBase (dolfin/fem/AssemblerBase.h:46-67). */
#define DFA_ADD_VALUES 1       /* do not zero the tensor before assembly               */

typedef struct dfa_assembler_s* dfa_assembler;

/* Last error message of the calling thread ("" if none). */
const char* dfa_last_error(void);

/* Library version string. */
const char* dfa_version(void);

/*
 * Create an assembler for one (mesh, function space) pair and build the CSR structure once.
 * Replaces AssemblerBase::init_global_tensor's first-call path: SparsityPatternBuilder::build
 * (dolfin/fem/SparsityPatternBuilder.cpp:42-269) + EigenMatrix::init
 * (dolfin/la/EigenMatrix.cpp:384-410).
 *   gdim == tdim in {1,2,3}; h_cells [num_cells x (gdim+1)] vertex ids (Mesh::cells()).
 *   element: DFA_ELEMENT_*; h_cell_dofs [num_cells x ndofs_cell] (DofMap::cell_dofs,
 *     dolfin/fem/DofMap.h:200-206), num_dofs = DofMap::global_dimension().
 *   h_coeff_dofs [num_cells x (gdim+1)]: dofmap of the P1 space holding the coefficients
 *     (kappa, f, mu, lmbda); for P1 forms this is normally h_cell_dofs itself.
 *   h_x [num_vertices x gdim]: coordinates used only to build a locality-preserving work
 *     partition (no numerical use; later calls may pass different coordinates).
 *   keep_diagonal: AssemblerBase::keep_diagonal (SparsityPatternBuilder.cpp:245-264).
 * The assembler copies what it needs; the caller's arrays may be freed afterwards.
 */
int dfa_assembler_create(dfa_assembler* out, int32_t gdim, int64_t num_vertices,
                         int64_t num_cells, const int32_t* h_cells, int32_t element,
                         int32_t ndofs_cell, const int32_t* h_cell_dofs, int64_t num_dofs,
                         const int32_t* h_coeff_dofs, int64_t num_coeff_dofs,
                         const double* h_x, int32_t keep_diagonal, void* stream);

int dfa_assembler_destroy(dfa_assembler h);

/* Tensor layout: rows (= columns) and number of stored entries. GenericMatrix::size, nnz
   (dolfin/la/GenericMatrix.h, EigenMatrix.cpp). */
int dfa_csr_dims(dfa_assembler h, int64_t* num_rows, int64_t* nnz);

/* Copy the device CSR structure (row_ptr [num_rows+1], col [nnz]) into caller buffers
   (device pointers). Equivalent of EigenMatrix::data() (EigenMatrix.cpp:341-355). */
int dfa_csr_structure(dfa_assembler h, int32_t* d_row_ptr, int32_t* d_col, void* stream);

/*
 * Bilinear form -> CSR values (device, [nnz]). Assembler::assemble for a rank-2 form:
 * zero (unless DFA_ADD_VALUES) + assemble_cells (dolfin/fem/Assembler.cpp:105-194).
 *   d_x [num_vertices x gdim]; d_c0, d_c1: P1 coefficient vectors (kappa; mu, lmbda).
 */
int dfa_assemble_matrix(dfa_assembler h, int32_t form, const double* d_x, const double* d_c0,
                        const double* d_c1, double* d_values, int32_t scatter, int32_t flags,
                        void* stream);

/* Linear form -> vector (device, [num_dofs]). f3 (host, 3 doubles) for the body force. */
int dfa_assemble_vector(dfa_assembler h, int32_t form, const double* d_x, const double* d_c0,
                        const double* h_f3, double* d_b, int32_t scatter, int32_t flags,
                        void* stream);

/*
 * SystemAssembler::assemble(A, b) for one square bilinear form a and linear form L, fused in
 * one pass over the cells, with the result of the per-cell symmetric Dirichlet modification
 * of SystemAssembler::apply_bc (dolfin/fem/SystemAssembler.cpp:303-492,1051-1075): when the
 * matrix is assembled without DFA_ADD_VALUES on the blocked scatter, the pass runs without
 * the conditions and a second step applies what the per-cell elimination sums up to to the
 * assembled rows (a Dirichlet row i: cells holding i on the diagonal, 0 elsewhere,
 * b_i = that count times g_i; a free row j: b_j -= A_ji g_i and A_ji = 0); otherwise (and
 * with the environment variable DFA_BC_INLINE=1) every cell tensor is modified in the pass.
 * The two agree up to rounding.
 *   d_bc_mask [num_dofs] uint8 (nonzero = Dirichlet dof), d_bc_value [num_dofs] (may be NULL
 *   when there are no boundary conditions). d_values or d_b may be NULL (assemble(A) or
 *   assemble(b) only). flags: DFA_ADD_VALUES adds to A and b instead of overwriting them
 *   (AssemblerBase::add_values, which SystemAssembler honours through init_global_tensor,
 *   dolfin/fem/AssemblerBase.cpp:137-285).
 */
int dfa_assemble_system(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                        const double* d_a_c0, const double* d_a_c1, const double* d_L_c0,
                        const double* h_f3, const uint8_t* d_bc_mask, const double* d_bc_value,
                        double* d_values, double* d_b, int32_t scatter, int32_t flags,
                        void* stream);

/*
 * SystemAssembler::assemble(A, b, x0) / assemble(b, x0) (dolfin/fem/SystemAssembler.h:79,84):
 * the same pass as dfa_assemble_system for an incremental (Newton) solve. Every Dirichlet
 * value g becomes x0[dof] - g before the cell loop (SystemAssembler.cpp:252-275), so the
 * solution increment du of A du = b makes x0 - du meet the boundary condition.
 *   d_x0 [num_dofs] (device). Without Dirichlet data (d_bc_mask NULL) x0 changes nothing.
 * The shifted values live in a buffer of the handle (allocated on the first call).
 */
int dfa_assemble_system_x0(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                           const double* d_a_c0, const double* d_a_c1, const double* d_L_c0,
                           const double* h_f3, const uint8_t* d_bc_mask, const double* d_bc_value,
                           const double* d_x0, double* d_values, double* d_b, int32_t scatter,
                           int32_t flags, void* stream);

/*
 * Allocation-free variant of dfa_assemble_system_x0 (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture (a Newton loop
 * captured once and replayed), tight solver loops, harnesses that forbid hipMalloc. Same
 * arguments and results; the shifted Dirichlet values live in a fixed device array of the
 * library (DFA_NOALLOC_MAX_DOFS doubles = 32 MiB, held by each GPU the library runs on until
 * the process ends) instead of a buffer allocated on the handle's first x0 call. Handles with
 * more dofs are refused (DFA_ERR_LIMIT); so is the coloured scatter before its colour plan
 * exists (DFA_ERR_STATE: it is built, in device memory, on first use). The call performs no
 * device allocation. All calls on one GPU share the array: calls that overlap on one GPU
 * (different streams or threads) race and must be serialised by the caller.
 */
#define DFA_NOALLOC_MAX_DOFS 4194304
int dfa_assemble_system_x0_noalloc(dfa_assembler h, int32_t a_form, int32_t L_form,
                                   const double* d_x, const double* d_a_c0, const double* d_a_c1,
                                   const double* d_L_c0, const double* h_f3,
                                   const uint8_t* d_bc_mask, const double* d_bc_value,
                                   const double* d_x0, double* d_values, double* d_b,
                                   int32_t scatter, int32_t flags, void* stream);

/* The boundary-value modification of SystemAssembler.cpp:252-275 on its own, for callers
   that pass the modified values to several passes (cell and exterior-facet):
   d_out[i] = d_bc_mask[i] ? d_x0[i] - d_bc_value[i] : 0, i < n (device arrays). */
int dfa_dirichlet_x0(int64_t n, const uint8_t* d_bc_mask, const double* d_bc_value,
                     const double* d_x0, double* d_out, void* stream);

/* Introspection for tests and benchmarks (no counterpart in the reference): the name and
   template arguments of the kernel dfa_assemble_system (a_form, L_form >= 0), or
   dfa_assemble_matrix (L_form -1) / dfa_assemble_vector (a_form -1), would launch for this
   handle, with or without Dirichlet data, written to name[len]. Launches nothing. */
int dfa_pass_kernel(dfa_assembler h, int32_t a_form, int32_t L_form, int32_t has_bc,
                    int32_t scatter, char* name, int32_t len);

/* Introspection (no counterpart in the reference): where the handle's block plan (P1, or
   vector P2 with its rank table) was built, *on_device = 1 (the device planner) or 0 (the
   host builder, or no plan). With
   DFA_PLAN_VERIFY=1 in the environment at creation the plan is built both ways and compared
   element by element; a difference makes this call fail (DFA_ERR_STATE) with the first
   differing array named in dfa_last_error(). */
int dfa_plan_builder(dfa_assembler h, int32_t* on_device);

/*
 * Adjoint (reverse-mode) assembly of J = u^T K(c, x) u - F(x)^T u without forming K.
 * Geometry is recomputed per cell. Outputs are overwritten (not accumulated).
 *   P1 Poisson:  c0 = kappa, F = P1 load of d_f;  d_dJdc0 = dJ/dkappa [num_coeff_dofs],
 *                d_dJdc1 unused (NULL).
 *   vector P2 elasticity: c0 = mu, c1 = lmbda, F = body force h_f3;
 *                d_dJdc0 = dJ/dmu, d_dJdc1 = dJ/dlmbda.
 *   d_dJdx [num_vertices x gdim] = dJ/d(vertex coordinates); d_J (1 double) = J.
 * Counterpart of assembling UFL derivative(J, kappa) / derivative(J, SpatialCoordinate)
 * (python/test/unit/fem/test_coordinate_derivative.py:29-63).
 */
int dfa_adjoint(dfa_assembler h, int32_t a_form, const double* d_x, const double* d_c0,
                const double* d_c1, const double* d_f, const double* h_f3, const double* d_u,
                double* d_J, double* d_dJdc0, double* d_dJdc1, double* d_dJdx, void* stream);

/* Plan parameters for assemblers created afterwards by the calling thread (0 = built-in
   default): P1 rows per block, P1 workgroup size (256 or 512), vector-P2 LDS values per
   block. Used by the autotuner; no counterpart in the reference. */
int dfa_set_plan_defaults(int32_t p1_block_rows, int32_t p1_threads, int32_t p2_block_vals);

/*
 * Batched assembly (no counterpart in the reference, which assembles one form per
 * Assembler::assemble call): n independent assemblers, each with one bilinear form
 * (DFA_FORM_P1_POISSON on tetrahedra or DFA_FORM_VP2_ELASTICITY), assembled by two kernels
 * (vector-P2 members, then P1 members) whose launches are captured once in a hipGraph. Every
 * member needs its blocked plan (built at creation when coordinates are given). d_x, d_c0,
 * d_c1 (lmbda, P2 only) and d_values are per member and are read/written at every launch, so
 * coefficients may change between launches. grid_per_cu = 0 launches one workgroup per row
 * block; grid_per_cu > 0 launches persistent workgroups, at most that many per CU.
 */
typedef struct dfa_batch_s* dfa_batch;
int dfa_batch_create(dfa_batch* out, int32_t n, const dfa_assembler* hs, const int32_t* forms,
                     const double* const* d_x, const double* const* d_c0,
                     const double* const* d_c1, double* const* d_values, int32_t grid_per_cu,
                     void* stream);
/* Replay the captured launch on `stream`. */
int dfa_batch_launch(dfa_batch b, void* stream);
int dfa_batch_info(dfa_batch b, int64_t* units, int32_t* grid, int32_t* lds_bytes, int64_t* cells);
int dfa_batch_destroy(dfa_batch b);

/* Introspection for benchmarks: algorithmic bytes of one matrix assembly (the compulsory
   traffic: connectivity, coordinates, coefficients, CSR values written once, col/row_ptr
   read once) and the bytes the plan of the chosen strategy actually stores. */
int dfa_plan_info(dfa_assembler h, int32_t scatter, int64_t* plan_bytes, int64_t* num_blocks,
                  int64_t* cell_appearances, int32_t* lds_bytes);

/*
 * Reverse Cuthill-McKee renumbering of the nodes of a mesh's cells (two nodes are adjacent
 * when they share a cell): old_to_new[node] receives the node's new index. This is the
 * numbering DOLFIN gives the dofmap nodes with parameters["reorder_dofs_serial"] = true and
 * the Boost ordering library: DofMapBuilder::compute_node_reordering (dolfin/fem/
 * DofMapBuilder.cpp:1473-1569) builds this graph and BoostGraphOrdering::compute_cuthill_mckee
 * (dolfin/graph/BoostGraphOrdering.cpp:34-75) orders it. Host computation, no device use.
 */
int dfa_node_reordering(int64_t num_cells, int32_t nodes_per_cell, const int32_t* cell_nodes,
                        int64_t num_nodes, int32_t* old_to_new);

/*
 * The same ordering of a graph given as sorted adjacency lists without self loops
 * (graph_ptr[num_nodes + 1], graph_adj): BoostGraphOrdering::compute_cuthill_mckee(graph,
 * true) for a dolfin::Graph (BoostGraphOrdering.cpp:34-75). Host computation.
 */
int dfa_graph_reordering(int64_t num_nodes, const int32_t* graph_ptr, const int32_t* graph_adj,
                         int32_t* old_to_new);

/*
 * Exterior facets of a P1 handle (Facet::exterior: one incident cell), sorted by cell and
 * local facet index (local facet k lies opposite local vertex k). Writes their number and,
 * when the arrays are non-NULL, each facet's cell and local index -- the order d_facet_mask
 * below uses (the counterpart of a MeshFunction over the exterior facets). Built on first
 * use and kept in the handle.
 */
int dfa_exterior_facets(dfa_assembler h, int64_t* num_facets, int32_t* facet_cell,
                        int32_t* facet_local);

/*
 * Exterior-facet integrals of P1 forms, ADDED to d_values and d_b: replaces
 * Assembler::assemble_exterior_facets (dolfin/fem/Assembler.cpp:196-298) and the facet terms
 * SystemAssembler merges into each cell tensor (dolfin/fem/SystemAssembler.cpp:421-481).
 * a_form: DFA_FORM_P1_FACET_MASS (alpha*u*v*ds) or -1; L_form: DFA_FORM_P1_FACET_LOAD
 * (g*v*ds, d_g on the coefficient dofs) or -1. A vector-P2 handle takes a_form -1 and
 * L_form DFA_FORM_VP2_FACET_TRACTION (inner(t, v)*ds, the Neumann term of elasticity) with
 * d_g pointing to the three traction components on the device. d_facet_mask (NULL = every exterior facet)
 * selects facets as an exterior-facet subdomain does. With d_bc_mask / d_bc_value the pass
 * completes dfa_assemble_system with the same Dirichlet data: Dirichlet rows are left as the
 * cell pass wrote them and Dirichlet columns are lifted into d_b. d_values may be NULL (the
 * lifting of the facet matrix into d_b is still applied); d_b may be NULL.
 */
int dfa_assemble_facets(dfa_assembler h, int32_t a_form, int32_t L_form, const double* d_x,
                        double alpha, const double* d_g, const uint8_t* d_facet_mask,
                        const uint8_t* d_bc_mask, const double* d_bc_value, double* d_values,
                        double* d_b, void* stream);

/*
 * d_y = alpha * A d_x + beta * d_y for the matrix with the handle's CSR structure and values
 * d_values (device): EigenMatrix::mult (dolfin/la/EigenMatrix.cpp:230-255) without leaving
 * HBM. beta = 0 does not read d_y.
 */
int dfa_csr_spmv(dfa_assembler h, const double* d_values, double alpha, const double* d_x,
                 double beta, double* d_y, void* stream);

/*
 * GenericMatrix::add_local(block, rows) from the host (dolfin/la/GenericMatrix.h:106-112,
 * EigenMatrix::add, dolfin/la/EigenMatrix.cpp:156-166): d_values[(h_rows[i], h_cols[j])] +=
 * h_block[i*n + j] for the dense m x n block. Every entry must be in the handle's sparsity
 * pattern (the pattern is fixed once built; nothing is inserted): otherwise the call fails and
 * changes nothing. Synchronous (returns after the addition is done on `stream`).
 */
int dfa_csr_add_local(dfa_assembler h, double* d_values, int32_t m, const int32_t* h_rows,
                      int32_t n, const int32_t* h_cols, const double* h_block, void* stream);

/* GenericVector::add_local(block, m, rows) from the host (EigenVector::add_local,
   dolfin/la/EigenVector.cpp:160-165): d_b[h_rows[i]] += h_block[i], 0 <= h_rows[i] < size.
   Synchronous. */
int dfa_vector_add_local(double* d_b, int64_t size, int32_t m, const int32_t* h_rows,
                         const double* h_block, void* stream);

/* The diagonal of that matrix (0 where the pattern has no diagonal entry): the Jacobi
   preconditioner of EigenKrylovSolver (dolfin/la/EigenKrylovSolver.cpp:52-56). */
int dfa_csr_diagonal(dfa_assembler h, const double* d_values, double* d_diag, void* stream);

/*
 * Conjugate gradients on the device for the matrix d_values on the handle's structure: the
 * "cg" method of EigenKrylovSolver with the "jacobi" preconditioner (jacobi != 0: diag(A)^-1,
 * 1 where the diagonal is 0) or none (dolfin/la/EigenKrylovSolver.cpp:43-56, reached through
 * dolfin::solve(A, x, b, "cg", "jacobi")). d_x (device) holds the initial guess and receives
 * the solution; d_b is not modified. Stops at the first iteration k with
 * |r_k| <= max(rtol |b|, atol) and sets *iterations = k (0 when the initial residual already
 * satisfies it). Fails with DFA_ERR_STATE on breakdown (p.Ap = 0 with r != 0) or when maxiter
 * iterations (maxiter < 0: 10 x rows) do not converge, *iterations then holding the iterations
 * done. The convergence state lives on the device and is read every check_every iterations;
 * the iterate is frozen from the converged iteration on, so neither the answer nor the count
 * depends on check_every. Reductions are summed in a fixed order (deterministic). The work
 * vectors (4 x rows doubles) are allocated on the handle's first solve and kept. Synchronous.
 */
int dfa_cg_solve(dfa_assembler h, const double* d_values, double* d_x, const double* d_b, int32_t jacobi,
                 double rtol, double atol, int64_t maxiter, int32_t check_every, int64_t* iterations,
                 void* stream);

/* Device properties the roofline figures rest on (no counterpart in the reference):
   info[0] compute units, [1] peak engine clock (kHz), [2] peak memory clock (kHz),
   [3] memory bus width (bits), [4] L2 size (bytes), [5] device memory (bytes). */
int dfa_device_info(int32_t device, int64_t* info);

/*
 * Greedy colouring of mesh cells: MeshColoring::color(mesh, {tdim, entity_dim, tdim}) /
 * MeshColoring::color_cells(mesh, "vertex" | "edge" | "facet") (dolfin/mesh/
 * MeshColoring.cpp:45-121). Two cells are neighbours when they share an entity of dimension
 * entity_dim (GraphBuilder::local_graph(mesh, tdim, entity_dim), dolfin/graph/
 * GraphBuilder.cpp:124-153); cells are visited in index order and each takes the smallest
 * colour no coloured neighbour holds (Boost's sequential_vertex_coloring, called from
 * dolfin/graph/BoostGraphColoring.h:87-113). Simplicial meshes: h_cells [num_cells x nvc]
 * vertex ids, nvc = tdim + 1, 0 <= entity_dim <= tdim. Writes h_colors [num_cells] and the
 * number of colours. Host computation.
 */
int dfa_mesh_coloring(int64_t num_cells, int32_t nvc, const int32_t* h_cells, int64_t num_vertices,
                      int32_t entity_dim, int32_t* h_colors, int32_t* num_colors);

/*
 * Colours of a P1 handle's coloured scatter (DFA_SCATTER_COLORED; built on first use): cells
 * sharing a dof get different colours. Writes the number of colours and, when h_color_ptr is
 * non-NULL (room for num_colors + 1 entries), the offsets of each colour's cells.
 */
int dfa_color_info(dfa_assembler h, int32_t* num_colors, int64_t* h_color_ptr, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DOLFIN_AMD_H */
