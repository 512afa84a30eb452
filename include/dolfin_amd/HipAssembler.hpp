// C++ host adapter over the dolfin_amd C ABI (include/dolfin_amd.h), header-only.
//
// Mirrors the DOLFIN 2019.1 assembler interface for the cell-integral path so that a
// maintainer can route dolfin::Assembler::assemble(GenericTensor&, const Form&)
// (dolfin/fem/Assembler.h:66) and SystemAssembler::assemble(A, b) (SystemAssembler.h:63-84)
// to the GPU:
//   HipAssembler         add_values / finalize_tensor / keep_diagonal flags
//                        (dolfin/fem/AssemblerBase.h:46-67), assemble(tensor, form)
//   HipSystemAssembler   (a, L, Dirichlet values) -> assemble(A, b) / (A) / (b) and the
//                        incremental assemble(A, b, x0) / (b, x0) (SystemAssembler.h:63-84),
//                        with the same three flags
//   TensorLayout         the sparsity a tensor is initialised from (GenericTensor::init,
//                        dolfin/la/GenericTensor.h:58; dolfin/la/TensorLayout.h)
//   DeviceCSRMatrix      the GenericMatrix subset the assembler uses: init(TensorLayout), empty,
//                        size, local_range, nnz, zero, add_local, apply, norm("frobenius"),
//                        data(), mult, diagonal (dolfin/la/GenericMatrix.h:56-255,
//                        EigenMatrix.cpp:77-105, 156-186, 230-255, 341-355, 384-420)
//   DeviceVector         GenericVector subset: size, local_range, zero, add_local, apply,
//                        norm("l2"), get_local
// Errors from the library are rethrown as std::runtime_error with the library's message,
// like dolfin_error (dolfin/log/Logger.cpp:142-185).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../dolfin_amd.h"

namespace dolfin_amd {

inline void check(int rc) {
  if (rc != DFA_OK) throw std::runtime_error(std::string("dolfin_amd: ") + dfa_last_error());
}

inline void hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owned device allocation.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() { release(); }
  void resize(size_t n) {
    release();
    n_ = n;
    if (n) hip(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
  }
  void upload(const T* h, size_t n, hipStream_t s = nullptr) {
    if (n != n_) resize(n);
    if (n) hip(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  }
  std::vector<T> download(hipStream_t s = nullptr) const {
    std::vector<T> h(n_);
    if (n_) hip(hipMemcpyAsync(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    return h;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// One (mesh, function space) pair: the engine handle with the CSR structure and plans.
// The caller passes what DOLFIN's Mesh / DofMap hold: Mesh::cells(), DofMap::cell_dofs
// (flattened), the coefficient space's cell dofs and the coordinates (used for the work
// partition only).
class Space {
 public:
  Space(int32_t gdim, int64_t num_vertices, const std::vector<int32_t>& cells, int32_t element,
        int32_t ndofs_cell, const std::vector<int32_t>& cell_dofs, int64_t num_dofs,
        const std::vector<int32_t>& coeff_dofs, int64_t num_coeff_dofs,
        const std::vector<double>& x, bool keep_diagonal = false, hipStream_t s = nullptr)
      : gdim_(gdim), element_(element), num_vertices_(num_vertices) {
    const int64_t nc = static_cast<int64_t>(cells.size()) / (gdim + 1);
    dfa_assembler h = nullptr;
    check(dfa_assembler_create(&h, gdim, num_vertices, nc, cells.data(), element, ndofs_cell,
                               cell_dofs.data(), num_dofs, coeff_dofs.data(), num_coeff_dofs,
                               x.data(), keep_diagonal ? 1 : 0, s));
    h_.reset(h, [](dfa_assembler p) { dfa_assembler_destroy(p); });
    check(dfa_csr_dims(h, &num_rows_, &nnz_));
  }
  dfa_assembler handle() const { return h_.get(); }
  int64_t num_rows() const { return num_rows_; }
  int64_t nnz() const { return nnz_; }
  int32_t gdim() const { return gdim_; }
  int32_t element() const { return element_; }
  int64_t num_vertices() const { return num_vertices_; }

 private:
  std::shared_ptr<dfa_assembler_s> h_;
  int32_t gdim_, element_;
  int64_t num_vertices_;
  int64_t num_rows_ = 0, nnz_ = 0;
};

// A registered cell-integral form: its kind, space and device coefficient vectors (the
// counterpart of dolfin::Form with an FFC-generated ufc::form, dolfin/fem/Form.h:85).
struct Form {
  std::shared_ptr<const Space> space;
  int32_t kind = -1;              // DFA_FORM_*
  int32_t rank = 2;               // 2 bilinear, 1 linear
  const double* x = nullptr;      // device coordinates [num_vertices x gdim]
  const double* c0 = nullptr;     // kappa / f / mu
  const double* c1 = nullptr;     // lmbda
  double f3[3] = {0, 0, 0};       // constant body force
};

// The layout a matrix is initialised from: DOLFIN's TensorLayout holds the index maps and the
// sparsity pattern (dolfin/la/TensorLayout.h, built by AssemblerBase::init_global_tensor,
// dolfin/fem/AssemblerBase.cpp:137-285); here the Space holds the device CSR structure built
// once from the dofmap, and the layout names it.
struct TensorLayout {
  std::shared_ptr<const Space> space;
  explicit TensorLayout(std::shared_ptr<const Space> s) : space(std::move(s)) {}
  int64_t size(int) const { return space->num_rows(); }
};

class DeviceCSRMatrix {
 public:
  bool empty() const { return space_ == nullptr; }
  // GenericTensor::init(const TensorLayout&) (dolfin/la/GenericTensor.h:58): device values
  // for the layout's pattern, zeroed (EigenMatrix::init, EigenMatrix.cpp:384-410).
  void init(const TensorLayout& layout) { init(layout.space); }
  void init(std::shared_ptr<const Space> space) {
    space_ = std::move(space);
    values_.resize(static_cast<size_t>(space_->nnz()));
    zero();
  }
  const Space* space() const { return space_.get(); }
  int64_t size(int) const { return space_ ? space_->num_rows() : 0; }
  // GenericTensor::local_range: one process owns every row (serial DOLFIN semantics).
  std::pair<int64_t, int64_t> local_range(int dim) const { return {0, size(dim)}; }
  int64_t nnz() const { return space_ ? space_->nnz() : 0; }
  void zero(hipStream_t s = nullptr) {
    if (values_.size()) hip(hipMemsetAsync(values_.get(), 0, values_.size() * sizeof(double), s), "hipMemsetAsync");
  }
  // GenericMatrix::add_local(block, rows) (GenericMatrix.h:106-112) from host data: the dense
  // block[i * cols.size() + j] is added at (rows[i], cols[j]); every entry must be in the
  // pattern. Synchronous. Assembly itself never calls this: it scatters on the device.
  void add_local(const double* block, const std::vector<int32_t>& rows,
                 const std::vector<int32_t>& cols, hipStream_t s = nullptr) {
    check(dfa_csr_add_local(space_->handle(), values_.get(), static_cast<int32_t>(rows.size()),
                            rows.data(), static_cast<int32_t>(cols.size()), cols.data(), block, s));
  }
  void apply(const std::string&) {}
  double* values() { return values_.get(); }
  double norm(const std::string& type) const {
    if (type != "frobenius") throw std::runtime_error("only the Frobenius norm is provided");
    double s = 0.0;
    for (double v : values_.download()) s += v * v;
    return std::sqrt(s);
  }
  // y = alpha A x + beta y on the device (EigenMatrix::mult, EigenMatrix.cpp:230-255);
  // x and y hold size(0) doubles in device memory.
  void mult(const double* x, double* y, double alpha = 1.0, double beta = 0.0, hipStream_t s = nullptr) const {
    check(dfa_csr_spmv(space_->handle(), values_.get(), alpha, x, beta, y, s));
  }
  // The diagonal into d (size(0) doubles, device memory): the Jacobi preconditioner's input.
  void diagonal(double* d, hipStream_t s = nullptr) const {
    check(dfa_csr_diagonal(space_->handle(), values_.get(), d, s));
  }
  // Conjugate gradients A x = b on the device (x: initial guess in, solution out; both
  // size(0) doubles in device memory), dolfin::solve(A, x, b, "cg", "jacobi" | "none"); returns
  // the converged iteration, throws on breakdown or when maxiter iterations do not converge.
  int64_t solve_cg(double* x, const double* b, bool jacobi = true, double rtol = 1e-10, double atol = 0.0,
                   int64_t maxiter = -1, hipStream_t s = nullptr) const {
    int64_t its = 0;
    check(dfa_cg_solve(space_->handle(), values_.get(), x, b, jacobi ? 1 : 0, rtol, atol, maxiter, 10, &its, s));
    return its;
  }
  // CSR export (EigenMatrix::data): row_ptr, col, values on the host.
  void data(std::vector<int32_t>& row_ptr, std::vector<int32_t>& col, std::vector<double>& vals) const {
    DeviceArray<int32_t> rp(static_cast<size_t>(space_->num_rows() + 1)), cc(static_cast<size_t>(space_->nnz()));
    check(dfa_csr_structure(space_->handle(), rp.get(), cc.get(), nullptr));
    row_ptr = rp.download();
    col = cc.download();
    vals = values_.download();
  }

 private:
  std::shared_ptr<const Space> space_;
  DeviceArray<double> values_;
};

class DeviceVector {
 public:
  bool empty() const { return values_.size() == 0; }
  void init(int64_t n) {
    values_.resize(static_cast<size_t>(n));
    zero();
  }
  int64_t size() const { return static_cast<int64_t>(values_.size()); }
  std::pair<int64_t, int64_t> local_range() const { return {0, size()}; }
  void zero(hipStream_t s = nullptr) {
    if (values_.size()) hip(hipMemsetAsync(values_.get(), 0, values_.size() * sizeof(double), s), "hipMemsetAsync");
  }
  // GenericVector::add_local(block, m, rows) (EigenVector.cpp:160-165) from host data.
  void add_local(const double* block, const std::vector<int32_t>& rows, hipStream_t s = nullptr) {
    check(dfa_vector_add_local(values_.get(), size(), static_cast<int32_t>(rows.size()), rows.data(),
                               block, s));
  }
  void apply(const std::string&) {}
  double* values() { return values_.get(); }
  const double* values() const { return values_.get(); }
  std::vector<double> get_local() const { return values_.download(); }
  double norm(const std::string& type) const {
    if (type != "l2") throw std::runtime_error("only the l2 norm is provided");
    double s = 0.0;
    for (double v : values_.download()) s += v * v;
    return std::sqrt(s);
  }

 private:
  DeviceArray<double> values_;
};

// Exterior-facet terms of P1 forms (Assembler::assemble_exterior_facets,
// dolfin/fem/Assembler.cpp:196-298): alpha*u*v*ds in the bilinear form and g*v*ds in the
// linear form, over the facets selected by a device mask in dfa_exterior_facets order (an
// exterior-facet subdomain; NULL = all exterior facets).
struct FacetTerms {
  bool robin = false;                 // alpha*u*v*ds present (P1)
  double alpha = 0.0;
  const double* g = nullptr;          // g on the coefficient dofs (device); NULL: no g*v*ds (P1)
  const double* traction = nullptr;   // 3 components (device): inner(t, v)*ds (vector P2)
  const uint8_t* facet_mask = nullptr;
  bool empty() const { return !robin && !g && !traction; }
  int32_t a_form() const { return robin ? DFA_FORM_P1_FACET_MASS : -1; }
  int32_t L_form() const {
    return traction ? DFA_FORM_VP2_FACET_TRACTION : (g ? DFA_FORM_P1_FACET_LOAD : -1);
  }
  const double* L_data() const { return traction ? traction : g; }
};

// Exterior facets of a space: (cell, local facet) pairs, the order facet masks use.
inline void exterior_facets(const Space& V, std::vector<int32_t>& facet_cell,
                            std::vector<int32_t>& facet_local) {
  int64_t n = 0;
  check(dfa_exterior_facets(V.handle(), &n, nullptr, nullptr));
  facet_cell.resize(static_cast<size_t>(n));
  facet_local.resize(static_cast<size_t>(n));
  check(dfa_exterior_facets(V.handle(), &n, facet_cell.data(), facet_local.data()));
}

// dolfin::AssemblerBase flags (dolfin/fem/AssemblerBase.h:46-67).
struct AssemblerFlags {
  bool add_values = false;
  bool finalize_tensor = true;
  bool keep_diagonal = false;  // fixed when the Space is created
};

class HipAssembler : public AssemblerFlags {
 public:
  explicit HipAssembler(int32_t scatter = DFA_SCATTER_AUTO, hipStream_t s = nullptr)
      : scatter_(scatter), stream_(s) {}

  // Assembler::assemble for a rank-2 form: structure on first use, then zero + cell loop.
  void assemble(DeviceCSRMatrix& A, const Form& a) {
    if (a.rank != 2) throw std::runtime_error("matrix assembly needs a bilinear form");
    if (A.empty() || A.space() != a.space.get()) A.init(a.space);
    check(dfa_assemble_matrix(a.space->handle(), a.kind, a.x, a.c0, a.c1, A.values(), scatter_,
                              add_values ? DFA_ADD_VALUES : 0, stream_));
    if (finalize_tensor) A.apply("add");
  }

  void assemble(DeviceVector& b, const Form& L) {
    if (L.rank != 1) throw std::runtime_error("vector assembly needs a linear form");
    if (b.empty() || b.size() != L.space->num_rows()) b.init(L.space->num_rows());
    check(dfa_assemble_vector(L.space->handle(), L.kind, L.x, L.c0, L.f3, b.values(), scatter_,
                              add_values ? DFA_ADD_VALUES : 0, stream_));
    if (finalize_tensor) b.apply("add");
  }

  // Cell integrals, then exterior-facet terms (the order of Assembler::assemble,
  // dolfin/fem/Assembler.cpp:52-103).
  void assemble(DeviceCSRMatrix& A, const Form& a, const FacetTerms& ds) {
    const bool fin = finalize_tensor;
    finalize_tensor = false;
    assemble(A, a);
    finalize_tensor = fin;
    if (ds.robin)
      check(dfa_assemble_facets(a.space->handle(), DFA_FORM_P1_FACET_MASS, -1, a.x, ds.alpha,
                                nullptr, ds.facet_mask, nullptr, nullptr, A.values(), nullptr,
                                stream_));
    if (finalize_tensor) A.apply("add");
  }

  void assemble(DeviceVector& b, const Form& L, const FacetTerms& ds) {
    const bool fin = finalize_tensor;
    finalize_tensor = false;
    assemble(b, L);
    finalize_tensor = fin;
    if (ds.L_form() >= 0)
      check(dfa_assemble_facets(L.space->handle(), -1, ds.L_form(), L.x, 0.0, ds.L_data(),
                                ds.facet_mask, nullptr, nullptr, nullptr, b.values(), stream_));
    if (finalize_tensor) b.apply("add");
  }

 private:
  int32_t scatter_;
  hipStream_t stream_;
};

// SystemAssembler with one set of Dirichlet values (bc_mask/bc_value on the device, one
// entry per dof), applied per cell as in SystemAssembler::apply_bc. The AssemblerBase flags
// add_values / finalize_tensor / keep_diagonal apply as in the reference (keep_diagonal is
// fixed when the Space is created).
class HipSystemAssembler : public AssemblerFlags {
 public:
  HipSystemAssembler(Form a, Form L, const uint8_t* d_bc_mask = nullptr,
                     const double* d_bc_value = nullptr, int32_t scatter = DFA_SCATTER_AUTO,
                     hipStream_t s = nullptr, FacetTerms ds = FacetTerms())
      : a_(std::move(a)), L_(std::move(L)), mask_(d_bc_mask), value_(d_bc_value),
        scatter_(scatter), stream_(s), ds_(ds) {
    if (a_.space != L_.space) throw std::runtime_error("a and L must share the function space");
    if ((mask_ == nullptr) != (value_ == nullptr))
      throw std::runtime_error("Dirichlet mask and values go together");
  }

  void assemble(DeviceCSRMatrix& A, DeviceVector& b) { run(&A, &b, nullptr); }
  void assemble(DeviceCSRMatrix& A) { run(&A, nullptr, nullptr); }
  void assemble(DeviceVector& b) { run(nullptr, &b, nullptr); }
  // Incremental variants (SystemAssembler.h:79,84): Dirichlet values become x0[dof] - g
  // (SystemAssembler.cpp:252-275).
  void assemble(DeviceCSRMatrix& A, DeviceVector& b, const DeviceVector& x0) { run(&A, &b, &x0); }
  void assemble(DeviceVector& b, const DeviceVector& x0) { run(nullptr, &b, &x0); }

 private:
  void run(DeviceCSRMatrix* A, DeviceVector* b, const DeviceVector* x0) {
    if (A && (A->empty() || A->space() != a_.space.get())) A->init(a_.space);
    if (b && (b->empty() || b->size() != a_.space->num_rows())) b->init(a_.space->num_rows());
    if (x0 && x0->size() != a_.space->num_rows())
      throw std::runtime_error("x0 must have one entry per dof");
    const int32_t flags = add_values ? DFA_ADD_VALUES : 0;
    const double* bcv = value_;
    if (x0 && mask_) {
      // the cell pass and the facet pass use the same shifted values
      shifted_.resize(static_cast<size_t>(a_.space->num_rows()));
      check(dfa_dirichlet_x0(a_.space->num_rows(), mask_, value_, x0->values(), shifted_.get(),
                             stream_));
      bcv = shifted_.get();
    }
    check(dfa_assemble_system(a_.space->handle(), a_.kind, L_.kind, a_.x, a_.c0, a_.c1, L_.c0,
                              L_.f3, mask_, bcv, A ? A->values() : nullptr,
                              b ? b->values() : nullptr, scatter_, flags, stream_));
    // exterior-facet tensors join their cell's tensor before apply_bc
    // (dolfin/fem/SystemAssembler.cpp:421-481)
    if (!ds_.empty())
      check(dfa_assemble_facets(a_.space->handle(), ds_.a_form(), ds_.L_form(), a_.x, ds_.alpha,
                                ds_.L_data(), ds_.facet_mask, mask_, bcv,
                                A ? A->values() : nullptr, b ? b->values() : nullptr, stream_));
    if (finalize_tensor) {
      if (A) A->apply("add");
      if (b) b->apply("add");
    }
  }
  Form a_, L_;
  const uint8_t* mask_;
  const double* value_;
  int32_t scatter_;
  hipStream_t stream_;
  FacetTerms ds_;
  DeviceArray<double> shifted_;
};

}  // namespace dolfin_amd
