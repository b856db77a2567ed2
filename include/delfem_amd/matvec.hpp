// delfem_amd/matvec.hpp -- C++ classes with the names and semantics of
// DelFEM's MatVec layer over the dfm.h C-ABI: block vectors, Dirichlet flags,
// the block-CRS matrix with separate diagonal and its ILU(k) factor.  Values
// live in device memory; host arrays are copied in and out explicitly.
//
//   MatVec::CVector_Blk        include/delfem/matvec/vector_blk.h:41-215
//   MatVec::CBCFlag            include/delfem/matvec/bcflag_blk.h:36-140
//   MatVec::CMatDia_BlkCrs     include/delfem/matvec/matdia_blkcrs.h:44-118
//   MatVec::CMatDiaFrac_BlkCrs include/delfem/matvec/matdiafrac_blkcrs.h:40-137
//   MatVec::CPrecond_Blk       include/delfem/matvec/matprecond_blk.h:37-41
//   MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
//                              include/delfem/matvec/solver_mat_iter.h:41-60
//
// The classes are declared in namespace delfem_amd::MatVec; unless
// DELFEM_AMD_NO_ALIASES is defined, namespace MatVec aliases it so that code
// written against the reference compiles unchanged.  Failures of the library
// (bad arguments, device errors) throw delfem_amd::Error; numerical outcomes
// are returned as bool like the reference.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../dfm.h"

namespace delfem_amd {

struct Error : std::runtime_error {
  int code;
  Error(const char* what, int rc) : std::runtime_error(std::string(what) + ": " + dfm_last_error()), code(rc) {}
};

// Throws on library failures; passes DFM_OK, DFM_NOT_CONVERGED and DFM_SINGULAR through.
inline int check(int rc, const char* what) {
  if (rc < 0) throw Error(what, rc);
  return rc;
}

namespace MatVec {

class CMatDia_BlkCrs;

class CVector_Blk {
 public:
  CVector_Blk(unsigned nblk, unsigned len) : nblk_(nblk), len_(len) {
    check(dfm_malloc(reinterpret_cast<void**>(&v_), sizeof(double) * size()), "CVector_Blk");
    SetVectorZero();
  }
  // a view of nblk*len device doubles owned elsewhere (one segment of a
  // multi-segment system's vector); it does not free them
  CVector_Blk(unsigned nblk, unsigned len, double* view) : nblk_(nblk), len_(len), v_(view), owns_(false) {}
  ~CVector_Blk() {
    if (v_ && owns_) dfm_free(v_);
  }
  CVector_Blk(const CVector_Blk&) = delete;
  CVector_Blk(CVector_Blk&& o) noexcept : nblk_(o.nblk_), len_(o.len_), v_(o.v_), owns_(o.owns_) { o.v_ = nullptr; }
  // operator= copies the values (vector_blk.cpp:62-71)
  CVector_Blk& operator=(const CVector_Blk& rhs) {
    check(dfm_vec_copy(size(), rhs.v_, v_), "CVector_Blk::operator=");
    return *this;
  }
  CVector_Blk& operator*=(double d) {
    check(dfm_vec_scal(size(), d, v_), "CVector_Blk::operator*=");
    return *this;
  }
  CVector_Blk& AXPY(const double& alpha, const CVector_Blk& rhs) {
    check(dfm_vec_axpy(size(), alpha, rhs.v_, v_), "CVector_Blk::AXPY");
    return *this;
  }
  // operator+= (vector_blk.cpp:73-81): y + 1.0 * x is y + x exactly
  CVector_Blk& operator+=(const CVector_Blk& rhs) { return AXPY(1.0, rhs); }
  void SetVectorZero() { check(dfm_vec_fill(size(), 0.0, v_), "CVector_Blk::SetVectorZero"); }
  double GetSquaredVectorNorm() const { return (*this) * (*this); }
  unsigned NBlk() const { return nblk_; }
  unsigned Len() const { return len_; }
  unsigned GetTotalDofSize() const { return nblk_ * len_; }
  friend double operator*(const CVector_Blk& a, const CVector_Blk& b) {
    double r = 0.0;
    check(dfm_vec_dot(a.size(), a.v_, b.v_, &r), "CVector_Blk dot");
    return r;
  }
  // host copies of the whole vector (GetValue / SetValue over all entries)
  void SetValues(const double* host) { check(dfm_memcpy_htod(v_, host, sizeof(double) * size()), "SetValues"); }
  void GetValues(double* host) const { check(dfm_memcpy_dtoh(host, v_, sizeof(double) * size()), "GetValues"); }
  std::vector<double> GetValues() const {
    std::vector<double> h(size());
    GetValues(h.data());
    return h;
  }
  double* device_ptr() { return v_; }
  const double* device_ptr() const { return v_; }
  size_t size() const { return (size_t)nblk_ * len_; }

 private:
  unsigned nblk_, len_;
  double* v_ = nullptr;
  bool owns_ = true;
};

class CBCFlag {
 public:
  CBCFlag(unsigned nblk, unsigned len) : nblk_(nblk), len_(len), flag_((size_t)nblk * len, 0) {}
  void SetAllFlagZero() { std::fill(flag_.begin(), flag_.end(), 0); }
  void SetBC(unsigned iblk, unsigned idof, int flag = 1) { flag_[(size_t)iblk * len_ + idof] = flag; }
  int GetBCFlag(unsigned iblk, unsigned idof) const { return flag_[(size_t)iblk * len_ + idof]; }
  unsigned NBlk() const { return nblk_; }
  unsigned LenBlk() const { return len_; }
  // CBCFlag::SetZeroToBCDof (bcflag_blk.h:102-129)
  bool SetZeroToBCDof(CVector_Blk& v) const {
    check(dfm_vec_zero_bc(v.size(), flag_.data(), v.device_ptr()), "SetZeroToBCDof");
    return true;
  }
  const int* data() const { return flag_.data(); }

 private:
  unsigned nblk_, len_;
  std::vector<int> flag_;
};

class CMatDia_BlkCrs {
 public:
  CMatDia_BlkCrs(unsigned nblk, unsigned len) { check(dfm_bcrs_create(nblk, len, &h_), "CMatDia_BlkCrs"); }
  ~CMatDia_BlkCrs() {
    if (h_) dfm_bcrs_destroy(h_);
  }
  CMatDia_BlkCrs(const CMatDia_BlkCrs&) = delete;
  CMatDia_BlkCrs& operator=(const CMatDia_BlkCrs&) = delete;
  unsigned NBlkMatCol() const { return info().nblk; }
  unsigned NBlkMatRow() const { return info().nblk; }
  unsigned LenBlkCol() const { return info().len; }
  unsigned LenBlkRow() const { return info().len; }
  unsigned NCrs() const { return info().ncrs; }
  // AddPattern(const Com::CIndexedArray&): index = row pointer, array = columns
  bool AddPattern(const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    return check(dfm_bcrs_add_pattern(h_, (unsigned)index.size() - 1, index.data(), array.data()),
                 "AddPattern") == DFM_OK;
  }
  bool SetZero() { return check(dfm_bcrs_set_zero(h_), "SetZero") == DFM_OK; }
  // Mearge of one element matrix; the diagonal-block matrix takes square
  // elements (the same node list for rows and columns)
  bool Mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
              unsigned blksize, const double* emat) {
    if (nblkel_col != nblkel_row) return false;
    for (unsigned i = 0; i < nblkel_col; i++)
      if (blkel_col[i] != blkel_row[i]) return false;
    return check(dfm_bcrs_mearge(h_, nblkel_col, blkel_col, blksize, emat), "Mearge") == DFM_OK;
  }
  // y = beta*y + alpha*A*x
  bool MatVec(double alpha, const CVector_Blk& x, double beta, CVector_Blk& y) const {
    return check(dfm_bcrs_matvec(h_, alpha, x.device_ptr(), beta, y.device_ptr()), "MatVec") == DFM_OK;
  }
  bool SetBoundaryCondition(const CBCFlag& flag) {
    return check(dfm_bcrs_set_bc(h_, flag.data()), "SetBoundaryCondition") == DFM_OK;
  }
  // host copies of the values (GetPtrValPSuP / GetPtrValDia)
  void GetValues(std::vector<double>& val, std::vector<double>& dia) const {
    const auto i = info();
    val.resize((size_t)i.ncrs * i.len * i.len);
    dia.resize((size_t)i.nblk * i.len * i.len);
    check(dfm_bcrs_get_values(h_, val.data(), dia.data()), "GetValues");
  }
  dfm_bcrs* handle() const { return h_; }

 private:
  struct Info {
    unsigned nblk, len, ncrs;
  };
  Info info() const {
    Info i{};
    check(dfm_bcrs_info(h_, &i.nblk, &i.len, &i.ncrs), "CMatDia_BlkCrs info");
    return i;
  }
  dfm_bcrs* h_ = nullptr;
};

// MatVec::CMat_BlkCrs (include/delfem/matvec/mat_blkcrs.h:45-195): a
// rectangular block CRS matrix, nblk_col block rows of len_col dofs by
// nblk_row block columns of len_row dofs; the off-diagonal blocks of a
// multi-segment system and CLinearSystem_Save's boundary matrix.
class CMat_BlkCrs {
 public:
  CMat_BlkCrs(unsigned nblk_col, unsigned len_col, unsigned nblk_row, unsigned len_row)
      : nblk_col_(nblk_col), len_col_(len_col), nblk_row_(nblk_row), len_row_(len_row) {
    check(dfm_bmat_create(nblk_col, len_col, nblk_row, len_row, &h_), "CMat_BlkCrs");
  }
  ~CMat_BlkCrs() {
    if (h_) dfm_bmat_destroy(h_);
  }
  CMat_BlkCrs(const CMat_BlkCrs&) = delete;
  CMat_BlkCrs& operator=(const CMat_BlkCrs&) = delete;
  unsigned NBlkMatCol() const { return nblk_col_; }
  unsigned NBlkMatRow() const { return nblk_row_; }
  unsigned LenBlkCol() const { return len_col_; }
  unsigned LenBlkRow() const { return len_row_; }
  unsigned NCrs() const {
    unsigned nb = 0, len = 0, ncrs = 0;
    check(dfm_bmat_info(h_, &nb, &len, &ncrs), "CMat_BlkCrs info");
    return ncrs;
  }
  // AddPattern(const Com::CIndexedArray&) (mat_blkcrs.cpp:280-414)
  bool AddPattern(const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    static const unsigned none = 0;
    return check(dfm_bmat_add_pattern(h_, (unsigned)index.size() - 1, index.data(),
                                      array.empty() ? &none : array.data()),
                 "CMat_BlkCrs::AddPattern") == DFM_OK;
  }
  bool SetZero() { return check(dfm_bmat_set_zero(h_), "CMat_BlkCrs::SetZero") == DFM_OK; }
  // Mearge of one element (mat_blkcrs.cpp:754-828); square elements only
  bool Mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
              unsigned blksize, const double* emat) {
    if (nblkel_col != nblkel_row) return false;
    for (unsigned i = 0; i < nblkel_col; i++)
      if (blkel_col[i] != blkel_row[i]) return false;
    return check(dfm_bmat_mearge(h_, nblkel_col, blkel_col, blksize, emat), "CMat_BlkCrs::Mearge") == DFM_OK;
  }
  // y = beta*y + alpha*A*x (mat_blkcrs.cpp:955-1174, isnt_trans = true)
  bool MatVec(double alpha, const CVector_Blk& x, double beta, CVector_Blk& y, bool isnt_trans = true) const {
    if (!isnt_trans) throw Error("CMat_BlkCrs::MatVec: the transposed product is not provided", DFM_NOT_IMPLEMENTED);
    return check(dfm_bmat_matvec(h_, alpha, x.device_ptr(), beta, y.device_ptr()), "CMat_BlkCrs::MatVec") == DFM_OK;
  }
  bool SetBoundaryCondition_Row(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_row(h_, bc.data()), "SetBoundaryCondition_Row") == DFM_OK;
  }
  bool SetBoundaryCondition_Colum(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_col(h_, bc.data()), "SetBoundaryCondition_Colum") == DFM_OK;
  }
  bool SetBoundaryConditionInverse_Colum(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_inverse_col(h_, bc.data()), "SetBoundaryConditionInverse_Colum") == DFM_OK;
  }
  // SetPatternBoundary (mat_blkcrs.cpp:1177-1265): the boundary matrix pattern of CLinearSystem_Save
  bool SetPatternBoundary(const CMatDia_BlkCrs& A, const CBCFlag& bc_col, const CBCFlag& bc_row) {
    return check(dfm_bmat_set_pattern_boundary(h_, A.handle(), bc_col.data(), bc_row.data()),
                 "SetPatternBoundary") == DFM_OK;
  }
  void GetValues(std::vector<double>& val) const {
    val.resize((size_t)NCrs() * len_col_ * len_row_);
    check(dfm_bmat_get_values(h_, val.data()), "CMat_BlkCrs::GetValues");
  }
  dfm_bmat* handle() const { return h_; }

 private:
  unsigned nblk_col_, len_col_, nblk_row_, len_row_;
  dfm_bmat* h_ = nullptr;
};

// Abstract preconditioner of the matrix layer (matprecond_blk.h:37-41).
class CPrecond_Blk {
 public:
  virtual ~CPrecond_Blk() {}
  virtual bool SolvePrecond(const CMatDia_BlkCrs& mat, CVector_Blk& vec) const = 0;
};

class CMatDiaFrac_BlkCrs : public CPrecond_Blk {
 public:
  // CMatDiaFrac_BlkCrs(A): ILU(0) pattern of A
  explicit CMatDiaFrac_BlkCrs(const CMatDia_BlkCrs& A) : CMatDiaFrac_BlkCrs(0, A) {}
  // CMatDiaFrac_BlkCrs(lev, A): symbolic ILU(lev) of A's pattern (lev = -1: unlimited)
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A) : src_(&A) {
    check(dfm_ilu_create(A.handle(), lev_fill, DFM_ORDER_NATURAL, nullptr, 0, &h_), "CMatDiaFrac_BlkCrs");
  }
  // ordered variant, order[new] = old
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A, const std::vector<int>& order)
      : src_(&A), ordering_(DFM_ORDER_GIVEN) {
    check(dfm_ilu_create(A.handle(), lev_fill, DFM_ORDER_GIVEN, order.data(), 0, &h_), "CMatDiaFrac_BlkCrs");
  }
  // device orderings: DFM_ORDER_MULTICOLOUR, DFM_ORDER_BLOCK_MULTICOLOUR
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A, int ordering, unsigned block_nodes = 0)
      : src_(&A), ordering_(ordering) {
    check(dfm_ilu_create(A.handle(), lev_fill, ordering, nullptr, block_nodes, &h_), "CMatDiaFrac_BlkCrs");
  }
  ~CMatDiaFrac_BlkCrs() override {
    if (h_) dfm_ilu_destroy(h_);
  }
  CMatDiaFrac_BlkCrs(const CMatDiaFrac_BlkCrs&) = delete;
  CMatDiaFrac_BlkCrs& operator=(const CMatDiaFrac_BlkCrs&) = delete;
  // SetValue(A) + DoILUDecomp(): false when more than 10 pivots were singular
  bool SetValue(const CMatDia_BlkCrs& A) {
    int nsing = 0;
    return check(dfm_ilu_set_value(h_, A.handle(), &nsing), "SetValue") == DFM_OK;
  }
  bool DoILUDecomp() { return true; }  // the factorisation runs in SetValue
  bool ForwardSubstitution(CVector_Blk& v) const {
    return check(dfm_ilu_forward(h_, v.device_ptr()), "ForwardSubstitution") == DFM_OK;
  }
  bool BackwardSubstitution(CVector_Blk& v) const {
    return check(dfm_ilu_backward(h_, v.device_ptr()), "BackwardSubstitution") == DFM_OK;
  }
  // Solve (matdiafrac_blkcrs.h:60-64): forward then backward substitution in
  // the factor's own numbering
  bool Solve(CVector_Blk& v) const {
    ForwardSubstitution(v);
    BackwardSubstitution(v);
    return true;
  }
  // CPrecond_Blk interface (matdiafrac_blkcrs.h:69-71): the matrix argument is not used
  bool SolvePrecond(const CMatDia_BlkCrs& mat, CVector_Blk& v) const override {
    (void)mat;
    return Solve(v);
  }
  const CMatDia_BlkCrs* source() const { return src_; }
  int ordering() const { return ordering_; }
  unsigned NCrs() const {
    unsigned n = 0;
    int a, b, c;
    check(dfm_ilu_info(h_, &n, &a, &b, &c), "NCrs");
    return n;
  }
  dfm_ilu* handle() const { return h_; }

 private:
  const CMatDia_BlkCrs* src_ = nullptr;
  int ordering_ = DFM_ORDER_NATURAL;
  dfm_ilu* h_ = nullptr;
};

// ---------------------------------------------------------------------------
// Matrix-layer Krylov solvers (src/matvec/solver_mat_iter.cpp).  They solve
// mat * u = r from u = 0; r is overwritten with the final residual.  With a
// natural-order CMatDiaFrac_BlkCrs built on `mat` as the preconditioner the
// preconditioned solvers run the library's device-resident solver; any other
// CPrecond_Blk (or an ordered factor, whose Solve works in its own numbering)
// runs the reference's loop over device vectors, calling SolvePrecond.
namespace Sol {
namespace detail {
inline const CMatDiaFrac_BlkCrs* device_factor(const CMatDia_BlkCrs& mat, const CPrecond_Blk& precond) {
  const auto* f = dynamic_cast<const CMatDiaFrac_BlkCrs*>(&precond);
  return f && f->source() == &mat && f->ordering() == DFM_ORDER_NATURAL ? f : nullptr;
}
}  // namespace detail

// Solve_CG (solver_mat_iter.cpp:37-109).  As in the reference, the
// convergence test reads the residual norm from before the current step, and
// a converged return leaves that step out of u (it is already in r).
inline bool Solve_CG(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat, CVector_Blk& r_vec,
                     CVector_Blk& u_vec) {
  const double tol = conv_ratio;
  const unsigned mx_iter = iteration;
  const unsigned nblk = r_vec.NBlk(), len = r_vec.Len();
  u_vec.SetVectorZero();
  double rr = r_vec * r_vec;
  if (rr < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr;
  CVector_Blk p_vec(nblk, len), Ap_vec(nblk, len);
  p_vec = r_vec;
  iteration = mx_iter;
  for (unsigned it = 1; it < mx_iter; it++) {
    mat.MatVec(1.0, p_vec, 0.0, Ap_vec);
    const double alpha = rr / (p_vec * Ap_vec);
    r_vec.AXPY(-alpha, Ap_vec);
    const double rr1 = r_vec * r_vec;
    if (rr * inv_rr0 < tol * tol) {
      conv_ratio = std::sqrt(rr * inv_rr0);
      iteration = it;
      return true;
    }
    u_vec.AXPY(alpha, p_vec);
    const double beta = rr1 / rr;
    rr = rr1;
    p_vec *= beta;
    p_vec += r_vec;
  }
  return true;
}

// Solve_PCG (solver_mat_iter.cpp:111-184): true when converged (conv_ratio =
// ||r||/||r0||, iteration = 0-based index); false at max_iter with both left
// at their inputs.
inline bool Solve_PCG(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat, CVector_Blk& r_vec,
                      CVector_Blk& u_vec, const CPrecond_Blk& precond, const CMatDia_BlkCrs& mat_p) {
  const double tol = conv_ratio;
  const unsigned mx_iter = iteration;
  if (const CMatDiaFrac_BlkCrs* f = detail::device_factor(mat, precond)) {
    const int rc = check(dfm_solve_pcg(mat.handle(), f->handle(), r_vec.device_ptr(), u_vec.device_ptr(),
                                       &conv_ratio, &iteration, nullptr),
                         "Sol::Solve_PCG");
    if (rc == DFM_OK) return true;
    conv_ratio = tol;
    iteration = mx_iter;
    return false;
  }
  const unsigned nblk = r_vec.NBlk(), len = r_vec.Len();
  u_vec.SetVectorZero();
  const double rr0 = r_vec * r_vec;
  if (rr0 < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  CVector_Blk p_vec(nblk, len), z_vec(nblk, len);
  z_vec = r_vec;
  precond.SolvePrecond(mat_p, z_vec);
  p_vec = z_vec;
  for (unsigned it = 0; it < mx_iter; it++) {
    const double rz = r_vec * z_vec;
    mat.MatVec(1.0, p_vec, 0.0, z_vec);
    const double alpha = rz / (p_vec * z_vec);
    r_vec.AXPY(-alpha, z_vec);
    u_vec.AXPY(alpha, p_vec);
    const double rr = r_vec * r_vec;
    if (rr * inv_rr0 < tol * tol) {
      conv_ratio = std::sqrt(rr * inv_rr0);
      iteration = it;
      return true;
    }
    z_vec = r_vec;
    precond.SolvePrecond(mat_p, z_vec);
    const double beta = (r_vec * z_vec) / rz;
    p_vec *= beta;
    p_vec += z_vec;
  }
  return false;
}

namespace detail {
// Solve_BiCGSTAB / Solve_PBiCGSTAB loop (solver_mat_iter.cpp:190-288, 294-423);
// precond == nullptr is the unpreconditioned form.  Always true; at max_iter
// conv_ratio keeps its input and iteration = max_iter.
inline bool bicgstab_loop(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat, CVector_Blk& r_vec,
                          CVector_Blk& u_vec, const CPrecond_Blk* precond) {
  const double tol = conv_ratio;
  const unsigned mx_iter = iteration;
  const unsigned nblk = r_vec.NBlk(), len = r_vec.Len();
  u_vec.SetVectorZero();
  const double rr0 = r_vec * r_vec;
  if (rr0 < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  CVector_Blk s_vec(nblk, len), p_vec(nblk, len), AMp_vec(nblk, len), AMs_vec(nblk, len), r0_vec(nblk, len);
  std::unique_ptr<CVector_Blk> Mp_vec, Ms_vec;
  if (precond) {
    Mp_vec.reset(new CVector_Blk(nblk, len));
    Ms_vec.reset(new CVector_Blk(nblk, len));
  }
  r0_vec = r_vec;
  p_vec = r_vec;
  iteration = mx_iter;
  for (unsigned it = 1; it < mx_iter; it++) {
    CVector_Blk* mp = &p_vec;
    if (precond) {
      *Mp_vec = p_vec;
      precond->SolvePrecond(mat, *Mp_vec);
      mp = Mp_vec.get();
    }
    const double r_r0 = r_vec * r0_vec;
    mat.MatVec(1.0, *mp, 0.0, AMp_vec);
    const double alpha = r_r0 / (AMp_vec * r0_vec);
    s_vec = r_vec;
    s_vec.AXPY(-alpha, AMp_vec);
    CVector_Blk* ms = &s_vec;
    if (precond) {
      *Ms_vec = s_vec;
      precond->SolvePrecond(mat, *Ms_vec);
      ms = Ms_vec.get();
    }
    mat.MatVec(1.0, *ms, 0.0, AMs_vec);
    const double den = AMs_vec * AMs_vec;
    const double omega = (s_vec * AMs_vec) / den;
    u_vec.AXPY(alpha, *mp);
    u_vec.AXPY(omega, *ms);
    r_vec = s_vec;
    r_vec.AXPY(-omega, AMs_vec);
    const double rr = r_vec * r_vec;
    if (rr * inv_rr0 < tol * tol) {
      conv_ratio = std::sqrt(rr * inv_rr0);
      iteration = it;
      return true;
    }
    const double beta = (r_vec * r0_vec) * alpha / (r_r0 * omega);
    p_vec *= beta;
    p_vec += r_vec;
    p_vec.AXPY(-beta * omega, AMp_vec);
  }
  return true;
}
}  // namespace detail

// Solve_BiCGSTAB (solver_mat_iter.cpp:190-288)
inline bool Solve_BiCGSTAB(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat, CVector_Blk& r_vec,
                           CVector_Blk& u_vec) {
  return detail::bicgstab_loop(conv_ratio, iteration, mat, r_vec, u_vec, nullptr);
}

// Solve_PBiCGSTAB (solver_mat_iter.cpp:294-423).  As in the reference the
// preconditioner is applied with `mat`; mat_p is not used.
inline bool Solve_PBiCGSTAB(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat,
                            CVector_Blk& r_vec, CVector_Blk& u_vec, const CPrecond_Blk& precond,
                            const CMatDia_BlkCrs& mat_p) {
  (void)mat_p;
  if (const CMatDiaFrac_BlkCrs* f = detail::device_factor(mat, precond)) {
    // the matrix layer returns early below 1e-30 (the ls layer: 1e-60)
    if (r_vec * r_vec < 1.0e-30) {
      u_vec.SetVectorZero();
      conv_ratio = 0.0;
      iteration = 0;
      return true;
    }
    check(dfm_solve_pbicgstab(mat.handle(), f->handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio,
                              &iteration, nullptr),
          "Sol::Solve_PBiCGSTAB");
    return true;
  }
  return detail::bicgstab_loop(conv_ratio, iteration, mat, r_vec, u_vec, &precond);
}

}  // namespace Sol
}  // namespace MatVec
}  // namespace delfem_amd

#ifndef DELFEM_AMD_NO_ALIASES
namespace MatVec = delfem_amd::MatVec;
#endif
