// delfem_amd/zmatvec.hpp -- C++ classes with the names and semantics of
// DelFEM's complex MatVec layer over the dfm.h C-ABI (scalar blocks, as the
// reference's complex solvers use them).  Values live in device memory;
// host arrays are std::complex<double>, whose layout is Com::Complex's
// (two doubles, real first).
//
//   MatVec::CZVector_Blk        include/delfem/matvec/zvector_blk.h:39-95
//   MatVec::CZMatDia_BlkCrs     include/delfem/matvec/zmatdia_blkcrs.h:39-100
//   MatVec::CZMatPrecond_Blk    include/delfem/matvec/zmatprecond_blk.h:34-38
//   MatVec::CZMatDiaFrac_BlkCrs include/delfem/matvec/zmatdiafrac_blkcrs.h:44-96
//   MatVec::Sol::Solve_CG / Solve_PCG (Hermitian) / Solve_CGNR / Solve_COCG / Solve_PCOCG /
//   Solve_BiCGSTAB / Solve_BiCGStabP
//                               include/delfem/matvec/zsolver_mat_iter.h:51-80
//
// The preconditioned solvers take the factor class itself (built on the same
// matrix) and run on the device; another CZMatPrecond_Blk is rejected with
// delfem_amd::Error.  Like the reference, every solver returns true, also
// when max_iter stops it.
#pragma once

#include <complex>
#include <vector>

#include "matvec.hpp"

namespace delfem_amd {
namespace MatVec {

class CZMatDia_BlkCrs;

class CZVector_Blk {
 public:
  CZVector_Blk(unsigned nblk, unsigned len) : nblk_(nblk) {
    if (len != 1) throw Error("CZVector_Blk: scalar blocks only", DFM_NOT_IMPLEMENTED);
    check(dfm_malloc(reinterpret_cast<void**>(&v_), 2 * sizeof(double) * std::max(1u, nblk)), "CZVector_Blk");
    SetVectorZero();
  }
  ~CZVector_Blk() {
    if (v_) dfm_free(v_);
  }
  CZVector_Blk(const CZVector_Blk&) = delete;
  CZVector_Blk(CZVector_Blk&& o) noexcept : nblk_(o.nblk_), v_(o.v_) { o.v_ = nullptr; }
  CZVector_Blk& operator=(const CZVector_Blk& rhs) {
    check(dfm_memcpy_dtod(v_, rhs.v_, 2 * sizeof(double) * nblk_), "CZVector_Blk::operator=");
    return *this;
  }
  void SetVectorZero() { check(dfm_vec_fill(2 * (size_t)nblk_, 0.0, v_), "CZVector_Blk::SetVectorZero"); }
  double GetSquaredVectorNorm() const {
    double r = 0.0;
    check(dfm_zvec_norm2(nblk_, v_, &r), "GetSquaredVectorNorm");
    return r;
  }
  // operator* (zvector_blk.cpp:39-60): sum a_i b_i, no conjugate
  friend std::complex<double> operator*(const CZVector_Blk& a, const CZVector_Blk& b) {
    double r[2] = {0.0, 0.0};
    check(dfm_zvec_dot(a.nblk_, a.v_, b.v_, r), "CZVector_Blk dot");
    return {r[0], r[1]};
  }
  unsigned BlkLen() const { return 1; }
  unsigned BlkVecLen() const { return nblk_; }
  void SetValues(const std::complex<double>* host) {
    check(dfm_memcpy_htod(v_, host, 2 * sizeof(double) * nblk_), "SetValues");
  }
  void GetValues(std::complex<double>* host) const {
    check(dfm_memcpy_dtoh(host, v_, 2 * sizeof(double) * nblk_), "GetValues");
  }
  std::vector<std::complex<double>> GetValues() const {
    std::vector<std::complex<double>> h(nblk_);
    GetValues(h.data());
    return h;
  }
  double* device_ptr() { return v_; }
  const double* device_ptr() const { return v_; }

 private:
  unsigned nblk_;
  double* v_ = nullptr;
};

class CZMatDia_BlkCrs {
 public:
  CZMatDia_BlkCrs(unsigned nblk, unsigned len) { check(dfm_zbcrs_create(nblk, len, &h_), "CZMatDia_BlkCrs"); }
  ~CZMatDia_BlkCrs() {
    if (h_) dfm_zbcrs_destroy(h_);
  }
  CZMatDia_BlkCrs(const CZMatDia_BlkCrs&) = delete;
  CZMatDia_BlkCrs& operator=(const CZMatDia_BlkCrs&) = delete;
  unsigned NBlkMatCol() const { return info().nblk; }
  unsigned NBlkMatRow() const { return info().nblk; }
  unsigned LenBlkCol() const { return 1; }
  unsigned LenBlkRow() const { return 1; }
  unsigned NCrs() const { return info().ncrs; }
  bool AddPattern(const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    return check(dfm_zbcrs_add_pattern(h_, (unsigned)index.size() - 1, index.data(), array.data()),
                 "AddPattern") == DFM_OK;
  }
  bool SetZero() { return check(dfm_zbcrs_set_zero(h_), "SetZero") == DFM_OK; }
  // Mearge of one square element matrix (nblkel^2 values, row-major)
  bool Mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
              unsigned blksize, const std::complex<double>* emat) {
    if (nblkel_col != nblkel_row) return false;
    for (unsigned i = 0; i < nblkel_col; i++)
      if (blkel_col[i] != blkel_row[i]) return false;
    return check(dfm_zbcrs_mearge(h_, nblkel_col, blkel_col, blksize, reinterpret_cast<const double*>(emat)),
                 "Mearge") == DFM_OK;
  }
  bool SetBoundaryCondition(const CBCFlag& flag) {
    return check(dfm_zbcrs_set_bc(h_, flag.data()), "SetBoundaryCondition") == DFM_OK;
  }
  // y = beta*y + alpha*A*x
  bool MatVec(double alpha, const CZVector_Blk& x, double beta, CZVector_Blk& y) const {
    return check(dfm_zbcrs_matvec(h_, alpha, x.device_ptr(), beta, y.device_ptr()), "MatVec") == DFM_OK;
  }
  // y = beta*y + alpha*A^H*x (zmatdia_blkcrs.cpp:803-847)
  bool MatVec_Hermitian(double alpha, const CZVector_Blk& x, double beta, CZVector_Blk& y) const {
    return check(dfm_zbcrs_matvec_hermitian(h_, alpha, x.device_ptr(), beta, y.device_ptr()),
                 "MatVec_Hermitian") == DFM_OK;
  }
  void GetValues(std::vector<std::complex<double>>& val, std::vector<std::complex<double>>& dia) const {
    const auto i = info();
    val.resize(i.ncrs);
    dia.resize(i.nblk);
    check(dfm_zbcrs_get_values(h_, reinterpret_cast<double*>(val.data()), reinterpret_cast<double*>(dia.data())),
          "GetValues");
  }
  dfm_zbcrs* handle() const { return h_; }

 private:
  struct Info {
    unsigned nblk, len, ncrs;
  };
  Info info() const {
    Info i{};
    check(dfm_zbcrs_info(h_, &i.nblk, &i.len, &i.ncrs), "CZMatDia_BlkCrs info");
    return i;
  }
  dfm_zbcrs* h_ = nullptr;
};

class CZMatPrecond_Blk {
 public:
  virtual ~CZMatPrecond_Blk() {}
  virtual bool SolvePrecond(const CZMatDia_BlkCrs& mat, CZVector_Blk& vec) const = 0;
};

class CZMatDiaFrac_BlkCrs : public CZMatPrecond_Blk {
 public:
  explicit CZMatDiaFrac_BlkCrs(const CZMatDia_BlkCrs& A) : CZMatDiaFrac_BlkCrs(0, A) {}
  // CZMatDiaFrac_BlkCrs(lev, A): symbolic ILU(lev) of A's pattern (lev = -1: unlimited)
  CZMatDiaFrac_BlkCrs(int lev_fill, const CZMatDia_BlkCrs& A) : src_(&A) {
    check(dfm_zilu_create(A.handle(), lev_fill, DFM_ORDER_NATURAL, nullptr, &h_), "CZMatDiaFrac_BlkCrs");
  }
  // device orderings (DFM_ORDER_MULTICOLOUR); the factor then works in its own numbering
  CZMatDiaFrac_BlkCrs(int lev_fill, const CZMatDia_BlkCrs& A, int ordering) : src_(&A), ordering_(ordering) {
    check(dfm_zilu_create(A.handle(), lev_fill, ordering, nullptr, &h_), "CZMatDiaFrac_BlkCrs");
  }
  ~CZMatDiaFrac_BlkCrs() override {
    if (h_) dfm_zilu_destroy(h_);
  }
  CZMatDiaFrac_BlkCrs(const CZMatDiaFrac_BlkCrs&) = delete;
  CZMatDiaFrac_BlkCrs& operator=(const CZMatDiaFrac_BlkCrs&) = delete;
  // SetValue(A) = copy + DoILUDecomp (zmatdiafrac_blkcrs.cpp:466-486, 192-262)
  bool SetValue(const CZMatDia_BlkCrs& A) {
    int nsmall = 0;
    return check(dfm_zilu_set_value(h_, A.handle(), &nsmall), "SetValue") == DFM_OK;
  }
  // SetValue_Initialize + DoILUDecomp as zpreconditioner.h calls them: the
  // device factorisation runs in the first
  bool SetValue_Initialize(const CZMatDia_BlkCrs& A) { return SetValue(A); }
  bool DoILUDecomp() { return true; }
  bool ForwardSubstitution(CZVector_Blk& v) const {
    return check(dfm_zilu_forward(h_, v.device_ptr()), "ForwardSubstitution") == DFM_OK;
  }
  bool BackwardSubstitution(CZVector_Blk& v) const {
    return check(dfm_zilu_backward(h_, v.device_ptr()), "BackwardSubstitution") == DFM_OK;
  }
  bool Solve(CZVector_Blk& v) const {
    ForwardSubstitution(v);
    BackwardSubstitution(v);
    return true;
  }
  bool SolvePrecond(const CZMatDia_BlkCrs& mat, CZVector_Blk& v) const override {
    (void)mat;
    return Solve(v);
  }
  const CZMatDia_BlkCrs* source() const { return src_; }
  int ordering() const { return ordering_; }
  dfm_zilu* handle() const { return h_; }

 private:
  const CZMatDia_BlkCrs* src_ = nullptr;
  int ordering_ = DFM_ORDER_NATURAL;
  dfm_zilu* h_ = nullptr;
};

namespace Sol {
namespace detail {
inline const CZMatDiaFrac_BlkCrs& device_zfactor(const CZMatDia_BlkCrs& mat, const CZMatPrecond_Blk& precond) {
  const auto* f = dynamic_cast<const CZMatDiaFrac_BlkCrs*>(&precond);
  if (!f || f->source() != &mat)
    throw Error("complex preconditioned solvers take a CZMatDiaFrac_BlkCrs built on the matrix", DFM_BAD_ARG);
  return *f;
}
}  // namespace detail

// Solve_CG for a Hermitian matrix (zsolver_mat_iter.cpp:32-119)
inline bool Solve_CG(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                     CZVector_Blk& u_vec) {
  check(dfm_zsolve_cg(mat.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio, &iteration),
        "Sol::Solve_CG");
  return true;
}
// Solve_PCG for a Hermitian matrix (:121-218); false at max_iter
inline bool Solve_PCG(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                      CZVector_Blk& u_vec, const CZMatPrecond_Blk& precond, const CZMatDia_BlkCrs& mat_p) {
  (void)mat_p;
  return check(dfm_zsolve_pcg(mat.handle(), detail::device_zfactor(mat, precond).handle(), r_vec.device_ptr(),
                              u_vec.device_ptr(), &conv_ratio, &iteration),
               "Sol::Solve_PCG") == DFM_OK;
}
// Solve_CGNR (zsolver_mat_iter.cpp:394-483)
inline bool Solve_CGNR(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                       CZVector_Blk& u_vec) {
  check(dfm_zsolve_cgnr(mat.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio, &iteration),
        "Sol::Solve_CGNR");
  return true;
}
// Solve_COCG (zsolver_mat_iter.cpp:220-299)
inline bool Solve_COCG(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                       CZVector_Blk& u_vec) {
  check(dfm_zsolve_cocg(mat.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio, &iteration),
        "Sol::Solve_COCG");
  return true;
}
// Solve_PCOCG (:303-392)
inline bool Solve_PCOCG(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                        CZVector_Blk& u_vec, const CZMatPrecond_Blk& precond, const CZMatDia_BlkCrs& mat_p) {
  (void)mat_p;
  const CZMatDiaFrac_BlkCrs& f = detail::device_zfactor(mat, precond);
  check(dfm_zsolve_pcocg(mat.handle(), f.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio, &iteration),
        "Sol::Solve_PCOCG");
  return true;
}
// Solve_BiCGSTAB (:487-616)
inline bool Solve_BiCGSTAB(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                           CZVector_Blk& u_vec) {
  check(dfm_zsolve_bicgstab(mat.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio, &iteration),
        "Sol::Solve_BiCGSTAB");
  return true;
}
// Solve_BiCGStabP (:622-739)
inline bool Solve_BiCGStabP(double& conv_ratio, unsigned& iteration, const CZMatDia_BlkCrs& mat, CZVector_Blk& r_vec,
                            CZVector_Blk& u_vec, const CZMatPrecond_Blk& precond, CZMatDia_BlkCrs& mat_p) {
  (void)mat_p;
  const CZMatDiaFrac_BlkCrs& f = detail::device_zfactor(mat, precond);
  check(dfm_zsolve_bicgstabp(mat.handle(), f.handle(), r_vec.device_ptr(), u_vec.device_ptr(), &conv_ratio,
                             &iteration),
        "Sol::Solve_BiCGStabP");
  return true;
}
}  // namespace Sol

}  // namespace MatVec
}  // namespace delfem_amd
