// delfem_amd/matvec.hpp -- C++ classes with the names and semantics of
// DelFEM's MatVec layer over the dfm.h C-ABI: block vectors, Dirichlet flags,
// the block-CRS matrix with separate diagonal and its ILU(k) factor.  Values
// live in device memory; host arrays are copied in and out explicitly.
//
//   MatVec::CVector_Blk        include/delfem/matvec/vector_blk.h:41-215
//   MatVec::CBCFlag            include/delfem/matvec/bcflag_blk.h:36-140
//   MatVec::CMatDia_BlkCrs     include/delfem/matvec/matdia_blkcrs.h:44-118
//   MatVec::CMatDiaFrac_BlkCrs include/delfem/matvec/matdiafrac_blkcrs.h:40-137
//
// The classes are declared in namespace delfem_amd::MatVec; unless
// DELFEM_AMD_NO_ALIASES is defined, namespace MatVec aliases it so that code
// written against the reference compiles unchanged.  Failures of the library
// (bad arguments, device errors) throw delfem_amd::Error; numerical outcomes
// are returned as bool like the reference.
#pragma once

#include <algorithm>
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
  ~CVector_Blk() {
    if (v_) dfm_free(v_);
  }
  CVector_Blk(const CVector_Blk&) = delete;
  CVector_Blk(CVector_Blk&& o) noexcept : nblk_(o.nblk_), len_(o.len_), v_(o.v_) { o.v_ = nullptr; }
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

class CMatDiaFrac_BlkCrs {
 public:
  // CMatDiaFrac_BlkCrs(lev, A): symbolic ILU(lev) of A's pattern (lev = -1: unlimited)
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A) {
    check(dfm_ilu_create(A.handle(), lev_fill, DFM_ORDER_NATURAL, nullptr, 0, &h_), "CMatDiaFrac_BlkCrs");
  }
  // ordered variant, order[new] = old
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A, const std::vector<int>& order) {
    check(dfm_ilu_create(A.handle(), lev_fill, DFM_ORDER_GIVEN, order.data(), 0, &h_), "CMatDiaFrac_BlkCrs");
  }
  // device orderings: DFM_ORDER_MULTICOLOUR, DFM_ORDER_BLOCK_MULTICOLOUR
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A, int ordering, unsigned block_nodes = 0) {
    check(dfm_ilu_create(A.handle(), lev_fill, ordering, nullptr, block_nodes, &h_), "CMatDiaFrac_BlkCrs");
  }
  ~CMatDiaFrac_BlkCrs() {
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
  unsigned NCrs() const {
    unsigned n = 0;
    int a, b, c;
    check(dfm_ilu_info(h_, &n, &a, &b, &c), "NCrs");
    return n;
  }
  dfm_ilu* handle() const { return h_; }

 private:
  dfm_ilu* h_ = nullptr;
};

}  // namespace MatVec
}  // namespace delfem_amd

#ifndef DELFEM_AMD_NO_ALIASES
namespace MatVec = delfem_amd::MatVec;
#endif
