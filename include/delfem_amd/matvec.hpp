// delfem_amd/matvec.hpp -- C++ classes with the names and semantics of
// DelFEM's MatVec layer over the dfm.h C-ABI: block vectors, Dirichlet flags,
// the block-CRS matrix with separate diagonal and its ILU(k) factor.
//
//   MatVec::CVector_Blk        include/delfem/matvec/vector_blk.h:41-215
//   MatVec::CBCFlag            include/delfem/matvec/bcflag_blk.h:36-140
//   MatVec::CMatDia_BlkCrs     include/delfem/matvec/matdia_blkcrs.h:44-118
//   MatVec::CMat_BlkCrs        include/delfem/matvec/mat_blkcrs.h:45-195
//   MatVec::CMatDiaFrac_BlkCrs include/delfem/matvec/matdiafrac_blkcrs.h:40-137
//   MatVec::CPrecond_Blk       include/delfem/matvec/matprecond_blk.h:37-41
//   MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
//                              include/delfem/matvec/solver_mat_iter.h:41-60
//
// Values live in device memory.  Element loops written against the reference
// work unchanged: per-entry vector access (GetValue / SetValue / AddValue /
// GetValuePtr) and per-element Mearge act on host copies that are kept in
// step lazily -- a vector's host copy is fetched whole on the first host
// access after device work and sent back whole before the next device
// operation; a matrix's Mearge calls accumulate into a host copy of its values
// that is sent to the device, once, when the matrix is next used there.
// Host pointers handed out (GetValue references, GetValuePtr, GetPtrValPSuP)
// stay valid, but see device results only after being fetched again.
//
// The classes are declared in namespace delfem_amd::MatVec; unless
// DELFEM_AMD_NO_ALIASES is defined, namespace MatVec aliases it so that code
// written against the reference compiles unchanged.  Failures of the library
// (bad arguments, device errors) throw delfem_amd::Error; numerical outcomes
// are returned as bool like the reference.  Variable ("flex") block sizes
// (the reference's aLen constructors) are provided for vectors, flags, the
// diagonal matrix and its factor; a flex CMat_BlkCrs (off-diagonal blocks
// between segments) is not.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../dfm.h"
#include "indexed_array.hpp"

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

namespace detail {
// One device array of doubles and its host copy.  host_ok / dev_ok say which
// side holds the current values (at least one does); the other side is
// brought up to date, whole, when it is next used.  A multi-segment system's
// whole-vector and per-segment vectors share one store.
class VecStore {
 public:
  explicit VecStore(size_t n) : n_(n) {
    if (n_) check(dfm_malloc(reinterpret_cast<void**>(&dev_), sizeof(double) * n_), "CVector_Blk");
  }
  ~VecStore() {
    if (dev_) dfm_free(dev_);
  }
  VecStore(const VecStore&) = delete;
  VecStore& operator=(const VecStore&) = delete;
  size_t size() const { return n_; }
  const double* host_read() {
    pull();
    return host_.data();
  }
  double* host_write() {
    pull();
    dev_ok_ = false;
    return host_.data();
  }
  const double* dev_read() {
    push();
    return dev_;
  }
  double* dev_write() {
    push();
    host_ok_ = false;
    return dev_;
  }
  // the device side, about to be overwritten whole: nothing is sent over
  double* dev_overwrite() {
    dev_ok_ = true;
    host_ok_ = false;
    return dev_;
  }

 private:
  void pull() {
    if (host_ok_) return;
    host_.resize(n_);
    if (n_) check(dfm_memcpy_dtoh(host_.data(), dev_, sizeof(double) * n_), "CVector_Blk (device to host)");
    host_ok_ = true;
  }
  void push() {
    if (dev_ok_) return;
    if (n_) check(dfm_memcpy_htod(dev_, host_.data(), sizeof(double) * n_), "CVector_Blk (host to device)");
    dev_ok_ = true;
  }
  size_t n_;
  double* dev_ = nullptr;
  std::vector<double> host_;
  bool host_ok_ = false, dev_ok_ = true;
};
}  // namespace detail

class CMatDia_BlkCrs;

typedef std::shared_ptr<const std::vector<unsigned>> DofPtr;
// first dof of each block (m_DofPtr) of per-block lengths aLen
inline DofPtr make_dofptr(const std::vector<unsigned>& aLen) {
  auto d = std::make_shared<std::vector<unsigned>>(aLen.size() + 1, 0);
  for (size_t i = 0; i < aLen.size(); i++) (*d)[i + 1] = (*d)[i] + aLen[i];
  return d;
}

class CVector_Blk {
 public:
  CVector_Blk(unsigned nblk, unsigned len)
      : nblk_(nblk), len_(len), st_(std::make_shared<detail::VecStore>((size_t)nblk * len)) {
    SetVectorZero();
  }
  // per-block lengths (vector_blk.h:57-60, Initialize :84-99): Len() is -1
  CVector_Blk(unsigned nblk, const std::vector<unsigned>& aLen) : CVector_Blk(nblk, make_dofptr(aLen)) {}
  CVector_Blk(unsigned nblk, DofPtr dofptr)
      : nblk_(nblk), len_(0), dofptr_(std::move(dofptr)), st_(std::make_shared<detail::VecStore>(dofptr_->back())) {
    SetVectorZero();
  }
  // a copy of the values (vector_blk.h:62-72)
  CVector_Blk(const CVector_Blk& v)
      : nblk_(v.nblk_), len_(v.len_), dofptr_(v.dofptr_), st_(std::make_shared<detail::VecStore>(v.size())) {
    SetVectorZero();
    *this = v;
  }
  // a view of nblk*len entries of `st` from `off` (one segment of a
  // multi-segment system's vector)
  CVector_Blk(unsigned nblk, unsigned len, std::shared_ptr<detail::VecStore> st, size_t off)
      : nblk_(nblk), len_(len), st_(std::move(st)), off_(off) {}
  // ... and of a segment of per-block lengths
  CVector_Blk(unsigned nblk, DofPtr dofptr, std::shared_ptr<detail::VecStore> st, size_t off)
      : nblk_(nblk), len_(0), dofptr_(std::move(dofptr)), st_(std::move(st)), off_(off) {}
  virtual ~CVector_Blk() = default;
  // operator= copies the values (vector_blk.cpp:62-71)
  CVector_Blk& operator=(const CVector_Blk& rhs) {
    if (this != &rhs) {
      const double* x = rhs.dread();
      check(dfm_vec_copy(size(), x, dwrite()), "CVector_Blk::operator=");
    }
    return *this;
  }
  CVector_Blk& operator*=(double d) {
    check(dfm_vec_scal(size(), d, dwrite()), "CVector_Blk::operator*=");
    return *this;
  }
  CVector_Blk& AXPY(const double& alpha, const CVector_Blk& rhs) {
    const double* x = rhs.dread();
    check(dfm_vec_axpy(size(), alpha, x, dwrite()), "CVector_Blk::AXPY");
    return *this;
  }
  // operator+= (vector_blk.cpp:73-81): y + 1.0 * x is y + x exactly
  CVector_Blk& operator+=(const CVector_Blk& rhs) { return AXPY(1.0, rhs); }
  void SetVectorZero() {
    double* p = whole() ? st_->dev_overwrite() : st_->dev_write() + off_;
    check(dfm_vec_fill(size(), 0.0, p), "CVector_Blk::SetVectorZero");
  }
  double GetSquaredVectorNorm() const { return (*this) * (*this); }
  unsigned NBlk() const { return nblk_; }
  // the block length, -1 with per-block lengths (vector_blk.h:121-127)
  int Len() const { return dofptr_ ? -1 : (int)len_; }
  unsigned Len(unsigned iblk) const { return dofptr_ ? (*dofptr_)[iblk + 1] - (*dofptr_)[iblk] : len_; }
  unsigned GetTotalDofSize() const { return (unsigned)size(); }
  friend double operator*(const CVector_Blk& a, const CVector_Blk& b) {
    double r = 0.0;
    const double* x = a.dread();
    const double* y = b.dread();
    check(dfm_vec_dot(a.size(), x, y, &r), "CVector_Blk dot");
    return r;
  }
  // per-entry access on the host copy (vector_blk.h:135-181)
  const double& GetValue(unsigned iblk, unsigned idofblk) const { return st_->host_read()[pos(iblk) + idofblk]; }
  void SetValue(unsigned iblk, unsigned idofblk, double val) { st_->host_write()[pos(iblk) + idofblk] = val; }
  void AddValue(unsigned iblk, unsigned idofblk, double val) { st_->host_write()[pos(iblk) + idofblk] += val; }
  double* GetValuePtr(unsigned iblk) { return st_->host_write() + pos(iblk); }
  const double* GetValuePtr(unsigned iblk) const { return st_->host_read() + pos(iblk); }
  // a zero vector of the same block layout
  static CVector_Blk like(const CVector_Blk& v) {
    return v.dofptr_ ? CVector_Blk(v.nblk_, v.dofptr_) : CVector_Blk(v.nblk_, v.len_);
  }
  // host copies of the whole vector
  void SetValues(const double* host) {
    double* p = whole() ? st_->dev_overwrite() : st_->dev_write() + off_;
    check(dfm_memcpy_htod(p, host, sizeof(double) * size()), "SetValues");
  }
  void GetValues(double* host) const {
    check(dfm_memcpy_dtoh(host, dread(), sizeof(double) * size()), "GetValues");
  }
  std::vector<double> GetValues() const {
    std::vector<double> h(size());
    GetValues(h.data());
    return h;
  }
  // device memory of the values, current; the non-const form counts as a write
  double* device_ptr() { return dwrite(); }
  const double* device_ptr() const { return dread(); }
  size_t size() const { return dofptr_ ? dofptr_->back() : (size_t)nblk_ * len_; }
  // the first dof of each block with per-block lengths, else NULL
  const DofPtr& dofptr() const { return dofptr_; }
  // re-point this vector at another store (the vector keeps its identity);
  // the caller has copied the values over
  void rebind(std::shared_ptr<detail::VecStore> st, size_t off) {
    st_ = std::move(st);
    off_ = off;
  }
  const std::shared_ptr<detail::VecStore>& store() const { return st_; }
  size_t offset() const { return off_; }

 private:
  bool whole() const { return off_ == 0 && size() == st_->size(); }
  size_t pos(unsigned iblk) const { return off_ + (dofptr_ ? (size_t)(*dofptr_)[iblk] : (size_t)iblk * len_); }
  const double* dread() const { return st_->dev_read() + off_; }
  double* dwrite() const { return st_->dev_write() + off_; }
  unsigned nblk_, len_;
  DofPtr dofptr_;
  std::shared_ptr<detail::VecStore> st_;
  size_t off_ = 0;
};

class CBCFlag {
 public:
  CBCFlag(unsigned nblk, unsigned len) : nblk_(nblk), len_(len), flag_((size_t)nblk * len, 0) {}
  // per-block lengths (bcflag_blk.h:46-58)
  CBCFlag(unsigned nblk, const std::vector<unsigned>& aLen)
      : nblk_(nblk), len_(0), dofptr_(make_dofptr(aLen)), flag_(dofptr_->back(), 0) {}
  virtual ~CBCFlag() = default;
  void SetAllFlagZero() { std::fill(flag_.begin(), flag_.end(), 0); }
  // SetBC (bcflag_blk.h:153-162): false, and nothing set, with per-block
  // lengths -- the reference refuses them (its flex flags stay zero)
  bool SetBC(unsigned iblk, unsigned idof, int flag = 1) {
    if (dofptr_) return false;
    flag_[(size_t)iblk * len_ + idof] = flag;
    return true;
  }
  int GetBCFlag(unsigned iblk, unsigned idof) const {
    return flag_[(dofptr_ ? (size_t)(*dofptr_)[iblk] : (size_t)iblk * len_) + idof];
  }
  unsigned NBlk() const { return nblk_; }
  int LenBlk() const { return dofptr_ ? -1 : (int)len_; }
  unsigned LenBlk(unsigned iblk) const { return dofptr_ ? (*dofptr_)[iblk + 1] - (*dofptr_)[iblk] : len_; }
  // CBCFlag::SetZeroToBCDof (bcflag_blk.h:102-129)
  bool SetZeroToBCDof(CVector_Blk& v) const {
    check(dfm_vec_zero_bc(v.size(), flag_.data(), v.device_ptr()), "SetZeroToBCDof");
    return true;
  }
  const int* data() const { return flag_.data(); }

 private:
  unsigned nblk_, len_;
  DofPtr dofptr_;
  std::vector<int> flag_;
};

namespace detail {
// Host copy of a block-CRS matrix's values for element loops: filled from
// the device on the first host access (or set to zero by SetZero), written
// back whole before the matrix is next used on the device.
struct ValMirror {
  std::vector<unsigned> rowptr, col;  // pattern (host)
  std::vector<unsigned> valptr, diaptr;  // per-block lengths: value offsets (m_ValPtr, m_DiaValPtr)
  std::vector<double> val, dia;
  bool valid = false;                 // val / dia hold the current values
  bool dirty = false;                 // ... and the device does not
};
}  // namespace detail

class CMatDia_BlkCrs {
 public:
  CMatDia_BlkCrs(unsigned nblk, unsigned len) : nblk_(nblk), len_(len) {
    check(dfm_bcrs_create(nblk, len, &h_), "CMatDia_BlkCrs");
    m_.rowptr.assign(nblk + 1, 0);
  }
  // per-block lengths (Initialize(nblk, aLen), matdia_blkcrs.cpp:70-104):
  // LenBlkCol() is -1, Mearge takes one block at a time
  CMatDia_BlkCrs(unsigned nblk, const std::vector<unsigned>& aLen)
      : nblk_(nblk), len_(0), dofptr_(make_dofptr(aLen)) {
    if (aLen.size() != nblk) throw Error("CMatDia_BlkCrs: aLen needs one length per block", DFM_BAD_ARG);
    check(dfm_fbcrs_create(nblk, aLen.data(), &fh_), "CMatDia_BlkCrs");
    m_.rowptr.assign(nblk + 1, 0);
    flex_layout();
  }
  virtual ~CMatDia_BlkCrs() {
    if (h_) dfm_bcrs_destroy(h_);
    if (fh_) dfm_fbcrs_destroy(fh_);
  }
  CMatDia_BlkCrs(const CMatDia_BlkCrs&) = delete;
  CMatDia_BlkCrs& operator=(const CMatDia_BlkCrs&) = delete;
  unsigned NBlkMatCol() const { return nblk_; }
  unsigned NBlkMatRow() const { return nblk_; }
  int LenBlkCol() const { return fh_ ? -1 : (int)len_; }
  int LenBlkRow() const { return fh_ ? -1 : (int)len_; }
  unsigned LenBlkCol(unsigned iblk) const { return fh_ ? (*dofptr_)[iblk + 1] - (*dofptr_)[iblk] : len_; }
  unsigned LenBlkRow(unsigned iblk) const { return LenBlkCol(iblk); }
  unsigned NCrs() const { return (unsigned)m_.col.size(); }
  // with per-block lengths: the first dof of each block, else NULL
  const DofPtr& dofptr() const { return dofptr_; }
  // AddPattern(const Com::CIndexedArray&) (matdia_blkcrs.cpp:346-505): the
  // diagonal is excluded, rows are merged with the existing pattern
  bool AddPattern(const Com::CIndexedArray& crs) { return AddPattern(crs.index, crs.array); }
  bool AddPattern(const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    flush();
    static const unsigned none = 0;
    if (fh_) {
      const bool ok = check(dfm_fbcrs_add_pattern(fh_, (unsigned)index.size() - 1, index.data(),
                                                  array.empty() ? &none : array.data()),
                            "AddPattern") == DFM_OK;
      flex_layout();
      return ok;
    }
    const bool ok = check(dfm_bcrs_add_pattern(h_, (unsigned)index.size() - 1, index.data(),
                                               array.empty() ? &none : array.data()),
                          "AddPattern") == DFM_OK;
    unsigned nb = 0, len = 0, ncrs = 0;
    check(dfm_bcrs_info(h_, &nb, &len, &ncrs), "AddPattern");
    m_.col.resize(ncrs);
    check(dfm_bcrs_get_pattern(h_, m_.rowptr.data(), m_.col.data()), "AddPattern");
    m_.valid = m_.dirty = false;
    return ok;
  }
  // SetZero (matdia_blkcrs.cpp:139-163): on the host copy, which the element
  // loop that follows fills
  bool SetZero() {
    m_.val.assign(nval(), 0.0);
    m_.dia.assign(ndia(), 0.0);
    m_.valid = m_.dirty = true;
    return true;
  }
  // Mearge (matdia_blkcrs.cpp:165-258): element block (i, j) at
  // emat[(i*nblkel_row + j)*blksize]; position i == j goes to the diagonal
  // block of node blkel_col[i], other blocks to their CRS entry of row
  // blkel_col[i], and blocks outside the pattern are dropped.  With
  // per-block lengths (:173-207) one block (i, j) of len_i x len_j values per
  // call, which must be in the pattern.
  bool Mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
              unsigned blksize, const double* emat) {
    if (fh_) return flex_mearge(nblkel_col, blkel_col, nblkel_row, blkel_row, blksize, emat);
    const size_t bb = (size_t)len_ * len_;
    if (blksize != bb) return false;
    for (unsigned i = 0; i < nblkel_col; i++)
      if (blkel_col[i] >= nblk_) return false;
    for (unsigned j = 0; j < nblkel_row; j++)
      if (blkel_row[j] >= nblk_) return false;
    host_values(true);
    for (unsigned ie = 0; ie < nblkel_col; ie++) {
      const unsigned i = blkel_col[ie];
      const unsigned* c0 = m_.col.data() + m_.rowptr[i];
      const unsigned* c1 = m_.col.data() + m_.rowptr[i + 1];
      for (unsigned je = 0; je < nblkel_row; je++) {
        const double* in = emat + ((size_t)ie * nblkel_row + je) * bb;
        double* out;
        if (ie == je) {
          out = m_.dia.data() + (size_t)i * bb;
        } else {
          const unsigned* q = std::lower_bound(c0, c1, blkel_row[je]);
          if (q == c1 || *q != blkel_row[je]) continue;
          out = m_.val.data() + (size_t)(q - m_.col.data()) * bb;
        }
        for (size_t t = 0; t < bb; t++) out[t] += in[t];
      }
    }
    return true;
  }
  // y = beta*y + alpha*A*x (matdia_blkcrs.cpp:921-1094; per-block lengths :931-960)
  bool MatVec(double alpha, const CVector_Blk& x, double beta, CVector_Blk& y) const {
    if (fh_) {
      dfm_fbcrs* h = fhandle();
      const double* xp = x.device_ptr();
      return check(dfm_fbcrs_matvec(h, alpha, xp, beta, y.device_ptr()), "MatVec") == DFM_OK;
    }
    dfm_bcrs* h = handle();
    const double* xp = x.device_ptr();
    return check(dfm_bcrs_matvec(h, alpha, xp, beta, y.device_ptr()), "MatVec") == DFM_OK;
  }
  bool SetBoundaryCondition(const CBCFlag& flag) {
    if (fh_) return check(dfm_fbcrs_set_bc(fhandle(), flag.data()), "SetBoundaryCondition") == DFM_OK;
    return check(dfm_bcrs_set_bc(handle(), flag.data()), "SetBoundaryCondition") == DFM_OK;
  }
  // the pattern and values of one row on the host (mat_blkcrs.h:150-173,
  // matdia_blkcrs.h:102-111); the non-const forms count as writes.  With
  // per-block lengths the values of a row are its blocks one after the other.
  const unsigned* GetPtrIndPSuP(unsigned ipoin, unsigned& npsup) const {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return m_.col.data() + m_.rowptr[ipoin];
  }
  const double* GetPtrValPSuP(unsigned ipoin, unsigned& npsup) const {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return host_values(false).data() + val_offset(m_.rowptr[ipoin]);
  }
  double* GetPtrValPSuP(unsigned ipoin, unsigned& npsup) {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return host_values(true).data() + val_offset(m_.rowptr[ipoin]);
  }
  const double* GetPtrValDia(unsigned ipoin) const {
    host_values(false);
    return m_.dia.data() + dia_offset(ipoin);
  }
  double* GetPtrValDia(unsigned ipoin) {
    host_values(true);
    return m_.dia.data() + dia_offset(ipoin);
  }
  // host copies of the values (CRS order) and of the pattern
  void GetValues(std::vector<double>& val, std::vector<double>& dia) const {
    host_values(false);
    val = m_.val;
    dia = m_.dia;
  }
  const std::vector<unsigned>& RowPtr() const { return m_.rowptr; }
  const std::vector<unsigned>& Col() const { return m_.col; }
  // the C-ABI handle with the device values current; the caller may change
  // them, so the host copy is dropped.  handle(): fixed block length only;
  // fhandle(): per-block lengths only.
  dfm_bcrs* handle() const {
    if (fh_) throw Error("CMatDia_BlkCrs: per-block lengths have no fixed-length handle", DFM_NOT_IMPLEMENTED);
    flush();
    m_.valid = false;
    return h_;
  }
  dfm_fbcrs* fhandle() const {
    if (!fh_) throw Error("CMatDia_BlkCrs: a fixed block length has no per-block-length handle", DFM_BAD_ARG);
    flush();
    m_.valid = false;
    return fh_;
  }
  bool flex() const { return fh_ != nullptr; }

 private:
  size_t nval() const { return fh_ ? m_.valptr.back() : m_.col.size() * len_ * len_; }
  size_t ndia() const { return fh_ ? m_.diaptr.back() : (size_t)nblk_ * len_ * len_; }
  size_t val_offset(unsigned q) const { return fh_ ? m_.valptr[q] : (size_t)q * len_ * len_; }
  size_t dia_offset(unsigned i) const { return fh_ ? m_.diaptr[i] : (size_t)i * len_ * len_; }
  void flex_layout() {
    unsigned nb = 0, ndof = 0, ncrs = 0, nv = 0, nd = 0;
    check(dfm_fbcrs_info(fh_, &nb, &ndof, &ncrs, &nv, &nd), "CMatDia_BlkCrs");
    m_.col.resize(ncrs);
    m_.valptr.resize(ncrs + 1);
    m_.diaptr.resize(nblk_ + 1);
    check(dfm_fbcrs_layout(fh_, nullptr, m_.rowptr.data(), m_.col.data(), m_.valptr.data(), m_.diaptr.data()),
          "CMatDia_BlkCrs");
    m_.valid = m_.dirty = false;
  }
  bool flex_mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
                   unsigned blksize, const double* emat) {
    if (nblkel_col != 1 || nblkel_row != 1) return false;
    const unsigned i = blkel_col[0], j = blkel_row[0];
    if (i >= nblk_ || j >= nblk_ || blksize != LenBlkCol(i) * LenBlkCol(j)) return false;
    host_values(true);
    double* out;
    if (i == j) {
      out = m_.dia.data() + m_.diaptr[i];
    } else {
      const unsigned* c0 = m_.col.data() + m_.rowptr[i];
      const unsigned* c1 = m_.col.data() + m_.rowptr[i + 1];
      const unsigned* q = std::lower_bound(c0, c1, j);
      if (q == c1 || *q != j) return false;
      out = m_.val.data() + m_.valptr[q - m_.col.data()];
    }
    for (unsigned t = 0; t < blksize; t++) out[t] += emat[t];
    return true;
  }
  std::vector<double>& host_values(bool write) const {
    if (!m_.valid) {
      m_.val.resize(nval());
      m_.dia.resize(ndia());
      if (fh_) check(dfm_fbcrs_get_values(fh_, m_.val.data(), m_.dia.data()), "CMatDia_BlkCrs values");
      else check(dfm_bcrs_get_values(h_, m_.val.data(), m_.dia.data()), "CMatDia_BlkCrs values");
      m_.valid = true;
      m_.dirty = false;
    }
    if (write) m_.dirty = true;
    return m_.val;
  }
  void flush() const {
    if (!m_.dirty) return;
    if (fh_) check(dfm_fbcrs_set_values(fh_, m_.val.data(), m_.dia.data()), "CMatDia_BlkCrs values");
    else check(dfm_bcrs_set_values(h_, m_.val.data(), m_.dia.data()), "CMatDia_BlkCrs values");
    m_.dirty = false;
  }
  unsigned nblk_, len_;
  DofPtr dofptr_;
  dfm_bcrs* h_ = nullptr;
  dfm_fbcrs* fh_ = nullptr;
  mutable detail::ValMirror m_;
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
    m_.rowptr.assign(nblk_col + 1, 0);
  }
  virtual ~CMat_BlkCrs() {
    if (h_) dfm_bmat_destroy(h_);
  }
  CMat_BlkCrs(const CMat_BlkCrs&) = delete;
  CMat_BlkCrs& operator=(const CMat_BlkCrs&) = delete;
  unsigned NBlkMatCol() const { return nblk_col_; }
  unsigned NBlkMatRow() const { return nblk_row_; }
  int LenBlkCol() const { return (int)len_col_; }
  int LenBlkRow() const { return (int)len_row_; }
  unsigned NCrs() const { return (unsigned)m_.col.size(); }
  // AddPattern(const Com::CIndexedArray&) (mat_blkcrs.cpp:280-414)
  bool AddPattern(const Com::CIndexedArray& crs) { return AddPattern(crs.index, crs.array); }
  bool AddPattern(const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    flush();
    static const unsigned none = 0;
    const bool ok = check(dfm_bmat_add_pattern(h_, (unsigned)index.size() - 1, index.data(),
                                               array.empty() ? &none : array.data()),
                          "CMat_BlkCrs::AddPattern") == DFM_OK;
    refresh_pattern();
    return ok;
  }
  bool SetZero() {
    m_.val.assign(m_.col.size() * len_col_ * len_row_, 0.0);
    m_.valid = m_.dirty = true;
    return true;
  }
  // Mearge (mat_blkcrs.cpp:754-828): element block (i, j) at
  // emat[(i*nblkel_row + j)*blksize] into the entry (blkel_col[i],
  // blkel_row[j]) where the pattern has it
  bool Mearge(unsigned nblkel_col, const unsigned* blkel_col, unsigned nblkel_row, const unsigned* blkel_row,
              unsigned blksize, const double* emat) {
    const size_t bs = (size_t)len_col_ * len_row_;
    if (blksize != bs) return false;
    for (unsigned i = 0; i < nblkel_col; i++)
      if (blkel_col[i] >= nblk_col_) return false;
    for (unsigned j = 0; j < nblkel_row; j++)
      if (blkel_row[j] >= nblk_row_) return false;
    host_values(true);
    for (unsigned ie = 0; ie < nblkel_col; ie++) {
      const unsigned i = blkel_col[ie];
      const unsigned* c0 = m_.col.data() + m_.rowptr[i];
      const unsigned* c1 = m_.col.data() + m_.rowptr[i + 1];
      for (unsigned je = 0; je < nblkel_row; je++) {
        const unsigned* q = std::lower_bound(c0, c1, blkel_row[je]);
        if (q == c1 || *q != blkel_row[je]) continue;
        const double* in = emat + ((size_t)ie * nblkel_row + je) * bs;
        double* out = m_.val.data() + (size_t)(q - m_.col.data()) * bs;
        for (size_t t = 0; t < bs; t++) out[t] += in[t];
      }
    }
    return true;
  }
  // y = beta*y + alpha*A*x (mat_blkcrs.cpp:955-1174, isnt_trans = true)
  bool MatVec(double alpha, const CVector_Blk& x, double beta, CVector_Blk& y, bool isnt_trans = true) const {
    if (!isnt_trans) throw Error("CMat_BlkCrs::MatVec: the transposed product is not provided", DFM_NOT_IMPLEMENTED);
    dfm_bmat* h = handle();
    const double* xp = x.device_ptr();
    return check(dfm_bmat_matvec(h, alpha, xp, beta, y.device_ptr()), "CMat_BlkCrs::MatVec") == DFM_OK;
  }
  bool SetBoundaryCondition_Row(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_row(handle(), bc.data()), "SetBoundaryCondition_Row") == DFM_OK;
  }
  bool SetBoundaryCondition_Colum(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_col(handle(), bc.data()), "SetBoundaryCondition_Colum") == DFM_OK;
  }
  bool SetBoundaryConditionInverse_Colum(const CBCFlag& bc) {
    return check(dfm_bmat_set_bc_inverse_col(handle(), bc.data()), "SetBoundaryConditionInverse_Colum") == DFM_OK;
  }
  // SetPatternBoundary (mat_blkcrs.cpp:1177-1265): the boundary matrix pattern of CLinearSystem_Save
  bool SetPatternBoundary(const CMatDia_BlkCrs& A, const CBCFlag& bc_col, const CBCFlag& bc_row) {
    flush();
    const bool ok = check(dfm_bmat_set_pattern_boundary(h_, A.handle(), bc_col.data(), bc_row.data()),
                          "SetPatternBoundary") == DFM_OK;
    refresh_pattern();
    return ok;
  }
  const unsigned* GetPtrIndPSuP(unsigned ipoin, unsigned& npsup) const {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return m_.col.data() + m_.rowptr[ipoin];
  }
  const double* GetPtrValPSuP(unsigned ipoin, unsigned& npsup) const {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return host_values(false).data() + (size_t)m_.rowptr[ipoin] * len_col_ * len_row_;
  }
  double* GetPtrValPSuP(unsigned ipoin, unsigned& npsup) {
    npsup = m_.rowptr[ipoin + 1] - m_.rowptr[ipoin];
    return host_values(true).data() + (size_t)m_.rowptr[ipoin] * len_col_ * len_row_;
  }
  void GetValues(std::vector<double>& val) const { val = host_values(false); }
  dfm_bmat* handle() const {
    flush();
    m_.valid = false;
    return h_;
  }

 private:
  void refresh_pattern() {
    unsigned nb = 0, len = 0, ncrs = 0;
    check(dfm_bmat_info(h_, &nb, &len, &ncrs), "CMat_BlkCrs pattern");
    m_.col.resize(ncrs);
    check(dfm_bmat_get_pattern(h_, m_.rowptr.data(), m_.col.data()), "CMat_BlkCrs pattern");
    m_.valid = m_.dirty = false;
  }
  std::vector<double>& host_values(bool write) const {
    if (!m_.valid) {
      m_.val.resize(m_.col.size() * len_col_ * len_row_);
      if (!m_.val.empty()) check(dfm_bmat_get_values(h_, m_.val.data()), "CMat_BlkCrs values");
      m_.valid = true;
      m_.dirty = false;
    }
    if (write) m_.dirty = true;
    return m_.val;
  }
  void flush() const {
    if (!m_.dirty) return;
    check(dfm_bmat_set_values(h_, m_.val.empty() ? nullptr : m_.val.data()), "CMat_BlkCrs values");
    m_.dirty = false;
  }
  unsigned nblk_col_, len_col_, nblk_row_, len_row_;
  dfm_bmat* h_ = nullptr;
  mutable detail::ValMirror m_;
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
  // (with per-block lengths: the flex factor, matdiafrac_blkcrs.cpp:227-245)
  CMatDiaFrac_BlkCrs(int lev_fill, const CMatDia_BlkCrs& A) : src_(&A) {
    if (A.flex()) {
      check(dfm_filu_create(A.fhandle(), lev_fill, &fh_), "CMatDiaFrac_BlkCrs");
      return;
    }
    check(dfm_ilu_create(A.handle(), lev_fill, DFM_ORDER_NATURAL, nullptr, 0, &h_), "CMatDiaFrac_BlkCrs");
  }
  // ordered variant, order[new] = old (fixed block length)
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
    if (fh_) dfm_filu_destroy(fh_);
  }
  CMatDiaFrac_BlkCrs(const CMatDiaFrac_BlkCrs&) = delete;
  CMatDiaFrac_BlkCrs& operator=(const CMatDiaFrac_BlkCrs&) = delete;
  // SetValue(A) (matdiafrac_blkcrs.cpp:1518-1537): the values of A, then the
  // factorisation; true whatever the factorisation met, as the reference,
  // which ignores DoILUDecomp's result there
  bool SetValue(const CMatDia_BlkCrs& A) {
    nsing_ = 0;
    if (fh_) factored_ = check(dfm_filu_set_value(fh_, A.fhandle(), &nsing_), "SetValue") == DFM_OK;
    else factored_ = check(dfm_ilu_set_value(h_, A.handle(), &nsing_), "SetValue") == DFM_OK;
    return true;
  }
  // DoILUDecomp (matdiafrac_blkcrs.cpp:769-1102): the device factorises in
  // SetValue; this reports its outcome -- false when it stopped at the 11th
  // singular pivot, leaving the factor as the reference leaves it
  bool DoILUDecomp() { return factored_; }
  // singular pivots met by the last factorisation
  int NSingular() const { return nsing_; }
  bool ForwardSubstitution(CVector_Blk& v) const {
    if (fh_) return check(dfm_filu_forward(fh_, v.device_ptr()), "ForwardSubstitution") == DFM_OK;
    return check(dfm_ilu_forward(h_, v.device_ptr()), "ForwardSubstitution") == DFM_OK;
  }
  bool BackwardSubstitution(CVector_Blk& v) const {
    if (fh_) return check(dfm_filu_backward(fh_, v.device_ptr()), "BackwardSubstitution") == DFM_OK;
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
    unsigned n = 0, nv = 0;
    int a, b, c;
    if (fh_) check(dfm_filu_info(fh_, &n, &nv, &a, &b), "NCrs");
    else check(dfm_ilu_info(h_, &n, &a, &b, &c), "NCrs");
    return n;
  }
  dfm_ilu* handle() const { return h_; }
  dfm_filu* fhandle() const { return fh_; }

 private:
  const CMatDia_BlkCrs* src_ = nullptr;
  int ordering_ = DFM_ORDER_NATURAL;
  dfm_ilu* h_ = nullptr;
  dfm_filu* fh_ = nullptr;
  int nsing_ = 0;
  bool factored_ = false;
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
  return f && !mat.flex() && f->source() == &mat && f->ordering() == DFM_ORDER_NATURAL ? f : nullptr;
}
}  // namespace detail

// Solve_CG (solver_mat_iter.cpp:37-109).  As in the reference, the
// convergence test reads the residual norm from before the current step, and
// a converged return leaves that step out of u (it is already in r).
inline bool Solve_CG(double& conv_ratio, unsigned& iteration, const CMatDia_BlkCrs& mat, CVector_Blk& r_vec,
                     CVector_Blk& u_vec) {
  const double tol = conv_ratio;
  const unsigned mx_iter = iteration;
  u_vec.SetVectorZero();
  double rr = r_vec * r_vec;
  if (rr < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr;
  CVector_Blk p_vec = CVector_Blk::like(r_vec), Ap_vec = CVector_Blk::like(r_vec);
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
  u_vec.SetVectorZero();
  const double rr0 = r_vec * r_vec;
  if (rr0 < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  CVector_Blk p_vec = CVector_Blk::like(r_vec), z_vec = CVector_Blk::like(r_vec);
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
  u_vec.SetVectorZero();
  const double rr0 = r_vec * r_vec;
  if (rr0 < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  CVector_Blk s_vec = CVector_Blk::like(r_vec), p_vec = CVector_Blk::like(r_vec);
  CVector_Blk AMp_vec = CVector_Blk::like(r_vec), AMs_vec = CVector_Blk::like(r_vec), r0_vec = CVector_Blk::like(r_vec);
  std::unique_ptr<CVector_Blk> Mp_vec, Ms_vec;
  if (precond) {
    Mp_vec.reset(new CVector_Blk(CVector_Blk::like(r_vec)));
    Ms_vec.reset(new CVector_Blk(CVector_Blk::like(r_vec)));
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
