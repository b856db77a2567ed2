// delfem_amd/ls.hpp -- C++ classes with the names and semantics of DelFEM's
// LsSol layer over the dfm.h C-ABI: one or several segments (each a square
// block-CRS matrix with its residual, update and boundary-condition flags,
// rectangular CMat_BlkCrs blocks between segments):
//
//   LsSol::CLinearSystem                 include/delfem/ls/linearsystem.h:51-176
//   LsSol::CPreconditioner_ILU           include/delfem/ls/preconditioner.h:59-108
//   LsSol::CLinearSystemPreconditioner   include/delfem/ls/preconditioner.h:115-145
//   LsSol::Solve_PCG / Solve_PBiCGSTAB   include/delfem/ls/solver_ls_iter.h:38-46
//
// Vector ids follow the reference: -1 residual (right-hand side), -2 update
// (solution), 0.. scratch vectors (ReSizeTmpVecSolver).  The solvers solve
// A * update = residual from update = 0; conv_ratio in = tolerance on
// ||r|| / ||r0||, out = the ratio reached; iteration in = maximum, out = the
// reference's count.  An extra `precision` argument (DFM_PREC_*) selects the
// mixed-precision paths; the default is the reference's FP64.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "matvec.hpp"

namespace delfem_amd {
namespace LsSol {

using delfem_amd::MatVec::CBCFlag;
using delfem_amd::MatVec::CMat_BlkCrs;
using delfem_amd::MatVec::CMatDia_BlkCrs;
using delfem_amd::MatVec::CVector_Blk;

class CLinearSystem {
 public:
  CLinearSystem() = default;
  ~CLinearSystem() {
    if (ms_) dfm_mseg_destroy(ms_);
  }
  CLinearSystem(const CLinearSystem&) = delete;
  CLinearSystem& operator=(const CLinearSystem&) = delete;
  // AddLinSysSeg(nnode, len) (linearsystem.cpp:138-167): a new segment with its
  // diagonal matrix, residual, update and flags; returns its index.  Vectors
  // are one device array per vector id, segments one after the other.
  int AddLinSysSeg(unsigned nnode, unsigned len) {
    Seg g;
    g.nnode = nnode;
    g.len = len;
    g.off = total_;
    g.mat.reset(new CMatDia_BlkCrs(nnode, len));
    g.bc.reset(new CBCFlag(nnode, len));
    segs_.push_back(std::move(g));
    total_ += (size_t)nnode * len;
    for (auto& g2 : segs_) g2.nd.resize(segs_.size());
    reset_mseg();
    res_ = make_vec();
    upd_ = make_vec();
    const size_t ntmp = tmp_.size();
    tmp_.clear();
    ReSizeTmpVecSolver((unsigned)ntmp);
    return (int)segs_.size() - 1;
  }
  unsigned GetNLinSysSeg() const { return (unsigned)segs_.size(); }
  // AddMat_Dia(ilss, CIndexedArray): the pattern of segment ilss's matrix
  bool AddMat_Dia(unsigned ilss, const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    return ilss < segs_.size() && segs_[ilss].mat->AddPattern(index, array);
  }
  // AddMat_NonDia(ilss, jlss, CIndexedArray) (linearsystem.cpp:450-520): the
  // block coupling segment ilss's rows to segment jlss's columns
  bool AddMat_NonDia(unsigned ilss, unsigned jlss, const std::vector<unsigned>& index,
                     const std::vector<unsigned>& array) {
    if (ilss >= segs_.size() || jlss >= segs_.size() || ilss == jlss) return false;
    auto& m = segs_[ilss].nd[jlss];
    if (!m) m.reset(new CMat_BlkCrs(segs_[ilss].nnode, segs_[ilss].len, segs_[jlss].nnode, segs_[jlss].len));
    reset_mseg();
    return m->AddPattern(index, array);
  }
  bool IsMatrix(unsigned ilss, unsigned jlss) const {
    return ilss == jlss ? ilss < segs_.size() : (bool)segs_.at(ilss).nd.at(jlss);
  }
  // InitializeMarge (linearsystem.cpp:74-88): zero matrices, residual, update
  void InitializeMarge() {
    for (auto& g : segs_) {
      g.mat->SetZero();
      for (auto& m : g.nd)
        if (m) m->SetZero();
    }
    res_.whole->SetVectorZero();
    upd_.whole->SetVectorZero();
  }
  // FinalizeMarge (linearsystem.cpp:92-128): Dirichlet conditions on the
  // diagonal matrices, then the rows and the columns of the off-diagonal
  // blocks, the residual zeroed at fixed dofs; returns ||residual||
  double FinalizeMarge() {
    const unsigned n = GetNLinSysSeg();
    for (auto& g : segs_) g.mat->SetBoundaryCondition(*g.bc);
    for (unsigned i = 0; i < n; i++)
      for (unsigned j = 0; j < n; j++)
        if (segs_[i].nd[j]) segs_[i].nd[j]->SetBoundaryCondition_Row(*segs_[i].bc);
    for (unsigned j = 0; j < n; j++)
      for (unsigned i = 0; i < n; i++)
        if (segs_[i].nd[j]) segs_[i].nd[j]->SetBoundaryCondition_Colum(*segs_[j].bc);
    double sq = 0.0;
    for (unsigned i = 0; i < n; i++) {
      segs_[i].bc->SetZeroToBCDof(*res_.seg[i]);
      sq += res_.seg[i]->GetSquaredVectorNorm();
    }
    return std::sqrt(sq);
  }
  CMatDia_BlkCrs& GetMatrix(unsigned ilss = 0) { return *segs_.at(ilss).mat; }
  const CMatDia_BlkCrs& GetMatrix(unsigned ilss = 0) const { return *segs_.at(ilss).mat; }
  CMat_BlkCrs& GetMatrix(unsigned ilss, unsigned jlss) { return *segs_.at(ilss).nd.at(jlss); }
  // GetVector(iv, ilss): -1 residual, -2 update, 0.. scratch
  CVector_Blk& GetVector(int iv, unsigned ilss) { return *vecs(iv).seg.at(ilss); }
  CVector_Blk& GetResidual(unsigned ilss = 0) { return *res_.seg.at(ilss); }
  CVector_Blk& GetUpdate(unsigned ilss = 0) { return *upd_.seg.at(ilss); }
  CBCFlag& GetBCFlag(unsigned ilss = 0) { return *segs_.at(ilss).bc; }

  unsigned GetTmpVectorArySize() const { return (unsigned)tmp_.size(); }
  bool ReSizeTmpVecSolver(unsigned n) {
    while (tmp_.size() < n) tmp_.push_back(make_vec());
    tmp_.resize(n);
    return true;
  }
  // DOT (linearsystem.cpp:335-364): the segments' dot products added in order
  double DOT(int iv1, int iv2) {
    double s = 0.0;
    for (unsigned i = 0; i < GetNLinSysSeg(); i++) s += *vecs(iv1).seg[i] * *vecs(iv2).seg[i];
    return s;
  }
  bool COPY(int iv1, int iv2) {
    if (iv1 != iv2) *vecs(iv2).whole = *vecs(iv1).whole;
    return true;
  }
  bool SCAL(double alpha, int iv) {
    *vecs(iv).whole *= alpha;
    return true;
  }
  bool AXPY(double alpha, int iv1, int iv2) {
    if (alpha != 0.0) vecs(iv2).whole->AXPY(alpha, *vecs(iv1).whole);
    return true;
  }
  // MATVEC (linearsystem.cpp:250-296)
  bool MATVEC(double alpha, int iv1, double beta, int iv2) {
    if (GetNLinSysSeg() == 1) return segs_[0].mat->MatVec(alpha, *vecs(iv1).seg[0], beta, *vecs(iv2).seg[0]);
    return check(dfm_mseg_matvec(mseg(), alpha, vecs(iv1).whole->device_ptr(), beta, vecs(iv2).whole->device_ptr()),
                 "CLinearSystem::MATVEC") == DFM_OK;
  }
  // the whole vector of an id (all segments)
  CVector_Blk& vec(int iv) { return *vecs(iv).whole; }
  // the C-ABI multi-segment handle over the segments' matrices
  dfm_mseg* mseg() {
    if (!ms_) {
      const unsigned n = GetNLinSysSeg();
      std::vector<unsigned> nb(n), len(n);
      for (unsigned i = 0; i < n; i++) {
        nb[i] = segs_[i].nnode;
        len[i] = segs_[i].len;
      }
      check(dfm_mseg_create(n, nb.data(), len.data(), &ms_), "CLinearSystem segments");
      for (unsigned i = 0; i < n; i++) {
        check(dfm_mseg_set_dia(ms_, i, segs_[i].mat->handle()), "CLinearSystem segments");
        for (unsigned j = 0; j < n; j++)
          if (segs_[i].nd[j]) check(dfm_mseg_set_nondia(ms_, i, j, segs_[i].nd[j]->handle()), "CLinearSystem segments");
      }
    }
    return ms_;
  }

 private:
  struct Seg {
    unsigned nnode = 0, len = 0;
    size_t off = 0;
    std::unique_ptr<CMatDia_BlkCrs> mat;
    std::unique_ptr<CBCFlag> bc;
    std::vector<std::unique_ptr<CMat_BlkCrs>> nd;
  };
  struct Vec {
    std::unique_ptr<CVector_Blk> whole;
    std::vector<std::unique_ptr<CVector_Blk>> seg;  // views into whole
  };
  Vec make_vec() {
    Vec v;
    v.whole.reset(new CVector_Blk((unsigned)total_, 1));
    for (auto& g : segs_) v.seg.emplace_back(new CVector_Blk(g.nnode, g.len, v.whole->device_ptr() + g.off));
    return v;
  }
  Vec& vecs(int iv) {
    if (iv == -1) return res_;
    if (iv == -2) return upd_;
    return tmp_.at((size_t)iv);
  }
  void reset_mseg() {
    if (ms_) dfm_mseg_destroy(ms_);
    ms_ = nullptr;
  }
  std::vector<Seg> segs_;
  size_t total_ = 0;
  Vec res_, upd_;
  std::vector<Vec> tmp_;
  dfm_mseg* ms_ = nullptr;
};

class CPreconditioner_ILU {
 public:
  CPreconditioner_ILU() = default;
  CPreconditioner_ILU(CLinearSystem& ls, unsigned nlev = 0) {
    SetFillInLevel((int)nlev);
    SetLinearSystem(ls);
  }
  ~CPreconditioner_ILU() { Clear(); }
  CPreconditioner_ILU(const CPreconditioner_ILU&) = delete;
  CPreconditioner_ILU& operator=(const CPreconditioner_ILU&) = delete;
  void Clear() {
    if (h_) dfm_ilu_destroy(h_);
    if (ms_) dfm_mseg_ilu_destroy(ms_);
    h_ = nullptr;
    ms_ = nullptr;
  }
  // SetFillInLevel(lev, ilss): one level for every block (ilss = -1)
  void SetFillInLevel(int lev, int ilss0 = -1) {
    if (ilss0 != -1) throw Error("SetFillInLevel: per-segment levels are not provided", DFM_NOT_IMPLEMENTED);
    lev_ = lev;
  }
  // SetOrdering(order[new] = old); empty = natural (one segment only)
  void SetOrdering(const std::vector<int>& aind) {
    order_ = aind;
    ordering_ = aind.empty() ? DFM_ORDER_NATURAL : DFM_ORDER_GIVEN;
  }
  // device-computed orderings (DFM_ORDER_MULTICOLOUR, DFM_ORDER_BLOCK_MULTICOLOUR; one segment)
  void SetDeviceOrdering(int ordering, unsigned block_nodes = 0) {
    order_.clear();
    ordering_ = ordering;
    block_nodes_ = block_nodes;
  }
  // symbolic factorisation (preconditioner.cpp:57-211); several segments:
  // ILU(lev) across them in natural order
  void SetLinearSystem(CLinearSystem& ls) {
    Clear();
    if (ls.GetNLinSysSeg() == 1) {
      check(dfm_ilu_create(ls.GetMatrix().handle(), lev_, ordering_, order_.empty() ? nullptr : order_.data(),
                           block_nodes_, &h_),
            "CPreconditioner_ILU::SetLinearSystem");
      return;
    }
    if (ordering_ != DFM_ORDER_NATURAL)
      throw Error("CPreconditioner_ILU: orderings need a single segment", DFM_NOT_IMPLEMENTED);
    check(dfm_mseg_ilu_create(ls.mseg(), lev_, &ms_), "CPreconditioner_ILU::SetLinearSystem");
  }
  bool SetValue(CLinearSystem& ls) {
    int nsing = 0;
    if (ms_) return check(dfm_mseg_ilu_set_value(ms_, &nsing), "CPreconditioner_ILU::SetValue") == DFM_OK;
    return check(dfm_ilu_set_value(h_, ls.GetMatrix().handle(), &nsing), "CPreconditioner_ILU::SetValue") == DFM_OK;
  }
  bool SolvePrecond(CLinearSystem& ls, unsigned iv) {
    if (ms_) return check(dfm_mseg_ilu_precond(ms_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
    return check(dfm_ilu_solve_precond(h_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
  }
  dfm_ilu* handle() const { return h_; }
  dfm_msilu* mseg_handle() const { return ms_; }

 private:
  int lev_ = 0, ordering_ = DFM_ORDER_NATURAL;
  unsigned block_nodes_ = 0;
  std::vector<int> order_;
  dfm_ilu* h_ = nullptr;
  dfm_msilu* ms_ = nullptr;
};

class CLinearSystemPreconditioner {
 public:
  CLinearSystemPreconditioner(CLinearSystem& ls, CPreconditioner_ILU& prec) : ls(ls), prec(prec) {}
  CLinearSystem& ls;
  CPreconditioner_ILU& prec;
};

namespace detail {
inline bool solve(bool bicg, double& conv_ratio, unsigned& iteration, CLinearSystemPreconditioner& lsp,
                  int precision) {
  int rc;
  if (lsp.prec.mseg_handle()) {
    if (precision != DFM_PREC_FP64)
      throw Error("several segments: FP64 solvers only", DFM_NOT_IMPLEMENTED);
    rc = (bicg ? dfm_mseg_solve_pbicgstab : dfm_mseg_solve_pcg)(lsp.ls.mseg(), lsp.prec.mseg_handle(),
                                                                lsp.ls.vec(-1).device_ptr(),
                                                                lsp.ls.vec(-2).device_ptr(), &conv_ratio, &iteration);
  } else {
    dfm_solve_opts o{};
    o.precision = precision;
    rc = (bicg ? dfm_solve_pbicgstab : dfm_solve_pcg)(lsp.ls.GetMatrix().handle(), lsp.prec.handle(),
                                                      lsp.ls.GetResidual().device_ptr(),
                                                      lsp.ls.GetUpdate().device_ptr(), &conv_ratio, &iteration, &o);
  }
  // Solve_PCG is false at max_iter; Solve_PBiCGSTAB is true either way
  // (solver_ls_iter.cpp:168-171, 407)
  return check(rc, bicg ? "Solve_PBiCGSTAB" : "Solve_PCG") == DFM_OK || bicg;
}
}  // namespace detail

// LsSol::Solve_PCG (src/ls/solver_ls_iter.cpp:100-172)
inline bool Solve_PCG(double& conv_ratio, unsigned& iteration, CLinearSystemPreconditioner& lsp,
                      int precision = DFM_PREC_FP64) {
  return detail::solve(false, conv_ratio, iteration, lsp, precision);
}
// LsSol::Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:291-408)
inline bool Solve_PBiCGSTAB(double& conv_ratio, unsigned& iteration, CLinearSystemPreconditioner& lsp,
                            int precision = DFM_PREC_FP64) {
  return detail::solve(true, conv_ratio, iteration, lsp, precision);
}

}  // namespace LsSol
}  // namespace delfem_amd

#ifndef DELFEM_AMD_NO_ALIASES
namespace LsSol = delfem_amd::LsSol;
#endif
