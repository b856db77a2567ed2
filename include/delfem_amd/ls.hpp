// delfem_amd/ls.hpp -- C++ classes with the names and semantics of DelFEM's
// LsSol layer over the dfm.h C-ABI, for one segment (one square block-CRS
// matrix with its residual, update and boundary-condition flags):
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
using delfem_amd::MatVec::CMatDia_BlkCrs;
using delfem_amd::MatVec::CVector_Blk;

class CLinearSystem {
 public:
  CLinearSystem() = default;
  // AddLinSysSeg(nnode, len): one segment; returns its index (0)
  int AddLinSysSeg(unsigned nnode, unsigned len) {
    if (mat_) return -1;
    nnode_ = nnode;
    len_ = len;
    mat_.reset(new CMatDia_BlkCrs(nnode, len));
    res_.reset(new CVector_Blk(nnode, len));
    upd_.reset(new CVector_Blk(nnode, len));
    bc_.reset(new CBCFlag(nnode, len));
    return 0;
  }
  // AddMat_Dia(ilss, CIndexedArray): the pattern of the diagonal matrix
  bool AddMat_Dia(unsigned ilss, const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    return ilss == 0 && mat_ && mat_->AddPattern(index, array);
  }
  // InitializeMarge (linearsystem.cpp:74-88): zero matrix, residual, update
  void InitializeMarge() {
    mat_->SetZero();
    res_->SetVectorZero();
    upd_->SetVectorZero();
  }
  // FinalizeMarge (linearsystem.cpp:92-128): Dirichlet rows/columns, zero the
  // residual at fixed dofs, return ||residual||
  double FinalizeMarge() {
    mat_->SetBoundaryCondition(*bc_);
    bc_->SetZeroToBCDof(*res_);
    return std::sqrt(res_->GetSquaredVectorNorm());
  }
  CMatDia_BlkCrs& GetMatrix(unsigned ilss = 0) { return (void)ilss, *mat_; }
  const CMatDia_BlkCrs& GetMatrix(unsigned ilss = 0) const { return (void)ilss, *mat_; }
  CVector_Blk& GetResidual(unsigned ilss = 0) { return (void)ilss, *res_; }
  CVector_Blk& GetUpdate(unsigned ilss = 0) { return (void)ilss, *upd_; }
  CBCFlag& GetBCFlag(unsigned ilss = 0) { return (void)ilss, *bc_; }

  unsigned GetTmpVectorArySize() const { return (unsigned)tmp_.size(); }
  bool ReSizeTmpVecSolver(unsigned n) {
    while (tmp_.size() < n) tmp_.emplace_back(new CVector_Blk(nnode_, len_));
    tmp_.resize(n);
    return true;
  }
  double DOT(int iv1, int iv2) { return vec(iv1) * vec(iv2); }
  bool COPY(int iv1, int iv2) {
    if (iv1 != iv2) vec(iv2) = vec(iv1);
    return true;
  }
  bool SCAL(double alpha, int iv) {
    vec(iv) *= alpha;
    return true;
  }
  bool AXPY(double alpha, int iv1, int iv2) {
    if (alpha != 0.0) vec(iv2).AXPY(alpha, vec(iv1));
    return true;
  }
  bool MATVEC(double alpha, int iv1, double beta, int iv2) { return mat_->MatVec(alpha, vec(iv1), beta, vec(iv2)); }
  CVector_Blk& vec(int iv) {
    if (iv == -1) return *res_;
    if (iv == -2) return *upd_;
    return *tmp_.at((size_t)iv);
  }

 private:
  unsigned nnode_ = 0, len_ = 0;
  std::unique_ptr<CMatDia_BlkCrs> mat_;
  std::unique_ptr<CVector_Blk> res_, upd_;
  std::unique_ptr<CBCFlag> bc_;
  std::vector<std::unique_ptr<CVector_Blk>> tmp_;
};

class CPreconditioner_ILU {
 public:
  CPreconditioner_ILU() = default;
  CPreconditioner_ILU(const CLinearSystem& ls, unsigned nlev = 0) {
    SetFillInLevel((int)nlev);
    SetLinearSystem(ls);
  }
  ~CPreconditioner_ILU() { Clear(); }
  CPreconditioner_ILU(const CPreconditioner_ILU&) = delete;
  CPreconditioner_ILU& operator=(const CPreconditioner_ILU&) = delete;
  void Clear() {
    if (h_) dfm_ilu_destroy(h_);
    h_ = nullptr;
  }
  void SetFillInLevel(int lev, int ilss0 = -1) {
    (void)ilss0;
    lev_ = lev;
  }
  // SetOrdering(order[new] = old); empty = natural
  void SetOrdering(const std::vector<int>& aind) {
    order_ = aind;
    ordering_ = aind.empty() ? DFM_ORDER_NATURAL : DFM_ORDER_GIVEN;
  }
  // device-computed orderings (DFM_ORDER_MULTICOLOUR, DFM_ORDER_BLOCK_MULTICOLOUR)
  void SetDeviceOrdering(int ordering, unsigned block_nodes = 0) {
    order_.clear();
    ordering_ = ordering;
    block_nodes_ = block_nodes;
  }
  void SetLinearSystem(const CLinearSystem& ls) {
    Clear();
    check(dfm_ilu_create(ls.GetMatrix().handle(), lev_, ordering_, order_.empty() ? nullptr : order_.data(),
                         block_nodes_, &h_),
          "CPreconditioner_ILU::SetLinearSystem");
  }
  bool SetValue(const CLinearSystem& ls) {
    int nsing = 0;
    return check(dfm_ilu_set_value(h_, ls.GetMatrix().handle(), &nsing), "CPreconditioner_ILU::SetValue") == DFM_OK;
  }
  bool SolvePrecond(CLinearSystem& ls, unsigned iv) {
    return check(dfm_ilu_solve_precond(h_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
  }
  dfm_ilu* handle() const { return h_; }

 private:
  int lev_ = 0, ordering_ = DFM_ORDER_NATURAL;
  unsigned block_nodes_ = 0;
  std::vector<int> order_;
  dfm_ilu* h_ = nullptr;
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
  dfm_solve_opts o{};
  o.precision = precision;
  const int rc = (bicg ? dfm_solve_pbicgstab : dfm_solve_pcg)(
      lsp.ls.GetMatrix().handle(), lsp.prec.handle(), lsp.ls.GetResidual().device_ptr(),
      lsp.ls.GetUpdate().device_ptr(), &conv_ratio, &iteration, &o);
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
