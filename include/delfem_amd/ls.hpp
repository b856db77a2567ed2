// delfem_amd/ls.hpp -- C++ classes with the names and semantics of DelFEM's
// LsSol layer over the dfm.h C-ABI: one or several segments (each a square
// block-CRS matrix with its residual, update and boundary-condition flags,
// rectangular CMat_BlkCrs blocks between segments):
//
//   LsSol::ILinearSystem_Sol,
//   LsSol::ILinearSystemPreconditioner_Sol include/delfem/ls/linearsystem_interface_solver.h:39-82
//   LsSol::CLinearSystem                   include/delfem/ls/linearsystem.h:51-176
//   LsSol::CPreconditioner                 include/delfem/ls/preconditioner.h:46-52
//   LsSol::CPreconditioner_ILU             include/delfem/ls/preconditioner.h:59-108
//   LsSol::CLinearSystemPreconditioner     include/delfem/ls/preconditioner.h:115-145
//   LsSol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
//                                          include/delfem/ls/solver_ls_iter.h:33-43
//
// Vector ids follow the reference: -1 residual (right-hand side), -2 update
// (solution), 0.. scratch vectors (ReSizeTmpVecSolver).  The solvers solve
// A * update = residual from update = 0; conv_ratio in = tolerance on
// ||r|| / ||r0||, out = the ratio reached; iteration in = maximum, out = the
// reference's count.  Solve_PCG / Solve_PBiCGSTAB on a
// CLinearSystemPreconditioner whose preconditioner is a CPreconditioner_ILU
// set up on its system run the library's device-resident solvers (an extra
// `precision` argument, DFM_PREC_*, selects the mixed-precision paths; the
// default is the reference's FP64); any other ILinearSystemPreconditioner_Sol
// runs the reference's loop over its virtual BLAS-1 calls.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <utility>
#include <vector>

#include "indexed_array.hpp"
#include "matvec.hpp"

namespace delfem_amd {
namespace LsSol {

using delfem_amd::MatVec::CBCFlag;
using delfem_amd::MatVec::CMat_BlkCrs;
using delfem_amd::MatVec::CMatDia_BlkCrs;
using delfem_amd::MatVec::CVector_Blk;

// The solvers' view of a linear system (linearsystem_interface_solver.h:39-58).
class ILinearSystem_Sol {
 public:
  virtual ~ILinearSystem_Sol() {}
  virtual unsigned GetTmpVectorArySize() const = 0;
  virtual bool ReSizeTmpVecSolver(unsigned size_new) = 0;
  virtual double DOT(int iv1, int iv2) = 0;                             // {v1} . {v2}
  virtual bool COPY(int iv1, int iv2) = 0;                              // {v2} := {v1}
  virtual bool SCAL(double alpha, int iv1) = 0;                         // {v1} := alpha {v1}
  virtual bool AXPY(double alpha, int iv1, int iv2) = 0;                // {v2} += alpha {v1}
  virtual bool MATVEC(double alpha, int iv1, double beta, int iv2) = 0; // {v2} := alpha A {v1} + beta {v2}
};

// ... with a preconditioner (linearsystem_interface_solver.h:63-82).
class ILinearSystemPreconditioner_Sol {
 public:
  virtual ~ILinearSystemPreconditioner_Sol() {}
  virtual unsigned GetTmpVectorArySize() const = 0;
  virtual bool ReSizeTmpVecSolver(unsigned size_new) = 0;
  virtual double DOT(int iv1, int iv2) = 0;
  virtual bool COPY(int iv1, int iv2) = 0;
  virtual bool SCAL(double alpha, int iv1) = 0;
  virtual bool AXPY(double alpha, int iv1, int iv2) = 0;
  virtual bool MATVEC(double alpha, int iv1, double beta, int iv2) = 0;
  virtual bool SolvePrecond(int iv) = 0;  // {v} := M^-1 {v}
};

namespace detail {
struct MsegDeleter {
  void operator()(dfm_mseg* p) const {
    if (p) dfm_mseg_destroy(p);
  }
};
}  // namespace detail

class CLinearSystem : public ILinearSystem_Sol {
 private:
  struct Seg {
    unsigned nnode = 0, len = 0;
    MatVec::DofPtr dofptr;  // per-block lengths, else NULL
    size_t off = 0;
    std::unique_ptr<CMatDia_BlkCrs> mat;
    std::unique_ptr<CBCFlag> bc;
    std::vector<std::unique_ptr<CMat_BlkCrs>> nd;
  };
 public:
  CLinearSystem() = default;
  ~CLinearSystem() override = default;
  CLinearSystem(const CLinearSystem&) = delete;
  CLinearSystem& operator=(const CLinearSystem&) = delete;

  // AddLinSysSeg(nnode, len) (linearsystem.cpp:138-167): a new segment with
  // its diagonal matrix, residual, update and flags; returns its index.  The
  // vectors of the segments already there keep their values and identity
  // (references handed out stay valid); all segments of one vector id share
  // one device array.
  int AddLinSysSeg(unsigned nnode, unsigned len) {
    Seg g;
    g.nnode = nnode;
    g.len = len;
    g.off = total_;
    g.mat.reset(new CMatDia_BlkCrs(nnode, len));
    g.bc.reset(new CBCFlag(nnode, len));
    return add_segment(std::move(g), (size_t)nnode * len);
  }
  // AddLinSysSeg(nnode, aLen) (linearsystem.cpp:170-201): a segment whose
  // block i has aLen[i] dofs (the matrix, vectors and flags with per-block
  // lengths).  Such a segment is the system's only one here: the
  // off-diagonal blocks that would couple it to others are not provided.
  int AddLinSysSeg(unsigned nnode, const std::vector<unsigned>& aLen) {
    if (aLen.size() != nnode) throw Error("CLinearSystem::AddLinSysSeg: aLen needs one length per block", DFM_BAD_ARG);
    Seg g;
    g.nnode = nnode;
    g.len = 0;
    g.off = total_;
    g.mat.reset(new CMatDia_BlkCrs(nnode, aLen));
    g.bc.reset(new CBCFlag(nnode, aLen));
    g.dofptr = g.mat->dofptr();
    const size_t ndof = g.dofptr->back();
    return add_segment(std::move(g), ndof);
  }

 private:
  int add_segment(Seg&& g, size_t ndof) {
    if (!segs_.empty() && (g.dofptr || segs_[0].dofptr))
      throw Error("CLinearSystem::AddLinSysSeg: a segment of per-block lengths must be the only one",
                  DFM_NOT_IMPLEMENTED);
    segs_.push_back(std::move(g));
    const size_t total_old = total_;
    total_ += ndof;
    for (auto& s : segs_) s.nd.resize(segs_.size());
    grow(res_, total_old);
    grow(upd_, total_old);
    for (auto& t : tmp_) grow(t, total_old);
    structure_changed();
    return (int)segs_.size() - 1;
  }

 public:
  unsigned GetNLinSysSeg() const { return (unsigned)segs_.size(); }
  unsigned GetBlkSizeLinSeg(unsigned ilss) const { return segs_.at(ilss).nnode; }
  // AddMat_Dia(ilss, crs) (linearsystem.cpp:419-448): the pattern of segment ilss's matrix
  bool AddMat_Dia(unsigned ilss, const Com::CIndexedArray& crs) { return AddMat_Dia(ilss, crs.index, crs.array); }
  bool AddMat_Dia(unsigned ilss, const std::vector<unsigned>& index, const std::vector<unsigned>& array) {
    if (ilss >= segs_.size()) return false;
    structure_changed();
    return segs_[ilss].mat->AddPattern(index, array);
  }
  // AddMat_NonDia(ilss, jlss, crs) (linearsystem.cpp:450-520): the block
  // coupling segment ilss's rows to segment jlss's columns
  bool AddMat_NonDia(unsigned ilss, unsigned jlss, const Com::CIndexedArray& crs) {
    return AddMat_NonDia(ilss, jlss, crs.index, crs.array);
  }
  bool AddMat_NonDia(unsigned ilss, unsigned jlss, const std::vector<unsigned>& index,
                     const std::vector<unsigned>& array) {
    if (ilss >= segs_.size() || jlss >= segs_.size() || ilss == jlss) return false;
    if (segs_[ilss].dofptr || segs_[jlss].dofptr)
      throw Error("CLinearSystem::AddMat_NonDia: blocks of per-block lengths between segments are not provided",
                  DFM_NOT_IMPLEMENTED);
    auto& m = segs_[ilss].nd[jlss];
    if (!m) m.reset(new CMat_BlkCrs(segs_[ilss].nnode, segs_[ilss].len, segs_[jlss].nnode, segs_[jlss].len));
    structure_changed();
    return m->AddPattern(index, array);
  }
  bool IsMatrix(unsigned ilss, unsigned jlss) const {
    return ilss == jlss ? ilss < segs_.size() : (bool)segs_.at(ilss).nd.at(jlss);
  }
  // InitializeMarge (linearsystem.cpp:74-88): zero matrices, residual, update
  virtual void InitializeMarge() {
    for (auto& g : segs_) {
      g.mat->SetZero();
      for (auto& m : g.nd)
        if (m) m->SetZero();
    }
    if (res_.whole) res_.whole->SetVectorZero();
    if (upd_.whole) upd_.whole->SetVectorZero();
  }
  // FinalizeMarge (linearsystem.cpp:92-128): Dirichlet conditions on the
  // diagonal matrices, then the rows and the columns of the off-diagonal
  // blocks, the residual zeroed at fixed dofs; returns ||residual||
  virtual double FinalizeMarge() {
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
  void ClearFixedBoundaryCondition() {
    for (auto& g : segs_) g.bc->SetAllFlagZero();
  }
  CMatDia_BlkCrs& GetMatrix(unsigned ilss) { return *segs_.at(ilss).mat; }
  const CMatDia_BlkCrs& GetMatrix(unsigned ilss) const { return *segs_.at(ilss).mat; }
  CMat_BlkCrs& GetMatrix(unsigned ilss, unsigned jlss) { return *segs_.at(ilss).nd.at(jlss); }
  const CMat_BlkCrs& GetMatrix(unsigned ilss, unsigned jlss) const { return *segs_.at(ilss).nd.at(jlss); }
  // GetVector(iv, ilss): -1 residual, -2 update, 0.. scratch
  CVector_Blk& GetVector(int iv, unsigned ilss) { return *vecs(iv).seg.at(ilss); }
  const CVector_Blk& GetVector(int iv, unsigned ilss) const {
    return *const_cast<CLinearSystem*>(this)->vecs(iv).seg.at(ilss);
  }
  CVector_Blk& GetResidual(unsigned ilss = 0) { return *res_.seg.at(ilss); }
  CVector_Blk& GetUpdate(unsigned ilss = 0) { return *upd_.seg.at(ilss); }
  CBCFlag& GetBCFlag(unsigned ilss) { return *segs_.at(ilss).bc; }

  // ---- ILinearSystem_Sol (linearsystem.cpp:208-415) ----
  unsigned GetTmpVectorArySize() const override { return (unsigned)tmp_.size(); }
  bool ReSizeTmpVecSolver(unsigned n) override {
    while (tmp_.size() < n) tmp_.push_back(make_vec());
    tmp_.resize(n);
    return true;
  }
  // the segments' dot products added in order; 1.0 (true) with no segments, as the reference
  double DOT(int iv1, int iv2) override {
    if (segs_.empty()) return 1.0;
    double s = 0.0;
    for (unsigned i = 0; i < GetNLinSysSeg(); i++) s += *vecs(iv1).seg[i] * *vecs(iv2).seg[i];
    return s;
  }
  bool COPY(int iv1, int iv2) override {
    if (iv1 != iv2 && !segs_.empty()) *vecs(iv2).whole = *vecs(iv1).whole;
    return true;
  }
  bool SCAL(double alpha, int iv) override {
    if (!segs_.empty()) *vecs(iv).whole *= alpha;
    return true;
  }
  bool AXPY(double alpha, int iv1, int iv2) override {
    if (alpha != 0.0 && !segs_.empty()) vecs(iv2).whole->AXPY(alpha, *vecs(iv1).whole);
    return true;
  }
  // MATVEC (linearsystem.cpp:250-296)
  bool MATVEC(double alpha, int iv1, double beta, int iv2) override {
    if (segs_.empty()) return true;
    if (alpha == 0.0) return SCAL(beta, iv2);  // :270-273
    if (GetNLinSysSeg() == 1) return segs_[0].mat->MatVec(alpha, *vecs(iv1).seg[0], beta, *vecs(iv2).seg[0]);
    const double* x = vecs(iv1).whole->device_ptr();
    return check(dfm_mseg_matvec(mseg().get(), alpha, x, beta, vecs(iv2).whole->device_ptr()),
                 "CLinearSystem::MATVEC") == DFM_OK;
  }

  // ---- beyond the reference ----
  // the whole vector of an id (all segments, one after the other); replaced,
  // not kept, when a segment is added
  CVector_Blk& vec(int iv) { return *vecs(iv).whole; }
  // the C-ABI multi-segment handle over the segments' blocks, their device
  // values current; renewed when the structure changes
  std::shared_ptr<dfm_mseg> mseg() const {
    if (!ms_) {
      const unsigned n = GetNLinSysSeg();
      std::vector<unsigned> nb(n), len(n);
      for (unsigned i = 0; i < n; i++) {
        nb[i] = segs_[i].nnode;
        len[i] = segs_[i].len;
      }
      dfm_mseg* h = nullptr;
      check(dfm_mseg_create(n, nb.data(), len.data(), &h), "CLinearSystem segments");
      ms_.reset(h, detail::MsegDeleter());
    }
    for (unsigned i = 0; i < GetNLinSysSeg(); i++) {
      check(dfm_mseg_set_dia(ms_.get(), i, segs_[i].mat->handle()), "CLinearSystem segments");
      for (unsigned j = 0; j < GetNLinSysSeg(); j++)
        if (segs_[i].nd[j]) check(dfm_mseg_set_nondia(ms_.get(), i, j, segs_[i].nd[j]->handle()), "CLinearSystem segments");
    }
    return ms_;
  }
  // bumped whenever a segment, a pattern or a block is added: a
  // preconditioner set up before is then out of date
  unsigned long long structure() const { return gen_; }

  // public data of the reference (linearsystem.h:167-170), kept in step with
  // the segments; the objects are owned by the system
  std::vector<std::vector<CMat_BlkCrs*>> m_Matrix_NonDia;
  std::vector<CMatDia_BlkCrs*> m_Matrix_Dia;
  std::vector<CVector_Blk*> m_Residual, m_Update;

 private:
  struct Vec {
    std::unique_ptr<CVector_Blk> whole;
    std::vector<std::unique_ptr<CVector_Blk>> seg;  // views into whole's store
  };
  Vec make_vec() {
    Vec v;
    v.whole.reset(new CVector_Blk((unsigned)total_, 1));
    for (auto& g : segs_) v.seg.emplace_back(seg_view(g, v.whole->store()));
    return v;
  }
  static CVector_Blk* seg_view(const Seg& g, const std::shared_ptr<MatVec::detail::VecStore>& st) {
    return g.dofptr ? new CVector_Blk(g.nnode, g.dofptr, st, g.off) : new CVector_Blk(g.nnode, g.len, st, g.off);
  }
  // a vector id after a segment was added: a longer store, the old values
  // copied in, the old segment views re-pointed, a view for the new segment
  void grow(Vec& v, size_t total_old) {
    Vec nv;
    nv.whole.reset(new CVector_Blk((unsigned)total_, 1));
    if (v.whole && total_old)
      check(dfm_memcpy_dtod(nv.whole->device_ptr(), v.whole->device_ptr(), sizeof(double) * total_old),
            "CLinearSystem::AddLinSysSeg");
    for (size_t i = 0; i < segs_.size(); i++) {
      if (i < v.seg.size()) {
        v.seg[i]->rebind(nv.whole->store(), segs_[i].off);
        nv.seg.push_back(std::move(v.seg[i]));
      } else {
        nv.seg.emplace_back(seg_view(segs_[i], nv.whole->store()));
      }
    }
    v = std::move(nv);
  }
  Vec& vecs(int iv) {
    if (iv == -1) return res_;
    if (iv == -2) return upd_;
    return tmp_.at((size_t)iv);
  }
  void structure_changed() {
    ++gen_;
    ms_.reset();
    const size_t n = segs_.size();
    m_Matrix_Dia.resize(n);
    m_Residual.resize(n);
    m_Update.resize(n);
    m_Matrix_NonDia.assign(n, std::vector<CMat_BlkCrs*>(n, nullptr));
    for (size_t i = 0; i < n; i++) {
      m_Matrix_Dia[i] = segs_[i].mat.get();
      m_Residual[i] = res_.seg.size() > i ? res_.seg[i].get() : nullptr;
      m_Update[i] = upd_.seg.size() > i ? upd_.seg[i].get() : nullptr;
      for (size_t j = 0; j < n; j++) m_Matrix_NonDia[i][j] = segs_[i].nd[j].get();
    }
  }
  std::vector<Seg> segs_;
  size_t total_ = 0;
  Vec res_, upd_;
  std::vector<Vec> tmp_;
  mutable std::shared_ptr<dfm_mseg> ms_;
  unsigned long long gen_ = 0;
};

// The preconditioner interface of the LsSol layer (preconditioner.h:46-52).
class CPreconditioner {
 public:
  virtual ~CPreconditioner() {}
  virtual void SetLinearSystem(const CLinearSystem& ls) = 0;
  virtual bool SetValue(const CLinearSystem& ls) = 0;
  virtual bool SolvePrecond(CLinearSystem& ls, unsigned iv) = 0;
};

class CPreconditioner_ILU : public CPreconditioner {
 public:
  CPreconditioner_ILU() = default;
  CPreconditioner_ILU(const CLinearSystem& ls, unsigned nlev = 0) {
    SetFillInLevel((int)nlev);
    SetLinearSystem(ls);
  }
  ~CPreconditioner_ILU() override { release(); }
  CPreconditioner_ILU(const CPreconditioner_ILU&) = delete;
  CPreconditioner_ILU& operator=(const CPreconditioner_ILU&) = delete;
  // Clear (preconditioner.cpp:37-54): the factors, the fill-level requests
  // and the ordering
  void Clear() {
    release();
    lev_inputs_.clear();
    order_.clear();
    ordering_ = DFM_ORDER_NATURAL;
  }
  // SetFillInLevel(lev, ilss) (preconditioner.h:77): a request kept until
  // Clear(); at SetLinearSystem the requests, in order, set the fill level of
  // every block (ilss = -1) or of segment ilss's block row and column
  // (preconditioner.cpp:108-122); no request means level 0.
  void SetFillInLevel(int lev, int ilss0 = -1) { lev_inputs_.push_back({lev, ilss0}); }
  // the level of each block (i, j), row-major nseg x nseg, the requests resolve to
  std::vector<int> FillLevels(unsigned nseg) const {
    std::vector<int> alev((size_t)nseg * nseg, 0);
    for (const auto& in : lev_inputs_) {
      if (in.second == -1) {
        std::fill(alev.begin(), alev.end(), in.first);
        continue;
      }
      if (in.second < 0 || (unsigned)in.second >= nseg)
        throw Error("SetFillInLevel: segment index beyond the linear system", DFM_BAD_ARG);
      for (unsigned i = 0; i < nseg; i++) {
        alev[(size_t)i * nseg + in.second] = in.first;
        alev[(size_t)in.second * nseg + i] = in.first;
      }
    }
    return alev;
  }
  // SetFillBlk(aBlk) (preconditioner.h:80): block rows (and columns) of the
  // first segment's factor that take fill of any level (AddFracPtn(lev, aBlk)),
  // with one segment or several
  void SetFillBlk(const std::vector<unsigned>& aBlk) { fill_blk_ = aBlk; }
  // SetOrdering(order[new] = old) (preconditioner.h:99-103); empty = natural.
  // As the reference, the ordered factor is ILU(0) whatever the fill level
  // (preconditioner.cpp:60-87), and a single segment is required.
  void SetOrdering(const std::vector<int>& aind) {
    order_ = aind;
    ordering_ = aind.empty() ? DFM_ORDER_NATURAL : DFM_ORDER_GIVEN;
  }
  // device-computed orderings (DFM_ORDER_MULTICOLOUR, DFM_ORDER_BLOCK_MULTICOLOUR;
  // one segment), which keep the fill level
  void SetDeviceOrdering(int ordering, unsigned block_nodes = 0) {
    order_.clear();
    ordering_ = ordering;
    block_nodes_ = block_nodes;
  }
  // symbolic factorisation (preconditioner.cpp:57-211); several segments:
  // ILU(lev) across them in natural order
  void SetLinearSystem(const CLinearSystem& ls) override {
    release();
    const std::vector<int> alev = FillLevels(ls.GetNLinSysSeg());
    if (ls.GetNLinSysSeg() == 1 && ls.GetMatrix(0).flex()) {
      // per-block lengths: natural-order ILU(lev) (matdiafrac_blkcrs.cpp:227-287)
      if (ordering_ != DFM_ORDER_NATURAL || !fill_blk_.empty())
        throw Error("CPreconditioner_ILU: per-block lengths take neither orderings nor fill blocks",
                    DFM_NOT_IMPLEMENTED);
      check(dfm_filu_create(ls.GetMatrix(0).fhandle(), alev[0], &fh_), "CPreconditioner_ILU::SetLinearSystem");
    } else if (ls.GetNLinSysSeg() == 1) {
      const dfm_bcrs* A = ls.GetMatrix(0).handle();
      if (!fill_blk_.empty() && ordering_ == DFM_ORDER_NATURAL) {
        check(dfm_ilu_create_fill(A, alev[0], fill_blk_.data(), (unsigned)fill_blk_.size(), &h_),
              "CPreconditioner_ILU::SetLinearSystem");
      } else {
        const int lev = ordering_ == DFM_ORDER_GIVEN ? 0 : alev[0];
        check(dfm_ilu_create(A, lev, ordering_, order_.empty() ? nullptr : order_.data(), block_nodes_, &h_),
              "CPreconditioner_ILU::SetLinearSystem");
      }
    } else {
      if (ordering_ != DFM_ORDER_NATURAL)
        throw Error("CPreconditioner_ILU: orderings need a single segment", DFM_NOT_IMPLEMENTED);
      mseg_ = ls.mseg();
      check(dfm_mseg_ilu_create_levels(mseg_.get(), alev.data(), fill_blk_.empty() ? nullptr : fill_blk_.data(),
                                       (unsigned)fill_blk_.size(), &ms_),
            "CPreconditioner_ILU::SetLinearSystem");
    }
    ls_ = &ls;
    gen_ = ls.structure();
  }
  // SetValue (preconditioner.cpp:217-297): false when the factorisation met
  // more than 10 singular pivots (the factor then as the reference leaves it)
  bool SetValue(const CLinearSystem& ls) override {
    bound(ls, "SetValue");
    int nsing = 0;
    if (ms_) {
      ls.mseg();  // the blocks' device values current
      return check(dfm_mseg_ilu_set_value(ms_, &nsing), "CPreconditioner_ILU::SetValue") == DFM_OK;
    }
    if (fh_)
      return check(dfm_filu_set_value(fh_, ls.GetMatrix(0).fhandle(), &nsing), "CPreconditioner_ILU::SetValue") ==
             DFM_OK;
    return check(dfm_ilu_set_value(h_, ls.GetMatrix(0).handle(), &nsing), "CPreconditioner_ILU::SetValue") == DFM_OK;
  }
  // SolvePrecond (preconditioner.cpp:301-335)
  bool SolvePrecond(CLinearSystem& ls, unsigned iv) override {
    bound(ls, "SolvePrecond");
    if (ms_) return check(dfm_mseg_ilu_precond(ms_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
    if (fh_) return check(dfm_filu_solve_precond(fh_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
    return check(dfm_ilu_solve_precond(h_, ls.vec((int)iv).device_ptr()), "SolvePrecond") == DFM_OK;
  }
  dfm_ilu* handle() const { return h_; }
  dfm_msilu* mseg_handle() const { return ms_; }
  dfm_filu* flex_handle() const { return fh_; }
  // set up on this system, whose structure has not changed since
  bool bound_to(const CLinearSystem& ls) const {
    return ls_ == &ls && gen_ == ls.structure() && (h_ || ms_ || fh_);
  }

 private:
  void bound(const CLinearSystem& ls, const char* what) const {
    if (!(h_ || ms_ || fh_)) throw Error(what, DFM_BAD_ARG);
    if (!bound_to(ls))
      throw Error("CPreconditioner_ILU: the linear system differs from, or changed since, SetLinearSystem",
                  DFM_BAD_ARG);
    (void)what;
  }
  void release() {
    if (h_) dfm_ilu_destroy(h_);
    if (ms_) dfm_mseg_ilu_destroy(ms_);
    if (fh_) dfm_filu_destroy(fh_);
    h_ = nullptr;
    ms_ = nullptr;
    fh_ = nullptr;
    mseg_.reset();
    ls_ = nullptr;
  }
  std::vector<std::pair<int, int>> lev_inputs_;  // SetFillInLevel requests (lev, ilss)
  int ordering_ = DFM_ORDER_NATURAL;
  unsigned block_nodes_ = 0;
  std::vector<int> order_;
  std::vector<unsigned> fill_blk_;
  dfm_ilu* h_ = nullptr;
  dfm_msilu* ms_ = nullptr;
  dfm_filu* fh_ = nullptr;  // a segment of per-block lengths
  std::shared_ptr<dfm_mseg> mseg_;  // keeps the handle ms_ was built on alive
  const CLinearSystem* ls_ = nullptr;
  unsigned long long gen_ = 0;
};

// preconditioner.h:115-145
class CLinearSystemPreconditioner : public ILinearSystemPreconditioner_Sol {
 public:
  CLinearSystemPreconditioner(CLinearSystem& ls, CPreconditioner& prec) : ls_(ls), prec_(prec) {}
  unsigned GetTmpVectorArySize() const override { return ls_.GetTmpVectorArySize(); }
  bool ReSizeTmpVecSolver(unsigned size_new) override { return ls_.ReSizeTmpVecSolver(size_new); }
  double DOT(int iv1, int iv2) override { return ls_.DOT(iv1, iv2); }
  bool COPY(int iv1, int iv2) override { return ls_.COPY(iv1, iv2); }
  bool SCAL(double alpha, int iv1) override { return ls_.SCAL(alpha, iv1); }
  bool AXPY(double alpha, int iv1, int iv2) override { return ls_.AXPY(alpha, iv1, iv2); }
  bool MATVEC(double alpha, int iv1, double beta, int iv2) override { return ls_.MATVEC(alpha, iv1, beta, iv2); }
  bool SolvePrecond(int iv) override { return prec_.SolvePrecond(ls_, (unsigned)iv); }
  CLinearSystem& linear_system() { return ls_; }
  CPreconditioner& preconditioner() { return prec_; }

 private:
  CLinearSystem& ls_;
  CPreconditioner& prec_;
};

namespace detail {
// The device-resident solvers, when lsp is a CLinearSystemPreconditioner over
// a CPreconditioner_ILU set up on its system; NULL otherwise.
inline CPreconditioner_ILU* device_ilu(ILinearSystemPreconditioner_Sol& lsp, CLinearSystem*& ls) {
  auto* p = dynamic_cast<CLinearSystemPreconditioner*>(&lsp);
  if (!p) return nullptr;
  auto* ilu = dynamic_cast<CPreconditioner_ILU*>(&p->preconditioner());
  if (!ilu || !ilu->bound_to(p->linear_system())) return nullptr;
  ls = &p->linear_system();
  return ilu;
}

inline bool device_solve(bool bicg, double& conv_ratio, unsigned& iteration, CLinearSystem& ls,
                         CPreconditioner_ILU& prec, int precision) {
  int rc;
  if (prec.flex_handle()) {
    if (precision != DFM_PREC_FP64) throw Error("per-block lengths: FP64 solvers only", DFM_NOT_IMPLEMENTED);
    dfm_fbcrs* A = ls.GetMatrix(0).fhandle();
    double* r = ls.GetResidual(0).device_ptr();
    rc = dfm_fsolve(bicg ? 1 : 0, A, prec.flex_handle(), r, ls.GetUpdate(0).device_ptr(), &conv_ratio, &iteration);
  } else if (prec.mseg_handle()) {
    if (precision != DFM_PREC_FP64) throw Error("several segments: FP64 solvers only", DFM_NOT_IMPLEMENTED);
    auto S = ls.mseg();
    double* r = ls.vec(-1).device_ptr();
    rc = (bicg ? dfm_mseg_solve_pbicgstab : dfm_mseg_solve_pcg)(S.get(), prec.mseg_handle(), r,
                                                                ls.vec(-2).device_ptr(), &conv_ratio, &iteration);
  } else {
    dfm_solve_opts o{};
    o.precision = precision;
    dfm_bcrs* A = ls.GetMatrix(0).handle();
    double* r = ls.GetResidual(0).device_ptr();
    rc = (bicg ? dfm_solve_pbicgstab : dfm_solve_pcg)(A, prec.handle(), r, ls.GetUpdate(0).device_ptr(),
                                                      &conv_ratio, &iteration, &o);
  }
  // Solve_PCG is false at max_iter; Solve_PBiCGSTAB is true either way
  // (solver_ls_iter.cpp:166-171, 407)
  return check(rc, bicg ? "Solve_PBiCGSTAB" : "Solve_PCG") == DFM_OK || bicg;
}

// Solve_PCG over the interface (solver_ls_iter.cpp:100-172): ids -2 x,
// -1 r, 0 p, 1 z
inline bool pcg_loop(double& conv_ratio, unsigned& iteration, ILinearSystemPreconditioner_Sol& ls) {
  const double tol = conv_ratio;
  const unsigned mx_iter = iteration;
  if (ls.GetTmpVectorArySize() < 2) ls.ReSizeTmpVecSolver(2);
  const int ix = -2, ir = -1, ip = 0, iz = 1;
  ls.SCAL(0.0, ix);
  const double rr0 = ls.DOT(ir, ir);
  if (rr0 < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  ls.COPY(ir, iz);
  ls.SolvePrecond(iz);
  ls.COPY(iz, ip);
  double rz = ls.DOT(ir, iz);
  for (unsigned it = 0; it < mx_iter; it++) {
    ls.MATVEC(1.0, ip, 0.0, iz);
    const double alpha = rz / ls.DOT(ip, iz);
    ls.AXPY(-alpha, iz, ir);
    ls.AXPY(alpha, ip, ix);
    const double rr = ls.DOT(ir, ir);
    if (rr * inv_rr0 < tol * tol) {
      conv_ratio = std::sqrt(rr * inv_rr0);
      iteration = it;
      return true;
    }
    ls.COPY(ir, iz);
    ls.SolvePrecond(iz);
    const double rz_new = ls.DOT(ir, iz);
    const double beta = rz_new / rz;
    rz = rz_new;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, iz, ip);
  }
  conv_ratio = std::sqrt(ls.DOT(ir, ir) * inv_rr0);
  return false;
}

// Solve_BiCGSTAB / Solve_PBiCGSTAB over the interface (solver_ls_iter.cpp:
// 177-286, 291-408).  Unpreconditioned: ids -2 x, -1 r, 0 s, 1 As, 2 p,
// 3 Ap, 4 r2, x zeroed before the 1e-60 test; preconditioned: 0 s, 1 Ms,
// 2 AMs, 3 p, 4 Mp, 5 AMp, 6 r2, x zeroed after it.  True in both cases;
// at max_iter iteration = max_iter and conv_ratio unchanged.
inline void apply_precond(ILinearSystemPreconditioner_Sol& ls, int iv) { ls.SolvePrecond(iv); }
inline void apply_precond(ILinearSystem_Sol&, int) {}

template <class LS>
inline bool bicgstab_loop(double& conv_ratio, unsigned& num_iter, LS& ls, bool precond) {
  const double tol = conv_ratio;
  const unsigned max_iter = num_iter;
  const unsigned ntmp = precond ? 7u : 5u;
  if (ls.GetTmpVectorArySize() < ntmp) ls.ReSizeTmpVecSolver(ntmp);
  const int ix = -2, ir = -1;
  const int is = 0, iMs = precond ? 1 : 0, iAMs = precond ? 2 : 1;
  const int ip = precond ? 3 : 2, iMp = precond ? 4 : 2, iAMp = precond ? 5 : 3, ir2 = precond ? 6 : 4;
  if (!precond) ls.SCAL(0.0, ix);
  const double rr0 = ls.DOT(ir, ir);
  if (rr0 < 1.0e-60) {
    conv_ratio = std::sqrt(rr0);
    num_iter = 0;
    return true;
  }
  const double inv_rr0 = 1.0 / rr0;
  if (precond) ls.SCAL(0.0, ix);
  ls.COPY(ir, ir2);
  ls.COPY(ir, ip);
  double r_r2 = precond ? 0.0 : ls.DOT(ir, ir2);
  num_iter = max_iter;
  for (unsigned it = 1; it < max_iter; it++) {
    if (precond) {
      ls.COPY(ip, iMp);
      apply_precond(ls, iMp);
      r_r2 = ls.DOT(ir, ir2);
    }
    ls.MATVEC(1.0, iMp, 0.0, iAMp);
    const double alpha = r_r2 / ls.DOT(iAMp, ir2);
    ls.COPY(ir, is);
    ls.AXPY(-alpha, iAMp, is);
    if (precond) {
      ls.COPY(is, iMs);
      apply_precond(ls, iMs);
    }
    ls.MATVEC(1.0, iMs, 0.0, iAMs);
    const double den = ls.DOT(iAMs, iAMs);
    const double num = precond ? ls.DOT(is, iAMs) : ls.DOT(iAMs, is);
    const double omega = num / den;
    ls.AXPY(alpha, iMp, ix);
    ls.AXPY(omega, iMs, ix);
    ls.COPY(is, ir);
    ls.AXPY(-omega, iAMs, ir);
    const double sq = ls.DOT(ir, ir) * inv_rr0;
    if (sq < tol * tol) {
      conv_ratio = std::sqrt(sq);
      num_iter = it;
      return true;
    }
    const double tmp1 = ls.DOT(ir, ir2);
    const double beta = tmp1 * alpha / (r_r2 * omega);
    if (!precond) r_r2 = tmp1;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
    ls.AXPY(-beta * omega, iAMp, ip);
  }
  return true;
}
}  // namespace detail

// LsSol::Solve_CG (solver_ls_iter.cpp:31-97): true always; num_iter = the
// 0-based index of the converging step, or max_iter (conv_ratio unchanged)
inline bool Solve_CG(double& conv_ratio, unsigned& num_iter, ILinearSystem_Sol& ls) {
  const unsigned max_iter = num_iter;
  const double tol = conv_ratio;
  if (ls.GetTmpVectorArySize() < 2) ls.ReSizeTmpVecSolver(2);
  const int ix = -2, ir = -1, ip = 0, iAp = 1;
  ls.SCAL(0.0, ix);
  double rr = ls.DOT(ir, ir);
  if (rr < 1.0e-60) {
    num_iter = 0;
    conv_ratio = std::sqrt(rr);
    return true;
  }
  const double inv_rr0 = 1.0 / rr;
  ls.COPY(ir, ip);
  for (num_iter = 0; num_iter < max_iter; num_iter++) {
    ls.MATVEC(1.0, ip, 0.0, iAp);
    const double alpha = rr / ls.DOT(ip, iAp);
    ls.AXPY(alpha, ip, ix);
    ls.AXPY(-alpha, iAp, ir);
    const double rr_new = ls.DOT(ir, ir);
    if (rr_new * inv_rr0 < tol * tol) {
      conv_ratio = std::sqrt(rr_new * inv_rr0);
      break;
    }
    const double beta = rr_new / rr;
    rr = rr_new;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
  }
  return true;
}

// LsSol::Solve_PCG (solver_ls_iter.cpp:100-172)
inline bool Solve_PCG(double& conv_ratio, unsigned& iteration, ILinearSystemPreconditioner_Sol& lsp,
                      int precision = DFM_PREC_FP64) {
  CLinearSystem* ls = nullptr;
  if (CPreconditioner_ILU* ilu = detail::device_ilu(lsp, ls))
    return detail::device_solve(false, conv_ratio, iteration, *ls, *ilu, precision);
  if (precision != DFM_PREC_FP64) throw Error("Solve_PCG: mixed precision needs a CPreconditioner_ILU", DFM_BAD_ARG);
  return detail::pcg_loop(conv_ratio, iteration, lsp);
}

// LsSol::Solve_BiCGSTAB (solver_ls_iter.cpp:177-286)
inline bool Solve_BiCGSTAB(double& conv_ratio, unsigned& num_iter, ILinearSystem_Sol& ls) {
  return detail::bicgstab_loop(conv_ratio, num_iter, ls, false);
}

// LsSol::Solve_PBiCGSTAB (solver_ls_iter.cpp:291-408)
inline bool Solve_PBiCGSTAB(double& conv_ratio, unsigned& num_iter, ILinearSystemPreconditioner_Sol& lsp,
                            int precision = DFM_PREC_FP64) {
  CLinearSystem* ls = nullptr;
  if (CPreconditioner_ILU* ilu = detail::device_ilu(lsp, ls))
    return detail::device_solve(true, conv_ratio, num_iter, *ls, *ilu, precision);
  if (precision != DFM_PREC_FP64)
    throw Error("Solve_PBiCGSTAB: mixed precision needs a CPreconditioner_ILU", DFM_BAD_ARG);
  return detail::bicgstab_loop(conv_ratio, num_iter, lsp, true);
}

}  // namespace LsSol
}  // namespace delfem_amd

#ifndef DELFEM_AMD_NO_ALIASES
namespace LsSol = delfem_amd::LsSol;
#endif
