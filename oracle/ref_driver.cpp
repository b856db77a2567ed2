// ref_driver.cpp -- C entry points over the compiled DelFEM reference.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
// matvec/ls sources from the DelFEM tree together with this driver into
// oracle/_ref/libdfmref.so.  The library is used to (1) pin the C restatement
// in dfm_oracle.c, (2) write the golden fixtures in tests/golden, and (3) time
// the reference CPU solver for bench.py's cpu_baseline.  The product never
// links it.
//
// Reference classes driven here (DelFEM include tree):
//   LsSol::CLinearSystem        include/delfem/ls/linearsystem.h:51-174
//   LsSol::CPreconditioner_ILU  include/delfem/ls/preconditioner.h:59-108
//   LsSol::Solve_PCG / Solve_PBiCGSTAB  include/delfem/ls/solver_ls_iter.h:35-45
//   MatVec::CMatDiaFrac_BlkCrs  include/delfem/matvec/matdiafrac_blkcrs.h:40-137
//   MatVec::Sol::Solve_*        include/delfem/matvec/solver_mat_iter.h:41-60
//   MatVec::CZMatDia_BlkCrs, CZMatDiaFrac_BlkCrs, CZVector_Blk, Sol::Solve_PCOCG
//                               include/delfem/matvec/z*.h (complex twin, len 1)
#include <malloc.h>

// The work vectors of the reference's solvers are CVector_Blk / CZVector_Blk,
// whose constructors leave the values uninitialised (vector_blk.h:53-56,
// zvector_blk.h:48-51), and MatVec(1, x, 0, y) multiplies the old y by 0
// (matdia_blkcrs.cpp:939, zmatdia_blkcrs.cpp:767): heap garbage that is Inf
// or NaN turns into NaN, depending on what the process freed before.
// glibc's M_PERTURB = 0xff makes every malloc'd block start as zero bytes (and
// fills freed blocks with 0xff), so every run sees the zero-initialised
// vectors the algorithms assume and is reproducible.  Set when this library
// is loaded.
__attribute__((constructor)) static void ref_zero_new_blocks() { mallopt(M_PERTURB, 0xff); }
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "delfem/indexed_array.h"
#include "delfem/matvec/matdia_blkcrs.h"
#include "delfem/matvec/matdiafrac_blkcrs.h"
#include "delfem/matvec/vector_blk.h"
#include "delfem/matvec/bcflag_blk.h"
#include "delfem/matvec/ordering_blk.h"
#include "delfem/ls/linearsystem.h"
#include "delfem/ls/preconditioner.h"
#include "delfem/ls/solver_ls_iter.h"
#include "delfem/matvec/solver_mat_iter.h"
#include "delfem/matvec/zmatdia_blkcrs.h"
#include "delfem/matvec/zmatdiafrac_blkcrs.h"
#include "delfem/matvec/zvector_blk.h"
#include "delfem/matvec/zsolver_mat_iter.h"
#include "delfem/femeqn/ker_emat_tri.h"
#include "delfem/femeqn/ker_emat_tet.h"

namespace {

struct RefSys {
  LsSol::CLinearSystem ls;
  unsigned nb = 0, b = 0;
};

// Factor object with access to the protected pattern arrays.
class FracPeek : public MatVec::CMatDiaFrac_BlkCrs {
 public:
  FracPeek(int lev, const MatVec::CMatDia_BlkCrs& A) : MatVec::CMatDiaFrac_BlkCrs(lev, A) {}
  FracPeek(int lev, const MatVec::CMatDia_BlkCrs& A, const MatVec::COrdering_Blk& o)
      : MatVec::CMatDiaFrac_BlkCrs(lev, A, o) {}
  const unsigned* RowPtr() const { return m_colInd_Blk; }
  const unsigned* Col() const { return m_rowPtr_Blk; }
  const unsigned* DiaInd() const { return m_DiaInd; }
  const double* Val() const { return m_valCrs_Blk; }
  const double* Dia() const { return m_valDia_Blk; }
};

struct RefFrac {
  FracPeek* f = nullptr;
  MatVec::COrdering_Blk order;
  bool ordered = false;
  unsigned nb = 0, b = 0;
};

// Preconditioner over a (possibly ordered) ILU(k) factor of the matrix layer:
// the ls-layer CPreconditioner_ILU honours an ordering for ILU(0) only
// (src/ls/preconditioner.cpp:60-87), this one uses CMatDiaFrac_BlkCrs(lev,A,order).
class FracPrecond : public LsSol::CPreconditioner {
 public:
  explicit FracPrecond(RefFrac* rf) : rf_(rf), tmp_(rf->nb, rf->b) {}
  void SetLinearSystem(const LsSol::CLinearSystem&) override {}
  bool SetValue(const LsSol::CLinearSystem& ls) override {
    if (rf_->ordered) return rf_->f->SetValue(ls.GetMatrix(0), rf_->order);
    return rf_->f->SetValue(ls.GetMatrix(0));
  }
  bool SolvePrecond(LsSol::CLinearSystem& ls, unsigned int iv) override {
    if (rf_->ordered) {
      rf_->order.OrderingVector_OldToNew(tmp_, ls.GetVector(iv, 0));
      rf_->f->ForwardSubstitution(tmp_);
      rf_->f->BackwardSubstitution(tmp_);
      rf_->order.OrderingVector_NewToOld(ls.GetVector(iv, 0), tmp_);
    } else {
      rf_->f->ForwardSubstitution(ls.GetVector(iv, 0));
      rf_->f->BackwardSubstitution(ls.GetVector(iv, 0));
    }
    return true;
  }

 private:
  RefFrac* rf_;
  MatVec::CVector_Blk tmp_;
};

// Complex factor with access to its pattern arrays.
class ZFracPeek : public MatVec::CZMatDiaFrac_BlkCrs {
 public:
  ZFracPeek(int lev, const MatVec::CZMatDia_BlkCrs& A) : MatVec::CZMatDiaFrac_BlkCrs(lev, A) {}
  const unsigned* RowPtr() const { return m_colInd_Blk; }
  const unsigned* Col() const { return m_rowPtr_Blk; }
  const Com::Complex* Val() const { return m_valCrs_Blk; }
  const Com::Complex* Dia() const { return m_valDia_Blk; }
};
struct RefZSys {
  MatVec::CZMatDia_BlkCrs* A = nullptr;
  unsigned nb = 0;
  std::vector<int> tmp;
};
void zcopy_in(MatVec::CZVector_Blk& v, unsigned n, const double* src) {
  for (unsigned i = 0; i < n; i++) v.SetValue(i, 0, Com::Complex(src[2 * i], src[2 * i + 1]));
}
void zcopy_out(const MatVec::CZVector_Blk& v, unsigned n, double* dst) {
  for (unsigned i = 0; i < n; i++) {
    dst[2 * i] = v.GetValue(i, 0).Real();
    dst[2 * i + 1] = v.GetValue(i, 0).Imag();
  }
}
std::ofstream& null_stream() {
  static std::ofstream s("/dev/null");
  return s;
}

void copy_vec_in(MatVec::CVector_Blk& v, unsigned nb, unsigned b, const double* src) {
  for (unsigned i = 0; i < nb; i++)
    for (unsigned r = 0; r < b; r++) v.SetValue(i, r, src[(size_t)i * b + r]);
}
void copy_vec_out(const MatVec::CVector_Blk& v, unsigned nb, unsigned b, double* dst) {
  for (unsigned i = 0; i < nb; i++)
    for (unsigned r = 0; r < b; r++) dst[(size_t)i * b + r] = v.GetValue(i, r);
}

// A system of one segment with per-block sizes (AddLinSysSeg(nnode, aLen),
// src/ls/linearsystem.cpp:170-201).
struct RefFx {
  LsSol::CLinearSystem ls;
  unsigned nb = 0;
  std::vector<unsigned> dofptr;
};
// Flex factor (or, at level 0 with no decomposition, a value copy of the
// matrix) with access to its pattern and value-offset arrays.
class FxPeek : public MatVec::CMatDiaFrac_BlkCrs {
 public:
  FxPeek(int lev, const MatVec::CMatDia_BlkCrs& A) : MatVec::CMatDiaFrac_BlkCrs(lev, A) {}
  const unsigned* RowPtr() const { return m_colInd_Blk; }
  const unsigned* Col() const { return m_rowPtr_Blk; }
  const unsigned* DiaInd() const { return m_DiaInd; }
  const unsigned* ValPtr() const { return m_ValPtr; }
  const unsigned* DiaValPtr() const { return m_DiaValPtr; }
  const double* Val() const { return m_valCrs_Blk; }
  const double* Dia() const { return m_valDia_Blk; }
};
void fx_vec_in(MatVec::CVector_Blk& v, const RefFx* s, const double* src) {
  for (unsigned i = 0; i < s->nb; i++)
    for (unsigned r = 0; r < s->dofptr[i + 1] - s->dofptr[i]; r++) v.SetValue(i, r, src[s->dofptr[i] + r]);
}
void fx_vec_out(const MatVec::CVector_Blk& v, const RefFx* s, double* dst) {
  for (unsigned i = 0; i < s->nb; i++)
    for (unsigned r = 0; r < s->dofptr[i + 1] - s->dofptr[i]; r++) dst[s->dofptr[i] + r] = v.GetValue(i, r);
}
// pattern, value offsets and values of a flex factor (or copy)
void fx_peek_out(const FxPeek* f, unsigned nb, unsigned* rowptr, unsigned* col, unsigned* valptr,
                 unsigned* diaptr, double* val, double* dia) {
  const unsigned ncrs = f->NCrs();
  if (rowptr) std::memcpy(rowptr, f->RowPtr(), sizeof(unsigned) * (nb + 1));
  if (col && ncrs) std::memcpy(col, f->Col(), sizeof(unsigned) * ncrs);
  if (valptr) std::memcpy(valptr, f->ValPtr(), sizeof(unsigned) * (ncrs + 1));
  if (diaptr) std::memcpy(diaptr, f->DiaValPtr(), sizeof(unsigned) * (nb + 1));
  if (val && f->ValPtr()[ncrs]) std::memcpy(val, f->Val(), sizeof(double) * f->ValPtr()[ncrs]);
  if (dia) std::memcpy(dia, f->Dia(), sizeof(double) * f->DiaValPtr()[nb]);
}

}  // namespace

extern "C" {

// The reference prints progress text to std::cout on the hot path
// (e.g. src/matvec/matdiafrac_blkcrs.cpp:1364); route it to /dev/null.
void ref_quiet(int on) {
  static std::streambuf* saved = nullptr;
  if (on) {
    if (!saved) saved = std::cout.rdbuf(null_stream().rdbuf());
  } else if (saved) {
    std::cout.rdbuf(saved);
    saved = nullptr;
  }
}

// ---- linear system ---------------------------------------------------------
void* ref_sys_create(unsigned nb, unsigned b, const unsigned* index, const unsigned* array) {
  RefSys* s = new RefSys;
  s->nb = nb;
  s->b = b;
  Com::CIndexedArray crs;
  crs.InitializeSize(nb);
  for (unsigned i = 0; i <= nb; i++) crs.index[i] = index[i];
  crs.array.assign(array, array + index[nb]);
  s->ls.AddLinSysSeg(nb, b);
  s->ls.AddMat_Dia(0, crs);
  return s;
}
void ref_sys_free(void* p) { delete static_cast<RefSys*>(p); }
void ref_sys_initialize_marge(void* p) { static_cast<RefSys*>(p)->ls.InitializeMarge(); }
unsigned ref_sys_ncrs(void* p) { return static_cast<RefSys*>(p)->ls.GetMatrix(0).NCrs(); }

void ref_sys_merge(void* p, unsigned nelem, unsigned npe, const unsigned* conn, const double* emat) {
  RefSys* s = static_cast<RefSys*>(p);
  MatVec::CMatDia_BlkCrs& A = s->ls.GetMatrix(0);
  const unsigned bb = s->b * s->b;
  for (unsigned e = 0; e < nelem; e++)
    A.Mearge(npe, conn + (size_t)e * npe, npe, conn + (size_t)e * npe, bb, emat + (size_t)e * npe * npe * bb);
}
void ref_sys_add_residual(void* p, unsigned nelem, unsigned npe, const unsigned* conn, const double* eres) {
  RefSys* s = static_cast<RefSys*>(p);
  MatVec::CVector_Blk& r = s->ls.GetVector(-1, 0);
  for (unsigned e = 0; e < nelem; e++)
    for (unsigned k = 0; k < npe; k++)
      for (unsigned d = 0; d < s->b; d++)
        r.AddValue(conn[(size_t)e * npe + k], d, eres[((size_t)e * npe + k) * s->b + d]);
}
void ref_sys_set_bc(void* p, const int* flags) {
  RefSys* s = static_cast<RefSys*>(p);
  for (unsigned i = 0; i < s->nb; i++)
    for (unsigned d = 0; d < s->b; d++)
      if (flags[(size_t)i * s->b + d]) s->ls.GetBCFlag(0).SetBC(i, d);
}
double ref_sys_finalize(void* p) { return static_cast<RefSys*>(p)->ls.FinalizeMarge(); }

void ref_sys_get_matrix(void* p, unsigned* rowptr, unsigned* col, double* val, double* dia) {
  RefSys* s = static_cast<RefSys*>(p);
  const MatVec::CMatDia_BlkCrs& A = s->ls.GetMatrix(0);
  const unsigned bb = s->b * s->b;
  unsigned n = 0;
  rowptr[0] = 0;
  for (unsigned i = 0; i < s->nb; i++) {
    unsigned np = 0;
    const unsigned* ind = A.GetPtrIndPSuP(i, np);
    unsigned nv = 0;
    const double* v = A.GetPtrValPSuP(i, nv);
    for (unsigned k = 0; k < np; k++) {
      col[n + k] = ind[k];
      std::memcpy(val + (size_t)(n + k) * bb, v + (size_t)k * bb, sizeof(double) * bb);
    }
    n += np;
    rowptr[i + 1] = n;
    std::memcpy(dia + (size_t)i * bb, A.GetPtrValDia(i), sizeof(double) * bb);
  }
}
void ref_sys_set_matrix(void* p, const double* val, const double* dia) {
  RefSys* s = static_cast<RefSys*>(p);
  MatVec::CMatDia_BlkCrs& A = s->ls.GetMatrix(0);
  const unsigned bb = s->b * s->b;
  unsigned n = 0;
  for (unsigned i = 0; i < s->nb; i++) {
    unsigned np = 0;
    double* v = A.GetPtrValPSuP(i, np);
    std::memcpy(v, val + (size_t)n * bb, sizeof(double) * bb * np);
    n += np;
    std::memcpy(A.GetPtrValDia(i), dia + (size_t)i * bb, sizeof(double) * bb);
  }
}
void ref_sys_resize_tmp(void* p, unsigned ntmp) { static_cast<RefSys*>(p)->ls.ReSizeTmpVecSolver(ntmp); }
void ref_sys_set_vector(void* p, int iv, const double* src) {
  RefSys* s = static_cast<RefSys*>(p);
  copy_vec_in(s->ls.GetVector(iv, 0), s->nb, s->b, src);
}
void ref_sys_get_vector(void* p, int iv, double* dst) {
  RefSys* s = static_cast<RefSys*>(p);
  copy_vec_out(s->ls.GetVector(iv, 0), s->nb, s->b, dst);
}
void ref_sys_matvec(void* p, double alpha, int iv1, double beta, int iv2) {
  static_cast<RefSys*>(p)->ls.MATVEC(alpha, iv1, beta, iv2);
}

// ---- ls-layer ILU preconditioner + Krylov ---------------------------------
void* ref_prec_create(void* p, int lev, const int* order) {
  RefSys* s = static_cast<RefSys*>(p);
  LsSol::CPreconditioner_ILU* pr = new LsSol::CPreconditioner_ILU;
  pr->SetFillInLevel(lev);
  if (order) pr->SetOrdering(std::vector<int>(order, order + s->nb));
  pr->SetLinearSystem(s->ls);
  return pr;
}
void ref_prec_free(void* q) { delete static_cast<LsSol::CPreconditioner_ILU*>(q); }
int ref_prec_set_value(void* q, void* p) {
  return static_cast<LsSol::CPreconditioner_ILU*>(q)->SetValue(static_cast<RefSys*>(p)->ls) ? 1 : 0;
}
void ref_prec_solve(void* q, void* p, int iv) {
  static_cast<LsSol::CPreconditioner_ILU*>(q)->SolvePrecond(static_cast<RefSys*>(p)->ls, iv);
}
int ref_solve_pcg(void* p, void* q, double* conv, unsigned* iter) {
  LsSol::CLinearSystemPreconditioner lsp(static_cast<RefSys*>(p)->ls, *static_cast<LsSol::CPreconditioner_ILU*>(q));
  return LsSol::Solve_PCG(*conv, *iter, lsp) ? 1 : 0;
}
// LsSol::Solve_CG (kind 0) / Solve_BiCGSTAB (kind 2) on the system itself
int ref_ls_solve(void* p, int kind, double* conv, unsigned* iter) {
  LsSol::CLinearSystem& ls = static_cast<RefSys*>(p)->ls;
  const bool ok = kind == 0 ? LsSol::Solve_CG(*conv, *iter, ls) : LsSol::Solve_BiCGSTAB(*conv, *iter, ls);
  return ok ? 1 : 0;
}
int ref_solve_pbicgstab(void* p, void* q, double* conv, unsigned* iter) {
  LsSol::CLinearSystemPreconditioner lsp(static_cast<RefSys*>(p)->ls, *static_cast<LsSol::CPreconditioner_ILU*>(q));
  return LsSol::Solve_PBiCGSTAB(*conv, *iter, lsp) ? 1 : 0;
}

// ---- matrix-layer factor (natural or ordered ILU(k)) ----------------------
void* ref_frac_create(void* p, int lev, const int* order) {
  RefSys* s = static_cast<RefSys*>(p);
  RefFrac* rf = new RefFrac;
  rf->nb = s->nb;
  rf->b = s->b;
  if (order) {
    rf->ordered = true;
    rf->order.SetOrdering(std::vector<int>(order, order + s->nb));
    rf->f = new FracPeek(lev, s->ls.GetMatrix(0), rf->order);
  } else {
    rf->f = new FracPeek(lev, s->ls.GetMatrix(0));
  }
  return rf;
}
// CPreconditioner_ILU after SetFillBlk(blk) on one segment (preconditioner.cpp:88-90,
// 149-150, 198-210): ILU(0) pattern, AddFracPtn(lev, blk), MakePatternFinalize
void* ref_frac_create_fill(void* p, int lev, const unsigned* blk, unsigned nblk) {
  RefSys* s = static_cast<RefSys*>(p);
  RefFrac* rf = new RefFrac;
  rf->nb = s->nb;
  rf->b = s->b;
  rf->f = new FracPeek(0, s->ls.GetMatrix(0));
  rf->f->AddFracPtn(lev, std::vector<unsigned>(blk, blk + nblk));
  rf->f->MakePatternFinalize();
  return rf;
}
void ref_frac_free(void* q) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  delete rf->f;
  delete rf;
}
unsigned ref_frac_ncrs(void* q) { return static_cast<RefFrac*>(q)->f->NCrs(); }
void ref_frac_get_pattern(void* q, unsigned* rowptr, unsigned* col, unsigned* diaind) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  std::memcpy(rowptr, rf->f->RowPtr(), sizeof(unsigned) * (rf->nb + 1));
  std::memcpy(col, rf->f->Col(), sizeof(unsigned) * rf->f->NCrs());
  std::memcpy(diaind, rf->f->DiaInd(), sizeof(unsigned) * rf->nb);
}
int ref_frac_set_value(void* q, void* p) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  RefSys* s = static_cast<RefSys*>(p);
  if (rf->ordered) {
    rf->f->SetValue_Initialize(s->ls.GetMatrix(0), rf->order);
  } else {
    rf->f->SetValue_Initialize(s->ls.GetMatrix(0));
  }
  return rf->f->DoILUDecomp() ? 1 : 0;
}
void ref_frac_get_values(void* q, double* val, double* dia) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  const unsigned bb = rf->b * rf->b;
  std::memcpy(val, rf->f->Val(), sizeof(double) * bb * rf->f->NCrs());
  std::memcpy(dia, rf->f->Dia(), sizeof(double) * bb * rf->nb);
}
void ref_frac_forward(void* q, double* vec) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  MatVec::CVector_Blk v(rf->nb, rf->b);
  copy_vec_in(v, rf->nb, rf->b, vec);
  rf->f->ForwardSubstitution(v);
  copy_vec_out(v, rf->nb, rf->b, vec);
}
void ref_frac_backward(void* q, double* vec) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  MatVec::CVector_Blk v(rf->nb, rf->b);
  copy_vec_in(v, rf->nb, rf->b, vec);
  rf->f->BackwardSubstitution(v);
  copy_vec_out(v, rf->nb, rf->b, vec);
}
// Krylov solve preconditioned by the matrix-layer factor (ordered ILU(k)).
// kind 0 = Solve_PCG, 1 = Solve_PBiCGSTAB.  The factor must already hold values.
int ref_frac_solve(void* q, void* p, int kind, double* conv, unsigned* iter) {
  RefFrac* rf = static_cast<RefFrac*>(q);
  RefSys* s = static_cast<RefSys*>(p);
  FracPrecond prec(rf);
  LsSol::CLinearSystemPreconditioner lsp(s->ls, prec);
  bool ok = kind == 0 ? LsSol::Solve_PCG(*conv, *iter, lsp) : LsSol::Solve_PBiCGSTAB(*conv, *iter, lsp);
  return ok ? 1 : 0;
}

// Matrix-layer solvers on the system's matrix (MatVec::Sol, solver_mat_iter.cpp):
// kind 0 Solve_CG, 1 Solve_PCG, 2 Solve_BiCGSTAB, 3 Solve_PBiCGSTAB; q is a
// natural-order factor holding values (kinds 1 and 3), used as the
// CPrecond_Blk with the system matrix as mat_p.  r: in right-hand side,
// out final residual; u: out solution.
int ref_mat_solve(void* p, void* q, int kind, double* r, double* u, double* conv, unsigned* iter) {
  RefSys* s = static_cast<RefSys*>(p);
  const MatVec::CMatDia_BlkCrs& A = s->ls.GetMatrix(0);
  MatVec::CVector_Blk rv(s->nb, s->b), uv(s->nb, s->b);
  copy_vec_in(rv, s->nb, s->b, r);
  bool ok = false;
  if (kind == 0) {
    ok = MatVec::Sol::Solve_CG(*conv, *iter, A, rv, uv);
  } else if (kind == 2) {
    ok = MatVec::Sol::Solve_BiCGSTAB(*conv, *iter, A, rv, uv);
  } else {
    const MatVec::CPrecond_Blk& pc = *static_cast<RefFrac*>(q)->f;
    ok = kind == 1 ? MatVec::Sol::Solve_PCG(*conv, *iter, A, rv, uv, pc, A)
                   : MatVec::Sol::Solve_PBiCGSTAB(*conv, *iter, A, rv, uv, pc, A);
  }
  copy_vec_out(rv, s->nb, s->b, r);
  copy_vec_out(uv, s->nb, s->b, u);
  return ok ? 1 : 0;
}

// ---- complex twin (len 1) ------------------------------------------------------
void* ref_zsys_create(unsigned nb, const unsigned* index, const unsigned* array) {
  RefZSys* z = new RefZSys;
  z->nb = nb;
  z->A = new MatVec::CZMatDia_BlkCrs(nb, 1);
  Com::CIndexedArray crs;
  crs.InitializeSize(nb);
  crs.index.assign(index, index + nb + 1);
  crs.array.assign(array, array + index[nb]);
  z->A->AddPattern(crs);
  z->A->SetZero();
  z->tmp.assign(nb, -1);
  return z;
}
void ref_zsys_free(void* p) {
  RefZSys* z = static_cast<RefZSys*>(p);
  delete z->A;
  delete z;
}
// CZMatDia_BlkCrs::Mearge of each element in order (emat: npe*npe complex per element)
void ref_zsys_merge(void* p, unsigned nelem, unsigned npe, const unsigned* conn, const double* emat) {
  RefZSys* z = static_cast<RefZSys*>(p);
  std::vector<Com::Complex> em(npe * npe);
  for (unsigned e = 0; e < nelem; e++) {
    for (unsigned k = 0; k < npe * npe; k++)
      em[k] = Com::Complex(emat[2 * ((size_t)e * npe * npe + k)], emat[2 * ((size_t)e * npe * npe + k) + 1]);
    z->A->Mearge(npe, conn + (size_t)e * npe, npe, conn + (size_t)e * npe, 1, em.data(), z->tmp.data());
  }
}
void ref_zsys_set_bc(void* p, const int* flags) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CBCFlag bc(z->nb, 1);
  for (unsigned i = 0; i < z->nb; i++)
    if (flags[i]) bc.SetBC(i, 0);
  z->A->SetBoundaryCondition(bc);
}
void ref_zsys_get(void* p, double* val, double* dia) {
  RefZSys* z = static_cast<RefZSys*>(p);
  const unsigned nb = z->nb;
  for (unsigned i = 0; i < nb; i++) {
    unsigned n = 0;
    const Com::Complex* v = z->A->GetPtrValPSuP(i, n);
    unsigned n0 = 0;
    const unsigned* c0 = z->A->GetPtrIndPSuP(0, n0);
    unsigned ni = 0;
    const unsigned* ci = z->A->GetPtrIndPSuP(i, ni);
    const size_t off = (size_t)(ci - c0);
    for (unsigned k = 0; k < n; k++) {
      val[2 * (off + k)] = v[k].Real();
      val[2 * (off + k) + 1] = v[k].Imag();
    }
    const Com::Complex* d = z->A->GetPtrValDia(i);
    dia[2 * i] = d->Real();
    dia[2 * i + 1] = d->Imag();
  }
}
void ref_zsys_matvec(void* p, double alpha, const double* x, double beta, double* y) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk xv(z->nb, 1), yv(z->nb, 1);
  zcopy_in(xv, z->nb, x);
  zcopy_in(yv, z->nb, y);
  z->A->MatVec(alpha, xv, beta, yv);
  zcopy_out(yv, z->nb, y);
}
void ref_zsys_matvec_hermitian(void* p, double alpha, const double* x, double beta, double* y) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk xv(z->nb, 1), yv(z->nb, 1);
  zcopy_in(xv, z->nb, x);
  zcopy_in(yv, z->nb, y);
  z->A->MatVec_Hermitian(alpha, xv, beta, yv);
  zcopy_out(yv, z->nb, y);
}
void* ref_zfrac_create(void* p, int lev) { return new ZFracPeek(lev, *static_cast<RefZSys*>(p)->A); }
void ref_zfrac_free(void* f) { delete static_cast<ZFracPeek*>(f); }
unsigned ref_zfrac_ncrs(void* f) { return static_cast<ZFracPeek*>(f)->NCrs(); }
void ref_zfrac_get_pattern(void* f, unsigned nb, unsigned* rowptr, unsigned* col) {
  ZFracPeek* z = static_cast<ZFracPeek*>(f);
  std::memcpy(rowptr, z->RowPtr(), sizeof(unsigned) * (nb + 1));
  std::memcpy(col, z->Col(), sizeof(unsigned) * z->NCrs());
}
// CZMatDiaFrac_BlkCrs::SetValue: copy of A and DoILUDecomp
int ref_zfrac_set_value(void* f, void* p) {
  return static_cast<ZFracPeek*>(f)->SetValue(*static_cast<RefZSys*>(p)->A) ? 1 : 0;
}
void ref_zfrac_get_values(void* f, unsigned nb, double* val, double* dia) {
  ZFracPeek* z = static_cast<ZFracPeek*>(f);
  for (unsigned k = 0; k < z->NCrs(); k++) {
    val[2 * k] = z->Val()[k].Real();
    val[2 * k + 1] = z->Val()[k].Imag();
  }
  for (unsigned i = 0; i < nb; i++) {
    dia[2 * i] = z->Dia()[i].Real();
    dia[2 * i + 1] = z->Dia()[i].Imag();
  }
}
void ref_zfrac_solve(void* f, unsigned nb, double* vec) {
  MatVec::CZVector_Blk v(nb, 1);
  zcopy_in(v, nb, vec);
  static_cast<ZFracPeek*>(f)->Solve(v);
  zcopy_out(v, nb, vec);
}
// MatVec::Sol::Solve_PCOCG with the factor as preconditioner; r in/out, u out
int ref_zsolve_pcocg(void* p, void* f, double* r, double* u, double* conv, unsigned* iter) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk rv(z->nb, 1), uv(z->nb, 1);
  zcopy_in(rv, z->nb, r);
  const bool ok = MatVec::Sol::Solve_PCOCG(*conv, *iter, *z->A, rv, uv, *static_cast<ZFracPeek*>(f), *z->A);
  zcopy_out(rv, z->nb, r);
  zcopy_out(uv, z->nb, u);
  return ok ? 1 : 0;
}

// ---- multi-segment LsSol::CLinearSystem --------------------------------------
struct RefMS {
  LsSol::CLinearSystem ls;
  std::vector<unsigned> nb, b;
  LsSol::CPreconditioner_ILU* prec = nullptr;
  ~RefMS() { delete prec; }
};
static Com::CIndexedArray make_crs(unsigned nrow, const unsigned* index, const unsigned* array) {
  Com::CIndexedArray crs;
  crs.InitializeSize(nrow);
  for (unsigned i = 0; i <= nrow; i++) crs.index[i] = index[i];
  crs.array.assign(array, array + index[nrow]);
  return crs;
}
void* ref_ms_create(unsigned nseg, const unsigned* nb, const unsigned* b) {
  RefMS* m = new RefMS;
  for (unsigned i = 0; i < nseg; i++) {
    m->nb.push_back(nb[i]);
    m->b.push_back(b[i]);
    m->ls.AddLinSysSeg(nb[i], b[i]);
  }
  return m;
}
void ref_ms_free(void* p) { delete static_cast<RefMS*>(p); }
void ref_ms_add_pattern(void* p, unsigned i, unsigned j, const unsigned* index, const unsigned* array) {
  RefMS* m = static_cast<RefMS*>(p);
  const Com::CIndexedArray crs = make_crs(m->nb[i], index, array);
  if (i == j) m->ls.AddMat_Dia(i, crs);
  else m->ls.AddMat_NonDia(i, j, crs);
}
void ref_ms_initialize(void* p) { static_cast<RefMS*>(p)->ls.InitializeMarge(); }
// Mearge of every element into block (i, j): emat[e][a][c] blocks of b_i x b_j
void ref_ms_merge(void* p, unsigned i, unsigned j, unsigned nelem, unsigned npe, const unsigned* conn,
                  const double* emat) {
  RefMS* m = static_cast<RefMS*>(p);
  const unsigned bs = m->b[i] * m->b[j];
  for (unsigned e = 0; e < nelem; e++) {
    const unsigned* no = conn + (size_t)e * npe;
    const double* em = emat + (size_t)e * npe * npe * bs;
    if (i == j) m->ls.GetMatrix(i).Mearge(npe, no, npe, no, bs, em);
    else m->ls.GetMatrix(i, j).Mearge(npe, no, npe, no, bs, em);
  }
}
void ref_ms_set_bc(void* p, unsigned i, const int* flags) {
  RefMS* m = static_cast<RefMS*>(p);
  for (unsigned k = 0; k < m->nb[i]; k++)
    for (unsigned d = 0; d < m->b[i]; d++)
      if (flags[(size_t)k * m->b[i] + d]) m->ls.GetBCFlag(i).SetBC(k, d);
}
// residual (vector id -1) from / to a concatenated vector
void ref_ms_set_vector(void* p, int iv, const double* src) {
  RefMS* m = static_cast<RefMS*>(p);
  size_t off = 0;
  for (unsigned i = 0; i < m->nb.size(); i++) {
    copy_vec_in(m->ls.GetVector(iv, i), m->nb[i], m->b[i], src + off);
    off += (size_t)m->nb[i] * m->b[i];
  }
}
void ref_ms_get_vector(void* p, int iv, double* dst) {
  RefMS* m = static_cast<RefMS*>(p);
  size_t off = 0;
  for (unsigned i = 0; i < m->nb.size(); i++) {
    copy_vec_out(m->ls.GetVector(iv, i), m->nb[i], m->b[i], dst + off);
    off += (size_t)m->nb[i] * m->b[i];
  }
}
double ref_ms_finalize(void* p) { return static_cast<RefMS*>(p)->ls.FinalizeMarge(); }
// values of block (i, j) in CRS order (dia: off-diagonal part + diagonal blocks)
void ref_ms_get_block(void* p, unsigned i, unsigned j, double* val, double* dia) {
  RefMS* m = static_cast<RefMS*>(p);
  const MatVec::CMat_BlkCrs& A = i == j ? static_cast<const MatVec::CMat_BlkCrs&>(m->ls.GetMatrix(i))
                                        : m->ls.GetMatrix(i, j);
  const unsigned bs = m->b[i] * m->b[j];
  size_t n = 0;
  for (unsigned k = 0; k < m->nb[i]; k++) {
    unsigned np = 0;
    const double* v = A.GetPtrValPSuP(k, np);
    std::memcpy(val + n * bs, v, sizeof(double) * bs * np);
    n += np;
  }
  if (i == j && dia)
    for (unsigned k = 0; k < m->nb[i]; k++)
      std::memcpy(dia + (size_t)k * bs, m->ls.GetMatrix(i).GetPtrValDia(k), sizeof(double) * bs);
}
void ref_ms_matvec(void* p, double alpha, const double* x, double beta, double* y) {
  RefMS* m = static_cast<RefMS*>(p);
  m->ls.ReSizeTmpVecSolver(2);
  ref_ms_set_vector(p, 0, x);
  ref_ms_set_vector(p, 1, y);
  m->ls.MATVEC(alpha, 0, beta, 1);
  ref_ms_get_vector(p, 1, y);
}
// CPreconditioner_ILU(ls, lev) across the segments, SetValue; returns its bool
int ref_ms_ilu(void* p, int lev) {
  RefMS* m = static_cast<RefMS*>(p);
  delete m->prec;
  m->prec = new LsSol::CPreconditioner_ILU;
  m->prec->SetFillInLevel(lev);
  m->prec->SetLinearSystem(m->ls);
  return m->prec->SetValue(m->ls) ? 1 : 0;
}
// The same with SetFillInLevel(lev[t], ilss[t]) for t < nin, in that order,
// and SetFillBlk(fill_blk[0 .. nfill)) when nfill > 0
int ref_ms_ilu_inputs(void* p, int nin, const int* lev, const int* ilss, const unsigned* fill_blk, int nfill) {
  RefMS* m = static_cast<RefMS*>(p);
  delete m->prec;
  m->prec = new LsSol::CPreconditioner_ILU;
  for (int t = 0; t < nin; t++) m->prec->SetFillInLevel(lev[t], ilss[t]);
  if (nfill > 0) m->prec->SetFillBlk(std::vector<unsigned>(fill_blk, fill_blk + nfill));
  m->prec->SetLinearSystem(m->ls);
  return m->prec->SetValue(m->ls) ? 1 : 0;
}
void ref_ms_precond(void* p, double* z) {
  RefMS* m = static_cast<RefMS*>(p);
  m->ls.ReSizeTmpVecSolver(1);
  ref_ms_set_vector(p, 0, z);
  m->prec->SolvePrecond(m->ls, 0);
  ref_ms_get_vector(p, 0, z);
}
// LsSol::Solve_PCG / Solve_PBiCGSTAB: r in = right-hand side, out = residual; x out
int ref_ms_solve(void* p, int bicg, double* r, double* x, double* conv, unsigned* iter) {
  RefMS* m = static_cast<RefMS*>(p);
  ref_ms_set_vector(p, -1, r);
  LsSol::CLinearSystemPreconditioner lsp(m->ls, *m->prec);
  const bool ok = bicg ? LsSol::Solve_PBiCGSTAB(*conv, *iter, lsp) : LsSol::Solve_PCG(*conv, *iter, lsp);
  ref_ms_get_vector(p, -1, r);
  ref_ms_get_vector(p, -2, x);
  return ok ? 1 : 0;
}

// ---- CMat_BlkCrs as CLinearSystem_Save's boundary matrix ---------------------
struct RefBMat {
  MatVec::CMat_BlkCrs* M = nullptr;
  unsigned nb = 0, b = 0;
};
// a CMat_BlkCrs with SetPatternBoundary(A of the system, bc_col, bc_row)
void* ref_bmat_create_boundary(void* sys, const int* flag_col, const int* flag_row) {
  RefSys* s = static_cast<RefSys*>(sys);
  RefBMat* m = new RefBMat;
  m->nb = s->nb;
  m->b = s->b;
  m->M = new MatVec::CMat_BlkCrs(s->nb, s->b, s->nb, s->b);
  MatVec::CBCFlag bc_col(s->nb, s->b), bc_row(s->nb, s->b);
  for (unsigned i = 0; i < s->nb; i++)
    for (unsigned d = 0; d < s->b; d++) {
      if (flag_col[i * s->b + d]) bc_col.SetBC(i, d);
      if (flag_row[i * s->b + d]) bc_row.SetBC(i, d);
    }
  m->M->SetPatternBoundary(s->ls.GetMatrix(0), bc_col, bc_row);
  m->M->SetZero();
  return m;
}
void ref_bmat_free(void* p) {
  RefBMat* m = static_cast<RefBMat*>(p);
  delete m->M;
  delete m;
}
unsigned ref_bmat_ncrs(void* p) { return static_cast<RefBMat*>(p)->M->NCrs(); }
// rowptr: nb+1, col: ncrs, val: ncrs*b*b (any may be NULL)
void ref_bmat_get(void* p, unsigned* rowptr, unsigned* col, double* val) {
  RefBMat* m = static_cast<RefBMat*>(p);
  const unsigned bb = m->b * m->b;
  unsigned n = 0;
  if (rowptr) rowptr[0] = 0;
  for (unsigned i = 0; i < m->nb; i++) {
    unsigned np = 0;
    const unsigned* ind = m->M->GetPtrIndPSuP(i, np);
    const double* v = m->M->GetPtrValPSuP(i, np);
    for (unsigned k = 0; k < np; k++) {
      if (col) col[n + k] = ind[k];
      if (val)
        for (unsigned t = 0; t < bb; t++) val[(size_t)(n + k) * bb + t] = v[k * bb + t];
    }
    n += np;
    if (rowptr) rowptr[i + 1] = n;
  }
}
void ref_bmat_merge(void* p, unsigned nelem, unsigned npe, const unsigned* conn, const double* emat) {
  RefBMat* m = static_cast<RefBMat*>(p);
  const unsigned bb = m->b * m->b;
  for (unsigned e = 0; e < nelem; e++)
    m->M->Mearge(npe, conn + (size_t)e * npe, npe, conn + (size_t)e * npe, bb, emat + (size_t)e * npe * npe * bb);
}
void ref_bmat_bc(void* p, const int* flag, int inverse_col) {
  RefBMat* m = static_cast<RefBMat*>(p);
  MatVec::CBCFlag bc(m->nb, m->b);
  for (unsigned i = 0; i < m->nb; i++)
    for (unsigned d = 0; d < m->b; d++)
      if (flag[i * m->b + d]) bc.SetBC(i, d);
  if (inverse_col) m->M->SetBoundaryConditionInverse_Colum(bc);
  else m->M->SetBoundaryCondition_Row(bc);
}
void ref_bmat_matvec(void* p, double alpha, const double* x, double beta, double* y) {
  RefBMat* m = static_cast<RefBMat*>(p);
  MatVec::CVector_Blk xv(m->nb, m->b), yv(m->nb, m->b);
  for (unsigned i = 0; i < m->nb; i++)
    for (unsigned d = 0; d < m->b; d++) {
      xv.SetValue(i, d, x[i * m->b + d]);
      yv.SetValue(i, d, y[i * m->b + d]);
    }
  m->M->MatVec(alpha, xv, beta, yv, true);
  for (unsigned i = 0; i < m->nb; i++)
    for (unsigned d = 0; d < m->b; d++) y[i * m->b + d] = yv.GetValue(i, d);
}

// MatVec::Sol::Solve_COCG / Solve_BiCGSTAB / Solve_BiCGStabP (the last with the factor).
static void zero_new_blocks() { mallopt(M_PERTURB, 0xff); }
// MatVec::Sol::Solve_CG / Solve_PCG (with the factor f, or f null) for a Hermitian matrix
int ref_zsolve_cg_herm(void* p, void* f, double* r, double* u, double* conv, unsigned* iter) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk rv(z->nb, 1), uv(z->nb, 1);
  zcopy_in(rv, z->nb, r);
  const bool ok = f ? MatVec::Sol::Solve_PCG(*conv, *iter, *z->A, rv, uv, *static_cast<ZFracPeek*>(f), *z->A)
                    : MatVec::Sol::Solve_CG(*conv, *iter, *z->A, rv, uv);
  zcopy_out(rv, z->nb, r);
  zcopy_out(uv, z->nb, u);
  return ok ? 1 : 0;
}
// MatVec::Sol::Solve_CGNR; r in/out, u out
int ref_zsolve_cgnr(void* p, double* r, double* u, double* conv, unsigned* iter) {
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk rv(z->nb, 1), uv(z->nb, 1);
  zcopy_in(rv, z->nb, r);
  const bool ok = MatVec::Sol::Solve_CGNR(*conv, *iter, *z->A, rv, uv);
  zcopy_out(rv, z->nb, r);
  zcopy_out(uv, z->nb, u);
  return ok ? 1 : 0;
}
int ref_zsolve_cocg(void* p, double* r, double* u, double* conv, unsigned* iter) {
  zero_new_blocks();
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk rv(z->nb, 1), uv(z->nb, 1);
  zcopy_in(rv, z->nb, r);
  const bool ok = MatVec::Sol::Solve_COCG(*conv, *iter, *z->A, rv, uv);
  zcopy_out(rv, z->nb, r);
  zcopy_out(uv, z->nb, u);
  return ok ? 1 : 0;
}
int ref_zsolve_bicgstab(void* p, void* f, double* r, double* u, double* conv, unsigned* iter) {
  zero_new_blocks();
  RefZSys* z = static_cast<RefZSys*>(p);
  MatVec::CZVector_Blk rv(z->nb, 1), uv(z->nb, 1);
  zcopy_in(rv, z->nb, r);
  const bool ok = f ? MatVec::Sol::Solve_BiCGStabP(*conv, *iter, *z->A, rv, uv, *static_cast<ZFracPeek*>(f), *z->A)
                    : MatVec::Sol::Solve_BiCGSTAB(*conv, *iter, *z->A, rv, uv);
  zcopy_out(rv, z->nb, r);
  zcopy_out(uv, z->nb, u);
  return ok ? 1 : 0;
}

// ---- element matrices via the reference kernels ----------------------------
// Loop bodies follow src/femeqn/eqn_poisson.cpp:99-117 and
// src/femeqn/eqn_linear_solid3d.cpp:285-320 using the reference's
// TriArea/TriDlDx (src/femeqn/ker_emat_tri.cpp) and TetVolume/TetDlDx
// (include/delfem/femeqn/ker_emat_tet.h).
void ref_emat_poisson2d_p1(unsigned nelem, const double* coord, const unsigned* conn, const double* alpha,
                           double source, const double* value, double* emat, double* eres) {
  for (unsigned e = 0; e < nelem; e++) {
    const unsigned* no = conn + (size_t)e * 3;
    double c[3][2];
    for (int k = 0; k < 3; k++) {
      c[k][0] = coord[no[k] * 2];
      c[k][1] = coord[no[k] * 2 + 1];
    }
    const double area = TriArea(c[0], c[1], c[2]);
    double dldx[3][2], ct[3];
    TriDlDx(dldx, ct, c[0], c[1], c[2]);
    double* em = emat + (size_t)e * 9;
    double* er = eres + (size_t)e * 3;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        em[i * 3 + j] = alpha[e] * area * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1]);
    for (int i = 0; i < 3; i++) er[i] = source * area * 0.33333333333333333;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) er[i] -= em[i * 3 + j] * value[no[j]];
  }
}
void ref_emat_solid3d_p1(unsigned nelem, const double* coord, const unsigned* conn, const double* lambda,
                         const double* myu, double rho, double gx, double gy, double gz, const double* disp,
                         double* emat, double* eres) {
  for (unsigned e = 0; e < nelem; e++) {
    const unsigned* no = conn + (size_t)e * 4;
    double c[4][3], u[4][3];
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) {
        c[k][d] = coord[no[k] * 3 + d];
        u[k][d] = disp[no[k] * 3 + d];
      }
    const double vol = TetVolume(c[0], c[1], c[2], c[3]);
    double dldx[4][3], zt[4];
    TetDlDx(dldx, zt, c[0], c[1], c[2], c[3]);
    const double lam = lambda[e], mu = myu[e];
    double em[4][4][3][3];
    for (int a = 0; a < 4; a++)
      for (int bq = 0; bq < 4; bq++) {
        for (int x = 0; x < 3; x++)
          for (int y = 0; y < 3; y++)
            em[a][bq][x][y] = vol * (lam * dldx[a][x] * dldx[bq][y] + mu * dldx[bq][x] * dldx[a][y]);
        const double d = dldx[a][0] * dldx[bq][0] + dldx[a][1] * dldx[bq][1] + dldx[a][2] * dldx[bq][2];
        em[a][bq][0][0] += vol * mu * d;
        em[a][bq][1][1] += vol * mu * d;
        em[a][bq][2][2] += vol * mu * d;
      }
    double er[4][3];
    for (int a = 0; a < 4; a++) {
      er[a][0] = vol * rho * gx * 0.25;
      er[a][1] = vol * rho * gy * 0.25;
      er[a][2] = vol * rho * gz * 0.25;
      for (int bq = 0; bq < 4; bq++) {
        er[a][0] -= em[a][bq][0][0] * u[bq][0] + em[a][bq][0][1] * u[bq][1] + em[a][bq][0][2] * u[bq][2];
        er[a][1] -= em[a][bq][1][0] * u[bq][0] + em[a][bq][1][1] * u[bq][1] + em[a][bq][1][2] * u[bq][2];
        er[a][2] -= em[a][bq][2][0] * u[bq][0] + em[a][bq][2][1] * u[bq][1] + em[a][bq][2][2] * u[bq][2];
      }
    }
    std::memcpy(emat + (size_t)e * 144, &em[0][0][0][0], sizeof(double) * 144);
    std::memcpy(eres + (size_t)e * 12, &er[0][0], sizeof(double) * 12);
  }
}


// ---- variable block sizes (flex) -------------------------------------------
void* ref_fx_create(unsigned nb, const unsigned* alen, const unsigned* index, const unsigned* array) {
  RefFx* s = new RefFx;
  s->nb = nb;
  s->dofptr.assign(nb + 1, 0);
  for (unsigned i = 0; i < nb; i++) s->dofptr[i + 1] = s->dofptr[i] + alen[i];
  Com::CIndexedArray crs;
  crs.InitializeSize(nb);
  for (unsigned i = 0; i <= nb; i++) crs.index[i] = index[i];
  crs.array.assign(array, array + index[nb]);
  s->ls.AddLinSysSeg(nb, std::vector<unsigned>(alen, alen + nb));
  s->ls.AddMat_Dia(0, crs);
  s->ls.InitializeMarge();
  return s;
}
void ref_fx_free(void* p) { delete static_cast<RefFx*>(p); }
// CMatDia_BlkCrs::Mearge of single blocks (i, j), in order
void ref_fx_merge(void* p, unsigned n, const unsigned* iblk, const unsigned* jblk, const double* emat) {
  RefFx* s = static_cast<RefFx*>(p);
  MatVec::CMatDia_BlkCrs& A = s->ls.GetMatrix(0);
  size_t off = 0;
  for (unsigned e = 0; e < n; e++) {
    const unsigned bs = (s->dofptr[iblk[e] + 1] - s->dofptr[iblk[e]]) * (s->dofptr[jblk[e] + 1] - s->dofptr[jblk[e]]);
    A.Mearge(1, iblk + e, 1, jblk + e, bs, emat + off);
    off += bs;
  }
}
void ref_fx_set_bc(void* p, const int* flags) {
  RefFx* s = static_cast<RefFx*>(p);
  for (unsigned i = 0; i < s->nb; i++)
    for (unsigned d = 0; d < s->dofptr[i + 1] - s->dofptr[i]; d++)
      if (flags[s->dofptr[i] + d]) s->ls.GetBCFlag(0).SetBC(i, d);
}
double ref_fx_finalize(void* p) { return static_cast<RefFx*>(p)->ls.FinalizeMarge(); }
unsigned ref_fx_ncrs(void* p) { return static_cast<RefFx*>(p)->ls.GetMatrix(0).NCrs(); }
// the matrix through a level-0 flex copy (SetValue_Initialize copies the
// values block by block into the same pattern)
void ref_fx_get_matrix(void* p, unsigned* rowptr, unsigned* col, unsigned* valptr, unsigned* diaptr, double* val,
                       double* dia) {
  RefFx* s = static_cast<RefFx*>(p);
  FxPeek f(0, s->ls.GetMatrix(0));
  f.SetValue_Initialize(s->ls.GetMatrix(0));
  fx_peek_out(&f, s->nb, rowptr, col, valptr, diaptr, val, dia);
}
void ref_fx_set_vector(void* p, int iv, const double* src) {
  RefFx* s = static_cast<RefFx*>(p);
  fx_vec_in(s->ls.GetVector(iv, 0), s, src);
}
void ref_fx_get_vector(void* p, int iv, double* dst) {
  RefFx* s = static_cast<RefFx*>(p);
  fx_vec_out(s->ls.GetVector(iv, 0), s, dst);
}
void ref_fx_resize_tmp(void* p, unsigned ntmp) { static_cast<RefFx*>(p)->ls.ReSizeTmpVecSolver(ntmp); }
void ref_fx_matvec(void* p, double alpha, int iv1, double beta, int iv2) {
  static_cast<RefFx*>(p)->ls.MATVEC(alpha, iv1, beta, iv2);
}
// CMatDiaFrac_BlkCrs(lev, A) over the flex matrix
void* ref_fx_frac_create(void* p, int lev) {
  RefFx* s = static_cast<RefFx*>(p);
  return new FxPeek(lev, s->ls.GetMatrix(0));
}
void ref_fx_frac_free(void* q) { delete static_cast<FxPeek*>(q); }
unsigned ref_fx_frac_ncrs(void* q) { return static_cast<FxPeek*>(q)->NCrs(); }
int ref_fx_frac_set_value(void* q, void* p) {
  FxPeek* f = static_cast<FxPeek*>(q);
  f->SetValue_Initialize(static_cast<RefFx*>(p)->ls.GetMatrix(0));
  return f->DoILUDecomp() ? 1 : 0;
}
void ref_fx_frac_get(void* q, void* p, unsigned* rowptr, unsigned* col, unsigned* diaind, unsigned* valptr,
                     unsigned* diaptr, double* val, double* dia) {
  FxPeek* f = static_cast<FxPeek*>(q);
  const unsigned nb = static_cast<RefFx*>(p)->nb;
  fx_peek_out(f, nb, rowptr, col, valptr, diaptr, val, dia);
  if (diaind) std::memcpy(diaind, f->DiaInd(), sizeof(unsigned) * nb);
}
void ref_fx_frac_forward(void* q, void* p, double* vec) {
  RefFx* s = static_cast<RefFx*>(p);
  std::vector<unsigned> alen(s->nb);
  for (unsigned i = 0; i < s->nb; i++) alen[i] = s->dofptr[i + 1] - s->dofptr[i];
  MatVec::CVector_Blk w(s->nb, alen);
  fx_vec_in(w, s, vec);
  static_cast<FxPeek*>(q)->ForwardSubstitution(w);
  fx_vec_out(w, s, vec);
}
void ref_fx_frac_backward(void* q, void* p, double* vec) {
  RefFx* s = static_cast<RefFx*>(p);
  std::vector<unsigned> alen(s->nb);
  for (unsigned i = 0; i < s->nb; i++) alen[i] = s->dofptr[i + 1] - s->dofptr[i];
  MatVec::CVector_Blk w(s->nb, alen);
  fx_vec_in(w, s, vec);
  static_cast<FxPeek*>(q)->BackwardSubstitution(w);
  fx_vec_out(w, s, vec);
}
// LsSol solvers on the flex system: kind 0 Solve_PCG / 1 Solve_PBiCGSTAB with
// CPreconditioner_ILU(ls, lev) (SetValue inside), 2 Solve_CG / 3
// Solve_BiCGSTAB without.  The residual (-1) must hold the right-hand side;
// r out: the final residual, x out: the update (-2).
int ref_fx_solve(void* p, int kind, int lev, double* r, double* x, double* conv, unsigned* iter) {
  RefFx* s = static_cast<RefFx*>(p);
  fx_vec_in(s->ls.GetVector(-1, 0), s, r);
  bool ok;
  if (kind <= 1) {
    LsSol::CPreconditioner_ILU prec(s->ls, lev);
    if (!prec.SetValue(s->ls)) return -1;
    LsSol::CLinearSystemPreconditioner lsp(s->ls, prec);
    ok = kind == 0 ? LsSol::Solve_PCG(*conv, *iter, lsp) : LsSol::Solve_PBiCGSTAB(*conv, *iter, lsp);
  } else {
    ok = kind == 2 ? LsSol::Solve_CG(*conv, *iter, s->ls) : LsSol::Solve_BiCGSTAB(*conv, *iter, s->ls);
  }
  fx_vec_out(s->ls.GetVector(-1, 0), s, r);
  fx_vec_out(s->ls.GetVector(-2, 0), s, x);
  return ok ? 1 : 0;
}
}  // extern "C"
