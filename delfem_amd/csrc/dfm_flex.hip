// dfm_flex.hip -- blocks of different sizes in one matrix ("flex" blocks):
// MatVec::CMatDia_BlkCrs / CMatDiaFrac_BlkCrs constructed with per-block
// lengths aLen (src/matvec/mat_blkcrs.cpp:63-92, matdia_blkcrs.cpp:70-104,
// matdiafrac_blkcrs.cpp:227-287) and the LsSol solvers over a segment of
// them (LsSol::CLinearSystem::AddLinSysSeg(nnode, aLen),
// src/ls/linearsystem.cpp:170-201).  The reference uses them for its rigid
// bodies and constraints (6-dof bodies, constraints of a few dofs); the
// systems are small, so the kernels give one thread to a block row (or to a
// dof row) and run the factorisation and the substitutions level set by
// level set, in FP64 and in the reference's operation order, which makes
// every result bitwise that of the reference's flex branches.
//
// Layout (the reference's): dofptr[i] = first dof of block i; the block
// pattern in CRS without the diagonal; valptr[q] = offset of CRS block q =
// (i, j), len_i x len_j row-major, offsets in CRS order (m_ValPtr);
// diaptr[i] = offset of the len_i x len_i diagonal block (m_DiaValPtr).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dfm_kernels.h"

using namespace dfm;

namespace {

// Longest block handled (the per-thread arrays of the kernels are this long).
constexpr unsigned kFxMaxLen = 32;

#define DFM_FLAUNCH(kernel, n, ...)                                                                          \
  do {                                                                                                       \
    if ((n) > 0) {                                                                                           \
      hipLaunchKernelGGL((kernel), dim3((unsigned)(((n) + kNT - 1) / kNT)), dim3(kNT), 0, ::dfm::stream(), \
                         __VA_ARGS__);                                                                       \
      DFM_HIP(hipGetLastError());                                                                            \
    }                                                                                                        \
  } while (0)

// CMatDia_BlkCrs::MatVec, flex branch (matdia_blkcrs.cpp:931-960), one
// thread per dof row: y = beta y, then alpha * a * x term by term over the
// CRS blocks in order and the diagonal block last.
__global__ void __launch_bounds__(kNT) k_fx_matvec(int ndof, const int* __restrict__ dofblk,
                                                  const int* __restrict__ dofptr, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const int* __restrict__ valptr,
                                                  const double* __restrict__ val, const int* __restrict__ diaptr,
                                                  const double* __restrict__ dia, double alpha,
                                                  const double* __restrict__ x, double beta, double* y) {
  const int d = blockIdx.x * kNT + threadIdx.x;
  if (d >= ndof) return;
  const int i = dofblk[d];
  const int r = d - dofptr[i], li = dofptr[i + 1] - dofptr[i];
  double acc = y[d] * beta;
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
    const int j = col[q];
    if (j == i) continue;
    const int lj = dofptr[j + 1] - dofptr[j];
    const double* a = val + valptr[q] + (size_t)r * lj;
    const double* xj = x + dofptr[j];
    for (int c = 0; c < lj; c++) acc += alpha * a[c] * xj[c];
  }
  const double* a = dia + diaptr[i] + (size_t)r * li;
  const double* xi = x + dofptr[i];
  for (int c = 0; c < li; c++) acc += alpha * a[c] * xi[c];
  y[d] = acc;
}

// CMatDia_BlkCrs::SetBoundaryCondition, flex branch (matdia_blkcrs.cpp:268-306),
// one thread per block row: rows of fixed dofs zeroed with a unit diagonal,
// columns of fixed dofs zeroed in the row's CRS blocks (zeros and ones only,
// so the order of the two passes does not matter).
__global__ void __launch_bounds__(kNT) k_fx_bc(int nb, const int* __restrict__ dofptr, const int* __restrict__ rowptr,
                                              const int* __restrict__ col, const int* __restrict__ valptr,
                                              double* __restrict__ val, const int* __restrict__ diaptr,
                                              double* __restrict__ dia, const int* __restrict__ flag) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  const int li = dofptr[i + 1] - dofptr[i];
  for (int r = 0; r < li; r++) {
    if (flag[dofptr[i] + r] == 0) continue;
    double* d = dia + diaptr[i];
    for (int k = 0; k < li; k++) d[k * li + r] = 0.0;
    for (int c = 0; c < li; c++) d[r * li + c] = 0.0;
    d[r * li + r] = 1.0;
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const int lj = dofptr[col[q] + 1] - dofptr[col[q]];
      for (int c = 0; c < lj; c++) val[valptr[q] + r * lj + c] = 0.0;
    }
  }
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
    const int j = col[q], lj = dofptr[j + 1] - dofptr[j];
    for (int c = 0; c < lj; c++) {
      if (flag[dofptr[j] + c] == 0) continue;
      for (int r = 0; r < li; r++) val[valptr[q] + r * lj + c] = 0.0;
    }
  }
}

// SetValue_Initialize (matdiafrac_blkcrs.h:90-95) for factor rows >= row0:
// CRS blocks from the matrix (fsrc[q] = matrix CRS block, -1 for fill,
// which stays zero) and the diagonal blocks; one thread per factor row.
__global__ void __launch_bounds__(kNT) k_fx_copy_rows(int nb, int row0, const int* __restrict__ dofptr,
                                                     const int* __restrict__ frow, const int* __restrict__ fcol,
                                                     const int* __restrict__ fvalptr,
                                                     const long long* __restrict__ fsrc,
                                                     const int* __restrict__ valptr, const double* __restrict__ val,
                                                     const int* __restrict__ diaptr, const double* __restrict__ dia,
                                                     double* __restrict__ fval, double* __restrict__ fdia) {
  const int i = row0 + blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  const int li = dofptr[i + 1] - dofptr[i];
  for (int q = frow[i]; q < frow[i + 1]; q++) {
    const int lj = dofptr[fcol[q] + 1] - dofptr[fcol[q]];
    const long long s = fsrc[q];
    double* out = fval + fvalptr[q];
    for (int t = 0; t < li * lj; t++) out[t] = s >= 0 ? val[valptr[s] + t] : 0.0;
  }
  for (int t = 0; t < li * li; t++) fdia[diaptr[i] + t] = dia[diaptr[i] + t];
}

// DoILUDecomp, flex branch (matdiafrac_blkcrs.cpp:777-846), one thread per
// row of a level set: A_ij -= A_ik A_kj with k innermost per entry
// (CalcSubMatPr, :62-73), Gauss-Jordan inversion of the diagonal block
// (CalcInvMat, :107-136; a pivot below 1e-30 stops it there and counts the
// row), D_i^-1 applied to the U blocks (CalcMatPr, :45-60; column by column,
// which forms every entry in the same order).  partial: the row the
// reference stops in (11th singular pivot) -- no U scaling.
__global__ void __launch_bounds__(kNT) k_fx_ilu(int nrow, const int* __restrict__ rows, const int* __restrict__ dofptr,
                                               const int* __restrict__ frow, const int* __restrict__ fcol,
                                               const int* __restrict__ fdiaind, const int* __restrict__ fvalptr,
                                               const int* __restrict__ diaptr, double* fval, double* fdia,
                                               int* __restrict__ nsing, int* __restrict__ singrow, int partial) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int i = rows[t];
  const int li = dofptr[i + 1] - dofptr[i];
  const int r0 = frow[i], rd = fdiaind[i], r1 = frow[i + 1];
  double* D = fdia + diaptr[i];
  for (int qk = r0; qk < rd; qk++) {
    const int k = fcol[qk], lk = dofptr[k + 1] - dofptr[k];
    const double* aik = fval + fvalptr[qk];
    int p = qk + 1;
    for (int qj = fdiaind[k]; qj < frow[k + 1]; qj++) {
      const int j = fcol[qj], lj = dofptr[j + 1] - dofptr[j];
      double* aij;
      if (j == i) {
        aij = D;
      } else {
        while (p < r1 && fcol[p] < j) p++;
        if (p >= r1 || fcol[p] != j) continue;
        aij = fval + fvalptr[p];
      }
      const double* akj = fval + fvalptr[qj];
      for (int r = 0; r < li; r++)
        for (int c = 0; c < lj; c++)
          for (int m = 0; m < lk; m++) aij[r * lj + c] -= aik[r * lk + m] * akj[m * lj + c];
    }
  }
  bool singular = false;
  for (int r = 0; r < li; r++) {
    if (fabs(D[r * li + r]) < 1.0e-30) {
      singular = true;
      break;
    }
    double tt = 1.0 / D[r * li + r];
    D[r * li + r] = 1.0;
    for (int k = 0; k < li; k++) D[r * li + k] *= tt;
    for (int j = 0; j < li; j++) {
      if (j == r) continue;
      tt = D[j * li + r];
      D[j * li + r] = 0.0;
      for (int k = 0; k < li; k++) D[j * li + k] -= tt * D[r * li + k];
    }
  }
  if (singular) atomicAdd(nsing, 1);
  singrow[i] = singular ? 1 : 0;
  if (partial && singular) return;
  double tmp[kFxMaxLen];
  for (int q = rd; q < r1; q++) {
    double* u = fval + fvalptr[q];
    const int lj = dofptr[fcol[q] + 1] - dofptr[fcol[q]];
    for (int c = 0; c < lj; c++) {
      for (int r = 0; r < li; r++) {
        double s = 0.0;
        for (int m = 0; m < li; m++) s += D[r * li + m] * u[m * lj + c];
        tmp[r] = s;
      }
      for (int r = 0; r < li; r++) u[r * lj + c] = tmp[r];
    }
  }
}

// ForwardSubstitution, flex branch (matdiafrac_blkcrs.cpp:339-379), one
// thread per row of a lower level set: t = v_i - sum L_ij v_j, v_i = D_i^-1 t.
__global__ void __launch_bounds__(kNT) k_fx_lower(int nrow, const int* __restrict__ rows,
                                                 const int* __restrict__ dofptr, const int* __restrict__ frow,
                                                 const int* __restrict__ fcol, const int* __restrict__ fdiaind,
                                                 const int* __restrict__ fvalptr, const double* __restrict__ fval,
                                                 const int* __restrict__ diaptr, const double* __restrict__ fdia,
                                                 double* vec) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int i = rows[t];
  const int li = dofptr[i + 1] - dofptr[i];
  double tv[kFxMaxLen];
  for (int r = 0; r < li; r++) tv[r] = vec[dofptr[i] + r];
  for (int q = frow[i]; q < fdiaind[i]; q++) {
    const int j = fcol[q], lj = dofptr[j + 1] - dofptr[j];
    const double* a = fval + fvalptr[q];
    const double* vj = vec + dofptr[j];
    for (int r = 0; r < li; r++)
      for (int c = 0; c < lj; c++) tv[r] -= a[r * lj + c] * vj[c];
  }
  const double* d = fdia + diaptr[i];
  for (int r = 0; r < li; r++) {
    double s = 0.0;
    for (int c = 0; c < li; c++) s += d[r * li + c] * tv[c];
    vec[dofptr[i] + r] = s;
  }
}

// BackwardSubstitution, flex branch (matdiafrac_blkcrs.cpp:512-553), one
// thread per row of an upper level set: v_i -= sum U_ij v_j.
__global__ void __launch_bounds__(kNT) k_fx_upper(int nrow, const int* __restrict__ rows,
                                                 const int* __restrict__ dofptr, const int* __restrict__ frow,
                                                 const int* __restrict__ fcol, const int* __restrict__ fdiaind,
                                                 const int* __restrict__ fvalptr, const double* __restrict__ fval,
                                                 double* vec) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int i = rows[t];
  const int li = dofptr[i + 1] - dofptr[i];
  double tv[kFxMaxLen];
  for (int r = 0; r < li; r++) tv[r] = vec[dofptr[i] + r];
  for (int q = fdiaind[i]; q < frow[i + 1]; q++) {
    const int j = fcol[q], lj = dofptr[j + 1] - dofptr[j];
    const double* a = fval + fvalptr[q];
    const double* vj = vec + dofptr[j];
    for (int r = 0; r < li; r++)
      for (int c = 0; c < lj; c++) tv[r] -= a[r * lj + c] * vj[c];
  }
  for (int r = 0; r < li; r++) vec[dofptr[i] + r] = tv[r];
}

// value offsets of a block pattern (m_ValPtr / m_DiaValPtr)
void fx_layout(unsigned nb, const std::vector<unsigned>& dofptr, const std::vector<unsigned>& rowptr,
               const std::vector<unsigned>& col, std::vector<unsigned>& valptr, std::vector<unsigned>& diaptr) {
  valptr.assign(col.size() + 1, 0);
  diaptr.assign(nb + 1, 0);
  unsigned v = 0, d = 0;
  for (unsigned i = 0; i < nb; i++) {
    const unsigned li = dofptr[i + 1] - dofptr[i];
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
      valptr[q] = v;
      v += li * (dofptr[col[q] + 1] - dofptr[col[q]]);
    }
    diaptr[i] = d;
    d += li * li;
  }
  valptr[col.size()] = v;
  diaptr[nb] = d;
}

template <typename T>
int upload_as_int(DevBuf<int>& d, const std::vector<T>& h) {
  std::vector<int> v(h.begin(), h.end());
  if (v.empty()) v.push_back(0);
  return d.upload(v);
}

}  // namespace

struct dfm_fbcrs_s {
  unsigned nb = 0;
  std::vector<unsigned> dofptr;
  std::vector<unsigned> rowptr, col, valptr, diaptr;
  DevBuf<int> d_dofptr, d_dofblk, d_rowptr, d_col, d_valptr, d_diaptr;
  DevBuf<double> d_val, d_dia;
  unsigned long long version = 0, pattern_version = 0;
  unsigned ndof() const { return dofptr[nb]; }
};

struct dfm_filu_s {
  unsigned nb = 0;
  int lev_fill = 0;
  const dfm_fbcrs_s* A = nullptr;
  unsigned long long pattern_version = 0;  // of A when the pattern was made
  std::vector<unsigned> frow, fcol, fdiaind, fvalptr;
  int nlev_lower = 0, nlev_upper = 0;
  std::vector<int> lev_ptr_L, lev_ptr_U;
  DevBuf<int> d_frow, d_fcol, d_fdiaind, d_fvalptr, d_rows_L, d_rows_U, d_nsing, d_singrow, d_fixrow;
  DevBuf<long long> d_fsrc;
  DevBuf<double> d_fval, d_fdia;
  std::vector<double*> work;  // solver vectors, zeroed when made
  bool factored = false;
  ~dfm_filu_s() {
    for (double* p : work) (void)hipFree(p);
  }
};

namespace {

int fx_refresh_layout(dfm_fbcrs_s* A) {
  fx_layout(A->nb, A->dofptr, A->rowptr, A->col, A->valptr, A->diaptr);
  DFM_TRY(upload_as_int(A->d_rowptr, A->rowptr));
  DFM_TRY(upload_as_int(A->d_col, A->col));
  DFM_TRY(upload_as_int(A->d_valptr, A->valptr));
  DFM_TRY(upload_as_int(A->d_diaptr, A->diaptr));
  DFM_TRY(A->d_val.alloc(std::max<size_t>(1, A->valptr.back())));
  DFM_TRY(A->d_dia.alloc(std::max<size_t>(1, A->diaptr.back())));
  DFM_HIP(hipMemsetAsync(A->d_val.p, 0, A->d_val.n * sizeof(double), stream()));
  DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double), stream()));
  A->version = next_version();
  A->pattern_version = next_version();
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

void level_rows(unsigned nb, const std::vector<int>& level, int nlev, std::vector<int>& ptr, std::vector<int>& rows) {
  ptr.assign(nlev + 1, 0);
  for (unsigned i = 0; i < nb; i++) ptr[level[i] + 1]++;
  for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
  rows.assign(nb, 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (unsigned i = 0; i < nb; i++) rows[fill[level[i]]++] = (int)i;
}

int fx_factor_rows(dfm_filu_s* M, int nrow, const int* rows_dev, int partial) {
  const dfm_fbcrs_s* A = M->A;
  DFM_FLAUNCH(k_fx_ilu, nrow, nrow, rows_dev, A->d_dofptr.p, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p,
              M->d_fvalptr.p, A->d_diaptr.p, M->d_fval.p, M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
  return DFM_OK;
}

int fx_sweep(dfm_filu_s* M, double* vec, bool lower, bool upper) {
  const dfm_fbcrs_s* A = M->A;
  if (lower)
    for (int l = 0; l < M->nlev_lower; l++) {
      const int off = M->lev_ptr_L[l], n = M->lev_ptr_L[l + 1] - off;
      DFM_FLAUNCH(k_fx_lower, n, n, M->d_rows_L.p + off, A->d_dofptr.p, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p,
                  M->d_fvalptr.p, M->d_fval.p, A->d_diaptr.p, M->d_fdia.p, vec);
    }
  if (upper)
    for (int l = 0; l < M->nlev_upper; l++) {
      const int off = M->lev_ptr_U[l], n = M->lev_ptr_U[l + 1] - off;
      DFM_FLAUNCH(k_fx_upper, n, n, M->d_rows_U.p + off, A->d_dofptr.p, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p,
                  M->d_fvalptr.p, M->d_fval.p, vec);
    }
  return sync_stream();
}

int fx_matvec(dfm_fbcrs_s* A, double alpha, const double* x, double beta, double* y) {
  const int n = (int)A->ndof();
  DFM_FLAUNCH(k_fx_matvec, n, n, A->d_dofblk.p, A->d_dofptr.p, A->d_rowptr.p, A->d_col.p, A->d_valptr.p, A->d_val.p,
              A->d_diaptr.p, A->d_dia.p, alpha, x, beta, y);
  return sync_stream();
}

// host loop with deferred synchronisation, synchronised once at its end
template <typename F>
int deferred(F&& body) {
  int rc;
  {
    SyncDefer defer;
    rc = body();
  }
  if (hipStreamSynchronize(stream()) != hipSuccess && rc >= 0) rc = DFM_HIP_ERROR;
  return rc;
}

int fx_work(dfm_filu_s* M, size_t n, int nvec) {
  while ((int)M->work.size() < nvec) {
    double* p = nullptr;
    DFM_TRY(dfm_malloc((void**)&p, std::max<size_t>(1, n) * sizeof(double)));
    M->work.push_back(p);
    DFM_TRY(dfm_vec_fill(std::max<size_t>(1, n), 0.0, p));
  }
  return DFM_OK;
}

// The LsSol solvers over one flex segment (src/ls/solver_ls_iter.cpp):
// MATVEC = the flex MatVec (alpha == 0 scales only, linearsystem.cpp:270-273),
// SolvePrecond = forward then backward substitution, DOT and the updates flat.
struct FxOps {
  dfm_fbcrs_s* A;
  dfm_filu_s* M;
  size_t n;
  int matvec(double alpha, const double* x, double beta, double* y) const {
    if (alpha == 0.0) return dfm_vec_scal(n, beta, y);
    return fx_matvec(A, alpha, x, beta, y);
  }
  int precond(double* z) const { return fx_sweep(M, z, true, true); }
  int dot(const double* a, const double* b, double* r) const { return dfm_vec_dot(n, a, b, r); }
};

// Solve_PCG (solver_ls_iter.cpp:100-172)
int fx_pcg(const FxOps& o, double* r, double* x, double* p, double* z, double* conv_ratio, unsigned* iteration) {
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  const size_t n = o.n;
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  double rr0 = 0.0;
  DFM_TRY(o.dot(r, r, &rr0));
  if (rr0 < 1.0e-30) {
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  const double inv_rr0 = 1.0 / rr0;
  DFM_TRY(dfm_vec_copy(n, r, z));
  DFM_TRY(o.precond(z));
  DFM_TRY(dfm_vec_copy(n, z, p));
  double rz = 0.0;
  DFM_TRY(o.dot(r, z, &rz));
  for (unsigned it = 0; it < mx_iter; it++) {
    DFM_TRY(o.matvec(1.0, p, 0.0, z));
    double pAp = 0.0;
    DFM_TRY(o.dot(p, z, &pAp));
    const double alpha = rz / pAp;
    DFM_TRY(dfm_vec_axpy(n, -alpha, z, r));
    DFM_TRY(dfm_vec_axpy(n, alpha, p, x));
    double rr = 0.0;
    DFM_TRY(o.dot(r, r, &rr));
    if (rr * inv_rr0 < tol * tol) {
      *conv_ratio = std::sqrt(rr * inv_rr0);
      *iteration = it;
      return DFM_OK;
    }
    DFM_TRY(dfm_vec_copy(n, r, z));
    DFM_TRY(o.precond(z));
    double rz_new = 0.0;
    DFM_TRY(o.dot(r, z, &rz_new));
    const double beta = rz_new / rz;
    rz = rz_new;
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(dfm_vec_axpy(n, 1.0, z, p));
  }
  double rr = 0.0;
  DFM_TRY(o.dot(r, r, &rr));
  *conv_ratio = std::sqrt(rr * inv_rr0);
  return DFM_NOT_CONVERGED;
}

// Solve_PBiCGSTAB (solver_ls_iter.cpp:291-408); w: 7 work vectors
int fx_pbicgstab(const FxOps& o, double* r, double* x, double* const* w, double* conv_ratio, unsigned* iteration) {
  const double tol = *conv_ratio;
  const unsigned max_iter = *iteration;
  const size_t n = o.n;
  double *vs = w[0], *Ms = w[1], *AMs = w[2], *p = w[3], *Mp = w[4], *AMp = w[5], *r2 = w[6];
  double rr0 = 0.0;
  DFM_TRY(o.dot(r, r, &rr0));
  if (rr0 < 1.0e-60) {
    *conv_ratio = std::sqrt(rr0);
    *iteration = 0;
    return DFM_OK;
  }
  const double inv_rr0 = 1.0 / rr0;
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  DFM_TRY(dfm_vec_copy(n, r, r2));
  DFM_TRY(dfm_vec_copy(n, r, p));
  *iteration = max_iter;
  for (unsigned it = 1; it < max_iter; it++) {
    DFM_TRY(dfm_vec_copy(n, p, Mp));
    DFM_TRY(o.precond(Mp));
    double r_r2 = 0.0;
    DFM_TRY(o.dot(r, r2, &r_r2));
    DFM_TRY(o.matvec(1.0, Mp, 0.0, AMp));
    double den = 0.0;
    DFM_TRY(o.dot(AMp, r2, &den));
    const double alpha = r_r2 / den;
    DFM_TRY(dfm_vec_copy(n, r, vs));
    DFM_TRY(dfm_vec_axpy(n, -alpha, AMp, vs));
    DFM_TRY(dfm_vec_copy(n, vs, Ms));
    DFM_TRY(o.precond(Ms));
    DFM_TRY(o.matvec(1.0, Ms, 0.0, AMs));
    double den2 = 0.0, num = 0.0;
    DFM_TRY(o.dot(AMs, AMs, &den2));
    DFM_TRY(o.dot(vs, AMs, &num));
    const double omega = num / den2;
    DFM_TRY(dfm_vec_axpy(n, alpha, Mp, x));
    DFM_TRY(dfm_vec_axpy(n, omega, Ms, x));
    DFM_TRY(dfm_vec_copy(n, vs, r));
    DFM_TRY(dfm_vec_axpy(n, -omega, AMs, r));
    double rr = 0.0;
    DFM_TRY(o.dot(r, r, &rr));
    if (rr * inv_rr0 < tol * tol) {
      *conv_ratio = std::sqrt(rr * inv_rr0);
      *iteration = it;
      return DFM_OK;
    }
    double tmp1 = 0.0;
    DFM_TRY(o.dot(r, r2, &tmp1));
    const double beta = tmp1 * alpha / (r_r2 * omega);
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(dfm_vec_axpy(n, 1.0, r, p));
    DFM_TRY(dfm_vec_axpy(n, -beta * omega, AMp, p));
  }
  return DFM_NOT_CONVERGED;
}

// Solve_CG (solver_ls_iter.cpp:31-97): no preconditioner; w: 2 work vectors.
// x zeroed first; 1e-60 early exit; num_iter = the 0-based index of the
// converging step, or max_iter (conv_ratio unchanged, DFM_NOT_CONVERGED;
// the reference returns true in both cases).
int fx_cg(const FxOps& o, double* r, double* x, double* const* w, double* conv_ratio, unsigned* num_iter) {
  const unsigned max_iter = *num_iter;
  const double tol = *conv_ratio;
  const size_t n = o.n;
  double *p = w[0], *Ap = w[1];
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  double rr = 0.0;
  DFM_TRY(o.dot(r, r, &rr));
  if (rr < 1.0e-60) {
    *num_iter = 0;
    *conv_ratio = std::sqrt(rr);
    return DFM_OK;
  }
  const double inv_rr0 = 1.0 / rr;
  DFM_TRY(dfm_vec_copy(n, r, p));
  for (unsigned it = 0; it < max_iter; it++) {
    DFM_TRY(o.matvec(1.0, p, 0.0, Ap));
    double pAp = 0.0;
    DFM_TRY(o.dot(p, Ap, &pAp));
    const double alpha = rr / pAp;
    DFM_TRY(dfm_vec_axpy(n, alpha, p, x));
    DFM_TRY(dfm_vec_axpy(n, -alpha, Ap, r));
    double rr_new = 0.0;
    DFM_TRY(o.dot(r, r, &rr_new));
    if (rr_new * inv_rr0 < tol * tol) {
      *conv_ratio = std::sqrt(rr_new * inv_rr0);
      *num_iter = it;
      return DFM_OK;
    }
    const double beta = rr_new / rr;
    rr = rr_new;
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(dfm_vec_axpy(n, 1.0, r, p));
  }
  *num_iter = max_iter;
  return DFM_NOT_CONVERGED;
}

// Solve_BiCGSTAB (solver_ls_iter.cpp:177-286): no preconditioner; w: 5
// vectors.  x zeroed first; (r, r2) carried from one iteration to the next;
// at max_iter iteration = max_iter, conv_ratio unchanged, DFM_NOT_CONVERGED
// (the reference returns true).
int fx_bicgstab(const FxOps& o, double* r, double* x, double* const* w, double* conv_ratio, unsigned* num_iter) {
  const unsigned max_iter = *num_iter;
  const double tol = *conv_ratio;
  const size_t n = o.n;
  double *vs = w[0], *As = w[1], *p = w[2], *Ap = w[3], *r2 = w[4];
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  double rr0 = 0.0;
  DFM_TRY(o.dot(r, r, &rr0));
  if (rr0 < 1.0e-60) {
    *conv_ratio = std::sqrt(rr0);
    *num_iter = 0;
    return DFM_OK;
  }
  const double inv_rr0 = 1.0 / rr0;
  DFM_TRY(dfm_vec_copy(n, r, r2));
  DFM_TRY(dfm_vec_copy(n, r, p));
  double r_r2 = 0.0;
  DFM_TRY(o.dot(r, r2, &r_r2));
  *num_iter = max_iter;
  for (unsigned it = 1; it < max_iter; it++) {
    DFM_TRY(o.matvec(1.0, p, 0.0, Ap));
    double den0 = 0.0;
    DFM_TRY(o.dot(Ap, r2, &den0));
    const double alpha = r_r2 / den0;
    DFM_TRY(dfm_vec_copy(n, r, vs));
    DFM_TRY(dfm_vec_axpy(n, -alpha, Ap, vs));
    DFM_TRY(o.matvec(1.0, vs, 0.0, As));
    double den = 0.0, num = 0.0;
    DFM_TRY(o.dot(As, As, &den));
    DFM_TRY(o.dot(As, vs, &num));
    const double omega = num / den;
    DFM_TRY(dfm_vec_axpy(n, alpha, p, x));
    DFM_TRY(dfm_vec_axpy(n, omega, vs, x));
    DFM_TRY(dfm_vec_copy(n, vs, r));
    DFM_TRY(dfm_vec_axpy(n, -omega, As, r));
    double rr = 0.0;
    DFM_TRY(o.dot(r, r, &rr));
    const double sq = rr * inv_rr0;
    if (sq < tol * tol) {
      *conv_ratio = std::sqrt(sq);
      *num_iter = it;
      return DFM_OK;
    }
    double tmp1 = 0.0;
    DFM_TRY(o.dot(r, r2, &tmp1));
    const double beta = (tmp1 * alpha) / (r_r2 * omega);
    r_r2 = tmp1;
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(dfm_vec_axpy(n, 1.0, r, p));
    DFM_TRY(dfm_vec_axpy(n, -beta * omega, Ap, p));
  }
  return DFM_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int dfm_fbcrs_create(unsigned nblk, const unsigned* alen, dfm_fbcrs** out) {
  DFM_CHECK(out && (nblk == 0 || alen), DFM_BAD_ARG, "bad arguments to dfm_fbcrs_create");
  DFM_TRY(ensure_init());
  for (unsigned i = 0; i < nblk; i++)
    DFM_CHECK(alen[i] >= 1 && alen[i] <= kFxMaxLen, DFM_BAD_ARG, "block %u has %u dofs (1..%u handled)", i, alen[i],
              kFxMaxLen);
  auto* A = new dfm_fbcrs_s;
  A->nb = nblk;
  A->dofptr.assign(nblk + 1, 0);
  for (unsigned i = 0; i < nblk; i++) A->dofptr[i + 1] = A->dofptr[i] + alen[i];
  A->rowptr.assign(nblk + 1, 0);
  std::vector<int> dofblk(std::max<unsigned>(1, A->dofptr[nblk]), 0);
  for (unsigned i = 0; i < nblk; i++)
    for (unsigned d = A->dofptr[i]; d < A->dofptr[i + 1]; d++) dofblk[d] = (int)i;
  int rc = upload_as_int(A->d_dofptr, A->dofptr);
  if (rc == DFM_OK) rc = A->d_dofblk.upload(dofblk);
  if (rc == DFM_OK) rc = fx_refresh_layout(A);
  if (rc != DFM_OK) {
    delete A;
    return rc;
  }
  *out = A;
  return DFM_OK;
}

int dfm_fbcrs_destroy(dfm_fbcrs* A) {
  delete A;
  return DFM_OK;
}

int dfm_fbcrs_add_pattern(dfm_fbcrs* A, unsigned nrow, const unsigned* index, const unsigned* array) {
  DFM_CHECK(A && index && nrow <= A->nb, DFM_BAD_ARG, "bad arguments to dfm_fbcrs_add_pattern");
  DFM_CHECK(index[0] == 0, DFM_BAD_ARG, "index[0] must be 0");
  for (unsigned i = 0; i < nrow; i++) {
    DFM_CHECK(index[i + 1] >= index[i], DFM_BAD_ARG, "index not monotone at row %u", i);
    for (unsigned q = index[i]; q < index[i + 1]; q++) {
      DFM_CHECK(array[q] < A->nb, DFM_BAD_ARG, "column %u out of range in row %u", array[q], i);
      DFM_CHECK(array[q] != i, DFM_BAD_ARG, "diagonal entry in row %u (the CRS part excludes the diagonal)", i);
    }
  }
  if (index[nrow] == 0) return DFM_OK;
  // union with the pattern already there, rows sorted ascending
  std::vector<unsigned> rp(A->nb + 1, 0), cl;
  for (unsigned i = 0; i < A->nb; i++) {
    std::vector<unsigned> row(A->col.begin() + A->rowptr[i], A->col.begin() + A->rowptr[i + 1]);
    if (i < nrow) row.insert(row.end(), array + index[i], array + index[i + 1]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    cl.insert(cl.end(), row.begin(), row.end());
    rp[i + 1] = (unsigned)cl.size();
  }
  A->rowptr.swap(rp);
  A->col.swap(cl);
  return fx_refresh_layout(A);
}

int dfm_fbcrs_info(const dfm_fbcrs* A, unsigned* nblk, unsigned* ndof, unsigned* ncrs, unsigned* nval,
                   unsigned* ndia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (nblk) *nblk = A->nb;
  if (ndof) *ndof = A->dofptr[A->nb];
  if (ncrs) *ncrs = (unsigned)A->col.size();
  if (nval) *nval = A->valptr.back();
  if (ndia) *ndia = A->diaptr.back();
  return DFM_OK;
}

int dfm_fbcrs_layout(const dfm_fbcrs* A, unsigned* dofptr, unsigned* rowptr, unsigned* col, unsigned* valptr,
                     unsigned* diaptr) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (dofptr) std::memcpy(dofptr, A->dofptr.data(), sizeof(unsigned) * (A->nb + 1));
  if (rowptr) std::memcpy(rowptr, A->rowptr.data(), sizeof(unsigned) * (A->nb + 1));
  if (col && !A->col.empty()) std::memcpy(col, A->col.data(), sizeof(unsigned) * A->col.size());
  if (valptr) std::memcpy(valptr, A->valptr.data(), sizeof(unsigned) * A->valptr.size());
  if (diaptr) std::memcpy(diaptr, A->diaptr.data(), sizeof(unsigned) * A->diaptr.size());
  return DFM_OK;
}

int dfm_fbcrs_set_zero(dfm_fbcrs* A) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  DFM_HIP(hipMemsetAsync(A->d_val.p, 0, A->d_val.n * sizeof(double), stream()));
  DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double), stream()));
  A->version = next_version();
  return sync_stream();
}

int dfm_fbcrs_set_values(dfm_fbcrs* A, const double* val, const double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val && A->valptr.back())
    DFM_HIP(hipMemcpyAsync(A->d_val.p, val, A->valptr.back() * sizeof(double), hipMemcpyHostToDevice, stream()));
  if (dia && A->diaptr.back())
    DFM_HIP(hipMemcpyAsync(A->d_dia.p, dia, A->diaptr.back() * sizeof(double), hipMemcpyHostToDevice, stream()));
  A->version = next_version();
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

int dfm_fbcrs_get_values(const dfm_fbcrs* A, double* val, double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val && A->valptr.back())
    DFM_HIP(hipMemcpyAsync(val, A->d_val.p, A->valptr.back() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  if (dia && A->diaptr.back())
    DFM_HIP(hipMemcpyAsync(dia, A->d_dia.p, A->diaptr.back() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

int dfm_fbcrs_device_values(dfm_fbcrs* A, double** val_dev, double** dia_dev) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val_dev) *val_dev = A->d_val.p;
  if (dia_dev) *dia_dev = A->d_dia.p;
  A->version = next_version();
  return DFM_OK;
}

int dfm_fbcrs_mearge(dfm_fbcrs* A, unsigned n, const unsigned* iblk, const unsigned* jblk, const double* emat) {
  DFM_CHECK(A && (n == 0 || (iblk && jblk && emat)), DFM_BAD_ARG, "bad arguments to dfm_fbcrs_mearge");
  if (n == 0) return DFM_OK;
  std::vector<long> dst(n);
  for (unsigned e = 0; e < n; e++) {
    const unsigned i = iblk[e], j = jblk[e];
    DFM_CHECK(i < A->nb && j < A->nb, DFM_BAD_ARG, "block (%u, %u) out of range", i, j);
    if (i == j) {
      dst[e] = -1 - (long)A->diaptr[i];
      continue;
    }
    long q = -1;
    for (unsigned k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
      if (A->col[k] == j) {
        q = (long)k;
        break;
      }
    DFM_CHECK(q >= 0, DFM_BAD_ARG, "block (%u, %u) is not in the pattern", i, j);
    dst[e] = (long)A->valptr[q];
  }
  std::vector<double> val(A->valptr.back()), dia(A->diaptr.back());
  DFM_TRY(dfm_fbcrs_get_values(A, val.data(), dia.data()));
  size_t off = 0;
  for (unsigned e = 0; e < n; e++) {
    const unsigned bs = (A->dofptr[iblk[e] + 1] - A->dofptr[iblk[e]]) * (A->dofptr[jblk[e] + 1] - A->dofptr[jblk[e]]);
    double* out = dst[e] < 0 ? dia.data() + (-1 - dst[e]) : val.data() + dst[e];
    for (unsigned t = 0; t < bs; t++) out[t] += emat[off + t];
    off += bs;
  }
  return dfm_fbcrs_set_values(A, val.data(), dia.data());
}

int dfm_fbcrs_set_bc(dfm_fbcrs* A, const int* flag) {
  DFM_CHECK(A && flag, DFM_BAD_ARG, "bad arguments to dfm_fbcrs_set_bc");
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, std::max<size_t>(1, A->ndof())));
  DFM_FLAUNCH(k_fx_bc, (int)A->nb, (int)A->nb, A->d_dofptr.p, A->d_rowptr.p, A->d_col.p, A->d_valptr.p, A->d_val.p,
              A->d_diaptr.p, A->d_dia.p, f.p);
  A->version = next_version();
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

int dfm_fbcrs_matvec(dfm_fbcrs* A, double alpha, const double* x_dev, double beta, double* y_dev) {
  DFM_CHECK(A && x_dev && y_dev, DFM_BAD_ARG, "bad arguments to dfm_fbcrs_matvec");
  return fx_matvec(A, alpha, x_dev, beta, y_dev);
}

int dfm_filu_create(const dfm_fbcrs* A, int lev_fill, dfm_filu** out) {
  DFM_CHECK(A && out, DFM_BAD_ARG, "bad arguments to dfm_filu_create");
  auto* M = new dfm_filu_s;
  M->nb = A->nb;
  M->lev_fill = lev_fill;
  M->A = A;
  M->pattern_version = A->pattern_version;
  std::vector<int> flev;
  int rc = iluk_symbolic(A->nb, A->rowptr, A->col, lev_fill, nullptr, M->frow, M->fcol, flev, M->fdiaind);
  if (rc != DFM_OK) {
    delete M;
    return rc;
  }
  std::vector<unsigned> fdiaptr;
  fx_layout(A->nb, A->dofptr, M->frow, M->fcol, M->fvalptr, fdiaptr);
  std::vector<int> levL, levU, rowsL, rowsU;
  M->nlev_lower = level_sets(A->nb, M->frow, M->fcol, M->fdiaind, true, levL);
  M->nlev_upper = level_sets(A->nb, M->frow, M->fcol, M->fdiaind, false, levU);
  level_rows(A->nb, levL, M->nlev_lower, M->lev_ptr_L, rowsL);
  level_rows(A->nb, levU, M->nlev_upper, M->lev_ptr_U, rowsU);
  // the matrix block of each factor entry (-1 for fill)
  std::vector<long long> fsrc(std::max<size_t>(1, M->fcol.size()), -1);
  for (unsigned i = 0; i < A->nb; i++) {
    unsigned p = A->rowptr[i];
    for (unsigned q = M->frow[i]; q < M->frow[i + 1]; q++) {
      while (p < A->rowptr[i + 1] && A->col[p] < M->fcol[q]) p++;
      if (p < A->rowptr[i + 1] && A->col[p] == M->fcol[q]) fsrc[q] = (long long)p;
    }
  }
  rc = upload_as_int(M->d_frow, M->frow);
  if (rc == DFM_OK) rc = upload_as_int(M->d_fcol, M->fcol);
  if (rc == DFM_OK) rc = upload_as_int(M->d_fdiaind, M->fdiaind);
  if (rc == DFM_OK) rc = upload_as_int(M->d_fvalptr, M->fvalptr);
  if (rc == DFM_OK) rc = upload_as_int(M->d_rows_L, rowsL);
  if (rc == DFM_OK) rc = upload_as_int(M->d_rows_U, rowsU);
  if (rc == DFM_OK) rc = M->d_fsrc.upload(fsrc);
  if (rc == DFM_OK) rc = M->d_fval.alloc(std::max<size_t>(1, M->fvalptr.back()));
  if (rc == DFM_OK) rc = M->d_fdia.alloc(std::max<size_t>(1, A->diaptr.back()));
  if (rc == DFM_OK) rc = M->d_nsing.alloc(1);
  if (rc == DFM_OK) rc = M->d_singrow.alloc(std::max<size_t>(1, A->nb));
  if (rc == DFM_OK) rc = M->d_fixrow.alloc(1);
  if (rc == DFM_OK) rc = hipStreamSynchronize(stream()) == hipSuccess ? DFM_OK : DFM_HIP_ERROR;
  if (rc != DFM_OK) {
    delete M;
    return rc;
  }
  *out = M;
  return DFM_OK;
}

int dfm_filu_destroy(dfm_filu* M) {
  delete M;
  return DFM_OK;
}

int dfm_filu_info(const dfm_filu* M, unsigned* ncrs, unsigned* nval, int* nlevel_lower, int* nlevel_upper) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (ncrs) *ncrs = (unsigned)M->fcol.size();
  if (nval) *nval = M->fvalptr.back();
  if (nlevel_lower) *nlevel_lower = M->nlev_lower;
  if (nlevel_upper) *nlevel_upper = M->nlev_upper;
  return DFM_OK;
}

int dfm_filu_get_pattern(const dfm_filu* M, unsigned* rowptr, unsigned* col, unsigned* diaind, unsigned* valptr) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (rowptr) std::memcpy(rowptr, M->frow.data(), sizeof(unsigned) * (M->nb + 1));
  if (col && !M->fcol.empty()) std::memcpy(col, M->fcol.data(), sizeof(unsigned) * M->fcol.size());
  if (diaind && M->nb) std::memcpy(diaind, M->fdiaind.data(), sizeof(unsigned) * M->nb);
  if (valptr) std::memcpy(valptr, M->fvalptr.data(), sizeof(unsigned) * M->fvalptr.size());
  return DFM_OK;
}

// SetValue_Initialize + DoILUDecomp.  With more than 10 singular pivots the
// factor is left as the reference's row-by-row loop leaves it: rows before
// the 11th singular one (in row order) factored, that row eliminated and
// half inverted without its U scaling, later rows holding the copied values.
int dfm_filu_set_value(dfm_filu* M, dfm_fbcrs* A, int* nsingular) {
  DFM_CHECK(M && A && A == M->A, DFM_BAD_ARG, "the factor was made on another matrix");
  DFM_CHECK(A->pattern_version == M->pattern_version, DFM_BAD_ARG,
            "the matrix pattern changed since the factor was made");
  const int nb = (int)M->nb;
  DFM_HIP(hipMemsetAsync(M->d_nsing.p, 0, sizeof(int), stream()));
  DFM_FLAUNCH(k_fx_copy_rows, nb, nb, 0, A->d_dofptr.p, M->d_frow.p, M->d_fcol.p, M->d_fvalptr.p, M->d_fsrc.p,
              A->d_valptr.p, A->d_val.p, A->d_diaptr.p, A->d_dia.p, M->d_fval.p, M->d_fdia.p);
  std::vector<int> rowsL_all;
  for (int l = 0; l < M->nlev_lower; l++) {
    const int off = M->lev_ptr_L[l], n = M->lev_ptr_L[l + 1] - off;
    DFM_TRY(fx_factor_rows(M, n, M->d_rows_L.p + off, 0));
  }
  int ns = 0;
  DFM_HIP(hipMemcpyAsync(&ns, M->d_nsing.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  if (nsingular) *nsingular = ns;
  M->factored = true;
  if (ns > 10) {
    std::vector<int> sing(std::max(1, nb));
    DFM_TRY(M->d_singrow.download(sing.data(), nb));
    int stop = -1, cnt = 0;
    for (int i = 0; i < nb && stop < 0; i++)
      if (sing[i] && ++cnt == 11) stop = i;
    DFM_CHECK(stop >= 0, DFM_HIP_ERROR, "singular pivot count and row flags disagree");
    DFM_FLAUNCH(k_fx_copy_rows, nb - stop, nb, stop, A->d_dofptr.p, M->d_frow.p, M->d_fcol.p, M->d_fvalptr.p,
                M->d_fsrc.p, A->d_valptr.p, A->d_val.p, A->d_diaptr.p, A->d_dia.p, M->d_fval.p, M->d_fdia.p);
    DFM_TRY(M->d_fixrow.upload(&stop, 1));
    DFM_TRY(fx_factor_rows(M, 1, M->d_fixrow.p, 1));
    DFM_HIP(hipStreamSynchronize(stream()));
    set_error("ILU factorisation met more than 10 singular pivots (stopped in row %d)", stop);
    return DFM_SINGULAR;
  }
  return DFM_OK;
}

int dfm_filu_get_values(const dfm_filu* M, double* val, double* dia) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (val && M->fvalptr.back())
    DFM_HIP(hipMemcpyAsync(val, M->d_fval.p, M->fvalptr.back() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  if (dia && M->A->diaptr.back())
    DFM_HIP(hipMemcpyAsync(dia, M->d_fdia.p, M->A->diaptr.back() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

int dfm_filu_forward(dfm_filu* M, double* vec_dev) {
  DFM_CHECK(M && vec_dev && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_filu_set_value)");
  return fx_sweep(M, vec_dev, true, false);
}
int dfm_filu_backward(dfm_filu* M, double* vec_dev) {
  DFM_CHECK(M && vec_dev && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_filu_set_value)");
  return fx_sweep(M, vec_dev, false, true);
}
int dfm_filu_solve_precond(dfm_filu* M, double* vec_dev) {
  DFM_CHECK(M && vec_dev && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_filu_set_value)");
  return fx_sweep(M, vec_dev, true, true);
}

// kind: 0 Solve_PCG, 1 Solve_PBiCGSTAB (M required, holding values),
// 2 Solve_CG, 3 Solve_BiCGSTAB (M may be NULL).  r in: right-hand side,
// out: final residual; x out: solution.  conv_ratio in tolerance, out
// achieved ratio; iteration in max, out count.  DFM_OK when converged,
// DFM_NOT_CONVERGED at max_iter (the reference's Solve_PCG returns false
// there, the other three true).
int dfm_fsolve(int kind, dfm_fbcrs* A, dfm_filu* M, double* r_dev, double* x_dev, double* conv_ratio,
               unsigned* iteration) {
  DFM_CHECK(A && r_dev && x_dev && conv_ratio && iteration && r_dev != x_dev && kind >= 0 && kind <= 3, DFM_BAD_ARG,
            "bad arguments to dfm_fsolve");
  DFM_CHECK(kind >= 2 || (M && M->A == A && M->factored), DFM_BAD_ARG,
            "the preconditioned solvers need a factor of this matrix holding values");
  DFM_TRY(ensure_init());
  const size_t n = A->ndof();
  dfm_filu_s tmp;  // work vectors of the unpreconditioned solvers when no factor is given
  dfm_filu_s* W = M ? M : &tmp;
  DFM_TRY(fx_work(W, n, 7));
  FxOps o{A, M, n};
  double* const* w = W->work.data();
  const int rc = deferred([&] {
    switch (kind) {
      case 0: return fx_pcg(o, r_dev, x_dev, w[0], w[1], conv_ratio, iteration);
      case 1: return fx_pbicgstab(o, r_dev, x_dev, w, conv_ratio, iteration);
      case 2: return fx_cg(o, r_dev, x_dev, w, conv_ratio, iteration);
      default: return fx_bicgstab(o, r_dev, x_dev, w, conv_ratio, iteration);
    }
  });
  return rc;
}

}  // extern "C"
