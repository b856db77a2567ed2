// dfm_bmat.hip -- MatVec::CMat_BlkCrs as the boundary matrix of the
// saved-stiffness linear system (Fem::Ls::CLinearSystem_Save,
// src/femls/linearsystem_fieldsave.cpp): square b x b blocks with the
// diagonal block kept inside the CRS rows, the pattern made by
// SetPatternBoundary from a diagonal-block matrix, the two boundary
// conditions and the MatVec that MakeResidual runs (r = f - K_b u).  All
// floating-point work keeps the reference's operation order.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "dfm_kernels.h"

using namespace dfm;

struct dfm_bmat_s {
  unsigned nb = 0, b = 0;
  std::vector<unsigned> rowptr, col;    // host pattern
  DevBuf<int> d_rowptr, d_col;
  DevBuf<double> d_val;                 // ncrs * b * b, row-major blocks
  DevBuf<long long> d_src;              // per entry: CRS index in the source matrix, or -1-i for its diagonal block i
  unsigned long long src_hash = 0;      // pattern hash of the matrix the pattern was made from
};

#define DFM_BLAUNCH(kernel, grid, ...)                                                      \
  do {                                                                                      \
    if ((grid) > 0) {                                                                       \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(kNT), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                           \
    }                                                                                       \
  } while (0)

namespace {

inline int bgrid(long n) { return (int)std::max<long>(1, (n + kNT - 1) / kNT); }

// values from the diagonal-block matrix the pattern was made from: the same
// element matrices merged in the same order give the same sums
__global__ void __launch_bounds__(kNT) k_bmat_copy(long n, int bb, const long long* __restrict__ src,
                                                  const double* __restrict__ aval, const double* __restrict__ adia,
                                                  double* __restrict__ val) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= n) return;
  const long q = t / bb, e = t % bb;
  const long long s = src[q];
  val[t] = s >= 0 ? aval[s * bb + e] : adia[(-1 - s) * bb + e];
}
// SetBoundaryConditionInverse_Colum (mat_blkcrs.cpp:876-907): one thread per
// block entry zeroes the columns of free dofs
__global__ void __launch_bounds__(kNT) k_bmat_bc_inv_col(long ncrs, int b, const int* __restrict__ col,
                                                        const int* __restrict__ flag, double* __restrict__ val) {
  const long q = (long)blockIdx.x * kNT + threadIdx.x;
  if (q >= ncrs) return;
  const int j = col[q];
  for (int jd = 0; jd < b; jd++) {
    if (flag[(long)j * b + jd] != 0) continue;
    for (int id = 0; id < b; id++) val[(q * b + id) * b + jd] = 0.0;
  }
}
// SetBoundaryCondition_Row (:910-952): one thread per row dof
__global__ void __launch_bounds__(kNT) k_bmat_bc_row(long ndof, int b, const int* __restrict__ rowptr,
                                                    const int* __restrict__ flag, double* __restrict__ val) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= ndof || flag[t] == 0) return;
  const long i = t / b, id = t % b;
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++)
    for (int jd = 0; jd < b; jd++) val[((long)q * b + id) * b + jd] = 0.0;
}
// MatVec (:955-1176, isnt_trans): one thread per row dof; y = y beta, then
// each (alpha a) x added straight into y; scalar blocks with beta = 0 start
// from 0
__global__ void __launch_bounds__(kNT) k_bmat_matvec(long ndof, int b, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    double alpha, const double* __restrict__ x, double beta,
                                                    double* __restrict__ y) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= ndof) return;
  const long i = t / b, id = t % b;
  double acc = (b == 1 && beta == 0.0) ? 0.0 : y[t] * beta;
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
    const double* a = val + ((long)q * b + id) * b;
    const double* xj = x + (long)col[q] * b;
    for (int jd = 0; jd < b; jd++) acc += alpha * a[jd] * xj[jd];
  }
  y[t] = acc;
}
__global__ void __launch_bounds__(kNT) k_copy_free(long n, const int* __restrict__ flag, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    if (flag[i] == 0) y[i] = x[i];
}

}  // namespace

int dfm_bmat_create(unsigned nblk_col, unsigned len_col, unsigned nblk_row, unsigned len_row, dfm_bmat** out) {
  DFM_CHECK(out && len_col > 0 && len_col <= 8, DFM_BAD_ARG, "bad arguments to dfm_bmat_create");
  DFM_CHECK(nblk_col == nblk_row && len_col == len_row, DFM_NOT_IMPLEMENTED,
            "dfm_bmat: square matrices of square blocks only (%ux%u blocks of %ux%u given)", nblk_col, nblk_row,
            len_col, len_row);
  DFM_TRY(ensure_init());
  std::unique_ptr<dfm_bmat_s> M(new dfm_bmat_s);
  M->nb = nblk_col;
  M->b = len_col;
  M->rowptr.assign(nblk_col + 1, 0);
  std::vector<int> rp(nblk_col + 1, 0);
  DFM_TRY(M->d_rowptr.upload(rp));
  DFM_TRY(M->d_val.alloc(1));
  DFM_HIP(hipStreamSynchronize(stream()));
  *out = M.release();
  return DFM_OK;
}
int dfm_bmat_destroy(dfm_bmat* M) {
  delete M;
  return DFM_OK;
}
int dfm_bmat_set_pattern_boundary(dfm_bmat* M, const dfm_bcrs* A, const int* flag_col, const int* flag_row) {
  DFM_CHECK(M && A && flag_col && flag_row && A->nb == M->nb && A->b == M->b, DFM_BAD_ARG,
            "dfm_bmat_set_pattern_boundary: matrix sizes differ");
  const unsigned nb = M->nb, b = M->b;
  std::vector<unsigned char> nz_col(nb, 0), nz_row(nb, 0);
  for (unsigned i = 0; i < nb; i++)
    for (unsigned d = 0; d < b; d++) {
      if (flag_col[(size_t)i * b + d] == 0) nz_col[i] = 1;
      if (flag_row[(size_t)i * b + d] != 0) nz_row[i] = 1;
    }
  std::vector<unsigned> rp(nb + 1, 0), cl;
  std::vector<long long> src;
  for (unsigned i = 0; i < nb; i++) {
    if (nz_col[i]) {
      if (nz_row[i]) {
        cl.push_back(i);
        src.push_back(-1 - (long long)i);
      }
      for (unsigned q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
        if (nz_row[A->col[q]]) {
          cl.push_back(A->col[q]);
          src.push_back(q);
        }
    }
    rp[i + 1] = (unsigned)cl.size();
  }
  M->rowptr.swap(rp);
  M->col.swap(cl);
  std::vector<int> rpi(M->rowptr.begin(), M->rowptr.end()), cli(M->col.begin(), M->col.end());
  DFM_TRY(M->d_rowptr.upload(rpi));
  DFM_TRY(M->d_col.upload(cli));
  DFM_TRY(M->d_src.upload(src));
  const size_t nval = std::max<size_t>(1, M->col.size() * b * b);
  DFM_TRY(M->d_val.alloc(nval));
  DFM_HIP(hipMemsetAsync(M->d_val.p, 0, nval * sizeof(double), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  M->src_hash = A->pattern_hash;
  return DFM_OK;
}
int dfm_bmat_info(const dfm_bmat* M, unsigned* nblk, unsigned* len, unsigned* ncrs) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL matrix");
  if (nblk) *nblk = M->nb;
  if (len) *len = M->b;
  if (ncrs) *ncrs = (unsigned)M->col.size();
  return DFM_OK;
}
int dfm_bmat_get_pattern(const dfm_bmat* M, unsigned* rowptr, unsigned* col) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL matrix");
  if (rowptr) std::memcpy(rowptr, M->rowptr.data(), sizeof(unsigned) * (M->nb + 1));
  if (col && !M->col.empty()) std::memcpy(col, M->col.data(), sizeof(unsigned) * M->col.size());
  return DFM_OK;
}
int dfm_bmat_set_zero(dfm_bmat* M) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL matrix");
  DFM_HIP(hipMemsetAsync(M->d_val.p, 0, M->d_val.n * sizeof(double), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_get_values(const dfm_bmat* M, double* val) {
  DFM_CHECK(M && val, DFM_BAD_ARG, "bad arguments to dfm_bmat_get_values");
  const size_t n = M->col.size() * M->b * M->b;
  if (n) DFM_HIP(hipMemcpyAsync(val, M->d_val.p, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_mearge(dfm_bmat* M, unsigned nblkel, const unsigned* blkel, unsigned blksize, const double* emat) {
  DFM_CHECK(M && blkel && emat && blksize == M->b * M->b, DFM_BAD_ARG, "bad arguments to dfm_bmat_mearge");
  const unsigned bb = blksize;
  for (unsigned ie = 0; ie < nblkel; ie++) DFM_CHECK(blkel[ie] < M->nb, DFM_BAD_ARG, "node out of range");
  for (unsigned ie = 0; ie < nblkel; ie++) {
    const unsigned i = blkel[ie];
    const unsigned q0 = M->rowptr[i], q1 = M->rowptr[i + 1];
    if (q1 == q0) continue;
    std::vector<double> rowv((size_t)(q1 - q0) * bb);
    DFM_HIP(hipMemcpyAsync(rowv.data(), M->d_val.p + (size_t)q0 * bb, rowv.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
    for (unsigned je = 0; je < nblkel; je++) {
      const unsigned j = blkel[je];
      for (unsigned q = q0; q < q1; q++)
        if (M->col[q] == j) {
          const double* in = emat + ((size_t)ie * nblkel + je) * bb;
          for (unsigned t = 0; t < bb; t++) rowv[(size_t)(q - q0) * bb + t] += in[t];
          break;
        }
    }
    DFM_HIP(hipMemcpyAsync(M->d_val.p + (size_t)q0 * bb, rowv.data(), rowv.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
  }
  return DFM_OK;
}
int dfm_bmat_copy_from(dfm_bmat* M, const dfm_bcrs* A) {
  DFM_CHECK(M && A && A->nb == M->nb && A->b == M->b, DFM_BAD_ARG, "dfm_bmat_copy_from: matrix sizes differ");
  DFM_CHECK(A->pattern_hash == M->src_hash, DFM_BAD_ARG,
            "dfm_bmat_copy_from: the matrix pattern differs from the one the boundary pattern was made from");
  const int bb = (int)(M->b * M->b);
  const long n = (long)M->col.size() * bb;
  if (n) DFM_BLAUNCH(k_bmat_copy, bgrid(n), n, bb, M->d_src.p, A->d_val.p, A->d_dia.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_set_bc_inverse_col(dfm_bmat* M, const int* flag) {
  DFM_CHECK(M && flag, DFM_BAD_ARG, "bad arguments to dfm_bmat_set_bc_inverse_col");
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, (size_t)M->nb * M->b));
  const long ncrs = (long)M->col.size();
  if (ncrs) DFM_BLAUNCH(k_bmat_bc_inv_col, bgrid(ncrs), ncrs, (int)M->b, M->d_col.p, f.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_set_bc_row(dfm_bmat* M, const int* flag) {
  DFM_CHECK(M && flag, DFM_BAD_ARG, "bad arguments to dfm_bmat_set_bc_row");
  DevBuf<int> f;
  const long ndof = (long)M->nb * M->b;
  DFM_TRY(f.upload(flag, ndof));
  if (ndof) DFM_BLAUNCH(k_bmat_bc_row, bgrid(ndof), ndof, (int)M->b, M->d_rowptr.p, f.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_matvec(dfm_bmat* M, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(M && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_bmat_matvec");
  const long ndof = (long)M->nb * M->b;
  if (ndof)
    DFM_BLAUNCH(k_bmat_matvec, bgrid(ndof), ndof, (int)M->b, M->d_rowptr.p, M->col.empty() ? nullptr : M->d_col.p,
                M->d_val.p, alpha, x, beta, y);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_vec_copy_free(size_t n, const int* flag, const double* x, double* y) {
  DFM_CHECK(flag && x && y, DFM_BAD_ARG, "bad arguments to dfm_vec_copy_free");
  DFM_TRY(ensure_init());
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, n));
  if (n) DFM_BLAUNCH(k_copy_free, (int)std::min<long>(bgrid((long)n), 2048), (long)n, f.p, x, y);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
