// dfm_bmat.hip -- MatVec::CMat_BlkCrs: block CRS with the diagonal block (if
// any) kept inside the rows, rectangular blocks allowed (rows of nbc blocks
// of bc dofs, columns of nbr blocks of br dofs).  Used as the boundary
// matrix of the saved-stiffness system (Fem::Ls::CLinearSystem_Save,
// src/femls/linearsystem_fieldsave.cpp; pattern by SetPatternBoundary) and
// as the off-diagonal blocks of multi-segment systems
// (src/ls/linearsystem.cpp).  All floating-point work keeps the reference's
// operation order, including MatVec's shape-specific forms.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "dfm_kernels.h"

using namespace dfm;


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
__global__ void __launch_bounds__(kNT) k_bmat_bc_row(long ndof, int bc, int br, const int* __restrict__ rowptr,
                                                    const int* __restrict__ flag, double* __restrict__ val) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= ndof || flag[t] == 0) return;
  const long i = t / bc, id = t % bc;
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++)
    for (int jd = 0; jd < br; jd++) val[((long)q * bc + id) * br + jd] = 0.0;
}
// MatVec (:955-1176, isnt_trans): one thread per row dof, with the
// reference's forms per block shape: scalar blocks (beta = 0 starts from 0),
// one row dof per block (alpha (a x) summed per entry), one column dof per
// block ((x alpha) scaled per entry) and the general (alpha a) x
__global__ void __launch_bounds__(kNT) k_bmat_matvec(long ndof, int bc, int br, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    double alpha, const double* __restrict__ x, double beta,
                                                    double* __restrict__ y) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= ndof) return;
  const long i = t / bc, id = t % bc;
  const long bs = (long)bc * br;
  // the specialised forms take CH entries at a time: indices, values and x
  // loaded before the first is used, then added one by one in row order
  constexpr int CH = 8;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  if (bc == 1 && br == 1) {
    double acc = beta != 0.0 ? y[t] * beta : 0.0;
    for (int qb = q0; qb < q1; qb += CH) {
      double v[CH], xv[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int q = qb + u < q1 ? qb + u : q0;
        v[u] = val[q];
        xv[u] = x[col[q]];
      }
#pragma unroll
      for (int u = 0; u < CH; u++)
        if (qb + u < q1) acc += alpha * v[u] * xv[u];
    }
    y[t] = acc;
  } else if (bc == 1 && (br == 2 || br == 3)) {
    double acc = y[t] * beta;
    for (int qb = q0; qb < q1; qb += CH) {
      double a[CH][3], xv[CH][3];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const long q = qb + u < q1 ? qb + u : q0;
        const long c = col[q];
#pragma unroll
        for (int e = 0; e < 3; e++) {
          a[u][e] = e < br ? val[q * br + e] : 0.0;
          xv[u][e] = e < br ? x[c * br + e] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        if (qb + u >= q1) continue;
        double s = a[u][0] * xv[u][0] + a[u][1] * xv[u][1];
        if (br == 3) s = s + a[u][2] * xv[u][2];
        acc += alpha * s;
      }
    }
    y[t] = acc;
  } else if (br == 1 && bc > 1) {
    double acc = y[t] * beta;
    for (int qb = q0; qb < q1; qb += CH) {
      double v[CH], xv[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const long q = qb + u < q1 ? qb + u : q0;
        v[u] = val[q * bc + id];
        xv[u] = x[col[q]];
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        if (qb + u >= q1) continue;
        const double ajx = xv[u] * alpha;
        acc += v[u] * ajx;
      }
    }
    y[t] = acc;
  } else {
    double acc = y[t] * beta;
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const double* a = val + (long)q * bs + id * br;
      const double* xj = x + (long)col[q] * br;
      for (int jd = 0; jd < br; jd++) acc += alpha * a[jd] * xj[jd];
    }
    y[t] = acc;
  }
}
// SetBoundaryCondition_Colum (:829-875): one thread per entry zeroes the
// columns of fixed dofs
__global__ void __launch_bounds__(kNT) k_bmat_bc_col(long ncrs, int bc, int br, const int* __restrict__ col,
                                                    const int* __restrict__ flag, double* __restrict__ val) {
  const long q = (long)blockIdx.x * kNT + threadIdx.x;
  if (q >= ncrs) return;
  const int j = col[q];
  for (int jd = 0; jd < br; jd++) {
    if (flag[(long)j * br + jd] == 0) continue;
    for (int id = 0; id < bc; id++) val[(q * bc + id) * br + jd] = 0.0;
  }
}
// values of a block taken from a combined matrix (b dofs per node) assembled
// by the same element loop: entry q's sub-block rows dof_r, columns dof_c
__global__ void __launch_bounds__(kNT) k_bmat_extract(long n, int bc, int br, int b, const long long* __restrict__ src,
                                                     const int* __restrict__ dof_r, const int* __restrict__ dof_c,
                                                     const double* __restrict__ aval,
                                                     const double* __restrict__ adia, double* __restrict__ val) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= n) return;
  const long bs = (long)bc * br;
  const long q = t / bs, e = t % bs, id = e / br, jd = e % br;
  const long long s = src[q];
  const long off = (long)dof_r[id] * b + dof_c[jd];
  val[t] = s >= 0 ? aval[s * b * b + off] : adia[(-1 - s) * b * b + off];
}
__global__ void __launch_bounds__(kNT) k_gather(long n, const long long* __restrict__ idx, const double* __restrict__ x,
                                               double* __restrict__ y) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) y[i] = x[idx[i]];
}
__global__ void __launch_bounds__(kNT) k_copy_free(long n, const int* __restrict__ flag, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    if (flag[i] == 0) y[i] = x[i];
}

}  // namespace

int dfm_bmat_create(unsigned nblk_col, unsigned len_col, unsigned nblk_row, unsigned len_row, dfm_bmat** out) {
  DFM_CHECK(out && len_col > 0 && len_col <= 8, DFM_BAD_ARG, "bad arguments to dfm_bmat_create");
  DFM_CHECK(len_row > 0 && len_row <= 8, DFM_BAD_ARG, "bad block size in dfm_bmat_create");
  DFM_TRY(ensure_init());
  std::unique_ptr<dfm_bmat_s> M(new dfm_bmat_s);
  M->nb = nblk_col;
  M->b = len_col;
  M->nbr = nblk_row;
  M->br = len_row;
  M->rowptr.assign(nblk_col + 1, 0);
  std::vector<int> rp(nblk_col + 1, 0);
  DFM_TRY(M->d_rowptr.upload(rp));
  DFM_TRY(M->d_val.alloc(1));
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  M->pattern_key = next_version();
  *out = M.release();
  return DFM_OK;
}
int dfm_bmat_destroy(dfm_bmat* M) {
  delete M;
  return DFM_OK;
}
int dfm_bmat_set_pattern_boundary(dfm_bmat* M, const dfm_bcrs* A, const int* flag_col, const int* flag_row) {
  DFM_CHECK(M && A && flag_col && flag_row && A->nb == M->nb && A->b == M->b && M->nbr == M->nb && M->br == M->b,
            DFM_BAD_ARG, "dfm_bmat_set_pattern_boundary: matrix sizes differ");
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
  const size_t nval = std::max<size_t>(1, M->col.size() * b * M->br);
  DFM_TRY(M->d_val.alloc(nval));
  DFM_HIP(hipMemsetAsync(M->d_val.p, 0, nval * sizeof(double), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  M->src_hash = A->pattern_hash;
  M->version = next_version();
  M->pattern_key = next_version();
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
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_get_values(const dfm_bmat* M, double* val) {
  DFM_CHECK(M && val, DFM_BAD_ARG, "bad arguments to dfm_bmat_get_values");
  const size_t n = M->col.size() * M->b * M->br;
  if (n) DFM_HIP(hipMemcpyAsync(val, M->d_val.p, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_bmat_set_values(dfm_bmat* M, const double* val) {
  DFM_CHECK(M && (val || M->col.empty()), DFM_BAD_ARG, "bad arguments to dfm_bmat_set_values");
  const size_t n = M->col.size() * M->b * M->br;
  if (n) DFM_HIP(hipMemcpyAsync(M->d_val.p, val, n * sizeof(double), hipMemcpyHostToDevice, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_mearge(dfm_bmat* M, unsigned nblkel, const unsigned* blkel, unsigned blksize, const double* emat) {
  DFM_CHECK(M && blkel && emat && blksize == M->b * M->br, DFM_BAD_ARG, "bad arguments to dfm_bmat_mearge");
  const unsigned bb = blksize;
  for (unsigned ie = 0; ie < nblkel; ie++)
    DFM_CHECK(blkel[ie] < M->nb && blkel[ie] < M->nbr, DFM_BAD_ARG, "node out of range");
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
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_copy_from(dfm_bmat* M, const dfm_bcrs* A) {
  DFM_CHECK(M && A && A->nb == M->nb && A->b == M->b && M->br == M->b, DFM_BAD_ARG,
            "dfm_bmat_copy_from: matrix sizes differ");
  DFM_CHECK(A->pattern_hash == M->src_hash, DFM_BAD_ARG,
            "dfm_bmat_copy_from: the matrix pattern differs from the one the boundary pattern was made from");
  const int bb = (int)(M->b * M->b);
  const long n = (long)M->col.size() * bb;
  if (n) DFM_BLAUNCH(k_bmat_copy, bgrid(n), n, bb, M->d_src.p, A->d_val.p, A->d_dia.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_set_bc_inverse_col(dfm_bmat* M, const int* flag) {
  DFM_CHECK(M && flag && M->br == M->b && M->nbr == M->nb, DFM_BAD_ARG, "bad arguments to dfm_bmat_set_bc_inverse_col");
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, (size_t)M->nb * M->b));
  const long ncrs = (long)M->col.size();
  if (ncrs) DFM_BLAUNCH(k_bmat_bc_inv_col, bgrid(ncrs), ncrs, (int)M->b, M->d_col.p, f.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_set_bc_row(dfm_bmat* M, const int* flag) {
  DFM_CHECK(M && flag, DFM_BAD_ARG, "bad arguments to dfm_bmat_set_bc_row");
  DevBuf<int> f;
  const long ndof = (long)M->nb * M->b;
  DFM_TRY(f.upload(flag, ndof));
  if (ndof) DFM_BLAUNCH(k_bmat_bc_row, bgrid(ndof), ndof, (int)M->b, (int)M->br, M->d_rowptr.p, f.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_matvec(dfm_bmat* M, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(M && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_bmat_matvec");
  const long ndof = (long)M->nb * M->b;
  if (ndof && !M->col.empty() && M->b <= 3 && M->br <= 3 && mv_sell_on()) {
    DFM_TRY(mv_sell_prepare(M->mv, M->nb, M->rowptr, M->col, (int)(M->b * M->br), M->d_val.p, M->pattern_key,
                            M->version));
    if (mv_sell_launch(M->mv, M->nb, false, (int)M->b, (int)M->br, nullptr, alpha, x, beta, y)) {
      DFM_HIP(hipGetLastError());
      return sync_stream();
    }
  }
  if (ndof)
    DFM_BLAUNCH(k_bmat_matvec, bgrid(ndof), ndof, (int)M->b, (int)M->br, M->d_rowptr.p,
                M->col.empty() ? nullptr : M->d_col.p, M->d_val.p, alpha, x, beta, y);
  return sync_stream();
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

int dfm_bmat_add_pattern(dfm_bmat* M, unsigned nrow, const unsigned* index, const unsigned* array) {
  DFM_CHECK(M && index && nrow == M->nb && index[0] == 0, DFM_BAD_ARG, "bad arguments to dfm_bmat_add_pattern");
  DFM_CHECK(M->col.empty(), DFM_NOT_IMPLEMENTED, "dfm_bmat_add_pattern: the matrix already has a pattern");
  for (unsigned i = 0; i < nrow; i++) {
    DFM_CHECK(index[i + 1] >= index[i], DFM_BAD_ARG, "index not monotone at row %u", i);
    for (unsigned q = index[i]; q < index[i + 1]; q++)
      DFM_CHECK(array[q] < M->nbr, DFM_BAD_ARG, "column %u out of range in row %u", array[q], i);
  }
  M->rowptr.assign(index, index + nrow + 1);
  M->col.assign(array, array + index[nrow]);
  std::vector<int> rpi(M->rowptr.begin(), M->rowptr.end()), cli(M->col.begin(), M->col.end());
  DFM_TRY(M->d_rowptr.upload(rpi));
  DFM_TRY(M->d_col.upload(cli));
  const size_t nval = std::max<size_t>(1, M->col.size() * M->b * M->br);
  DFM_TRY(M->d_val.alloc(nval));
  DFM_HIP(hipMemsetAsync(M->d_val.p, 0, nval * sizeof(double), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  M->src_hash = 0;
  M->version = next_version();
  M->pattern_key = next_version();
  return DFM_OK;
}
int dfm_bmat_set_bc_col(dfm_bmat* M, const int* flag) {
  DFM_CHECK(M && flag, DFM_BAD_ARG, "bad arguments to dfm_bmat_set_bc_col");
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, (size_t)M->nbr * M->br));
  const long ncrs = (long)M->col.size();
  if (ncrs) DFM_BLAUNCH(k_bmat_bc_col, bgrid(ncrs), ncrs, (int)M->b, (int)M->br, M->d_col.p, f.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bmat_extract(dfm_bmat* M, const dfm_bcrs* A, const unsigned* dofs_row, const unsigned* dofs_col) {
  DFM_CHECK(M && A && dofs_row && dofs_col && A->nb == M->nb && A->nb == M->nbr, DFM_BAD_ARG,
            "dfm_bmat_extract: block counts differ");
  for (unsigned d = 0; d < M->b; d++) DFM_CHECK(dofs_row[d] < A->b, DFM_BAD_ARG, "row dof out of range");
  for (unsigned d = 0; d < M->br; d++) DFM_CHECK(dofs_col[d] < A->b, DFM_BAD_ARG, "column dof out of range");
  std::vector<long long> src(M->col.size());
  for (unsigned i = 0; i < M->nb; i++)
    for (unsigned q = M->rowptr[i]; q < M->rowptr[i + 1]; q++) {
      const unsigned j = M->col[q];
      if (j == i) {
        src[q] = -1 - (long long)i;
        continue;
      }
      long long s = -1;
      for (unsigned p = A->rowptr[i]; p < A->rowptr[i + 1]; p++)
        if (A->col[p] == j) {
          s = p;
          break;
        }
      DFM_CHECK(s >= 0, DFM_BAD_ARG, "dfm_bmat_extract: entry (%u, %u) is not in the combined pattern", i, j);
      src[q] = s;
    }
  DevBuf<long long> d_src;
  DevBuf<int> dr, dc;
  std::vector<int> vr(dofs_row, dofs_row + M->b), vc(dofs_col, dofs_col + M->br);
  DFM_TRY(d_src.upload(src));
  DFM_TRY(dr.upload(vr));
  DFM_TRY(dc.upload(vc));
  const long n = (long)M->col.size() * M->b * M->br;
  if (n)
    DFM_BLAUNCH(k_bmat_extract, bgrid(n), n, (int)M->b, (int)M->br, (int)A->b, d_src.p, dr.p, dc.p, A->d_val.p,
                A->d_dia.p, M->d_val.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  M->version = next_version();
  return DFM_OK;
}
int dfm_bcrs_extract(dfm_bcrs* D, const dfm_bcrs* A, const unsigned* dofs) {
  DFM_CHECK(D && A && dofs && D->nb == A->nb && D->rowptr == A->rowptr && D->col == A->col, DFM_BAD_ARG,
            "dfm_bcrs_extract: the segment matrix needs the combined matrix's pattern");
  for (unsigned d = 0; d < D->b; d++) DFM_CHECK(dofs[d] < A->b, DFM_BAD_ARG, "dof out of range");
  const unsigned nb = D->nb, ncrs = D->rowptr[nb];
  std::vector<long long> src(ncrs);
  for (unsigned q = 0; q < ncrs; q++) src[q] = q;
  std::vector<long long> srcd(nb);
  for (unsigned i = 0; i < nb; i++) srcd[i] = -1 - (long long)i;
  DevBuf<long long> d_src, d_srcd;
  DevBuf<int> dr;
  std::vector<int> vr(dofs, dofs + D->b);
  DFM_TRY(d_src.upload(src));
  DFM_TRY(d_srcd.upload(srcd));
  DFM_TRY(dr.upload(vr));
  const int b = (int)D->b;
  const long nv = (long)ncrs * b * b, nd = (long)nb * b * b;
  if (nv)
    DFM_BLAUNCH(k_bmat_extract, bgrid(nv), nv, b, b, (int)A->b, d_src.p, dr.p, dr.p, A->d_val.p, A->d_dia.p,
                D->d_val.p);
  if (nd)
    DFM_BLAUNCH(k_bmat_extract, bgrid(nd), nd, b, b, (int)A->b, d_srcd.p, dr.p, dr.p, A->d_val.p, A->d_dia.p,
                D->d_dia.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  D->version = next_version();
  return DFM_OK;
}
int dfm_vec_gather(size_t n, const long long* idx, const double* x, double* y) {
  DFM_CHECK(idx && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_vec_gather");
  DFM_TRY(ensure_init());
  DevBuf<long long> d;
  DFM_TRY(d.upload(idx, n));
  if (n) DFM_BLAUNCH(k_gather, (int)std::min<long>(bgrid((long)n), 2048), (long)n, d.p, x, y);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
