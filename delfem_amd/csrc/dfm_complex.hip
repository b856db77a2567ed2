// dfm_complex.hip -- complex twin of the linear-solve path, scalar blocks as
// in the reference's complex classes:
//   MatVec::CZMatDia_BlkCrs      (src/matvec/zmatdia_blkcrs.cpp): pattern,
//     element assembly of the complex Helmholtz equation, Mearge, boundary
//     conditions, MatVec;
//   MatVec::CZMatDiaFrac_BlkCrs  (src/matvec/zmatdiafrac_blkcrs.cpp): ILU(k)
//     on the host-built symbolic pattern, numeric factorisation and
//     substitutions level-scheduled on the device;
//   MatVec::Sol::Solve_PCOCG     (src/matvec/zsolver_mat_iter.cpp:303-392):
//     the preconditioned conjugate orthogonal CG loop kept on the device and
//     replayed from hipGraphs of 16 iterations.
// A complex value is a double2 (re, im), the layout of Com::Complex and of
// std::complex<double>.  Every complex operation is written in the operation
// order of Com::Complex (include/delfem/complex.h:82-160), so assembly,
// MatVec, factorisation and substitutions reproduce the reference bit for bit
// in natural ordering.  The symbolic factor (fill pattern, orderings, level
// sets, sliced-ELL plans) is the real path's, built on the scalar pattern.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dfm_kernels.h"

using namespace dfm;

namespace {

// Device state of one complex Krylov run.
struct ZState {
  double2 rw, alpha, beta;  // (r, w) of the current iteration (COCG / PCOCG), step sizes
  double2 omega, r_r0;      // BiCGSTAB: omega, InnerProduct(r, r0*)
  double inv_rr0, tol, ratio;
  double d0, d1;            // generic reduction outputs
  int done;                 // 0 running, 1 converged, 2 stopped at max_iter, 3 zero right-hand side
  int iter, conv_iter, maxit;
};

// Workspace of the device-resident solvers: vectors, reduction partials,
// state (with a pinned host mirror) and the instantiated hipGraphs of
// 16-iteration blocks, keyed by their operands.
struct ZCtx {
  dfm::DevBuf<double2> work;
  dfm::DevBuf<double> partials;
  dfm::DevBuf<ZState> state;
  ZState* h = nullptr;
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
  ZCtx() = default;
  ZCtx(const ZCtx&) = delete;
  ~ZCtx() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (h) (void)hipHostFree(h);
  }
};

// A sliced-ELL operator: slot k of lane l at col / val[k*64+l], the lane's
// row at rows[s*64+l], its diagonal at dval[s*64+l].
struct ZOp {
  int nslice;
  const int *rows, *sptr, *col;
  const double2 *val, *dval;
};

}  // namespace

struct dfm_zbcrs_s {
  dfm_bcrs_s* pat = nullptr;            // scalar pattern: rowptr / col on host and device
  DevBuf<double2> d_val, d_dia;         // CRS order, diagonal separate
  unsigned long long version = 0;       // renewed whenever values change
  unsigned long long val_hash = 0;      // pattern d_val was sized for
  // natural-order sliced-ELL copy for the unpreconditioned solvers (built on first use)
  unsigned long long sell_hash = 0, sell_version = 0;
  int sell_nslice = 0;
  DevBuf<int> s_sptr, s_rows, s_col;
  DevBuf<long long> s_src;
  DevBuf<double2> s_val, s_dval;
  // transposed pattern for MatVec_Hermitian (built on first use): for column
  // i the rows j holding it, ascending, and the entry positions
  unsigned long long t_hash = 0;
  DevBuf<int> t_ptr, t_row, t_src;
  ZCtx ctx;
  ~dfm_zbcrs_s() {
    if (pat) dfm_bcrs_destroy(pat);
  }
};

struct dfm_zilu_s {
  dfm_ilu_s* M = nullptr;               // symbolic factor on the scalar pattern
  DevBuf<double2> d_fval, d_fdia;       // factor values, factor numbering, CRS order
  DevBuf<double2> zL, zLd, zU;          // sliced-ELL copies of L (with the inverted pivots) and U
  DevBuf<double2> zA, zAd;              // sliced-ELL copy of A in factor numbering
  unsigned long long a_version = 0;     // version of A mirrored in zA
  bool factored = false;
  ZCtx ctx;
  ~dfm_zilu_s() {
    if (M) dfm_ilu_destroy(M);
  }
};

#define DFM_ZLAUNCH(kernel, grid, nt, ...)                                                   \
  do {                                                                                      \
    if ((grid) > 0) {                                                                       \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(nt), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                           \
    }                                                                                       \
  } while (0)

namespace {

inline int zgrid_for(long n, int per = kNT) { return (int)std::max<long>(1, (n + per - 1) / per); }
inline int zgrid_vec(long n) { return (int)std::min<long>(std::max<long>(1, (n + kNT - 1) / kNT), 2048); }

// ---------------------------------------------------------------------------
// Com::Complex arithmetic (complex.h:82-160)
__device__ inline double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 zdiv(double2 a, double2 b) {
  const double f = 1.0 / (b.x * b.x + b.y * b.y);
  return make_double2((a.x * b.x + a.y * b.y) * f, (-a.x * b.y + a.y * b.x) * f);
}
__device__ inline double2 zrdiv(double d, double2 c) {  // double / Complex
  const double f = d / (c.x * c.x + c.y * c.y);
  return make_double2(f * c.x, -f * c.y);
}
__device__ inline double2 zdscale(double d, double2 c) { return make_double2(d * c.x, d * c.y); }
__device__ inline double2 zzero() { return make_double2(0.0, 0.0); }

typedef double f64x2 __attribute__((ext_vector_type(2)));
// streamed once per pass: non-temporal 16-byte load
__device__ inline double2 zload_nt(const double2* p) {
  const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p));
  return make_double2(v.x, v.y);
}

// One lane's row of a sliced-ELL matrix, slots [k0, k1): t -= a_k x[c_k]
// in slot order (padding slots, column < 0, are skipped), with the column,
// value and gathered x loads of four slots issued before their operations.
__device__ inline double2 zrow_sub(double2 t, int k0, int k1, int lane, const int* __restrict__ col,
                                   const double2* __restrict__ val, const double2* x) {
  int k = k0;
  for (; k + 4 <= k1; k += 4) {
    int c[4];
    double2 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) c[u] = __builtin_nontemporal_load(col + (size_t)(k + u) * 64 + lane);
#pragma unroll
    for (int u = 0; u < 4; u++) a[u] = zload_nt(val + (size_t)(k + u) * 64 + lane);
#pragma unroll
    for (int u = 0; u < 4; u++) b[u] = c[u] >= 0 ? x[c[u]] : zzero();
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (c[u] >= 0) t = zsub(t, zmul(a[u], b[u]));
  }
  for (; k < k1; k++) {
    const int c = __builtin_nontemporal_load(col + (size_t)k * 64 + lane);
    if (c < 0) break;
    t = zsub(t, zmul(zload_nt(val + (size_t)k * 64 + lane), x[c]));
  }
  return t;
}

// ---------------------------------------------------------------------------
// Element assembly, gathered per row: thread i adds the row-i entries of the
// elements around node i in ascending element order, the order in which
// sequential CZMatDia_BlkCrs::Mearge calls (zmatdia_blkcrs.cpp:198-237) add
// into each entry.

// Helmholtz triangle, src/femeqn/eqn_helmholtz.cpp:51-133: emat = area
// grad.grad - (k^2 area / 12)(1 + delta_ij), k = 2*3.1416 / wave_length,
// eres = -emat value.
__global__ void __launch_bounds__(kNT) k_zasm_helmholtz2d(int nnode, const unsigned* __restrict__ conn,
                                                         const int* __restrict__ eptr,
                                                         const int* __restrict__ elist,
                                                         const int* __restrict__ rowptr,
                                                         const int* __restrict__ col, double2* __restrict__ val,
                                                         double2* __restrict__ dia,
                                                         const double* __restrict__ coord, double wave_length,
                                                         const double2* __restrict__ value,
                                                         double2* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double2 racc = res ? res[i] : zzero();
  double2 dacc = dia[i];
  for (int t = eptr[i]; t < eptr[i + 1]; t++) {
    const int e = elist[t];
    unsigned no[3];
    double p[3][2];
    int a = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      no[k] = conn[(size_t)e * 3 + k];
      p[k][0] = coord[(size_t)no[k] * 2];
      p[k][1] = coord[(size_t)no[k] * 2 + 1];
      if ((int)no[k] == i) a = k;
    }
    const double area = d_tri_area(p[0], p[1], p[2]);
    double dl[3][2];
    d_tri_dldx(dl, p[0], p[1], p[2]);
    double2 em[3];
#pragma unroll
    for (int j = 0; j < 3; j++) em[j] = make_double2(area * (dl[a][0] * dl[j][0] + dl[a][1] * dl[j][1]), 0.0);
    const double k = 2 * 3.1416 / wave_length;
    const double tm = k * k * area / 12.0;
    em[a].x -= tm;
    em[a].y -= 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      em[j].x -= tm;
      em[j].y -= 0.0;
    }
    double2 er = zzero();
#pragma unroll
    for (int j = 0; j < 3; j++) er = zsub(er, zmul(em[j], value[no[j]]));
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j == a) {
        dacc = zadd(dacc, em[j]);
      } else {
        const int q = find_col(col, q0, q1, (int)no[j]);
        if (q >= 0) val[q] = zadd(val[q], em[j]);
      }
    }
    racc = zadd(racc, er);
  }
  dia[i] = dacc;
  if (res) res[i] = racc;
}

// Sommerfeld radiation line, eqn_helmholtz.cpp:437-497:
// emat = i k elen/6 [2 1; 1 2] + i/(2 elen k) [-1 1; 1 -1], eres = -emat value.
__global__ void __launch_bounds__(kNT) k_zasm_sommerfeld(int nnode, const unsigned* __restrict__ conn,
                                                        const int* __restrict__ eptr, const int* __restrict__ elist,
                                                        const int* __restrict__ rowptr,
                                                        const int* __restrict__ col, double2* __restrict__ val,
                                                        double2* __restrict__ dia, const double* __restrict__ coord,
                                                        double wave_length, const double2* __restrict__ value,
                                                        double2* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double2 racc = res ? res[i] : zzero();
  double2 dacc = dia[i];
  for (int t = eptr[i]; t < eptr[i + 1]; t++) {
    const int e = elist[t];
    const unsigned n0 = conn[(size_t)e * 2], n1 = conn[(size_t)e * 2 + 1];
    const int a = ((int)n0 == i) ? 0 : 1;
    const double x0 = coord[(size_t)n0 * 2], y0 = coord[(size_t)n0 * 2 + 1];
    const double x1 = coord[(size_t)n1 * 2], y1 = coord[(size_t)n1 * 2 + 1];
    const double elen = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1));
    const double k = 2 * 3.1416 / wave_length;
    const double2 unit_i = make_double2(0.0, 1.0);
    const double2 t1 = zdscale(k / 6.0 * elen, unit_i);
    const double2 t2 = zdscale(-1 / (2.0 * elen * k), unit_i);
    const double2 ed = zadd(zdscale(2, t1), t2), eo = zsub(t1, t2);
    const double2 em[2] = {a == 0 ? ed : eo, a == 0 ? eo : ed};
    double2 er = zzero();
    er = zsub(er, zmul(em[0], value[n0]));
    er = zsub(er, zmul(em[1], value[n1]));
    dacc = zadd(dacc, ed);
    const int q = find_col(col, q0, q1, (int)(a == 0 ? n1 : n0));
    if (q >= 0) val[q] = zadd(val[q], eo);
    racc = zadd(racc, er);
  }
  dia[i] = dacc;
  if (res) res[i] = racc;
}

// SetBoundaryCondition, zmatdia_blkcrs.cpp:239-268 (scalar blocks): a fixed
// row becomes the unit row, a fixed column is zeroed.
__global__ void __launch_bounds__(kNT) k_zbc_rows(int nb, const int* __restrict__ rowptr, double2* __restrict__ val,
                                                 double2* __restrict__ dia, const int* __restrict__ flag) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb || flag[i] == 0) return;
  dia[i] = make_double2(1.0, 0.0);
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++) val[q] = zzero();
}
__global__ void __launch_bounds__(kNT) k_zbc_cols(long ncrs, const int* __restrict__ col, double2* __restrict__ val,
                                                 const int* __restrict__ flag) {
  const long q = (long)blockIdx.x * kNT + threadIdx.x;
  if (q < ncrs && flag[col[q]]) val[q] = zzero();
}
__global__ void __launch_bounds__(kNT) k_zzero_bc(long n, const int* __restrict__ flag, double2* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    if (flag[i]) x[i] = zzero();
}

// MatVec, zmatdia_blkcrs.cpp:737-770 (scalar blocks): y = beta y + alpha A x
// with the off-diagonal sum accumulated as alpha * (a x) per entry.
__global__ void __launch_bounds__(kNT) k_zcrs_matvec(int nb, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double2* __restrict__ val,
                                                    const double2* __restrict__ dia, double alpha,
                                                    const double2* __restrict__ x, double beta,
                                                    double2* __restrict__ y) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  double re = 0.0, im = 0.0;
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
    const double2 a = val[q], b = x[col[q]];
    re += alpha * (a.x * b.x - a.y * b.y);
    im += alpha * (a.x * b.y + a.y * b.x);
  }
  const double2 yi = y[i];
  double2 t = make_double2(yi.x * beta, yi.y * beta);
  t.x += re;
  t.y += im;
  y[i] = zadd(t, zmul(zdscale(alpha, dia[i]), x[i]));
}

// MatVec_Hermitian (zmatdia_blkcrs.cpp:803-847): y = alpha A^H x + beta y.
// The reference scales y by beta, then walks the rows j of A adding
// alpha*InnerProduct(x_j, a_ji) = alpha*(x_j conj(a_ji)) to y_i, and the
// diagonal term to y_j after row j's entries.  Per y_i that order is: rows
// j < i holding column i (ascending), the diagonal, rows j > i; one thread
// per y_i gathers them so through the transposed pattern.
__device__ inline double2 zinner(double2 a, double2 b) {  // a conj(b), complex.h:153-156
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__global__ void __launch_bounds__(kNT) k_zcrs_matvec_h(int nb, const int* __restrict__ tptr,
                                                      const int* __restrict__ trow, const int* __restrict__ tsrc,
                                                      const double2* __restrict__ val,
                                                      const double2* __restrict__ dia, double alpha,
                                                      const double2* __restrict__ x, double beta,
                                                      double2* __restrict__ y) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  const double2 yi = y[i];
  double2 t = make_double2(yi.x * beta, yi.y * beta);
  bool diag = false;
  for (int q = tptr[i]; q < tptr[i + 1]; q++) {
    const int j = trow[q];
    if (!diag && j > i) {
      t = zadd(t, zdscale(alpha, zinner(x[i], dia[i])));
      diag = true;
    }
    t = zadd(t, zdscale(alpha, zinner(x[j], val[tsrc[q]])));
  }
  if (!diag) t = zadd(t, zdscale(alpha, zinner(x[i], dia[i])));
  y[i] = t;
}

// ---------------------------------------------------------------------------
// ILU.  Copy of A into the factor pattern (CZMatDiaFrac_BlkCrs::SetValue,
// zmatdiafrac_blkcrs.cpp:466-486): fill entries start at zero.
__global__ void __launch_bounds__(kNT) k_zilu_copy(long nf, const long long* __restrict__ fsrc,
                                                  const double2* __restrict__ aval, double2* __restrict__ fval) {
  const long q = (long)blockIdx.x * kNT + threadIdx.x;
  if (q >= nf) return;
  const long long s = fsrc[q];
  fval[q] = s >= 0 ? aval[s] : zzero();
}
__global__ void __launch_bounds__(kNT) k_zilu_copy_dia(int nb, const int* __restrict__ order,
                                                      const double2* __restrict__ adia, double2* __restrict__ fdia) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i < nb) fdia[i] = adia[order ? order[i] : i];
}

// DoILUDecomp, zmatdiafrac_blkcrs.cpp:192-262, for the rows of one level of
// the lower factor: every row it reads (the U part of rows k < i) is final.
// Pivots with |d|^2 < 1e-20 are counted; the reference reports them and
// divides anyway.
__global__ void __launch_bounds__(kNT) k_zilu_factor(int nrow, const int* __restrict__ rows,
                                                    const int* __restrict__ frow, const int* __restrict__ fcol,
                                                    const int* __restrict__ fdiaind, double2* __restrict__ fval,
                                                    double2* __restrict__ fdia, int* __restrict__ nsmall) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int i = rows[t];
  const int q0 = frow[i], q1 = frow[i + 1];
  double2 di = fdia[i];
  for (int ik = q0; ik < q1; ik++) {
    const int k = fcol[ik];
    if (k >= i) break;
    const double2 ikv = fval[ik];
    const int kend = frow[k + 1];
    for (int kj = fdiaind[k]; kj < kend; kj++) {
      const int j = fcol[kj];
      const double2 kjv = fval[kj];
      if (j != i) {
        const int ij = find_col(fcol, q0, q1, j);
        if (ij < 0) continue;
        fval[ij] = zsub(fval[ij], zmul(ikv, kjv));
      } else {
        di = zsub(di, zmul(ikv, kjv));
      }
    }
  }
  if (di.x * di.x + di.y * di.y < 1.0e-20) atomicAdd(nsmall, 1);
  const double2 inv = zrdiv(1.0, di);
  fdia[i] = inv;
  for (int q = fdiaind[i]; q < q1; q++) fval[q] = zmul(fval[q], inv);
}

// Sliced-ELL value copies: slot k lane l holds src[k*64+l] (-1: padding);
// the per-lane diagonal is dia[perm[row]] (perm NULL: identity).
__global__ void __launch_bounds__(kNT) k_zsell_gather(long nslot, const long long* __restrict__ src,
                                                     const double2* __restrict__ v, double2* __restrict__ out) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= nslot) return;
  const long long s = src[t];
  out[t] = s >= 0 ? v[s] : zzero();
}
__global__ void __launch_bounds__(kNT) k_zsell_dia(long nlane, const int* __restrict__ rows,
                                                  const int* __restrict__ perm, const double2* __restrict__ dia,
                                                  double2* __restrict__ out) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= nlane) return;
  const int row = rows[t];
  out[t] = row >= 0 ? dia[perm ? perm[row] : row] : zzero();
}

// Forward substitution (zmatdiafrac_blkcrs.cpp:116-138) of the slices
// [s0, s1) of one level, one wavefront per slice, lane = row:
// x_i = d_i (src_i - sum_q L_iq x_col(q)), the sum taken in CRS order.  src
// may be x (in place) or another vector (the Krylov loop reads r and writes
// w = M^-1 r without a copy: every row belongs to exactly one level).
__global__ void __launch_bounds__(64) k_zfwd(int s0, int s1, const int* __restrict__ rows,
                                            const int* __restrict__ sptr, const int* __restrict__ col,
                                            const double2* __restrict__ val, const double2* __restrict__ dval,
                                            const double2* src, double2* x, const ZState* __restrict__ st) {
  if (st && st->done) return;
  const int s = s0 + xcd_block();
  if (s >= s1) return;
  const int lane = threadIdx.x;
  const int row = rows[(size_t)s * 64 + lane];
  if (row < 0) return;
  const double2 t = zrow_sub(src[row], sptr[s], sptr[s + 1], lane, col, val, x);
  x[row] = zmul(dval[(size_t)s * 64 + lane], t);
}
// Backward substitution (:140-157): x_i -= sum_q U_iq x_col(q).  DOT: also
// the slice's share of (r, x) for the final x (partials[s], partials[nparts+s]):
// every row is finished by exactly one slice.
template <bool DOT>
__global__ void __launch_bounds__(64) k_zbwd(int s0, int s1, const int* __restrict__ rows,
                                            const int* __restrict__ sptr, const int* __restrict__ col,
                                            const double2* __restrict__ val, double2* __restrict__ x,
                                            const double2* __restrict__ r, double* __restrict__ partials, int nparts,
                                            const ZState* __restrict__ st) {
  if (st && st->done) return;
  const int s = s0 + xcd_block();
  if (s >= s1) return;
  const int lane = threadIdx.x;
  const int row = rows[(size_t)s * 64 + lane];
  double v0 = 0.0, v1 = 0.0;
  if (row >= 0) {
    const double2 t = zrow_sub(x[row], sptr[s], sptr[s + 1], lane, col, val, x);
    x[row] = t;
    if (DOT) {
      const double2 ri = r[row];
      v0 = ri.x * t.x - ri.y * t.y;
      v1 = ri.x * t.y + ri.y * t.x;
    }
  }
  if (DOT) {
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    if (lane == 0) {
      partials[s] = v0;
      partials[(size_t)nparts + s] = v1;
    }
  }
}

// ---------------------------------------------------------------------------
// Vectors (CZVector_Blk, src/matvec/zvector_blk.cpp).  Reductions write one
// partial per workgroup and a one-workgroup finalize kernel adds them in
// index order: results do not depend on timing.
__global__ void __launch_bounds__(kNT) k_zgather(long n, const int* __restrict__ idx, const double2* __restrict__ src,
                                                double2* __restrict__ dst) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    dst[i] = src[idx ? idx[i] : i];
}
__global__ void __launch_bounds__(kNT) k_zfill(long n, double2* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) x[i] = zzero();
}
// operator* (zvector_blk.cpp:39-60): sum a_i b_i without conjugation
__global__ void __launch_bounds__(kNT) k_zdot(long n, const double2* __restrict__ a, const double2* __restrict__ b,
                                             double* __restrict__ partials, const ZState* __restrict__ st) {
  if (st && st->done) return;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const double2 x = a[i], y = b[i];
    v[0] += x.x * y.x - x.y * y.y;
    v[1] += x.x * y.y + x.y * y.x;
  }
  write_partials<2>(v, partials, blockIdx.x, gridDim.x);
}
// InnerProduct (zvector_blk.cpp:62-80): sum conj(a_i) b_i
__global__ void __launch_bounds__(kNT) k_zdotc(long n, const double2* __restrict__ a, const double2* __restrict__ b,
                                              double* __restrict__ partials) {
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const double2 x = a[i], y = b[i];
    v[0] += x.x * y.x + x.y * y.y;
    v[1] += x.x * y.y - x.y * y.x;
  }
  write_partials<2>(v, partials, blockIdx.x, gridDim.x);
}
// GetSquaredVectorNorm (:154-161)
__global__ void __launch_bounds__(kNT) k_znorm2(long n, const double2* __restrict__ a, double* __restrict__ partials) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const double2 x = a[i];
    v[0] += x.x * x.x + x.y * x.y;
  }
  write_partials<1>(v, partials, blockIdx.x, gridDim.x);
}

enum {
  ZFIN_STORE = 0, ZFIN_RR0 = 1, ZFIN_RW0 = 2, ZFIN_RR0C = 3,             // run also after the stop
  ZFIN_ALPHA = 4, ZFIN_RATIO = 5, ZFIN_BETA = 6,                         // COCG / PCOCG
  ZFIN_BALPHA = 7, ZFIN_BOMEGA = 8, ZFIN_BCONV = 9, ZFIN_BCONVP = 10     // BiCGSTAB / BiCGStabP
};

// Epilogues of Solve_PCOCG (zsolver_mat_iter.cpp:303-392) on the reduced sums.
template <int ND, int FIN>
__global__ void __launch_bounds__(kFinNT) k_zfinalize(const double* __restrict__ partials, int nparts,
                                                     ZState* __restrict__ st) {
  if (FIN >= ZFIN_ALPHA && st->done) return;
  __shared__ double sh[kFinNT / 64];
  double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < ND; d++) s[d] = block_sum_n<kFinNT>(strided_sum<kFinNT>(partials + (size_t)d * nparts, nparts), sh);
  if (threadIdx.x != 0) return;
  if constexpr (FIN == ZFIN_STORE) {
    st->d0 = s[0];
    st->d1 = s[1];
  } else if constexpr (FIN == ZFIN_RR0) {
    st->d0 = s[0];
    if (s[0] < 1.0e-30) {
      st->done = 3;
    } else {
      st->inv_rr0 = 1.0 / s[0];
    }
  } else if constexpr (FIN == ZFIN_RW0) {
    st->rw = make_double2(s[0], s[1]);
  } else if constexpr (FIN == ZFIN_RR0C) {
    st->r_r0 = make_double2(s[0], s[1]);
  } else if constexpr (FIN == ZFIN_BALPHA) {
    if (st->iter >= st->maxit) {
      st->done = 2;
      return;
    }
    st->alpha = zdiv(st->r_r0, make_double2(s[0], s[1]));
  } else if constexpr (FIN == ZFIN_BOMEGA) {
    // Complex / double (complex.h:148-150)
    st->omega = make_double2(s[1] / s[0], s[2] / s[0]);
  } else if constexpr (FIN == ZFIN_BCONV || FIN == ZFIN_BCONVP) {
    const double sq_ratio = s[0] * st->inv_rr0;
    if (FIN == ZFIN_BCONVP) st->ratio = sqrt(s[0] * st->inv_rr0);
    if (sq_ratio < st->tol * st->tol) {
      if (FIN == ZFIN_BCONV) st->ratio = sqrt(s[0] * st->inv_rr0);
      st->conv_iter = st->iter;
      st->done = 1;
      return;
    }
    const double2 t1 = make_double2(s[1], s[2]);
    st->beta = FIN == ZFIN_BCONVP ? zmul(zdiv(st->alpha, st->omega), zdiv(t1, st->r_r0))
                                  : zmul(zdiv(t1, st->r_r0), zdiv(st->alpha, st->omega));
    st->r_r0 = t1;
    st->iter++;
  } else if constexpr (FIN == ZFIN_ALPHA) {
    if (st->iter >= st->maxit) {  // the loop runs iterations 1 .. max_iter-1
      st->done = 2;
      return;
    }
    st->alpha = zdiv(st->rw, make_double2(s[0], s[1]));
  } else if constexpr (FIN == ZFIN_RATIO) {
    const double ratio = sqrt(s[0] * st->inv_rr0);
    if (ratio < st->tol) {
      st->ratio = ratio;
      st->conv_iter = st->iter;
      st->done = 1;
    }
  } else if constexpr (FIN == ZFIN_BETA) {
    const double2 rw2 = make_double2(s[0], s[1]);
    st->beta = zdiv(rw2, st->rw);
    st->rw = rw2;
    st->iter++;
  }
}

// y = A x over a sliced-ELL operator, one wavefront per slice (as MatVec(1,
// x, 0, y) except that the old y is not read: 0 times a finite value), with
// this workgroup's partials of
//   EP_PAP     x . y                          (Solve_COCG / Solve_PCOCG)
//   EP_CONJ    InnerProduct(y, aux) = sum conj(y) aux       (BiCGSTAB alpha)
//   EP_OMEGA   ||y||^2 and InnerProduct(aux, y)             (BiCGSTAB omega)
enum { EP_PAP = 0, EP_CONJ = 1, EP_OMEGA = 2 };
template <int EP>
__global__ void __launch_bounds__(kNT) k_zspmv(int nslice, const int* __restrict__ rows, const int* __restrict__ sptr,
                                              const int* __restrict__ col, const double2* __restrict__ val,
                                              const double2* __restrict__ dval, const double2* __restrict__ x,
                                              double2* __restrict__ y, const double2* __restrict__ aux,
                                              double* __restrict__ partials, const ZState* __restrict__ st) {
  if (st->done) return;
  constexpr int ND = EP == EP_OMEGA ? 3 : 2;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = xcd_block() * kWaves + wave;
  double v[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) v[d] = 0.0;
  if (s < nslice) {
    const int row = rows[(size_t)s * 64 + lane];
    if (row >= 0) {
      double re = 0.0, im = 0.0;
      const int k1 = sptr[s + 1];
      int k = sptr[s];
      for (; k + 4 <= k1; k += 4) {
        int c[4];
        double2 a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) c[u] = __builtin_nontemporal_load(col + (size_t)(k + u) * 64 + lane);
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = zload_nt(val + (size_t)(k + u) * 64 + lane);
#pragma unroll
        for (int u = 0; u < 4; u++) b[u] = c[u] >= 0 ? x[c[u]] : zzero();
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (c[u] >= 0) {
            re += a[u].x * b[u].x - a[u].y * b[u].y;
            im += a[u].x * b[u].y + a[u].y * b[u].x;
          }
      }
      for (; k < k1; k++) {
        const int c = __builtin_nontemporal_load(col + (size_t)k * 64 + lane);
        if (c < 0) break;
        const double2 a = zload_nt(val + (size_t)k * 64 + lane), b = x[c];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
      }
      const double2 xr = x[row];
      const double2 yr = zadd(make_double2(re, im), zmul(dval[(size_t)s * 64 + lane], xr));
      y[row] = yr;
      if constexpr (EP == EP_PAP) {
        v[0] = xr.x * yr.x - xr.y * yr.y;
        v[1] = xr.x * yr.y + xr.y * yr.x;
      } else if constexpr (EP == EP_CONJ) {
        const double2 a = aux[row];
        v[0] = yr.x * a.x + yr.y * a.y;
        v[1] = yr.x * a.y - yr.y * a.x;
      } else {
        const double2 a = aux[row];
        v[0] = yr.x * yr.x + yr.y * yr.y;
        v[1] = a.x * yr.x + a.y * yr.y;
        v[2] = a.x * yr.y - a.y * yr.x;
      }
    }
  }
  write_partials<ND>(v, partials, blockIdx.x, gridDim.x);
}
// r += (-alpha) Ap, with the partial of ||r||^2
__global__ void __launch_bounds__(kNT) k_zupd_r(long n, const double2* __restrict__ Ap, double2* __restrict__ r,
                                               double* __restrict__ partials, const ZState* __restrict__ st) {
  if (st->done) return;
  const double2 ma = make_double2(-st->alpha.x, -st->alpha.y);
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const double2 ri = zadd(r[i], zmul(ma, Ap[i]));
    r[i] = ri;
    v[0] += ri.x * ri.x + ri.y * ri.y;
  }
  write_partials<1>(v, partials, blockIdx.x, gridDim.x);
}
// u += alpha p (after the convergence test, as the reference), then
// p = p beta + w
__global__ void __launch_bounds__(kNT) k_zupd_up(long n, const double2* __restrict__ w, double2* __restrict__ p,
                                                double2* __restrict__ u, const ZState* __restrict__ st) {
  if (st->done) return;
  const double2 al = st->alpha, be = st->beta;
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const double2 pi = p[i];
    u[i] = zadd(u[i], zmul(al, pi));
    p[i] = zadd(zmul(pi, be), w[i]);
  }
}

// BiCGSTAB (zsolver_mat_iter.cpp:487-616, 622-739) vector steps
// s = r + (-alpha) AMp
__global__ void __launch_bounds__(kNT) k_zbicg_s(long n, const double2* __restrict__ r,
                                                const double2* __restrict__ AMp, double2* __restrict__ sv,
                                                const ZState* __restrict__ st) {
  if (st->done) return;
  const double2 ma = make_double2(-st->alpha.x, -st->alpha.y);
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    sv[i] = zadd(r[i], zmul(ma, AMp[i]));
}
// u += alpha Mp; u += omega Ms; r = s + (-omega) AMs; partials of ||r||^2
// and InnerProduct(r, r0*)
__global__ void __launch_bounds__(kNT) k_zbicg_ur(long n, const double2* __restrict__ Mp,
                                                 const double2* __restrict__ Ms, const double2* __restrict__ sv,
                                                 const double2* __restrict__ AMs, const double2* __restrict__ r0,
                                                 double2* __restrict__ u, double2* __restrict__ r,
                                                 double* __restrict__ partials, const ZState* __restrict__ st) {
  if (st->done) return;
  const double2 al = st->alpha, om = st->omega, mo = make_double2(-st->omega.x, -st->omega.y);
  double v[3] = {0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    double2 ui = zadd(u[i], zmul(al, Mp[i]));
    u[i] = zadd(ui, zmul(om, Ms[i]));
    const double2 ri = zadd(sv[i], zmul(mo, AMs[i]));
    r[i] = ri;
    const double2 a = r0[i];
    v[0] += ri.x * ri.x + ri.y * ri.y;
    v[1] += ri.x * a.x + ri.y * a.y;
    v[2] += ri.x * a.y - ri.y * a.x;
  }
  write_partials<3>(v, partials, blockIdx.x, gridDim.x);
}
// p = p beta; p += r; p += (-beta omega) AMp
__global__ void __launch_bounds__(kNT) k_zbicg_p(long n, const double2* __restrict__ r,
                                                const double2* __restrict__ AMp, double2* __restrict__ p,
                                                const ZState* __restrict__ st) {
  if (st->done) return;
  const double2 be = st->beta;
  const double2 c = zmul(make_double2(-be.x, -be.y), st->omega);
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    double2 t = zmul(p[i], be);
    t = zadd(t, r[i]);
    p[i] = zadd(t, zmul(c, AMp[i]));
  }
}
// x = conj(x) (SetVectorConjugate, zvector_blk.cpp:149-152)
__global__ void __launch_bounds__(kNT) k_zconj(long n, double2* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) x[i].y = -x[i].y;
}

// ---------------------------------------------------------------------------
// host helpers
int zrefresh_a(dfm_zilu_s* Z, const dfm_zbcrs_s* A) {
  if (Z->a_version == A->version) return DFM_OK;
  dfm_ilu_s* M = Z->M;
  const long nslot = M->A64.nk * kSlice, nlane = (long)M->A64.nslice * kSlice;
  DFM_ZLAUNCH(k_zsell_gather, zgrid_for(nslot), kNT, nslot, M->d_srcA.p, A->d_val.p, Z->zA.p);
  DFM_ZLAUNCH(k_zsell_dia, zgrid_for(nlane), kNT, nlane, M->A64.rows.p, M->permuted ? M->d_order.p : nullptr,
              A->d_dia.p, Z->zAd.p);
  Z->a_version = A->version;
  return DFM_OK;
}

// M^-1 in factor numbering, one launch per level: forward from src into x
// (src may be x), then backward on x.  dot_r: also (r, x) into partials
// (one entry per backward slice).
int zsweep(dfm_zilu_s* Z, const double2* src, double2* x, const ZState* st, bool fwd, bool bwd,
           const double2* dot_r = nullptr, double* partials = nullptr) {
  dfm_ilu_s* M = Z->M;
  if (fwd)
    for (int l = 0; l < M->nlev_lower; l++) {
      const int s0 = M->L64.group_slice_ptr[l], s1 = M->L64.group_slice_ptr[l + 1];
      DFM_ZLAUNCH(k_zfwd, s1 - s0, 64, s0, s1, M->L64.rows.p, M->L64.slice_ptr.p, M->L64.col.p, Z->zL.p, Z->zLd.p,
                  src, x, st);
    }
  if (bwd)
    for (int l = 0; l < M->nlev_upper; l++) {
      const int s0 = M->U64.group_slice_ptr[l], s1 = M->U64.group_slice_ptr[l + 1];
      if (dot_r)
        DFM_ZLAUNCH(k_zbwd<true>, s1 - s0, 64, s0, s1, M->U64.rows.p, M->U64.slice_ptr.p, M->U64.col.p, Z->zU.p, x,
                    dot_r, partials, M->U64.nslice, st);
      else
        DFM_ZLAUNCH(k_zbwd<false>, s1 - s0, 64, s0, s1, M->U64.rows.p, M->U64.slice_ptr.p, M->U64.col.p, Z->zU.p, x,
                    (const double2*)nullptr, (double*)nullptr, 0, st);
    }
  return DFM_OK;
}

int zread_state(ZCtx& c) {
  DFM_HIP(hipMemcpyAsync(c.h, c.state.p, sizeof(ZState), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

// Sizes the workspace for nvec vectors of n entries and nparts partials per
// reduction (three reductions at most), and writes the initial state.
int zctx_begin(ZCtx& c, long n, int nvec, int nparts, double tol, unsigned mx_iter) {
  if (c.work.n < (size_t)(nvec * n)) DFM_TRY(c.work.alloc(std::max<long>(1, nvec * n)));
  if (c.partials.n < (size_t)(3 * nparts)) DFM_TRY(c.partials.alloc(3 * nparts));
  if (!c.state.n) DFM_TRY(c.state.alloc(1));
  if (!c.h) DFM_HIP(hipHostMalloc((void**)&c.h, sizeof(ZState), hipHostMallocDefault));
  std::memset(c.h, 0, sizeof(ZState));
  c.h->tol = tol;
  c.h->maxit = (int)std::min<unsigned>(mx_iter, 0x7fffffffu);
  c.h->iter = 1;
  DFM_HIP(hipMemcpyAsync(c.state.p, c.h, sizeof(ZState), hipMemcpyHostToDevice, stream()));
  return DFM_OK;
}

// Runs `nblock` iterations of `body` as one hipGraph, instantiated once per
// operand set and cached (directly where the stream cannot be captured).
template <class Body>
int zrun_block(ZCtx& c, const std::string& key, int nblock, Body&& body) {
  hipStream_t s = stream();
  static const bool use_graphs = [] {
    const char* e = getenv("DFM_GRAPHS");
    return !(e && e[0] == '0');
  }();
  if (s == nullptr || !use_graphs) {
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
    return DFM_OK;
  }
  for (auto& g : c.graphs)
    if (g.first == key) {
      DFM_HIP(hipGraphLaunch(g.second, s));
      return DFM_OK;
    }
  DFM_TRY(evict_graph(c.graphs));
  hipGraph_t graph = nullptr;
  DFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = DFM_OK;
  for (int i = 0; i < nblock && rc == DFM_OK; i++) rc = body();
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != DFM_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
    return DFM_OK;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DFM_HIP(e);
  c.graphs.emplace_back(key, exec);
  DFM_HIP(hipGraphLaunch(exec, s));
  return DFM_OK;
}

// Replays 16-iteration blocks until the state says stop.
template <class Body>
int zloop(ZCtx& c, const char* tag, const void* k1, const void* k2, Body&& body) {
  const int check_every = 16;
  char key[192];
  snprintf(key, sizeof key, "%s:%p:%p:%p:%p:%d", tag, (void*)c.work.p, (void*)c.partials.p, k1, k2, check_every);
  for (;;) {
    DFM_TRY(zrun_block(c, key, check_every, body));
    DFM_TRY(zread_state(c));
    if (c.h->done) return DFM_OK;
  }
}

// The natural-order sliced-ELL copy of A (one group, rows in order, entries
// in CRS order), for the unpreconditioned solvers.
int znatural_op(dfm_zbcrs_s* A, ZOp& op) {
  const dfm_bcrs_s* P = A->pat;
  const unsigned nb = P->nb;
  if (A->sell_hash != P->pattern_hash || !A->s_rows.n) {
    std::vector<int> rows(nb), gptr = {0, (int)nb};
    std::vector<unsigned> beg(nb), end(nb);
    for (unsigned i = 0; i < nb; i++) {
      rows[i] = (int)i;
      beg[i] = P->rowptr[i];
      end[i] = P->rowptr[i + 1];
    }
    SellPlan plan;
    build_sell_plan(rows, gptr, beg, end, P->col, nullptr, plan);
    DFM_TRY(A->s_sptr.upload(plan.slice_ptr));
    DFM_TRY(A->s_rows.upload(plan.rows));
    DFM_TRY(A->s_col.upload(plan.col));
    DFM_TRY(A->s_src.upload(plan.src));
    DFM_TRY(A->s_val.alloc(std::max<size_t>(1, (size_t)plan.nk * kSlice)));
    DFM_TRY(A->s_dval.alloc(std::max<size_t>(1, (size_t)plan.nslice * kSlice)));
    A->sell_nslice = plan.nslice;
    A->sell_hash = P->pattern_hash;
    A->sell_version = 0;
  }
  if (A->sell_version != A->version) {
    const long nslot = (long)A->s_src.n, nlane = (long)A->sell_nslice * kSlice;
    if (nslot) DFM_ZLAUNCH(k_zsell_gather, zgrid_for(nslot), kNT, nslot, A->s_src.p, A->d_val.p, A->s_val.p);
    if (nlane)
      DFM_ZLAUNCH(k_zsell_dia, zgrid_for(nlane), kNT, nlane, A->s_rows.p, (const int*)nullptr, A->d_dia.p, A->s_dval.p);
    A->sell_version = A->version;
  }
  op = ZOp{A->sell_nslice, A->s_rows.p, A->s_sptr.p, A->s_col.p, A->s_val.p, A->s_dval.p};
  return DFM_OK;
}

ZOp zfactor_op(const dfm_zilu_s* Z) {
  const dfm_ilu_s* M = Z->M;
  return ZOp{M->A64.nslice, M->A64.rows.p, M->A64.slice_ptr.p, M->A64.col.p, Z->zA.p, Z->zAd.p};
}

template <int EP>
int zspmv(const ZOp& A, const double2* x, double2* y, const double2* aux, double* parts, const ZState* st) {
  const int g = std::max(1, (A.nslice + kWaves - 1) / kWaves);
  DFM_ZLAUNCH(k_zspmv<EP>, g, kNT, A.nslice, A.rows, A.sptr, A.col, A.val, A.dval, x, y, aux, parts, st);
  return DFM_OK;
}
inline int zspmv_grid(const ZOp& A) { return std::max(1, (A.nslice + kWaves - 1) / kWaves); }

// MatVec::Sol::Solve_PCOCG (zsolver_mat_iter.cpp:303-392) with Z's factor, or
// Solve_COCG (:220-299) when Z is null (w = r).  Vectors in the operator's
// numbering: perm / inv map it to the caller's (null: identity).
int zcocg_core(ZCtx& c, const ZOp& A, dfm_zilu_s* Z, long n, const int* perm, const int* inv, double2* r_dev,
               double2* u_dev, double* conv_ratio, unsigned* iteration) {
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  const int gv = zgrid_vec(n), gA = zspmv_grid(A), nsU = Z ? Z->M->U64.nslice : 0;
  DFM_TRY(zctx_begin(c, n, Z ? 5 : 4, std::max(std::max(gv, gA), nsU), tol, mx_iter));
  double2 *r = c.work.p, *u = r + n, *p = u + n, *Ap = p + n, *w = Z ? Ap + n : r;
  double* parts = c.partials.p;
  ZState* st = c.state.p;
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, perm, (const double2*)r_dev, r);
  DFM_ZLAUNCH(k_zfill, gv, kNT, n, u);
  DFM_ZLAUNCH(k_znorm2, gv, kNT, n, (const double2*)r, parts);
  DFM_ZLAUNCH((k_zfinalize<1, ZFIN_RR0>), 1, kFinNT, parts, gv, st);
  DFM_TRY(zread_state(c));
  if (c.h->done == 3) {  // ||r||^2 < 1e-30: conv_ratio 0, iteration 0, u = 0
    DFM_ZLAUNCH(k_zfill, gv, kNT, n, u_dev);
    DFM_HIP(hipStreamSynchronize(stream()));
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  // w = M^-1 r (with the partials of (r, w)), p = w
  if (Z) {
    DFM_TRY(zsweep(Z, r, w, nullptr, true, true, r, parts));
    DFM_ZLAUNCH((k_zfinalize<2, ZFIN_RW0>), 1, kFinNT, parts, nsU, st);
  } else {
    DFM_ZLAUNCH(k_zdot, gv, kNT, n, (const double2*)r, (const double2*)r, parts, (const ZState*)nullptr);
    DFM_ZLAUNCH((k_zfinalize<2, ZFIN_RW0>), 1, kFinNT, parts, gv, st);
  }
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, (const int*)nullptr, (const double2*)w, p);
  if (mx_iter > 1) {
    auto body = [&]() -> int {
      DFM_TRY(zspmv<EP_PAP>(A, p, Ap, nullptr, parts, st));
      DFM_ZLAUNCH((k_zfinalize<2, ZFIN_ALPHA>), 1, kFinNT, parts, gA, st);
      DFM_ZLAUNCH(k_zupd_r, gv, kNT, n, (const double2*)Ap, r, parts, (const ZState*)st);
      DFM_ZLAUNCH((k_zfinalize<1, ZFIN_RATIO>), 1, kFinNT, parts, gv, st);
      if (Z) {
        DFM_TRY(zsweep(Z, r, w, st, true, true, r, parts));
        DFM_ZLAUNCH((k_zfinalize<2, ZFIN_BETA>), 1, kFinNT, parts, nsU, st);
      } else {
        DFM_ZLAUNCH(k_zdot, gv, kNT, n, (const double2*)r, (const double2*)r, parts, (const ZState*)st);
        DFM_ZLAUNCH((k_zfinalize<2, ZFIN_BETA>), 1, kFinNT, parts, gv, st);
      }
      DFM_ZLAUNCH(k_zupd_up, gv, kNT, n, (const double2*)w, p, u, (const ZState*)st);
      return DFM_OK;
    };
    DFM_TRY(zloop(c, Z ? "pcocg" : "cocg", A.val, A.rows, body));
  }
  DFM_TRY(zread_state(c));
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, inv, (const double2*)r, r_dev);
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, inv, (const double2*)u, u_dev);
  DFM_HIP(hipStreamSynchronize(stream()));
  if (c.h->done == 1) {
    *conv_ratio = c.h->ratio;
    *iteration = (unsigned)c.h->conv_iter;
  } else {
    *iteration = mx_iter;
  }
  return DFM_OK;
}

// MatVec::Sol::Solve_BiCGStabP (zsolver_mat_iter.cpp:622-739) with Z's factor,
// or Solve_BiCGSTAB (:487-616) when Z is null.  The preconditioned form
// starts from r = conj(b) with r0* = b and stores conv_ratio at every
// iteration (at max_iter it holds the last ratio); both return true.
int zbicg_core(ZCtx& c, const ZOp& A, dfm_zilu_s* Z, long n, const int* perm, const int* inv, double2* r_dev,
               double2* u_dev, double* conv_ratio, unsigned* iteration) {
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  const int gv = zgrid_vec(n), gA = zspmv_grid(A);
  DFM_TRY(zctx_begin(c, n, Z ? 9 : 7, std::max(gv, gA), tol, mx_iter));
  double2 *r = c.work.p, *u = r + n, *p = u + n, *sv = p + n, *r0 = sv + n, *AMp = r0 + n, *AMs = AMp + n;
  double2 *Mp = Z ? AMs + n : p, *Ms = Z ? Mp + n : sv;
  double* parts = c.partials.p;
  ZState* st = c.state.p;
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, perm, (const double2*)r_dev, r);
  DFM_ZLAUNCH(k_zfill, gv, kNT, n, u);
  DFM_ZLAUNCH(k_znorm2, gv, kNT, n, (const double2*)r, parts);
  DFM_ZLAUNCH((k_zfinalize<1, ZFIN_RR0>), 1, kFinNT, parts, gv, st);
  DFM_TRY(zread_state(c));
  if (c.h->done == 3) {
    DFM_ZLAUNCH(k_zfill, gv, kNT, n, u_dev);
    DFM_HIP(hipStreamSynchronize(stream()));
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, (const int*)nullptr, (const double2*)r, r0);
  if (Z) DFM_ZLAUNCH(k_zconj, gv, kNT, n, r);
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, (const int*)nullptr, (const double2*)r, p);
  // InnerProduct(r, r0*) = (r0*, r) with a conjugate on r: k_zspmv's EP_CONJ form
  // over the vectors, here as a plain reduction
  DFM_ZLAUNCH(k_zdotc, gv, kNT, n, (const double2*)r, (const double2*)r0, parts);
  DFM_ZLAUNCH((k_zfinalize<2, ZFIN_RR0C>), 1, kFinNT, parts, gv, st);
  if (mx_iter > 1) {
    auto body = [&]() -> int {
      if (Z) DFM_TRY(zsweep(Z, p, Mp, st, true, true));
      DFM_TRY(zspmv<EP_CONJ>(A, Mp, AMp, r0, parts, st));
      DFM_ZLAUNCH((k_zfinalize<2, ZFIN_BALPHA>), 1, kFinNT, parts, gA, st);
      DFM_ZLAUNCH(k_zbicg_s, gv, kNT, n, (const double2*)r, (const double2*)AMp, sv, (const ZState*)st);
      if (Z) DFM_TRY(zsweep(Z, sv, Ms, st, true, true));
      DFM_TRY(zspmv<EP_OMEGA>(A, Ms, AMs, sv, parts, st));
      DFM_ZLAUNCH((k_zfinalize<3, ZFIN_BOMEGA>), 1, kFinNT, parts, gA, st);
      DFM_ZLAUNCH(k_zbicg_ur, gv, kNT, n, (const double2*)Mp, (const double2*)Ms, (const double2*)sv,
                  (const double2*)AMs, (const double2*)r0, u, r, parts, (const ZState*)st);
      if (Z)
        DFM_ZLAUNCH((k_zfinalize<3, ZFIN_BCONVP>), 1, kFinNT, parts, gv, st);
      else
        DFM_ZLAUNCH((k_zfinalize<3, ZFIN_BCONV>), 1, kFinNT, parts, gv, st);
      DFM_ZLAUNCH(k_zbicg_p, gv, kNT, n, (const double2*)r, (const double2*)AMp, p, (const ZState*)st);
      return DFM_OK;
    };
    DFM_TRY(zloop(c, Z ? "bicgstabp" : "bicgstab", A.val, A.rows, body));
  }
  DFM_TRY(zread_state(c));
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, inv, (const double2*)r, r_dev);
  DFM_ZLAUNCH(k_zgather, gv, kNT, n, inv, (const double2*)u, u_dev);
  DFM_HIP(hipStreamSynchronize(stream()));
  if (c.h->done == 1) {
    *conv_ratio = c.h->ratio;
    *iteration = (unsigned)c.h->conv_iter;
  } else {
    if (Z && mx_iter > 1 && c.h->iter > 1) *conv_ratio = c.h->ratio;
    *iteration = mx_iter;
  }
  return DFM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// MatVec::CZMatDia_BlkCrs
int dfm_zbcrs_create(unsigned nblk, unsigned len, dfm_zbcrs** out) {
  DFM_CHECK(out && len == 1, DFM_BAD_ARG,
            "dfm_zbcrs_create: the complex matrix has scalar blocks (len=%u given)", len);
  std::unique_ptr<dfm_zbcrs_s> Z(new dfm_zbcrs_s);
  DFM_TRY(dfm_bcrs_create(nblk, 1, &Z->pat));
  DFM_TRY(Z->d_dia.alloc(std::max<unsigned>(1, nblk)));
  DFM_TRY(Z->d_val.alloc(1));
  DFM_HIP(hipMemsetAsync(Z->d_dia.p, 0, Z->d_dia.n * sizeof(double2), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  Z->val_hash = Z->pat->pattern_hash;
  Z->version = next_version();
  *out = Z.release();
  return DFM_OK;
}
int dfm_zbcrs_destroy(dfm_zbcrs* A) {
  delete A;
  return DFM_OK;
}
int dfm_zbcrs_add_pattern(dfm_zbcrs* A, unsigned nrow, const unsigned* index, const unsigned* array) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  DFM_TRY(dfm_bcrs_add_pattern(A->pat, nrow, index, array));
  if (A->pat->pattern_hash == A->val_hash) return DFM_OK;
  // the pattern grew: values restart from zero, as the real matrix's
  const size_t ncrs = A->pat->rowptr[A->pat->nb];
  DFM_TRY(A->d_val.alloc(std::max<size_t>(1, ncrs)));
  DFM_HIP(hipMemsetAsync(A->d_val.p, 0, A->d_val.n * sizeof(double2), stream()));
  DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double2), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  A->val_hash = A->pat->pattern_hash;
  A->version = next_version();
  return DFM_OK;
}
int dfm_zbcrs_info(const dfm_zbcrs* A, unsigned* nblk, unsigned* len, unsigned* ncrs) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  return dfm_bcrs_info(A->pat, nblk, len, ncrs);
}
int dfm_zbcrs_get_pattern(const dfm_zbcrs* A, unsigned* rowptr, unsigned* col) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  return dfm_bcrs_get_pattern(A->pat, rowptr, col);
}
int dfm_zbcrs_set_zero(dfm_zbcrs* A) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  DFM_HIP(hipMemsetAsync(A->d_val.p, 0, A->d_val.n * sizeof(double2), stream()));
  DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double2), stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  A->version = next_version();
  return DFM_OK;
}
int dfm_zbcrs_set_values(dfm_zbcrs* A, const double* val, const double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  const size_t ncrs = A->pat->rowptr[A->pat->nb];
  if (val && ncrs) DFM_HIP(hipMemcpyAsync(A->d_val.p, val, ncrs * sizeof(double2), hipMemcpyHostToDevice, stream()));
  if (dia && A->pat->nb)
    DFM_HIP(hipMemcpyAsync(A->d_dia.p, dia, A->pat->nb * sizeof(double2), hipMemcpyHostToDevice, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  A->version = next_version();
  return DFM_OK;
}
int dfm_zbcrs_get_values(const dfm_zbcrs* A, double* val, double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  const size_t ncrs = A->pat->rowptr[A->pat->nb];
  if (val && ncrs) DFM_HIP(hipMemcpyAsync(val, A->d_val.p, ncrs * sizeof(double2), hipMemcpyDeviceToHost, stream()));
  if (dia && A->pat->nb)
    DFM_HIP(hipMemcpyAsync(dia, A->d_dia.p, A->pat->nb * sizeof(double2), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_zbcrs_device_values(dfm_zbcrs* A, double** val, double** dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val) *val = (double*)A->d_val.p;
  if (dia) *dia = (double*)A->d_dia.p;
  A->version = next_version();  // the caller may write through the pointers
  return DFM_OK;
}
int dfm_zbcrs_mearge(dfm_zbcrs* A, unsigned nblkel, const unsigned* blkel, unsigned blksize, const double* emat) {
  DFM_CHECK(A && blkel && emat && blksize == 1, DFM_BAD_ARG, "bad arguments to dfm_zbcrs_mearge");
  const dfm_bcrs_s* P = A->pat;
  for (unsigned ie = 0; ie < nblkel; ie++) DFM_CHECK(blkel[ie] < P->nb, DFM_BAD_ARG, "node out of range");
  const double2* em = (const double2*)emat;
  for (unsigned ie = 0; ie < nblkel; ie++) {
    const unsigned i = blkel[ie];
    const unsigned q0 = P->rowptr[i], q1 = P->rowptr[i + 1];
    std::vector<double2> rowv(q1 - q0);
    double2 d;
    if (q1 > q0)
      DFM_HIP(hipMemcpyAsync(rowv.data(), A->d_val.p + q0, rowv.size() * sizeof(double2), hipMemcpyDeviceToHost,
                             stream()));
    DFM_HIP(hipMemcpyAsync(&d, A->d_dia.p + i, sizeof(double2), hipMemcpyDeviceToHost, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
    for (unsigned je = 0; je < nblkel; je++) {
      const double2 x = em[(size_t)ie * nblkel + je];
      if (ie == je) {
        d.x = d.x + x.x;
        d.y = d.y + x.y;
      } else {
        const unsigned j = blkel[je];
        for (unsigned q = q0; q < q1; q++)
          if (P->col[q] == j) {
            rowv[q - q0].x = rowv[q - q0].x + x.x;
            rowv[q - q0].y = rowv[q - q0].y + x.y;
            break;
          }
      }
    }
    if (q1 > q0)
      DFM_HIP(hipMemcpyAsync(A->d_val.p + q0, rowv.data(), rowv.size() * sizeof(double2), hipMemcpyHostToDevice,
                             stream()));
    DFM_HIP(hipMemcpyAsync(A->d_dia.p + i, &d, sizeof(double2), hipMemcpyHostToDevice, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
  }
  A->version = next_version();
  return DFM_OK;
}
int dfm_zbcrs_set_bc(dfm_zbcrs* A, const int* flag) {
  DFM_CHECK(A && flag, DFM_BAD_ARG, "bad arguments to dfm_zbcrs_set_bc");
  const dfm_bcrs_s* P = A->pat;
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, P->nb));
  DFM_ZLAUNCH(k_zbc_rows, zgrid_for(P->nb), kNT, (int)P->nb, P->d_rowptr.p, A->d_val.p, A->d_dia.p, f.p);
  const long ncrs = P->rowptr[P->nb];
  if (ncrs) DFM_ZLAUNCH(k_zbc_cols, zgrid_for(ncrs), kNT, ncrs, P->d_col.p, A->d_val.p, f.p);
  DFM_HIP(hipStreamSynchronize(stream()));
  A->version = next_version();
  return DFM_OK;
}
int dfm_zbcrs_matvec(dfm_zbcrs* A, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(A && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_zbcrs_matvec");
  const dfm_bcrs_s* P = A->pat;
  DFM_ZLAUNCH(k_zcrs_matvec, zgrid_for(P->nb), kNT, (int)P->nb, P->d_rowptr.p, P->d_col.p, A->d_val.p, A->d_dia.p,
              alpha, (const double2*)x, beta, (double2*)y);
  return sync_stream();
}

static int ztranspose(dfm_zbcrs* A) {
  const dfm_bcrs_s* P = A->pat;
  if (A->t_hash == P->pattern_hash && A->t_ptr.p) return DFM_OK;
  const unsigned nb = P->nb;
  std::vector<int> ptr(nb + 1, 0), row(std::max<size_t>(1, P->col.size())), src(row.size());
  for (unsigned q = 0; q < P->col.size(); q++) ptr[P->col[q] + 1]++;
  for (unsigned i = 0; i < nb; i++) ptr[i + 1] += ptr[i];
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (unsigned j = 0; j < nb; j++)  // rows in ascending order: each column's list ascends
    for (unsigned q = P->rowptr[j]; q < P->rowptr[j + 1]; q++) {
      const int k = fill[P->col[q]]++;
      row[k] = (int)j;
      src[k] = (int)q;
    }
  DFM_TRY(A->t_ptr.upload(ptr));
  DFM_TRY(A->t_row.upload(row));
  DFM_TRY(A->t_src.upload(src));
  A->t_hash = P->pattern_hash;
  return DFM_OK;
}
int dfm_zbcrs_matvec_hermitian(dfm_zbcrs* A, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(A && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_zbcrs_matvec_hermitian");
  DFM_TRY(ensure_init());
  DFM_TRY(ztranspose(A));
  const dfm_bcrs_s* P = A->pat;
  DFM_ZLAUNCH(k_zcrs_matvec_h, zgrid_for(P->nb), kNT, (int)P->nb, A->t_ptr.p, A->t_row.p, A->t_src.p, A->d_val.p,
              A->d_dia.p, alpha, (const double2*)x, beta, (double2*)y);
  return sync_stream();
}

// CZVector_Blk
static int zreduce(int nd, size_t n, const double2* x, const double2* y, double* out) {
  static DevBuf<double> partials;
  static DevBuf<ZState> st;
  static ZState* h = nullptr;
  DFM_TRY(ensure_init());
  const int g = zgrid_vec((long)n);
  if (partials.n < (size_t)(2 * g)) DFM_TRY(partials.alloc(2 * g));
  if (!st.n) DFM_TRY(st.alloc(1));
  if (!h) DFM_HIP(hipHostMalloc((void**)&h, sizeof(ZState), hipHostMallocDefault));
  if (nd == 2) {
    DFM_ZLAUNCH(k_zdot, g, kNT, (long)n, x, y, partials.p, (const ZState*)nullptr);
    DFM_ZLAUNCH((k_zfinalize<2, ZFIN_STORE>), 1, kFinNT, partials.p, g, st.p);
  } else if (nd == 3) {  // conjugated: sum conj(x_i) y_i
    DFM_ZLAUNCH(k_zdotc, g, kNT, (long)n, x, y, partials.p);
    DFM_ZLAUNCH((k_zfinalize<2, ZFIN_STORE>), 1, kFinNT, partials.p, g, st.p);
  } else {
    DFM_ZLAUNCH(k_znorm2, g, kNT, (long)n, x, partials.p);
    DFM_ZLAUNCH((k_zfinalize<1, ZFIN_STORE>), 1, kFinNT, partials.p, g, st.p);
  }
  DFM_HIP(hipMemcpyAsync(h, st.p, sizeof(ZState), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  out[0] = h->d0;
  if (nd >= 2) out[1] = h->d1;
  return DFM_OK;
}
int dfm_zvec_inner(size_t n, const double* x, const double* y, double* result) {
  DFM_CHECK(x && y && result, DFM_BAD_ARG, "bad arguments to dfm_zvec_inner");
  return zreduce(3, n, (const double2*)x, (const double2*)y, result);
}
int dfm_zvec_dot(size_t n, const double* x, const double* y, double* result) {
  DFM_CHECK(x && y && result, DFM_BAD_ARG, "bad arguments to dfm_zvec_dot");
  return zreduce(2, n, (const double2*)x, (const double2*)y, result);
}
int dfm_zvec_norm2(size_t n, const double* x, double* result) {
  DFM_CHECK(x && result, DFM_BAD_ARG, "bad arguments to dfm_zvec_norm2");
  return zreduce(1, n, (const double2*)x, nullptr, result);
}
int dfm_zvec_zero_bc(size_t n, const int* flag, double* x) {
  DFM_CHECK(flag && x, DFM_BAD_ARG, "bad arguments to dfm_zvec_zero_bc");
  DFM_TRY(ensure_init());
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, n));
  DFM_ZLAUNCH(k_zzero_bc, zgrid_vec((long)n), kNT, (long)n, f.p, (double2*)x);
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Element assembly
int dfm_zasm_create(dfm_zbcrs* A, unsigned nnode, unsigned nelem, unsigned npe, const unsigned* conn, dfm_asm** out) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  return dfm_asm_create(A->pat, nnode, nelem, npe, conn, out);
}
int dfm_zassemble_helmholtz2d_p1(dfm_zbcrs* A, dfm_asm* p, const double* coord, double wave_length,
                                 const double* value, double* res) {
  DFM_CHECK(A && p && p->A == A->pat && p->npe == 3 && coord && value, DFM_BAD_ARG,
            "dfm_zassemble_helmholtz2d_p1 needs a 3-node plan made on this matrix");
  DFM_CHECK(wave_length > 0.0, DFM_BAD_ARG, "wave_length must be positive");
  const dfm_bcrs_s* P = A->pat;
  DFM_ZLAUNCH(k_zasm_helmholtz2d, zgrid_for(p->nnode), kNT, (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p,
              p->d_elsup.p, P->d_rowptr.p, P->d_col.p, A->d_val.p, A->d_dia.p, coord, wave_length,
              (const double2*)value, (double2*)res);
  DFM_HIP(hipStreamSynchronize(stream()));
  A->version = next_version();
  return DFM_OK;
}
int dfm_zassemble_sommerfeld_b1(dfm_zbcrs* A, dfm_asm* p, const double* coord, double wave_length,
                                const double* value, double* res) {
  DFM_CHECK(A && p && p->A == A->pat && p->npe == 2 && coord && value, DFM_BAD_ARG,
            "dfm_zassemble_sommerfeld_b1 needs a 2-node plan made on this matrix");
  DFM_CHECK(wave_length > 0.0, DFM_BAD_ARG, "wave_length must be positive");
  const dfm_bcrs_s* P = A->pat;
  DFM_ZLAUNCH(k_zasm_sommerfeld, zgrid_for(p->nnode), kNT, (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p,
              p->d_elsup.p, P->d_rowptr.p, P->d_col.p, A->d_val.p, A->d_dia.p, coord, wave_length,
              (const double2*)value, (double2*)res);
  DFM_HIP(hipStreamSynchronize(stream()));
  A->version = next_version();
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// MatVec::CZMatDiaFrac_BlkCrs
int dfm_zilu_create(const dfm_zbcrs* A, int lev_fill, int ordering, const int* order, dfm_zilu** out) {
  DFM_CHECK(A && out, DFM_BAD_ARG, "bad arguments to dfm_zilu_create");
  std::unique_ptr<dfm_zilu_s> Z(new dfm_zilu_s);
  DFM_TRY(dfm_ilu_create(A->pat, lev_fill, ordering, order, 0, &Z->M));
  const dfm_ilu_s* M = Z->M;
  const size_t nb = M->nb, nf = M->fcol.size();
  DFM_TRY(Z->d_fval.alloc(std::max<size_t>(1, nf)));
  DFM_TRY(Z->d_fdia.alloc(std::max<size_t>(1, nb)));
  DFM_TRY(Z->zL.alloc(std::max<size_t>(1, (size_t)M->L64.nk * kSlice)));
  DFM_TRY(Z->zLd.alloc(std::max<size_t>(1, (size_t)M->L64.nslice * kSlice)));
  DFM_TRY(Z->zU.alloc(std::max<size_t>(1, (size_t)M->U64.nk * kSlice)));
  DFM_TRY(Z->zA.alloc(std::max<size_t>(1, (size_t)M->A64.nk * kSlice)));
  DFM_TRY(Z->zAd.alloc(std::max<size_t>(1, (size_t)M->A64.nslice * kSlice)));
  *out = Z.release();
  return DFM_OK;
}
int dfm_zilu_destroy(dfm_zilu* Z) {
  delete Z;
  return DFM_OK;
}
int dfm_zilu_info(const dfm_zilu* Z, unsigned* ncrs, int* nlevel_lower, int* nlevel_upper, int* ncolour) {
  DFM_CHECK(Z, DFM_BAD_ARG, "NULL factor");
  return dfm_ilu_info(Z->M, ncrs, nlevel_lower, nlevel_upper, ncolour);
}
int dfm_zilu_get_pattern(const dfm_zilu* Z, unsigned* rowptr, unsigned* col, int* lev, unsigned* diaind, int* order) {
  DFM_CHECK(Z, DFM_BAD_ARG, "NULL factor");
  return dfm_ilu_get_pattern(Z->M, rowptr, col, lev, diaind, order);
}
int dfm_zilu_set_value(dfm_zilu* Z, dfm_zbcrs* A, int* nsmall) {
  DFM_CHECK(Z && A && A->pat->nb == Z->M->nb, DFM_BAD_ARG, "factor and matrix do not match");
  dfm_ilu_s* M = Z->M;
  DFM_CHECK(A->pat->pattern_hash == M->pattern_hash, DFM_BAD_ARG,
            "matrix pattern differs from the one the factor was built on");
  const int nb = (int)M->nb;
  const long nf = (long)M->fcol.size();
  if (nf) DFM_ZLAUNCH(k_zilu_copy, zgrid_for(nf), kNT, nf, M->d_fsrc.p, A->d_val.p, Z->d_fval.p);
  DFM_ZLAUNCH(k_zilu_copy_dia, zgrid_for(nb), kNT, nb, M->permuted ? M->d_order.p : nullptr, A->d_dia.p, Z->d_fdia.p);
  DFM_HIP(hipMemsetAsync(M->d_nsing.p, 0, sizeof(int), stream()));
  for (int l = 0; l < M->nlev_lower; l++) {
    const int off = M->lev_ptr_L[l], nrow = M->lev_ptr_L[l + 1] - off;
    DFM_ZLAUNCH(k_zilu_factor, zgrid_for(nrow), kNT, nrow, M->d_rows_L.p + off, M->d_frow.p, M->d_fcol.p,
                M->d_fdiaind.p, Z->d_fval.p, Z->d_fdia.p, M->d_nsing.p);
  }
  {
    const long nL = M->L64.nk * kSlice, nU = M->U64.nk * kSlice, nl = (long)M->L64.nslice * kSlice;
    if (nL) DFM_ZLAUNCH(k_zsell_gather, zgrid_for(nL), kNT, nL, M->d_srcL.p, Z->d_fval.p, Z->zL.p);
    if (nU) DFM_ZLAUNCH(k_zsell_gather, zgrid_for(nU), kNT, nU, M->d_srcU.p, Z->d_fval.p, Z->zU.p);
    if (nl) DFM_ZLAUNCH(k_zsell_dia, zgrid_for(nl), kNT, nl, M->L64.rows.p, (const int*)nullptr, Z->d_fdia.p, Z->zLd.p);
  }
  int ns = 0;
  DFM_HIP(hipMemcpyAsync(&ns, M->d_nsing.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  if (nsmall) *nsmall = ns;
  Z->factored = true;
  return DFM_OK;
}
int dfm_zilu_get_values(const dfm_zilu* Z, double* val, double* dia) {
  DFM_CHECK(Z, DFM_BAD_ARG, "NULL factor");
  const size_t nf = Z->M->fcol.size(), nb = Z->M->nb;
  if (val && nf) DFM_HIP(hipMemcpyAsync(val, Z->d_fval.p, nf * sizeof(double2), hipMemcpyDeviceToHost, stream()));
  if (dia && nb) DFM_HIP(hipMemcpyAsync(dia, Z->d_fdia.p, nb * sizeof(double2), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_zilu_forward(dfm_zilu* Z, double* vec) {
  DFM_CHECK(Z && vec && Z->factored, DFM_BAD_ARG, "factor has no values (call dfm_zilu_set_value)");
  DFM_TRY(zsweep(Z, (const double2*)vec, (double2*)vec, nullptr, true, false));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_zilu_backward(dfm_zilu* Z, double* vec) {
  DFM_CHECK(Z && vec && Z->factored, DFM_BAD_ARG, "factor has no values (call dfm_zilu_set_value)");
  DFM_TRY(zsweep(Z, (const double2*)vec, (double2*)vec, nullptr, false, true));
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}
int dfm_zilu_solve_precond(dfm_zilu* Z, double* vec) {
  DFM_CHECK(Z && vec && Z->factored, DFM_BAD_ARG, "factor has no values (call dfm_zilu_set_value)");
  dfm_ilu_s* M = Z->M;
  const long n = M->nb;
  double2* v = (double2*)vec;
  if (!M->permuted) {
    DFM_TRY(zsweep(Z, v, v, nullptr, true, true));
  } else {
    DevBuf<double2> tmp;
    DFM_TRY(tmp.alloc(std::max<long>(1, n)));
    DFM_ZLAUNCH(k_zgather, zgrid_vec(n), kNT, n, M->d_order.p, (const double2*)v, tmp.p);
    DFM_TRY(zsweep(Z, tmp.p, tmp.p, nullptr, true, true));
    DFM_ZLAUNCH(k_zgather, zgrid_vec(n), kNT, n, M->d_inv.p, (const double2*)tmp.p, v);
  }
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

// Synchronises the stream when it goes out of scope: declared after a local
// workspace of a loop that defers synchronisation, it keeps the workspace
// alive until the kernels queued on it have run.
struct StreamDrain {
  ~StreamDrain() { (void)hipStreamSynchronize(stream()); }
};

// ---------------------------------------------------------------------------
// MatVec::Sol complex solvers.  r: right-hand side in, final residual out;
// u: solution out.  The preconditioned ones run in the factor's numbering
// with its permuted copy of A.  All return DFM_OK, as the reference returns
// true also when max_iter stops them.
static int zsolve_checks(dfm_zbcrs* A, dfm_zilu* Z, double* r, double* u, double* conv, unsigned* iter) {
  DFM_CHECK(A && r && u && conv && iter && r != u, DFM_BAD_ARG, "bad arguments to a complex solver");
  if (Z) {
    DFM_CHECK(Z->factored, DFM_BAD_ARG, "factor has no values (call dfm_zilu_set_value)");
    DFM_CHECK(A->pat->nb == Z->M->nb && A->pat->pattern_hash == Z->M->pattern_hash, DFM_BAD_ARG,
              "matrix pattern differs from the one the factor was built on");
  }
  return DFM_OK;
}
int dfm_zsolve_pcocg(dfm_zbcrs* A, dfm_zilu* Z, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(Z, DFM_BAD_ARG, "dfm_zsolve_pcocg needs a factor");
  DFM_TRY(zsolve_checks(A, Z, r, u, conv_ratio, iteration));
  DFM_TRY(zrefresh_a(Z, A));
  const dfm_ilu_s* M = Z->M;
  return zcocg_core(Z->ctx, zfactor_op(Z), Z, M->nb, M->permuted ? M->d_order.p : nullptr,
                    M->permuted ? M->d_inv.p : nullptr, (double2*)r, (double2*)u, conv_ratio, iteration);
}
int dfm_zsolve_cocg(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_TRY(zsolve_checks(A, nullptr, r, u, conv_ratio, iteration));
  ZOp op;
  DFM_TRY(znatural_op(A, op));
  return zcocg_core(A->ctx, op, nullptr, A->pat->nb, nullptr, nullptr, (double2*)r, (double2*)u, conv_ratio,
                    iteration);
}
// MatVec::Sol::Solve_CG / Solve_PCG for a Hermitian matrix
// (zsolver_mat_iter.cpp:32-119, 121-218): CG with the real part of the
// conjugated inner product; host-driven loops as Solve_CGNR.  Solve_CG stops
// when |r|^2/|r0|^2 < tol^2 and counts from 1; Solve_PCG applies the factor
// (SolvePrecond), counts from 0 and reports DFM_NOT_CONVERGED at max_iter
// (the reference's false) with conv_ratio and iteration as given.
static int zcg_herm(dfm_zbcrs* A, dfm_zilu* Z, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_TRY(zsolve_checks(A, Z, r, u, conv_ratio, iteration));
  DFM_TRY(ensure_init());
  const size_t nb = A->pat->nb, n2 = 2 * nb;
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  DFM_TRY(dfm_vec_fill(n2, 0.0, u));
  double sq0 = 0.0;
  DFM_TRY(dfm_zvec_norm2(nb, r, &sq0));
  if (sq0 < 1.0e-30) {
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  const double inv = 1.0 / sq0;
  DevBuf<double> w;
  StreamDrain drain_w;  // destroyed before w: queued kernels finish before w is freed
  DFM_TRY(w.alloc(3 * std::max<size_t>(1, n2)));
  double *p = w.p, *Ap = w.p + n2, *z = w.p + 2 * n2;
  DFM_TRY(dfm_vec_fill(3 * n2, 0.0, w.p));
  double c[2];
  if (!Z) {
    double sq_res0 = sq0;
    DFM_TRY(dfm_vec_copy(n2, r, p));
    *iteration = mx_iter;
    for (unsigned it = 1; it < mx_iter; it++) {
      DFM_TRY(dfm_zbcrs_matvec(A, 1.0, p, 0.0, Ap));
      DFM_TRY(dfm_zvec_inner(nb, p, Ap, c));
      const double alpha = sq_res0 / c[0];
      DFM_TRY(dfm_vec_axpy(n2, alpha, p, u));
      DFM_TRY(dfm_vec_axpy(n2, -alpha, Ap, r));
      double sq1 = 0.0;
      DFM_TRY(dfm_zvec_norm2(nb, r, &sq1));
      if (sq1 * inv < tol * tol) {
        *conv_ratio = std::sqrt(sq1 * inv);
        *iteration = it;
        return DFM_OK;
      }
      const double beta = sq1 / sq_res0;
      sq_res0 = sq1;
      DFM_TRY(dfm_vec_scal(n2, beta, p));
      DFM_TRY(dfm_vec_axpy(n2, 1.0, r, p));
    }
    return DFM_OK;
  }
  DFM_TRY(dfm_vec_copy(n2, r, z));
  DFM_TRY(dfm_zilu_solve_precond(Z, z));
  DFM_TRY(dfm_vec_copy(n2, z, p));
  DFM_TRY(dfm_zvec_inner(nb, r, z, c));
  double rz0 = c[0];
  for (unsigned it = 0; it < mx_iter; it++) {
    DFM_TRY(dfm_zbcrs_matvec(A, 1.0, p, 0.0, Ap));
    DFM_TRY(dfm_zvec_inner(nb, p, Ap, c));
    const double alpha = rz0 / c[0];
    DFM_TRY(dfm_vec_axpy(n2, alpha, p, u));
    DFM_TRY(dfm_vec_axpy(n2, -alpha, Ap, r));
    double sq = 0.0;
    DFM_TRY(dfm_zvec_norm2(nb, r, &sq));
    if (sq * inv < tol * tol) {
      *conv_ratio = std::sqrt(sq * inv);
      *iteration = it;
      return DFM_OK;
    }
    DFM_TRY(dfm_vec_copy(n2, r, z));
    DFM_TRY(dfm_zilu_solve_precond(Z, z));
    DFM_TRY(dfm_zvec_inner(nb, r, z, c));
    const double beta = c[0] / rz0;
    rz0 = c[0];
    DFM_TRY(dfm_vec_scal(n2, beta, p));
    DFM_TRY(dfm_vec_axpy(n2, 1.0, z, p));
  }
  return DFM_NOT_CONVERGED;
}
// a host loop with deferred synchronisation, synchronised once at its end
template <typename F>
static int deferred(F&& body) {
  int rc;
  {
    SyncDefer defer;
    rc = body();
  }
  if (hipStreamSynchronize(stream()) != hipSuccess && rc >= 0) rc = DFM_HIP_ERROR;
  return rc;
}
int dfm_zsolve_cg(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  return deferred([&] { return zcg_herm(A, nullptr, r, u, conv_ratio, iteration); });
}
int dfm_zsolve_pcg(dfm_zbcrs* A, dfm_zilu* Z, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(Z, DFM_BAD_ARG, "dfm_zsolve_pcg needs a factor");
  return deferred([&] { return zcg_herm(A, Z, r, u, conv_ratio, iteration); });
}
// MatVec::Sol::Solve_CGNR (zsolver_mat_iter.cpp:394-483): CG on the normal
// equations with z = A^H r; alpha and beta are ratios of squared norms.  A
// host-driven loop over the matrix-vector kernels (squared norms are device
// reductions); the vector updates are AXPY on the interleaved doubles, which
// is the reference's complex AXPY with a real factor, component by component.
static int zcgnr(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration);
int dfm_zsolve_cgnr(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  return deferred([&] { return zcgnr(A, r, u, conv_ratio, iteration); });
}
static int zcgnr(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_TRY(zsolve_checks(A, nullptr, r, u, conv_ratio, iteration));
  DFM_TRY(ensure_init());
  const size_t nb = A->pat->nb, n2 = 2 * nb;
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  DFM_TRY(dfm_vec_fill(n2, 0.0, u));
  double sq0 = 0.0;
  DFM_TRY(dfm_zvec_norm2(nb, r, &sq0));
  if (sq0 < 1.0e-30) {
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  const double inv = 1.0 / sq0;
  DevBuf<double> w;
  StreamDrain drain_w;  // destroyed before w: queued kernels finish before w is freed
  DFM_TRY(w.alloc(3 * std::max<size_t>(1, n2)));
  double *p = w.p, *z = w.p + n2, *Ap = w.p + 2 * n2;
  DFM_TRY(dfm_vec_fill(3 * n2, 0.0, w.p));
  DFM_TRY(dfm_zbcrs_matvec_hermitian(A, 1.0, r, 0.0, z));
  double sqz0 = 0.0;
  DFM_TRY(dfm_zvec_norm2(nb, z, &sqz0));
  DFM_TRY(dfm_vec_copy(n2, z, p));
  *iteration = mx_iter;
  for (unsigned it = 1; it < mx_iter; it++) {
    DFM_TRY(dfm_zbcrs_matvec(A, 1.0, p, 0.0, Ap));
    double sqAp = 0.0;
    DFM_TRY(dfm_zvec_norm2(nb, Ap, &sqAp));
    const double alpha = sqz0 / sqAp;
    DFM_TRY(dfm_vec_axpy(n2, alpha, p, u));
    DFM_TRY(dfm_vec_axpy(n2, -alpha, Ap, r));
    double sqr = 0.0;
    DFM_TRY(dfm_zvec_norm2(nb, r, &sqr));
    const double ratio = std::sqrt(sqr * inv);
    if (ratio < tol) {
      *conv_ratio = ratio;
      *iteration = it;
      return DFM_OK;
    }
    DFM_TRY(dfm_zbcrs_matvec_hermitian(A, 1.0, r, 0.0, z));
    double sqz1 = 0.0;
    DFM_TRY(dfm_zvec_norm2(nb, z, &sqz1));
    const double beta = sqz1 / sqz0;
    sqz0 = sqz1;
    DFM_TRY(dfm_vec_scal(n2, beta, p));
    DFM_TRY(dfm_vec_axpy(n2, 1.0, z, p));
  }
  return DFM_OK;
}
int dfm_zsolve_bicgstab(dfm_zbcrs* A, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_TRY(zsolve_checks(A, nullptr, r, u, conv_ratio, iteration));
  ZOp op;
  DFM_TRY(znatural_op(A, op));
  return zbicg_core(A->ctx, op, nullptr, A->pat->nb, nullptr, nullptr, (double2*)r, (double2*)u, conv_ratio,
                    iteration);
}
int dfm_zsolve_bicgstabp(dfm_zbcrs* A, dfm_zilu* Z, double* r, double* u, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(Z, DFM_BAD_ARG, "dfm_zsolve_bicgstabp needs a factor");
  DFM_TRY(zsolve_checks(A, Z, r, u, conv_ratio, iteration));
  DFM_TRY(zrefresh_a(Z, A));
  const dfm_ilu_s* M = Z->M;
  return zbicg_core(Z->ctx, zfactor_op(Z), Z, M->nb, M->permuted ? M->d_order.p : nullptr,
                    M->permuted ? M->d_inv.p : nullptr, (double2*)r, (double2*)u, conv_ratio, iteration);
}
