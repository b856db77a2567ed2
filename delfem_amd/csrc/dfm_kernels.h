// dfm_kernels.h -- gfx950 device kernels of delfem_amd.
//
// All floating-point expressions are written in the evaluation order of the
// DelFEM reference routine they replace and the library is compiled with
// -ffp-contract=off, so the FP64 kernels reproduce the reference bit for bit
// when they see the same operands in the same order (natural ordering).
// The kernels are HBM-bandwidth bound (0.1-0.25 flop/byte): the layouts keep
// every wavefront's loads contiguous (sliced ELL with values stored
// [slot][entry][lane]) and reductions are deterministic (fixed tree, last
// workgroup sums the per-workgroup partials in index order).
#pragma once

#include <type_traits>

#include "dfm_internal.h"

namespace dfm {

constexpr int kNT = 256;             // threads per workgroup
constexpr int kWaves = kNT / 64;     // wavefronts (= slices) per workgroup
constexpr int kSpmvNT = 256;         // threads per workgroup of the solver SpMV
constexpr int kSpmvWaves = kSpmvNT / 64;
// Triangular sweeps launch one small workgroup per slice: a colour of a
// multicolour ordering holds only a few hundred slices, and one-wave
// workgroups spread them over all CUs.
constexpr int kTrsvWaves = 1;
constexpr int kTrsvNT = 64 * kTrsvWaves;

// ---------------------------------------------------------------------------
// Deterministic two-phase reductions.  Every workgroup writes its partial sums
// with plain stores; a one-workgroup finalize kernel launched next adds them in
// index order and applies the Krylov epilogue.  The kernel boundary makes the
// partials visible, so no device-scope fence (an L2 write-back on a
// multi-XCD part) is needed per workgroup.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over a workgroup of NT threads, valid in thread 0.
template <int NT>
__device__ inline double block_sum_n(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; i++) s += sh[i];
  __syncthreads();
  return s;
}

// partials[d*nparts + part] = workgroup sum of v[d]
template <int ND, int NT = kNT>
__device__ inline void write_partials(const double (&v)[ND], double* partials, int part, int nparts) {
  __shared__ double sh[NT / 64];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const double s = block_sum_n<NT>(v[d], sh);
    if (threadIdx.x == 0) partials[(size_t)d * nparts + part] = s;
  }
}

// Krylov epilogues applied by the finalize kernel to the reduced sums s[0..2].
enum {
  FIN_STORE = 0,      // d0,d1,d2 = s            (always runs)
  FIN_PCG_PAP = 1,    // alpha = rz / <p,Ap>      (solver_ls_iter.cpp:135-139)
  FIN_PCG_RR = 2,     // convergence test         (:144-154)
  FIN_PCG_RZ0 = 3,    // rz = <r,z>               (:131)
  FIN_PCG_RZ = 4,     // beta = <r,z>_new / rz    (:157-161)
  FIN_BICG_ALPHA = 5, // r_r2, alpha              (:334-347)
  FIN_BICG_OMEGA = 6, // omega = <s,AMs>/<AMs,AMs> (:366-372)
  FIN_BICG_XR = 7     // convergence test, beta   (:383-398)
};

constexpr int kFinNT = 1024;

// This thread's share of sum(p[0..n)): p[t] + p[t + NT] + p[t + 2 NT] + ...
// added in that order, with U loads issued before their adds so that a long
// partial array costs a few memory round trips rather than one per element.
template <int NT, int U = 8>
__device__ inline double strided_sum(const double* __restrict__ p, int n) {
  double acc = 0.0;
  for (int base = (int)threadIdx.x; base < n; base += U * NT) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = base + u * NT;
      v[u] = i < n ? p[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc += v[u];
  }
  return acc;
}

// strided_sum of ND arrays p[d][0..n) at once: every array's loads of a
// round are issued before any is added (one memory round trip per round for
// all of them, not one per array); each array's sum is formed in the same
// order as strided_sum's.
template <int NT, int ND, int U = 8>
__device__ inline void strided_sums(const double* const (&p)[ND], const int (&n)[ND], double (&acc)[ND]) {
  int nmax = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    acc[d] = 0.0;
    nmax = n[d] > nmax ? n[d] : nmax;
  }
  for (int base = (int)threadIdx.x; base < nmax; base += U * NT) {
    double v[ND][U];
#pragma unroll
    for (int d = 0; d < ND; d++)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = base + u * NT;
        v[d][u] = i < n[d] ? p[d][i] : 0.0;
      }
#pragma unroll
    for (int d = 0; d < ND; d++)
#pragma unroll
      for (int u = 0; u < U; u++) acc[d] += v[d][u];
  }
}

template <int ND, int FIN>
__global__ void __launch_bounds__(kFinNT) k_finalize(const double* __restrict__ partials, int nparts,
                                                    KState* __restrict__ st) {
  if (FIN != FIN_STORE && st->done) return;
  __shared__ double sh[kFinNT / 64];
  double s[3] = {0.0, 0.0, 0.0};
  {
    const double* p[ND];
    int n[ND];
    double t[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      p[d] = partials + (size_t)d * nparts;
      n[d] = nparts;
    }
    strided_sums<kFinNT, ND>(p, n, t);
#pragma unroll
    for (int d = 0; d < ND; d++) s[d] = block_sum_n<kFinNT>(t[d], sh);
  }
  if (threadIdx.x != 0) return;
  if constexpr (FIN == FIN_STORE) {
    st->d0 = s[0];
    st->d1 = s[1];
    st->d2 = s[2];
  } else if constexpr (FIN == FIN_PCG_PAP) {
    st->pAp = s[0];
    st->alpha = st->rz / s[0];
  } else if constexpr (FIN == FIN_PCG_RR || FIN == FIN_BICG_XR) {
    st->rr = s[0];
    if (s[0] * st->inv_rr0 < st->tol2) {
      st->done = 1;
      st->conv_iter = st->iter;
      st->conv_rr = s[0];
    } else if constexpr (FIN == FIN_BICG_XR) {
      st->beta = s[1] * st->alpha / (st->r_r2 * st->omega);
    }
    st->iter += 1;
    if (!st->done && st->iter >= st->maxit) st->done = 2;
  } else if constexpr (FIN == FIN_PCG_RZ0) {
    st->rz = s[0];
  } else if constexpr (FIN == FIN_PCG_RZ) {
    st->rz_new = s[0];
    st->beta = s[0] / st->rz;
    st->rz = s[0];
  } else if constexpr (FIN == FIN_BICG_ALPHA) {
    st->r_r2 = s[2];
    st->alpha = s[2] / s[0];
  } else if constexpr (FIN == FIN_BICG_OMEGA) {
    st->omega = s[0] / s[1];
  }
}
// ---------------------------------------------------------------------------
// Storage types of the sliced-ELL copies: double, float, or bfloat16 held as
// its 16 raw bits (bf16s).  Values are converted to the arithmetic type of the
// vectors when loaded.
template <typename TV>
__device__ inline TV cvt(double v) { return (TV)v; }
template <typename TV>
__device__ inline TV cvt(float v) { return (TV)v; }
template <typename TV>
__device__ inline TV cvt(bf16s v) {
  return (TV)__uint_as_float((unsigned)v.bits << 16);
}
// double -> bf16 with round-to-nearest-even through float
__device__ inline bf16s to_bf16(double d) {
  const unsigned u = __float_as_uint((float)d);
  bf16s r;
  r.bits = (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  return r;
}
template <typename TS>
__device__ inline TS store_as(double v) { return (TS)v; }
template <>
__device__ inline bf16s store_as<bf16s>(double v) { return to_bf16(v); }

// Streamed (read-once) operands of the sliced-ELL kernels are loaded with the
// non-temporal hint so the gathered vectors keep the caches.
template <typename T>
__device__ inline T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ inline bf16s ld_stream<bf16s>(const bf16s* p) {
  bf16s r;
  r.bits = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
  return r;
}

// Slot layout of a block in a sliced-ELL array.  Each lane's block of one
// slot is E 32-bit words (float: bb words, double: 2 bb, bf16: entries paired
// into ceil(bb/2) words).  A slot stores the first 4 floor(E/4) words of all
// 64 lanes as 16-byte quads, quad-major ((q*64 + lane)*4 + j), then the
// remaining E mod 4 words of each lane contiguously (Q*256 + lane*R + r), so a
// lane fetches its block with 16-byte loads plus one 4- or 8-byte load and
// every load instruction of a wave reads one contiguous 1 KiB or 256/512 B
// run.  A slot occupies 256 E bytes, a multiple of 256.
template <typename TS>
__host__ __device__ constexpr int slot_words(int bb) {
  return std::is_same<TS, bf16s>::value ? (bb + 1) / 2 : bb * (int)sizeof(TS) / 4;
}
// TS elements allocated per lane per slot (bf16 pads an odd bb by one)
template <typename TS>
__host__ __device__ constexpr int slot_elems(int bb) {
  return slot_words<TS>(bb) * 4 / (int)sizeof(TS);
}
// Position (in TS elements) of entry e of the block of `lane` in slot k.
template <typename TS>
__host__ __device__ inline size_t slot_index(size_t k, int bb, int e, int lane) {
  const int E = slot_words<TS>(bb), Q = E / 4, R = E % 4;
  const bool half = std::is_same<TS, bf16s>::value;
  const int w = half ? (e >> 1) : e * (int)sizeof(TS) / 4;  // first word of the entry
  size_t word = k * 64 * (size_t)E;
  if (w < 4 * Q) word += ((size_t)(w >> 2) * 64 + lane) * 4 + (w & 3);
  else word += (size_t)Q * 256 + (size_t)lane * R + (w - 4 * Q);
  if (half) return word * 2 + (e & 1);
  return word * 4 / sizeof(TS);
}
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
// The E raw 32-bit words of one lane's block in slot k (non-temporal loads:
// the matrix streams once per use, the gathered vectors keep the caches).
template <int BB, typename TS>
__device__ inline void load_words(const TS* __restrict__ val, size_t k, int lane,
                                  unsigned (&w)[slot_words<TS>(BB)]) {
  constexpr int E = slot_words<TS>(BB), Q = E / 4, R = E % 4;
  const unsigned* base = reinterpret_cast<const unsigned*>(val) + k * 64 * E;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(base + ((size_t)q * 64 + lane) * 4));
    w[4 * q] = v.x;
    w[4 * q + 1] = v.y;
    w[4 * q + 2] = v.z;
    w[4 * q + 3] = v.w;
  }
  const unsigned* rem = base + Q * 256 + (size_t)lane * R;
  if constexpr (R == 2) {
    const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(rem));
    w[4 * Q] = v.x;
    w[4 * Q + 1] = v.y;
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) w[4 * Q + r] = __builtin_nontemporal_load(rem + r);
  }
}
// ... and those words as BB values of type TV
template <int BB, typename TS, typename TV>
__device__ inline void words_to_block(const unsigned (&w)[slot_words<TS>(BB)], TV (&a)[BB]) {
#pragma unroll
  for (int e = 0; e < BB; e++) {
    if constexpr (std::is_same<TS, bf16s>::value) {
      const unsigned u = w[e >> 1];
      a[e] = (TV)__uint_as_float((e & 1) ? (u & 0xFFFF0000u) : (u << 16));
    } else if constexpr (std::is_same<TS, double>::value) {
      a[e] = (TV)__hiloint2double((int)w[2 * e + 1], (int)w[2 * e]);
    } else {
      a[e] = (TV)__uint_as_float(w[e]);
    }
  }
}
// All BB values of one block slot, converted to TV.
template <int BB, typename TS, typename TV>
__device__ inline void load_block(const TS* __restrict__ val, size_t k, int lane, TV (&a)[BB]) {
  unsigned w[slot_words<TS>(BB)];
  load_words<BB, TS>(val, k, lane, w);
  words_to_block<BB, TS, TV>(w, a);
}

// Workgroup index remapped so that the workgroups dispatched to one XCD
// (dispatch is round-robin over the 8 XCDs) cover one contiguous range of
// slices: neighbouring rows, which gather overlapping parts of x, then share
// that XCD's L2 instead of being spread over all eight.
__device__ inline int xcd_block() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G < 16) return b;
  const int x = b & 7, i = b >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + i;
}

// Holds three wave-uniform values that were loaded independently at one
// point of the program, so that the compiler neither sinks a load below a
// branch nor waits for each of them in turn: the loads leave together and
// the wave waits for them once.
__device__ __forceinline__ void issue_together(int a, int b, int c) {
  asm volatile("" ::"s"(a), "s"(b), "s"(c));
}
// The x block of column c (B entries), zeros for a padding slot (c < 0).  The
// loads are issued whatever c is (block 0 for padding, always in bounds), so
// that the B words become one multi-word load and no branch splits the wave;
// the zero is selected afterwards.
template <int B, typename T>
__device__ __forceinline__ void gather_x(const T* x, int c, T (&xv)[B]) {
  const size_t cb = (size_t)(c >= 0 ? c : 0) * B;
  T v[B];
#pragma unroll
  for (int e = 0; e < B; e++) v[e] = x[cb + e];
#pragma unroll
  for (int e = 0; e < B; e++) xv[e] = c >= 0 ? v[e] : (T)0;
}
// One block slot of a sliced-ELL row: values a[] (storage TS converted to TV),
// gathered x block xv[] (zeros for a padding slot, column < 0).
template <int B, typename TS, typename TV>
__device__ inline void sell_slot(const int* __restrict__ col, const TS* __restrict__ val, const TV* x, long k,
                                 int lane, TV (&a)[B * B], TV (&xv)[B]) {
  const int c = ld_stream(col + (size_t)k * 64 + lane);
  load_block<B * B, TS, TV>(val, (size_t)k, lane, a);
  gather_x<B, TV>(x, c, xv);
}
// ---------------------------------------------------------------------------
// Block row kernels in the reference's summation order.
// acc[r] += sum_c A[r*B+c]*x[c], the B-term sum formed first (matdia_blkcrs.cpp:1027-1058)
template <int B, typename T>
__device__ inline void blk_madd(T* acc, const T* a, const T* x) {
  if constexpr (B == 1) {
    acc[0] += a[0] * x[0];
  } else if constexpr (B == 2) {
    acc[0] += a[0] * x[0] + a[1] * x[1];
    acc[1] += a[2] * x[0] + a[3] * x[1];
  } else {
#pragma unroll
    for (int r = 0; r < B; r++) {
      T s = a[r * B] * x[0];
#pragma unroll
      for (int c = 1; c < B; c++) s = s + a[r * B + c] * x[c];
      acc[r] += s;
    }
  }
}
// acc[r] -= sum_c A[r*B+c]*x[c] (matdiafrac_blkcrs.cpp:447-449, 613-615)
template <int B, typename T>
__device__ inline void blk_msub(T* acc, const T* a, const T* x) {
  if constexpr (B == 1) {
    acc[0] -= a[0] * x[0];
  } else {
#pragma unroll
    for (int r = 0; r < B; r++) {
      T s = a[r * B] * x[0];
#pragma unroll
      for (int c = 1; c < B; c++) s = s + a[r * B + c] * x[c];
      acc[r] -= s;
    }
  }
}
// y[r] = sum_c D[r*B+c]*t[c] (matdiafrac_blkcrs.cpp:451-456)
template <int B, typename T>
__device__ inline void blk_mul(T* y, const T* d, const T* t) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    T s = d[r * B] * t[0];
#pragma unroll
    for (int c = 1; c < B; c++) s = s + d[r * B + c] * t[c];
    y[r] = s;
  }
}

// ---------------------------------------------------------------------------
// CRS MatVec on the canonical storage: y = beta*y + alpha*A*x
// (CMatDia_BlkCrs::MatVec, src/matvec/matdia_blkcrs.cpp:921-1094).
template <int B>
static __global__ void __launch_bounds__(kNT) k_crs_matvec(int nb, const int* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ dia, double alpha,
                                                   const double* __restrict__ x, double beta,
                                                   double* __restrict__ y) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  constexpr int BB = B * B;
  // entries taken CH at a time: their indices, blocks and x loaded before the
  // first is used, then added one by one in row order
  constexpr int CH = B == 3 ? 4 : 8;
  double acc[B];
#pragma unroll
  for (int r = 0; r < B; r++) acc[r] = y[(size_t)i * B + r] * beta;
  auto add = [&](const double* A, const double* xj) {
    if constexpr (B == 1) {
      acc[0] += alpha * A[0] * xj[0];
    } else {
#pragma unroll
      for (int r = 0; r < B; r++) {
        double s = A[r * B] * xj[0];
#pragma unroll
        for (int c = 1; c < B; c++) s = s + A[r * B + c] * xj[c];
        acc[r] += alpha * s;
      }
    }
  };
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  for (int qb = q0; qb < q1; qb += CH) {
    int c[CH];
    double a[CH][BB], xv[CH][B];
#pragma unroll
    for (int u = 0; u < CH; u++) c[u] = qb + u < q1 ? col[qb + u] : i;
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const size_t q = qb + u < q1 ? (size_t)(qb + u) : (size_t)q0;
#pragma unroll
      for (int e = 0; e < BB; e++) a[u][e] = val[q * BB + e];
#pragma unroll
      for (int e = 0; e < B; e++) xv[u][e] = x[(size_t)c[u] * B + e];
    }
#pragma unroll
    for (int u = 0; u < CH; u++)
      if (qb + u < q1) add(a[u], xv[u]);
  }
  {
    double d[BB], xi[B];
#pragma unroll
    for (int e = 0; e < BB; e++) d[e] = dia[(size_t)i * BB + e];
#pragma unroll
    for (int e = 0; e < B; e++) xi[e] = x[(size_t)i * B + e];
    add(d, xi);
  }
#pragma unroll
  for (int r = 0; r < B; r++) y[(size_t)i * B + r] = acc[r];
}

// MatVec over the sliced-ELL copy (MvSell), one lane per block row, entries
// in CRS order, eight (nine-value blocks: four) slots' columns, blocks and x
// loaded before the first is added.  Forms (the sums of the one-thread-per-
// row kernels, term for term): MV_SQ square blocks, diagonal block last
// (CMatDia_BlkCrs::MatVec); rectangular CMat_BlkCrs::MatVec forms MV_SCALAR
// (1x1, beta == 0 starts from 0), MV_ROW (one row dof: alpha (a . x)), MV_COL
// (one column dof: a_id (x alpha)) and MV_GEN (alpha a_{id,jd} x_jd each).
enum { MV_SQ, MV_SCALAR, MV_ROW, MV_COL, MV_GEN };
static __global__ void k_mv_gather(long n, int bs, const long long* __restrict__ src, const double* __restrict__ aval,
                                   double* __restrict__ val) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= n) return;
  const long kk = t / (bs * 64), rem = t % (bs * 64), e = rem / 64, l = rem % 64;
  const long long q = src[kk * 64 + l];
  val[t] = q >= 0 ? aval[q * bs + e] : 0.0;
}
template <int FORM, int BC, int BR>
static __global__ void __launch_bounds__(256) k_sell_mv(int nrow, const int* __restrict__ sptr,
                                                        const int* __restrict__ col, const double* __restrict__ val,
                                                        const double* __restrict__ dia, double alpha,
                                                        const double* __restrict__ x, double beta,
                                                        double* __restrict__ y) {
  constexpr int BS = BC * BR;
  constexpr int CH = BS >= 9 ? 4 : 8;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  const long i = (long)s * 64 + l;
  if (i >= nrow) return;
  const int k0 = sptr[s], k1 = sptr[s + 1];
  double acc[BC];
#pragma unroll
  for (int r = 0; r < BC; r++) {
    if constexpr (FORM == MV_SCALAR) acc[r] = beta != 0.0 ? y[i] * beta : 0.0;
    else acc[r] = y[i * BC + r] * beta;
  }
  auto add = [&](const double (&a)[BS], const double (&xj)[BR]) {
    if constexpr (FORM == MV_SQ && BC == 1) {
      acc[0] += alpha * a[0] * xj[0];
    } else if constexpr (FORM == MV_SQ) {
#pragma unroll
      for (int r = 0; r < BC; r++) {
        double t = a[r * BC] * xj[0];
#pragma unroll
        for (int c = 1; c < BC; c++) t = t + a[r * BC + c] * xj[c];
        acc[r] += alpha * t;
      }
    } else if constexpr (FORM == MV_SCALAR) {
      acc[0] += alpha * a[0] * xj[0];
    } else if constexpr (FORM == MV_ROW) {
      double t = a[0] * xj[0] + a[1] * xj[1];
      if constexpr (BR == 3) t = t + a[2] * xj[2];
      acc[0] += alpha * t;
    } else if constexpr (FORM == MV_COL) {
      const double ajx = xj[0] * alpha;
#pragma unroll
      for (int r = 0; r < BC; r++) acc[r] += a[r] * ajx;
    } else {
#pragma unroll
      for (int r = 0; r < BC; r++)
#pragma unroll
        for (int c = 0; c < BR; c++) acc[r] += alpha * a[r * BR + c] * xj[c];
    }
  };
  for (int kb = k0; kb < k1; kb += CH) {
    int c[CH];
    double a[CH][BS], xv[CH][BR];
#pragma unroll
    for (int u = 0; u < CH; u++) c[u] = kb + u < k1 ? col[(size_t)(kb + u) * 64 + l] : -1;
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const size_t k = kb + u < k1 ? (size_t)(kb + u) : (size_t)k0;
      const size_t cc = c[u] >= 0 ? (size_t)c[u] : 0;
#pragma unroll
      for (int e = 0; e < BS; e++) a[u][e] = val[(k * BS + e) * 64 + l];
#pragma unroll
      for (int e = 0; e < BR; e++) xv[u][e] = x[cc * BR + e];
    }
#pragma unroll
    for (int u = 0; u < CH; u++)
      if (c[u] >= 0) add(a[u], xv[u]);
  }
  if constexpr (FORM == MV_SQ) {
    double d[BS], xi[BR];
#pragma unroll
    for (int e = 0; e < BS; e++) d[e] = dia[i * BS + e];
#pragma unroll
    for (int e = 0; e < BR; e++) xi[e] = x[i * BR + e];
    add(d, xi);
  }
#pragma unroll
  for (int r = 0; r < BC; r++) {
    if constexpr (FORM == MV_SCALAR) y[i] = acc[0];
    else y[i * BC + r] = acc[r];
  }
}

// Generic block size (b > 3): each product added separately (:1072-1085).
static __global__ void __launch_bounds__(kNT) k_crs_matvec_gen(int nb, int b, const int* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ dia, double alpha,
                                                       const double* __restrict__ x, double beta,
                                                       double* __restrict__ y) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  const int bb = b * b;
  for (int r = 0; r < b; r++) y[(size_t)i * b + r] *= beta;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  for (int q = q0; q <= q1; q++) {
    const double* A = q < q1 ? val + (size_t)q * bb : dia + (size_t)i * bb;
    const double* xj = x + (size_t)(q < q1 ? col[q] : i) * b;
    for (int r = 0; r < b; r++) {
      double a = y[(size_t)i * b + r];
      for (int c = 0; c < b; c++) a += alpha * A[r * b + c] * xj[c];
      y[(size_t)i * b + r] = a;
    }
  }
}

// ---------------------------------------------------------------------------
// Dirichlet conditions (CMatDia_BlkCrs::SetBoundaryCondition, matdia_blkcrs.cpp:318-342)
static __global__ void k_bc_rows(int nb, int b, const int* __restrict__ rowptr, double* __restrict__ val,
                          double* __restrict__ dia, const int* __restrict__ flag) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nb) return;
  const int bb = b * b;
  for (int d = 0; d < b; d++) {
    if (flag[(size_t)i * b + d] == 0) continue;
    double* D = dia + (size_t)i * bb;
    for (int k = 0; k < b; k++) D[k * b + d] = 0.0;
    for (int k = 0; k < b; k++) D[d * b + k] = 0.0;
    D[d * b + d] = 1.0;
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++)
      for (int k = 0; k < b; k++) val[(size_t)q * bb + d * b + k] = 0.0;
  }
}
static __global__ void k_bc_cols(long ncrs, int b, const int* __restrict__ col, double* __restrict__ val,
                          const int* __restrict__ flag) {
  const long q = (long)blockIdx.x * kNT + threadIdx.x;
  if (q >= ncrs) return;
  const int j = col[q], bb = b * b;
  for (int d = 0; d < b; d++) {
    if (flag[(size_t)j * b + d] == 0) continue;
    for (int k = 0; k < b; k++) val[(size_t)q * bb + k * b + d] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Element assembly, gathered by matrix row.  Node i visits its elements in
// ascending element order and adds its row of each element matrix, so every
// entry receives its element contributions in the order sequential Mearge
// calls would add them (matdia_blkcrs.cpp:226-256); the residual likewise
// (vector_blk.h:170-181).  No atomics, no colouring.
__device__ inline int find_col(const int* __restrict__ col, int q0, int q1, int j) {
  for (int q = q0; q < q1; q++)
    if (col[q] == j) return q;
  return -1;
}

// src/femeqn/ker_emat_tri.cpp:12-14, 36-69
__device__ inline double d_tri_area(const double* p0, const double* p1, const double* p2) {
  return 0.5 * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}
__device__ inline void d_tri_dldx(double dl[3][2], const double* p0, const double* p1, const double* p2) {
  const double area = d_tri_area(p0, p1, p2);
  const double t = 0.5 / area;
  dl[0][0] = t * (p1[1] - p2[1]);
  dl[1][0] = t * (p2[1] - p0[1]);
  dl[2][0] = t * (p0[1] - p1[1]);
  dl[0][1] = t * (p2[0] - p1[0]);
  dl[1][1] = t * (p0[0] - p2[0]);
  dl[2][1] = t * (p1[0] - p0[0]);
}
// include/delfem/femeqn/ker_emat_tet.h:32-77
__device__ inline double d_tet_volume(const double* p0, const double* p1, const double* p2, const double* p3) {
  double v = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p3[1] - p0[1]) * (p2[2] - p0[2])) -
             (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p3[0] - p0[0]) * (p2[2] - p0[2])) +
             (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p3[0] - p0[0]) * (p2[1] - p0[1]));
  v *= 0.1666666666666;
  return v;
}
__device__ inline void d_tet_dldx(double d[4][3], const double* p0, const double* p1, const double* p2,
                                  const double* p3) {
  const double vol = d_tet_volume(p0, p1, p2, p3);
  const double t = 1.0 / (vol * 6.0);
  d[0][0] = -t * ((p2[1] - p1[1]) * (p3[2] - p1[2]) - (p3[1] - p1[1]) * (p2[2] - p1[2]));
  d[0][1] = +t * ((p2[0] - p1[0]) * (p3[2] - p1[2]) - (p3[0] - p1[0]) * (p2[2] - p1[2]));
  d[0][2] = -t * ((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]));
  d[1][0] = +t * ((p3[1] - p2[1]) * (p0[2] - p2[2]) - (p0[1] - p2[1]) * (p3[2] - p2[2]));
  d[1][1] = -t * ((p3[0] - p2[0]) * (p0[2] - p2[2]) - (p0[0] - p2[0]) * (p3[2] - p2[2]));
  d[1][2] = +t * ((p3[0] - p2[0]) * (p0[1] - p2[1]) - (p0[0] - p2[0]) * (p3[1] - p2[1]));
  d[2][0] = -t * ((p0[1] - p3[1]) * (p1[2] - p3[2]) - (p1[1] - p3[1]) * (p0[2] - p3[2]));
  d[2][1] = +t * ((p0[0] - p3[0]) * (p1[2] - p3[2]) - (p1[0] - p3[0]) * (p0[2] - p3[2]));
  d[2][2] = -t * ((p0[0] - p3[0]) * (p1[1] - p3[1]) - (p1[0] - p3[0]) * (p0[1] - p3[1]));
  d[3][0] = +t * ((p1[1] - p0[1]) * (p2[2] - p0[2]) - (p2[1] - p0[1]) * (p1[2] - p0[2]));
  d[3][1] = -t * ((p1[0] - p0[0]) * (p2[2] - p0[2]) - (p2[0] - p0[0]) * (p1[2] - p0[2]));
  d[3][2] = +t * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

// AddLinearSystem_Poisson2D_P1 (src/femeqn/eqn_poisson.cpp:49-125), row of node i.
static __global__ void __launch_bounds__(kNT) k_asm_poisson2d(int nnode, const unsigned* __restrict__ conn,
                                                      const int* __restrict__ eptr, const int* __restrict__ elist,
                                                      const int* __restrict__ rowptr, const int* __restrict__ col,
                                                      double* __restrict__ val, double* __restrict__ dia,
                                                      const double* __restrict__ coord,
                                                      const double* __restrict__ alpha, double source,
                                                      const double* __restrict__ value, double* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double racc = res ? res[i] : 0.0;
  double dacc = dia[i];
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
    double em[3];
#pragma unroll
    for (int j = 0; j < 3; j++) em[j] = alpha[e] * area * (dl[a][0] * dl[j][0] + dl[a][1] * dl[j][1]);
    double er = source * area * 0.33333333333333333;
#pragma unroll
    for (int j = 0; j < 3; j++) er -= em[j] * value[no[j]];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j == a) {
        dacc += em[j];
      } else {
        const int q = find_col(col, q0, q1, (int)no[j]);
        if (q >= 0) val[q] += em[j];
      }
    }
    racc += er;
  }
  dia[i] = dacc;
  if (res) res[i] = racc;
}

// AddLinSys_LinearSolid3D_Static_P1 (src/femeqn/eqn_linear_solid3d.cpp:241-334), row of node i.
static __global__ void __launch_bounds__(kNT) k_asm_solid3d(int nnode, const unsigned* __restrict__ conn,
                                                    const int* __restrict__ eptr, const int* __restrict__ elist,
                                                    const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    double* __restrict__ val, double* __restrict__ dia,
                                                    const double* __restrict__ coord,
                                                    const double* __restrict__ lambda,
                                                    const double* __restrict__ myu, double rho, double gx,
                                                    double gy, double gz, const double* __restrict__ disp,
                                                    double* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double racc[3] = {0.0, 0.0, 0.0};
  if (res)
    for (int x = 0; x < 3; x++) racc[x] = res[(size_t)i * 3 + x];
  for (int t = eptr[i]; t < eptr[i + 1]; t++) {
    const int e = elist[t];
    unsigned no[4];
    double p[4][3];
    int a = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      no[k] = conn[(size_t)e * 4 + k];
#pragma unroll
      for (int d = 0; d < 3; d++) p[k][d] = coord[(size_t)no[k] * 3 + d];
      if ((int)no[k] == i) a = k;
    }
    const double vol = d_tet_volume(p[0], p[1], p[2], p[3]);
    double dl[4][3];
    d_tet_dldx(dl, p[0], p[1], p[2], p[3]);
    const double lam = lambda[e], mu = myu[e];
    double er[3] = {vol * rho * gx * 0.25, vol * rho * gy * 0.25, vol * rho * gz * 0.25};
    for (int c = 0; c < 4; c++) {
      double E[9];
#pragma unroll
      for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) E[x * 3 + y] = vol * (lam * dl[a][x] * dl[c][y] + mu * dl[c][x] * dl[a][y]);
      const double dd = dl[a][0] * dl[c][0] + dl[a][1] * dl[c][1] + dl[a][2] * dl[c][2];
      E[0] += vol * mu * dd;
      E[4] += vol * mu * dd;
      E[8] += vol * mu * dd;
      const double* u = disp + (size_t)no[c] * 3;
      er[0] -= E[0] * u[0] + E[1] * u[1] + E[2] * u[2];
      er[1] -= E[3] * u[0] + E[4] * u[1] + E[5] * u[2];
      er[2] -= E[6] * u[0] + E[7] * u[1] + E[8] * u[2];
      double* out;
      if (c == a) {
        out = dia + (size_t)i * 9;
      } else {
        const int q = find_col(col, q0, q1, (int)no[c]);
        if (q < 0) continue;
        out = val + (size_t)q * 9;
      }
#pragma unroll
      for (int s = 0; s < 9; s++) out[s] += E[s];
    }
#pragma unroll
    for (int x = 0; x < 3; x++) racc[x] += er[x];
  }
  if (res)
    for (int x = 0; x < 3; x++) res[(size_t)i * 3 + x] = racc[x];
}

// Steady Oseen P1-P1 with PSPG, combined (u,v,p) 3x3 blocks: viscous,
// Galerkin advection, pressure/continuity and PSPG terms of
// src/femeqn/eqn_navier_stokes.cpp:73-110, 231-253 with a prescribed advection
// field, stabilisation parameter of :137-158 without the Courant term; block
// layout of the combined Stokes assembly (src/femeqn/eqn_stokes.cpp:317-331).
// Row of node i: the 3 blocks E[a][j] of each element around i, and its
// residual share -E[a][.] * value.
__device__ inline void oseen_row(const double* p0, const double* p1, const double* p2, const double (&a)[3][2],
                                 int ia, double myu, double rho, double (&E)[3][9]) {
  const double area = d_tri_area(p0, p1, p2);
  double dl[3][2];
  d_tri_dldx(dl, p0, p1, p2);
  const int i = ia;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    E[j][0] = area * myu * dl[i][0] * dl[j][0];
    E[j][1] = area * myu * dl[i][1] * dl[j][0];
    E[j][3] = area * myu * dl[i][0] * dl[j][1];
    E[j][4] = area * myu * dl[i][1] * dl[j][1];
    const double t = area * myu * (dl[j][0] * dl[i][0] + dl[j][1] * dl[i][1]);
    E[j][0] += t;
    E[j][4] += t;
  }
  {
    const double s0 = a[0][0] + a[1][0] + a[2][0], s1 = a[0][1] + a[1][1] + a[2][1];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double t1 = dl[j][0] * s0 + dl[j][1] * s1;
      double t2 = t1 + (dl[j][0] * a[i][0] + dl[j][1] * a[i][1]);
      t2 *= area * rho * 0.083333333333333;
      E[j][0] += t2;
      E[j][4] += t2;
    }
  }
  {
    const double t = area * 0.33333333333333333;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      E[j][2] = -t * dl[i][0];
      E[j][5] = -t * dl[i][1];
      E[j][6] = t * dl[j][0];
      E[j][7] = t * dl[j][1];
    }
  }
  double tau;
  {
    const double va0 = (a[0][0] + a[1][0] + a[2][0]) / 3.0, va1 = (a[0][1] + a[1][1] + a[2][1]) / 3.0;
    const double norm_v = sqrt(va0 * va0 + va1 * va1);
    const double h = sqrt(area / 3.14) * 2;
    if (norm_v * h * rho * 1.0e-30 > myu) {
      tau = h * 0.5 / norm_v;
    } else if (norm_v * h * rho < myu * 1.0e-30) {
      tau = h * h * rho * 0.5 / myu;
    } else {
      const double re_c = 0.5 * norm_v * h * rho / myu;
      tau = (h * 0.5 / norm_v) / sqrt(1 + 1 / (re_c * re_c));
    }
  }
  {
    double av0 = a[0][0] + a[1][0] + a[2][0], av1 = a[0][1] + a[1][1] + a[2][1];
    av0 *= 0.3333333333333333333;
    av1 *= 0.3333333333333333333;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double t1 = (dl[j][0] * av0 + dl[j][1] * av1) * tau * area;
      E[j][6] += t1 * dl[i][0];
      E[j][7] += t1 * dl[i][1];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; j++) E[j][8] = area * tau / rho * (dl[j][0] * dl[i][0] + dl[j][1] * dl[i][1]);
}

static __global__ void __launch_bounds__(kNT) k_asm_oseen2d(int nnode, const unsigned* __restrict__ conn,
                                                           const int* __restrict__ eptr, const int* __restrict__ elist,
                                                           const int* __restrict__ rowptr, const int* __restrict__ col,
                                                           double* __restrict__ val, double* __restrict__ dia,
                                                           const double* __restrict__ coord, double myu, double rho,
                                                           const double* __restrict__ adv,
                                                           const double* __restrict__ value, double* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double racc[3] = {0.0, 0.0, 0.0};
  if (res)
    for (int r = 0; r < 3; r++) racc[r] = res[(size_t)i * 3 + r];
  for (int t = eptr[i]; t < eptr[i + 1]; t++) {
    const int e = elist[t];
    unsigned no[3];
    double p[3][2], a[3][2];
    int ia = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      no[k] = conn[(size_t)e * 3 + k];
      p[k][0] = coord[(size_t)no[k] * 2];
      p[k][1] = coord[(size_t)no[k] * 2 + 1];
      a[k][0] = adv[(size_t)no[k] * 2];
      a[k][1] = adv[(size_t)no[k] * 2 + 1];
      if ((int)no[k] == i) ia = k;
    }
    double E[3][9];
    oseen_row(p[0], p[1], p[2], a, ia, myu, rho, E);
    double er[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; j++) {
      const double* u = value + (size_t)no[j] * 3;
#pragma unroll
      for (int r = 0; r < 3; r++) er[r] -= E[j][r * 3 + 0] * u[0] + E[j][r * 3 + 1] * u[1] + E[j][r * 3 + 2] * u[2];
      double* out;
      if (j == ia) {
        out = dia + (size_t)i * 9;
      } else {
        const int q = find_col(col, q0, q1, (int)no[j]);
        if (q < 0) continue;
        out = val + (size_t)q * 9;
      }
#pragma unroll
      for (int s = 0; s < 9; s++) out[s] += E[j][s];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) racc[r] += er[r];
  }
  if (res)
    for (int r = 0; r < 3; r++) res[(size_t)i * 3 + r] = racc[r];
}

// Positive Helmholtz / implicit-diffusion form: alpha*K + c*M with the P1
// consistent mass of src/femeqn/eqn_diffusion.cpp:302-314 (emat = M + 1*K),
// per-element alpha and c, source f; eres = f*area/3 - emat * value.
static __global__ void __launch_bounds__(kNT) k_asm_helmholtz2d(int nnode, const unsigned* __restrict__ conn,
                                                               const int* __restrict__ eptr,
                                                               const int* __restrict__ elist,
                                                               const int* __restrict__ rowptr,
                                                               const int* __restrict__ col, double* __restrict__ val,
                                                               double* __restrict__ dia,
                                                               const double* __restrict__ coord,
                                                               const double* __restrict__ alpha,
                                                               const double* __restrict__ cmass, double source,
                                                               const double* __restrict__ value,
                                                               double* __restrict__ res) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= nnode) return;
  const int q0 = rowptr[i], q1 = rowptr[i + 1];
  double racc = res ? res[i] : 0.0;
  double dacc = dia[i];
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
    const double m = cmass[e] * area * 0.08333333333333333;
    double em[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double kc = alpha[e] * area * (dl[a][0] * dl[j][0] + dl[a][1] * dl[j][1]);
      const double mm = (j == a) ? m + m : m;
      em[j] = mm + 1.0 * kc;
    }
    double er = source * area * 0.333333333333333333;
#pragma unroll
    for (int j = 0; j < 3; j++) er -= em[j] * value[no[j]];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j == a) {
        dacc += em[j];
      } else {
        const int q = find_col(col, q0, q1, (int)no[j]);
        if (q >= 0) val[q] += em[j];
      }
    }
    racc += er;
  }
  dia[i] = dacc;
  if (res) res[i] = racc;
}

// ---------------------------------------------------------------------------
// ILU numeric factorisation (CMatDiaFrac_BlkCrs::SetValue_Initialize + DoILUDecomp,
// include/delfem/matvec/matdiafrac_blkcrs.h:90-117, src/matvec/matdiafrac_blkcrs.cpp:769-1102).
static __global__ void k_ilu_copy(long nf, int bb, const long long* __restrict__ fsrc, const double* __restrict__ aval,
                           double* __restrict__ fval) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= nf * bb) return;
  const long q = t / bb;
  const int e = (int)(t - q * bb);
  const long long s = fsrc[q];
  fval[t] = s >= 0 ? aval[(size_t)s * bb + e] : 0.0;
}
// rows [row0, nb) of the factor's diagonal
static __global__ void k_ilu_copy_dia(int nb, int row0, int bb, const int* __restrict__ order,
                                      const double* __restrict__ adia, double* __restrict__ fdia) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= (long)(nb - row0) * bb) return;
  const long i = row0 + t / bb;
  const int e = (int)(t - (i - row0) * bb);
  const long io = order ? order[i] : i;
  fdia[(size_t)i * bb + e] = adia[(size_t)io * bb + e];
}

// One thread per row of a level of the lower DAG: IKJ update with the
// finished rows k in L(i), then invert the pivot block and scale the U row.
// singrow[i] records whether row i's pivot was singular.  partial != 0
// reproduces the state DoILUDecomp leaves in the row where it gives up (the
// 11th singular pivot, matdiafrac_blkcrs.cpp:880-890): the row is updated,
// its pivot left as the inversion attempt left it, its U part not scaled.
template <int B>
static __global__ void __launch_bounds__(kNT) k_ilu_factor(int nrow, const int* __restrict__ rows,
                                                   const int* __restrict__ frow, const int* __restrict__ fcol,
                                                   const int* __restrict__ fdiaind, double* __restrict__ fval,
                                                   double* __restrict__ fdia, int* __restrict__ nsing,
                                                   int* __restrict__ singrow, int partial) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  constexpr int BB = B * B;
  const int i = rows[t];
  const int r0 = frow[i], rd = fdiaind[i], r1 = frow[i + 1];
  double* D = fdia + (size_t)i * BB;
  for (int qk = r0; qk < rd; qk++) {
    const int k = fcol[qk];
    double L[BB];
#pragma unroll
    for (int s = 0; s < BB; s++) L[s] = fval[(size_t)qk * BB + s];
    int p = qk + 1;
    const int u0 = fdiaind[k], u1 = frow[k + 1];
    for (int qj = u0; qj < u1; qj++) {
      const int j = fcol[qj];
      double* Aij;
      if (j == i) {
        Aij = D;
      } else {
        while (p < r1 && fcol[p] < j) p++;
        if (p >= r1 || fcol[p] != j) continue;
        Aij = fval + (size_t)p * BB;
      }
      const double* U = fval + (size_t)qj * BB;
      double u[BB];
#pragma unroll
      for (int s = 0; s < BB; s++) u[s] = U[s];
      if constexpr (B == 1) {
        Aij[0] -= L[0] * u[0];
      } else {
#pragma unroll
        for (int r = 0; r < B; r++)
#pragma unroll
          for (int c = 0; c < B; c++) {
            double s = L[r * B] * u[c];
#pragma unroll
            for (int m = 1; m < B; m++) s = s + L[r * B + m] * u[m * B + c];
            Aij[r * B + c] -= s;
          }
      }
    }
  }
  double d[BB];
#pragma unroll
  for (int s = 0; s < BB; s++) d[s] = D[s];
  bool singular = false;
  if constexpr (B == 1) {
    if (fabs(d[0]) > 1.0e-30) d[0] = 1.0 / d[0];
    else singular = true;
  } else if constexpr (B == 2) {
    const double det = d[0] * d[3] - d[1] * d[2];
    if (fabs(det) > 1.0e-30) {
      const double id = 1.0 / det;
      const double d0 = d[0];
      d[0] = id * d[3];
      d[1] = -id * d[1];
      d[2] = -id * d[2];
      d[3] = id * d0;
    } else {
      singular = true;
    }
  } else {
    const double det = d[0] * d[4] * d[8] + d[3] * d[7] * d[2] + d[6] * d[1] * d[5] - d[0] * d[7] * d[5] -
                       d[6] * d[4] * d[2] - d[3] * d[1] * d[8];
    if (fabs(det) > 1.0e-30) {
      double s[9];
#pragma unroll
      for (int z = 0; z < 9; z++) s[z] = d[z];
      const double id = 1.0 / det;
      d[0] = id * (s[4] * s[8] - s[5] * s[7]);
      d[1] = id * (s[2] * s[7] - s[1] * s[8]);
      d[2] = id * (s[1] * s[5] - s[2] * s[4]);
      d[3] = id * (s[5] * s[6] - s[3] * s[8]);
      d[4] = id * (s[0] * s[8] - s[2] * s[6]);
      d[5] = id * (s[2] * s[3] - s[0] * s[5]);
      d[6] = id * (s[3] * s[7] - s[4] * s[6]);
      d[7] = id * (s[1] * s[6] - s[0] * s[7]);
      d[8] = id * (s[0] * s[4] - s[1] * s[3]);
    } else {
      singular = true;
    }
  }
  if (singular) atomicAdd(nsing, 1);
  singrow[i] = singular ? 1 : 0;
  if (partial && singular) return;  // the pivot block is unchanged when singular (b <= 3)
#pragma unroll
  for (int s = 0; s < BB; s++) D[s] = d[s];
  for (int q = rd; q < r1; q++) {
    double* U = fval + (size_t)q * BB;
    double u[BB];
#pragma unroll
    for (int s = 0; s < BB; s++) u[s] = U[s];
    if constexpr (B == 1) {
      U[0] = u[0] * d[0];
    } else {
#pragma unroll
      for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
          double s = d[r * B] * u[c];
#pragma unroll
          for (int m = 1; m < B; m++) s = s + d[r * B + m] * u[m * B + c];
          U[r * B + c] = s;
        }
    }
  }
}

// The same factorisation with one wave per row (rows of at most kIluWaveRow
// entries): the row is staged in LDS, the updates of step k (one per entry
// of U(k), each to a different entry of row i) are spread over the lanes and
// the steps run in order, so every entry receives the same subtractions in
// the same order as in k_ilu_factor (a wave's LDS accesses are performed in
// program order); pivot and U scaling as there.
constexpr int kIluWaveRow = 64;
constexpr int kIluWaves = 4;
template <int B>
__device__ inline void ilu_invert_pivot(double (&d)[B * B], bool& singular) {
  singular = false;
  if constexpr (B == 1) {
    if (fabs(d[0]) > 1.0e-30) d[0] = 1.0 / d[0];
    else singular = true;
  } else if constexpr (B == 2) {
    const double det = d[0] * d[3] - d[1] * d[2];
    if (fabs(det) > 1.0e-30) {
      const double id = 1.0 / det;
      const double d0 = d[0];
      d[0] = id * d[3];
      d[1] = -id * d[1];
      d[2] = -id * d[2];
      d[3] = id * d0;
    } else {
      singular = true;
    }
  } else {
    const double det = d[0] * d[4] * d[8] + d[3] * d[7] * d[2] + d[6] * d[1] * d[5] - d[0] * d[7] * d[5] -
                       d[6] * d[4] * d[2] - d[3] * d[1] * d[8];
    if (fabs(det) > 1.0e-30) {
      double s[9];
#pragma unroll
      for (int z = 0; z < 9; z++) s[z] = d[z];
      const double id = 1.0 / det;
      d[0] = id * (s[4] * s[8] - s[5] * s[7]);
      d[1] = id * (s[2] * s[7] - s[1] * s[8]);
      d[2] = id * (s[1] * s[5] - s[2] * s[4]);
      d[3] = id * (s[5] * s[6] - s[3] * s[8]);
      d[4] = id * (s[0] * s[8] - s[2] * s[6]);
      d[5] = id * (s[2] * s[3] - s[0] * s[5]);
      d[6] = id * (s[3] * s[7] - s[4] * s[6]);
      d[7] = id * (s[1] * s[6] - s[0] * s[7]);
      d[8] = id * (s[0] * s[4] - s[1] * s[3]);
    } else {
      singular = true;
    }
  }
}
template <int B>
static __global__ void __launch_bounds__(64 * kIluWaves) k_ilu_factor_w(
    int nrow, const int* __restrict__ rows, const int* __restrict__ frow, const int* __restrict__ fcol,
    const int* __restrict__ fdiaind, double* __restrict__ fval, double* __restrict__ fdia, int* __restrict__ nsing,
    int* __restrict__ singrow, int partial) {
  constexpr int BB = B * B;
  __shared__ double sv_all[kIluWaves][kIluWaveRow * BB + BB];
  __shared__ int sc_all[kIluWaves][kIluWaveRow];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * kIluWaves + w;
  if (t >= nrow) return;  // the whole wave
  double* sv = sv_all[w];
  double* sD = sv + kIluWaveRow * BB;
  int* sc = sc_all[w];
  const int i = __builtin_amdgcn_readfirstlane(rows[t]);
  const int r0 = __builtin_amdgcn_readfirstlane(frow[i]), rd = __builtin_amdgcn_readfirstlane(fdiaind[i]),
            r1 = __builtin_amdgcn_readfirstlane(frow[i + 1]);
  const int n = r1 - r0;
  for (int e = lane; e < n * BB; e += 64) sv[e] = fval[(size_t)r0 * BB + e];
  for (int e = lane; e < n; e += 64) sc[e] = fcol[r0 + e];
  if (lane < BB) sD[lane] = fdia[(size_t)i * BB + lane];
  __builtin_amdgcn_wave_barrier();
  for (int qk = r0; qk < rd; qk++) {
    const int k = __builtin_amdgcn_readfirstlane(fcol[qk]);
    double L[BB];
#pragma unroll
    for (int s2 = 0; s2 < BB; s2++) L[s2] = sv[(qk - r0) * BB + s2];
    const int u0 = __builtin_amdgcn_readfirstlane(fdiaind[k]), u1 = __builtin_amdgcn_readfirstlane(frow[k + 1]);
    for (int qj = u0 + lane; qj < u1; qj += 64) {
      const int j = fcol[qj];
      double* Aij;
      if (j == i) {
        Aij = sD;
      } else {
        int lo = qk + 1 - r0, hi = n;  // first column >= j in row i after entry qk
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (sc[mid] < j) lo = mid + 1;
          else hi = mid;
        }
        if (lo >= n || sc[lo] != j) continue;
        Aij = sv + lo * BB;
      }
      const double* U = fval + (size_t)qj * BB;
      double u[BB];
#pragma unroll
      for (int s2 = 0; s2 < BB; s2++) u[s2] = U[s2];
      if constexpr (B == 1) {
        Aij[0] -= L[0] * u[0];
      } else {
#pragma unroll
        for (int r = 0; r < B; r++)
#pragma unroll
          for (int c = 0; c < B; c++) {
            double s2 = L[r * B] * u[c];
#pragma unroll
            for (int m = 1; m < B; m++) s2 = s2 + L[r * B + m] * u[m * B + c];
            Aij[r * B + c] -= s2;
          }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  double d[BB];
#pragma unroll
  for (int s2 = 0; s2 < BB; s2++) d[s2] = sD[s2];
  bool singular;
  ilu_invert_pivot<B>(d, singular);
  if (lane == 0) {
    if (singular) atomicAdd(nsing, 1);
    singrow[i] = singular ? 1 : 0;
  }
  if (partial && singular) {
    // the row as updated, pivot as the inversion attempt left it, U not scaled
    for (int e = lane; e < n * BB; e += 64) fval[(size_t)r0 * BB + e] = sv[e];
    if (lane < BB) fdia[(size_t)i * BB + lane] = sD[lane];
    return;
  }
  if (lane < BB) fdia[(size_t)i * BB + lane] = d[lane];
  for (int e = lane; e < (rd - r0) * BB; e += 64) fval[(size_t)r0 * BB + e] = sv[e];
  for (int q = rd + lane; q < r1; q += 64) {
    double u[BB], o[BB];
#pragma unroll
    for (int s2 = 0; s2 < BB; s2++) u[s2] = sv[(q - r0) * BB + s2];
    if constexpr (B == 1) {
      o[0] = u[0] * d[0];
    } else {
#pragma unroll
      for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
          double s2 = d[r * B] * u[c];
#pragma unroll
          for (int m = 1; m < B; m++) s2 = s2 + d[r * B + m] * u[m * B + c];
          o[r * B + c] = s2;
        }
    }
#pragma unroll
    for (int s2 = 0; s2 < BB; s2++) fval[(size_t)q * BB + s2] = o[s2];
  }
}

// Generic block size (b > 3, b <= 8): CalcSubMatPr / CalcInvMat / CalcMatPr
// (matdiafrac_blkcrs.cpp:45-136).
static __global__ void __launch_bounds__(kNT) k_ilu_factor_gen(int nrow, int b, const int* __restrict__ rows,
                                                       const int* __restrict__ frow, const int* __restrict__ fcol,
                                                       const int* __restrict__ fdiaind, double* __restrict__ fval,
                                                       double* __restrict__ fdia, int* __restrict__ nsing,
                                                       int* __restrict__ singrow, int partial) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int bb = b * b;
  const int i = rows[t];
  const int r0 = frow[i], rd = fdiaind[i], r1 = frow[i + 1];
  double* D = fdia + (size_t)i * bb;
  for (int qk = r0; qk < rd; qk++) {
    const int k = fcol[qk];
    const double* L = fval + (size_t)qk * bb;
    int p = qk + 1;
    for (int qj = fdiaind[k]; qj < frow[k + 1]; qj++) {
      const int j = fcol[qj];
      double* Aij;
      if (j == i) {
        Aij = D;
      } else {
        while (p < r1 && fcol[p] < j) p++;
        if (p >= r1 || fcol[p] != j) continue;
        Aij = fval + (size_t)p * bb;
      }
      const double* U = fval + (size_t)qj * bb;
      for (int r = 0; r < b; r++)
        for (int c = 0; c < b; c++)
          for (int m = 0; m < b; m++) Aij[r * b + c] -= L[r * b + m] * U[m * b + c];
    }
  }
  bool singular = false;
  for (int r = 0; r < b && !singular; r++) {
    if (fabs(D[r * b + r]) < 1.0e-30) {
      singular = true;
      break;
    }
    double tt = 1.0 / D[r * b + r];
    D[r * b + r] = 1.0;
    for (int k = 0; k < b; k++) D[r * b + k] *= tt;
    for (int j = 0; j < b; j++) {
      if (j == r) continue;
      tt = D[j * b + r];
      D[j * b + r] = 0.0;
      for (int k = 0; k < b; k++) D[j * b + k] -= tt * D[r * b + k];
    }
  }
  if (singular) atomicAdd(nsing, 1);
  singrow[i] = singular ? 1 : 0;
  // CalcInvMat leaves the pivot block half-inverted where it stops (in place, as here)
  if (partial && singular) return;
  double tmp[64];
  for (int q = rd; q < r1; q++) {
    double* U = fval + (size_t)q * bb;
    for (int r = 0; r < b; r++)
      for (int c = 0; c < b; c++) {
        tmp[r * b + c] = 0.0;
        for (int m = 0; m < b; m++) tmp[r * b + c] += D[r * b + m] * U[m * b + c];
      }
    for (int s = 0; s < bb; s++) U[s] = tmp[s];
  }
}

// ---------------------------------------------------------------------------
// Sliced-ELL value refresh: to[(k*bb+e)*64+lane] = from[src*bb+e] (0 for padding).
template <typename T>
static __global__ void k_sell_gather(long nslot, int bb, const long long* __restrict__ src, const double* __restrict__ from,
                              T* __restrict__ to) {
  const long slot = (long)blockIdx.x * kNT + threadIdx.x;
  if (slot >= nslot) return;
  const long k = slot >> 6;
  const int lane = (int)(slot & 63);
  const long long s = src[slot];
  for (int e = 0; e < bb; e++) to[slot_index<T>(k, bb, e, lane)] = store_as<T>(s >= 0 ? from[(size_t)s * bb + e] : 0.0);
  if (slot_elems<T>(bb) > bb) to[slot_index<T>(k, bb, bb, lane)] = store_as<T>(0.0);
}
// Diagonal blocks per slice: to[(s*bb+e)*64+lane] = from[map(row)*bb+e].
template <typename T>
static __global__ void k_sell_gather_dia(long nslot, int bb, const int* __restrict__ rows, const int* __restrict__ order,
                                  const double* __restrict__ from, T* __restrict__ to) {
  const long slot = (long)blockIdx.x * kNT + threadIdx.x;
  if (slot >= nslot) return;
  const long s = slot >> 6;
  const int lane = (int)(slot & 63);
  const int row = rows[slot];
  const long src = row < 0 ? -1 : (order ? order[row] : row);
  for (int e = 0; e < bb; e++)
    to[slot_index<T>(s, bb, e, lane)] = store_as<T>(src >= 0 ? from[(size_t)src * bb + e] : 0.0);
  if (slot_elems<T>(bb) > bb) to[slot_index<T>(s, bb, bb, lane)] = store_as<T>(0.0);
}


// s[r] = sum_c A[r*B+c]*x[c], the sum blk_msub forms before subtracting it
template <int B, typename T>
__device__ inline void blk_prod(T* s, const T* a, const T* x) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    T v = a[r * B] * x[0];
#pragma unroll
    for (int c = 1; c < B; c++) v = v + a[r * B + c] * x[c];
    s[r] = v;
  }
}

// What the slot helpers do with each slot's block a and gathered x.
enum { SLOT_MSUB = 0, SLOT_PROD = 1, SLOT_MADD = 2 };
// Slots [k0 + base, k0 + base + m) (m <= 16, wave-uniform) of a lane whose
// columns are complete in x, in order: SLOT_MSUB t -= block * x (blk_msub),
// SLOT_MADD t += block * x (blk_madd), SLOT_PROD the products of slot k
// (counted from k0) left in prod[k] for k < NP and in
// ovf[(ovf0 + k - NP) * B ..] for NP <= k < nreal (the lane's real slots).
// The m column indices are fetched at once, then blocks and gathers in chunks
// of CH slots, so a lane waits for two memory round trips only once.
// FIRST: base == 0, so slot numbers are constants.  init(t) runs once, in
// the FIRST call after the first chunk's loads have been issued and before
// its first block product, so that a caller's own arithmetic for t does not
// hold back the slot loads.
struct NoInit {
  template <typename A>
  __device__ __forceinline__ void operator()(A&) const {}
};
template <int B, typename TS, typename T, int OP, int NP, bool FIRST, typename Init = NoInit>
__device__ __forceinline__ void sell_slots16(const int* __restrict__ col, const TS* __restrict__ val, const T* x,
                                             int k0, int base, int m, int lane, T (&t)[B], T (&prod)[NP][B],
                                             T* ovf, int ovf0, int nreal, const Init& init = Init()) {
  constexpr int BB = B * B;
  constexpr int E = slot_words<TS>(BB);
  constexpr int CH = E * 4 <= 20 ? 8 : 4;
  int cc[16];
#pragma unroll
  for (int u = 0; u < 16; u++)
    if (u < m) cc[u] = ld_stream(col + (size_t)(k0 + base + u) * 64 + lane);
#pragma unroll
  for (int c0 = 0; c0 < 16; c0 += CH) {
    if (c0 < m) {
      unsigned w[CH][E];
      T xv[CH][B];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        if (c0 + u < m) {
          load_words<BB, TS>(val, (size_t)(k0 + base + c0 + u), lane, w[u]);
          gather_x<B, T>(x, cc[c0 + u], xv[u]);
        }
      }
      if constexpr (FIRST)
        if (c0 == 0) init(t);
#pragma unroll
      for (int u = 0; u < CH; u++) {
        if (c0 + u < m) {
          T a[BB];
          words_to_block<BB, TS, T>(w[u], a);
          if constexpr (OP == SLOT_PROD) {
            const int k = (FIRST ? 0 : base) + c0 + u;
            if (FIRST && c0 + u < NP) {
              blk_prod<B, T>(prod[(c0 + u) % NP], a, xv[u]);
            } else if (k < nreal) {
              T p[B];
              blk_prod<B, T>(p, a, xv[u]);
#pragma unroll
              for (int r = 0; r < B; r++) ovf[(size_t)(ovf0 + k - NP) * B + r] = p[r];
            }
          } else if constexpr (OP == SLOT_MADD) {
            blk_madd<B, T>(t, a, xv[u]);
          } else {
            blk_msub<B, T>(t, a, xv[u]);
          }
        }
      }
    }
  }
}
// All n (wave-uniform) slots from k0, 16 at a time (init(t) as above; with
// no slots it runs at once).
template <int B, typename TS, typename T, int OP, int NP, typename Init = NoInit>
__device__ __forceinline__ void sell_slots(const int* __restrict__ col, const TS* __restrict__ val, const T* x,
                                           int k0, int n, int lane, T (&t)[B], T (&prod)[NP][B], T* ovf,
                                           int ovf0, int nreal, const Init& init = Init()) {
  static_assert(NP <= 16, "register products must come from the first 16 slots");
  if (n > 0)
    sell_slots16<B, TS, T, OP, NP, true, Init>(col, val, x, k0, 0, n < 16 ? n : 16, lane, t, prod, ovf, ovf0, nreal,
                                               init);
  else
    init(t);
  for (int base = 16; base < n; base += 16)
    sell_slots16<B, TS, T, OP, NP, false>(col, val, x, k0, base, n - base < 16 ? n - base : 16, lane, t, prod, ovf,
                                          ovf0, nreal);
}

// acc += sum over slots k0 <= k < k1 (in that order) of block(k) * x[col(k)];
// k0 and k1 are the same across the wave (one slice per wave).
template <int B, typename TS, typename T>
__device__ __forceinline__ void sell_madd_range(const int* __restrict__ col, const TS* __restrict__ val,
                                                const T* __restrict__ x, int k0, int k1, int lane, T (&acc)[B]) {
  T none[1][B];
  k0 = __builtin_amdgcn_readfirstlane(k0);
  k1 = __builtin_amdgcn_readfirstlane(k1);
  sell_slots<B, TS, T, SLOT_MADD, 1>(col, val, x, k0, k1 - k0, lane, acc, none, nullptr, 0, 0);
}

// ---------------------------------------------------------------------------
// SpMV on the sliced-ELL copy of A (rows = slice*64+lane, diagonal last):
//   y = A x            (res == nullptr)
//   y = res - A x      (true residual form)
// Optional partial sums (DOTS): d0 = <y,w0>, d1 = <y,y>, d2 = <u2,v2>, in double.
// BiCGSTAB omega = <s,AMs>/<AMs,AMs> uses w0 = s.  GATE: skip once the solver
// has stopped (Krylov-internal products); the true-residual form runs always.
template <int B, typename TS, typename T, bool DOTS, bool GATE>
__global__ void __launch_bounds__(kSpmvNT) k_spmv(int nslice, int nb, const int* __restrict__ sptr,
                                             const int* __restrict__ col, const TS* __restrict__ val,
                                             const TS* __restrict__ dval, const T* __restrict__ x,
                                             T* __restrict__ y, const T* __restrict__ res,
                                             const T* __restrict__ w0, const T* __restrict__ u2,
                                             const T* __restrict__ v2, double* __restrict__ partials,
                                             const KState* __restrict__ st) {
  constexpr int BB = B * B;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int s = xcd_block() * kSpmvWaves + wave;
  // the slice's slot range (scalar loads) is fetched in the same round trip
  // as the stop flag
  const int sl = s < nslice ? s : nslice - 1;
  const int k0 = sptr[sl], k1 = sptr[sl + 1];
  const int done = GATE ? st->done : 0;
  issue_together(k0, k1, done);
  if (done) return;
  double dsum[3] = {0.0, 0.0, 0.0};
  if (s < nslice) {
    const int row = s * 64 + lane;
    if (row < nb) {
      T acc[B];
#pragma unroll
      for (int r = 0; r < B; r++) acc[r] = (T)0;
      sell_madd_range<B, TS, T>(col, val, x, k0, k1, lane, acc);
      {
        T a[BB], xv[B];
        load_block<BB, TS, T>(dval, (size_t)s, lane, a);
#pragma unroll
        for (int e = 0; e < B; e++) xv[e] = x[(size_t)row * B + e];
        blk_madd<B, T>(acc, a, xv);
      }
#pragma unroll
      for (int r = 0; r < B; r++) {
        const size_t o = (size_t)row * B + r;
        const T yv = res ? (T)(res[o] - acc[r]) : acc[r];
        y[o] = yv;
        if (DOTS) {
          if (w0) dsum[0] += (double)yv * (double)w0[o];
          dsum[1] += (double)yv * (double)yv;
          if (u2) dsum[2] += (double)u2[o] * (double)v2[o];
        }
      }
    }
  }
  if constexpr (DOTS) write_partials<3, kSpmvNT>(dsum, partials, blockIdx.x, gridDim.x);
}

// t -= sum over slots k0 <= k < k1 (in that order) of block(k) * v[col(k)];
// k0 and k1 are the same across the wave (one slice per wave).
template <int B, typename TS, typename T, typename Init = NoInit>
__device__ __forceinline__ void sell_msub_range(const int* __restrict__ col, const TS* __restrict__ val,
                                                const T* v, int k0, int k1, int lane, T (&t)[B],
                                                const Init& init = Init()) {
  T none[1][B];
  k0 = __builtin_amdgcn_readfirstlane(k0);
  k1 = __builtin_amdgcn_readfirstlane(k1);
  sell_slots<B, TS, T, SLOT_MSUB, 1, Init>(col, val, v, k0, k1 - k0, lane, t, none, nullptr, 0, 0, init);
}

// What a one-slice wave of a level kernel reads before anything else: its
// lane's row, the slice's slot range and the stop flag (false without a
// state), fetched in one round trip rather than one after the other.
struct SliceHead {
  int row, k0, k1, done;
};
__device__ __forceinline__ SliceHead slice_head(int s, int lane, const int* __restrict__ rows,
                                                const int* __restrict__ sptr, const KState* __restrict__ st) {
  SliceHead h;
  h.k0 = sptr[s];
  h.k1 = sptr[s + 1];
  h.done = st ? st->done : 0;
  h.row = rows[(size_t)s * 64 + lane];
  issue_together(h.k0, h.k1, h.done);
  asm volatile("" ::"v"(h.row));
  return h;
}

// ---------------------------------------------------------------------------
// Triangular solves over one level set (ForwardSubstitution / BackwardSubstitution,
// src/matvec/matdiafrac_blkcrs.cpp:332-663).  Lower: out_i = Dinv_i (rhs_i - sum L_ij out_j);
// rhs may alias out.  Upper: v_i -= sum U_ij v_j in place; with DOT the rows
// finished here contribute <r,v> partials (slot part0 + workgroup) to a
// reduction that spans all level launches of the sweep.
// One 64-row slice of a lower sweep (lane = row of the slice).
// row: the lane's row (-1 padding); [k0, k1): the slice's slots.
template <int B, typename TS, typename T>
__device__ __forceinline__ void trsv_lower_slice(int s, int lane, int row, int k0, int k1,
                                                 const int* __restrict__ col, const TS* __restrict__ val,
                                                 const TS* __restrict__ dval, const T* rhs, T* out) {
  constexpr int BB = B * B;
  if (row < 0) return;
  T t[B];
#pragma unroll
  for (int r = 0; r < B; r++) t[r] = rhs[(size_t)row * B + r];
  sell_msub_range<B, TS, T>(col, val, out, k0, k1, lane, t);
  T d[BB], y[B];
  load_block<BB, TS, T>(dval, (size_t)s, lane, d);
  blk_mul<B, T>(y, d, t);
#pragma unroll
  for (int r = 0; r < B; r++) out[(size_t)row * B + r] = y[r];
}

// One slice of an upper sweep; with DOT adds <rvec, v> of its rows to dsum.
template <int B, typename TS, typename T, bool DOT>
__device__ __forceinline__ void trsv_upper_slice(int lane, int row, int k0, int k1, const int* __restrict__ col,
                                                 const TS* __restrict__ val, T* vec, const T* __restrict__ rvec,
                                                 double& dsum) {
  if (row < 0) return;
  T t[B];
#pragma unroll
  for (int r = 0; r < B; r++) t[r] = vec[(size_t)row * B + r];
  sell_msub_range<B, TS, T>(col, val, vec, k0, k1, lane, t);
#pragma unroll
  for (int r = 0; r < B; r++) {
    vec[(size_t)row * B + r] = t[r];
    if (DOT) dsum += (double)rvec[(size_t)row * B + r] * (double)t[r];
  }
}

template <int B, typename TS, typename T>
__global__ void __launch_bounds__(kTrsvNT) k_trsv_lower(int s0, int s1, const int* __restrict__ rows,
                                                   const int* __restrict__ sptr, const int* __restrict__ col,
                                                   const TS* __restrict__ val, const TS* __restrict__ dval,
                                                   const T* rhs, T* out, const KState* __restrict__ st) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int s = s0 + xcd_block() * kTrsvWaves + wave;
  const int sl = s < s1 ? s : s1 - 1;
  const SliceHead h = slice_head(sl, lane, rows, sptr, st);
  if (h.done || s >= s1) return;
  trsv_lower_slice<B, TS, T>(s, lane, h.row, h.k0, h.k1, col, val, dval, rhs, out);
}

template <int B, typename TS, typename T, bool DOT>
__global__ void __launch_bounds__(kTrsvNT) k_trsv_upper(int s0, int s1, const int* __restrict__ rows,
                                                   const int* __restrict__ sptr, const int* __restrict__ col,
                                                   const TS* __restrict__ val, T* vec, const T* __restrict__ rvec,
                                                   double* __restrict__ partials, int part0, int nparts,
                                                   const KState* __restrict__ st) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int s = s0 + xcd_block() * kTrsvWaves + wave;
  const int sl = s < s1 ? s : s1 - 1;
  const SliceHead h = slice_head(sl, lane, rows, sptr, st);
  if (h.done) return;
  double dsum[1] = {0.0};
  if (s < s1) trsv_upper_slice<B, TS, T, DOT>(lane, h.row, h.k0, h.k1, col, val, vec, rvec, dsum[0]);
  if constexpr (DOT) write_partials<1, kTrsvNT>(dsum, partials, part0 + blockIdx.x, nparts);
}

// ---------------------------------------------------------------------------
// Persistent sweeps: every level set of a sweep in one launch, for factors
// with many small levels (ILU(k >= 1) under a block ordering), where a launch
// per level costs more than the level's work.  One-wave workgroups take
// 64-row slices in level order from a work counter (counter 0; counter i is
// sync[i * kSyncStride]).  A wave first
// loads what no earlier level writes -- its rows' right-hand side, the column
// of each of its first persist_cols() slots, the blocks of its first
// persist_blocks() slots (held as raw words) and its diagonal block --
// then waits until all slices of the previous level are done (counter
// 2 + g - 1 reaching that level's slice count).  After the wait one
// batch of loads -- every gather, and the blocks not yet held -- is issued at
// once, so the dependent part of a slice is one memory round trip; the
// subtractions then run in slot order, as in the per-level kernels, so FP64
// results are bitwise the same.  The slice is counted done in counter 2 + g.
// Slices are taken in level order by waves that are running, so a wave waits
// only on slices held by running waves: the sweep cannot deadlock whatever
// the grid and residency.  The wait is bounded: after kPersistMaxPolls polls
// the wave records st->sweep_fault and goes on (and every other wave stops
// waiting when it sees the fault), so the kernel always ends and the host
// reports the fault.  The count is published after an agent-scope release
// fence and read before an agent-scope acquire fence (the XCDs' L2 caches are
// not coherent with each other).  The last wave to leave (counter 1 counts
// them) resets the counters for the next launch.
constexpr int kPersistMaxPolls = 1 << 18;
// Counters of a persistent sweep sit kSyncStride ints apart (256 bytes), so
// that pollers of one level and producers of another do not share a line.
constexpr int kSyncStride = 64;
// Slots whose columns a wave loads before the wait (their gathered values,
// about 72 words, are held after it), and slots whose blocks it holds across
// the wait (about 96 words): register budgets that leave no spills.
template <int B, typename T>
__host__ __device__ constexpr int persist_cols() {
  return 72 / (B * (int)sizeof(T) / 4) < 16 ? 72 / (B * (int)sizeof(T) / 4) : 16;
}
template <typename TS, int BB, int PC>
__host__ __device__ constexpr int persist_blocks() {
  return 96 / slot_words<TS>(BB) < PC ? 96 / slot_words<TS>(BB) : PC;
}

template <int B, typename TS, typename T, bool UPPER, bool DOT>
__global__ void __launch_bounds__(64) k_trsv_persist(int nslice, int ngroup, const int* __restrict__ slev,
                                                     const int* __restrict__ gsp, const int* __restrict__ rows,
                                                     const int* __restrict__ sptr, const int* __restrict__ col,
                                                     const TS* __restrict__ val, const TS* __restrict__ dval,
                                                     const T* rhs, T* out, const T* __restrict__ rvec,
                                                     double* __restrict__ partials, KState* __restrict__ st,
                                                     int* __restrict__ sync) {
  constexpr int BB = B * B;
  constexpr int E = slot_words<TS>(BB);
  constexpr int PC = persist_cols<B, T>();
  constexpr int PB = persist_blocks<TS, BB, PC>();
  const int lane = threadIdx.x;
  double dsum[1] = {0.0};
  if (!st->done) {
    for (;;) {
      int s = 0;
      if (lane == 0) s = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s = __shfl(s, 0, 64);
      if (s >= nslice) break;
      const int g = slev[s];
      const int row = rows[(size_t)s * 64 + lane];
      const int k0 = sptr[s], nk = sptr[s + 1] - k0;  // nk is the same for the whole wave
      T t[B], d[BB];
      int c[PC];
      unsigned w[PB][E];
      if (row >= 0) {
#pragma unroll
        for (int e = 0; e < B; e++) t[e] = rhs[(size_t)row * B + e];
#pragma unroll
        for (int u = 0; u < PC; u++) c[u] = u < nk ? ld_stream(col + (size_t)(k0 + u) * 64 + lane) : -1;
#pragma unroll
        for (int u = 0; u < PB; u++)
          if (u < nk) load_words<BB, TS>(val, (size_t)(k0 + u), lane, w[u]);
        if constexpr (!UPPER) load_block<BB, TS, T>(dval, (size_t)s, lane, d);
      }
      if (g > 0) {
        const int need = gsp[g] - gsp[g - 1];
        int polls = 0;
        while (__hip_atomic_load(&sync[(1 + g) * kSyncStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++polls > kPersistMaxPolls) {
            __hip_atomic_store(&st->sweep_fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          // once any wave has given up, the others stop waiting too
          if (__hip_atomic_load(&st->sweep_fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      if (row >= 0) {
        // one batch: the gathers of the first PC slots and the blocks of slots PB..PC-1
        T xv[PC][B];
        unsigned w2[PC > PB ? PC - PB : 1][E];
#pragma unroll
        for (int u = 0; u < PC; u++) {
          if (u < nk) {
            gather_x<B, T>(out, c[u], xv[u]);
            if (u >= PB) load_words<BB, TS>(val, (size_t)(k0 + u), lane, w2[u - PB]);
          }
        }
#pragma unroll
        for (int u = 0; u < PC; u++) {
          if (u < nk) {
            T a[BB];
            if (u < PB) words_to_block<BB, TS, T>(w[u], a);
            else words_to_block<BB, TS, T>(w2[u - PB], a);
            blk_msub<B, T>(t, a, xv[u]);
          }
        }
        if (nk > PC) sell_msub_range<B, TS, T>(col, val, out, k0 + PC, k0 + nk, lane, t);
        if constexpr (UPPER) {
#pragma unroll
          for (int e = 0; e < B; e++) {
            out[(size_t)row * B + e] = t[e];
            if (DOT) dsum[0] += (double)rvec[(size_t)row * B + e] * (double)t[e];
          }
        } else {
          T y[B];
          blk_mul<B, T>(y, d, t);
#pragma unroll
          for (int e = 0; e < B; e++) out[(size_t)row * B + e] = y[e];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (lane == 0) __hip_atomic_fetch_add(&sync[(2 + g) * kSyncStride], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if constexpr (DOT) write_partials<1, 64>(dsum, partials, blockIdx.x, 0);
  if (lane == 0) {
    const int left = __hip_atomic_fetch_add(&sync[kSyncStride], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (left == (int)gridDim.x - 1)
      for (int i = 0; i < ngroup + 2; i++)
        __hip_atomic_store(&sync[i * kSyncStride], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
// Phase-tiled sweeps: one launch per phase (a colour of the ordering), one
// workgroup per tile of whole dependency components, thread t = tile row t
// (layout: TilePlan).  Before the first barrier every thread loads what does
// not depend on this launch: its right-hand side, the blocks and tile
// positions of its in-phase columns -- the first kTileInSlots in registers,
// the rest in the workgroup's overflow LDS (the row's places there: rov) --
// and, for a lower row, the subtraction of its columns in earlier phases
// (complete in memory; they come first in column order) and its diagonal
// block, for an upper row the products of its columns in later phases
// (subtracted last, in order; kTileOutSlots in registers, the rest in the
// overflow LDS).  Those columns' indices are fetched ahead of their blocks
// and gathers, so the prologue costs two dependent memory round trips per
// chunk of slots only once.  Then the tile runs its local levels with no
// global memory access: the rows of level l gather their in-phase columns
// from LDS (where the rows of earlier levels left them; padding slots read a
// row of zeros) in one batch, subtract them, finish and publish to LDS; a
// barrier ends the level.  Slice bounds are wave-uniform (a slice is one
// wave), so the per-slot branches are scalar.  The results go to the vector
// after the last level.  Every subtraction is the one of the per-level
// kernels in the same order (the library is built without contraction, so a
// block product formed first and subtracted after is exactly blk_msub's), and
// padding slots subtract exact zeros, so FP64 results are bitwise those of
// the per-level sweeps.  With DOT the tile's <rvec, v> partial goes to
// partials[tile].

template <int B, typename TS, typename T, int NT, bool UPPER, bool DOT>
__global__ void __launch_bounds__(NT) k_trsv_tile(int tile0, const int* __restrict__ rows,
                                                  const int* __restrict__ rlev, const int* __restrict__ depth,
                                                  const int* __restrict__ sptr, const int* __restrict__ split,
                                                  const int* __restrict__ col, const TS* __restrict__ val,
                                                  const TS* __restrict__ dval, const int* __restrict__ rov,
                                                  int ovf_in, int ovf_out, const T* rhs, T* out,
                                                  const T* __restrict__ rvec, double* __restrict__ partials,
                                                  const KState* __restrict__ st) {
  if (st && st->done) return;
  constexpr int BB = B * B;
  constexpr int E = slot_words<TS>(BB);
  constexpr int PI = kTileInSlots;
  constexpr int PO = UPPER ? kTileOutSlots : 1;
  __shared__ T xs[(NT + 1) * B];  // tile row NT: zeros
  // overflow: ovf_in in-phase blocks, ovf_out later-phase products (upper),
  // then the in-phase blocks' tile positions
  extern __shared__ __align__(16) unsigned char tile_lds[];
  T* ov_val = reinterpret_cast<T*>(tile_lds);
  T* ov_prod = ov_val + (size_t)ovf_in * BB;
  int* ov_col = reinterpret_cast<int*>(ov_prod + (size_t)ovf_out * B);
  const int tile = tile0 + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int s = tile * (NT / 64) + (tid >> 6);
  const size_t rix = (size_t)tile * NT + tid;
  const int row = rows[rix];
  const int lev = rlev[rix];
  const int nlev = depth[tile];
  // a slice is one wave: its bounds are uniform
  const int k0 = __builtin_amdgcn_readfirstlane(sptr[s]);
  const int km = __builtin_amdgcn_readfirstlane(split[s]);
  const int k1 = __builtin_amdgcn_readfirstlane(sptr[s + 1]);
  const int ki = UPPER ? k0 : km, ni = UPPER ? km - k0 : k1 - km;  // in-phase part
  const int ko = UPPER ? km : k0, no = UPPER ? k1 - km : km - k0;  // the other phases
  const int4 ov = reinterpret_cast<const int4*>(rov)[rix];  // in: first, count; out: first, count
  if (tid < B) xs[NT * B + tid] = (T)0;
  T t[B], d[BB];
  int ci[PI];
  unsigned wi[PI][E];
  T so[PO][B];  // upper: products of the first later-phase slots
  double dsum[1] = {0.0};
#pragma unroll
  for (int u = 0; u < PI; u++) ci[u] = NT;
  if (row >= 0) {
#pragma unroll
    for (int e = 0; e < B; e++) t[e] = rhs[(size_t)row * B + e];
#pragma unroll
    for (int u = 0; u < PI; u++) {
      if (u < ni) {
        const int c = col[(size_t)(ki + u) * 64 + lane];
        ci[u] = c >= 0 ? c : NT;
        load_words<BB, TS>(val, (size_t)(ki + u), lane, wi[u]);
      }
    }
    if constexpr (!UPPER) load_block<BB, TS, T>(dval, (size_t)s, lane, d);
    for (int u = 0; u < ov.y; u++) {
      T a[BB];
      load_block<BB, TS, T>(val, (size_t)(ki + PI + u), lane, a);
      const int o = ov.x + u;
#pragma unroll
      for (int e = 0; e < BB; e++) ov_val[(size_t)o * BB + e] = a[e];
      const int c = col[(size_t)(ki + PI + u) * 64 + lane];
      ov_col[o] = c >= 0 ? c : NT;
    }
    sell_slots<B, TS, T, UPPER ? SLOT_PROD : SLOT_MSUB, PO>(col, val, out, ko, no, lane, t, so, ov_prod, ov.z,
                                                       PO + ov.w);
  }
  __syncthreads();  // the zero row
  for (int l = 0; l < nlev; l++) {
    if (row >= 0 && lev == l) {
      T xv[PI][B];
#pragma unroll
      for (int u = 0; u < PI; u++)
#pragma unroll
        for (int e = 0; e < B; e++) xv[u][e] = xs[ci[u] * B + e];
#pragma unroll
      for (int u = 0; u < PI; u++) {
        if (u < ni) {
          T a[BB];
          words_to_block<BB, TS, T>(wi[u], a);
          blk_msub<B, T>(t, a, xv[u]);
        }
      }
      for (int u = 0; u < ov.y; u++) {
        T a[BB], x2[B];
        const int o = ov.x + u;
        const int c = ov_col[o];
#pragma unroll
        for (int e = 0; e < BB; e++) a[e] = ov_val[(size_t)o * BB + e];
#pragma unroll
        for (int e = 0; e < B; e++) x2[e] = xs[c * B + e];
        blk_msub<B, T>(t, a, x2);
      }
      if constexpr (UPPER) {
#pragma unroll
        for (int u = 0; u < PO; u++) {
          if (u < no) {
#pragma unroll
            for (int r = 0; r < B; r++) t[r] -= so[u][r];
          }
        }
        for (int u = 0; u < ov.w; u++) {
          const int o = ov.z + u;
#pragma unroll
          for (int r = 0; r < B; r++) t[r] -= ov_prod[(size_t)o * B + r];
        }
      }
      T y[B];
      if constexpr (UPPER) {
#pragma unroll
        for (int e = 0; e < B; e++) y[e] = t[e];
      } else {
        blk_mul<B, T>(y, d, t);
      }
#pragma unroll
      for (int e = 0; e < B; e++) {
        xs[tid * B + e] = y[e];
        t[e] = y[e];
      }
    }
    __syncthreads();
  }
  if (row >= 0) {
#pragma unroll
    for (int e = 0; e < B; e++) {
      out[(size_t)row * B + e] = t[e];
      if (DOT) dsum[0] += (double)rvec[(size_t)row * B + e] * (double)t[e];
    }
  }
  if constexpr (DOT) write_partials<1, NT>(dsum, partials, tile, 0);
}

// ---------------------------------------------------------------------------
// Split-lane phase-tiled sweeps (b x b blocks, b = 2 or 3; layout: TileSplit):
// G lanes per tile row, lane r computing row r of every block product
// (s = sum_c A[r][c] x[c], then t_r -= s, as blk_msub does for its row r),
// so a level costs a wave a third of the arithmetic of one lane per row.  The
// lower sweep's D^-1 product needs all of the row's t: the G lanes of a row
// are one quad of the wave, and each reads the others' values through DPP
// quad permutes.  Slot order, products and padding are those of k_trsv_tile,
// so FP64 results are bitwise the per-level sweeps'.  An alternative form
// (sweep mode 4): G times the threads per tile make its prologue and barriers
// cost more than the arithmetic it saves on the Oseen factors.
template <int C>
__device__ inline double quad_lane(double v) {
  constexpr int ctrl = C | (C << 2) | (C << 4) | (C << 6);
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int C>
__device__ inline float quad_lane(float v) {
  constexpr int ctrl = C | (C << 2) | (C << 4) | (C << 6);
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
}
// the value of v in lane c of this lane's group of G (a quad, or a half quad for G = 2)
template <int G, typename T>
__device__ inline T group_lane(T v, int c, int lane) {
  if constexpr (G == 4) {
    return c == 0 ? quad_lane<0>(v) : (c == 1 ? quad_lane<1>(v) : (c == 2 ? quad_lane<2>(v) : quad_lane<3>(v)));
  } else {
    // lanes (0,1) and (2,3) of a quad: lane c of the pair
    const T a = (lane & 2) ? quad_lane<2>(v) : quad_lane<0>(v);
    const T b = (lane & 2) ? quad_lane<3>(v) : quad_lane<1>(v);
    return c == 0 ? a : b;
  }
}
// t -= sum over B-value row slots [k0, k0 + n) (wave-uniform n) of
// a_r . x[col]; PROD: the products of the first NP slots to prod[], the
// lane's further real ones (k < nreal) to ovf[(ovf0 + k - NP) * B], nothing to t
template <int B, typename TS, typename T, bool PROD, int NP>
__device__ __forceinline__ void row_slots(const int* __restrict__ col, const TS* __restrict__ val, const T* x,
                                          int k0, int n, int lane, T& t, T (&prod)[NP], T* ovf, int ovf0,
                                          int nreal) {
  constexpr int E = slot_words<TS>(B);
  constexpr int CH = 8;
  for (int base = 0; base < n; base += 16) {
    const int m = n - base < 16 ? n - base : 16;
    int cc[16];
#pragma unroll
    for (int u = 0; u < 16; u++)
      if (u < m) cc[u] = ld_stream(col + (size_t)(k0 + base + u) * 64 + lane);
#pragma unroll
    for (int c0 = 0; c0 < 16; c0 += CH) {
      if (c0 < m) {
        unsigned w[CH][E];
        T xv[CH][B];
#pragma unroll
        for (int u = 0; u < CH; u++) {
          if (c0 + u < m) {
            load_words<B, TS>(val, (size_t)(k0 + base + c0 + u), lane, w[u]);
            gather_x<B, T>(x, cc[c0 + u], xv[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
          if (c0 + u < m) {
            T a[B];
            words_to_block<B, TS, T>(w[u], a);
            T v = a[0] * xv[u][0];
#pragma unroll
            for (int e = 1; e < B; e++) v = v + a[e] * xv[u][e];
            if constexpr (PROD) {
              const int k = base + c0 + u;
              if (k < NP) {
#pragma unroll
                for (int q = 0; q < NP; q++)
                  if (q == k) prod[q] = v;
              } else if (k < nreal) {
                ovf[(size_t)(ovf0 + k - NP) * B] = v;
              }
            } else {
              t -= v;
            }
          }
        }
      }
    }
  }
}

template <int B, typename TS, typename T, int NTS, bool UPPER, bool DOT>
__global__ void __launch_bounds__(NTS) k_trsv_tile_split(int tile0, const int* __restrict__ rows,
                                                         const int* __restrict__ rlev, const int* __restrict__ depth,
                                                         const int* __restrict__ sptr, const int* __restrict__ split,
                                                         const int* __restrict__ col, const TS* __restrict__ val,
                                                         const TS* __restrict__ dval, const int* __restrict__ rov,
                                                         int ovf_in, int ovf_out, const T* rhs, T* out,
                                                         const T* __restrict__ rvec, double* __restrict__ partials,
                                                         const KState* __restrict__ st) {
  if (st && st->done) return;
  constexpr int G = B == 2 ? 2 : 4;
  constexpr int NT = NTS / G;  // tile rows
  constexpr int E = slot_words<TS>(B);
  constexpr int PI = 2 * kTileInSlots;
  constexpr int PO = UPPER ? kTileOutSlots : 1;
  __shared__ T xs[(NT + 1) * B];  // tile row NT: zeros
  // overflow: ovf_in in-phase blocks (B rows of B values), ovf_out products
  // (B per slot), then the in-phase blocks' tile positions
  extern __shared__ __align__(16) unsigned char tile_lds[];
  T* ov_val = reinterpret_cast<T*>(tile_lds);
  T* ov_prod = ov_val + (size_t)ovf_in * B * B;
  int* ov_col = reinterpret_cast<int*>(ov_prod + (size_t)ovf_out * B);
  const int tile = tile0 + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int trow = tid / G, r = tid % G;
  const bool act = r < B;
  const int s = tile * (NTS / 64) + (tid >> 6);
  const size_t rix = (size_t)tile * NT + trow;
  const int row = rows[rix];
  const int lev = rlev[rix];
  const int nlev = depth[tile];
  const int k0 = __builtin_amdgcn_readfirstlane(sptr[s]);
  const int km = __builtin_amdgcn_readfirstlane(split[s]);
  const int k1 = __builtin_amdgcn_readfirstlane(sptr[s + 1]);
  const int ki = UPPER ? k0 : km, ni = UPPER ? km - k0 : k1 - km;  // in-phase part
  const int ko = UPPER ? km : k0, no = UPPER ? k1 - km : km - k0;  // the other phases
  const int4 ov = reinterpret_cast<const int4*>(rov)[rix];
  // the plan places a row's in-phase slots beyond kTileInSlots at overflow
  // slots [ov.x, ov.x + ov.y); this kernel holds PI in registers, so it
  // stages only the ones beyond PI, at their places in that range
  const int ovx = ov.x + (PI - kTileInSlots), ovn = ov.y > PI - kTileInSlots ? ov.y - (PI - kTileInSlots) : 0;
  if (tid < B) xs[NT * B + tid] = (T)0;
  T t = (T)0, d[B];
  int ci[PI];
  unsigned wi[PI][E];
  T so[PO];
  double dsum[1] = {0.0};
#pragma unroll
  for (int u = 0; u < PI; u++) ci[u] = NT;
  const bool live = row >= 0 && act;
  if (live) {
    t = rhs[(size_t)row * B + r];
#pragma unroll
    for (int u = 0; u < PI; u++) {
      if (u < ni) {
        const int c = col[(size_t)(ki + u) * 64 + lane];
        ci[u] = c >= 0 ? c : NT;
        load_words<B, TS>(val, (size_t)(ki + u), lane, wi[u]);
      }
    }
    if constexpr (!UPPER) load_block<B, TS, T>(dval, (size_t)s, lane, d);
    for (int u = 0; u < ovn; u++) {
      T a[B];
      load_block<B, TS, T>(val, (size_t)(ki + PI + u), lane, a);
      const int o = ovx + u;
#pragma unroll
      for (int e = 0; e < B; e++) ov_val[((size_t)o * B + r) * B + e] = a[e];
      if (r == 0) {
        const int c = col[(size_t)(ki + PI + u) * 64 + lane];
        ov_col[o] = c >= 0 ? c : NT;
      }
    }
    row_slots<B, TS, T, UPPER, PO>(col, val, out, ko, no, lane, t, so, ov_prod + r, ov.z, PO + ov.w);
  }
  __syncthreads();  // the zero row, the staged overflow
  for (int l = 0; l < nlev; l++) {
    const bool mine = row >= 0 && lev == l;  // the same for the G lanes of a row
    if (mine) {
      if (act) {
        T xv[PI][B];
#pragma unroll
        for (int u = 0; u < PI; u++)
#pragma unroll
          for (int e = 0; e < B; e++) xv[u][e] = xs[ci[u] * B + e];
#pragma unroll
        for (int u = 0; u < PI; u++) {
          if (u < ni) {
            T a[B];
            words_to_block<B, TS, T>(wi[u], a);
            T v = a[0] * xv[u][0];
#pragma unroll
            for (int e = 1; e < B; e++) v = v + a[e] * xv[u][e];
            t -= v;
          }
        }
        for (int u = 0; u < ovn; u++) {
          const int o = ovx + u;
          const int c = ov_col[o];
          T v = ov_val[((size_t)o * B + r) * B] * xs[c * B];
#pragma unroll
          for (int e = 1; e < B; e++) v = v + ov_val[((size_t)o * B + r) * B + e] * xs[c * B + e];
          t -= v;
        }
        if constexpr (UPPER) {
#pragma unroll
          for (int u = 0; u < PO; u++)
            if (u < no) t -= so[u];
          for (int u = 0; u < ov.w; u++) t -= ov_prod[((size_t)ov.z + u) * B + r];
        }
      }
      if constexpr (!UPPER) {
        // y_r = sum_c D[r][c] t_c, t_c from lane c of the row's group
        T tc[B];
#pragma unroll
        for (int c = 0; c < B; c++) tc[c] = group_lane<G, T>(t, c, lane);
        T y = d[0] * tc[0];
#pragma unroll
        for (int c = 1; c < B; c++) y = y + d[c] * tc[c];
        t = y;
      }
      if (act) xs[trow * B + r] = t;
    }
    __syncthreads();
  }
  if (live) {
    out[(size_t)row * B + r] = t;
    if (DOT) dsum[0] += (double)rvec[(size_t)row * B + r] * (double)t;
  }
  if constexpr (DOT) write_partials<1, NTS>(dsum, partials, tile, 0);
}

// Forward sweep of one level fused with the PCG update that precedes it
// (solver_ls_iter.cpp:140-145, then :158): for each row of the level
//   r_i += (-alpha) Ap_i ; x_i += alpha p_i ; z_i = Dinv_i (r_i - sum L_ij z_j).
// Ap arrives in z.  A row's own r, x, p and Ap are all the update needs and the
// substitution reads r only at its own row, so every row is updated just
// before it is solved and the separate vector pass disappears; rows of later
// levels still hold Ap in z, rows of earlier levels already hold their z.
// Partial <r,r> of the level's rows go to prr[part0 + workgroup].  With RZ,
// rows flagged in rzmask (those without U entries, which the backward sweep
// leaves unchanged) also add <r,z> to prz[part0 + workgroup], so the backward
// sweep need not visit them.
template <int B, typename TS, typename T, bool RZ>
__global__ void __launch_bounds__(kTrsvNT) k_trsv_lower_xr(int s0, int s1, const int* __restrict__ rows,
                                                      const int* __restrict__ sptr, const int* __restrict__ col,
                                                      const TS* __restrict__ val, const TS* __restrict__ dval,
                                                      const T* __restrict__ p, T* __restrict__ r,
                                                      T* __restrict__ x, T* z,
                                                      const unsigned char* __restrict__ rzmask,
                                                      double* __restrict__ prr, int part0,
                                                      double* __restrict__ prz, const KState* __restrict__ st) {
  constexpr int BB = B * B;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int s = s0 + xcd_block() * kTrsvWaves + wave;
  const int sl = s < s1 ? s : s1 - 1;
  const SliceHead h = slice_head(sl, lane, rows, sptr, st);
  if (h.done) return;
  const T a = (T)st->alpha, ma = (T)(-st->alpha);
  double drr[1] = {0.0}, drz[1] = {0.0};
  const int row = s < s1 ? h.row : -1;
  if (row >= 0) {
    const bool rz = RZ && rzmask[(size_t)s * 64 + lane];
    // the row's r, z (= Ap), x, p are loaded here and combined only after the
    // slot loads have left (init), and r, x stored at the end
    T rr[B], zz[B], xx[B], pp[B];
#pragma unroll
    for (int e = 0; e < B; e++) {
      const size_t o = (size_t)row * B + e;
      rr[e] = r[o];
      zz[e] = z[o];
      xx[e] = x[o];
      pp[e] = p[o];
    }
    T t[B], rin[B];
    auto init = [&](T(&tt)[B]) {
#pragma unroll
      for (int e = 0; e < B; e++) {
        // pins the arithmetic here: without it the compiler moves it ahead
        // of the slot loads, which then wait for r and z
        asm volatile("" : "+v"(rr[e]), "+v"(zz[e]));
        const T rv = rr[e] + ma * zz[e];
        tt[e] = rv;
        rin[e] = rv;
      }
    };
    sell_msub_range<B, TS, T>(col, val, z, h.k0, h.k1, lane, t, init);
    T d[BB], y[B];
    load_block<BB, TS, T>(dval, (size_t)s, lane, d);
    blk_mul<B, T>(y, d, t);
#pragma unroll
    for (int e = 0; e < B; e++) {
      const size_t o = (size_t)row * B + e;
      r[o] = rin[e];
      x[o] = xx[e] + a * pp[e];
      drr[0] += (double)rin[e] * (double)rin[e];
      z[o] = y[e];
      if (rz) drz[0] += (double)rin[e] * (double)y[e];
    }
  }
  write_partials<1, kTrsvNT>(drr, prr, part0 + blockIdx.x, 0);
  if constexpr (RZ) write_partials<1, kTrsvNT>(drz, prz, part0 + blockIdx.x, 0);
}

// Finalize of a PCG iteration whose <r,r> and <r,z> partials were left by
// different launches: the convergence test of solver_ls_iter.cpp:144-154 on
// sum(prr), then beta = <r,z>_new / rz (:157-161) on sum(prz) unless the
// iteration ended the solve.
static __global__ void __launch_bounds__(kFinNT) k_finalize_rrz(const double* __restrict__ prr, int nrr,
                                                        const double* __restrict__ prz, int nrz,
                                                        KState* __restrict__ st) {
  if (st->done) return;
  __shared__ double sh[kFinNT / 64];
  double t[2];
  {
    // both partial arrays in the same load rounds
    const double* p[2] = {prr, prz};
    const int n[2] = {nrr, nrz};
    strided_sums<kFinNT, 2>(p, n, t);
  }
  const double rr = block_sum_n<kFinNT>(t[0], sh);
  const double rz = block_sum_n<kFinNT>(t[1], sh);
  if (threadIdx.x != 0) return;
  st->rr = rr;
  if (rr * st->inv_rr0 < st->tol2) {
    st->done = 1;
    st->conv_iter = st->iter;
    st->conv_rr = rr;
  }
  st->iter += 1;
  if (!st->done && st->iter >= st->maxit) st->done = 2;
  if (st->done) return;
  st->rz_new = rz;
  st->beta = rz / st->rz;
  st->rz = rz;
}

// ---------------------------------------------------------------------------
// Krylov vector updates (LsSol::Solve_PCG / Solve_PBiCGSTAB, solver_ls_iter.cpp)
// PCG: r += (-alpha) z ; x += alpha p ; partial <r,r> (:140-145)
template <typename T>
__global__ void __launch_bounds__(kNT) k_pcg_update_xr(long n, const T* __restrict__ z, const T* __restrict__ p,
                                                      T* __restrict__ r, T* __restrict__ x,
                                                      double* __restrict__ partials, const KState* __restrict__ st) {
  if (st->done) return;
  const double alpha = st->alpha;
  const T a = (T)alpha, ma = (T)(-alpha);
  double dsum[1] = {0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    const T rv = r[i] + ma * z[i];
    r[i] = rv;
    x[i] = x[i] + a * p[i];
    dsum[0] += (double)rv * (double)rv;
  }
  write_partials<1>(dsum, partials, blockIdx.x, gridDim.x);
}
// PCG: p = beta p + z (:163-164)
template <typename T>
__global__ void __launch_bounds__(kNT) k_pcg_update_p(long n, const T* __restrict__ z, T* __restrict__ p,
                                                     const KState* __restrict__ st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) p[i] = beta * p[i] + z[i];
}
// BiCGSTAB: s = r + (-alpha) AMp (:355-357)
template <typename T>
__global__ void __launch_bounds__(kNT) k_bicg_s(long n, const T* __restrict__ r, const T* __restrict__ AMp,
                                               T* __restrict__ s, const KState* __restrict__ st) {
  if (st->done) return;
  const T ma = (T)(-st->alpha);
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) s[i] = r[i] + ma * AMp[i];
}
// BiCGSTAB: x += alpha Mp; x += omega Ms; r = s + (-omega) AMs; partial <r,r>, <r,r2> (:376-392)
template <typename T>
__global__ void __launch_bounds__(kNT) k_bicg_xr(long n, const T* __restrict__ Mp, const T* __restrict__ Ms,
                                                const T* __restrict__ s, const T* __restrict__ AMs,
                                                const T* __restrict__ r2, T* __restrict__ x, T* __restrict__ r,
                                                double* __restrict__ partials, const KState* __restrict__ st) {
  if (st->done) return;
  const T a = (T)st->alpha, w = (T)st->omega, mw = (T)(-st->omega);
  double dsum[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    T xv = x[i] + a * Mp[i];
    xv = xv + w * Ms[i];
    x[i] = xv;
    const T rv = s[i] + mw * AMs[i];
    r[i] = rv;
    dsum[0] += (double)rv * (double)rv;
    dsum[1] += (double)rv * (double)r2[i];
  }
  write_partials<2>(dsum, partials, blockIdx.x, gridDim.x);
}
// BiCGSTAB: p = beta p; p += r; p += (-beta*omega) AMp (:401-403)
template <typename T>
__global__ void __launch_bounds__(kNT) k_bicg_p(long n, const T* __restrict__ r, const T* __restrict__ AMp,
                                               T* __restrict__ p, const KState* __restrict__ st) {
  if (st->done) return;
  const T beta = (T)st->beta, mbw = (T)(-st->beta * st->omega);
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    T pv = beta * p[i];
    pv = pv + r[i];
    pv = pv + mbw * AMp[i];
    p[i] = pv;
  }
}

// ---------------------------------------------------------------------------
// Plain vector kernels
template <typename T>
__global__ void k_dot(long n, const T* __restrict__ x, const T* __restrict__ y, double* __restrict__ partials) {
  double dsum[1] = {0.0};
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    dsum[0] += (double)x[i] * (double)y[i];
  write_partials<1>(dsum, partials, blockIdx.x, gridDim.x);
}
static __global__ void k_axpy(long n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) y[i] += alpha * x[i];
}
static __global__ void k_scal(long n, double alpha, double* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) x[i] *= alpha;
}
static __global__ void k_zero_bc(long n, const int* __restrict__ flag, double* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT)
    if (flag[i] == 1) x[i] = 0.0;
}
// out[new] = in[order[new]] (OrderingVector_OldToNew, ordering_blk.cpp:151-166) with type conversion and scale
template <typename TI, typename TO>
__global__ void k_permute(int nb, int b, const int* __restrict__ order, const TI* __restrict__ in,
                          TO* __restrict__ out, double scale) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= (long)nb * b) return;
  const long i = t / b;
  const int r = (int)(t - i * b);
  const long src = order ? order[i] : i;
  out[t] = (TO)(scale * (double)in[(size_t)src * b + r]);
}
// out[order[new]] = in[new] (OrderingVector_NewToOld, ordering_blk.cpp:136-149)
template <typename T>
__global__ void k_unpermute(int nb, int b, const int* __restrict__ order, const T* __restrict__ in, T* __restrict__ out) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= (long)nb * b) return;
  const long i = t / b;
  const int r = (int)(t - i * b);
  const long dst = order ? order[i] : i;
  out[(size_t)dst * b + r] = in[t];
}
// Reliable update of the mixed-precision solvers: x64 += (double)xr; xr = 0.
template <typename T>
__global__ void k_reliable_update(long n, T* __restrict__ xr, double* __restrict__ x64) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) {
    x64[i] = x64[i] + (double)xr[i];
    xr[i] = (T)0;
  }
}
// y = (T)x
template <typename T>
__global__ void k_convert(long n, const double* __restrict__ x, T* __restrict__ y) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) y[i] = (T)x[i];
}
// Plain copy used to calibrate the bandwidth ceiling and the HBM counters.
static __global__ void __launch_bounds__(kNT) k_copy4(long n4, const float4* __restrict__ in, float4* __restrict__ out) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n4; i += (long)gridDim.x * kNT) out[i] = in[i];
}
// Streaming read of n4 16-byte words: eight non-temporal 16-byte loads in
// flight per lane (the fastest plain read measured on this part); the sum is
// stored only on an impossible value so that the loads stay live without a
// write stream.
static __global__ void __launch_bounds__(kNT) k_read4(long n4, const float4* __restrict__ in, float* __restrict__ out) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  constexpr int U = 8;
  const f32x4* p = reinterpret_cast<const f32x4*>(in);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const long stride = (long)gridDim.x * kNT;
  long i = (long)blockIdx.x * kNT + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    f32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(p + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; u++) acc += v[u];
  }
  for (; i < n4; i += stride) acc += __builtin_nontemporal_load(p + i);
  if (acc.x + acc.y + acc.z + acc.w == 12345.f) out[0] = 1.f;
}
template <typename T>
__global__ void k_fill(long n, T v, T* __restrict__ x) {
  for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < n; i += (long)gridDim.x * kNT) x[i] = v;
}

}  // namespace dfm
