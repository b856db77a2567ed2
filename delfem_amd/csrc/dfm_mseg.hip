// dfm_mseg.hip -- multi-segment linear systems (LsSol::CLinearSystem with
// several segments, src/ls/linearsystem.cpp): diagonal blocks are the
// library's block-CRS matrices (one block size per segment), off-diagonal
// blocks rectangular CMat_BlkCrs (dfm_bmat); vectors are the segments'
// vectors one after the other.  The preconditioner is CPreconditioner_ILU(k)
// across the segments (src/ls/preconditioner.cpp:57-335): per segment the
// natural-order ILU of its diagonal block, per off-diagonal block a
// CMatFrac_BlkCrs copy (fill patterns from the host symbolic below) updated
// by DoILUDecompUp / Low / LowUp
// (src/matvec/matfrac_blkcrs.cpp:540-783) and the diagonal blocks by
// CMatDiaFrac_BlkCrs::DoILUDecompLowUp (matdiafrac_blkcrs.cpp:665-767), in
// the reference's block order and operation order.  The Krylov loops
// (LsSol::Solve_PCG / Solve_PBiCGSTAB, src/ls/solver_ls_iter.cpp) run on the
// host over device kernels; the dot products add the segments' sums in
// segment order as CLinearSystem::DOT does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "dfm_kernels.h"

using namespace dfm;

struct dfm_mseg_s {
  unsigned nseg = 0;
  std::vector<unsigned> nb, b;
  std::vector<size_t> off;                          // dof offset of each segment (nseg + 1)
  std::vector<dfm_bcrs_s*> dia;                     // borrowed
  std::vector<std::vector<dfm_bmat_s*>> nd;         // borrowed, null where there is no block
  // DOT: per-segment partial sums and results, read back with one copy
  DevBuf<double> dot_partials;
  DevBuf<KState> dot_state;
  KState* dot_host = nullptr;  // pinned readback of the per-segment dots, dot_host_n slots
  unsigned dot_host_n = 0;
  ~dfm_mseg_s() {
    if (dot_host) (void)hipHostFree(dot_host);
  }
};

struct dfm_msilu_s {
  dfm_mseg_s* S = nullptr;
  int lev_fill = 0;
  std::vector<dfm_ilu_s*> dia;                      // owned: natural order, pattern with cross-segment fill
  std::vector<std::vector<dfm_bmat_s*>> F;          // owned: off-diagonal factor blocks (block pattern + fill)
  std::vector<std::vector<std::vector<unsigned>>> src_col;  // the blocks' patterns the factor was built on
  std::vector<std::vector<DevBuf<long long>>> src;  // factor entry -> block entry, -1 for fill
  std::vector<double*> work;                        // Krylov vectors
  DevBuf<unsigned char> dstate;                     // device-driven loops: their scalars (MsState)
  void* hstate = nullptr;                           // pinned copy, polled
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;  // their iteration blocks, by operand set
  ~dfm_msilu_s() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (hstate) (void)hipHostFree(hstate);
    for (auto* m : dia)
      if (m) dfm_ilu_destroy(m);
    for (auto& row : F)
      for (auto* f : row)
        if (f) dfm_bmat_destroy(f);
    for (double* w : work)
      if (w) dfm_free(w);
  }
};

constexpr int kMaxSeg = 16;  // segments of one system

#define DFM_MLAUNCH(kernel, grid, ...)                                                      \
  do {                                                                                      \
    if ((grid) > 0) {                                                                       \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(kNT), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                           \
    }                                                                                       \
  } while (0)

namespace {

inline int mgrid(long n) { return (int)std::max<long>(1, (n + kNT - 1) / kNT); }
// the grid dfm_vec_dot uses, so a segment's dot is the same sum
inline int grid_vec(long n) { return (int)std::min<long>(std::max<long>(1, (n + kNT - 1) / kNT), 2048); }

// CalcSubMatPr (matfrac_blkcrs.cpp:43-54): out -= a b, one product at a time
__device__ inline void msub(double* out, const double* a, const double* c, int ni, int nk, int nj) {
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++) out[i * nj + j] -= a[i * nk + k] * c[k * nj + j];
}
// CalcMatPr (:56-69): out = d out through a temporary
__device__ inline void mmul(double* out, const double* d, int ni, int nj) {
  double tmp[64];
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++) {
      tmp[i * nj + j] = 0.0;
      for (int k = 0; k < ni; k++) tmp[i * nj + j] += d[i * ni + k] * out[k * nj + j];
    }
  for (int i = 0; i < ni * nj; i++) out[i] = tmp[i];
}

// DoILUDecompUp (matfrac_blkcrs.cpp:601-701) for the rows of one lower level
// of factor i: U_ij row r -= L_rk U_ij row k, then scaled by the inverted pivot
__global__ void __launch_bounds__(kNT) k_ms_up(int nrow, const int* __restrict__ rows, const int* __restrict__ frow,
                                              const int* __restrict__ fcol, const int* __restrict__ fdiaind,
                                              const double* __restrict__ fval, const double* __restrict__ fdia,
                                              const int* __restrict__ rp, const int* __restrict__ cl,
                                              double* __restrict__ F, int bi, int bj) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t >= nrow) return;
  const int r = rows[t];
  const int bs = bi * bj, bii = bi * bi;
  for (int ik = frow[r]; ik < fdiaind[r]; ik++) {
    const int k = fcol[ik];
    for (int kj = rp[k]; kj < rp[k + 1]; kj++) {
      const int ij = find_col(cl, rp[r], rp[r + 1], cl[kj]);
      if (ij < 0) continue;
      msub(F + (long)ij * bs, fval + (long)ik * bii, F + (long)kj * bs, bi, bi, bj);
    }
  }
  for (int q = rp[r]; q < rp[r + 1]; q++) mmul(F + (long)q * bs, fdia + (long)r * bii, bi, bj);
}
// DoILUDecompLow (:703-783): L_ij row r -= L_ij(r, k) U_j row k, rows independent
__global__ void __launch_bounds__(kNT) k_ms_low(int nrow, const int* __restrict__ rp, const int* __restrict__ cl,
                                               double* __restrict__ F, const int* __restrict__ frow,
                                               const int* __restrict__ fcol, const int* __restrict__ fdiaind,
                                               const double* __restrict__ fval, int bi, int bj) {
  const int r = blockIdx.x * kNT + threadIdx.x;
  if (r >= nrow) return;
  const int bs = bi * bj, bjj = bj * bj;
  for (int ik = rp[r]; ik < rp[r + 1]; ik++) {
    const int k = cl[ik];
    for (int kj = fdiaind[k]; kj < frow[k + 1]; kj++) {
      const int ij = find_col(cl, rp[r], rp[r + 1], fcol[kj]);
      if (ij < 0) continue;
      msub(F + (long)ij * bs, F + (long)ik * bs, fval + (long)kj * bjj, bi, bj, bj);
    }
  }
}
// CMatFrac_BlkCrs::DoILUDecompLowUp (:540-599): T_ij -= L_ik U_kj on T's pattern
__global__ void __launch_bounds__(kNT) k_ms_lowup(int nrow, const int* __restrict__ trp, const int* __restrict__ tcl,
                                                 double* __restrict__ T, const int* __restrict__ lrp,
                                                 const int* __restrict__ lcl, const double* __restrict__ Lv,
                                                 const int* __restrict__ urp, const int* __restrict__ ucl,
                                                 const double* __restrict__ Uv, int bi, int bk, int bj) {
  const int r = blockIdx.x * kNT + threadIdx.x;
  if (r >= nrow) return;
  for (int ik = lrp[r]; ik < lrp[r + 1]; ik++) {
    const int k = lcl[ik];
    for (int kj = urp[k]; kj < urp[k + 1]; kj++) {
      const int ij = find_col(tcl, trp[r], trp[r + 1], ucl[kj]);
      if (ij < 0) continue;
      msub(T + (long)ij * bi * bj, Lv + (long)ik * bi * bk, Uv + (long)kj * bk * bj, bi, bk, bj);
    }
  }
}
// CMatDiaFrac_BlkCrs::DoILUDecompLowUp (matdiafrac_blkcrs.cpp:665-767): the
// copied diagonal block of segment i -= L_ik U_ki before its factorisation
__global__ void __launch_bounds__(kNT) k_ms_lowup_dia(int nrow, int row0, const int* __restrict__ frow,
                                                     const int* __restrict__ fcol, double* __restrict__ fval,
                                                     double* __restrict__ fdia, const int* __restrict__ lrp,
                                                     const int* __restrict__ lcl, const double* __restrict__ Lv,
                                                     const int* __restrict__ urp, const int* __restrict__ ucl,
                                                     const double* __restrict__ Uv, int bi, int bk) {
  const int r = row0 + (int)(blockIdx.x * kNT + threadIdx.x);
  if (r >= nrow) return;
  const int bii = bi * bi;
  for (int ik = lrp[r]; ik < lrp[r + 1]; ik++) {
    const int k = lcl[ik];
    for (int kj = urp[k]; kj < urp[k + 1]; kj++) {
      const int j = ucl[kj];
      const double* a = Lv + (long)ik * bi * bk;
      const double* c = Uv + (long)kj * bk * bi;
      if (j != r) {
        const int ij = find_col(fcol, frow[r], frow[r + 1], j);
        if (ij < 0) continue;
        msub(fval + (long)ij * bii, a, c, bi, bk, bi);
      } else {
        msub(fdia + (long)r * bii, a, c, bi, bk, bi);
      }
    }
  }
}

// One wave per row for the four kernels above (rows of at most kMsWaveRow
// entries, blocks of at most kMsWaveBlk values): the target row is staged in
// LDS, the products of one step (one L entry: its partner row's entries,
// each to a different target entry) spread over the lanes, the steps in
// order, so every target entry receives the same subtractions in the same
// order (a wave's LDS accesses are performed in program order).
constexpr int kMsWaveRow = 64, kMsWaveBlk = 9, kMsWaves = 4;
__device__ inline int lds_find(const int* sc, int n, int j) {
  for (int q = 0; q < n; q++)
    if (sc[q] == j) return q;
  return -1;
}
struct MsRow {  // one wave's LDS: the row's blocks, its diagonal block, its columns
  double v[kMsWaveRow * kMsWaveBlk + kMsWaveBlk];
  int c[kMsWaveRow];
};
// stage n blocks of bs values from val[q0..] and their columns, and (dia) a bs block
__device__ inline void ms_stage(MsRow& R, int lane, const double* val, const int* col, long q0, int n, int bs,
                                const double* dia) {
  for (int e = lane; e < n * bs; e += 64) R.v[e] = val[q0 * bs + e];
  for (int e = lane; e < n; e += 64) R.c[e] = col[q0 + e];
  if (dia && lane < bs) R.v[kMsWaveRow * kMsWaveBlk + lane] = dia[lane];
  __builtin_amdgcn_wave_barrier();
}
__device__ inline void ms_unstage(const MsRow& R, int lane, double* val, long q0, int n, int bs, double* dia) {
  __builtin_amdgcn_wave_barrier();
  for (int e = lane; e < n * bs; e += 64) val[q0 * bs + e] = R.v[e];
  if (dia && lane < bs) dia[lane] = R.v[kMsWaveRow * kMsWaveBlk + lane];
}
#define MS_WAVE_ROW(nrow)                                    \
  __shared__ MsRow rows_lds[kMsWaves];                       \
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63; \
  const int t = blockIdx.x * kMsWaves + wv;                  \
  if (t >= (nrow)) return;                                   \
  MsRow& R = rows_lds[wv];

__global__ void __launch_bounds__(64 * kMsWaves) k_ms_up_w(int nrow, const int* __restrict__ rows,
                                                           const int* __restrict__ frow, const int* __restrict__ fcol,
                                                           const int* __restrict__ fdiaind,
                                                           const double* __restrict__ fval,
                                                           const double* __restrict__ fdia, const int* __restrict__ rp,
                                                           const int* __restrict__ cl, double* __restrict__ F, int bi,
                                                           int bj) {
  MS_WAVE_ROW(nrow)
  const int r = __builtin_amdgcn_readfirstlane(rows[t]);
  const int bs = bi * bj, bii = bi * bi;
  const int q0 = __builtin_amdgcn_readfirstlane(rp[r]), n = __builtin_amdgcn_readfirstlane(rp[r + 1]) - q0;
  ms_stage(R, lane, F, cl, q0, n, bs, nullptr);
  const int i0 = __builtin_amdgcn_readfirstlane(frow[r]), i1 = __builtin_amdgcn_readfirstlane(fdiaind[r]);
  for (int ik = i0; ik < i1; ik++) {
    const int k = __builtin_amdgcn_readfirstlane(fcol[ik]);
    const int k0 = __builtin_amdgcn_readfirstlane(rp[k]), k1 = __builtin_amdgcn_readfirstlane(rp[k + 1]);
    for (int kj = k0 + lane; kj < k1; kj += 64) {
      const int ij = lds_find(R.c, n, cl[kj]);
      if (ij < 0) continue;
      msub(R.v + ij * bs, fval + (long)ik * bii, F + (long)kj * bs, bi, bi, bj);
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int q = lane; q < n; q += 64) mmul(R.v + q * bs, fdia + (long)r * bii, bi, bj);
  ms_unstage(R, lane, F, q0, n, bs, nullptr);
}
__global__ void __launch_bounds__(64 * kMsWaves) k_ms_low_w(int nrow, const int* __restrict__ rp,
                                                            const int* __restrict__ cl, double* __restrict__ F,
                                                            const int* __restrict__ frow,
                                                            const int* __restrict__ fcol,
                                                            const int* __restrict__ fdiaind,
                                                            const double* __restrict__ fval, int bi, int bj) {
  MS_WAVE_ROW(nrow)
  const int r = t;
  const int bs = bi * bj, bjj = bj * bj;
  const int q0 = __builtin_amdgcn_readfirstlane(rp[r]), n = __builtin_amdgcn_readfirstlane(rp[r + 1]) - q0;
  ms_stage(R, lane, F, cl, q0, n, bs, nullptr);
  for (int ik = 0; ik < n; ik++) {
    const int k = R.c[ik];
    const int k0 = __builtin_amdgcn_readfirstlane(fdiaind[k]), k1 = __builtin_amdgcn_readfirstlane(frow[k + 1]);
    for (int kj = k0 + lane; kj < k1; kj += 64) {
      const int ij = lds_find(R.c, n, fcol[kj]);
      if (ij < 0) continue;
      msub(R.v + ij * bs, R.v + ik * bs, fval + (long)kj * bjj, bi, bj, bj);
    }
    __builtin_amdgcn_wave_barrier();
  }
  ms_unstage(R, lane, F, q0, n, bs, nullptr);
}
__global__ void __launch_bounds__(64 * kMsWaves) k_ms_lowup_w(int nrow, const int* __restrict__ trp,
                                                              const int* __restrict__ tcl, double* __restrict__ T,
                                                              const int* __restrict__ lrp,
                                                              const int* __restrict__ lcl,
                                                              const double* __restrict__ Lv,
                                                              const int* __restrict__ urp,
                                                              const int* __restrict__ ucl,
                                                              const double* __restrict__ Uv, int bi, int bk, int bj) {
  MS_WAVE_ROW(nrow)
  const int r = t;
  const int bs = bi * bj;
  const int q0 = __builtin_amdgcn_readfirstlane(trp[r]), n = __builtin_amdgcn_readfirstlane(trp[r + 1]) - q0;
  ms_stage(R, lane, T, tcl, q0, n, bs, nullptr);
  const int i0 = __builtin_amdgcn_readfirstlane(lrp[r]), i1 = __builtin_amdgcn_readfirstlane(lrp[r + 1]);
  for (int ik = i0; ik < i1; ik++) {
    const int k = __builtin_amdgcn_readfirstlane(lcl[ik]);
    const int k0 = __builtin_amdgcn_readfirstlane(urp[k]), k1 = __builtin_amdgcn_readfirstlane(urp[k + 1]);
    for (int kj = k0 + lane; kj < k1; kj += 64) {
      const int ij = lds_find(R.c, n, ucl[kj]);
      if (ij < 0) continue;
      msub(R.v + ij * bs, Lv + (long)ik * bi * bk, Uv + (long)kj * bk * bj, bi, bk, bj);
    }
    __builtin_amdgcn_wave_barrier();
  }
  ms_unstage(R, lane, T, q0, n, bs, nullptr);
}
__global__ void __launch_bounds__(64 * kMsWaves) k_ms_lowup_dia_w(
    int nrow, int row0, const int* __restrict__ frow, const int* __restrict__ fcol, double* __restrict__ fval,
    double* __restrict__ fdia, const int* __restrict__ lrp, const int* __restrict__ lcl,
    const double* __restrict__ Lv, const int* __restrict__ urp, const int* __restrict__ ucl,
    const double* __restrict__ Uv, int bi, int bk) {
  MS_WAVE_ROW(nrow - row0)
  const int r = row0 + t;
  const int bii = bi * bi;
  const int q0 = __builtin_amdgcn_readfirstlane(frow[r]), n = __builtin_amdgcn_readfirstlane(frow[r + 1]) - q0;
  ms_stage(R, lane, fval, fcol, q0, n, bii, fdia + (long)r * bii);
  double* sD = R.v + kMsWaveRow * kMsWaveBlk;
  const int i0 = __builtin_amdgcn_readfirstlane(lrp[r]), i1 = __builtin_amdgcn_readfirstlane(lrp[r + 1]);
  for (int ik = i0; ik < i1; ik++) {
    const int k = __builtin_amdgcn_readfirstlane(lcl[ik]);
    const int k0 = __builtin_amdgcn_readfirstlane(urp[k]), k1 = __builtin_amdgcn_readfirstlane(urp[k + 1]);
    const double* a = Lv + (long)ik * bi * bk;
    for (int kj = k0 + lane; kj < k1; kj += 64) {
      const int j = ucl[kj];
      const double* c = Uv + (long)kj * bk * bi;
      if (j != r) {
        const int ij = lds_find(R.c, n, j);
        if (ij < 0) continue;
        msub(R.v + ij * bii, a, c, bi, bk, bi);
      } else {
        msub(sD, a, c, bi, bk, bi);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  ms_unstage(R, lane, fval, q0, n, bii, fdia + (long)r * bii);
}
#undef MS_WAVE_ROW

// SetValue_Initialize of an off-diagonal factor block (CMat_BlkCrs::SetValue,
// mat_blkcrs.cpp:557-694): the block's values where it has the entry, 0 on fill
__global__ void __launch_bounds__(kNT) k_ms_init(long n, int bs, const long long* __restrict__ src,
                                                const double* __restrict__ bval, double* __restrict__ fval) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= n) return;
  const long q = t / bs, e = t % bs;
  const long long s = src[q];
  fval[t] = s >= 0 ? bval[s * bs + e] : 0.0;
}

// ---- symbolic factorisation across segments (host) ---------------------------
// A block pattern with fill levels (the CRowLev lists of the reference's
// CMatFrac_BlkCrs / CMatDiaFrac_BlkCrs); diagonal patterns hold no diagonal.
struct LevPat {
  unsigned nrow = 0;
  std::vector<unsigned> rp, cl;
  std::vector<int> lev;
  void from(unsigned n, const std::vector<unsigned>& r, const std::vector<unsigned>& c) {
    nrow = n;
    rp = r;
    cl = c;
    lev.assign(c.size(), 0);
  }
  bool full(unsigned ncol) const { return (unsigned long long)cl.size() == (unsigned long long)nrow * ncol; }
  void diaind(std::vector<unsigned>& d) const {
    d.resize(nrow);
    for (unsigned r = 0; r < nrow; r++)
      d[r] = (unsigned)(std::upper_bound(cl.begin() + rp[r], cl.begin() + rp[r + 1], r) - cl.begin());
  }
};
// One row being extended, as a sorted linked list: an inserted column keeps
// the level it was first given (CListRLN::Insert, matfrac_blkcrs.h:153-255).
struct RowList {
  struct Node {
    unsigned c;
    int lev, next;
  };
  std::vector<Node> v;
  int head = -1;
  void init(const LevPat& P, unsigned r) {
    v.clear();
    for (unsigned q = P.rp[r]; q < P.rp[r + 1]; q++) v.push_back({P.cl[q], P.lev[q], (int)v.size() + 1});
    head = v.empty() ? -1 : 0;
    if (!v.empty()) v.back().next = -1;
  }
  void insert(unsigned j, int lev, int& cur) {
    if (head < 0 || j < v[head].c) {
      v.push_back({j, lev, head});
      head = cur = (int)v.size() - 1;
      return;
    }
    int c = (cur < 0 || v[cur].c > j) ? head : cur;
    while (v[c].next >= 0 && v[v[c].next].c <= j) c = v[c].next;
    if (v[c].c == j) {
      cur = c;
      return;
    }
    v.push_back({j, lev, v[c].next});
    v[c].next = (int)v.size() - 1;
    cur = v[c].next;
  }
  void emit(LevPat& P) const {
    for (int t = head; t >= 0; t = v[t].next) {
      P.cl.push_back(v[t].c);
      P.lev.push_back(v[t].lev);
    }
    P.rp.push_back((unsigned)P.cl.size());
  }
};
inline bool lev_skip(int l, int K) { return K != -1 && l + 1 > K; }
inline LevPat lp_start(unsigned nrow, size_t hint) {
  LevPat P;
  P.nrow = nrow;
  P.rp.reserve(nrow + 1);
  P.rp.push_back(0);
  P.cl.reserve(hint);
  P.lev.reserve(hint);
  return P;
}
// AddFracPtnLowUp: T gains L_ik U_kj fill (matfrac_blkcrs.cpp:156-249; the
// diagonal-block form, matdiafrac_blkcrs.cpp:1105-1247, leaves out column r)
void lp_lowup(LevPat& T, unsigned ncol, const LevPat& L, const LevPat& U, int K, bool skip_diag, RowList& rl) {
  if (T.full(ncol) || K == 0) return;
  LevPat P = lp_start(T.nrow, T.cl.size() * 2);
  for (unsigned r = 0; r < T.nrow; r++) {
    rl.init(T, r);
    for (unsigned ik = L.rp[r]; ik < L.rp[r + 1]; ik++) {
      const int lik = L.lev[ik];
      if (lev_skip(lik, K)) continue;
      const unsigned k = L.cl[ik];
      int cur = rl.head;
      for (unsigned kj = U.rp[k]; kj < U.rp[k + 1]; kj++) {
        const int lkj = U.lev[kj];
        if (lev_skip(lkj, K) || (skip_diag && U.cl[kj] == r)) continue;
        rl.insert(U.cl[kj], std::max(lik, lkj) + 1, cur);
      }
    }
    rl.emit(P);
  }
  T = std::move(P);
}
// AddFracPtnUp (matfrac_blkcrs.cpp:251-362): U_ij row r gains L_ii(r, k)
// U_ij(k, :) fill from the rows k < r already extended
void lp_up(LevPat& T, unsigned ncol, const LevPat& D, int K, RowList& rl) {
  if (T.full(ncol) || K == 0) return;
  std::vector<unsigned> dind;
  D.diaind(dind);
  LevPat P = lp_start(T.nrow, T.cl.size() * 2);
  for (unsigned r = 0; r < T.nrow; r++) {
    rl.init(T, r);
    for (unsigned ik = D.rp[r]; ik < dind[r]; ik++) {
      const int lik = D.lev[ik];
      if (lev_skip(lik, K)) continue;
      const unsigned k = D.cl[ik];
      int cur = rl.head;
      for (unsigned kj = P.rp[k]; kj < P.rp[k + 1]; kj++) {
        const int lkj = P.lev[kj];
        if (lev_skip(lkj, K)) continue;
        rl.insert(P.cl[kj], std::max(lik, lkj) + 1, cur);
      }
    }
    rl.emit(P);
  }
  T = std::move(P);
}
// AddFracPtnLow (matfrac_blkcrs.cpp:364-495): L_ij row r, walked in column
// order while it grows, gains L_ij(r, k) U_jj(k, :) fill
void lp_low(LevPat& T, unsigned ncol, const LevPat& D, int K, RowList& rl) {
  if (T.full(ncol) || K == 0) return;
  std::vector<unsigned> dind;
  D.diaind(dind);
  LevPat P = lp_start(T.nrow, T.cl.size() * 2);
  for (unsigned r = 0; r < T.nrow; r++) {
    rl.init(T, r);
    for (int t = rl.head; t >= 0; t = rl.v[t].next) {
      const int lik = rl.v[t].lev;
      if (lev_skip(lik, K)) continue;
      const unsigned k = rl.v[t].c;
      int cur = t;
      for (unsigned kj = dind[k]; kj < D.rp[k + 1]; kj++) {
        const int lkj = D.lev[kj];
        if (lev_skip(lkj, K)) continue;
        rl.insert(D.cl[kj], std::max(lik, lkj) + 1, cur);
      }
    }
    rl.emit(P);
  }
  T = std::move(P);
}
// AddFracPtn of a diagonal block (matdiafrac_blkcrs.cpp:1249-1377) from the
// levels it already holds
void lp_dia_fill(LevPat& D, int K, RowList& rl) {
  const unsigned nb = D.nrow;
  if (D.cl.empty() || D.full(nb - 1) || K == 0) return;
  LevPat P = lp_start(nb, D.cl.size() * 4);
  std::vector<unsigned> dind(nb);
  for (unsigned r = 0; r < nb; r++) {
    rl.init(D, r);
    for (int t = rl.head; t >= 0; t = rl.v[t].next) {
      const int lik = rl.v[t].lev;
      if (lev_skip(lik, K)) continue;
      const unsigned k = rl.v[t].c;
      if (k >= r) break;
      int cur = t;
      for (unsigned kj = dind[k]; kj < P.rp[k + 1]; kj++) {
        const int lkj = P.lev[kj];
        if (lev_skip(lkj, K) || P.cl[kj] == r) continue;
        rl.insert(P.cl[kj], std::max(lik, lkj) + 1, cur);
      }
    }
    rl.emit(P);
    dind[r] = (unsigned)(std::upper_bound(P.cl.begin() + P.rp[r], P.cl.end(), r) - P.cl.begin());
  }
  D = std::move(P);
}

// CLinearSystem::DOT: each segment's dot product (the same two-phase
// reduction as dfm_vec_dot), read back together, added in segment order
int seg_dot(dfm_mseg_s* S, const double* x, const double* y, double* out) {
  const unsigned n = S->nseg;
  std::vector<int> g(n);
  int gmax = 1;
  for (unsigned i = 0; i < n; i++) {
    g[i] = grid_vec((long)(S->off[i + 1] - S->off[i]));
    gmax = std::max(gmax, g[i]);
  }
  if (S->dot_partials.n < (size_t)gmax * n) DFM_TRY(S->dot_partials.alloc((size_t)gmax * n));
  if (S->dot_state.n < n) {
    DFM_TRY(S->dot_state.alloc(n));
    DFM_HIP(hipMemsetAsync(S->dot_state.p, 0, sizeof(KState) * n, stream()));
  }
  if (S->dot_host_n < n) {
    if (S->dot_host) DFM_HIP(hipHostFree(S->dot_host));
    S->dot_host = nullptr;
    S->dot_host_n = 0;
    DFM_HIP(hipHostMalloc((void**)&S->dot_host, sizeof(KState) * n, hipHostMallocDefault));
    S->dot_host_n = n;
  }
  for (unsigned i = 0; i < n; i++) {
    double* part = S->dot_partials.p + (size_t)i * gmax;
    DFM_MLAUNCH(k_dot<double>, g[i], (long)(S->off[i + 1] - S->off[i]), x + S->off[i], y + S->off[i], part);
    hipLaunchKernelGGL((k_finalize<1, FIN_STORE>), dim3(1), dim3(kFinNT), 0, stream(), part, g[i],
                       S->dot_state.p + i);
    DFM_HIP(hipGetLastError());
  }
  DFM_HIP(hipMemcpyAsync(S->dot_host, S->dot_state.p, sizeof(KState) * n, hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  double s = 0.0;
  for (unsigned i = 0; i < n; i++) s += S->dot_host[i].d0;
  *out = s;
  return DFM_OK;
}
// CLinearSystem::AXPY returns without touching y when alpha == 0 (linearsystem.cpp:302)
int ls_axpy(size_t n, double alpha, const double* x, double* y) {
  if (alpha == 0) return DFM_OK;
  return dfm_vec_axpy(n, alpha, x, y);
}

}  // namespace

int dfm_mseg_create(unsigned nseg, const unsigned* nblk, const unsigned* len, dfm_mseg** out) {
  DFM_CHECK(out && nblk && len && nseg > 0 && nseg <= (unsigned)kMaxSeg, DFM_BAD_ARG,
            "bad arguments to dfm_mseg_create");
  DFM_TRY(ensure_init());
  std::unique_ptr<dfm_mseg_s> S(new dfm_mseg_s);
  S->nseg = nseg;
  S->off.assign(1, 0);
  for (unsigned i = 0; i < nseg; i++) {
    DFM_CHECK(len[i] > 0 && len[i] <= 8, DFM_BAD_ARG, "segment %u: block size %u", i, len[i]);
    S->nb.push_back(nblk[i]);
    S->b.push_back(len[i]);
    S->off.push_back(S->off.back() + (size_t)nblk[i] * len[i]);
  }
  S->dia.assign(nseg, nullptr);
  S->nd.assign(nseg, std::vector<dfm_bmat_s*>(nseg, nullptr));
  *out = S.release();
  return DFM_OK;
}
int dfm_mseg_destroy(dfm_mseg* S) {
  delete S;
  return DFM_OK;
}
int dfm_mseg_set_dia(dfm_mseg* S, unsigned i, dfm_bcrs* A) {
  DFM_CHECK(S && A && i < S->nseg && A->nb == S->nb[i] && A->b == S->b[i], DFM_BAD_ARG,
            "dfm_mseg_set_dia: segment %u does not match the matrix", i);
  S->dia[i] = A;
  return DFM_OK;
}
int dfm_mseg_set_nondia(dfm_mseg* S, unsigned i, unsigned j, dfm_bmat* B) {
  DFM_CHECK(S && B && i < S->nseg && j < S->nseg && i != j && B->nb == S->nb[i] && B->b == S->b[i] &&
                B->nbr == S->nb[j] && B->br == S->b[j],
            DFM_BAD_ARG, "dfm_mseg_set_nondia: block (%u, %u) does not match the segments", i, j);
  S->nd[i][j] = B;
  return DFM_OK;
}
int dfm_mseg_matvec(dfm_mseg* S, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(S && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_mseg_matvec");
  if (alpha == 0.0) return dfm_vec_scal(S->off[S->nseg], beta, y);
  for (unsigned i = 0; i < S->nseg; i++) {
    DFM_CHECK(S->dia[i], DFM_BAD_ARG, "segment %u has no diagonal block", i);
    DFM_TRY(dfm_bcrs_matvec(S->dia[i], alpha, x + S->off[i], beta, y + S->off[i]));
    for (unsigned j = 0; j < S->nseg; j++)
      if (S->nd[i][j]) DFM_TRY(dfm_bmat_matvec(S->nd[i][j], alpha, x + S->off[j], 1.0, y + S->off[i]));
  }
  return DFM_OK;
}
int dfm_mseg_dot(dfm_mseg* S, const double* x, const double* y, double* out) {
  DFM_CHECK(S && x && y && out, DFM_BAD_ARG, "bad arguments to dfm_mseg_dot");
  return seg_dot(S, x, y, out);
}

// CPreconditioner_ILU(ls, lev_fill) across the segments: the symbolic part
// of SetLinearSystem (preconditioner.cpp:57-211).  Every pattern starts as
// its block's with level 0; for lev_fill != 0 they are extended in (i, j)
// order -- diagonal block: LowUp with each k < i, then its own fill; above
// the diagonal: LowUp with k < i, then Up with factor i; below: LowUp with
// k < j, then Low with factor j.
int dfm_mseg_ilu_create(dfm_mseg* S, int lev_fill, dfm_msilu** out) {
  DFM_CHECK(S && out && lev_fill >= -1, DFM_BAD_ARG, "bad arguments to dfm_mseg_ilu_create");
  std::vector<int> alev((size_t)S->nseg * S->nseg, lev_fill);
  return dfm_mseg_ilu_create_levels(S, alev.data(), nullptr, 0, out);
}

// The same with a fill level per block (i, j) (row-major nseg x nseg), what
// SetFillInLevel(lev, ilss) calls resolve to (preconditioner.cpp:108-122);
// all levels 0 and no fill blocks: no fill (:105, 124-130).  fill_blk:
// SetFillBlk's rows of segment 0, which take fill of any level in block
// (0, 0) (AddFracPtn(lev, aBlkFill), :143-148).  A block the fill would need
// but the system lacks stays out when its own level is 0 (the reference
// creates it empty, :171-185).
int dfm_mseg_ilu_create_levels(dfm_mseg* S, const int* alev, const unsigned* fill_blk, unsigned nfill,
                               dfm_msilu** out) {
  DFM_CHECK(S && out && alev && (nfill == 0 || fill_blk), DFM_BAD_ARG,
            "bad arguments to dfm_mseg_ilu_create_levels");
  const unsigned n = S->nseg;
  bool any = nfill > 0;
  for (unsigned t = 0; t < n * n; t++) {
    DFM_CHECK(alev[t] >= -1, DFM_BAD_ARG, "fill level %d of block %u below -1", alev[t], t);
    any = any || alev[t] != 0;
  }
  const int lev_fill = alev[0];
  std::unique_ptr<dfm_msilu_s> M(new dfm_msilu_s);
  M->S = S;
  M->lev_fill = lev_fill;
  M->dia.assign(n, nullptr);
  M->F.assign(n, std::vector<dfm_bmat_s*>(n, nullptr));
  M->src_col.assign(n, std::vector<std::vector<unsigned>>(n));
  M->src.resize(n);
  for (auto& row : M->src) row = std::vector<DevBuf<long long>>(n);
  std::vector<LevPat> D(n);
  std::vector<std::vector<LevPat>> F(n, std::vector<LevPat>(n));
  for (unsigned i = 0; i < n; i++) {
    DFM_CHECK(S->dia[i], DFM_BAD_ARG, "segment %u has no diagonal block", i);
    D[i].from(S->nb[i], S->dia[i]->rowptr, S->dia[i]->col);
    for (unsigned j = 0; j < n; j++)
      if (S->nd[i][j]) F[i][j].from(S->nb[i], S->nd[i][j]->rowptr, S->nd[i][j]->col);
  }
  if (any) {
    RowList rl;
    for (unsigned i = 0; i < n; i++)
      for (unsigned j = 0; j < n; j++) {
        const int K = alev[i * n + j];
        if (i == j) {
          for (unsigned k = 0; k < i; k++)
            if (S->nd[i][k] && S->nd[k][i]) lp_lowup(D[i], S->nb[i] - 1, F[i][k], F[k][i], K, true, rl);
          if (i == 0 && nfill > 0) {
            // block (0, 0) still holds A_00's pattern at level 0: AddFracPtn(K, aBlkFill) is the
            // fill-block symbolic of that pattern
            const std::vector<unsigned> blk(fill_blk, fill_blk + nfill);
            std::vector<unsigned> fr, fc, fd;
            std::vector<int> fl;
            DFM_TRY(iluk_symbolic(S->nb[0], S->dia[0]->rowptr, S->dia[0]->col, K, nullptr, fr, fc, fl, fd, &blk));
            D[0].rp = std::move(fr);
            D[0].cl = std::move(fc);
            D[0].lev = std::move(fl);
            continue;
          }
          lp_dia_fill(D[i], K, rl);
          continue;
        }
        const unsigned kmax = std::min(i, j);
        for (unsigned k = 0; k < kmax + 1; k++) {
          const bool have = S->nd[i][j] != nullptr;
          if (k == i && k < j) {
            if (have) lp_up(F[i][j], S->nb[j], D[k], K, rl);
            continue;
          }
          if (k == j && k < i) {
            if (have) lp_low(F[i][j], S->nb[j], D[k], K, rl);
            continue;
          }
          if (!S->nd[i][k] || !S->nd[k][j]) continue;
          if (!have && K == 0) continue;
          DFM_CHECK(have, DFM_BAD_ARG, "fill needs block (%u, %u), which the system does not have", i, j);
          lp_lowup(F[i][j], S->nb[j], F[i][k], F[k][j], K, false, rl);
        }
      }
  }
  for (unsigned i = 0; i < n; i++) {
    DFM_TRY(ilu_create_impl(S->dia[i], alev[i * n + i], DFM_ORDER_NATURAL, nullptr, 0, &D[i].rp, &D[i].cl,
                            &M->dia[i]));
    for (unsigned j = 0; j < n; j++) {
      const dfm_bmat_s* B = S->nd[i][j];
      if (!B) continue;
      const LevPat& P = F[i][j];
      DFM_TRY(dfm_bmat_create(B->nb, B->b, B->nbr, B->br, &M->F[i][j]));
      std::vector<unsigned> rp = P.rp, cl = P.cl;
      if (cl.empty()) cl.push_back(0);
      DFM_TRY(dfm_bmat_add_pattern(M->F[i][j], B->nb, rp.data(), cl.data()));
      DFM_CHECK(M->F[i][j]->col == P.cl, DFM_BAD_ARG, "block (%u, %u): factor pattern not kept", i, j);
      std::vector<long long> src(P.cl.size(), -1);
      for (unsigned r = 0; r < B->nb; r++)
        for (unsigned q = B->rowptr[r]; q < B->rowptr[r + 1]; q++) {
          const auto it = std::lower_bound(P.cl.begin() + P.rp[r], P.cl.begin() + P.rp[r + 1], B->col[q]);
          src[it - P.cl.begin()] = q;
        }
      DFM_TRY(M->src[i][j].upload(src));
      M->src_col[i][j] = B->col;
    }
  }
  *out = M.release();
  return DFM_OK;
}
int dfm_mseg_ilu_destroy(dfm_msilu* M) {
  delete M;
  return DFM_OK;
}
namespace {
// Longest row (entries) of a CRS pattern, and whether rows of that length
// and blocks of bs values take the wave-per-row kernels (DFM_MS_WAVE=0: never,
// for comparisons).
bool g_ms_wave = [] {
  const char* e = getenv("DFM_MS_WAVE");
  return !(e && e[0] == '0');
}();
template <typename V>
bool ms_wave_fits(const V& rowptr, int bs) {
  if (!g_ms_wave || bs > kMsWaveBlk) return false;
  for (size_t r = 0; r + 1 < rowptr.size(); r++)
    if ((long)rowptr[r + 1] - (long)rowptr[r] > kMsWaveRow) return false;
  return true;
}
inline int wgrid(long nrow) { return (int)std::max<long>(1, (nrow + kMsWaves - 1) / kMsWaves); }

struct DiaPrep {
  dfm_msilu* M;
  unsigned i;
};
// DoILUDecompLowUp of segment i's diagonal factor with each pair of
// off-diagonal factors (i, k), (k, i), k < i, on rows [row0, nb_i)
int dia_lowup(void* ctx, int row0) {
  const DiaPrep* P = static_cast<const DiaPrep*>(ctx);
  dfm_msilu* M = P->M;
  dfm_mseg_s* S = M->S;
  const unsigned i = P->i;
  dfm_ilu_s* D = M->dia[i];
  for (unsigned k = 0; k < i; k++) {
    const dfm_bmat_s *lo = M->F[i][k], *up = M->F[k][i];
    if (!lo || !up) continue;
    if ((long)S->nb[i] - row0 <= 0) continue;
    if (ms_wave_fits(D->frow, (int)(S->b[i] * S->b[i]))) {
      hipLaunchKernelGGL(k_ms_lowup_dia_w, dim3(wgrid((long)S->nb[i] - row0)), dim3(64 * kMsWaves), 0, stream(),
                         (int)S->nb[i], row0, D->d_frow.p, D->d_fcol.p, D->d_fval.p, D->d_fdia.p, lo->d_rowptr.p,
                         lo->col.empty() ? nullptr : lo->d_col.p, lo->d_val.p, up->d_rowptr.p,
                         up->col.empty() ? nullptr : up->d_col.p, up->d_val.p, (int)S->b[i], (int)S->b[k]);
      DFM_HIP(hipGetLastError());
      continue;
    }
    DFM_MLAUNCH(k_ms_lowup_dia, mgrid((long)S->nb[i] - row0), (int)S->nb[i], row0, D->d_frow.p, D->d_fcol.p,
                D->d_fval.p, D->d_fdia.p, lo->d_rowptr.p, lo->col.empty() ? nullptr : lo->d_col.p, lo->d_val.p,
                up->d_rowptr.p, up->col.empty() ? nullptr : up->d_col.p, up->d_val.p, (int)S->b[i], (int)S->b[k]);
  }
  return DFM_OK;
}
// the state of rows [row0, nb_i) before DoILUDecomp: copied values, then the LowUp updates
int dia_restore(void* ctx, int row0) {
  const DiaPrep* P = static_cast<const DiaPrep*>(ctx);
  DFM_TRY(ilu_numeric_copy_rows(P->M->dia[P->i], P->M->S->dia[P->i], row0));
  return dia_lowup(ctx, row0);
}
}  // namespace

int dfm_mseg_ilu_set_value(dfm_msilu* M, int* nsingular) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  dfm_mseg_s* S = M->S;
  const unsigned n = S->nseg;
  int nsing_total = 0;
  // every off-diagonal factor block gets new values below
  for (auto& row : M->F)
    for (dfm_bmat_s* F : row)
      if (F) F->version = next_version();
  // SetValue_Initialize of every block first (preconditioner.cpp:231-251)
  for (unsigned i = 0; i < n; i++) {
    DFM_CHECK(S->dia[i]->pattern_hash == M->dia[i]->pattern_hash, DFM_BAD_ARG,
              "segment %u: matrix pattern differs from the one the factor was built on", i);
    DFM_TRY(ilu_numeric_copy(M->dia[i], S->dia[i]));
    for (unsigned j = 0; j < n; j++) {
      dfm_bmat_s* F = M->F[i][j];
      if (!F) continue;
      const long bs = (long)F->b * F->br, nv = (long)F->col.size() * bs;
      DFM_CHECK(S->nd[i][j]->col == M->src_col[i][j], DFM_BAD_ARG, "block (%u, %u): pattern changed", i, j);
      if (nv) DFM_MLAUNCH(k_ms_init, mgrid(nv), nv, (int)bs, M->src[i][j].p, S->nd[i][j]->d_val.p, F->d_val.p);
    }
  }
  // the blocks in (i, j) order (:253-297)
  for (unsigned i = 0; i < n; i++)
    for (unsigned j = 0; j < n; j++) {
      if (i == j) {
        DiaPrep prep{M, i};
        DFM_TRY(dia_lowup(&prep, 0));
        int ns = 0;
        const int rc = ilu_numeric_factor(M->dia[i], S->dia[i], &ns, dia_restore, &prep);
        nsing_total += ns;
        if (rc != DFM_OK) {
          if (nsingular) *nsingular = nsing_total;
          return rc;
        }
        continue;
      }
      dfm_bmat_s* T = M->F[i][j];
      if (!T) continue;
      const int* tcl = T->col.empty() ? nullptr : T->d_col.p;
      const bool wave = ms_wave_fits(T->rowptr, (int)(T->b * T->br));
      const dim3 wblk(64 * kMsWaves);
      const unsigned kmax = std::min(i, j);
      for (unsigned k = 0; k < kmax + 1; k++) {
        if (k == i && k < j) {  // DoILUDecompUp with factor i, one launch per lower level
          const dfm_ilu_s* D = M->dia[i];
          for (int l = 0; l < D->nlev_lower; l++) {
            const int o = D->lev_ptr_L[l], nr = D->lev_ptr_L[l + 1] - o;
            if (wave && T->b * T->b <= (unsigned)kMsWaveBlk) {
              hipLaunchKernelGGL(k_ms_up_w, dim3(wgrid(nr)), wblk, 0, stream(), nr, D->d_rows_L.p + o, D->d_frow.p,
                                 D->d_fcol.p, D->d_fdiaind.p, D->d_fval.p, D->d_fdia.p, T->d_rowptr.p, tcl,
                                 T->d_val.p, (int)T->b, (int)T->br);
              DFM_HIP(hipGetLastError());
              continue;
            }
            DFM_MLAUNCH(k_ms_up, mgrid(nr), nr, D->d_rows_L.p + o, D->d_frow.p, D->d_fcol.p, D->d_fdiaind.p,
                        D->d_fval.p, D->d_fdia.p, T->d_rowptr.p, tcl, T->d_val.p, (int)T->b, (int)T->br);
          }
          continue;
        }
        if (k == j && k < i) {  // DoILUDecompLow with factor j
          const dfm_ilu_s* D = M->dia[j];
          if (wave && T->br * T->br <= (unsigned)kMsWaveBlk) {
            hipLaunchKernelGGL(k_ms_low_w, dim3(wgrid(T->nb)), wblk, 0, stream(), (int)T->nb, T->d_rowptr.p, tcl,
                               T->d_val.p, D->d_frow.p, D->d_fcol.p, D->d_fdiaind.p, D->d_fval.p, (int)T->b,
                               (int)T->br);
            DFM_HIP(hipGetLastError());
            continue;
          }
          DFM_MLAUNCH(k_ms_low, mgrid(T->nb), (int)T->nb, T->d_rowptr.p, tcl, T->d_val.p, D->d_frow.p, D->d_fcol.p,
                      D->d_fdiaind.p, D->d_fval.p, (int)T->b, (int)T->br);
          continue;
        }
        const dfm_bmat_s *lo = M->F[i][k], *up = M->F[k][j];
        if (!lo || !up) continue;
        if (wave) {
          hipLaunchKernelGGL(k_ms_lowup_w, dim3(wgrid(T->nb)), wblk, 0, stream(), (int)T->nb, T->d_rowptr.p, tcl,
                             T->d_val.p, lo->d_rowptr.p, lo->col.empty() ? nullptr : lo->d_col.p, lo->d_val.p,
                             up->d_rowptr.p, up->col.empty() ? nullptr : up->d_col.p, up->d_val.p, (int)T->b,
                             (int)lo->br, (int)T->br);
          DFM_HIP(hipGetLastError());
          continue;
        }
        DFM_MLAUNCH(k_ms_lowup, mgrid(T->nb), (int)T->nb, T->d_rowptr.p, tcl, T->d_val.p, lo->d_rowptr.p,
                    lo->col.empty() ? nullptr : lo->d_col.p, lo->d_val.p, up->d_rowptr.p,
                    up->col.empty() ? nullptr : up->d_col.p, up->d_val.p, (int)T->b, (int)lo->br, (int)T->br);
      }
    }
  DFM_HIP(hipStreamSynchronize(stream()));
  if (nsingular) *nsingular = nsing_total;
  return DFM_OK;
}
int dfm_mseg_ilu_get_pattern(const dfm_msilu* M, unsigned i, unsigned j, unsigned* ncrs, unsigned* rowptr,
                             unsigned* col) {
  DFM_CHECK(M && i < M->S->nseg && j < M->S->nseg && ncrs, DFM_BAD_ARG, "bad arguments to dfm_mseg_ilu_get_pattern");
  const std::vector<unsigned>* rp;
  const std::vector<unsigned>* cl;
  if (i == j) {
    rp = &M->dia[i]->frow;
    cl = &M->dia[i]->fcol;
  } else {
    DFM_CHECK(M->F[i][j], DFM_BAD_ARG, "no block (%u, %u)", i, j);
    rp = &M->F[i][j]->rowptr;
    cl = &M->F[i][j]->col;
  }
  *ncrs = (unsigned)cl->size();
  if (rowptr) std::memcpy(rowptr, rp->data(), sizeof(unsigned) * rp->size());
  if (col && !cl->empty()) std::memcpy(col, cl->data(), sizeof(unsigned) * cl->size());
  return DFM_OK;
}
int dfm_mseg_ilu_get_values(const dfm_msilu* M, unsigned i, unsigned j, double* val, double* dia) {
  DFM_CHECK(M && i < M->S->nseg && j < M->S->nseg, DFM_BAD_ARG, "bad arguments to dfm_mseg_ilu_get_values");
  if (i == j) return dfm_ilu_get_values(M->dia[i], val, dia);
  DFM_CHECK(M->F[i][j], DFM_BAD_ARG, "no block (%u, %u)", i, j);
  return dfm_bmat_get_values(M->F[i][j], val);
}
// SolvePrecond across segments (preconditioner.cpp:317-333)
int dfm_mseg_ilu_precond(dfm_msilu* M, double* vec) {
  DFM_CHECK(M && vec, DFM_BAD_ARG, "bad arguments to dfm_mseg_ilu_precond");
  dfm_mseg_s* S = M->S;
  const unsigned n = S->nseg;
  for (unsigned i = 0; i < n; i++) {
    for (unsigned j = 0; j < i; j++)
      if (M->F[i][j]) DFM_TRY(dfm_bmat_matvec(M->F[i][j], -1.0, vec + S->off[j], 1.0, vec + S->off[i]));
    DFM_TRY(ilu_sweep64(M->dia[i], vec + S->off[i], true, false));
  }
  for (int i = (int)n - 1; i >= 0; i--) {
    for (unsigned j = (unsigned)i + 1; j < n; j++)
      if (M->F[i][j]) DFM_TRY(dfm_bmat_matvec(M->F[i][j], -1.0, vec + S->off[j], 1.0, vec + S->off[i]));
    DFM_TRY(ilu_sweep64(M->dia[i], vec + S->off[i], false, true));
  }
  return sync_stream();
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

// Work vectors start at zero, as the reference's: MATVEC(1, x, 0, y) scales
// the old y by 0, which keeps a NaN that uninitialised memory may hold.
static int msilu_work(dfm_msilu_s* M, int nvec) {
  const size_t n = M->S->off[M->S->nseg];
  while ((int)M->work.size() < nvec) {
    double* p = nullptr;
    DFM_TRY(dfm_malloc((void**)&p, std::max<size_t>(1, n) * sizeof(double)));
    M->work.push_back(p);
    DFM_TRY(dfm_vec_fill(std::max<size_t>(1, n), 0.0, p));
  }
  return DFM_OK;
}

// LsSol::Solve_PCG (solver_ls_iter.cpp:100-172) on the segments: r in =
// right-hand side, out = residual; x out.  DFM_OK when converged,
// DFM_NOT_CONVERGED at max_iter (conv_ratio then the final ratio).
static int ms_solve_pcg(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(S && M && M->S == S && r && x && conv_ratio && iteration && r != x, DFM_BAD_ARG,
            "bad arguments to dfm_mseg_solve_pcg");
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  const size_t n = S->off[S->nseg];
  DFM_TRY(msilu_work(M, 2));
  double *p = M->work[0], *z = M->work[1];
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  double rr0 = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr0));
  if (rr0 < 1.0e-30) {
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  const double inv_rr0 = 1.0 / rr0;
  DFM_TRY(dfm_vec_copy(n, r, z));
  DFM_TRY(dfm_mseg_ilu_precond(M, z));
  DFM_TRY(dfm_vec_copy(n, z, p));
  double rz = 0.0;
  DFM_TRY(seg_dot(S, r, z, &rz));
  for (unsigned it = 0; it < mx_iter; it++) {
    DFM_TRY(dfm_mseg_matvec(S, 1.0, p, 0.0, z));
    double pAp = 0.0;
    DFM_TRY(seg_dot(S, p, z, &pAp));
    const double alpha = rz / pAp;
    DFM_TRY(ls_axpy(n, -alpha, z, r));
    DFM_TRY(ls_axpy(n, alpha, p, x));
    double rr = 0.0;
    DFM_TRY(seg_dot(S, r, r, &rr));
    if (rr * inv_rr0 < tol * tol) {
      *conv_ratio = std::sqrt(rr * inv_rr0);
      *iteration = it;
      return DFM_OK;
    }
    DFM_TRY(dfm_vec_copy(n, r, z));
    DFM_TRY(dfm_mseg_ilu_precond(M, z));
    double rz_new = 0.0;
    DFM_TRY(seg_dot(S, r, z, &rz_new));
    const double beta = rz_new / rz;
    rz = rz_new;
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(ls_axpy(n, 1.0, z, p));
  }
  double rr = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr));
  *conv_ratio = std::sqrt(rr * inv_rr0);
  return DFM_NOT_CONVERGED;
}

// LsSol::Solve_PBiCGSTAB (solver_ls_iter.cpp:291-408) on the segments.
// Returns DFM_OK when converged, DFM_NOT_CONVERGED at max_iter (the
// reference returns true in both cases; iteration = max_iter then).
static int ms_solve_pbicgstab(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio,
                             unsigned* iteration) {
  DFM_CHECK(S && M && M->S == S && r && x && conv_ratio && iteration && r != x, DFM_BAD_ARG,
            "bad arguments to dfm_mseg_solve_pbicgstab");
  const double tol = *conv_ratio;
  const unsigned max_iter = *iteration;
  const size_t n = S->off[S->nseg];
  DFM_TRY(msilu_work(M, 7));
  double *vs = M->work[0], *Ms = M->work[1], *AMs = M->work[2], *p = M->work[3], *Mp = M->work[4],
         *AMp = M->work[5], *r2 = M->work[6];
  double rr0 = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr0));
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
    DFM_TRY(dfm_mseg_ilu_precond(M, Mp));
    double r_r2 = 0.0;
    DFM_TRY(seg_dot(S, r, r2, &r_r2));
    DFM_TRY(dfm_mseg_matvec(S, 1.0, Mp, 0.0, AMp));
    double den = 0.0;
    DFM_TRY(seg_dot(S, AMp, r2, &den));
    const double alpha = r_r2 / den;
    DFM_TRY(dfm_vec_copy(n, r, vs));
    DFM_TRY(ls_axpy(n, -alpha, AMp, vs));
    DFM_TRY(dfm_vec_copy(n, vs, Ms));
    DFM_TRY(dfm_mseg_ilu_precond(M, Ms));
    DFM_TRY(dfm_mseg_matvec(S, 1.0, Ms, 0.0, AMs));
    double den2 = 0.0, num = 0.0;
    DFM_TRY(seg_dot(S, AMs, AMs, &den2));
    DFM_TRY(seg_dot(S, vs, AMs, &num));
    const double omega = num / den2;
    DFM_TRY(ls_axpy(n, alpha, Mp, x));
    DFM_TRY(ls_axpy(n, omega, Ms, x));
    DFM_TRY(dfm_vec_copy(n, vs, r));
    DFM_TRY(ls_axpy(n, -omega, AMs, r));
    double rr = 0.0;
    DFM_TRY(seg_dot(S, r, r, &rr));
    if (rr * inv_rr0 < tol * tol) {
      *conv_ratio = std::sqrt(rr * inv_rr0);
      *iteration = it;
      return DFM_OK;
    }
    double tmp1 = 0.0;
    DFM_TRY(seg_dot(S, r, r2, &tmp1));
    const double beta = tmp1 * alpha / (r_r2 * omega);
    DFM_TRY(dfm_vec_scal(n, beta, p));
    DFM_TRY(ls_axpy(n, 1.0, r, p));
    DFM_TRY(ls_axpy(n, -beta * omega, AMp, p));
  }
  return DFM_NOT_CONVERGED;
}
// ---------------------------------------------------------------------------
// Device-driven loops: the same operations in the same order, with the
// scalars (alpha, beta, omega and the convergence test) formed on the device
// from the segments' sums (added in segment order, as DOT does) and the host
// only enqueueing: blocks of kMsPoll iterations, each replayed from a hipGraph
// captured once per operand set, with the stop flag polled after each block.
// Updates of x, r and p are skipped once the flag is set (converged, or
// max_iter reached), so the iterations of a block after that change nothing
// the caller sees: the results are bitwise those of the host-driven loops.
namespace {

struct MsState {
  double inv_rr0, tol2, rz, alpha, beta, omega, r_r2, den2, rr, c_bo;
  int done;  // 0 running, 1 converged, 2 stopped at max_iter
  int iter, maxit, conv_iter;
};
constexpr int kMsPoll = 4;

enum { MS_RZ0, MS_PAP, MS_PCG_RR, MS_RZNEW, MS_RR2, MS_DEN, MS_DEN2_OMEGA, MS_RR_BETA };

// Segment table of the fused dots: segment i owns the blocks base[i] ..
// base[i+1] - 1 (as many as dfm_vec_dot gives a vector of its length) and the
// entries off[i] .. off[i+1] - 1.
struct SegTab {
  int nseg, G;
  int base[kMaxSeg + 1];
  long off[kMaxSeg + 1];
};

// One step of the loops' scalar arithmetic (written as the host loops write
// it) on the sums s of one or two dots.
template <int OP>
__device__ void ms_step(const double* s, MsState* st) {
  if (st->done) return;
  if constexpr (OP == MS_RZ0) {
    st->rz = s[0];
  } else if constexpr (OP == MS_PAP) {
    st->alpha = st->rz / s[0];
  } else if constexpr (OP == MS_PCG_RR) {
    st->rr = s[0];
    if (s[0] * st->inv_rr0 < st->tol2) {
      st->done = 1;
      st->conv_iter = st->iter;
    }
    st->iter += 1;
    if (!st->done && st->iter >= st->maxit) st->done = 2;
  } else if constexpr (OP == MS_RZNEW) {
    st->beta = s[0] / st->rz;
    st->rz = s[0];
  } else if constexpr (OP == MS_RR2) {
    st->r_r2 = s[0];
  } else if constexpr (OP == MS_DEN) {
    st->alpha = st->r_r2 / s[0];
  } else if constexpr (OP == MS_DEN2_OMEGA) {  // s = (AMs, AMs), (vs, AMs)
    st->den2 = s[0];
    st->omega = s[1] / s[0];
  } else if constexpr (OP == MS_RR_BETA) {  // s = (r, r), (r, r2)
    st->rr = s[0];
    if (s[0] * st->inv_rr0 < st->tol2) {
      st->done = 1;
      st->conv_iter = st->iter;
      return;
    }
    st->beta = s[1] * st->alpha / (st->r_r2 * st->omega);
    st->c_bo = -st->beta * st->omega;
    st->iter += 1;
    if (st->iter >= st->maxit) st->done = 2;
  }
}

// ND dots <x_d, y_d> of every segment in one launch: block b of segment i
// sums what block b - base[i] of dfm_vec_dot's kernel sums for that
// segment's vectors, and writes its partial at d * G + b.
template <int ND>
__global__ void __launch_bounds__(kNT) k_ms_dot(SegTab t, const double* __restrict__ x0,
                                                const double* __restrict__ y0, const double* __restrict__ x1,
                                                const double* __restrict__ y1, double* __restrict__ partials) {
  int i = 0;
  while (i + 1 < t.nseg && (int)blockIdx.x >= t.base[i + 1]) i++;
  const int lb = (int)blockIdx.x - t.base[i], g = t.base[i + 1] - t.base[i];
  const long o = t.off[i], n = t.off[i + 1] - o;
  double dsum[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) dsum[d] = 0.0;
  for (long k = (long)lb * kNT + threadIdx.x; k < n; k += (long)g * kNT) {
    dsum[0] += x0[o + k] * y0[o + k];
    if constexpr (ND > 1) dsum[1] += x1[o + k] * y1[o + k];
  }
  write_partials<ND>(dsum, partials, blockIdx.x, t.G);
}
// The segments' sums, each reduced as dfm_vec_dot's finalize reduces it, added
// in segment order (DOT, linearsystem.cpp:335-364), then the scalar step OP.
template <int ND, int OP>
__global__ void __launch_bounds__(kFinNT) k_ms_dot_fin(SegTab t, const double* __restrict__ partials,
                                                       MsState* __restrict__ st) {
  __shared__ double sh[kFinNT / 64];
  double tot[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) tot[d] = 0.0;
  for (int i = 0; i < t.nseg; i++) {
    const double* p[ND];
    int n[ND];
    double a[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      p[d] = partials + (size_t)d * t.G + t.base[i];
      n[d] = t.base[i + 1] - t.base[i];
    }
    strided_sums<kFinNT, ND>(p, n, a);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double s = block_sum_n<kFinNT>(a[d], sh);
      if (threadIdx.x == 0) tot[d] += s;
    }
  }
  if (threadIdx.x == 0) ms_step<OP>(tot, st);
}

// Fused vector updates, each the host loop's copies and AXPYs (an AXPY with a
// zero factor is skipped, linearsystem.cpp:302) element by element in the
// same order; nothing once the stop flag is set.
#define MS_LOOP(n) for (long i = (long)blockIdx.x * kNT + threadIdx.x; i < (n); i += (long)gridDim.x * kNT)
// PCG: r -= alpha Ap, x += alpha p, then z = r for the preconditioner
__global__ void k_ms_pcg_xr(long n, const MsState* __restrict__ st, const double* __restrict__ p,
                            double* __restrict__ z, double* __restrict__ r, double* __restrict__ x) {
  if (st->done) return;
  const double a = st->alpha, c = -a;
  MS_LOOP(n) {
    double ri = r[i];
    if (c != 0.0) ri += c * z[i];
    r[i] = ri;
    if (a != 0.0) x[i] += a * p[i];
    z[i] = ri;
  }
}
// PCG: p = beta p + z
__global__ void k_ms_pcg_p(long n, const MsState* __restrict__ st, const double* __restrict__ z,
                           double* __restrict__ p) {
  if (st->done) return;
  const double b = st->beta;
  MS_LOOP(n) {
    double pi = p[i] * b;
    pi += 1.0 * z[i];
    p[i] = pi;
  }
}
// BiCGSTAB: vs = r - alpha AMp, and Ms = vs for the preconditioner
__global__ void k_ms_bicg_vs(long n, const MsState* __restrict__ st, const double* __restrict__ r,
                             const double* __restrict__ AMp, double* __restrict__ vs, double* __restrict__ Ms) {
  if (st->done) return;
  const double c = -st->alpha;
  MS_LOOP(n) {
    double v = r[i];
    if (c != 0.0) v += c * AMp[i];
    vs[i] = v;
    Ms[i] = v;
  }
}
// BiCGSTAB: x += alpha Mp, x += omega Ms, r = vs - omega AMs
__global__ void k_ms_bicg_xr(long n, const MsState* __restrict__ st, const double* __restrict__ Mp,
                             const double* __restrict__ Ms, const double* __restrict__ vs,
                             const double* __restrict__ AMs, double* __restrict__ x, double* __restrict__ r) {
  if (st->done) return;
  const double a = st->alpha, w = st->omega, c = -w;
  MS_LOOP(n) {
    double xi = x[i];
    if (a != 0.0) xi += a * Mp[i];
    if (w != 0.0) xi += w * Ms[i];
    x[i] = xi;
    double ri = vs[i];
    if (c != 0.0) ri += c * AMs[i];
    r[i] = ri;
  }
}
// BiCGSTAB: p = beta p + r - beta omega AMp, and Mp = p for the next
// iteration's preconditioner
__global__ void k_ms_bicg_p(long n, const MsState* __restrict__ st, const double* __restrict__ r,
                            const double* __restrict__ AMp, double* __restrict__ p, double* __restrict__ Mp) {
  if (st->done) return;
  const double b = st->beta, c = st->c_bo;
  MS_LOOP(n) {
    double pi = p[i] * b;
    pi += 1.0 * r[i];
    if (c != 0.0) pi += c * AMp[i];
    p[i] = pi;
    Mp[i] = pi;
  }
}
__global__ void k_ms_copy(long n, const double* __restrict__ x, double* __restrict__ y) {
  MS_LOOP(n) y[i] = x[i];
}
#undef MS_LOOP

struct MsLoop {
  dfm_mseg_s* S;
  dfm_msilu_s* M;
  MsState* st;
  size_t n;
  int g;
  SegTab tab;
  // ND dots of all segments and the scalar step OP: two launches
  template <int ND, int OP>
  int dot(const double* x0, const double* y0, const double* x1 = nullptr, const double* y1 = nullptr) {
    DFM_MLAUNCH((k_ms_dot<ND>), tab.G, tab, x0, y0, x1, y1, S->dot_partials.p);
    hipLaunchKernelGGL((k_ms_dot_fin<ND, OP>), dim3(1), dim3(kFinNT), 0, stream(), tab, S->dot_partials.p, st);
    DFM_HIP(hipGetLastError());
    return DFM_OK;
  }
  int copy(const double* x, double* y) {  // a kernel, also inside graphs
    DFM_MLAUNCH(k_ms_copy, g, (long)n, x, y);
    return DFM_OK;
  }
  // the state's host copy after the work enqueued so far
  int poll(MsState& h) {
    DFM_HIP(hipMemcpyAsync(M->hstate, st, sizeof(MsState), hipMemcpyDeviceToHost, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
    std::memcpy(&h, M->hstate, sizeof(MsState));
    return DFM_OK;
  }
};

int ms_loop_init(dfm_mseg_s* S, dfm_msilu_s* M, double rr0, double tol, unsigned maxit, int iter0, MsLoop& L) {
  if (M->dstate.n < sizeof(MsState)) DFM_TRY(M->dstate.alloc(sizeof(MsState)));
  if (!M->hstate) DFM_HIP(hipHostMalloc(&M->hstate, sizeof(MsState), hipHostMallocDefault));
  SegTab& t = L.tab;
  t.nseg = (int)S->nseg;
  t.base[0] = 0;
  t.off[0] = 0;
  for (unsigned i = 0; i < S->nseg; i++) {
    t.base[i + 1] = t.base[i] + grid_vec((long)(S->off[i + 1] - S->off[i]));
    t.off[i + 1] = (long)S->off[i + 1];
  }
  t.G = t.base[S->nseg];
  // room for the fused dots and for seg_dot's layout (its largest grid per segment)
  int gmax = 1;
  for (unsigned i = 0; i < S->nseg; i++) gmax = std::max(gmax, t.base[i + 1] - t.base[i]);
  const size_t need = std::max((size_t)2 * t.G, (size_t)gmax * S->nseg);
  if (S->dot_partials.n < need) DFM_TRY(S->dot_partials.alloc(need));
  MsState h{};
  h.inv_rr0 = 1.0 / rr0;
  h.tol2 = tol * tol;
  h.maxit = (int)std::min<unsigned>(maxit, 0x7fffffffu);
  h.iter = iter0;
  std::memcpy(M->hstate, &h, sizeof(MsState));
  L.S = S;
  L.M = M;
  L.st = reinterpret_cast<MsState*>(M->dstate.p);
  L.n = S->off[S->nseg];
  L.g = grid_vec((long)L.n);
  DFM_HIP(hipMemcpyAsync(L.st, M->hstate, sizeof(MsState), hipMemcpyHostToDevice, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));  // the pinned copy is the poll buffer too
  return DFM_OK;
}

// The MatVec copies of every block the loops multiply by, made current before
// a block is captured (a capture may not allocate, and a refresh would be
// replayed with it).
int ms_prepare_matvecs(dfm_mseg_s* S, dfm_msilu_s* M) {
  if (!mv_sell_on()) return DFM_OK;
  for (unsigned i = 0; i < S->nseg; i++) {
    dfm_bcrs_s* A = S->dia[i];
    if (A->nb > 0 && A->b <= 3)
      DFM_TRY(mv_sell_prepare(A->mv, A->nb, A->rowptr, A->col, (int)(A->b * A->b), A->d_val.p, A->pattern_hash,
                              A->version));
    for (unsigned j = 0; j < S->nseg; j++)
      for (dfm_bmat_s* B : {S->nd[i][j], M->F[i][j]})
        if (B && B->nb > 0 && !B->col.empty() && B->b <= 3 && B->br <= 3)
          DFM_TRY(mv_sell_prepare(B->mv, B->nb, B->rowptr, B->col, (int)(B->b * B->br), B->d_val.p, B->pattern_key,
                                  B->version));
  }
  return DFM_OK;
}

// Operand set of a captured block: the vectors, the loop's buffers, every
// block's and factor's storage it reads, and the sweep mode.
std::string ms_graph_key(const char* tag, const MsLoop& L, std::initializer_list<const void*> ptrs) {
  std::string k = std::string(tag) + ":" + std::to_string(kMsPoll) + ":" + sweep_mode_key();
  char buf[24];
  auto add = [&](const void* p) {
    snprintf(buf, sizeof buf, ":%p", p);
    k += buf;
  };
  for (const void* p : ptrs) add(p);
  add(L.S->dot_partials.p);
  add(L.S->dot_state.p);
  add(L.st);
  for (unsigned i = 0; i < L.S->nseg; i++) {
    add(L.S->dia[i]);
    add(L.S->dia[i]->mv.val.p);
    add(L.S->dia[i]->mv.col.p);
    add(L.M->dia[i]);
    for (unsigned j = 0; j < L.S->nseg; j++)
      for (const dfm_bmat_s* B : {L.S->nd[i][j], L.M->F[i][j]}) {
        add(B);
        if (B) {
          add(B->mv.val.p);
          add(B->mv.col.p);
        }
      }
  }
  return k;
}
// One block of kMsPoll iterations: replayed from the cached graph, captured
// on first use, launched directly where the stream cannot be captured.
template <class Body>
int ms_run_block(dfm_msilu_s* M, const std::string& key, Body&& body) {
  hipStream_t s = stream();
  static const bool use_graphs = [] {
    const char* e = getenv("DFM_GRAPHS");
    return !(e && e[0] == '0');
  }();
  static bool capture_ok = true;
  if (s == nullptr || !use_graphs || !capture_ok) {
    for (int i = 0; i < kMsPoll; i++) DFM_TRY(body());
    return DFM_OK;
  }
  for (auto& g : M->graphs)
    if (g.first == key) {
      DFM_HIP(hipGraphLaunch(g.second, s));
      return DFM_OK;
    }
  DFM_TRY(evict_graph(M->graphs));
  hipGraph_t graph = nullptr;
  DFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = DFM_OK;
  for (int i = 0; i < kMsPoll && rc == DFM_OK; i++) rc = body();
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != DFM_OK || e != hipSuccess) {
    // nothing ran: launch directly from now on
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    capture_ok = false;
    for (int i = 0; i < kMsPoll; i++) DFM_TRY(body());
    return DFM_OK;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DFM_HIP(e);
  M->graphs.emplace_back(key, exec);
  DFM_HIP(hipGraphLaunch(exec, s));
  return DFM_OK;
}

// Solve_PCG as ms_solve_pcg
int ms_dev_pcg(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(S && M && M->S == S && r && x && conv_ratio && iteration && r != x, DFM_BAD_ARG,
            "bad arguments to dfm_mseg_solve_pcg");
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iteration;
  const size_t n = S->off[S->nseg];
  DFM_TRY(msilu_work(M, 2));
  double *p = M->work[0], *z = M->work[1];
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  double rr0 = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr0));  // also sizes the dot buffers
  if (rr0 < 1.0e-30) {
    *conv_ratio = 0.0;
    *iteration = 0;
    return DFM_OK;
  }
  MsLoop L;
  DFM_TRY(ms_loop_init(S, M, rr0, tol, mx_iter, 0, L));
  DFM_TRY(ms_prepare_matvecs(S, M));
  MsState* st = L.st;
  const long nn = (long)n;
  DFM_TRY(L.copy(r, z));
  DFM_TRY(dfm_mseg_ilu_precond(M, z));
  DFM_TRY(L.copy(z, p));
  DFM_TRY((L.dot<1, MS_RZ0>(r, z)));
  auto body = [&]() -> int {
    DFM_TRY(dfm_mseg_matvec(S, 1.0, p, 0.0, z));
    DFM_TRY((L.dot<1, MS_PAP>(p, z)));
    DFM_MLAUNCH(k_ms_pcg_xr, L.g, nn, st, p, z, r, x);
    DFM_TRY((L.dot<1, MS_PCG_RR>(r, r)));
    DFM_TRY(dfm_mseg_ilu_precond(M, z));
    DFM_TRY((L.dot<1, MS_RZNEW>(r, z)));
    DFM_MLAUNCH(k_ms_pcg_p, L.g, nn, st, z, p);
    return DFM_OK;
  };
  const std::string key = ms_graph_key("pcg", L, {r, x, p, z});
  MsState h{};
  for (unsigned it = 0; it < mx_iter; it += kMsPoll) {
    DFM_TRY(ms_run_block(M, key, body));
    DFM_TRY(L.poll(h));
    if (h.done) break;
  }
  if (h.done == 1) {
    *conv_ratio = std::sqrt(h.rr * h.inv_rr0);
    *iteration = (unsigned)h.conv_iter;
    return DFM_OK;
  }
  // max_iter: the ratio of the final residual
  double rr = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr));
  *conv_ratio = std::sqrt(rr * (1.0 / rr0));
  return DFM_NOT_CONVERGED;
}

// Solve_PBiCGSTAB as ms_solve_pbicgstab
int ms_dev_pbicgstab(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration) {
  DFM_CHECK(S && M && M->S == S && r && x && conv_ratio && iteration && r != x, DFM_BAD_ARG,
            "bad arguments to dfm_mseg_solve_pbicgstab");
  const double tol = *conv_ratio;
  const unsigned max_iter = *iteration;
  const size_t n = S->off[S->nseg];
  DFM_TRY(msilu_work(M, 7));
  double *vs = M->work[0], *Ms = M->work[1], *AMs = M->work[2], *p = M->work[3], *Mp = M->work[4],
         *AMp = M->work[5], *r2 = M->work[6];
  double rr0 = 0.0;
  DFM_TRY(seg_dot(S, r, r, &rr0));
  if (rr0 < 1.0e-60) {
    *conv_ratio = std::sqrt(rr0);
    *iteration = 0;
    return DFM_OK;
  }
  // the loop's counter starts at 1 (it = 1 .. max_iter - 1)
  MsLoop L;
  DFM_TRY(ms_loop_init(S, M, rr0, tol, max_iter, 1, L));
  DFM_TRY(ms_prepare_matvecs(S, M));
  MsState* st = L.st;
  const long nn = (long)n;
  DFM_TRY(dfm_vec_scal(n, 0.0, x));
  DFM_TRY(L.copy(r, r2));
  DFM_TRY(L.copy(r, p));
  DFM_TRY(L.copy(p, Mp));
  *iteration = max_iter;
  MsState h{};
  if (max_iter <= 1) return DFM_NOT_CONVERGED;
  auto body = [&]() -> int {
    DFM_TRY(dfm_mseg_ilu_precond(M, Mp));
    DFM_TRY((L.dot<1, MS_RR2>(r, r2)));
    DFM_TRY(dfm_mseg_matvec(S, 1.0, Mp, 0.0, AMp));
    DFM_TRY((L.dot<1, MS_DEN>(AMp, r2)));
    DFM_MLAUNCH(k_ms_bicg_vs, L.g, nn, st, r, AMp, vs, Ms);
    DFM_TRY(dfm_mseg_ilu_precond(M, Ms));
    DFM_TRY(dfm_mseg_matvec(S, 1.0, Ms, 0.0, AMs));
    DFM_TRY((L.dot<2, MS_DEN2_OMEGA>(AMs, AMs, vs, AMs)));
    DFM_MLAUNCH(k_ms_bicg_xr, L.g, nn, st, Mp, Ms, vs, AMs, x, r);
    DFM_TRY((L.dot<2, MS_RR_BETA>(r, r, r, r2)));
    DFM_MLAUNCH(k_ms_bicg_p, L.g, nn, st, r, AMp, p, Mp);
    return DFM_OK;
  };
  const std::string key = ms_graph_key("bicg", L, {r, x, vs, Ms, AMs, p, Mp, AMp, r2});
  for (unsigned it = 1; it < max_iter; it += kMsPoll) {
    DFM_TRY(ms_run_block(M, key, body));
    DFM_TRY(L.poll(h));
    if (h.done) break;
  }
  if (h.done == 1) {
    *conv_ratio = std::sqrt(h.rr * h.inv_rr0);
    *iteration = (unsigned)h.conv_iter;
    return DFM_OK;
  }
  return DFM_NOT_CONVERGED;
}

bool g_ms_device_loop = [] {
  const char* e = getenv("DFM_MSEG_LOOP");
  return !(e && std::strcmp(e, "host") == 0);
}();

}  // namespace

int dfm_mseg_solve_pcg(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration) {
  if (g_ms_device_loop) return deferred([&] { return ms_dev_pcg(S, M, r, x, conv_ratio, iteration); });
  return deferred([&] { return ms_solve_pcg(S, M, r, x, conv_ratio, iteration); });
}
int dfm_mseg_solve_pbicgstab(dfm_mseg* S, dfm_msilu* M, double* r, double* x, double* conv_ratio,
                             unsigned* iteration) {
  if (g_ms_device_loop) return deferred([&] { return ms_dev_pbicgstab(S, M, r, x, conv_ratio, iteration); });
  return deferred([&] { return ms_solve_pbicgstab(S, M, r, x, conv_ratio, iteration); });
}
int dfm_mseg_set_device_loop(int on) {
  g_ms_device_loop = on != 0;
  return DFM_OK;
}
