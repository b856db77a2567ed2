// dfm_batch.hip -- many independent scalar FE systems solved in one
// persistent launch: per system a multicolour ILU(0) numeric factorisation
// (DoILUDecomp, src/matvec/matdiafrac_blkcrs.cpp:769-1102, b = 1) and a
// preconditioned CG (LsSol::Solve_PCG, src/ls/solver_ls_iter.cpp:100-172),
// in FP64 or in FP32 with FP64 reliable updates.
//
// Layout.  The systems are the diagonal blocks of one assembled block-CRS
// matrix (1x1 blocks), rows [sys_ptr[s], sys_ptr[s+1]).  Each system is
// renumbered by a greedy colouring of its graph (rows of one colour are
// independent in both triangular sweeps) and its off-diagonal entries are
// stored column-major ELL, slot k of new row i at ell0 + k*n + i, slots in
// ascending new column, so the L part precedes the U part and a wavefront
// reads consecutive rows.  A workgroup owns one system at a time: it takes the
// next system from a work counter (largest first), factors it colour by
// colour, iterates with workgroup barriers only, writes its result and takes
// the next.  Dot products are workgroup reductions in FP64 in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <memory>
#include <utility>

#include "dfm_internal.h"

namespace dfm {

constexpr int kMaxColourBatch = 16;

struct SysDesc {
  int row0;       // first row in the concatenated (new) numbering
  int n;          // rows
  int w;          // ELL width
  int ncol;       // colours
  long long ell0; // first ELL slot
};

struct BatchOut {
  int iters;      // PCG iterations (FP64: the reference's 0-based index at convergence)
  int outer;      // reliable updates (FP32 mode)
  int status;     // 1 converged, 0 not converged, -1 singular pivots > 10
  int nsing;      // singular pivots met
  double relres;  // ||b - A x|| / ||b|| in FP64 at the end
  unsigned long long t0, t1;  // wall clock at pick-up and at completion
};

}  // namespace dfm

struct dfm_batch_s {
  unsigned nsys = 0, nrow = 0;
  int max_w = 0, max_ncol = 0;
  std::vector<unsigned> sys_ptr;
  std::vector<int> perm;                 // new (concatenated) -> old row
  std::vector<dfm::SysDesc> desc;
  std::vector<int> cptr;                 // per system kMaxColourBatch+1 colour offsets (local rows)
  dfm::DevBuf<dfm::SysDesc> d_sys;
  dfm::DevBuf<int> d_cptr, d_perm, d_ell_col;
  std::vector<int> sched;                // systems by decreasing work
  dfm::DevBuf<long long> d_ell_src;
  dfm::DevBuf<unsigned> d_next;
  dfm::DevBuf<double> d_a64, d_f64, d_adia64, d_fdia64, d_b64, d_x64, d_r64;
  dfm::DevBuf<float> d_a32, d_f32, d_adia32, d_fdia32;
  dfm::DevBuf<char> d_vec;               // solver vectors r, z, p, q, d (T = float or double)
  dfm::DevBuf<dfm::BatchOut> d_out;
  const dfm_bcrs_s* A = nullptr;
  int block = 512, wpc = 1;              // launch shape
  unsigned team_min_rows = 16384;        // systems at least this large get a team of workgroups
  int team_max = 4;                      // workgroups per team (1: no teams)
  std::string items_key;                 // launch shape the work items were built for
  dfm::DevBuf<char> d_items, d_teams;
  dfm::DevBuf<int> d_fail;
  int nitem = 0;
  bool local = false;                    // teams queued per XCD (xq: queue offsets, singles from xq[8])
  std::vector<int> xq;
  int resident = 0;
  bool tuned = false;
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
  ~dfm_batch_s() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  }
};

namespace dfm {

// ---------------------------------------------------------------------------
// Device side
constexpr int kMaxTeam = 16;               // workgroups sharing one system
constexpr unsigned kTeamSpinLimit = 1u << 21;

// Per system, reset every launch: the team barrier counter (alone on its
// cache line) and double-buffered slots for the members' partial sums.
struct TeamState {
  unsigned arrive;
  unsigned pad[31];
  double part[2][kMaxTeam];
};

// One unit of work: member `member` of a team of `team` workgroups on system `sys`.
struct Item {
  int sys, member, team, pad;
};

// Work counters: one per XCD for its teams, one for the single items, each
// on its own 128-byte line.
constexpr int kXcd = 8;
constexpr int kCounterStride = 32;
struct BatchArgs {
  int nitem;
  int local;          // teams XCD-local: items[xq[x] .. xq[x+1]) for XCD x, singles from xq[kXcd]
  int xq[kXcd + 1];
  const Item* items;
  unsigned* next;
  TeamState* teams;
  int* fail;
  const SysDesc* sys;
  const int* cptr;
  const int* ecol;
  const double* a64;
  double* f64;
  const double* adia64;
  double* fdia64;
  float* a32;
  float* f32;
  float* adia32;
  float* fdia32;
  const double* b64;
  double* x64;
  double* r64;
  void* vec;          // 5 vectors of nrow entries, type T
  long nrow;
  double tol2;
  int maxit;
  BatchOut* out;
};

// all-reduce over the workgroup, result in every thread; alternating LDS
// buffers make one barrier per reduction enough
__device__ inline double wg_sum(double v, double (*sh)[16], int& buf) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[buf][w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; i++) s += sh[buf][i];
  buf ^= 1;
  return s;
}

// The workgroups working on one system.  With K = 1 it is the workgroup
// alone (barriers are workgroup barriers).  With K > 1 every member owns a
// contiguous share of each row range; a team barrier is: every wave waits for
// its stores, a workgroup barrier, then wave 0 makes an agent-scope release,
// increments the system's counter, waits until all K members arrived
// (relaxed polls, bounded: a timeout sets *fail and the launch completes with
// the result marked invalid) and makes an agent-scope acquire, and a second
// workgroup barrier.  Team reductions write each member's workgroup sum to a
// slot, pass a team barrier and add the K slots in member order: identical
// in every member.
struct Team {
  int K, m;
  int local;    // all members on this XCD: barriers through its L2
  TeamState* ts;
  int* fail;
  unsigned nsync;
  int tbuf;
  double (*sh)[16];
  int buf;
  double* bc;   // LDS broadcast slot
};

// Thread 0's part of a team barrier, after every wave has waited for its
// own stores and a workgroup barrier: arrive on the system's counter, wait
// until all K members arrived (relaxed polls, bounded: a timeout sets *fail
// and the launch completes with the result marked invalid), and make the
// members' stores visible to this workgroup.
__device__ inline void team_meet(Team& t) {
  if (__hip_atomic_load(t.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (t.local) {
    // Members on one XCD share its L2: once every wave's stores are
    // acknowledged (vmcnt(0); the L1 writes through) they are in that L2, so
    // the arrival needs no L2 write-back (measured: without it the team
    // results are those of the agent-scope barrier); the departure
    // invalidates with buffer_inv sc1 (sc0 alone left stale L1 lines here:
    // wrong results).
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  } else {
    // one agent-scope release -- an L2 write-back -- covers the stores of
    // all waves, which have been acknowledged by L2
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  }
  __hip_atomic_fetch_add(&t.ts->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned target = t.nsync * (unsigned)t.K;
  unsigned spins = 0;
  while (__hip_atomic_load(&t.ts->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kTeamSpinLimit) {
      __hip_atomic_store(t.fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
  // the acquire invalidates the CU's L1 (shared by all waves of the
  // workgroup) and, across XCDs, the XCD's L2
  if (t.local) asm volatile("buffer_inv sc1" ::: "memory");
  else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Team barrier.  Every wave waits until L2 has acknowledged its own stores
// (a workgroup-scope barrier does not wait for them on this device), then a
// workgroup barrier, thread 0 meets the other members, and a second
// workgroup barrier holds the other waves until it has.
__device__ inline void tsync(Team& t) {
  if (t.K == 1) {
    __syncthreads();
    return;
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  t.nsync++;
  if (threadIdx.x == 0) team_meet(t);
  __syncthreads();
}

// Team reduction: the waves' sums to LDS in the first barrier of the team
// barrier, thread 0 adds them in wave order, publishes the workgroup's sum in
// its member slot, meets the other members and adds the K slots in member
// order (identical in every member); the second barrier broadcasts it.
__device__ inline double tsum(Team& t, double v) {
  if (t.K == 1) return wg_sum(v, t.sh, t.buf);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) t.sh[t.buf][threadIdx.x >> 6] = v;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  t.nsync++;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += t.sh[t.buf][i];
    t.ts->part[t.tbuf][t.m] = s;
    team_meet(t);
    double a = 0.0;
    for (int k = 0; k < t.K; k++)
      a += __hip_atomic_load(&t.ts->part[t.tbuf][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *t.bc = a;
  }
  __syncthreads();
  const double r = *t.bc;
  t.buf ^= 1;
  t.tbuf ^= 1;
  return r;
}

// this member's share [a, b) of the rows [lo, hi)
__device__ inline void tshare(const Team& t, int lo, int hi, int& a, int& b) {
  const long long len = hi - lo;
  a = lo + (int)(len * t.m / t.K);
  b = lo + (int)(len * (t.m + 1) / t.K);
}

// Row access.  A row's W column slots are loaded together (predicated on the
// system's width, padding -1), then the value and vector loads of every slot
// are issued together, so a wavefront has 2W independent loads in flight
// instead of a chain of dependent ones; the arithmetic then runs in slot
// (ascending column) order.
template <int W>
__device__ inline void ell_cols(const int* __restrict__ ecol, const SysDesc& d, int i, int (&c)[W]) {
#pragma unroll
  for (int k = 0; k < W; k++) c[k] = k < d.w ? ecol[d.ell0 + (long long)k * d.n + i] : -1;
}

// Rows are taken U at a time per thread (rows i, i+T, .., i+(U-1)T), their
// loads issued together before any arithmetic.
template <int W, typename T = double>
constexpr int rows_per_pass() { return W <= 8 ? 2 : 1; }
// rows per thread per pass of the vector updates (no gathers)
constexpr int kStreamRows = 4;

// y = A x over rows [lo, hi) (ELL + diagonal); returns the <x, y> partial
template <int W, typename V, typename T>
__device__ inline double ell_spmv(const SysDesc& d, int lo, int hi, const int* __restrict__ ecol,
                                  const V* __restrict__ a, const V* __restrict__ adia, const T* x, T* y) {
  constexpr int U = rows_per_pass<W, T>();
  const int NT = blockDim.x;
  double part = 0.0;
  for (int i0 = lo + (int)threadIdx.x; i0 < hi; i0 += U * NT) {
    int c[U][W];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
      if (i < hi) {
        ell_cols<W>(ecol, d, i, c[u]);
      } else {
#pragma unroll
        for (int k = 0; k < W; k++) c[u][k] = -1;
      }
    }
    T v[U][W], xi[U], di[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
#pragma unroll
      for (int k = 0; k < W; k++)
        v[u][k] = c[u][k] >= 0 ? (T)a[d.ell0 + (long long)k * d.n + i] * x[c[u][k]] : (T)0;
      xi[u] = i < hi ? x[i] : (T)0;
      di[u] = i < hi ? (T)adia[d.row0 + i] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
      if (i >= hi) continue;
      T acc = 0;
#pragma unroll
      for (int k = 0; k < W; k++) acc += v[u][k];
      acc += di[u] * xi[u];
      y[i] = acc;
      part += (double)xi[u] * (double)acc;
    }
  }
  return part;
}

// One substitution step over rows [lo, hi) of one colour.  Forward
// (LOWER): z_i = Dinv_i (r_i - sum_{j<i} L_ij z_j); backward: z_i -= sum_{j>i}
// U_ij z_j, returning the <r, z> partial of the finished rows.  Each sum in
// ascending column order, as the reference's substitutions.  With DOT a
// forward step also returns the <r, z> partial of its rows: the backward step
// of the last colour, whose rows have no later neighbours (the colouring is
// proper), so the backward step changes nothing there.
template <int W, bool LOWER, typename V, typename T, bool DOT = false>
__device__ inline double ell_subst(const SysDesc& d, int lo, int hi, const int* __restrict__ ecol,
                                   const V* __restrict__ f, const V* __restrict__ fdia, const T* r, T* z) {
  constexpr int U = rows_per_pass<W, T>();
  const int NT = blockDim.x;
  double part = 0.0;
  for (int i0 = lo + (int)threadIdx.x; i0 < hi; i0 += U * NT) {
    int c[U][W];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
      if (i < hi) {
        ell_cols<W>(ecol, d, i, c[u]);
      } else {
#pragma unroll
        for (int k = 0; k < W; k++) c[u][k] = -1;
      }
    }
    T v[U][W], t[U], di[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
#pragma unroll
      for (int k = 0; k < W; k++) {
        const bool use = LOWER ? (c[u][k] >= 0 && c[u][k] < i) : (c[u][k] > i);
        v[u][k] = use ? (T)f[d.ell0 + (long long)k * d.n + i] * z[c[u][k]] : (T)0;
      }
      t[u] = i < hi ? (LOWER ? r[i] : z[i]) : (T)0;
      di[u] = (LOWER && i < hi) ? (T)fdia[d.row0 + i] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * NT;
      if (i >= hi) continue;
      T tt = t[u];
#pragma unroll
      for (int k = 0; k < W; k++) {
        const bool use = LOWER ? (c[u][k] >= 0 && c[u][k] < i) : (c[u][k] > i);
        if (use) tt -= v[u][k];
      }
      if (LOWER) {
        const T zi = di[u] * tt;
        z[i] = zi;
        if (DOT) part += (double)t[u] * (double)zi;
      } else {
        z[i] = tt;
        part += (double)r[i] * (double)tt;
      }
    }
  }
  return part;
}

// z = (LU)^-1 r colour by colour (forward), then in reverse (backward);
// returns <r, z> over the team.  The last colour's forward step also takes
// its <r, z> partial (its backward step, which subtracts nothing) with no
// barrier between, as the same threads hold the same rows.  With c0 = 1 the
// forward step of colour 0 (z_i = Dinv_i r_i: its rows have no earlier
// neighbours) was made by xr_update; callers pass it only with two or more
// colours.  The partial sums are added in the order of the colour-by-colour
// backward sweep, so the results are bitwise those of separate steps.
template <int W, typename V, typename T>
__device__ inline double ell_precond(Team& tm, const SysDesc& d, const int* __restrict__ cp,
                                     const int* __restrict__ ecol, const V* __restrict__ f,
                                     const V* __restrict__ fdia, const T* r, T* z, int c0 = 0) {
  int a, b;
  const int last = d.ncol - 1;
  for (int c = c0; c < last; c++) {
    tshare(tm, cp[c], cp[c + 1], a, b);
    ell_subst<W, true, V, T>(d, a, b, ecol, f, fdia, r, z);
    tsync(tm);
  }
  tshare(tm, cp[last], cp[last + 1], a, b);
  double part = ell_subst<W, true, V, T, true>(d, a, b, ecol, f, fdia, r, z);
  for (int c = last - 1; c >= 0; c--) {
    tsync(tm);
    tshare(tm, cp[c], cp[c + 1], a, b);
    part += ell_subst<W, false, V, T>(d, a, b, ecol, f, fdia, r, z);
  }
  return tsum(tm, part);
}

// p = beta p + z over [a, b)
template <int W, typename T>
__device__ inline void p_update(int a, int b, T beta, const T* z, T* p) {
  constexpr int U = kStreamRows;
  for (int i0 = a + (int)threadIdx.x; i0 < b; i0 += U * (int)blockDim.x) {
    T pv[U], zv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < b) {
        pv[u] = p[i];
        zv[u] = z[i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < b) p[i] = beta * pv[u] + zv[u];
    }
  }
}

// x += alpha p; r -= alpha q over [a, b); returns the <r, r> partial.  With
// fdia0 (the system's inverted pivots) the rows below n0 (colour 0) also get
// z_i = Dinv_i r_i, the forward substitution step of colour 0, from the new r.
template <int W, typename T, typename V = T>
__device__ inline double xr_update(int a, int b, T alpha, const T* p, const T* q, T* x, T* r, int n0 = 0,
                                   const V* __restrict__ fdia0 = nullptr, T* z = nullptr) {
  constexpr int U = kStreamRows;
  const T malpha = -alpha;
  double rr = 0.0;
  for (int i0 = a + (int)threadIdx.x; i0 < b; i0 += U * (int)blockDim.x) {
    T xv[U], pv[U], rv[U], qv[U], dv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < b) {
        xv[u] = x[i];
        pv[u] = p[i];
        rv[u] = r[i];
        qv[u] = q[i];
        if (fdia0 && i < n0) dv[u] = (T)fdia0[i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i >= b) continue;
      x[i] = xv[u] + alpha * pv[u];
      const T rn = rv[u] + malpha * qv[u];
      r[i] = rn;
      rr += (double)rn * (double)rn;
      if (fdia0 && i < n0) z[i] = dv[u] * rn;
    }
  }
  return rr;
}

// The reference's PCG loop on this system (FP64 copies, x from 0).  Stops
// when <r,r> < tol2 <r0,r0>; returns the reference's iteration value and
// leaves <r,r>/<r0,r0> in rel2.
template <int W>
__device__ inline int wg_pcg64(Team& tm, const SysDesc& d, const int* cp, const int* ecol, const double* a,
                               const double* adia, const double* f, const double* fdia, double* r, double* x, double* z,
                               double* p, double* q, double tol2, int maxit, double& rel2) {
  int ra, rb;
  tshare(tm, 0, d.n, ra, rb);
  double rr0 = 0.0;
  for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
    x[i] = 0.0;
    rr0 += r[i] * r[i];
  }
  rr0 = tsum(tm, rr0);
  rel2 = 0.0;
  if (rr0 < 1.0e-30) return 0;
  const double inv_rr0 = 1.0 / rr0;
  // colour 0's forward step taken with the r update (two or more colours)
  const bool z0 = d.ncol > 1;
  double rz = ell_precond<W, double, double>(tm, d, cp, ecol, f, fdia, r, z);
  for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) p[i] = z[i];
  tsync(tm);
  rel2 = 1.0;
  for (int it = 0; it < maxit; it++) {
    const double pap = tsum(tm, ell_spmv<W, double, double>(d, ra, rb, ecol, a, adia, p, q));
    const double alpha = rz / pap;
    const double rr = tsum(tm, xr_update<W, double>(ra, rb, alpha, p, q, x, r, cp[1], z0 ? fdia + d.row0 : nullptr, z));
    rel2 = rr * inv_rr0;
    if (rel2 < tol2) return it;
    const double rz_new = ell_precond<W, double, double>(tm, d, cp, ecol, f, fdia, r, z, z0 ? 1 : 0);
    const double beta = rz_new / rz;
    rz = rz_new;
    p_update<W, double>(ra, rb, beta, z, p);
    tsync(tm);
  }
  return maxit;
}

// r64 = b - A64 x64 over this member's rows (r64 may be null); returns <r64, r64> over the team
template <int W>
__device__ inline double true_res(Team& tm, const SysDesc& d, const int* ecol, const double* a64,
                                  const double* adia64, const double* b, const double* x, double* r64) {
  int ra, rb;
  tshare(tm, 0, d.n, ra, rb);
  double part = 0.0;
  for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
    int c[W];
    ell_cols<W>(ecol, d, i, c);
    double v[W];
#pragma unroll
    for (int k = 0; k < W; k++) v[k] = c[k] >= 0 ? a64[d.ell0 + (long long)k * d.n + i] * x[c[k]] : 0.0;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < W; k++) acc += v[k];
    acc += adia64[d.row0 + i] * x[i];
    const double rv = b[i] - acc;
    if (r64) r64[i] = rv;
    part += rv * rv;
  }
  return tsum(tm, part);
}

// FP32 PCG with FP64 reliable updates: the FP32 recurrence runs on the FP32
// matrix and factor; when its residual has dropped by delta^2 since the last
// update (or reached the tolerance) the partial solution is added to the FP64
// solution, the FP64 true residual replaces the recursive one, and the
// iteration continues (restarted from the true residual when the recursion
// had run more than 10x ahead of it).  Converged on the FP64 true residual.
template <int W>
__device__ inline void wg_pcg_mixed(Team& tm, const SysDesc& d, const int* cp, const int* ecol, const BatchArgs& g,
                                    const double* b, double* x64, double* r64, float* r, float* xr, float* z,
                                    float* p, float* q, int& iters, int& nupd, int& status, double& relres) {
  int ra, rb;
  tshare(tm, 0, d.n, ra, rb);
  double bb = 0.0;
  for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
    x64[i] = 0.0;
    xr[i] = 0.0f;
    r[i] = (float)b[i];
    bb += b[i] * b[i];
  }
  bb = tsum(tm, bb);
  iters = 0;
  nupd = 0;
  if (bb < 1.0e-30) {
    status = 1;
    relres = 0.0;
    return;
  }
  const double delta2 = 0.01;
  const bool z0 = d.ncol > 1;  // colour 0's forward step taken with the r update
  double rz = ell_precond<W, float, float>(tm, d, cp, ecol, g.f32, g.fdia32, r, z);
  for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) p[i] = z[i];
  tsync(tm);
  double rr_ref = bb, rr_true = bb;
  bool done = false;
  int it = 0;
  for (; it < g.maxit && !done; it++) {
    const double pap = tsum(tm, ell_spmv<W, float, float>(d, ra, rb, ecol, g.a32, g.adia32, p, q));
    const float alpha = (float)(rz / pap);
    const double rr =
        tsum(tm, xr_update<W, float>(ra, rb, alpha, p, q, xr, r, cp[1], z0 ? g.fdia32 + d.row0 : nullptr, z));
    bool restart = false;
    int c0 = z0 ? 1 : 0;
    if (rr < delta2 * rr_ref || rr < g.tol2 * bb) {
      for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
        x64[i] += (double)xr[i];
        xr[i] = 0.0f;
      }
      tsync(tm);
      rr_true = true_res<W>(tm, d, ecol, g.a64, g.adia64, b, x64, r64);
      nupd++;
      if (rr_true < g.tol2 * bb) {
        done = true;
        continue;
      }
      for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) r[i] = (float)r64[i];
      tsync(tm);
      restart = rr_true > 100.0 * rr;
      rr_ref = rr_true;
      c0 = 0;  // r replaced: colour 0's forward step again
    }
    const double rz_new = ell_precond<W, float, float>(tm, d, cp, ecol, g.f32, g.fdia32, r, z, c0);
    const float beta = restart ? 0.0f : (float)(rz_new / rz);
    rz = rz_new;
    p_update<W, float>(ra, rb, beta, z, p);
    tsync(tm);
  }
  iters = it;
  if (!done) {
    for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) x64[i] += (double)xr[i];
    tsync(tm);
    rr_true = true_res<W>(tm, d, ecol, g.a64, g.adia64, b, x64, r64);
  }
  status = rr_true < g.tol2 * bb ? 1 : 0;
  relres = sqrt(rr_true / bb);
}

// ILU(0) numeric of one system in place (b = 1 branch of DoILUDecomp: for
// each k of L(i) in ascending order, with row k finished, a_ij -= a_ik u_kj
// over U(k); pivot inverted; U row scaled by it).  Row i and row k are held
// in registers, the fill positions matched by column.  Returns the number
// of singular pivots.
template <int W>
__device__ inline int wg_ilu0(Team& tm, const SysDesc& d, const int* cp, const int* __restrict__ ecol,
                              const double* __restrict__ a64, double* f64, const double* __restrict__ adia64,
                              double* fdia64) {
  double nsing = 0.0;
  int ra, rb;
  for (int c = 0; c < d.ncol; c++) {
    tshare(tm, cp[c], cp[c + 1], ra, rb);
    for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
      int ci[W];
      double fi[W];
      ell_cols<W>(ecol, d, i, ci);
#pragma unroll
      for (int k = 0; k < W; k++) fi[k] = ci[k] >= 0 ? a64[d.ell0 + (long long)k * d.n + i] : 0.0;
      double dii = adia64[d.row0 + i];
#pragma unroll
      for (int qk = 0; qk < W; qk++) {
        const int k = ci[qk];
        if (k < 0 || k > i) continue;
        const double L = fi[qk];
        int ck[W];
        double uk[W];
        ell_cols<W>(ecol, d, k, ck);
#pragma unroll
        for (int qj = 0; qj < W; qj++) uk[qj] = ck[qj] > k ? f64[d.ell0 + (long long)qj * d.n + k] : 0.0;
#pragma unroll
        for (int qj = 0; qj < W; qj++) {
          const int j = ck[qj];
          if (j <= k) continue;
          if (j == i) {
            dii -= L * uk[qj];
            continue;
          }
#pragma unroll
          for (int pp = 0; pp < W; pp++)
            if (pp > qk && ci[pp] == j) fi[pp] -= L * uk[qj];
        }
      }
      if (fabs(dii) > 1.0e-30) dii = 1.0 / dii;
      else nsing += 1.0;
      fdia64[d.row0 + i] = dii;
#pragma unroll
      for (int k = 0; k < W; k++)
        if (ci[k] >= 0) f64[d.ell0 + (long long)k * d.n + i] = ci[k] > i ? fi[k] * dii : fi[k];
    }
    tsync(tm);
  }
  return (int)tsum(tm, nsing);
}

template <int W, bool MIXED>
__global__ void __launch_bounds__(1024) k_batch_solve(BatchArgs g) {
  __shared__ double sh[2][16];
  __shared__ double s_bc;
  __shared__ Item s_item;
  __shared__ unsigned long long s_t0;
  int buf = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      Item item{-1, 0, 1, 0};
      if (g.local) {
        // this XCD's teams first, then the single items of every XCD
        const int x = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kXcd - 1);  // HW_REG_XCC_ID
        const unsigned nt = (unsigned)(g.xq[x + 1] - g.xq[x]);
        unsigned* cx = g.next + x * kCounterStride;
        unsigned t = nt;
        if (__hip_atomic_load(cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nt) t = atomicAdd(cx, 1u);
        if (t < nt) {
          item = g.items[g.xq[x] + t];
        } else {
          const unsigned ns = (unsigned)(g.nitem - g.xq[kXcd]);
          t = atomicAdd(g.next + kXcd * kCounterStride, 1u);
          if (t < ns) item = g.items[g.xq[kXcd] + t];
        }
      } else {
        const unsigned t = atomicAdd(g.next, 1u);
        if (t < (unsigned)g.nitem) item = g.items[t];
      }
      s_item = item;
      s_t0 = wall_clock64();
    }
    __syncthreads();
    const Item it = s_item;
    const unsigned long long t0 = s_t0;
    __syncthreads();
    if (it.sys < 0) break;
    const int s = it.sys;
    Team tm{it.team, it.member, g.local, g.teams + s, g.fail, 0u, 0, sh, buf, &s_bc};
    const SysDesc d = g.sys[s];
    const int* cp = g.cptr + (size_t)s * (kMaxColourBatch + 1);
    const int nsing = wg_ilu0<W>(tm, d, cp, g.ecol, g.a64, g.f64, g.adia64, g.fdia64);
    int iters = 0, outer = 0, status = 0;
    double relres = 0.0;
    const double* b = g.b64 + d.row0;
    double* x = g.x64 + d.row0;
    int ra, rb;
    tshare(tm, 0, d.n, ra, rb);
    if (nsing > 10) {
      status = -1;
    } else if (!MIXED) {
      double* vb = reinterpret_cast<double*>(g.vec);
      double *r = vb + d.row0, *z = vb + g.nrow + d.row0, *p = vb + 2 * g.nrow + d.row0,
             *q = vb + 3 * g.nrow + d.row0;
      for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) r[i] = b[i];
      tsync(tm);
      double rel2 = 0.0;
      iters = wg_pcg64<W>(tm, d, cp, g.ecol, g.a64, g.adia64, g.f64, g.fdia64, r, x, z, p, q, g.tol2, g.maxit, rel2);
      status = rel2 < g.tol2 ? 1 : 0;
      double bb = 0.0;
      for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) bb += b[i] * b[i];
      bb = tsum(tm, bb);
      const double rt = true_res<W>(tm, d, g.ecol, g.a64, g.adia64, b, x, nullptr);
      relres = bb > 0.0 ? sqrt(rt / bb) : 0.0;
    } else {
      // FP32 copies of this system's matrix and factor, then the refined solve
      for (int i = ra + (int)threadIdx.x; i < rb; i += blockDim.x) {
        for (int k = 0; k < d.w; k++) {
          const long long qq = d.ell0 + (long long)k * d.n + i;
          g.a32[qq] = (float)g.a64[qq];
          g.f32[qq] = (float)g.f64[qq];
        }
        g.adia32[d.row0 + i] = (float)g.adia64[d.row0 + i];
        g.fdia32[d.row0 + i] = (float)g.fdia64[d.row0 + i];
      }
      tsync(tm);
      float* vf = reinterpret_cast<float*>(g.vec);
      float *r = vf + d.row0, *z = vf + g.nrow + d.row0, *p = vf + 2 * g.nrow + d.row0, *q = vf + 3 * g.nrow + d.row0,
            *xr = vf + 4 * g.nrow + d.row0;
      wg_pcg_mixed<W>(tm, d, cp, g.ecol, g, b, x, g.r64 + d.row0, r, xr, z, p, q, iters, outer, status, relres);
    }
    buf = tm.buf;
    if (threadIdx.x == 0 && tm.m == 0) {
      BatchOut o;
      o.iters = iters;
      o.outer = outer;
      o.status = status;
      o.nsing = nsing;
      o.relres = relres;
      o.t0 = t0;
      o.t1 = wall_clock64();
      g.out[s] = o;
    }
    __syncthreads();
  }
}

// Work counter, failure flag and every system's team counter back to zero.
__global__ void k_batch_reset(unsigned* next, int* fail, TeamState* teams, int nsys) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= kXcd) next[t * kCounterStride] = 0u;
  if (t == 0) *fail = 0;
  if (t < nsys) teams[t].arrive = 0u;
}
__global__ void k_batch_gather(long long nslot, const long long* __restrict__ src, const double* __restrict__ val,
                               double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nslot) return;
  const long long s = src[t];
  out[t] = s >= 0 ? val[s] : 0.0;
}
__global__ void k_batch_permute(long n, const int* __restrict__ perm, const double* __restrict__ in,
                                double* __restrict__ out) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) out[t] = in[perm[t]];
}
__global__ void k_batch_unpermute(long n, const int* __restrict__ perm, const double* __restrict__ in,
                                  double* __restrict__ out) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) out[perm[t]] = in[t];
}

static int grid256(long long n) { return (int)std::max<long long>(1, (n + 255) / 256); }

// ---------------------------------------------------------------------------
// Host side
static int ncu_count();

// Work items for a launch shape: the largest systems (at least team_min_rows
// rows) get teams of 2..team_max workgroups while the team members fit in
// the workgroups resident at once less a reserve; every other system is one
// item.  Teams come first, so all their members are taken by resident
// workgroups at the start of the launch and no member waits for a partner
// that cannot start.  XCD-local mode (workgroups go to the XCDs round-robin,
// blockIdx % 8; DFM_BATCH_XCD=0 turns it off): each team is given to one XCD
// (the one with the least work so far, by rows^1.5, that still has room for
// it) and queued there; its members are then taken by that XCD's workgroups
// and meet through its L2.  The single items stay in one queue for all.
static bool batch_xcd_local() {
  static const bool on = [] {
    const char* e = getenv("DFM_BATCH_XCD");
    return !(e && e[0] == '0');
  }();
  return on && ncu_count() % kXcd == 0;
}
template <int W, bool MIXED>
static int build_items(dfm_batch_s* B) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_batch_solve<W, MIXED>),
                                                   B->block, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  // workgroups running at once: the launch has wg_per_cu per CU, occupancy permitting
  const int resident = std::min(per_cu, std::max(1, B->wpc)) * ncu_count();
  const bool local = batch_xcd_local();
  char key[96];
  snprintf(key, sizeof key, "%d:%d:%d:%d:%u:%d:%d", W, (int)MIXED, B->block, B->wpc, B->team_min_rows, B->team_max,
           (int)local);
  if (B->items_key == key) return DFM_OK;
  std::vector<Item> singles;
  std::vector<std::vector<Item>> xq(local ? kXcd : 1);
  std::vector<int> budget(xq.size());
  std::vector<double> cost(xq.size(), 0.0);
  const int per_q = resident / (int)xq.size();
  for (auto& bgt : budget) bgt = per_q - std::max(local ? 2 : 16, per_q / 4);
  for (int s : B->sched) {
    const int n = B->desc[s].n;
    int K = 1;
    if (B->team_max > 1 && (unsigned)n >= B->team_min_rows)
      K = std::min({B->team_max, kMaxTeam, std::max(2, n / (int)std::max(1u, B->team_min_rows / 4))});
    int x = -1;
    if (K > 1)
      for (int c = 0; c < (int)xq.size(); c++)
        if (budget[c] >= K && (x < 0 || cost[c] < cost[x])) x = c;
    if (x >= 0) {
      budget[x] -= K;
      cost[x] += (double)n * std::sqrt((double)n);
      for (int m = 0; m < K; m++) xq[x].push_back(Item{s, m, K, 0});
    } else {
      singles.push_back(Item{s, 0, 1, 0});
    }
  }
  std::vector<Item> items;
  B->xq.assign(kXcd + 1, 0);
  for (size_t c = 0; c < xq.size(); c++) {
    items.insert(items.end(), xq[c].begin(), xq[c].end());
    if (local) B->xq[c + 1] = (int)items.size();
  }
  if (!local)
    for (int c = 1; c <= kXcd; c++) B->xq[c] = (int)items.size();
  items.insert(items.end(), singles.begin(), singles.end());
  B->nitem = (int)items.size();
  B->local = local;
  B->resident = resident;
  DFM_TRY(B->d_items.upload(reinterpret_cast<const char*>(items.data()), items.size() * sizeof(Item)));
  DFM_HIP(hipStreamSynchronize(stream()));
  B->items_key = key;
  return DFM_OK;
}

static int batch_args(dfm_batch_s* B, bool mixed, double tol, unsigned maxit, BatchArgs& g) {
  g.nitem = B->nitem;
  g.local = B->local ? 1 : 0;
  for (int c = 0; c <= kXcd; c++) g.xq[c] = B->xq[c];
  g.items = reinterpret_cast<const Item*>(B->d_items.p);
  g.teams = reinterpret_cast<TeamState*>(B->d_teams.p);
  g.fail = B->d_fail.p;
  g.sys = B->d_sys.p;
  g.cptr = B->d_cptr.p;
  g.next = B->d_next.p;
  g.ecol = B->d_ell_col.p;
  g.a64 = B->d_a64.p;
  g.f64 = B->d_f64.p;
  g.adia64 = B->d_adia64.p;
  g.fdia64 = B->d_fdia64.p;
  g.a32 = B->d_a32.p;
  g.f32 = B->d_f32.p;
  g.adia32 = B->d_adia32.p;
  g.fdia32 = B->d_fdia32.p;
  g.b64 = B->d_b64.p;
  g.x64 = B->d_x64.p;
  g.r64 = B->d_r64.p;
  g.nrow = (long)B->nrow;
  const size_t vbytes = (size_t)B->nrow * 5 * (mixed ? sizeof(float) : sizeof(double));
  if (B->d_vec.n < vbytes) DFM_TRY(B->d_vec.alloc(vbytes));
  g.vec = B->d_vec.p;
  g.tol2 = tol * tol;
  g.maxit = (int)std::min<unsigned>(maxit, 0x7fffffffu);
  g.out = B->d_out.p;
  return DFM_OK;
}

static int ncu_count() {
  static int ncu = 0;
  if (ncu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
      ncu = 1;
  }
  return ncu;
}

// The solve step (work counter reset + persistent launch), replayed from a
// graph per (precision, launch shape, tolerance, iteration limit).
static int batch_launch(dfm_batch_s* B, bool mixed, double tol, unsigned maxit) {
  const int W = B->max_w <= 6 ? 6 : (B->max_w <= 8 ? 8 : 16);
  if (W == 6) DFM_TRY(mixed ? (build_items<6, true>(B)) : (build_items<6, false>(B)));
  else if (W == 8) DFM_TRY(mixed ? (build_items<8, true>(B)) : (build_items<8, false>(B)));
  else DFM_TRY(mixed ? (build_items<16, true>(B)) : (build_items<16, false>(B)));
  BatchArgs g;
  DFM_TRY(batch_args(B, mixed, tol, maxit, g));
  // XCD-local teams need every XCD's workgroups, so the grid is not cut to the item count
  const int grid = B->local ? ncu_count() * B->wpc : std::min<int>(B->nitem, ncu_count() * B->wpc);
  char key[200];
  snprintf(key, sizeof key, "%d:%d:%d:%.17g:%u:%p:%s", (int)mixed, B->block, B->wpc, tol, maxit, g.vec,
           B->items_key.c_str());
  hipStream_t s = stream();
  for (auto& e : B->graphs)
    if (e.first == key) {
      DFM_HIP(hipGraphLaunch(e.second, s));
      return DFM_OK;
    }
  auto body = [&]() -> int {
    hipLaunchKernelGGL(k_batch_reset, dim3((B->nsys + 255) / 256), dim3(256), 0, s, B->d_next.p, B->d_fail.p,
                       reinterpret_cast<TeamState*>(B->d_teams.p), (int)B->nsys);
    DFM_HIP(hipGetLastError());
    if (B->max_w <= 6) {
      if (mixed) hipLaunchKernelGGL((k_batch_solve<6, true>), dim3(grid), dim3(B->block), 0, s, g);
      else hipLaunchKernelGGL((k_batch_solve<6, false>), dim3(grid), dim3(B->block), 0, s, g);
    } else if (B->max_w <= 8) {
      if (mixed) hipLaunchKernelGGL((k_batch_solve<8, true>), dim3(grid), dim3(B->block), 0, s, g);
      else hipLaunchKernelGGL((k_batch_solve<8, false>), dim3(grid), dim3(B->block), 0, s, g);
    } else {
      if (mixed) hipLaunchKernelGGL((k_batch_solve<16, true>), dim3(grid), dim3(B->block), 0, s, g);
      else hipLaunchKernelGGL((k_batch_solve<16, false>), dim3(grid), dim3(B->block), 0, s, g);
    }
    DFM_HIP(hipGetLastError());
    return DFM_OK;
  };
  if (s == nullptr) return body();
  hipGraph_t graph = nullptr;
  DFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != DFM_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return body();
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DFM_HIP(e);
  B->graphs.emplace_back(key, exec);
  DFM_HIP(hipGraphLaunch(exec, s));
  return DFM_OK;
}

}  // namespace dfm

using namespace dfm;

extern "C" {

int dfm_batch_create(const dfm_bcrs* A, unsigned nsys, const unsigned* sys_ptr, dfm_batch** out) {
  DFM_CHECK(A && sys_ptr && out && nsys > 0, DFM_BAD_ARG, "bad arguments to dfm_batch_create");
  DFM_CHECK(A->b == 1, DFM_NOT_IMPLEMENTED, "the batched solver takes 1x1 blocks (b = %u)", A->b);
  DFM_CHECK(sys_ptr[0] == 0 && sys_ptr[nsys] == A->nb, DFM_BAD_ARG,
            "sys_ptr must start at 0 and end at the block count %u", A->nb);
  for (unsigned s = 0; s < nsys; s++)
    DFM_CHECK(sys_ptr[s + 1] > sys_ptr[s], DFM_BAD_ARG, "system %u is empty", s);
  DFM_TRY(ensure_init());
  std::unique_ptr<dfm_batch_s> B(new dfm_batch_s);
  B->A = A;
  B->nsys = nsys;
  B->nrow = A->nb;
  B->sys_ptr.assign(sys_ptr, sys_ptr + nsys + 1);
  B->perm.resize(A->nb);
  B->desc.resize(nsys);
  B->cptr.assign((size_t)nsys * (kMaxColourBatch + 1), 0);
  long long ell_total = 0;
  // first pass: colouring, widths
  std::vector<std::vector<int>> orders(nsys);
  for (unsigned s = 0; s < nsys; s++) {
    const unsigned r0 = sys_ptr[s], n = sys_ptr[s + 1] - r0;
    std::vector<unsigned> lrow(n + 1, 0), lcol;
    lcol.reserve(A->rowptr[r0 + n] - A->rowptr[r0]);
    int w = 0;
    for (unsigned i = 0; i < n; i++) {
      for (unsigned q = A->rowptr[r0 + i]; q < A->rowptr[r0 + i + 1]; q++) {
        const unsigned c = A->col[q];
        DFM_CHECK(c >= r0 && c < r0 + n, DFM_BAD_ARG,
                  "row %u couples system %u to row %u outside it: the systems must be diagonal blocks", r0 + i, s, c);
        lcol.push_back(c - r0);
      }
      lrow[i + 1] = (unsigned)lcol.size();
      w = std::max(w, (int)(lrow[i + 1] - lrow[i]));
    }
    std::vector<int> colour;
    const int nc = greedy_colour(n, lrow, lcol, colour);
    DFM_CHECK(nc <= kMaxColourBatch, DFM_NOT_IMPLEMENTED, "system %u needs %d colours (at most %d)", s, nc,
              kMaxColourBatch);
    // the colour-by-colour sweeps need no two coupled rows in one colour (a
    // structurally unsymmetric pattern could leave such a pair)
    for (unsigned i = 0; i < n; i++)
      for (unsigned q = lrow[i]; q < lrow[i + 1]; q++)
        DFM_CHECK(lcol[q] == i || colour[lcol[q]] != colour[i], DFM_NOT_IMPLEMENTED,
                  "system %u: rows %u and %u are coupled one way only (structurally unsymmetric pattern)", s, i,
                  lcol[q]);
    colour_order(n, colour, nc, orders[s]);
    int* cp = B->cptr.data() + (size_t)s * (kMaxColourBatch + 1);
    for (unsigned i = 0; i < n; i++) cp[colour[i] + 1]++;
    for (int c = 0; c < nc; c++) cp[c + 1] += cp[c];
    for (int c = nc; c < kMaxColourBatch; c++) cp[c + 1] = cp[c];
    DFM_CHECK(w <= 16, DFM_NOT_IMPLEMENTED, "system %u has a row with %d off-diagonal entries (at most 16)", s, w);
    B->desc[s] = SysDesc{(int)r0, (int)n, w, nc, ell_total};
    ell_total += (long long)w * n;
    B->max_w = std::max(B->max_w, w);
    B->max_ncol = std::max(B->max_ncol, nc);
  }
  // second pass: ELL columns (new numbering, ascending) and sources in A
  std::vector<int> ecol((size_t)std::max<long long>(1, ell_total), -1);
  std::vector<long long> esrc((size_t)std::max<long long>(1, ell_total), -1);
  for (unsigned s = 0; s < nsys; s++) {
    const SysDesc& d = B->desc[s];
    const std::vector<int>& ord = orders[s];
    std::vector<int> inv(d.n);
    for (int i = 0; i < d.n; i++) inv[ord[i]] = i;
    std::vector<std::pair<int, long long>> ent;
    for (int i = 0; i < d.n; i++) {
      const unsigned orow = (unsigned)d.row0 + (unsigned)ord[i];
      B->perm[d.row0 + i] = (int)orow;
      ent.clear();
      for (unsigned q = A->rowptr[orow]; q < A->rowptr[orow + 1]; q++)
        ent.emplace_back(inv[A->col[q] - d.row0], (long long)q);
      std::sort(ent.begin(), ent.end());
      for (size_t k = 0; k < ent.size(); k++) {
        const size_t slot = (size_t)d.ell0 + k * (size_t)d.n + i;
        ecol[slot] = ent[k].first;
        esrc[slot] = ent[k].second;
      }
    }
  }
  // schedule: most work (rows * width) first
  std::vector<int> sched(nsys);
  std::iota(sched.begin(), sched.end(), 0);
  std::stable_sort(sched.begin(), sched.end(), [&](int a, int b) {
    return (long long)B->desc[a].n * (B->desc[a].w + 1) > (long long)B->desc[b].n * (B->desc[b].w + 1);
  });
  const size_t nslot = (size_t)std::max<long long>(1, ell_total), nr = A->nb;
  DFM_TRY(B->d_sys.upload(B->desc));
  DFM_TRY(B->d_cptr.upload(B->cptr));
  B->sched = sched;
  DFM_TRY(B->d_teams.alloc((size_t)nsys * sizeof(TeamState)));
  DFM_TRY(B->d_fail.alloc(1));
  DFM_TRY(B->d_perm.upload(B->perm));
  DFM_TRY(B->d_ell_col.upload(ecol));
  DFM_TRY(B->d_ell_src.upload(esrc));
  DFM_TRY(B->d_next.alloc((kXcd + 1) * kCounterStride));
  DFM_TRY(B->d_a64.alloc(nslot));
  DFM_TRY(B->d_f64.alloc(nslot));
  DFM_TRY(B->d_a32.alloc(nslot));
  DFM_TRY(B->d_f32.alloc(nslot));
  DFM_TRY(B->d_adia64.alloc(nr));
  DFM_TRY(B->d_fdia64.alloc(nr));
  DFM_TRY(B->d_adia32.alloc(nr));
  DFM_TRY(B->d_fdia32.alloc(nr));
  DFM_TRY(B->d_b64.alloc(nr));
  DFM_TRY(B->d_x64.alloc(nr));
  DFM_TRY(B->d_r64.alloc(nr));
  DFM_TRY(B->d_out.alloc(nsys));
  DFM_HIP(hipStreamSynchronize(stream()));
  *out = B.release();
  return DFM_OK;
}

int dfm_batch_destroy(dfm_batch* B) {
  delete B;
  return DFM_OK;
}

int dfm_batch_info(const dfm_batch* B, unsigned* max_width, int* max_colours, int* block, int* wg_per_cu) {
  DFM_CHECK(B, DFM_BAD_ARG, "NULL batch");
  if (max_width) *max_width = (unsigned)B->max_w;
  if (max_colours) *max_colours = B->max_ncol;
  if (block) *block = B->block;
  if (wg_per_cu) *wg_per_cu = B->wpc;
  return DFM_OK;
}

int dfm_batch_set_value(dfm_batch* B, const dfm_bcrs* A) {
  DFM_CHECK(B && A && A == B->A, DFM_BAD_ARG, "dfm_batch_set_value needs the matrix the batch was made from");
  const long long nslot = (long long)B->d_a64.n;
  hipLaunchKernelGGL(k_batch_gather, dim3(grid256(nslot)), dim3(256), 0, stream(), nslot, B->d_ell_src.p,
                     A->d_val.p, B->d_a64.p);
  DFM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_batch_permute, dim3(grid256(B->nrow)), dim3(256), 0, stream(), (long)B->nrow, B->d_perm.p,
                     A->d_dia.p, B->d_adia64.p);
  DFM_HIP(hipGetLastError());
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

int dfm_batch_set_launch(dfm_batch* B, int block, int wg_per_cu) {
  DFM_CHECK(B && (block == 256 || block == 512 || block == 1024) && wg_per_cu >= 1 && wg_per_cu <= 8 &&
                block * wg_per_cu <= 2048,
            DFM_BAD_ARG, "launch shape: block 256/512/1024, 1..8 workgroups per CU, at most 2048 threads per CU");
  B->block = block;
  B->wpc = wg_per_cu;
  return DFM_OK;
}

int dfm_batch_set_teams(dfm_batch* B, unsigned min_rows, int max_team) {
  DFM_CHECK(B && max_team >= 1 && max_team <= kMaxTeam, DFM_BAD_ARG, "max_team must be 1..%d", kMaxTeam);
  B->team_min_rows = std::max(1u, min_rows);
  B->team_max = max_team;
  B->items_key.clear();
  return DFM_OK;
}

int dfm_batch_solve_pcg(dfm_batch* B, const double* rhs, double* x, double tol, unsigned maxit, int precision,
                        unsigned* iters, int* outer, int* status, double* relres, double* ms) {
  DFM_CHECK(B && rhs && x && tol > 0.0 && maxit > 0, DFM_BAD_ARG, "bad arguments to dfm_batch_solve_pcg");
  DFM_CHECK(precision == DFM_PREC_FP64 || precision == DFM_PREC_MIXED_F32, DFM_BAD_ARG,
            "the batched solver runs DFM_PREC_FP64 or DFM_PREC_MIXED_F32 (FP32 with FP64 refinement)");
  const bool mixed = precision == DFM_PREC_MIXED_F32;
  hipLaunchKernelGGL(k_batch_permute, dim3(grid256(B->nrow)), dim3(256), 0, stream(), (long)B->nrow, B->d_perm.p,
                     rhs, B->d_b64.p);
  DFM_HIP(hipGetLastError());
  DFM_TRY(batch_launch(B, mixed, tol, maxit));
  hipLaunchKernelGGL(k_batch_unpermute, dim3(grid256(B->nrow)), dim3(256), 0, stream(), (long)B->nrow,
                     B->d_perm.p, B->d_x64.p, x);
  DFM_HIP(hipGetLastError());
  std::vector<BatchOut> o(B->nsys);
  DFM_TRY(B->d_out.download(o.data(), B->nsys));
  int fail = 0;
  DFM_TRY(B->d_fail.download(&fail, 1));
  DFM_CHECK(!fail, DFM_HIP_ERROR, "a team barrier of the batched solve timed out (results invalid)");
  int dev = 0, khz = 0;
  DFM_HIP(hipGetDevice(&dev));
  DFM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  unsigned long long tmin = ~0ull;
  for (const BatchOut& e : o) tmin = std::min(tmin, e.t0);
  bool all = true;
  for (unsigned s = 0; s < B->nsys; s++) {
    if (iters) iters[s] = (unsigned)o[s].iters;
    if (outer) outer[s] = o[s].outer;
    if (status) status[s] = o[s].status;
    if (relres) relres[s] = o[s].relres;
    if (ms) {
      ms[2 * s] = khz > 0 ? (double)(o[s].t1 - o[s].t0) / khz : 0.0;     // solve time on its workgroup
      ms[2 * s + 1] = khz > 0 ? (double)(o[s].t1 - tmin) / khz : 0.0;   // completion since the launch began
    }
    all = all && o[s].status == 1;
  }
  return all ? DFM_OK : DFM_NOT_CONVERGED;
}

int dfm_batch_get_factor(const dfm_batch* B, unsigned sys, double* val, double* dia, int* order) {
  DFM_CHECK(B && sys < B->nsys, DFM_BAD_ARG, "bad arguments to dfm_batch_get_factor");
  const SysDesc& d = B->desc[sys];
  std::vector<double> f((size_t)d.w * d.n), fd(d.n);
  std::vector<int> c((size_t)d.w * d.n);
  if (!f.empty()) {
    DFM_HIP(hipMemcpy(f.data(), B->d_f64.p + d.ell0, f.size() * sizeof(double), hipMemcpyDeviceToHost));
    DFM_HIP(hipMemcpy(c.data(), B->d_ell_col.p + d.ell0, c.size() * sizeof(int), hipMemcpyDeviceToHost));
  }
  DFM_HIP(hipMemcpy(fd.data(), B->d_fdia64.p + d.row0, fd.size() * sizeof(double), hipMemcpyDeviceToHost));
  size_t t = 0;
  for (int i = 0; i < d.n; i++)
    for (int k = 0; k < d.w; k++) {
      const size_t slot = (size_t)k * d.n + i;
      if (c[slot] < 0) break;
      if (val) val[t] = f[slot];
      t++;
    }
  if (dia) std::copy(fd.begin(), fd.end(), dia);
  if (order)
    for (int i = 0; i < d.n; i++) order[i] = B->perm[d.row0 + i] - d.row0;
  return DFM_OK;
}

}  // extern "C"
