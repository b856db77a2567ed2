// dfm_solver.hip -- device-resident Krylov solvers: LsSol::Solve_PCG and
// LsSol::Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:100-172, 291-408) with
// the ILU preconditioner, in FP64 and in mixed precision.
//
// All vectors live in factor numbering; the right-hand side is permuted in
// once and the solution and residual are permuted out once.  Scalars
// (alpha, beta, omega, residual norms) never leave the device inside the
// iteration: each kernel reads them from a KState, every reduction is a
// deterministic two-phase sum, and blocks of iterations are replayed as one
// hipGraph.  The host polls the convergence flag once per block.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "dfm_kernels.h"

namespace dfm {

static inline int grid_for(long n, int per = kNT) { return (int)std::max<long>(1, (n + per - 1) / per); }
static inline int grid_vec(long n) { return (int)std::min<long>(std::max<long>(1, (n + kNT - 1) / kNT), 2048); }

#define DFM_LAUNCH(kernel, grid, ...)                                                       \
  do {                                                                                      \
    if ((grid) > 0) {                                                                       \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(kNT), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                           \
    }                                                                                       \
  } while (0)

#define DFM_LAUNCH_NT(kernel, grid, nt, ...)                                              \
  do {                                                                                    \
    if ((grid) > 0) {                                                                     \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(nt), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                         \
    }                                                                                     \
  } while (0)

#define DFM_LAUNCH_FIN(ND, FIN, partials, nparts, st)                                             \
  do {                                                                                            \
    hipLaunchKernelGGL((k_finalize<ND, FIN>), dim3(1), dim3(kFinNT), 0, ::dfm::stream(), partials, \
                       nparts, st);                                                               \
    DFM_HIP(hipGetLastError());                                                                   \
  } while (0)

// ---------------------------------------------------------------------------
// Value copies of the operators in the storage types the solves stream.
template <typename T>
static int sell_alloc_like(const Sell<double>& src, Sell<T>& dst) {
  dst.nslice = src.nslice;
  dst.nk = src.nk;
  dst.bb = src.bb;
  dst.group_slice_ptr = src.group_slice_ptr;
  DFM_TRY(dst.val.alloc(std::max<size_t>(1, (size_t)src.nk * slot_elems<T>(src.bb) * kSlice)));
  DFM_TRY(dst.dval.alloc(std::max<size_t>(1, (size_t)src.nslice * slot_elems<T>(src.bb) * kSlice)));
  return DFM_OK;
}

template <typename TS>
static int gather(Sell<TS>& S, const DevBuf<long long>& src, const double* from, int bb) {
  const long nslot = S.nk * kSlice;
  DFM_LAUNCH(k_sell_gather<TS>, grid_for(nslot), nslot, bb, src.p, from, S.val.p);
  return DFM_OK;
}
template <typename TS>
static int gather_dia(Sell<TS>& S, const int* rows, const int* order, const double* from, int bb) {
  const long nslot = (long)S.nslice * kSlice;
  DFM_LAUNCH(k_sell_gather_dia<TS>, grid_for(nslot), nslot, bb, rows, order, from, S.dval.p);
  return DFM_OK;
}

// Refresh after a factorisation (factor copies) and whenever A's values
// changed (matrix copies).  want32 / want16: create the FP32 / BF16 copies.
int ilu_refresh_sell(dfm_ilu_s* M, const dfm_bcrs_s* A, bool want32, bool want16) {
  const int bb = (int)(M->b * M->b);
  DFM_TRY(gather(M->L64, M->d_srcL, M->d_fval.p, bb));
  DFM_TRY(gather_dia(M->L64, M->L64.rows.p, nullptr, M->d_fdia.p, bb));
  DFM_TRY(gather(M->U64, M->d_srcU, M->d_fval.p, bb));
  if (want32 && !M->have32) {
    DFM_TRY(sell_alloc_like(M->L64, M->L32));
    DFM_TRY(sell_alloc_like(M->U64, M->U32));
    DFM_TRY(sell_alloc_like(M->A64, M->A32));
    M->have32 = true;
    M->a32_valid = false;
  }
  if (want16 && !M->have16) {
    DFM_TRY(sell_alloc_like(M->L64, M->L16));
    DFM_TRY(sell_alloc_like(M->U64, M->U16));
    M->have16 = true;
  }
  if (M->have32) {
    DFM_TRY(gather(M->L32, M->d_srcL, M->d_fval.p, bb));
    DFM_TRY(gather_dia(M->L32, M->L64.rows.p, nullptr, M->d_fdia.p, bb));
    DFM_TRY(gather(M->U32, M->d_srcU, M->d_fval.p, bb));
  }
  if (M->have16) {
    DFM_TRY(gather(M->L16, M->d_srcL, M->d_fval.p, bb));
    DFM_TRY(gather_dia(M->L16, M->L64.rows.p, nullptr, M->d_fdia.p, bb));
    DFM_TRY(gather(M->U16, M->d_srcU, M->d_fval.p, bb));
  }
  for (TileSweep* ts : {&M->tileL, &M->tileU}) {
    if (!ts->ntile) continue;
    const bool lower = ts == &M->tileL;
    DFM_TRY(gather(ts->v64, ts->src, M->d_fval.p, bb));
    if (lower) DFM_TRY(gather_dia(ts->v64, ts->v64.rows.p, nullptr, M->d_fdia.p, bb));
    if (M->have32) {
      if (ts->v32.nslice != ts->v64.nslice) DFM_TRY(sell_alloc_like(ts->v64, ts->v32));
      DFM_TRY(gather(ts->v32, ts->src, M->d_fval.p, bb));
      if (lower) DFM_TRY(gather_dia(ts->v32, ts->v64.rows.p, nullptr, M->d_fdia.p, bb));
    }
    if (M->have16) {
      if (ts->v16.nslice != ts->v64.nslice) DFM_TRY(sell_alloc_like(ts->v64, ts->v16));
      DFM_TRY(gather(ts->v16, ts->src, M->d_fval.p, bb));
      if (lower) DFM_TRY(gather_dia(ts->v16, ts->v64.rows.p, nullptr, M->d_fdia.p, bb));
    }
    if (ts->g) {
      // block row r of entry q: the b values at from[(q*b + r) * b]
      const int b = (int)M->b;
      DFM_TRY(gather(ts->s64, ts->ssrc, M->d_fval.p, b));
      if (lower) DFM_TRY(gather_dia(ts->s64, ts->s64.rows.p, nullptr, M->d_fdia.p, b));
      if (M->have32) {
        if (ts->s32.nslice != ts->s64.nslice) DFM_TRY(sell_alloc_like(ts->s64, ts->s32));
        DFM_TRY(gather(ts->s32, ts->ssrc, M->d_fval.p, b));
        if (lower) DFM_TRY(gather_dia(ts->s32, ts->s64.rows.p, nullptr, M->d_fdia.p, b));
      }
      if (M->have16) {
        if (ts->s16.nslice != ts->s64.nslice) DFM_TRY(sell_alloc_like(ts->s64, ts->s16));
        DFM_TRY(gather(ts->s16, ts->ssrc, M->d_fval.p, b));
        if (lower) DFM_TRY(gather_dia(ts->s16, ts->s64.rows.p, nullptr, M->d_fdia.p, b));
      }
    }
  }
  if (A && (M->a_version != A->version || (M->have32 && !M->a32_valid))) {
    const int* ord = M->permuted ? M->d_order.p : nullptr;
    DFM_TRY(gather(M->A64, M->d_srcA, A->d_val.p, bb));
    DFM_TRY(gather_dia(M->A64, M->A64.rows.p, ord, A->d_dia.p, bb));
    if (M->have32) {
      DFM_TRY(gather(M->A32, M->d_srcA, A->d_val.p, bb));
      DFM_TRY(gather_dia(M->A32, M->A64.rows.p, ord, A->d_dia.p, bb));
      M->a32_valid = true;
    }
    M->a_version = A->version;
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Level-set sweeps, one launch per level.  `plan` carries the layout (rows,
// slices, columns) of the FP64 copy; `L`/`U` the values in storage type TS.
// Plans with many levels take one of two other forms, both given the solver
// state (for the stop flag and, persistent sweeps, the fault report): the
// phase-tiled sweep (k_trsv_tile, one launch per colour) where the factor
// has a phase tiling, else the persistent sweep (k_trsv_persist, one launch).
// Sweep mode (DFM_PERSIST, dfm_set_sweep_mode): 0 one launch per level
// always; 1 (default) plans of at least kPersistMinLevels levels take the
// tiled form if there is one, else the persistent one; 2 the persistent form
// for every plan; 3 the tiled form wherever there is one (per level
// elsewhere).  Persistent grid: two waves per slice of the largest level
// (waves of later levels load their factor slots while the level before them
// finishes), at least 64; DFM_PERSIST_GRID overrides it.
constexpr int kPersistMinLevels = 8;
constexpr int kPersistMaxGrid = 4096;
static int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return e ? atoi(e) : dflt;
}
static int g_persist_mode = env_int("DFM_PERSIST", 1);      // dfm_set_sweep_mode
static int g_persist_grid = env_int("DFM_PERSIST_GRID", 0);
static bool use_tile(const TileSweep* ts, const SweepSync* ps, const KState*) {
  if (!ts || ts->ntile == 0) return false;
  return g_persist_mode >= 3 || (g_persist_mode == 1 && ps && ps->ngroup >= kPersistMinLevels);
}
static bool use_persist(const SweepSync* ps, const KState* st, const TileSweep* ts = nullptr) {
  if (!ps || !st || ps->nslice == 0 || use_tile(ts, ps, st)) return false;
  return g_persist_mode == 2 || (g_persist_mode == 1 && ps->ngroup >= kPersistMinLevels);
}
template <typename TS>
static const Sell<TS>& tile_vals(const TileSweep& ts) {
  if constexpr (std::is_same<TS, double>::value) return ts.v64;
  else if constexpr (std::is_same<TS, float>::value) return ts.v32;
  else return ts.v16;
}
template <typename TS>
static const Sell<TS>& split_vals(const TileSweep& ts) {
  if constexpr (std::is_same<TS, double>::value) return ts.s64;
  else if constexpr (std::is_same<TS, float>::value) return ts.s32;
  else return ts.s16;
}
// Sweep mode 4 takes the split-lane tiled kernels where the layout has them
// (2x2 and 3x3 blocks); modes 1 and 3 one lane per tile row, which measured
// faster (Oseen 256^2 FAC32: 51 / 61 us per lower / upper sweep against 59 /
// 69 us at the split form's best tile size).
static int g_tile_split = 0;
// One launch per phase, in the plan's order.
template <int B, typename TS, typename T, int NT, bool UPPER, bool DOT>
static int tile_sweep_nt(const TileSweep& ts, const T* rhs, T* out, const T* rvec, double* partials,
                         const KState* st) {
  const Sell<TS>& v = tile_vals<TS>(ts);
  const Sell<double>& lay = ts.v64;
  const size_t shmem = (size_t)ts.ovf_in * (B * B * sizeof(T) + sizeof(int)) + (size_t)ts.ovf_out * B * sizeof(T);
  for (size_t p = 0; p + 1 < ts.phase_tile.size(); p++) {
    const int t0 = ts.phase_tile[p], nt = ts.phase_tile[p + 1] - t0;
    if (nt == 0) continue;
    hipLaunchKernelGGL((k_trsv_tile<B, TS, T, NT, UPPER, DOT>), dim3(nt), dim3(NT), shmem, stream(), t0,
                       lay.rows.p, ts.rlev.p, ts.depth.p, lay.slice_ptr.p, ts.split.p, lay.col.p, v.val.p, v.dval.p,
                       ts.rov.p, ts.ovf_in, ts.ovf_out, rhs, out, rvec, partials, st);
    DFM_HIP(hipGetLastError());
  }
  return DFM_OK;
}
// Split-lane form: G lanes per tile row, G * nt threads per tile.
template <int B, typename TS, typename T, int NTS, bool UPPER, bool DOT>
static int tile_split_nt(const TileSweep& ts, const T* rhs, T* out, const T* rvec, double* partials,
                         const KState* st) {
  const Sell<TS>& v = split_vals<TS>(ts);
  const Sell<double>& lay = ts.s64;
  const size_t shmem = (size_t)ts.ovf_in * (B * B * sizeof(T) + sizeof(int)) + (size_t)ts.ovf_out * B * sizeof(T);
  for (size_t p = 0; p + 1 < ts.phase_tile.size(); p++) {
    const int t0 = ts.phase_tile[p], nt = ts.phase_tile[p + 1] - t0;
    if (nt == 0) continue;
    hipLaunchKernelGGL((k_trsv_tile_split<B, TS, T, NTS, UPPER, DOT>), dim3(nt), dim3(NTS), shmem, stream(), t0,
                       ts.v64.rows.p, ts.rlev.p, ts.depth.p, lay.slice_ptr.p, ts.ssplit.p, lay.col.p, v.val.p,
                       v.dval.p, ts.rov.p, ts.ovf_in, ts.ovf_out, rhs, out, rvec, partials, st);
    DFM_HIP(hipGetLastError());
  }
  return DFM_OK;
}
template <int B, typename TS, typename T, bool UPPER, bool DOT>
static int tile_sweep(const TileSweep& ts, const T* rhs, T* out, const T* rvec, double* partials,
                      const KState* st) {
  if constexpr (B == 2 || B == 3) {
    if (g_tile_split && ts.g) {
      constexpr int G = B == 2 ? 2 : 4;
      switch (ts.nt) {
        case 64: return tile_split_nt<B, TS, T, 64 * G, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
        case 128: return tile_split_nt<B, TS, T, 128 * G, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
        default:
          if constexpr (G == 4) {
            // 256 rows x 4 lanes would exceed a workgroup: one lane per row
            break;
          } else {
            return tile_split_nt<B, TS, T, 256 * G, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
          }
      }
    }
  }
  switch (ts.nt) {
    case 64: return tile_sweep_nt<B, TS, T, 64, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
    case 128: return tile_sweep_nt<B, TS, T, 128, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
    default: return tile_sweep_nt<B, TS, T, 256, UPPER, DOT>(ts, rhs, out, rvec, partials, st);
  }
}
static int persist_grid(const SweepSync& ps) {
  const int g = g_persist_grid > 0 ? g_persist_grid : std::max(64, 2 * ps.maxlev);
  return std::max(1, std::min(std::min(g, ps.nslice), kPersistMaxGrid));
}

template <int B, typename TS, typename T>
static int sweep_lower(const Sell<TS>& L, const Sell<double>& plan, const T* rhs, T* out, KState* st,
                       const SweepSync* ps = nullptr, const TileSweep* ts = nullptr) {
  if (use_tile(ts, ps, st)) return tile_sweep<B, TS, T, false, false>(*ts, rhs, out, nullptr, nullptr, st);
  if (use_persist(ps, st)) {
    DFM_LAUNCH_NT((k_trsv_persist<B, TS, T, false, false>), persist_grid(*ps), 64, ps->nslice, ps->ngroup,
                  ps->slev.p, ps->gsp.p, plan.rows.p, plan.slice_ptr.p, plan.col.p, L.val.p, L.dval.p, rhs, out,
                  nullptr, nullptr, st, ps->sync.p);
    return DFM_OK;
  }
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  for (int g = 0; g < ng; g++) {
    const int s0 = plan.group_slice_ptr[g], s1 = plan.group_slice_ptr[g + 1];
    DFM_LAUNCH_NT((k_trsv_lower<B, TS, T>), grid_for(s1 - s0, kTrsvWaves), kTrsvNT, s0, s1, plan.rows.p,
                  plan.slice_ptr.p, plan.col.p, L.val.p, L.dval.p, rhs, out, st);
  }
  return DFM_OK;
}
// Number of partial sums an upper sweep with DOT leaves for the finalize kernel.
static int upper_nparts(const Sell<double>& plan, const SweepSync* ps = nullptr, const KState* st = nullptr,
                        const TileSweep* ts = nullptr) {
  if (use_tile(ts, ps, st)) return ts->ntile;
  if (use_persist(ps, st)) return persist_grid(*ps);
  int n = 0;
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  for (int g = 0; g < ng; g++) n += grid_for(plan.group_slice_ptr[g + 1] - plan.group_slice_ptr[g], kTrsvWaves);
  return n;
}
// Backward sweep; with DOT the partial sums of <rvec, vec> over all rows are
// left in partials[0 .. upper_nparts) for a finalize kernel.
template <int B, typename TS, typename T, bool DOT>
static int sweep_upper(const Sell<TS>& U, const Sell<double>& plan, T* vec, const T* rvec, double* partials,
                       KState* st, const SweepSync* ps = nullptr, const TileSweep* ts = nullptr) {
  if (use_tile(ts, ps, st)) return tile_sweep<B, TS, T, true, DOT>(*ts, vec, vec, rvec, partials, st);
  if (use_persist(ps, st)) {
    DFM_LAUNCH_NT((k_trsv_persist<B, TS, T, true, DOT>), persist_grid(*ps), 64, ps->nslice, ps->ngroup, ps->slev.p,
                  ps->gsp.p, plan.rows.p, plan.slice_ptr.p, plan.col.p, U.val.p, U.val.p, vec, vec, rvec, partials,
                  st, ps->sync.p);
    return DFM_OK;
  }
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  const int nparts = DOT ? upper_nparts(plan) : 0;
  int part0 = 0;
  for (int g = 0; g < ng; g++) {
    const int s0 = plan.group_slice_ptr[g], s1 = plan.group_slice_ptr[g + 1];
    const int grid = grid_for(s1 - s0, kTrsvWaves);
    DFM_LAUNCH_NT((k_trsv_upper<B, TS, T, DOT>), grid, kTrsvNT, s0, s1, plan.rows.p, plan.slice_ptr.p, plan.col.p,
                  U.val.p, vec, rvec, partials, part0, nparts, st);
    part0 += grid;
  }
  return DFM_OK;
}

// Partial sums a fused lower sweep (k_trsv_lower_xr) leaves: one per workgroup.
static int lower_nparts(const Sell<double>& plan) { return upper_nparts(plan); }

int ilu_sweep64(dfm_ilu_s* M, double* vec, bool lower, bool upper) {
  // level sets, or the phase tiles where the factor has them (bitwise the same)
  const SweepSync *psL = &M->syncL, *psU = &M->syncU;
  const TileSweep *tsL = &M->tileL, *tsU = &M->tileU;
  switch (M->b) {
    case 1:
      if (lower) DFM_TRY((sweep_lower<1, double, double>(M->L64, M->L64, vec, vec, nullptr, psL, tsL)));
      if (upper)
        DFM_TRY((sweep_upper<1, double, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr, psU, tsU)));
      break;
    case 2:
      if (lower) DFM_TRY((sweep_lower<2, double, double>(M->L64, M->L64, vec, vec, nullptr, psL, tsL)));
      if (upper)
        DFM_TRY((sweep_upper<2, double, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr, psU, tsU)));
      break;
    case 3:
      if (lower) DFM_TRY((sweep_lower<3, double, double>(M->L64, M->L64, vec, vec, nullptr, psL, tsL)));
      if (upper)
        DFM_TRY((sweep_upper<3, double, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr, psU, tsU)));
      break;
    default:
      set_error("block size %u: the device substitutions support 1x1, 2x2 and 3x3 blocks", M->b);
      return DFM_NOT_IMPLEMENTED;
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Operator views: matrix storage TA, factor storage TF.
template <typename TA, typename TF>
struct Ops {
  const Sell<TA>* A;
  const Sell<TF>* L;
  const Sell<TF>* U;
  const Sell<double>* planA;
  const Sell<double>* planL;
  const Sell<double>* planU;
  int nb;
  const SweepSync* syncL;
  const SweepSync* syncU;
  const TileSweep* tileL;
  const TileSweep* tileU;
};
template <typename TS>
static const Sell<TS>& sel(const Sell<double>& s64, const Sell<float>& s32, const Sell<bf16s>& s16) {
  if constexpr (std::is_same<TS, double>::value) return s64;
  else if constexpr (std::is_same<TS, float>::value) return s32;
  else return s16;
}
template <typename TA, typename TF>
static Ops<TA, TF> ops_of(dfm_ilu_s* M) {
  static const Sell<bf16s> none{};
  return {&sel<TA>(M->A64, M->A32, none), &sel<TF>(M->L64, M->L32, M->L16), &sel<TF>(M->U64, M->U32, M->U16),
          &M->A64, &M->L64, &M->U64, (int)M->nb, &M->syncL, &M->syncU, &M->tileL, &M->tileU};
}

// In-solve timing of the Krylov iteration's SpMV (dfm_set_kernel_timing):
// while it is on, solves launch their iterations directly instead of
// replaying graphs, and each iteration SpMV is launched with a start / stop
// event pair that the runtime stamps when the kernel itself starts and ends.
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;  // start, stop per launch
  size_t used = 0;             // launches recorded since it was switched on
};
static KernelTimer g_ktimer;
constexpr size_t kTimerMaxLaunches = 1u << 16;

template <int B, typename TA, typename TF, typename T, bool DOTS, bool GATE>
static int spmv(const Ops<TA, TF>& o, const T* x, T* y, const T* res, const T* w0, const T* u2, const T* v2,
                double* partials, const KState* st) {
  const int nslice = o.planA->nslice;
  if constexpr (DOTS && GATE) {
    if (g_ktimer.on && g_ktimer.used < kTimerMaxLaunches && grid_for(nslice, kSpmvWaves) > 0) {
      while (g_ktimer.ev.size() < 2 * (g_ktimer.used + 1)) {
        hipEvent_t e = nullptr;
        DFM_HIP(hipEventCreate(&e));
        g_ktimer.ev.push_back(e);
      }
      hipEvent_t* e = &g_ktimer.ev[2 * g_ktimer.used];
      hipExtLaunchKernelGGL((k_spmv<B, TA, T, DOTS, GATE>), dim3(grid_for(nslice, kSpmvWaves)), dim3(kSpmvNT), 0,
                            stream(),
                            e[0], e[1], 0, nslice, (int)o.nb, (const int*)o.planA->slice_ptr.p,
                            (const int*)o.planA->col.p, (const TA*)o.A->val.p, (const TA*)o.A->dval.p, x, y, res, w0,
                            u2, v2, partials, st);
      DFM_HIP(hipGetLastError());
      g_ktimer.used++;
      return DFM_OK;
    }
  }
  DFM_LAUNCH_NT((k_spmv<B, TA, T, DOTS, GATE>), grid_for(nslice, kSpmvWaves), kSpmvNT, nslice, o.nb,
                o.planA->slice_ptr.p, o.planA->col.p, o.A->val.p, o.A->dval.p, x, y, res, w0, u2, v2, partials, st);
  return DFM_OK;
}
static int spmv_nparts(const Sell<double>& planA) { return grid_for(planA.nslice, kSpmvWaves); }

static int read_state(dfm_ilu_s* M) {
  DFM_HIP(hipMemcpyAsync(M->h_state, M->state.p, sizeof(KState), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
  DFM_CHECK(!M->h_state->sweep_fault, DFM_HIP_ERROR,
            "a persistent triangular sweep gave up waiting for a level set (results invalid)");
  return DFM_OK;
}
static int write_state(dfm_ilu_s* M) {
  DFM_HIP(hipMemcpyAsync(M->state.p, M->h_state, sizeof(KState), hipMemcpyHostToDevice, stream()));
  return DFM_OK;
}
template <typename T>
static int device_dot(dfm_ilu_s* M, long n, const T* x, const T* y, double* out) {
  const int g = grid_vec(n);
  DFM_LAUNCH(k_dot<T>, g, n, x, y, M->partials.p);
  DFM_LAUNCH_FIN(1, FIN_STORE, M->partials.p, g, M->state.p);
  DFM_TRY(read_state(M));
  *out = M->h_state->d0;
  return DFM_OK;
}

// Runs `nblock` iterations of `body` as one hipGraph (instantiated once per
// operand set and cached on the factor), or directly when the stream cannot
// be captured (the legacy default stream).
template <class Body>
static int run_block(dfm_ilu_s* M, const std::string& key, int nblock, Body&& body) {
  hipStream_t s = stream();
  // DFM_GRAPHS=0 launches the iterations directly instead of replaying a graph
  static const bool use_graphs = [] {
    const char* e = getenv("DFM_GRAPHS");
    return !(e && e[0] == '0');
  }();
  if (s == nullptr || !use_graphs || g_ktimer.on) {
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
    return DFM_OK;
  }
  for (auto& g : M->graphs)
    if (g.first == key) {
      DFM_HIP(hipGraphLaunch(g.second, s));
      return DFM_OK;
    }
  static bool capture_ok = true;
  if (!capture_ok) {
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
    return DFM_OK;
  }
  DFM_TRY(evict_graph(M->graphs));
  hipGraph_t graph = nullptr;
  DFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = DFM_OK;
  for (int i = 0; i < nblock && rc == DFM_OK; i++) rc = body();
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != DFM_OK || e != hipSuccess) {
    // a launch the runtime could not capture: nothing ran; launch the
    // iterations directly from now on
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    capture_ok = false;
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
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

int evict_graph(std::vector<std::pair<std::string, hipGraphExec_t>>& graphs) {
  if (graphs.size() < kMaxGraphs) return DFM_OK;
  DFM_HIP(hipStreamSynchronize(stream()));
  (void)hipGraphExecDestroy(graphs.front().second);
  graphs.erase(graphs.begin());
  return DFM_OK;
}

// The sweep mode as a graph key part, for graphs captured outside this file.
std::string sweep_mode_key() { return "p" + std::to_string(g_persist_mode) + "." + std::to_string(g_persist_grid); }

static std::string graph_key(const std::string& tag, int B, int nblock, std::initializer_list<const void*> ptrs) {
  // the sweep mode is part of what a graph captured
  std::string k = tag + ":" + std::to_string(B) + ":" + std::to_string(nblock) + ":p" +
                  std::to_string(g_persist_mode) + "." + std::to_string(g_persist_grid);
  char buf[32];
  for (const void* p : ptrs) {
    snprintf(buf, sizeof buf, ":%p", p);
    k += buf;
  }
  return k;
}
template <typename T>
static const char* tname() {
  if constexpr (std::is_same<T, double>::value) return "f64";
  else if constexpr (std::is_same<T, float>::value) return "f32";
  else return "bf16";
}

struct KResult {
  double conv;
  unsigned iter;
  bool converged;
};

static void init_state(KState& h, double rr0, double tol, unsigned maxit, int iter0) {
  std::memset(&h, 0, sizeof h);
  h.inv_rr0 = 1.0 / rr0;
  h.tol2 = tol * tol;
  h.iter = iter0;
  h.maxit = (int)std::min<unsigned>(maxit, 0x7fffffffu);
  if (h.iter >= h.maxit) h.done = 2;
}

// One PCG iteration (solver_ls_iter.cpp:134-165) on the device:
//   Ap (+ partial <p,Ap>) -> alpha -> forward sweep fused with the r, x update
//   (+ partial <r,r>) -> backward sweep (+ partial <r,z>) -> convergence test
//   and beta -> p = beta p + z.
// The convergence test reads <r,r> after the sweeps rather than before them;
// it sees the same value, and the sweeps of the final iteration are the only
// extra work.  The first backward level (rows without U entries, which it
// leaves unchanged) is not launched: the forward sweep adds their <r,z>.
template <int B, typename TA, typename TF, typename T>
static int pcg_body(dfm_ilu_s* M, const Ops<TA, TF>& o, T* r, T* x, T* p, T* z) {
  const long n = (long)o.nb * B;
  KState* st = M->state.p;
  double* part = M->partials.p;
  const int gv = grid_vec(n);
  DFM_TRY((spmv<B, TA, TF, T, true, true>(o, p, z, nullptr, p, nullptr, nullptr, part, st)));
  DFM_LAUNCH_FIN(1, FIN_PCG_PAP, part, spmv_nparts(*o.planA), st);
  const Sell<double>& pl = *o.planL;
  const Sell<double>& pu = *o.planU;
  const int ngL = (int)pl.group_slice_ptr.size() - 1, ngU = (int)pu.group_slice_ptr.size() - 1;
  const bool skip_u0 = M->upper0_empty;
  const int nrr = lower_nparts(pl);
  double* prr = part;
  double* prz = part + nrr;
  int part0 = 0;
  for (int g = 0; g < ngL; g++) {
    const int s0 = pl.group_slice_ptr[g], s1 = pl.group_slice_ptr[g + 1];
    const int grid = grid_for(s1 - s0, kTrsvWaves);
    if (skip_u0) {
      DFM_LAUNCH_NT((k_trsv_lower_xr<B, TF, T, true>), grid, kTrsvNT, s0, s1, pl.rows.p, pl.slice_ptr.p, pl.col.p,
                    o.L->val.p, o.L->dval.p, p, r, x, z, M->d_rzmaskL.p, prr, part0, prz, st);
    } else {
      DFM_LAUNCH_NT((k_trsv_lower_xr<B, TF, T, false>), grid, kTrsvNT, s0, s1, pl.rows.p, pl.slice_ptr.p,
                    pl.col.p, o.L->val.p, o.L->dval.p, p, r, x, z, nullptr, prr, part0, nullptr, st);
    }
    part0 += grid;
  }
  int nrz = skip_u0 ? nrr : 0;
  for (int g = skip_u0 ? 1 : 0; g < ngU; g++) {
    const int s0 = pu.group_slice_ptr[g], s1 = pu.group_slice_ptr[g + 1];
    const int grid = grid_for(s1 - s0, kTrsvWaves);
    DFM_LAUNCH_NT((k_trsv_upper<B, TF, T, true>), grid, kTrsvNT, s0, s1, pu.rows.p, pu.slice_ptr.p, pu.col.p,
                  o.U->val.p, z, r, prz, nrz, 0, st);
    nrz += grid;
  }
  hipLaunchKernelGGL(k_finalize_rrz, dim3(1), dim3(kFinNT), 0, stream(), prr, nrr, prz, nrz, st);
  DFM_HIP(hipGetLastError());
  DFM_LAUNCH(k_pcg_update_p<T>, gv, n, z, p, st);
  return DFM_OK;
}

// PCG start: z = M^-1 r, rz = <r,z>, p = z (:128-132)
template <int B, typename TA, typename TF, typename T>
static int pcg_start(dfm_ilu_s* M, const Ops<TA, TF>& o, T* r, T* p, T* z) {
  const long n = (long)o.nb * B;
  KState* st = M->state.p;
  DFM_TRY((sweep_lower<B, TF, T>(*o.L, *o.planL, r, z, st, o.syncL, o.tileL)));
  DFM_TRY((sweep_upper<B, TF, T, true>(*o.U, *o.planU, z, r, M->partials.p, st, o.syncU, o.tileU)));
  DFM_LAUNCH_FIN(1, FIN_PCG_RZ0, M->partials.p, upper_nparts(*o.planU, o.syncU, st, o.tileU), st);
  DFM_HIP(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));
  return DFM_OK;
}

// One BiCGSTAB iteration (solver_ls_iter.cpp:327-404).  w: r2, p, Mp, AMp, s, Ms, AMs.
template <int B, typename TA, typename TF, typename T>
static int bicg_body(dfm_ilu_s* M, const Ops<TA, TF>& o, T* r, T* x, T* w) {
  const long n = (long)o.nb * B;
  T *r2 = w, *p = w + n, *Mp = w + 2 * n, *AMp = w + 3 * n, *s = w + 4 * n, *Ms = w + 5 * n, *AMs = w + 6 * n;
  KState* st = M->state.p;
  double* part = M->partials.p;
  const int gv = grid_vec(n);
  DFM_TRY((sweep_lower<B, TF, T>(*o.L, *o.planL, p, Mp, st, o.syncL, o.tileL)));
  DFM_TRY((sweep_upper<B, TF, T, false>(*o.U, *o.planU, Mp, nullptr, nullptr, st, o.syncU, o.tileU)));
  DFM_TRY((spmv<B, TA, TF, T, true, true>(o, Mp, AMp, nullptr, r2, r, r2, part, st)));
  DFM_LAUNCH_FIN(3, FIN_BICG_ALPHA, part, spmv_nparts(*o.planA), st);
  DFM_LAUNCH(k_bicg_s<T>, gv, n, r, AMp, s, st);
  DFM_TRY((sweep_lower<B, TF, T>(*o.L, *o.planL, s, Ms, st, o.syncL, o.tileL)));
  DFM_TRY((sweep_upper<B, TF, T, false>(*o.U, *o.planU, Ms, nullptr, nullptr, st, o.syncU, o.tileU)));
  DFM_TRY((spmv<B, TA, TF, T, true, true>(o, Ms, AMs, nullptr, s, nullptr, nullptr, part, st)));
  DFM_LAUNCH_FIN(2, FIN_BICG_OMEGA, part, spmv_nparts(*o.planA), st);
  DFM_LAUNCH(k_bicg_xr<T>, gv, n, Mp, Ms, s, AMs, r2, x, r, part, st);
  DFM_LAUNCH_FIN(2, FIN_BICG_XR, part, gv, st);
  DFM_LAUNCH(k_bicg_p<T>, gv, n, r, AMp, p, st);
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// FP64 solvers with the reference's semantics.
// LsSol::Solve_PCG (solver_ls_iter.cpp:100-172).  r: in rhs, out final residual; x: out.
template <int B>
static int pcg64(dfm_ilu_s* M, double* r, double* x, double* p, double* z, double tol, unsigned maxit,
                 unsigned check_every, KResult& res) {
  const long n = (long)M->nb * B;
  Ops<double, double> o = ops_of<double, double>(M);
  double rr0 = 0.0;
  DFM_TRY(device_dot<double>(M, n, r, r, &rr0));
  DFM_LAUNCH(k_fill<double>, grid_vec(n), n, 0.0, x);
  if (rr0 < 1.0e-30) {
    res = {0.0, 0u, true};
    return DFM_OK;
  }
  init_state(*M->h_state, rr0, tol, maxit, 0);
  DFM_TRY(write_state(M));
  DFM_TRY((pcg_start<B, double, double, double>(M, o, r, p, z)));
  const std::string key = graph_key("pcg64", B, (int)check_every, {r, x, p, z, M->partials.p, M->state.p});
  if (!M->h_state->done) {
    for (;;) {
      DFM_TRY(run_block(M, key, (int)check_every,
                        [&]() { return pcg_body<B, double, double, double>(M, o, r, x, p, z); }));
      DFM_TRY(read_state(M));
      if (M->h_state->done) break;
    }
  }
  const KState& s = *M->h_state;
  if (s.done == 1) res = {std::sqrt(s.conv_rr * s.inv_rr0), (unsigned)s.conv_iter, true};
  else res = {std::sqrt(s.rr * s.inv_rr0), maxit, false};
  return DFM_OK;
}

// LsSol::Solve_PBiCGSTAB (solver_ls_iter.cpp:291-408).
template <int B>
static int bicg64(dfm_ilu_s* M, double* r, double* x, double* w, double tol, unsigned maxit, unsigned check_every,
                  KResult& res, bool& x_zeroed) {
  const long n = (long)M->nb * B;
  Ops<double, double> o = ops_of<double, double>(M);
  double rr0 = 0.0;
  DFM_TRY(device_dot<double>(M, n, r, r, &rr0));
  x_zeroed = false;
  if (rr0 < 1.0e-60) {
    res = {std::sqrt(rr0), 0u, true};
    return DFM_OK;
  }
  DFM_LAUNCH(k_fill<double>, grid_vec(n), n, 0.0, x);
  x_zeroed = true;
  DFM_HIP(hipMemcpyAsync(w, r, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));      // r2 = r
  DFM_HIP(hipMemcpyAsync(w + n, r, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));  // p = r
  init_state(*M->h_state, rr0, tol, maxit, 1);
  DFM_TRY(write_state(M));
  const std::string key = graph_key("bicg64", B, (int)check_every, {r, x, w, M->partials.p, M->state.p});
  if (!M->h_state->done) {
    for (;;) {
      DFM_TRY(run_block(M, key, (int)check_every,
                        [&]() { return bicg_body<B, double, double, double>(M, o, r, x, w); }));
      DFM_TRY(read_state(M));
      if (M->h_state->done) break;
    }
  }
  const KState& sv = *M->h_state;
  if (sv.done == 1) res = {std::sqrt(sv.conv_rr * sv.inv_rr0), (unsigned)sv.conv_iter, true};
  else res = {tol, maxit, false};  // the reference leaves conv_ratio and reports max_iter
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision with reliable updates.  The Krylov recurrence runs on
// vectors of type T against the matrix copy TA and factor copy TF, with FP64
// reduction accumulators.  Whenever the recursive residual has dropped by
// `delta` since the last update (or the recurrence reports convergence) the
// partial solution is folded into the FP64 solution and the true residual
// b - A x, computed with the FP64 matrix, replaces the recursive residual.
// The search direction is kept, so no Krylov information is thrown away as
// in restarted refinement.  Convergence is decided on the FP64 true residual.
template <int B>
static int true_residual64(dfm_ilu_s* M, const double* b, const double* x, double* r, double* rr) {
  Ops<double, double> o = ops_of<double, double>(M);
  DFM_TRY((spmv<B, double, double, double, true, false>(o, x, r, b, nullptr, nullptr, nullptr, M->partials.p,
                                                          M->state.p)));
  DFM_LAUNCH_FIN(3, FIN_STORE, M->partials.p, spmv_nparts(M->A64), M->state.p);
  DFM_TRY(read_state(M));
  *rr = M->h_state->d1;
  return DFM_OK;
}

template <int B, typename TA, typename TF, typename T>
static int mixed_core(dfm_ilu_s* M, bool bicg, const double* b, double* r64, double* x64, T* w, double tol,
                      unsigned maxit, unsigned check_every, double delta, KResult& res, unsigned& nupdates) {
  const long n = (long)M->nb * B;
  const double delta2 = delta * delta;
  Ops<TA, TF> o = ops_of<TA, TF>(M);
  T* rr_ = w;          // recursive residual
  T* xr = w + n;       // partial solution since the last reliable update
  T* ws = w + 2 * n;   // PCG: p, z ; BiCGSTAB: r2, p, Mp, AMp, s, Ms, AMs
  nupdates = 0;
  double rr0 = 0.0;
  DFM_TRY(device_dot<double>(M, n, b, b, &rr0));
  DFM_LAUNCH(k_fill<double>, grid_vec(n), n, 0.0, x64);
  DFM_HIP(hipMemcpyAsync(r64, b, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  if (rr0 < (bicg ? 1.0e-60 : 1.0e-30)) {
    res = {bicg ? std::sqrt(rr0) : 0.0, 0u, true};
    return DFM_OK;
  }
  DFM_LAUNCH(k_convert<T>, grid_vec(n), n, r64, rr_);
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, xr);
  init_state(*M->h_state, rr0, tol, maxit, bicg ? 1 : 0);
  DFM_TRY(write_state(M));
  if (!bicg) {
    DFM_TRY((pcg_start<B, TA, TF, T>(M, o, rr_, ws, ws + n)));
  } else {
    DFM_HIP(hipMemcpyAsync(ws, rr_, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));      // r2
    DFM_HIP(hipMemcpyAsync(ws + n, rr_, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));  // p
  }
  const std::string tag = std::string(bicg ? "bicgR" : "pcgR") + tname<T>() + tname<TA>() + tname<TF>();
  const std::string key = graph_key(tag, B, (int)check_every, {rr_, xr, ws, M->partials.p, M->state.p});
  double rr_ref = rr0, rr_true = rr0;
  double rr_low = rr0;  // smallest recursive <r,r> seen since the last update
  bool converged = false;
  for (;;) {
    if (!M->h_state->done) {
      DFM_TRY(run_block(M, key, (int)check_every, [&]() {
        return bicg ? bicg_body<B, TA, TF, T>(M, o, rr_, xr, ws) : pcg_body<B, TA, TF, T>(M, o, rr_, xr, ws, ws + n);
      }));
      DFM_TRY(read_state(M));
    }
    KState& h = *M->h_state;
    // A recurrence on FP32 vectors that has lost its way (BiCGSTAB can: the
    // residual grows by orders of magnitude or overflows) is stopped by an
    // update and restarted from the FP64 true residual.  With FP64 vectors
    // only an overflow counts: BiCGSTAB's residual swings by such factors on
    // its way down, and a restart throws away its Krylov space.
    const bool diverged = !std::isfinite(h.rr) || (std::is_same<T, float>::value && h.rr > 1.0e4 * rr_low);
    if (std::isfinite(h.rr)) rr_low = std::min(rr_low, h.rr);
    if (!(h.done || h.rr < delta2 * rr_ref || diverged)) continue;
    if (std::isfinite(h.rr)) {
      // reliable update
      DFM_LAUNCH(k_reliable_update<T>, grid_vec(n), n, xr, x64);
      DFM_TRY(true_residual64<B>(M, b, x64, r64, &rr_true));
      nupdates++;
    } else {
      // the partial solution overflowed: drop it and restart from the last
      // true residual (x64 and r64 still hold it)
      DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, xr);
    }
    KState& hs = *M->h_state;
    if (rr_true * hs.inv_rr0 < hs.tol2) {
      converged = true;
      break;
    }
    if (hs.done == 2 || !std::isfinite(rr_true)) break;
    DFM_LAUNCH(k_convert<T>, grid_vec(n), n, r64, rr_);
    // Residual replacement keeps the search directions while the recursive
    // residual still tracks the true one.  When the recursion has run ahead
    // by more than 10x in norm (it converged on the low-precision operator
    // within one check block) its directions no longer fit the replaced
    // residual: restart the recurrence from the true residual.
    const bool restart = diverged || rr_true > 100.0 * hs.rr;
    rr_ref = rr_true;
    rr_low = rr_true;
    hs.done = 0;
    hs.rr = rr_true;
    DFM_TRY(write_state(M));
    if (restart) {
      if (!bicg) {
        DFM_TRY((pcg_start<B, TA, TF, T>(M, o, rr_, ws, ws + n)));
      } else {
        DFM_HIP(hipMemcpyAsync(ws, rr_, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));      // r2
        DFM_HIP(hipMemcpyAsync(ws + n, rr_, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));  // p
      }
    }
  }
  const KState& s = *M->h_state;
  res = {std::sqrt(rr_true * s.inv_rr0), (unsigned)(bicg ? s.iter - 1 : s.iter), converged};
  return DFM_OK;
}

// ---------------------------------------------------------------------------
static int ensure_workspace(dfm_ilu_s* M, size_t bytes) {
  if (M->work.n < bytes) DFM_TRY(M->work.alloc(bytes));
  size_t need = 0;
  need = std::max<size_t>(need, (size_t)spmv_nparts(M->A64));
  need = std::max<size_t>(need, (size_t)upper_nparts(M->U64));
  need = std::max<size_t>(need, (size_t)lower_nparts(M->L64));
  need = std::max<size_t>(need, (size_t)kPersistMaxGrid);
  need = std::max<size_t>(need, (size_t)M->tileU.ntile);
  need = 3 * need + 64;
  if (M->partials.n < need) DFM_TRY(M->partials.alloc(need));
  return DFM_OK;
}

template <int B>
static int solve_dispatch(dfm_bcrs_s* A, dfm_ilu_s* M, bool bicg, double* r_user, double* x_user, double* conv_ratio,
                          unsigned* iteration, dfm_solve_opts* opts) {
  const int nb = (int)M->nb;
  const long n = (long)nb * B;
  const int prec = opts ? opts->precision : DFM_PREC_FP64;
  const bool mixed = prec != DFM_PREC_FP64;
  const bool vec32 = prec == DFM_PREC_MIXED_F32 || prec == DFM_PREC_MIXED_F32_BF16;
  const bool fac16 = prec == DFM_PREC_MIXED_BF16 || prec == DFM_PREC_MIXED_F32_BF16 || prec == DFM_PREC_FAC16;
  const unsigned check_every = (opts && opts->check_every) ? opts->check_every : 16;
  const bool want_true = opts && (opts->want_true_relres || mixed);
  const double tol = *conv_ratio;
  const unsigned maxit = *iteration;
  DFM_TRY(ilu_refresh_sell(M, A, mixed, fac16));
  const int* ord = M->permuted ? M->d_order.p : nullptr;
  // workspace: f64 r, x, b, then the solver's vectors (+2 for the mixed recurrence)
  const size_t nsolver = bicg ? 7 : 2;
  const size_t bytes = mixed ? 3 * n * sizeof(double) + (nsolver + 2) * n * (vec32 ? 4 : 8)
                             : (3 + nsolver) * n * sizeof(double);
  DFM_TRY(ensure_workspace(M, bytes));
  double* r = reinterpret_cast<double*>(M->work.p);
  double* x = r + n;
  double* b = r + 2 * n;
  char* wsolver = reinterpret_cast<char*>(r + 3 * n);
  DFM_LAUNCH((k_permute<double, double>), grid_for(n), nb, B, ord, r_user, mixed ? b : r, 1.0);
  if (want_true && !mixed) DFM_HIP(hipMemcpyAsync(b, r, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  KResult kr{0.0, 0u, false};
  int rc = DFM_OK;
  bool x_written = true;
  if (!mixed) {
    double* w = reinterpret_cast<double*>(wsolver);
    if (!bicg) {
      DFM_TRY((pcg64<B>(M, r, x, w, w + n, tol, maxit, check_every, kr)));
    } else {
      DFM_TRY((bicg64<B>(M, r, x, w, tol, maxit, check_every, kr, x_written)));
    }
    *conv_ratio = kr.conv;
    *iteration = kr.iter;
    rc = kr.converged ? DFM_OK : DFM_NOT_CONVERGED;
    if (want_true && x_written) {
      double rr = 0.0, bb2 = 0.0;
      DFM_TRY(device_dot<double>(M, n, b, b, &bb2));
      DFM_TRY(true_residual64<B>(M, b, x, w, &rr));  // solver vectors are free now
      opts->true_relres = bb2 > 0.0 ? std::sqrt(rr / bb2) : 0.0;
    }
  } else {
    unsigned nup = 0;
    const double delta = opts->reliable_delta > 0.0 ? opts->reliable_delta : 0.1;
    switch (prec) {
      case DFM_PREC_MIXED:
        DFM_TRY((mixed_core<B, float, float, double>(M, bicg, b, r, x, reinterpret_cast<double*>(wsolver), tol, maxit,
                                                     check_every, delta, kr, nup)));
        break;
      case DFM_PREC_MIXED_BF16:
        DFM_TRY((mixed_core<B, float, bf16s, double>(M, bicg, b, r, x, reinterpret_cast<double*>(wsolver), tol, maxit,
                                                     check_every, delta, kr, nup)));
        break;
      case DFM_PREC_MIXED_F32:
        DFM_TRY((mixed_core<B, float, float, float>(M, bicg, b, r, x, reinterpret_cast<float*>(wsolver), tol, maxit,
                                                    check_every, delta, kr, nup)));
        break;
      case DFM_PREC_FAC32:
        DFM_TRY((mixed_core<B, double, float, double>(M, bicg, b, r, x, reinterpret_cast<double*>(wsolver), tol,
                                                      maxit, check_every, delta, kr, nup)));
        break;
      case DFM_PREC_FAC16:
        DFM_TRY((mixed_core<B, double, bf16s, double>(M, bicg, b, r, x, reinterpret_cast<double*>(wsolver), tol,
                                                      maxit, check_every, delta, kr, nup)));
        break;
      default:
        DFM_TRY((mixed_core<B, float, bf16s, float>(M, bicg, b, r, x, reinterpret_cast<float*>(wsolver), tol, maxit,
                                                    check_every, delta, kr, nup)));
    }
    opts->inner_iters = kr.iter;
    opts->outer_iters = nup;
    opts->true_relres = kr.conv;
    *conv_ratio = kr.conv;
    *iteration = kr.iter;
    rc = kr.converged ? DFM_OK : DFM_NOT_CONVERGED;
  }
  if (x_written) DFM_LAUNCH(k_unpermute<double>, grid_for(n), nb, B, ord, x, x_user);
  DFM_LAUNCH(k_unpermute<double>, grid_for(n), nb, B, ord, r, r_user);
  DFM_HIP(hipStreamSynchronize(stream()));
  return rc;
}

static int solve_entry(dfm_bcrs_s* A, dfm_ilu_s* M, bool bicg, double* r, double* x, double* conv_ratio,
                       unsigned* iteration, dfm_solve_opts* opts) {
  DFM_CHECK(A && M && r && x && conv_ratio && iteration, DFM_BAD_ARG, "bad arguments to the solver");
  DFM_CHECK(A->nb == M->nb && A->b == M->b, DFM_BAD_ARG, "factor and matrix do not match");
  DFM_CHECK(A->pattern_hash == M->pattern_hash, DFM_BAD_ARG,
            "matrix pattern differs from the one the factor was built on");
  DFM_CHECK(M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  DFM_CHECK(!opts || (opts->precision >= DFM_PREC_FP64 && opts->precision <= DFM_PREC_FAC16), DFM_BAD_ARG,
            "unknown precision %d", opts ? opts->precision : 0);
  switch (A->b) {
    case 1: return solve_dispatch<1>(A, M, bicg, r, x, conv_ratio, iteration, opts);
    case 2: return solve_dispatch<2>(A, M, bicg, r, x, conv_ratio, iteration, opts);
    case 3: return solve_dispatch<3>(A, M, bicg, r, x, conv_ratio, iteration, opts);
    default:
      set_error("block size %u: the device solvers support 1x1, 2x2 and 3x3 blocks", A->b);
      return DFM_NOT_IMPLEMENTED;
  }
}

// ---------------------------------------------------------------------------
// Kernel timing on the real operands (dfm_bench_kernel).
template <int B, typename TA, typename TF, typename T>
static int bench_kernel_t(dfm_bcrs_s* A, dfm_ilu_s* M, int which, int nrep, double* avg_ms, double* bytes) {
  const int nb = (int)M->nb;
  const long n = (long)nb * B;
  Ops<TA, TF> o = ops_of<TA, TF>(M);
  DevBuf<T> xv, yv;
  DFM_TRY(xv.alloc(n));
  DFM_TRY(yv.alloc(n));
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)1, xv.p);
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, yv.p);
  const double sv = sizeof(T), sa = sizeof(TA), sf = sizeof(TF), bb = (double)B * B;
  const double nnzA = (double)A->col.size();
  double nnzL = 0, nnzU = 0;
  for (unsigned i = 0; i < M->nb; i++) {
    nnzL += M->fdiaind[i] - M->frow[i];
    nnzU += M->frow[i + 1] - M->fdiaind[i];
  }
  // the instantiations the PCG iteration launches: SpMV with the fused <p,Ap>
  // partials, lower sweep, upper sweep with the fused <r,z> partials
  DFM_TRY(ensure_workspace(M, 0));
  init_state(*M->h_state, 1.0, 0.0, 1u << 30, 0);
  DFM_TRY(write_state(M));
  KState* st = M->state.p;
  double* part = M->partials.p;
  auto run = [&]() -> int {
    if (which == 0) return spmv<B, TA, TF, T, true, true>(o, xv.p, yv.p, nullptr, xv.p, nullptr, nullptr, part, st);
    if (which == 1) return sweep_lower<B, TF, T>(*o.L, *o.planL, xv.p, yv.p, st, o.syncL, o.tileL);
    return sweep_upper<B, TF, T, true>(*o.U, *o.planU, yv.p, xv.p, part, st, o.syncU, o.tileU);
  };
  // algorithmic bytes: values, column indices, row ids (sweeps), each vector
  // element read or written once (x of SpMV and the dot operand are the same
  // vector)
  if (which == 0) *bytes = (nnzA + nb) * bb * sa + 4.0 * nnzA + 2.0 * n * sv;
  else if (which == 1) *bytes = (nnzL + nb) * bb * sf + 4.0 * nnzL + 4.0 * nb + 3.0 * n * sv;
  else *bytes = nnzU * bb * sf + 4.0 * nnzU + 4.0 * nb + 3.0 * n * sv;
  DFM_TRY(run());  // warm-up
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < nrep; r++) DFM_TRY(run());
  DFM_HIP(hipEventRecord(e1, stream()));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / nrep;
  return read_state(M);  // reports a persistent sweep's fault
}

template <int B>
static int bench_dispatch(dfm_bcrs_s* A, dfm_ilu_s* M, int which, int prec, int nrep, double* ms, double* bytes) {
  switch (prec) {
    case DFM_PREC_FP64: return bench_kernel_t<B, double, double, double>(A, M, which, nrep, ms, bytes);
    case DFM_PREC_MIXED: return bench_kernel_t<B, float, float, double>(A, M, which, nrep, ms, bytes);
    case DFM_PREC_MIXED_BF16: return bench_kernel_t<B, float, bf16s, double>(A, M, which, nrep, ms, bytes);
    case DFM_PREC_MIXED_F32: return bench_kernel_t<B, float, float, float>(A, M, which, nrep, ms, bytes);
    case DFM_PREC_FAC32: return bench_kernel_t<B, double, float, double>(A, M, which, nrep, ms, bytes);
    case DFM_PREC_FAC16: return bench_kernel_t<B, double, bf16s, double>(A, M, which, nrep, ms, bytes);
    default: return bench_kernel_t<B, float, bf16s, float>(A, M, which, nrep, ms, bytes);
  }
}

}  // namespace dfm

extern "C" {
int dfm_bench_kernel(dfm_bcrs* A, dfm_ilu* M, int which, int precision, int nrep, double* avg_ms, double* bytes) {
  using namespace dfm;
  DFM_CHECK(A && M && avg_ms && bytes && nrep > 0 && which >= 0 && which <= 2, DFM_BAD_ARG, "bad arguments");
  DFM_CHECK(precision >= DFM_PREC_FP64 && precision <= DFM_PREC_FAC16, DFM_BAD_ARG, "bad precision");
  DFM_CHECK(M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  const bool fac16 = precision == DFM_PREC_MIXED_BF16 || precision == DFM_PREC_MIXED_F32_BF16 ||
                     precision == DFM_PREC_FAC16;
  DFM_TRY(ilu_refresh_sell(M, A, precision != DFM_PREC_FP64, fac16));
  switch (A->b) {
    case 1: return bench_dispatch<1>(A, M, which, precision, nrep, avg_ms, bytes);
    case 2: return bench_dispatch<2>(A, M, which, precision, nrep, avg_ms, bytes);
    case 3: return bench_dispatch<3>(A, M, which, precision, nrep, avg_ms, bytes);
    default: return DFM_NOT_IMPLEMENTED;
  }
}

int dfm_bench_copy(size_t bytes, int nrep, double* avg_ms) {
  using namespace dfm;
  DFM_CHECK(avg_ms && nrep > 0 && bytes >= 16, DFM_BAD_ARG, "bad arguments");
  DFM_TRY(ensure_init());
  const long n4 = (long)(bytes / 16);
  DevBuf<float4> a, b;
  DFM_TRY(a.alloc(n4));
  DFM_TRY(b.alloc(n4));
  DFM_HIP(hipMemsetAsync(a.p, 0, n4 * 16, stream()));
  const int grid = (int)std::min<long>((n4 + kNT - 1) / kNT, 1 << 16);
  DFM_LAUNCH(k_copy4, grid, n4, a.p, b.p);
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < nrep; r++) DFM_LAUNCH(k_copy4, grid, n4, a.p, b.p);
  DFM_HIP(hipEventRecord(e1, stream()));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / nrep;
  return DFM_OK;
}

int dfm_bench_read(size_t bytes, int nrep, double* avg_ms) {
  using namespace dfm;
  DFM_CHECK(avg_ms && nrep > 0 && bytes >= 16, DFM_BAD_ARG, "bad arguments");
  DFM_TRY(ensure_init());
  const long n4 = (long)(bytes / 16);
  DevBuf<float4> a;
  DevBuf<float> out;
  DFM_TRY(a.alloc(n4));
  DFM_TRY(out.alloc(1));
  DFM_HIP(hipMemsetAsync(a.p, 0, n4 * 16, stream()));
  const int grid = (int)std::min<long>((n4 + kNT - 1) / kNT, 1 << 16);
  DFM_LAUNCH(k_read4, grid, n4, a.p, out.p);
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < nrep; r++) DFM_LAUNCH(k_read4, grid, n4, a.p, out.p);
  DFM_HIP(hipEventRecord(e1, stream()));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / nrep;
  return DFM_OK;
}

int dfm_set_kernel_timing(int on) {
  DFM_HIP(hipStreamSynchronize(dfm::stream()));
  dfm::g_ktimer.on = on != 0;
  if (on) dfm::g_ktimer.used = 0;
  return DFM_OK;
}

int dfm_get_kernel_timing(double* avg_all_ms, double* avg_working_ms, int* nlaunch, int* nworking) {
  DFM_CHECK(avg_all_ms && avg_working_ms && nlaunch && nworking, DFM_BAD_ARG, "NULL output");
  DFM_HIP(hipStreamSynchronize(dfm::stream()));
  const size_t n = dfm::g_ktimer.used;
  std::vector<double> ms(n);
  for (size_t i = 0; i < n; i++) {
    float t = 0.f;
    DFM_HIP(hipEventElapsedTime(&t, dfm::g_ktimer.ev[2 * i], dfm::g_ktimer.ev[2 * i + 1]));
    ms[i] = t;
  }
  double all = 0.0, work = 0.0;
  int nw = 0;
  if (n) {
    std::vector<double> sorted(ms);
    std::sort(sorted.begin(), sorted.end());
    const double half_median = 0.5 * sorted[n / 2];
    for (double v : ms) {
      all += v;
      if (v >= half_median) {
        work += v;
        nw++;
      }
    }
  }
  *avg_all_ms = n ? all / (double)n : 0.0;
  *avg_working_ms = nw ? work / nw : 0.0;
  *nlaunch = (int)n;
  *nworking = nw;
  return DFM_OK;
}

int dfm_set_sweep_mode(int mode, int grid) {
  DFM_CHECK(mode >= 0 && mode <= 4 && grid >= 0 && grid <= dfm::kPersistMaxGrid, DFM_BAD_ARG,
            "bad sweep mode %d / grid %d", mode, grid);
  dfm::g_persist_mode = mode;
  dfm::g_tile_split = mode == 4;
  dfm::g_persist_grid = grid;
  return DFM_OK;
}

int dfm_solve_pcg(dfm_bcrs* A, dfm_ilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration,
                  dfm_solve_opts* opts) {
  return dfm::solve_entry(A, M, false, r, x, conv_ratio, iteration, opts);
}
int dfm_solve_pbicgstab(dfm_bcrs* A, dfm_ilu* M, double* r, double* x, double* conv_ratio, unsigned* iteration,
                        dfm_solve_opts* opts) {
  return dfm::solve_entry(A, M, true, r, x, conv_ratio, iteration, opts);
}
}
