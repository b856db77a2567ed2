// dfm_solver.hip -- device-resident Krylov solvers: LsSol::Solve_PCG and
// LsSol::Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:100-172, 291-408) with
// the ILU preconditioner, plus the mixed-precision variant (FP32 matrix,
// factor and inner Krylov; FP64 reductions, true residual and refinement).
//
// All vectors live in factor numbering; the right-hand side is permuted in
// once and the solution and residual are permuted out once.  Scalars
// (alpha, beta, omega, residual norms) never leave the device inside the
// iteration: each kernel reads them from a KState and the host only polls the
// convergence flag every check_every iterations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

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

template <typename T>
static int sell_alloc_like(const Sell<double>& src, Sell<T>& dst) {
  dst.nslice = src.nslice;
  dst.nk = src.nk;
  dst.bb = src.bb;
  dst.group_slice_ptr = src.group_slice_ptr;
  DFM_TRY(dst.val.alloc(std::max<size_t>(1, (size_t)src.nk * src.bb * kSlice)));
  DFM_TRY(dst.dval.alloc(std::max<size_t>(1, (size_t)src.nslice * src.bb * kSlice)));
  return DFM_OK;
}

// Refresh the sliced-ELL value copies: factor (always) and A (when its
// values changed since the last refresh); FP32 copies when requested.
int ilu_refresh_sell(dfm_ilu_s* M, const dfm_bcrs_s* A, bool want32) {
  const int bb = (int)(M->b * M->b);
  auto gather = [&](auto& S, const DevBuf<long long>& src, const double* from) -> int {
    using TT = std::remove_pointer_t<decltype(S.val.p)>;
    const long nslot = S.nk * kSlice;
    DFM_LAUNCH(k_sell_gather<TT>, grid_for(nslot), nslot, bb, src.p, from, S.val.p);
    return DFM_OK;
  };
  auto gather_dia = [&](auto& S, const int* rows, const int* order, const double* from) -> int {
    using TT = std::remove_pointer_t<decltype(S.val.p)>;
    const long nslot = (long)S.nslice * kSlice;
    DFM_LAUNCH(k_sell_gather_dia<TT>, grid_for(nslot), nslot, bb, rows, order, from, S.dval.p);
    return DFM_OK;
  };
  DFM_TRY(gather(M->L64, M->d_srcL, M->d_fval.p));
  DFM_TRY(gather_dia(M->L64, M->L64.rows.p, nullptr, M->d_fdia.p));
  DFM_TRY(gather(M->U64, M->d_srcU, M->d_fval.p));
  if (want32) {
    if (!M->have32) {
      DFM_TRY(sell_alloc_like(M->L64, M->L32));
      DFM_TRY(sell_alloc_like(M->U64, M->U32));
      DFM_TRY(sell_alloc_like(M->A64, M->A32));
      M->have32 = true;
      M->a32_valid = false;
    }
    DFM_TRY(gather(M->L32, M->d_srcL, M->d_fval.p));
    DFM_TRY(gather_dia(M->L32, M->L64.rows.p, nullptr, M->d_fdia.p));
    DFM_TRY(gather(M->U32, M->d_srcU, M->d_fval.p));
  }
  if (A && (M->a_version != A->version || (want32 && !M->a32_valid))) {
    const int* ord = M->permuted ? M->d_order.p : nullptr;
    DFM_TRY(gather(M->A64, M->d_srcA, A->d_val.p));
    DFM_TRY(gather_dia(M->A64, M->A64.rows.p, ord, A->d_dia.p));
    if (want32) {
      DFM_TRY(gather(M->A32, M->d_srcA, A->d_val.p));
      DFM_TRY(gather_dia(M->A32, M->A64.rows.p, ord, A->d_dia.p));
      M->a32_valid = true;
    } else {
      M->a32_valid = false;
    }
    M->a_version = A->version;
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
#define DFM_LAUNCH_FIN(ND, FIN, partials, nparts, st)                                             \
  do {                                                                                            \
    hipLaunchKernelGGL((k_finalize<ND, FIN>), dim3(1), dim3(kFinNT), 0, ::dfm::stream(), partials, \
                       nparts, st);                                                               \
    DFM_HIP(hipGetLastError());                                                                   \
  } while (0)

template <int B, typename T>
static int sweep_lower(const Sell<T>& L, const Sell<double>& plan, const T* rhs, T* out, const KState* st) {
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  for (int g = 0; g < ng; g++) {
    const int s0 = plan.group_slice_ptr[g], s1 = plan.group_slice_ptr[g + 1];
    DFM_LAUNCH((k_trsv_lower<B, T>), grid_for(s1 - s0, kWaves), s0, s1, plan.rows.p, plan.slice_ptr.p, plan.col.p,
               L.val.p, L.dval.p, rhs, out, st);
  }
  return DFM_OK;
}
static int upper_nparts(const Sell<double>& plan) {
  int n = 0;
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  for (int g = 0; g < ng; g++) n += grid_for(plan.group_slice_ptr[g + 1] - plan.group_slice_ptr[g], kWaves);
  return n;
}
// Backward sweep; with DOT the partial sums of <rvec, vec> over all rows are
// left in partials[0 .. upper_nparts) for a finalize kernel.
template <int B, typename T, bool DOT>
static int sweep_upper(const Sell<T>& U, const Sell<double>& plan, T* vec, const T* rvec, double* partials,
                       const KState* st) {
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  const int nparts = DOT ? upper_nparts(plan) : 0;
  int part0 = 0;
  for (int g = 0; g < ng; g++) {
    const int s0 = plan.group_slice_ptr[g], s1 = plan.group_slice_ptr[g + 1];
    const int grid = grid_for(s1 - s0, kWaves);
    DFM_LAUNCH((k_trsv_upper<B, T, DOT>), grid, s0, s1, plan.rows.p, plan.slice_ptr.p, plan.col.p, U.val.p, vec,
               rvec, partials, part0, nparts, st);
    part0 += grid;
  }
  return DFM_OK;
}

int ilu_sweep64(dfm_ilu_s* M, double* vec, bool lower, bool upper) {
  switch (M->b) {
    case 1:
      if (lower) DFM_TRY((sweep_lower<1, double>(M->L64, M->L64, vec, vec, nullptr)));
      if (upper) DFM_TRY((sweep_upper<1, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr)));
      break;
    case 2:
      if (lower) DFM_TRY((sweep_lower<2, double>(M->L64, M->L64, vec, vec, nullptr)));
      if (upper) DFM_TRY((sweep_upper<2, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr)));
      break;
    case 3:
      if (lower) DFM_TRY((sweep_lower<3, double>(M->L64, M->L64, vec, vec, nullptr)));
      if (upper) DFM_TRY((sweep_upper<3, double, false>(M->U64, M->U64, vec, nullptr, nullptr, nullptr)));
      break;
    default:
      set_error("block size %u: the device substitutions support 1x1, 2x2 and 3x3 blocks", M->b);
      return DFM_NOT_IMPLEMENTED;
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Krylov cores.  Plain views of one precision's operators.
template <typename T>
struct Ops {
  const Sell<T>* A;
  const Sell<T>* L;
  const Sell<T>* U;
  const Sell<double>* planA;
  const Sell<double>* planL;
  const Sell<double>* planU;
  int nb;
};

template <typename T>
static Ops<T> ops_of(dfm_ilu_s* M);
template <>
Ops<double> ops_of<double>(dfm_ilu_s* M) {
  return {&M->A64, &M->L64, &M->U64, &M->A64, &M->L64, &M->U64, (int)M->nb};
}
template <>
Ops<float> ops_of<float>(dfm_ilu_s* M) {
  return {&M->A32, &M->L32, &M->U32, &M->A64, &M->L64, &M->U64, (int)M->nb};
}

template <int B, typename T, bool DOTS, bool GATE>
static int spmv(const Ops<T>& o, const T* x, T* y, const T* res, const T* w0, const T* u2, const T* v2,
                double* partials, const KState* st) {
  const int nslice = o.planA->nslice;
  DFM_LAUNCH((k_spmv<B, T, DOTS, GATE>), grid_for(nslice, kWaves), nslice, o.nb, o.planA->slice_ptr.p,
             o.planA->col.p, o.A->val.p, o.A->dval.p, x, y, res, w0, u2, v2, partials, st);
  return DFM_OK;
}
template <typename T>
static int spmv_nparts(const Ops<T>& o) {
  return grid_for(o.planA->nslice, kWaves);
}

static int read_state(dfm_ilu_s* M) {
  DFM_HIP(hipMemcpyAsync(M->h_state, M->state.p, sizeof(KState), hipMemcpyDeviceToHost, stream()));
  DFM_HIP(hipStreamSynchronize(stream()));
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
  if (s == nullptr) {
    for (int i = 0; i < nblock; i++) DFM_TRY(body());
    return DFM_OK;
  }
  for (auto& g : M->graphs)
    if (g.first == key) {
      DFM_HIP(hipGraphLaunch(g.second, s));
      return DFM_OK;
    }
  hipGraph_t graph = nullptr;
  DFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = DFM_OK;
  for (int i = 0; i < nblock && rc == DFM_OK; i++) rc = body();
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != DFM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  DFM_HIP(e);
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DFM_HIP(e);
  M->graphs.emplace_back(key, exec);
  DFM_HIP(hipGraphLaunch(exec, s));
  return DFM_OK;
}

static std::string graph_key(const char* tag, int B, int nblock, std::initializer_list<const void*> ptrs) {
  std::string k = std::string(tag) + ":" + std::to_string(B) + ":" + std::to_string(nblock);
  char buf[32];
  for (const void* p : ptrs) {
    snprintf(buf, sizeof buf, ":%p", p);
    k += buf;
  }
  return k;
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

// One PCG iteration (solver_ls_iter.cpp:134-165) on the device.
template <int B, typename T>
static int pcg_body(dfm_ilu_s* M, const Ops<T>& o, T* r, T* x, T* p, T* z) {
  const long n = (long)o.nb * B;
  KState* st = M->state.p;
  double* part = M->partials.p;
  const int gv = grid_vec(n);
  DFM_TRY((spmv<B, T, true, true>(o, p, z, nullptr, p, nullptr, nullptr, part, st)));
  DFM_LAUNCH_FIN(1, FIN_PCG_PAP, part, spmv_nparts(o), st);
  DFM_LAUNCH(k_pcg_update_xr<T>, gv, n, z, p, r, x, part, st);
  DFM_LAUNCH_FIN(1, FIN_PCG_RR, part, gv, st);
  DFM_TRY((sweep_lower<B, T>(*o.L, *o.planL, r, z, st)));
  DFM_TRY((sweep_upper<B, T, true>(*o.U, *o.planU, z, r, part, st)));
  DFM_LAUNCH_FIN(1, FIN_PCG_RZ, part, upper_nparts(*o.planU), st);
  DFM_LAUNCH(k_pcg_update_p<T>, gv, n, z, p, st);
  return DFM_OK;
}

// PCG start: z = M^-1 r, rz = <r,z>, p = z (:128-132)
template <int B, typename T>
static int pcg_start(dfm_ilu_s* M, const Ops<T>& o, T* r, T* p, T* z) {
  const long n = (long)o.nb * B;
  KState* st = M->state.p;
  DFM_TRY((sweep_lower<B, T>(*o.L, *o.planL, r, z, st)));
  DFM_TRY((sweep_upper<B, T, true>(*o.U, *o.planU, z, r, M->partials.p, st)));
  DFM_LAUNCH_FIN(1, FIN_PCG_RZ0, M->partials.p, upper_nparts(*o.planU), st);
  DFM_HIP(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));
  return DFM_OK;
}

// LsSol::Solve_PCG (solver_ls_iter.cpp:100-172).  r: in rhs, out final residual; x: out.
template <int B, typename T>
static int pcg_core(dfm_ilu_s* M, const Ops<T>& o, T* r, T* x, T* p, T* z, double tol, unsigned maxit,
                    unsigned check_every, KResult& res) {
  const long n = (long)o.nb * B;
  double rr0 = 0.0;
  DFM_TRY(device_dot<T>(M, n, r, r, &rr0));
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, x);
  if (rr0 < 1.0e-30) {
    res = {0.0, 0u, true};
    return DFM_OK;
  }
  init_state(*M->h_state, rr0, tol, maxit, 0);
  DFM_TRY(write_state(M));
  DFM_TRY((pcg_start<B, T>(M, o, r, p, z)));
  const std::string key = graph_key(sizeof(T) == 8 ? "pcg64" : "pcg32", B, (int)check_every, {r, x, p, z, M->partials.p, M->state.p});
  for (;;) {
    DFM_TRY(run_block(M, key, (int)check_every, [&]() { return pcg_body<B, T>(M, o, r, x, p, z); }));
    DFM_TRY(read_state(M));
    if (M->h_state->done) break;
  }
  const KState& s = *M->h_state;
  if (s.done == 1) {
    res = {std::sqrt(s.conv_rr * s.inv_rr0), (unsigned)s.conv_iter, true};
  } else {
    res = {std::sqrt(s.rr * s.inv_rr0), maxit, false};
  }
  return DFM_OK;
}

// One BiCGSTAB iteration (solver_ls_iter.cpp:327-404).
template <int B, typename T>
static int bicg_body(dfm_ilu_s* M, const Ops<T>& o, T* r, T* x, T* w) {
  const long n = (long)o.nb * B;
  T *r2 = w, *p = w + n, *Mp = w + 2 * n, *AMp = w + 3 * n, *s = w + 4 * n, *Ms = w + 5 * n, *AMs = w + 6 * n;
  KState* st = M->state.p;
  double* part = M->partials.p;
  const int gv = grid_vec(n);
  DFM_TRY((sweep_lower<B, T>(*o.L, *o.planL, p, Mp, st)));
  DFM_TRY((sweep_upper<B, T, false>(*o.U, *o.planU, Mp, nullptr, nullptr, st)));
  DFM_TRY((spmv<B, T, true, true>(o, Mp, AMp, nullptr, r2, r, r2, part, st)));
  DFM_LAUNCH_FIN(3, FIN_BICG_ALPHA, part, spmv_nparts(o), st);
  DFM_LAUNCH(k_bicg_s<T>, gv, n, r, AMp, s, st);
  DFM_TRY((sweep_lower<B, T>(*o.L, *o.planL, s, Ms, st)));
  DFM_TRY((sweep_upper<B, T, false>(*o.U, *o.planU, Ms, nullptr, nullptr, st)));
  DFM_TRY((spmv<B, T, true, true>(o, Ms, AMs, nullptr, s, nullptr, nullptr, part, st)));
  DFM_LAUNCH_FIN(2, FIN_BICG_OMEGA, part, spmv_nparts(o), st);
  DFM_LAUNCH(k_bicg_xr<T>, gv, n, Mp, Ms, s, AMs, r2, x, r, part, st);
  DFM_LAUNCH_FIN(2, FIN_BICG_XR, part, gv, st);
  DFM_LAUNCH(k_bicg_p<T>, gv, n, r, AMp, p, st);
  return DFM_OK;
}

// LsSol::Solve_PBiCGSTAB (solver_ls_iter.cpp:291-408).
template <int B, typename T>
static int bicg_core(dfm_ilu_s* M, const Ops<T>& o, T* r, T* x, T* w, double tol, unsigned maxit,
                     unsigned check_every, KResult& res, bool& x_zeroed) {
  const long n = (long)o.nb * B;
  double rr0 = 0.0;
  DFM_TRY(device_dot<T>(M, n, r, r, &rr0));
  x_zeroed = false;
  if (rr0 < 1.0e-60) {
    res = {std::sqrt(rr0), 0u, true};
    return DFM_OK;
  }
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, x);
  x_zeroed = true;
  DFM_HIP(hipMemcpyAsync(w, r, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));      // r2 = r
  DFM_HIP(hipMemcpyAsync(w + n, r, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));  // p = r
  init_state(*M->h_state, rr0, tol, maxit, 1);
  DFM_TRY(write_state(M));
  const std::string key = graph_key(sizeof(T) == 8 ? "bicg64" : "bicg32", B, (int)check_every, {r, x, w, M->partials.p, M->state.p});
  if (!M->h_state->done) {
    for (;;) {
      DFM_TRY(run_block(M, key, (int)check_every, [&]() { return bicg_body<B, T>(M, o, r, x, w); }));
      DFM_TRY(read_state(M));
      if (M->h_state->done) break;
    }
  }
  const KState& sv = *M->h_state;
  if (sv.done == 1) {
    res = {std::sqrt(sv.conv_rr * sv.inv_rr0), (unsigned)sv.conv_iter, true};
  } else {
    res = {tol, maxit, false};  // the reference leaves conv_ratio and reports max_iter
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision with reliable updates: the Krylov recurrence runs in FP32
// (matrix, factor, vectors) with FP64 reduction accumulators.  Whenever the
// recursive residual has dropped by kDelta since the last update (or the FP32
// recurrence reports convergence), the FP32 partial solution is folded into
// the FP64 solution, the true residual b - A x is recomputed with the FP64
// matrix and replaces the FP32 residual.  The search direction is kept, so the
// Krylov information is not lost as in restarted refinement.  Convergence is
// decided on the FP64 true residual.
constexpr double kDelta2 = 1.0e-2;  // squared residual drop between reliable updates

template <int B>
static int true_residual64(dfm_ilu_s* M, const double* b, const double* x, double* r, double* rr) {
  Ops<double> o = ops_of<double>(M);
  DFM_TRY((spmv<B, double, true, false>(o, x, r, b, nullptr, nullptr, nullptr, M->partials.p, M->state.p)));
  DFM_LAUNCH_FIN(3, FIN_STORE, M->partials.p, spmv_nparts(o), M->state.p);
  DFM_TRY(read_state(M));
  *rr = M->h_state->d1;
  return DFM_OK;
}

template <int B>
static int mixed_core(dfm_ilu_s* M, bool bicg, const double* b, double* r64, double* x64, float* w32, double tol,
                      unsigned maxit, unsigned check_every, KResult& res, unsigned& nupdates) {
  const long n = (long)M->nb * B;
  Ops<float> o = ops_of<float>(M);
  float* r32 = w32;
  float* x32 = w32 + n;
  float* ws = w32 + 2 * n;  // PCG: p, z ; BiCGSTAB: r2, p, Mp, AMp, s, Ms, AMs
  nupdates = 0;
  double rr0 = 0.0;
  DFM_TRY(device_dot<double>(M, n, b, b, &rr0));
  DFM_LAUNCH(k_fill<double>, grid_vec(n), n, 0.0, x64);
  DFM_HIP(hipMemcpyAsync(r64, b, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  if (rr0 < (bicg ? 1.0e-60 : 1.0e-30)) {
    res = {bicg ? std::sqrt(rr0) : 0.0, 0u, true};
    return DFM_OK;
  }
  DFM_LAUNCH(k_to_f32, grid_vec(n), n, r64, r32);
  DFM_LAUNCH(k_fill<float>, grid_vec(n), n, 0.0f, x32);
  init_state(*M->h_state, rr0, tol, maxit, bicg ? 1 : 0);
  DFM_TRY(write_state(M));
  if (!bicg) {
    DFM_TRY((pcg_start<B, float>(M, o, r32, ws, ws + n)));
  } else {
    DFM_HIP(hipMemcpyAsync(ws, r32, n * sizeof(float), hipMemcpyDeviceToDevice, stream()));      // r2
    DFM_HIP(hipMemcpyAsync(ws + n, r32, n * sizeof(float), hipMemcpyDeviceToDevice, stream()));  // p
  }
  const std::string key =
      graph_key(bicg ? "bicg32r" : "pcg32r", B, (int)check_every, {r32, x32, ws, M->partials.p, M->state.p});
  double rr_ref = rr0, rr_true = rr0;
  bool converged = false;
  for (;;) {
    if (!M->h_state->done) {
      DFM_TRY(run_block(M, key, (int)check_every, [&]() {
        return bicg ? bicg_body<B, float>(M, o, r32, x32, ws) : pcg_body<B, float>(M, o, r32, x32, ws, ws + n);
      }));
      DFM_TRY(read_state(M));
    }
    KState& h = *M->h_state;
    if (!(h.done || h.rr < kDelta2 * rr_ref)) continue;
    // reliable update
    DFM_LAUNCH(k_reliable_update, grid_vec(n), n, x32, x64);
    DFM_TRY(true_residual64<B>(M, b, x64, r64, &rr_true));
    nupdates++;
    KState& hs = *M->h_state;
    if (rr_true * hs.inv_rr0 < hs.tol2) {
      converged = true;
      break;
    }
    if (hs.done == 2 || !std::isfinite(rr_true)) break;
    DFM_LAUNCH(k_to_f32, grid_vec(n), n, r64, r32);
    rr_ref = rr_true;
    hs.done = 0;
    hs.rr = rr_true;
    DFM_TRY(write_state(M));
  }
  const KState& s = *M->h_state;
  res = {std::sqrt(rr_true * s.inv_rr0), (unsigned)(bicg ? s.iter - 1 : s.iter), converged};
  return DFM_OK;
}

// ---------------------------------------------------------------------------
static int ensure_workspace(dfm_ilu_s* M, size_t n64, size_t n32) {
  if (M->w64.n < n64) DFM_TRY(M->w64.alloc(n64));
  if (n32 && M->w32.n < n32) DFM_TRY(M->w32.alloc(n32));
  size_t need = 0;
  need = std::max<size_t>(need, (size_t)grid_for(M->A64.nslice, kWaves));
  need = std::max<size_t>(need, (size_t)upper_nparts(M->U64));
  need = std::max<size_t>(need, (size_t)2048);
  need = 3 * need + 64;
  if (M->partials.n < need) DFM_TRY(M->partials.alloc(need));
  return DFM_OK;
}

template <int B>
static int solve_dispatch(dfm_bcrs_s* A, dfm_ilu_s* M, bool bicg, double* r_user, double* x_user, double* conv_ratio,
                          unsigned* iteration, dfm_solve_opts* opts) {
  const int nb = (int)M->nb;
  const long n = (long)nb * B;
  const bool mixed = opts && opts->precision == DFM_PREC_MIXED;
  const unsigned check_every = (opts && opts->check_every) ? opts->check_every : 16;
  const bool want_true = opts && (opts->want_true_relres || mixed);
  const double tol = *conv_ratio;
  const unsigned maxit = *iteration;
  DFM_TRY(ilu_refresh_sell(M, A, mixed));
  const int* ord = M->permuted ? M->d_order.p : nullptr;
  const size_t nvec64 = mixed ? 3 : (bicg ? 10 : 5);
  const size_t nvec32 = mixed ? (bicg ? 9 : 4) : 0;
  DFM_TRY(ensure_workspace(M, nvec64 * n, nvec32 * n));
  double* r = M->w64.p;
  double* x = M->w64.p + n;
  double* b = M->w64.p + 2 * n;
  DFM_LAUNCH((k_permute<double, double>), grid_for(n), nb, B, ord, r_user, mixed ? b : r, 1.0);
  if (want_true && !mixed) DFM_HIP(hipMemcpyAsync(b, r, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  KResult kr{0.0, 0u, false};
  int rc = DFM_OK;
  bool x_written = true;
  if (!mixed) {
    Ops<double> o = ops_of<double>(M);
    if (!bicg) {
      DFM_TRY((pcg_core<B, double>(M, o, r, x, M->w64.p + 3 * n, M->w64.p + 4 * n, tol, maxit, check_every, kr)));
    } else {
      DFM_TRY((bicg_core<B, double>(M, o, r, x, M->w64.p + 3 * n, tol, maxit, check_every, kr, x_written)));
    }
    *conv_ratio = kr.conv;
    if (kr.converged || !bicg) *iteration = kr.iter;
    else *iteration = maxit;
    rc = kr.converged ? DFM_OK : DFM_NOT_CONVERGED;
    if (want_true && x_written) {
      double rr = 0.0, bb2 = 0.0;
      DFM_TRY(device_dot<double>(M, n, b, b, &bb2));
      double* tmp = M->w64.p + 4 * n;  // free after the solve
      DFM_TRY(true_residual64<B>(M, b, x, tmp, &rr));
      opts->true_relres = bb2 > 0.0 ? std::sqrt(rr / bb2) : 0.0;
    }
  } else {
    unsigned nup = 0;
    DFM_TRY(mixed_core<B>(M, bicg, b, r, x, M->w32.p, tol, maxit, check_every, kr, nup));
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
  DFM_CHECK(M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  DFM_CHECK(!opts || opts->precision == DFM_PREC_FP64 || opts->precision == DFM_PREC_MIXED, DFM_BAD_ARG,
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

template <int B, typename T>
static int bench_kernel_t(dfm_bcrs_s* A, dfm_ilu_s* M, int which, int nrep, double* avg_ms, double* bytes) {
  const int nb = (int)M->nb;
  const long n = (long)nb * B;
  Ops<T> o = ops_of<T>(M);
  DevBuf<T> xv, yv;
  DFM_TRY(xv.alloc(n));
  DFM_TRY(yv.alloc(n));
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)1, xv.p);
  DFM_LAUNCH(k_fill<T>, grid_vec(n), n, (T)0, yv.p);
  const double s = sizeof(T), bb = (double)B * B;
  const double nnzA = (double)A->col.size();
  double nnzL = 0, nnzU = 0;
  for (unsigned i = 0; i < M->nb; i++) {
    nnzL += M->fdiaind[i] - M->frow[i];
    nnzU += M->frow[i + 1] - M->fdiaind[i];
  }
  auto run = [&]() -> int {
    if (which == 0) return spmv<B, T, false, false>(o, xv.p, yv.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (which == 1) return sweep_lower<B, T>(*o.L, *o.planL, xv.p, yv.p, nullptr);
    return sweep_upper<B, T, false>(*o.U, *o.planU, yv.p, nullptr, nullptr, nullptr);
  };
  if (which == 0) *bytes = (nnzA + nb) * bb * s + 4.0 * nnzA + 2.0 * n * s;
  else if (which == 1) *bytes = (nnzL + nb) * bb * s + 4.0 * nnzL + 4.0 * nb + 3.0 * n * s;
  else *bytes = nnzU * bb * s + 4.0 * nnzU + 4.0 * nb + 3.0 * n * s;
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
  return DFM_OK;
}

}  // namespace dfm

extern "C" {
int dfm_bench_kernel(dfm_bcrs* A, dfm_ilu* M, int which, int precision, int nrep, double* avg_ms, double* bytes) {
  DFM_CHECK(A && M && avg_ms && bytes && nrep > 0 && which >= 0 && which <= 2, DFM_BAD_ARG, "bad arguments");
  DFM_CHECK(M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  const bool f32 = precision == DFM_PREC_MIXED;
  DFM_TRY(dfm::ilu_refresh_sell(M, A, f32 || M->have32));
  switch (A->b) {
    case 1: return f32 ? dfm::bench_kernel_t<1, float>(A, M, which, nrep, avg_ms, bytes)
                       : dfm::bench_kernel_t<1, double>(A, M, which, nrep, avg_ms, bytes);
    case 2: return f32 ? dfm::bench_kernel_t<2, float>(A, M, which, nrep, avg_ms, bytes)
                       : dfm::bench_kernel_t<2, double>(A, M, which, nrep, avg_ms, bytes);
    case 3: return f32 ? dfm::bench_kernel_t<3, float>(A, M, which, nrep, avg_ms, bytes)
                       : dfm::bench_kernel_t<3, double>(A, M, which, nrep, avg_ms, bytes);
    default: return DFM_NOT_IMPLEMENTED;
  }
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
