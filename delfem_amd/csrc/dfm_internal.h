// dfm_internal.h -- internal structures of the delfem_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "dfm.h"

namespace dfm {

void set_error(const char* fmt, ...);
hipStream_t stream();
// Host-driven solver loops (the multi-segment and complex CGNR / CG ones)
// defer the stream synchronisation that each entry point otherwise ends
// with: every call enqueues on the one stream, reductions read back by the
// host still wait for their own results, and the loop synchronises once at
// its end.
extern thread_local int g_sync_defer;
struct SyncDefer {
  SyncDefer() { ++g_sync_defer; }
  ~SyncDefer() { --g_sync_defer; }
};
int sync_stream();  // hipStreamSynchronize unless deferred
std::string sweep_mode_key();  // the triangular sweeps' mode, as part of a graph's key
// Graph caches (per factor) keep the iteration blocks of at most kMaxGraphs
// operand sets; before a new capture the oldest goes, once the stream has
// drained (a replay of it may still run).
constexpr size_t kMaxGraphs = 8;
int evict_graph(std::vector<std::pair<std::string, hipGraphExec_t>>& graphs);
int ensure_init();
// Value versions, unique across all matrices (real and complex).
unsigned long long next_version();

#define DFM_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      ::dfm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return DFM_HIP_ERROR;                                                              \
    }                                                                                    \
  } while (0)

#define DFM_CHECK(cond, code, ...)   \
  do {                               \
    if (!(cond)) {                   \
      ::dfm::set_error(__VA_ARGS__); \
      return (code);                 \
    }                                \
  } while (0)

#define DFM_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != DFM_OK) return rc_; \
  } while (0)

// Owning device buffer (hipMalloc on the library's device).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    if (count == n && p) return DFM_OK;
    release();
    if (count == 0) return DFM_OK;
    DFM_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
    return DFM_OK;
  }
  int upload(const T* h, size_t count) {
    DFM_TRY(alloc(count));
    if (count) DFM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
    return DFM_OK;
  }
  int upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
  int download(T* h, size_t count) const {
    if (count) DFM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
    DFM_HIP(hipStreamSynchronize(stream()));
    return DFM_OK;
  }
};

constexpr int kSlice = 64;  // rows per SELL slice = one wavefront

// bfloat16 storage (the upper 16 bits of an IEEE float)
struct bf16s {
  unsigned short bits;
};

// ---------------------------------------------------------------------------
// Host-side algorithms (dfm_host.cpp)
long fem_pattern(unsigned nnode, unsigned nelem, unsigned npe, const unsigned* conn,
                 std::vector<unsigned>& rowptr, std::vector<unsigned>& col);
// returns DFM_OK or DFM_UNSORTED
// fill_blk (may be NULL): AddFracPtn(lev, aBlkFill) -- rows and columns that take fill of any level
int iluk_symbolic(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
                  int lev_fill, const std::vector<int>* order, std::vector<unsigned>& frow,
                  std::vector<unsigned>& fcol, std::vector<int>& flev, std::vector<unsigned>& fdiaind,
                  const std::vector<unsigned>* fill_blk = nullptr);
int greedy_colour(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
                  std::vector<int>& colour);
void colour_order(unsigned nb, const std::vector<int>& colour, int ncolour, std::vector<int>& order);
// colour_ptr (may be NULL): ncolour + 1 offsets of the colours in the new order
int bmc_order(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
              unsigned block_nodes, std::vector<int>& order, int& nblock, std::vector<int>* colour_ptr = nullptr);
int level_sets(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
               const std::vector<unsigned>& fdiaind, bool lower, std::vector<int>& level);

// Phase-tiled sweep plan of one triangle of a factor whose rows fall into
// phases: contiguous row ranges in factor order (the colours of a multicolour
// or block-multicolour ordering), swept one launch per phase (the lower
// triangle first phase first, the upper last phase first).  Inside a phase,
// rows depend on each other only through small connected components of the
// phase's own entries; a tile (one workgroup of nt threads, thread t = tile
// row t) holds whole components, at most nt rows, and runs them level by
// level (a row's local level: 1 + the largest local level of the in-phase
// rows it reads) with a workgroup barrier between levels.  A tile row's slots
// keep the factor's column order and come in two parts, laid out as a
// sliced-ELL array whose slices are the tiles' 64-row waves: lower rows first
// their columns in earlier phases (col = factor row), then their in-phase
// columns (col = tile row of that column); upper rows first their in-phase
// columns, then those in later phases.  split[s] is the first slot of slice
// s's second part; -1 columns pad each part to the slice's widest row.
struct TilePlan {
  int nt = 0;                       // threads (= rows) per tile: 64, 128 or 256
  int ntile = 0;
  std::vector<int> phase_tile;      // launch p covers tiles [phase_tile[p], phase_tile[p + 1])
  std::vector<int> rows, rlev;      // ntile * nt: factor row and local level per tile row (-1: padding)
  std::vector<int> depth;           // per tile: number of local levels
  std::vector<int> slice_ptr;       // ntile * nt / 64 + 1
  std::vector<int> split;           // per slice
  std::vector<int> rov;             // ntile * nt * 4: per tile row, in the tile's overflow LDS, the first
                                    // slot and count of its in-phase slots beyond kTileInSlots, then of
                                    // (upper) its later-phase products beyond kTileOutSlots
  int ovf_in = 0, ovf_out = 0;      // largest overflow of a tile, slots
  long nk = 0;                      // slots per lane over all slices
  std::vector<int> col;             // nk * 64
  std::vector<long long> src;       // nk * 64: factor entry, -1 padding
};
// Split-lane form of a TilePlan for b x b blocks (b > 1): g = 2 (b = 2) or
// 4 (b = 3) lanes per tile row, lane t*g + r holding row r of the row's
// blocks (lanes r >= b idle).  Waves (64 lanes, 64/g tile rows) are the
// slices; per slice the two parts of TilePlan, padded to the widest of its
// rows.  Per slot-lane: col (the TilePlan encoding), src = q*b + r (the
// b values of block row r of factor entry q); rows[lane] = row*b + r (the
// diagonal block's row r), -1 for padding and idle lanes.
struct TileSplit {
  int g = 0, b = 0, nslice = 0;
  long nk = 0;
  std::vector<int> slice_ptr, split, rows, col;
  std::vector<long long> src;
};
void build_tile_split(const TilePlan& tp, int b, TileSplit& ts);

// Largest component a tile can hold.
constexpr int kTileMaxRows = 256;
// In-phase slots a tile row holds in registers, and (upper) later-phase
// products; a row's further slots go to the workgroup's overflow LDS, of at
// most kTileOverflowBytes (FP64 values) per tile.
constexpr int kTileInSlots = 4;
constexpr int kTileOutSlots = 8;
constexpr int kTileOverflowBytes = 48 * 1024;
// Builds the plan for b x b blocks; pack_rows: components are packed into one
// tile while it holds at most that many rows (a larger component takes a tile
// of its own).  Returns false (plan unusable) when a component exceeds
// kTileMaxRows rows or a tile's overflow exceeds kTileOverflowBytes.
bool build_tile_plan(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
                     const std::vector<unsigned>& fdiaind, const std::vector<int>& phase_ptr, bool upper,
                     int pack_rows, int b, TilePlan& tp);
// Phase boundaries (row offsets, first 0, last nb) for the tiled sweeps of a
// factor without a colour ordering; false when level-set sweeps (nlev level
// sets) are estimated to be as fast.
bool discover_phases(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
                     const std::vector<unsigned>& fdiaind, int nlev, std::vector<int>& phase_ptr);

// Sliced-ELL layout plan: rows grouped into "groups" (level sets, or one group
// for a plain SpMV), each group cut into slices of kSlice rows.  For slice s
// and block slot k (slice_ptr[s] <= k < slice_ptr[s+1]) lane l holds the
// block whose source index is src[k*kSlice+l] (-1: padding).
struct SellPlan {
  int nslice = 0;
  long nk = 0;                      // total block slots per lane
  std::vector<int> slice_ptr;       // nslice+1
  std::vector<int> rows;            // nslice*kSlice, -1 padding
  std::vector<int> col;             // nk*kSlice, -1 padding
  std::vector<long long> src;       // nk*kSlice, -1 padding
  std::vector<int> group_slice_ptr; // ngroup+1
};
// rows_by_group: rows listed group after group; group_ptr: ngroup+1 offsets.
// For each row, entries [beg(row), end(row)) of (colv, srcv) are laid out in order.
void build_sell_plan(const std::vector<int>& rows_by_group, const std::vector<int>& group_ptr,
                     const std::vector<unsigned>& beg, const std::vector<unsigned>& end,
                     const std::vector<unsigned>& colv, const std::vector<long long>* srcv,
                     SellPlan& plan);
// ---------------------------------------------------------------------------
template <typename T>
struct Sell {
  int nslice = 0;
  long nk = 0;
  int bb = 0;
  DevBuf<int> slice_ptr, rows, col;
  DevBuf<T> val, dval;  // val: nk*bb*kSlice, dval: nslice*bb*kSlice
  std::vector<int> group_slice_ptr;
};


// Deterministic reduction state shared by the Krylov kernels.
struct KState {
  double rz, pAp, rr, inv_rr0, tol2;
  double alpha, beta, omega, r_r2, rz_new;
  double d0, d1, d2;    // generic reduction outputs
  double conv_rr;       // rr at convergence
  int done;             // 0 running, 1 converged (recursive residual), 2 stopped at maxit
  int iter, conv_iter, maxit;
  int sweep_fault;      // a persistent sweep gave up waiting for a level (see k_trsv_persist)
};

// Work and level counters of the persistent sweep over one sliced-ELL plan:
// slev[slice] = its level, gsp = the plan's group_slice_ptr, sync = work
// counter, exit counter, one done-count per level, each on its own 256-byte
// line (kept zero between launches).
struct SweepSync {
  DevBuf<int> slev, gsp, sync;
  int nslice = 0, ngroup = 0, maxlev = 0;  // maxlev: slices of the largest level
};

// Device copy of a TilePlan: the layout (slice_ptr, rows, col of v64) and the
// value copies in each storage type (v32 / v16 made with the factor's
// FP32 / BF16 copies).  ntile == 0: the factor has no usable phase tiling.
struct TileSweep {
  int nt = 0, ntile = 0, ovf_in = 0, ovf_out = 0;
  std::vector<int> phase_tile;
  DevBuf<int> rlev, depth, split, rov;
  DevBuf<long long> src;
  Sell<double> v64;
  Sell<float> v32;
  Sell<bf16s> v16;
  // split-lane layout (TileSplit; g = 0: none): slice_ptr, rows, col of s64,
  // values per storage type
  int g = 0;
  DevBuf<int> ssplit;
  DevBuf<long long> ssrc;
  Sell<double> s64;
  Sell<float> s32;
  Sell<bf16s> s16;
};

// dfm_ilu_create, or (given_row / given_col, natural order) a factor on a
// pattern computed by the caller that holds A's (cross-segment fill).
int ilu_create_impl(const dfm_bcrs_s* A, int lev_fill, int ordering, const int* order, unsigned block_nodes,
                    const std::vector<unsigned>* given_row, const std::vector<unsigned>* given_col, dfm_ilu_s** out,
                    const std::vector<unsigned>* fill_blk = nullptr);
// Numeric ILU in two steps (dfm_ilu_set_value = copy + factor): between
// them a caller may change the copied values (cross-segment updates).
int ilu_numeric_copy(dfm_ilu_s* M, const dfm_bcrs_s* A);
// the copy of factor rows [row0, nb) only
int ilu_numeric_copy_rows(dfm_ilu_s* M, const dfm_bcrs_s* A, int row0);
// DoILUDecomp.  With more than 10 singular pivots the factor is left as the
// reference leaves it: rows before the 11th singular pivot (in factor
// numbering) factored, that row updated but not scaled, the rows after it as
// they were before the factorisation -- restore(row0) must put rows
// [row0, nb) back in that state (NULL: ilu_numeric_copy_rows from A).
typedef int (*ilu_restore_fn)(void* ctx, int row0);
int ilu_numeric_factor(dfm_ilu_s* M, dfm_bcrs_s* A, int* nsingular, ilu_restore_fn restore = nullptr,
                       void* restore_ctx = nullptr);
int ilu_sweep64(dfm_ilu_s* M, double* vec, bool lower, bool upper);

// Sliced-ELL copy of a CRS block matrix for its MatVec (built on first use,
// its values refreshed whenever the matrix's value version differs from the
// one copied): slices of 64 block rows, slot k of lane l = the row's k-th CRS
// entry (column, -1 for padding after the row's last entry), the entry's bs
// values at val[(k * bs + e) * 64 + l].  The MatVec kernel over it adds the
// entries in CRS order as the one-thread-per-row kernels do, bit for bit.
struct MvSell {
  int nslice = 0;
  long nk = 0;
  DevBuf<int> sptr, col;
  DevBuf<long long> src;
  DevBuf<double> val;
  unsigned long long plan_key = 0;  // pattern the plan was built for
  unsigned long long version = 0;   // values copied
};
bool mv_sell_on();  // DFM_MV_SELL=0: the one-thread-per-row kernels (comparisons)
// plan (if pattern_key changed) and values (if version changed) of nrow rows
int mv_sell_prepare(MvSell& m, unsigned nrow, const std::vector<unsigned>& rowptr,
                    const std::vector<unsigned>& col, int bs, const double* dval, unsigned long long pattern_key,
                    unsigned long long version);
// y = beta y + alpha A x over the copy; form as dfm_bcrs (square, diagonal
// blocks dia) or dfm_bmat (rectangular bc x br); false when no kernel form fits
bool mv_sell_launch(const MvSell& m, unsigned nrow, bool square, int bc, int br, const double* dia, double alpha,
                    const double* x, double beta, double* y);

}  // namespace dfm

struct dfm_bcrs_s {
  unsigned nb = 0, b = 0;
  std::vector<unsigned> rowptr, col;   // host pattern (CRS without diagonal)
  dfm::DevBuf<int> d_rowptr, d_col;
  dfm::DevBuf<double> d_val, d_dia;
  unsigned long long version = 0;       // unique across matrices, renewed whenever values change
  unsigned long long pattern_hash = 0;  // FNV-1a of (nb, rowptr, col), refreshed when the pattern changes
  dfm::MvSell mv;                       // MatVec copy, on first use
};

struct dfm_asm_s {
  dfm_bcrs_s* A = nullptr;
  unsigned nnode = 0, nelem = 0, npe = 0;
  dfm::DevBuf<unsigned> d_conn;
  dfm::DevBuf<int> d_elsup_ptr, d_elsup;  // node -> elements (ascending)
  dfm::DevBuf<int> d_slot;                 // per (node, element-slot, local j): CRS index, -1 diag, -2 dropped
};

struct dfm_bmat_s {
  unsigned nb = 0, b = 0;               // rows: nb blocks of b dofs
  unsigned nbr = 0, br = 0;             // columns: nbr blocks of br dofs
  std::vector<unsigned> rowptr, col;    // host pattern (the diagonal block, if any, inside the rows)
  dfm::DevBuf<int> d_rowptr, d_col;
  dfm::DevBuf<double> d_val;            // ncrs * b * br, row-major blocks
  dfm::DevBuf<long long> d_src;         // per entry: CRS index in the source matrix, or -1-i for its diagonal block i
  unsigned long long src_hash = 0;      // pattern hash of the matrix the pattern was made from
  unsigned long long version = 0;       // renewed whenever values change (next_version)
  unsigned long long pattern_key = 0;   // renewed whenever the pattern changes
  dfm::MvSell mv;                       // MatVec copy, on first use
};

struct dfm_ilu_s {
  unsigned nb = 0, b = 0;
  int lev_fill = 0, ordering = DFM_ORDER_NATURAL, ncolour = 0;
  bool permuted = false;
  std::vector<int> order, inv;          // order[new] = old
  std::vector<unsigned> frow, fcol, fdiaind;
  std::vector<int> flev;
  int nlev_lower = 0, nlev_upper = 0;
  // device factor in CRS (factor numbering)
  dfm::DevBuf<int> d_frow, d_fcol, d_fdiaind, d_order, d_inv;
  dfm::DevBuf<long long> d_fsrc;        // A block per factor entry, -1 for fill
  dfm::DevBuf<double> d_fval, d_fdia;
  dfm::DevBuf<int> d_rows_L;            // rows sorted by lower level
  std::vector<int> lev_ptr_L;
  std::vector<int> lev_maxrow_L;         // longest factor row (entries) of each lower level
  dfm::DevBuf<int> d_nsing;
  dfm::DevBuf<int> d_singrow;           // per factor row: the pivot was singular
  dfm::DevBuf<int> d_fixrow;            // the row DoILUDecomp stopped in (one entry)
  // triangular-solve layouts (value copies refreshed after each factorisation)
  dfm::SellPlan planL, planU;
  bool upper0_empty = false;            // first backward level has no U entries (true for level sets)
  dfm::DevBuf<unsigned char> d_rzmaskL; // per forward-sweep slice lane: row lies in the first backward level
  dfm::DevBuf<long long> d_srcL, d_srcU;
  dfm::Sell<double> L64, U64;
  dfm::SweepSync syncL, syncU;          // persistent-sweep counters of planL / planU
  std::vector<int> phase_ptr;           // colour offsets in factor order (multicolour orderings)
  dfm::TileSweep tileL, tileU;          // phase-tiled sweep layouts (ntile 0: none)
  dfm::Sell<float> L32, U32;
  dfm::Sell<dfm::bf16s> L16, U16;
  bool have32 = false, have16 = false;
  // permuted copy of A for the Krylov SpMV
  dfm::SellPlan planA;
  dfm::DevBuf<long long> d_srcA;
  dfm::Sell<double> A64;
  dfm::Sell<float> A32;
  unsigned long long a_version = 0;     // version of A mirrored in A64 / A32
  bool a32_valid = false;
  unsigned long long pattern_hash = 0;  // pattern of the matrix the factor was built on
  bool factored = false;
  // Krylov workspace (factor numbering)
  dfm::DevBuf<char> work;               // Krylov vectors (typed views carved per solve)
  dfm::DevBuf<double> partials;
  dfm::DevBuf<dfm::KState> state;
  dfm::KState* h_state = nullptr;       // pinned mirror
  // instantiated hipGraphs of blocks of Krylov iterations, keyed by their operands
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
  ~dfm_ilu_s() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (h_state) (void)hipHostFree(h_state);
  }
};
