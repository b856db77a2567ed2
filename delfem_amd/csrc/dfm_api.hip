// dfm_api.hip -- C-ABI of delfem_amd: runtime, CMatDia_BlkCrs, assembly and
// ILU (symbolic on the host, numeric and substitutions on the device).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <mutex>

#include "dfm_kernels.h"

namespace dfm {

static thread_local std::string g_err;
static int g_device = -1;
static hipStream_t g_stream = nullptr;
static bool g_own_stream = true;
static bool g_inited = false;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

hipStream_t stream() { return g_stream; }

// Scratch of the allocation-free entry points: a file-scope device array,
// one copy per device, whose address each device gets at set-up (dfm_init)
// through hipGetSymbolAddress.  Host code uses only those addresses, never
// the array's name.  Calls that overlap on one device share it and race.
constexpr size_t kScratchDoubles = (size_t)1 << 22;  // 4 Mi doubles = 32 MiB per device
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
static double* g_scratch[kMaxScratchDevices] = {};

// with `device` current
static int scratch_setup(int device) {
  if (device >= kMaxScratchDevices || g_scratch[device]) return DFM_OK;
  void* p = nullptr;
  DFM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  DFM_CHECK(p, DFM_HIP_ERROR, "no device address for the scratch array");
  g_scratch[device] = static_cast<double*>(p);
  return DFM_OK;
}

int ensure_init() {
  if (g_inited) return DFM_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  return dfm_init(dev);
}

static inline int grid_for(long n, int per = kNT) { return (int)std::max<long>(1, (n + per - 1) / per); }
static inline int grid_vec(long n) { return (int)std::min<long>(std::max<long>(1, (n + kNT - 1) / kNT), 2048); }

#define DFM_LAUNCH(kernel, grid, ...)                                                    \
  do {                                                                                   \
    if ((grid) > 0) {                                                                    \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(kNT), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                        \
    }                                                                                    \
  } while (0)
#define DFM_LAUNCH_NT(kernel, grid, nt, ...)                                              \
  do {                                                                                   \
    if ((grid) > 0) {                                                                    \
      hipLaunchKernelGGL((kernel), dim3(grid), dim3(nt), 0, ::dfm::stream(), __VA_ARGS__); \
      DFM_HIP(hipGetLastError());                                                        \
    }                                                                                    \
  } while (0)

}  // namespace dfm

using namespace dfm;

// Implemented in dfm_solver.hip
namespace dfm {
int ilu_refresh_sell(dfm_ilu_s* M, const dfm_bcrs_s* A, bool want32, bool want16);
int ilu_sweep64(dfm_ilu_s* M, double* vec, bool lower, bool upper);
}

// Value versions are unique across all matrices, so a factor's mirrored copy
// of A (tagged with the version it copied) is refreshed whenever it is given
// a different matrix or changed values.
thread_local int dfm::g_sync_defer = 0;
int dfm::sync_stream() {
  if (g_sync_defer) return DFM_OK;
  DFM_HIP(hipStreamSynchronize(stream()));
  return DFM_OK;
}

unsigned long long dfm::next_version() {
  static std::atomic<unsigned long long> counter{1};
  return ++counter;
}

bool dfm::mv_sell_on() {
  static const bool on = [] {
    const char* e = getenv("DFM_MV_SELL");
    return !(e && e[0] == '0');
  }();
  return on;
}
int dfm::mv_sell_prepare(MvSell& m, unsigned nrow, const std::vector<unsigned>& rowptr,
                         const std::vector<unsigned>& col, int bs, const double* dval, unsigned long long pattern_key,
                         unsigned long long version) {
  if (m.plan_key != pattern_key || m.nslice == 0) {
    const int ns = (int)((nrow + 63) / 64);
    std::vector<int> sptr(ns + 1, 0);
    for (int s2 = 0; s2 < ns; s2++) {
      unsigned w = 0;
      for (unsigned i = (unsigned)s2 * 64; i < std::min(nrow, (unsigned)s2 * 64 + 64); i++)
        w = std::max(w, rowptr[i + 1] - rowptr[i]);
      sptr[s2 + 1] = sptr[s2] + (int)w;
    }
    const long nk = sptr[ns];
    std::vector<int> cl((size_t)std::max<long>(1, nk) * 64, -1);
    std::vector<long long> src(cl.size(), -1);
    for (unsigned i = 0; i < nrow; i++) {
      const int s2 = (int)(i / 64), l = (int)(i % 64);
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const size_t k = (size_t)sptr[s2] + (q - rowptr[i]);
        cl[k * 64 + l] = (int)col[q];
        src[k * 64 + l] = q;
      }
    }
    DFM_TRY(m.sptr.upload(sptr));
    DFM_TRY(m.col.upload(cl));
    DFM_TRY(m.src.upload(src));
    DFM_TRY(m.val.alloc((size_t)std::max<long>(1, nk) * 64 * bs));
    m.nslice = ns;
    m.nk = nk;
    m.plan_key = pattern_key;
    m.version = 0;
  }
  if (m.version != version) {
    const long n = m.nk * 64 * bs;
    if (n) DFM_LAUNCH(k_mv_gather, grid_for(n), n, bs, m.src.p, dval, m.val.p);
    m.version = version;
  }
  return DFM_OK;
}
bool dfm::mv_sell_launch(const MvSell& m, unsigned nrow, bool square, int bc, int br, const double* dia, double alpha,
                         const double* x, double beta, double* y) {
  const dim3 grid((unsigned)std::max(1, (m.nslice + 3) / 4)), blk(256);
#define DFM_MVS(F, C, R) \
  hipLaunchKernelGGL((k_sell_mv<F, C, R>), grid, blk, 0, stream(), (int)nrow, m.sptr.p, m.col.p, m.val.p, dia, alpha, x, beta, y)
  if (square) {
    switch (bc) {
      case 1: DFM_MVS(MV_SQ, 1, 1); return true;
      case 2: DFM_MVS(MV_SQ, 2, 2); return true;
      case 3: DFM_MVS(MV_SQ, 3, 3); return true;
      default: return false;
    }
  }
  switch (bc * 4 + br) {
    case 5: DFM_MVS(MV_SCALAR, 1, 1); return true;   // 1 x 1
    case 6: DFM_MVS(MV_ROW, 1, 2); return true;      // 1 x 2
    case 7: DFM_MVS(MV_ROW, 1, 3); return true;      // 1 x 3
    case 9: DFM_MVS(MV_COL, 2, 1); return true;      // 2 x 1
    case 13: DFM_MVS(MV_COL, 3, 1); return true;     // 3 x 1
    case 10: DFM_MVS(MV_GEN, 2, 2); return true;
    case 11: DFM_MVS(MV_GEN, 2, 3); return true;
    case 14: DFM_MVS(MV_GEN, 3, 2); return true;
    case 15: DFM_MVS(MV_GEN, 3, 3); return true;
    default: return false;
  }
#undef DFM_MVS
}

extern "C" {

const char* dfm_last_error(void) { return dfm::g_err.c_str(); }

int dfm_init(int device) {
  int count = 0;
  DFM_HIP(hipGetDeviceCount(&count));
  DFM_CHECK(device >= 0 && device < count, DFM_BAD_ARG, "device %d out of range (%d devices)", device, count);
  if (g_inited && g_device == device) return DFM_OK;
  DFM_HIP(hipSetDevice(device));
  DFM_TRY(scratch_setup(device));
  if (g_stream && g_own_stream) (void)hipStreamDestroy(g_stream);
  g_own_stream = true;
  DFM_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  g_inited = true;
  return DFM_OK;
}

int dfm_synchronize(void) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

void* dfm_stream(void) {
  if (ensure_init() != DFM_OK) return nullptr;
  return (void*)g_stream;
}

int dfm_set_stream(void* s) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipStreamSynchronize(g_stream));
  if (g_own_stream) (void)hipStreamDestroy(g_stream);
  g_stream = (hipStream_t)s;
  g_own_stream = false;
  return DFM_OK;
}

int dfm_malloc(void** p, size_t bytes) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipMalloc(p, bytes));
  return DFM_OK;
}
int dfm_free(void* p) {
  DFM_HIP(hipFree(p));
  return DFM_OK;
}
int dfm_memcpy_htod(void* d, const void* h, size_t bytes) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}
int dfm_memcpy_dtoh(void* h, const void* d, size_t bytes) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}
int dfm_memcpy_dtod(void* d, const void* s, size_t bytes) {
  DFM_TRY(ensure_init());
  DFM_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

int dfm_fem_pattern(unsigned nnode, unsigned nelem, unsigned npe, const unsigned* conn, unsigned* rowptr,
                    unsigned* col, size_t cap, size_t* ncrs) {
  DFM_CHECK(conn || nelem == 0, DFM_BAD_ARG, "conn is NULL");
  for (size_t t = 0; t < (size_t)nelem * npe; t++)
    DFM_CHECK(conn[t] < nnode, DFM_BAD_ARG, "conn[%zu]=%u >= nnode=%u", t, conn[t], nnode);
  std::vector<unsigned> rp, cl;
  fem_pattern(nnode, nelem, npe, conn, rp, cl);
  if (ncrs) *ncrs = cl.size();
  if (rowptr) std::memcpy(rowptr, rp.data(), sizeof(unsigned) * rp.size());
  if (col) {
    DFM_CHECK(cap >= cl.size(), DFM_BAD_ARG, "col capacity %zu < %zu", cap, cl.size());
    std::memcpy(col, cl.data(), sizeof(unsigned) * cl.size());
  }
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// MatVec::CMatDia_BlkCrs

// FNV-1a over the block-CRS pattern: a factor or copy built on one pattern
// checks the matrix it is later given against it.
static unsigned long long hash_pattern(const dfm_bcrs_s* A) {
  unsigned long long h = 1469598103934665603ull;
  auto mix = [&h](unsigned v) {
    for (int k = 0; k < 4; k++) {
      h ^= (v >> (8 * k)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  mix(A->nb);
  for (unsigned v : A->rowptr) mix(v);
  for (unsigned v : A->col) mix(v);
  return h;
}

int dfm_bcrs_create(unsigned nblk, unsigned len, dfm_bcrs** out) {
  DFM_CHECK(out && len > 0 && len <= 8, DFM_BAD_ARG, "bad arguments to dfm_bcrs_create (len=%u)", len);
  DFM_TRY(ensure_init());
  dfm_bcrs_s* A = new dfm_bcrs_s;
  A->nb = nblk;
  A->b = len;
  A->rowptr.assign(nblk + 1, 0);
  A->pattern_hash = hash_pattern(A);
  A->version = next_version();
  int rc = A->d_rowptr.alloc(nblk + 1);
  if (rc == DFM_OK) rc = A->d_dia.alloc((size_t)nblk * len * len);
  if (rc == DFM_OK && nblk) {
    if (hipMemsetAsync(A->d_rowptr.p, 0, sizeof(int) * (nblk + 1), g_stream) != hipSuccess ||
        hipMemsetAsync(A->d_dia.p, 0, sizeof(double) * A->d_dia.n, g_stream) != hipSuccess)
      rc = DFM_HIP_ERROR;
  }
  if (rc != DFM_OK) {
    delete A;
    return rc;
  }
  *out = A;
  return DFM_OK;
}

int dfm_bcrs_destroy(dfm_bcrs* A) {
  delete A;
  return DFM_OK;
}

// AddPattern(CIndexedArray) (matdia_blkcrs.cpp:346-505): the first pattern is
// taken as given; a later one is merged and every row re-sorted ascending.
// Values are reset to zero whenever the pattern changes.
int dfm_bcrs_add_pattern(dfm_bcrs* A, unsigned nrow, const unsigned* index, const unsigned* array) {
  DFM_CHECK(A && index && nrow <= A->nb, DFM_BAD_ARG, "bad arguments to dfm_bcrs_add_pattern");
  DFM_CHECK(index[0] == 0, DFM_BAD_ARG, "index[0] must be 0");
  const unsigned nadd = index[nrow];
  for (unsigned i = 0; i < nrow; i++) {
    DFM_CHECK(index[i + 1] >= index[i], DFM_BAD_ARG, "index not monotone at row %u", i);
    for (unsigned q = index[i]; q < index[i + 1]; q++) {
      DFM_CHECK(array[q] < A->nb, DFM_BAD_ARG, "column %u out of range in row %u", array[q], i);
      DFM_CHECK(array[q] != i, DFM_BAD_ARG, "diagonal entry in row %u (the CRS part excludes the diagonal)", i);
    }
  }
  if (nadd == 0) return DFM_OK;
  const unsigned nb = A->nb;
  std::vector<unsigned> rp(nb + 1, 0), cl;
  if (A->rowptr[nb] == 0) {
    for (unsigned i = 0; i < nrow; i++) rp[i + 1] = index[i + 1];
    for (unsigned i = nrow; i < nb; i++) rp[i + 1] = index[nrow];
    cl.assign(array, array + nadd);
  } else {
    bool included = true;
    {
      std::vector<int> mark(nb, 0);
      for (unsigned i = 0; i < nrow && included; i++) {
        for (unsigned q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) mark[A->col[q]] = 1;
        for (unsigned q = index[i]; q < index[i + 1]; q++)
          if (!mark[array[q]]) included = false;
        for (unsigned q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) mark[A->col[q]] = 0;
      }
    }
    if (included) return DFM_OK;
    std::vector<unsigned> stamp(nb, 0);
    for (unsigned i = 0; i < nb; i++) {
      const size_t r0 = cl.size();
      for (unsigned q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
        if (stamp[A->col[q]] != i + 1) {
          stamp[A->col[q]] = i + 1;
          cl.push_back(A->col[q]);
        }
      if (i < nrow)
        for (unsigned q = index[i]; q < index[i + 1]; q++)
          if (stamp[array[q]] != i + 1) {
            stamp[array[q]] = i + 1;
            cl.push_back(array[q]);
          }
      std::sort(cl.begin() + r0, cl.end());
      rp[i + 1] = (unsigned)cl.size();
    }
  }
  A->rowptr.swap(rp);
  A->col.swap(cl);
  A->pattern_hash = hash_pattern(A);
  std::vector<int> rpi(A->rowptr.begin(), A->rowptr.end()), cli(A->col.begin(), A->col.end());
  DFM_TRY(A->d_rowptr.upload(rpi));
  DFM_TRY(A->d_col.upload(cli));
  const size_t nval = (size_t)A->rowptr[nb] * A->b * A->b;
  DFM_TRY(A->d_val.alloc(nval));
  if (nval) DFM_HIP(hipMemsetAsync(A->d_val.p, 0, nval * sizeof(double), g_stream));
  DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double), g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}

int dfm_bcrs_info(const dfm_bcrs* A, unsigned* nblk, unsigned* len, unsigned* ncrs) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (nblk) *nblk = A->nb;
  if (len) *len = A->b;
  if (ncrs) *ncrs = A->rowptr[A->nb];
  return DFM_OK;
}

int dfm_bcrs_get_pattern(const dfm_bcrs* A, unsigned* rowptr, unsigned* col) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (rowptr) std::memcpy(rowptr, A->rowptr.data(), sizeof(unsigned) * (A->nb + 1));
  if (col && !A->col.empty()) std::memcpy(col, A->col.data(), sizeof(unsigned) * A->col.size());
  return DFM_OK;
}

int dfm_bcrs_set_zero(dfm_bcrs* A) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (A->d_val.n) DFM_HIP(hipMemsetAsync(A->d_val.p, 0, A->d_val.n * sizeof(double), g_stream));
  if (A->d_dia.n) DFM_HIP(hipMemsetAsync(A->d_dia.p, 0, A->d_dia.n * sizeof(double), g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}

int dfm_bcrs_set_values(dfm_bcrs* A, const double* val, const double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val && A->d_val.n)
    DFM_HIP(hipMemcpyAsync(A->d_val.p, val, A->d_val.n * sizeof(double), hipMemcpyHostToDevice, g_stream));
  if (dia && A->d_dia.n)
    DFM_HIP(hipMemcpyAsync(A->d_dia.p, dia, A->d_dia.n * sizeof(double), hipMemcpyHostToDevice, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}

int dfm_bcrs_get_values(const dfm_bcrs* A, double* val, double* dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val && A->d_val.n)
    DFM_HIP(hipMemcpyAsync(val, A->d_val.p, A->d_val.n * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  if (dia && A->d_dia.n)
    DFM_HIP(hipMemcpyAsync(dia, A->d_dia.p, A->d_dia.n * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

int dfm_bcrs_device_values(dfm_bcrs* A, double** val, double** dia) {
  DFM_CHECK(A, DFM_BAD_ARG, "NULL matrix");
  if (val) *val = A->d_val.p;
  if (dia) *dia = A->d_dia.p;
  A->version = next_version();  // the caller may write through the pointers
  return DFM_OK;
}

// Mearge of one element given on the host (matdia_blkcrs.cpp:165-258):
// applied to a host copy of the touched rows, then written back.
int dfm_bcrs_mearge(dfm_bcrs* A, unsigned nblkel, const unsigned* blkel, unsigned blksize, const double* emat) {
  DFM_CHECK(A && blkel && emat && blksize == A->b * A->b, DFM_BAD_ARG, "bad arguments to dfm_bcrs_mearge");
  const unsigned bb = blksize;
  for (unsigned ie = 0; ie < nblkel; ie++) DFM_CHECK(blkel[ie] < A->nb, DFM_BAD_ARG, "node out of range");
  for (unsigned ie = 0; ie < nblkel; ie++) {
    const unsigned i = blkel[ie];
    const unsigned q0 = A->rowptr[i], q1 = A->rowptr[i + 1];
    std::vector<double> rowv((size_t)(q1 - q0) * bb), d(bb);
    if (q1 > q0)
      DFM_HIP(hipMemcpyAsync(rowv.data(), A->d_val.p + (size_t)q0 * bb, rowv.size() * sizeof(double),
                             hipMemcpyDeviceToHost, g_stream));
    DFM_HIP(hipMemcpyAsync(d.data(), A->d_dia.p + (size_t)i * bb, bb * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    DFM_HIP(hipStreamSynchronize(g_stream));
    for (unsigned je = 0; je < nblkel; je++) {
      const double* in = emat + ((size_t)ie * nblkel + je) * bb;
      if (ie == je) {
        for (unsigned t = 0; t < bb; t++) d[t] += in[t];
      } else {
        const unsigned j = blkel[je];
        for (unsigned q = q0; q < q1; q++)
          if (A->col[q] == j) {
            for (unsigned t = 0; t < bb; t++) rowv[(size_t)(q - q0) * bb + t] += in[t];
            break;
          }
      }
    }
    if (q1 > q0)
      DFM_HIP(hipMemcpyAsync(A->d_val.p + (size_t)q0 * bb, rowv.data(), rowv.size() * sizeof(double),
                             hipMemcpyHostToDevice, g_stream));
    DFM_HIP(hipMemcpyAsync(A->d_dia.p + (size_t)i * bb, d.data(), bb * sizeof(double), hipMemcpyHostToDevice, g_stream));
    DFM_HIP(hipStreamSynchronize(g_stream));
  }
  A->version = next_version();
  return DFM_OK;
}

int dfm_bcrs_set_bc(dfm_bcrs* A, const int* flag) {
  DFM_CHECK(A && flag, DFM_BAD_ARG, "bad arguments to dfm_bcrs_set_bc");
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, (size_t)A->nb * A->b));
  DFM_LAUNCH(k_bc_rows, grid_for(A->nb), (int)A->nb, (int)A->b, A->d_rowptr.p, A->d_val.p, A->d_dia.p, f.p);
  const long ncrs = A->rowptr[A->nb];
  DFM_LAUNCH(k_bc_cols, grid_for(ncrs), ncrs, (int)A->b, A->d_col.p, A->d_val.p, f.p);
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}

int dfm_bcrs_matvec(dfm_bcrs* A, double alpha, const double* x, double beta, double* y) {
  DFM_CHECK(A && x && y && x != y, DFM_BAD_ARG, "bad arguments to dfm_bcrs_matvec");
  const int nb = (int)A->nb, g = grid_for(nb);
  if (nb > 0 && A->b <= 3 && mv_sell_on()) {
    DFM_TRY(mv_sell_prepare(A->mv, A->nb, A->rowptr, A->col, (int)(A->b * A->b), A->d_val.p, A->pattern_hash,
                            A->version));
    if (mv_sell_launch(A->mv, A->nb, true, (int)A->b, (int)A->b, A->d_dia.p, alpha, x, beta, y)) {
      DFM_HIP(hipGetLastError());
      return sync_stream();
    }
  }
  switch (A->b) {
    case 1: DFM_LAUNCH(k_crs_matvec<1>, g, nb, A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, alpha, x, beta, y); break;
    case 2: DFM_LAUNCH(k_crs_matvec<2>, g, nb, A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, alpha, x, beta, y); break;
    case 3: DFM_LAUNCH(k_crs_matvec<3>, g, nb, A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, alpha, x, beta, y); break;
    default:
      DFM_LAUNCH(k_crs_matvec_gen, g, nb, (int)A->b, A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, alpha, x,
                 beta, y);
  }
  return sync_stream();
}

// ---------------------------------------------------------------------------
// BLAS-1
static int vec_dot_impl(size_t n, const double* x, const double* y, double* result) {
  static DevBuf<double> partials;
  static DevBuf<KState> st;
  DFM_TRY(ensure_init());
  const int g = grid_vec((long)n);
  if (partials.n < (size_t)g) DFM_TRY(partials.alloc(g));
  if (!st.p) {
    DFM_TRY(st.alloc(1));
    DFM_HIP(hipMemsetAsync(st.p, 0, sizeof(KState), g_stream));
  }
  DFM_LAUNCH(k_dot<double>, g, (long)n, x, y, partials.p);
  hipLaunchKernelGGL((k_finalize<1, FIN_STORE>), dim3(1), dim3(kFinNT), 0, g_stream, partials.p, g, st.p);
  DFM_HIP(hipGetLastError());
  KState h;
  DFM_HIP(hipMemcpyAsync(&h, st.p, sizeof(KState), hipMemcpyDeviceToHost, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  *result = h.d0;
  return DFM_OK;
}
int dfm_vec_dot(size_t n, const double* x, const double* y, double* result) {
  DFM_CHECK(result, DFM_BAD_ARG, "NULL result");
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  return vec_dot_impl(n, x, y, result);
}
int dfm_vec_axpy(size_t n, double alpha, const double* x, double* y) {
  DFM_TRY(ensure_init());
  if (alpha == 0.0) return DFM_OK;  // CLinearSystem::AXPY (linearsystem.cpp:312)
  DFM_LAUNCH(k_axpy, grid_vec((long)n), (long)n, alpha, x, y);
  return sync_stream();
}
int dfm_vec_scal(size_t n, double alpha, double* x) {
  DFM_TRY(ensure_init());
  DFM_LAUNCH(k_scal, grid_vec((long)n), (long)n, alpha, x);
  return sync_stream();
}
int dfm_vec_fill(size_t n, double value, double* x) {
  DFM_TRY(ensure_init());
  DFM_LAUNCH(k_fill<double>, grid_vec((long)n), (long)n, value, x);
  return sync_stream();
}
int dfm_vec_copy(size_t n, const double* x, double* y) {
  DFM_TRY(ensure_init());
  if (x == y || n == 0) return DFM_OK;
  DFM_HIP(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  return sync_stream();
}
int dfm_vec_zero_bc(size_t n, const int* flag, double* x) {
  DFM_TRY(ensure_init());
  DevBuf<int> f;
  DFM_TRY(f.upload(flag, n));
  DFM_LAUNCH(k_zero_bc, grid_vec((long)n), (long)n, f.p, x);
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// Element assembly plans
int dfm_asm_create(dfm_bcrs* A, unsigned nnode, unsigned nelem, unsigned npe, const unsigned* conn, dfm_asm** out) {
  DFM_CHECK(A && conn && out && nnode == A->nb && npe > 0, DFM_BAD_ARG, "bad arguments to dfm_asm_create");
  for (size_t t = 0; t < (size_t)nelem * npe; t++)
    DFM_CHECK(conn[t] < nnode, DFM_BAD_ARG, "conn[%zu]=%u >= nnode", t, conn[t]);
  dfm_asm_s* p = new dfm_asm_s;
  p->A = A;
  p->nnode = nnode;
  p->nelem = nelem;
  p->npe = npe;
  std::vector<int> eptr(nnode + 1, 0), elist((size_t)nelem * npe);
  for (size_t t = 0; t < (size_t)nelem * npe; t++) eptr[conn[t] + 1]++;
  for (unsigned i = 0; i < nnode; i++) eptr[i + 1] += eptr[i];
  {
    std::vector<int> fill(eptr.begin(), eptr.end() - 1);
    for (unsigned e = 0; e < nelem; e++)
      for (unsigned k = 0; k < npe; k++) {
        const unsigned nd = conn[(size_t)e * npe + k];
        // an element lists a node once; repeated nodes would be merged twice by Mearge
        elist[fill[nd]++] = (int)e;
      }
  }
  int rc = p->d_conn.upload(conn, (size_t)nelem * npe);
  if (rc == DFM_OK) rc = p->d_elsup_ptr.upload(eptr);
  if (rc == DFM_OK) rc = p->d_elsup.upload(elist);
  if (rc == DFM_OK && hipStreamSynchronize(g_stream) != hipSuccess) rc = DFM_HIP_ERROR;
  if (rc != DFM_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return DFM_OK;
}
int dfm_asm_destroy(dfm_asm* p) {
  delete p;
  return DFM_OK;
}
int dfm_assemble_poisson2d_p1(dfm_asm* p, const double* coord, const double* alpha, double source,
                              const double* value, double* res) {
  DFM_CHECK(p && p->npe == 3 && p->A->b == 1 && coord && alpha && value, DFM_BAD_ARG,
            "dfm_assemble_poisson2d_p1 needs a 3-node plan on a 1x1 block matrix");
  dfm_bcrs_s* A = p->A;
  DFM_LAUNCH(k_asm_poisson2d, grid_for(p->nnode), (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p, p->d_elsup.p,
             A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, coord, alpha, source, value, res);
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}
int dfm_assemble_solid3d_p1(dfm_asm* p, const double* coord, const double* lambda, const double* myu, double rho,
                            double gx, double gy, double gz, const double* disp, double* res) {
  DFM_CHECK(p && p->npe == 4 && p->A->b == 3 && coord && lambda && myu && disp, DFM_BAD_ARG,
            "dfm_assemble_solid3d_p1 needs a 4-node plan on a 3x3 block matrix");
  dfm_bcrs_s* A = p->A;
  DFM_LAUNCH(k_asm_solid3d, grid_for(p->nnode), (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p, p->d_elsup.p,
             A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, coord, lambda, myu, rho, gx, gy, gz, disp, res);
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}
int dfm_assemble_oseen2d_p1p1(dfm_asm* p, const double* coord, double myu, double rho, const double* adv,
                              const double* value, double* res) {
  DFM_CHECK(p && p->npe == 3 && p->A->b == 3 && coord && adv && value, DFM_BAD_ARG,
            "dfm_assemble_oseen2d_p1p1 needs a 3-node plan on a 3x3 block matrix");
  DFM_CHECK(myu > 0.0 && rho > 0.0, DFM_BAD_ARG, "dfm_assemble_oseen2d_p1p1 needs myu > 0 and rho > 0");
  dfm_bcrs_s* A = p->A;
  DFM_LAUNCH(k_asm_oseen2d, grid_for(p->nnode), (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p, p->d_elsup.p,
             A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, coord, myu, rho, adv, value, res);
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}
int dfm_assemble_helmholtz2d_p1(dfm_asm* p, const double* coord, const double* alpha, const double* cmass,
                                double source, const double* value, double* res) {
  DFM_CHECK(p && p->npe == 3 && p->A->b == 1 && coord && alpha && cmass && value, DFM_BAD_ARG,
            "dfm_assemble_helmholtz2d_p1 needs a 3-node plan on a 1x1 block matrix");
  dfm_bcrs_s* A = p->A;
  DFM_LAUNCH(k_asm_helmholtz2d, grid_for(p->nnode), (int)p->nnode, p->d_conn.p, p->d_elsup_ptr.p, p->d_elsup.p,
             A->d_rowptr.p, A->d_col.p, A->d_val.p, A->d_dia.p, coord, alpha, cmass, source, value, res);
  DFM_HIP(hipStreamSynchronize(g_stream));
  A->version = next_version();
  return DFM_OK;
}

// ---------------------------------------------------------------------------
// ILU
// slice -> level map and zeroed counters of a plan's persistent sweep
static int build_sweep_sync(const SellPlan& plan, SweepSync& ss) {
  const int ng = (int)plan.group_slice_ptr.size() - 1;
  std::vector<int> slev((size_t)std::max(1, plan.nslice), 0);
  int maxlev = 0;
  for (int g = 0; g < ng; g++) {
    const int s0 = plan.group_slice_ptr[g], s1 = plan.group_slice_ptr[g + 1];
    for (int sl = s0; sl < s1; sl++) slev[sl] = g;
    maxlev = std::max(maxlev, s1 - s0);
  }
  ss.nslice = plan.nslice;
  ss.ngroup = ng;
  ss.maxlev = maxlev;
  DFM_TRY(ss.slev.upload(slev));
  DFM_TRY(ss.gsp.upload(plan.group_slice_ptr));
  const size_t nsync = ((size_t)ng + 2) * kSyncStride;
  DFM_TRY(ss.sync.alloc(nsync));
  DFM_HIP(hipMemsetAsync(ss.sync.p, 0, sizeof(int) * nsync, g_stream));
  return DFM_OK;
}

static int upload_sell_plan(const SellPlan& plan, int bb, Sell<double>& s64, DevBuf<long long>& dsrc) {
  s64.nslice = plan.nslice;
  s64.nk = plan.nk;
  s64.bb = bb;
  s64.group_slice_ptr = plan.group_slice_ptr;
  DFM_TRY(s64.slice_ptr.upload(plan.slice_ptr));
  DFM_TRY(s64.rows.upload(plan.rows));
  DFM_TRY(s64.col.upload(plan.col));
  DFM_TRY(dsrc.upload(plan.src));
  DFM_TRY(s64.val.alloc(std::max<size_t>(1, (size_t)plan.nk * bb * kSlice)));
  DFM_TRY(s64.dval.alloc(std::max<size_t>(1, (size_t)plan.nslice * bb * kSlice)));
  return DFM_OK;
}

static int upload_tile_plan(const TilePlan& tp, int b, TileSweep& ts) {
  const int bb = b * b;
  ts.nt = tp.nt;
  ts.ntile = tp.ntile;
  ts.phase_tile = tp.phase_tile;
  Sell<double>& v = ts.v64;
  v.nslice = tp.ntile * tp.nt / kSlice;
  v.nk = tp.nk;
  v.bb = bb;
  v.group_slice_ptr = {0, v.nslice};
  DFM_TRY(v.slice_ptr.upload(tp.slice_ptr));
  DFM_TRY(v.rows.upload(tp.rows));
  DFM_TRY(v.col.upload(tp.col));
  DFM_TRY(ts.src.upload(tp.src));
  DFM_TRY(ts.rlev.upload(tp.rlev));
  DFM_TRY(ts.depth.upload(tp.depth));
  DFM_TRY(ts.split.upload(tp.split));
  DFM_TRY(ts.rov.upload(tp.rov));
  ts.ovf_in = tp.ovf_in;
  ts.ovf_out = tp.ovf_out;
  if (b == 2 || b == 3) {
    TileSplit sp;
    build_tile_split(tp, b, sp);
    ts.g = sp.g;
    Sell<double>& w = ts.s64;
    w.nslice = sp.nslice;
    w.nk = sp.nk;
    w.bb = b;
    w.group_slice_ptr = {0, w.nslice};
    DFM_TRY(w.slice_ptr.upload(sp.slice_ptr));
    DFM_TRY(w.rows.upload(sp.rows));
    DFM_TRY(w.col.upload(sp.col));
    DFM_TRY(ts.ssplit.upload(sp.split));
    DFM_TRY(ts.ssrc.upload(sp.src));
    DFM_TRY(w.val.alloc(std::max<size_t>(1, (size_t)sp.nk * b * kSlice)));
    DFM_TRY(w.dval.alloc(std::max<size_t>(1, (size_t)w.nslice * b * kSlice)));
  }
  DFM_TRY(v.val.alloc(std::max<size_t>(1, (size_t)tp.nk * bb * kSlice)));
  DFM_TRY(v.dval.alloc(std::max<size_t>(1, (size_t)v.nslice * bb * kSlice)));
  return DFM_OK;
}

}  // extern "C"

namespace dfm {
// Rows a tile of the phase-tiled sweeps is filled to (DFM_TILE_ROWS, 1..256):
// smaller tiles spread a phase over more compute units.
static int g_tile_rows = [] {
  const char* e = getenv("DFM_TILE_ROWS");
  const int v = e ? atoi(e) : 64;
  return v >= 1 && v <= kTileMaxRows ? v : 64;
}();

// Phases found for factors without a colour ordering (DFM_DISCOVER_PHASES=0:
// off, for comparisons).
static bool g_discover_phases = [] {
  const char* e = getenv("DFM_DISCOVER_PHASES");
  return !e || atoi(e) != 0;
}();

// dfm_ilu_create; with given_row / given_col (natural order only) the factor
// pattern is that one instead of the ILU(k) symbolic of A's: it must hold
// A's pattern, and its extra entries start at zero
int ilu_create_impl(const dfm_bcrs* A, int lev_fill, int ordering, const int* order, unsigned block_nodes,
                    const std::vector<unsigned>* given_row, const std::vector<unsigned>* given_col, dfm_ilu** out,
                    const std::vector<unsigned>* fill_blk) {
  DFM_CHECK(A && out && lev_fill >= -1, DFM_BAD_ARG, "bad arguments to dfm_ilu_create");
  DFM_CHECK(!given_row || (given_col && ordering == DFM_ORDER_NATURAL && given_row->size() == A->nb + 1 &&
                           given_col->size() == given_row->back()),
            DFM_BAD_ARG, "bad given factor pattern");
  DFM_CHECK(ordering == DFM_ORDER_NATURAL || ordering == DFM_ORDER_GIVEN || ordering == DFM_ORDER_MULTICOLOUR ||
                ordering == DFM_ORDER_BLOCK_MULTICOLOUR,
            DFM_BAD_ARG, "unknown ordering %d", ordering);
  DFM_CHECK(ordering != DFM_ORDER_GIVEN || order, DFM_BAD_ARG, "DFM_ORDER_GIVEN needs an order array");
  DFM_TRY(ensure_init());
  const unsigned nb = A->nb, b = A->b, bb = b * b;
  std::unique_ptr<dfm_ilu_s> M(new dfm_ilu_s);
  M->nb = nb;
  M->b = b;
  M->lev_fill = lev_fill;
  M->ordering = ordering;
  M->pattern_hash = A->pattern_hash;
  if (ordering == DFM_ORDER_GIVEN) {
    M->order.assign(order, order + nb);
    std::vector<int> seen(nb, 0);
    for (unsigned i = 0; i < nb; i++) {
      DFM_CHECK(M->order[i] >= 0 && (unsigned)M->order[i] < nb && !seen[M->order[i]], DFM_BAD_ARG,
                "order is not a permutation (entry %u)", i);
      seen[M->order[i]] = 1;
    }
    M->permuted = true;
  } else if (ordering == DFM_ORDER_MULTICOLOUR) {
    std::vector<int> colour;
    M->ncolour = greedy_colour(nb, A->rowptr, A->col, colour);
    colour_order(nb, colour, M->ncolour, M->order);
    M->phase_ptr.assign(M->ncolour + 1, 0);
    for (unsigned i = 0; i < nb; i++) M->phase_ptr[colour[i] + 1]++;
    for (int c = 0; c < M->ncolour; c++) M->phase_ptr[c + 1] += M->phase_ptr[c];
    M->permuted = true;
  } else if (ordering == DFM_ORDER_BLOCK_MULTICOLOUR) {
    int nblock = 0;
    M->ncolour = bmc_order(nb, A->rowptr, A->col, block_nodes ? block_nodes : 32u, M->order, nblock, &M->phase_ptr);
    M->permuted = true;
  } else {
    M->order.resize(nb);
    for (unsigned i = 0; i < nb; i++) M->order[i] = (int)i;
    M->permuted = false;
  }
  M->inv.assign(nb, 0);
  for (unsigned i = 0; i < nb; i++) M->inv[M->order[i]] = (int)i;
  if (given_row) {
    M->frow = *given_row;
    M->fcol = *given_col;
    M->flev.assign(M->fcol.size(), 0);
    M->fdiaind.assign(nb, 0);
    for (unsigned i = 0; i < nb; i++) {
      const auto b0 = M->fcol.begin() + M->frow[i], b1 = M->fcol.begin() + M->frow[i + 1];
      DFM_CHECK(std::is_sorted(b0, b1) && std::adjacent_find(b0, b1) == b1 && std::find(b0, b1, i) == b1,
                DFM_BAD_ARG, "given factor pattern: row %u not strictly ascending or holds the diagonal", i);
      for (unsigned q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
        DFM_CHECK(std::binary_search(b0, b1, A->col[q]), DFM_BAD_ARG,
                  "given factor pattern: row %u lacks column %u of the matrix", i, A->col[q]);
      M->fdiaind[i] = (unsigned)(std::upper_bound(b0, b1, i) - M->fcol.begin());
    }
  } else {
    DFM_CHECK(!fill_blk || fill_blk->empty() || !M->permuted, DFM_BAD_ARG,
              "fill blocks are given for natural-order factors only");
    DFM_TRY(iluk_symbolic(nb, A->rowptr, A->col, lev_fill, M->permuted ? &M->order : nullptr, M->frow, M->fcol,
                          M->flev, M->fdiaind, fill_blk));
  }
  const size_t nf = M->fcol.size();
  // factor entry -> block of A
  std::vector<long long> fsrc(nf, -1);
  {
    std::vector<long long> pos(nb, -1);
    for (unsigned i = 0; i < nb; i++) {
      const unsigned io = (unsigned)M->order[i];
      for (unsigned q = A->rowptr[io]; q < A->rowptr[io + 1]; q++) pos[A->col[q]] = q;
      for (unsigned q = M->frow[i]; q < M->frow[i + 1]; q++) fsrc[q] = pos[M->order[M->fcol[q]]];
      for (unsigned q = A->rowptr[io]; q < A->rowptr[io + 1]; q++) pos[A->col[q]] = -1;
    }
  }
  // level sets
  std::vector<int> levL, levU;
  M->nlev_lower = level_sets(nb, M->frow, M->fcol, M->fdiaind, true, levL);
  M->nlev_upper = level_sets(nb, M->frow, M->fcol, M->fdiaind, false, levU);
  auto group_rows = [&](const std::vector<int>& lev, int nlev, std::vector<int>& rows, std::vector<int>& gptr) {
    gptr.assign(nlev + 1, 0);
    for (unsigned i = 0; i < nb; i++) gptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) gptr[l + 1] += gptr[l];
    rows.assign(nb, 0);
    std::vector<int> fill(gptr.begin(), gptr.end() - 1);
    for (unsigned i = 0; i < nb; i++) rows[fill[lev[i]]++] = (int)i;
  };
  std::vector<int> rowsL, gptrL, rowsU, gptrU;
  group_rows(levL, M->nlev_lower, rowsL, gptrL);
  group_rows(levU, M->nlev_upper, rowsU, gptrU);
  M->lev_ptr_L = gptrL;
  M->lev_maxrow_L.assign(M->nlev_lower, 0);
  for (unsigned i = 0; i < nb; i++)
    M->lev_maxrow_L[levL[i]] = std::max(M->lev_maxrow_L[levL[i]], (int)(M->frow[i + 1] - M->frow[i]));
  {
    std::vector<unsigned> begL(nb), endL(nb), begU(nb), endU(nb);
    for (unsigned i = 0; i < nb; i++) {
      begL[i] = M->frow[i];
      endL[i] = M->fdiaind[i];
      begU[i] = M->fdiaind[i];
      endU[i] = M->frow[i + 1];
    }
    build_sell_plan(rowsL, gptrL, begL, endL, M->fcol, nullptr, M->planL);
    build_sell_plan(rowsU, gptrU, begU, endU, M->fcol, nullptr, M->planU);
    // rows of the first backward level (no U entries), flagged per forward
    // slice lane for the fused PCG sweep
    const SellPlan& pu = M->planU;
    M->upper0_empty = pu.group_slice_ptr.size() >= 2 && pu.slice_ptr[pu.group_slice_ptr[1]] == pu.slice_ptr[0];
    std::vector<unsigned char> mask(M->planL.rows.size(), 0);
    for (size_t i = 0; i < mask.size(); i++) {
      const int row = M->planL.rows[i];
      mask[i] = row >= 0 && levU[row] == 0;
    }
    DFM_TRY(M->d_rzmaskL.upload(mask));
  }
  // phase-tiled sweep layouts: phases are the colours of a colour ordering,
  // else found in the factor's own numbering (a mesh the caller numbered by
  // colours of tiles) where they beat the level sets
  TilePlan tpL, tpU;
  std::vector<int> phases = M->phase_ptr;
  if (phases.size() < 2 && g_discover_phases)
    discover_phases(nb, M->frow, M->fcol, M->fdiaind, std::max(M->nlev_lower, M->nlev_upper), phases);
  const bool tiled = phases.size() >= 2 &&
                     build_tile_plan(nb, M->frow, M->fcol, M->fdiaind, phases, false, g_tile_rows, (int)b, tpL) &&
                     build_tile_plan(nb, M->frow, M->fcol, M->fdiaind, phases, true, g_tile_rows, (int)b, tpU);
  // permuted copy of A: rows in factor numbering, columns renumbered and sorted
  {
    std::vector<unsigned> beg(nb), end(nb), colv(A->col.size());
    std::vector<long long> srcv(A->col.size());
    std::vector<std::pair<unsigned, long long>> tmp;
    unsigned n = 0;
    for (unsigned i = 0; i < nb; i++) {
      const unsigned io = (unsigned)M->order[i];
      tmp.clear();
      for (unsigned q = A->rowptr[io]; q < A->rowptr[io + 1]; q++) tmp.push_back({(unsigned)M->inv[A->col[q]], q});
      if (M->permuted) std::sort(tmp.begin(), tmp.end());
      beg[i] = n;
      for (auto& t : tmp) {
        colv[n] = t.first;
        srcv[n] = t.second;
        n++;
      }
      end[i] = n;
    }
    std::vector<int> rows(nb), gptr = {0, (int)nb};
    for (unsigned i = 0; i < nb; i++) rows[i] = (int)i;
    build_sell_plan(rows, gptr, beg, end, colv, &srcv, M->planA);
  }
  // device uploads
  {
    std::vector<int> frow(M->frow.begin(), M->frow.end()), fcol(M->fcol.begin(), M->fcol.end()),
        fdia(M->fdiaind.begin(), M->fdiaind.end());
    DFM_TRY(M->d_frow.upload(frow));
    DFM_TRY(M->d_fcol.upload(fcol));
    DFM_TRY(M->d_fdiaind.upload(fdia));
    DFM_TRY(M->d_fsrc.upload(fsrc));
    DFM_TRY(M->d_fval.alloc(std::max<size_t>(1, nf * bb)));
    DFM_TRY(M->d_fdia.alloc(std::max<size_t>(1, (size_t)nb * bb)));
    DFM_TRY(M->d_rows_L.upload(rowsL));
    DFM_TRY(M->d_order.upload(M->order));
    DFM_TRY(M->d_inv.upload(M->inv));
    DFM_TRY(M->d_nsing.alloc(1));
    DFM_TRY(M->d_singrow.alloc(std::max<size_t>(1, nb)));
    DFM_TRY(M->d_fixrow.alloc(1));
    DFM_TRY(upload_sell_plan(M->planL, (int)bb, M->L64, M->d_srcL));
    DFM_TRY(upload_sell_plan(M->planU, (int)bb, M->U64, M->d_srcU));
    DFM_TRY(build_sweep_sync(M->planL, M->syncL));
    DFM_TRY(build_sweep_sync(M->planU, M->syncU));
    if (tiled) {
      DFM_TRY(upload_tile_plan(tpL, (int)b, M->tileL));
      DFM_TRY(upload_tile_plan(tpU, (int)b, M->tileU));
    }
    DFM_TRY(upload_sell_plan(M->planA, (int)bb, M->A64, M->d_srcA));
    DFM_TRY(M->state.alloc(1));
    DFM_HIP(hipHostMalloc((void**)&M->h_state, sizeof(KState), hipHostMallocDefault));
    DFM_HIP(hipMemsetAsync(M->state.p, 0, sizeof(KState), g_stream));
    DFM_HIP(hipStreamSynchronize(g_stream));
  }
  *out = M.release();
  return DFM_OK;
}
}  // namespace dfm

extern "C" {

int dfm_ilu_create(const dfm_bcrs* A, int lev_fill, int ordering, const int* order, unsigned block_nodes,
                   dfm_ilu** out) {
  return dfm::ilu_create_impl(A, lev_fill, ordering, order, block_nodes, nullptr, nullptr, out);
}

int dfm_ilu_create_fill(const dfm_bcrs* A, int lev_fill, const unsigned* fill_blk, unsigned nfill, dfm_ilu** out) {
  DFM_CHECK(nfill == 0 || fill_blk, DFM_BAD_ARG, "NULL fill block list");
  const std::vector<unsigned> blk(fill_blk, fill_blk + nfill);
  return dfm::ilu_create_impl(A, lev_fill, DFM_ORDER_NATURAL, nullptr, 0, nullptr, nullptr, out, &blk);
}

int dfm_ilu_destroy(dfm_ilu* M) {
  delete M;
  return DFM_OK;
}

int dfm_ilu_info(const dfm_ilu* M, unsigned* ncrs, int* nlevel_lower, int* nlevel_upper, int* ncolour) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (ncrs) *ncrs = (unsigned)M->fcol.size();
  if (nlevel_lower) *nlevel_lower = M->nlev_lower;
  if (nlevel_upper) *nlevel_upper = M->nlev_upper;
  if (ncolour) *ncolour = M->ncolour;
  return DFM_OK;
}

int dfm_ilu_tile_info(const dfm_ilu* M, int* nphase, int* ntile_lower, int* ntile_upper, int* tile_threads) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (nphase) *nphase = M->tileL.ntile ? (int)M->tileL.phase_tile.size() - 1 : 0;
  if (ntile_lower) *ntile_lower = M->tileL.ntile;
  if (ntile_upper) *ntile_upper = M->tileU.ntile;
  if (tile_threads) *tile_threads = std::max(M->tileL.nt, M->tileU.nt);
  return DFM_OK;
}

int dfm_ilu_get_pattern(const dfm_ilu* M, unsigned* rowptr, unsigned* col, int* lev, unsigned* diaind, int* order) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  if (rowptr) std::memcpy(rowptr, M->frow.data(), sizeof(unsigned) * (M->nb + 1));
  if (col && !M->fcol.empty()) std::memcpy(col, M->fcol.data(), sizeof(unsigned) * M->fcol.size());
  if (lev && !M->flev.empty()) std::memcpy(lev, M->flev.data(), sizeof(int) * M->flev.size());
  if (diaind) std::memcpy(diaind, M->fdiaind.data(), sizeof(unsigned) * M->nb);
  if (order) std::memcpy(order, M->order.data(), sizeof(int) * M->nb);
  return DFM_OK;
}

}  // extern "C"

// SetValue_Initialize: A's values into the factor pattern (fill zero)
int dfm::ilu_numeric_copy(dfm_ilu_s* M, const dfm_bcrs_s* A) { return ilu_numeric_copy_rows(M, A, 0); }

int dfm::ilu_numeric_copy_rows(dfm_ilu_s* M, const dfm_bcrs_s* A, int row0) {
  const int nb = (int)M->nb, b = (int)M->b, bb = b * b;
  if (row0 >= nb) return DFM_OK;
  const long q0 = (long)M->frow[row0];
  const long nf = (long)M->fcol.size() - q0;
  DFM_LAUNCH(k_ilu_copy, grid_for(nf * bb), nf, bb, M->d_fsrc.p + q0, A->d_val.p, M->d_fval.p + q0 * bb);
  DFM_LAUNCH(k_ilu_copy_dia, grid_for((long)(nb - row0) * bb), nb, row0, bb, M->permuted ? M->d_order.p : nullptr,
             A->d_dia.p, M->d_fdia.p);
  return DFM_OK;
}

extern "C" {

int dfm_ilu_set_value(dfm_ilu* M, dfm_bcrs* A, int* nsingular) {
  DFM_CHECK(M && A && A->nb == M->nb && A->b == M->b, DFM_BAD_ARG, "factor and matrix do not match");
  DFM_CHECK(A->pattern_hash == M->pattern_hash, DFM_BAD_ARG,
            "matrix pattern differs from the one the factor was built on");
  DFM_TRY(ilu_numeric_copy(M, A));
  return ilu_numeric_factor(M, A, nsingular);
}

}  // extern "C"

namespace {
// Rows of at most kIluWaveRow entries are factored one wave per row
// (DFM_ILU_WAVE=0: one thread per row, for comparisons).
bool g_ilu_wave = [] {
  const char* e = getenv("DFM_ILU_WAVE");
  return !(e && e[0] == '0');
}();
// one launch of the factor kernel over nrow rows listed at rows_dev, whose
// longest row has maxrow entries (-1: not known)
int ilu_factor_launch(dfm_ilu_s* M, int nrow, const int* rows_dev, int partial, int maxrow = -1) {
  const int b = (int)M->b;
  const int g = grid_for(nrow);
  if (g_ilu_wave && b <= 3 && maxrow >= 0 && maxrow <= kIluWaveRow) {
    const int gw = (nrow + kIluWaves - 1) / kIluWaves;
    switch (b) {
      case 1:
        DFM_LAUNCH_NT(k_ilu_factor_w<1>, gw, 64 * kIluWaves, nrow, rows_dev, M->d_frow.p, M->d_fcol.p,
                      M->d_fdiaind.p, M->d_fval.p, M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
        break;
      case 2:
        DFM_LAUNCH_NT(k_ilu_factor_w<2>, gw, 64 * kIluWaves, nrow, rows_dev, M->d_frow.p, M->d_fcol.p,
                      M->d_fdiaind.p, M->d_fval.p, M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
        break;
      default:
        DFM_LAUNCH_NT(k_ilu_factor_w<3>, gw, 64 * kIluWaves, nrow, rows_dev, M->d_frow.p, M->d_fcol.p,
                      M->d_fdiaind.p, M->d_fval.p, M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
    }
    return DFM_OK;
  }
  switch (b) {
    case 1:
      DFM_LAUNCH(k_ilu_factor<1>, g, nrow, rows_dev, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p, M->d_fval.p,
                 M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
      break;
    case 2:
      DFM_LAUNCH(k_ilu_factor<2>, g, nrow, rows_dev, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p, M->d_fval.p,
                 M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
      break;
    case 3:
      DFM_LAUNCH(k_ilu_factor<3>, g, nrow, rows_dev, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p, M->d_fval.p,
                 M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
      break;
    default:
      DFM_LAUNCH(k_ilu_factor_gen, g, nrow, b, rows_dev, M->d_frow.p, M->d_fcol.p, M->d_fdiaind.p, M->d_fval.p,
                 M->d_fdia.p, M->d_nsing.p, M->d_singrow.p, partial);
  }
  return DFM_OK;
}
}  // namespace

// DoILUDecomp on the values in the factor, then the sliced-ELL copies.  The
// levels run in parallel, so every row is factored and the singular pivots
// counted; a row depends only on rows before it, so when there are more than
// 10 the rows before the 11th singular one (in factor numbering) already hold
// what the reference's row-by-row loop leaves there, and the rest is put
// back: later rows restored to their copied values, the stopping row redone
// without its U scaling.
int dfm::ilu_numeric_factor(dfm_ilu_s* M, dfm_bcrs_s* A, int* nsingular, ilu_restore_fn restore, void* restore_ctx) {
  DFM_HIP(hipMemsetAsync(M->d_nsing.p, 0, sizeof(int), g_stream));
  for (int l = 0; l < M->nlev_lower; l++) {
    const int off = M->lev_ptr_L[l], nrow = M->lev_ptr_L[l + 1] - off;
    DFM_TRY(ilu_factor_launch(M, nrow, M->d_rows_L.p + off, 0, M->lev_maxrow_L[l]));
  }
  DFM_TRY(ilu_refresh_sell(M, A, M->have32, M->have16));
  int ns = 0;
  DFM_HIP(hipMemcpyAsync(&ns, M->d_nsing.p, sizeof(int), hipMemcpyDeviceToHost, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  if (nsingular) *nsingular = ns;
  M->factored = true;
  if (ns > 10) {
    std::vector<int> sing(M->nb);
    DFM_TRY(M->d_singrow.download(sing.data(), M->nb));
    int stop = -1, cnt = 0;
    for (unsigned i = 0; i < M->nb && stop < 0; i++)
      if (sing[i] && ++cnt == 11) stop = (int)i;
    DFM_CHECK(stop >= 0, DFM_HIP_ERROR, "singular pivot count and row flags disagree");
    if (restore) DFM_TRY(restore(restore_ctx, stop));
    else DFM_TRY(ilu_numeric_copy_rows(M, A, stop));
    DFM_TRY(M->d_fixrow.upload(&stop, 1));
    DFM_TRY(ilu_factor_launch(M, 1, M->d_fixrow.p, 1));
    DFM_TRY(ilu_refresh_sell(M, A, M->have32, M->have16));
    DFM_HIP(hipStreamSynchronize(g_stream));
    set_error("ILU factorisation met more than 10 singular pivots (stopped in factor row %d)", stop);
    return DFM_SINGULAR;
  }
  return DFM_OK;
}

extern "C" {

int dfm_ilu_get_values(const dfm_ilu* M, double* val, double* dia) {
  DFM_CHECK(M, DFM_BAD_ARG, "NULL factor");
  const size_t bb = (size_t)M->b * M->b;
  if (val && !M->fcol.empty())
    DFM_HIP(hipMemcpyAsync(val, M->d_fval.p, M->fcol.size() * bb * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  if (dia) DFM_HIP(hipMemcpyAsync(dia, M->d_fdia.p, M->nb * bb * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

int dfm_ilu_forward(dfm_ilu* M, double* vec) {
  DFM_CHECK(M && vec && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  DFM_TRY(ilu_sweep64(M, vec, true, false));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}
int dfm_ilu_backward(dfm_ilu* M, double* vec) {
  DFM_CHECK(M && vec && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  DFM_TRY(ilu_sweep64(M, vec, false, true));
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}
int dfm_ilu_solve_precond_noalloc(dfm_ilu* M, double* vec) {
  DFM_CHECK(M && vec && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  DFM_CHECK(g_inited, DFM_BAD_ARG, "dfm_init has not been called");
  const int nb = (int)M->nb, b = (int)M->b;
  const long n = (long)nb * b;
  if (!M->permuted) {
    DFM_TRY(ilu_sweep64(M, vec, true, true));
  } else {
    DFM_CHECK((size_t)n <= kScratchDoubles, DFM_TOO_LARGE,
              "ordered factor of %ld values: the allocation-free SolvePrecond holds at most %zu", n,
              kScratchDoubles);
    double* tmp = g_device < kMaxScratchDevices ? g_scratch[g_device] : nullptr;
    DFM_CHECK(tmp, DFM_BAD_ARG, "device %d has no scratch array", g_device);
    DFM_LAUNCH((k_permute<double, double>), grid_for(n), nb, b, M->d_order.p, vec, tmp, 1.0);
    DFM_TRY(ilu_sweep64(M, tmp, true, true));
    DFM_LAUNCH(k_unpermute<double>, grid_for(n), nb, b, M->d_order.p, tmp, vec);
  }
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

int dfm_ilu_solve_precond(dfm_ilu* M, double* vec) {
  DFM_CHECK(M && vec && M->factored, DFM_BAD_ARG, "factor has no values (call dfm_ilu_set_value)");
  const int nb = (int)M->nb, b = (int)M->b;
  const long n = (long)nb * b;
  if (!M->permuted) {
    DFM_TRY(ilu_sweep64(M, vec, true, true));
  } else {
    DevBuf<double> tmp;
    DFM_TRY(tmp.alloc(n));
    DFM_LAUNCH((k_permute<double, double>), grid_for(n), nb, b, M->d_order.p, vec, tmp.p, 1.0);
    DFM_TRY(ilu_sweep64(M, tmp.p, true, true));
    DFM_LAUNCH(k_unpermute<double>, grid_for(n), nb, b, M->d_order.p, tmp.p, vec);
    DFM_HIP(hipStreamSynchronize(g_stream));
  }
  DFM_HIP(hipStreamSynchronize(g_stream));
  return DFM_OK;
}

}  // extern "C"
