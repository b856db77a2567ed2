// host_check.cpp -- the library's host-side integer setup (dfm_host.cpp)
// built on its own, for runs under AddressSanitizer / UndefinedBehavior-
// Sanitizer on the CPU: FE pattern, orderings, ILU(k) symbolic, level sets,
// sliced-ELL plans, phases found in a natural-order factor and phase-tiled
// sweep plans, with the tile plans' layout invariants checked here and the
// patterns written out for comparison with the oracle.
//
// usage: host_check <in> <out>
//   in : int32 nnode, nelem, npe, lev, ordering (0 natural, 2 multicolour,
//        3 block multicolour), block_nodes, pack_rows, block size b;
//        uint32 conn[nelem*npe]
//   out: int32 status (0 ok), nnz, nf, ncolour, nlev_lower, nlev_upper,
//        tile threads L / U, tiles L / U, nphase (phases found, natural
//        order); uint32 rowptr[nnode+1], col[nnz]; int32 order[nnode];
//        uint32 frow[nnode+1], fcol[nf]; int32 flev[nf];
//        int32 phase_ptr[ncolour+1 or nphase+1]
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dfm_internal.h"

namespace dfm {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
}
}  // namespace dfm

using namespace dfm;

static int g_fail = 0;
#define EXPECT(c, ...)                        \
  do {                                        \
    if (!(c)) {                               \
      if (g_fail < 20) {                      \
        fprintf(stderr, "check failed: " #c); \
        fprintf(stderr, " -- " __VA_ARGS__);  \
        fputc('\n', stderr);                  \
      }                                       \
      g_fail++;                               \
    }                                         \
  } while (0)

// The layout promises of TilePlan, for one triangle.
static void check_tiles(const TilePlan& tp, unsigned nb, const std::vector<unsigned>& frow,
                        const std::vector<unsigned>& fcol, const std::vector<unsigned>& fdia,
                        const std::vector<int>& phase_ptr, bool upper) {
  const int np = (int)phase_ptr.size() - 1;
  EXPECT(tp.nt == 64 || tp.nt == 128 || tp.nt == 256, "nt %d", tp.nt);
  EXPECT((int)tp.phase_tile.size() == np + 1 && tp.phase_tile[0] == 0 && tp.phase_tile[np] == tp.ntile, "phases");
  EXPECT(tp.rows.size() == (size_t)tp.ntile * tp.nt && tp.rlev.size() == tp.rows.size(), "rows");
  const int nslice = tp.ntile * tp.nt / kSlice;
  EXPECT((int)tp.slice_ptr.size() == nslice + 1 && (int)tp.split.size() == nslice, "slices");
  EXPECT(tp.col.size() == (size_t)tp.nk * kSlice && tp.src.size() == tp.col.size(), "slots");
  if (g_fail) return;
  std::vector<int> phase_of(nb), seen(nb, 0);
  for (int p = 0; p < np; p++)
    for (int i = phase_ptr[p]; i < phase_ptr[p + 1]; i++) phase_of[i] = p;
  for (int pp = 0; pp < np; pp++) {
    const int p = upper ? np - 1 - pp : pp;
    for (int t = tp.phase_tile[pp]; t < tp.phase_tile[pp + 1]; t++) {
      int depth = 0, prev = -1;
      for (int r = 0; r < tp.nt; r++) {
        const int i = tp.rows[(size_t)t * tp.nt + r];
        if (i < 0) {
          prev = (int)nb;  // padding only after the tile's rows
          continue;
        }
        EXPECT(i < (int)nb && i > prev, "tile %d row %d: %d after %d", t, r, i, prev);
        if (i >= (int)nb) continue;
        prev = i;
        EXPECT(phase_of[i] == p, "tile %d of phase %d holds row %d of phase %d", t, p, i, phase_of[i]);
        seen[i]++;
        depth = std::max(depth, tp.rlev[(size_t)t * tp.nt + r] + 1);
      }
      EXPECT(tp.depth[t] == depth, "tile %d depth %d, rows say %d", t, tp.depth[t], depth);
    }
  }
  for (unsigned i = 0; i < nb; i++) EXPECT(seen[i] == 1, "row %u in %d tiles", i, seen[i]);
  // slots: the factor row's entries in order, first part then second part
  for (int s = 0; s < nslice; s++) {
    const int t = s / (tp.nt / kSlice);
    EXPECT(tp.slice_ptr[s] <= tp.split[s] && tp.split[s] <= tp.slice_ptr[s + 1], "slice %d bounds", s);
    for (int l = 0; l < kSlice; l++) {
      const int r = (s % (tp.nt / kSlice)) * kSlice + l;
      const int i = tp.rows[(size_t)t * tp.nt + r];
      const int lev = tp.rlev[(size_t)t * tp.nt + r];
      unsigned q = 0, qend = 0;
      if (i >= 0) {
        q = upper ? fdia[i] : frow[i];
        qend = upper ? frow[i + 1] : fdia[i];
      }
      for (int part = 0; part < 2; part++) {
        const int k0 = part ? tp.split[s] : tp.slice_ptr[s], k1 = part ? tp.slice_ptr[s + 1] : tp.split[s];
        const bool in_part = upper ? part == 0 : part == 1;
        bool pad = false;
        for (int k = k0; k < k1; k++) {
          const int c = tp.col[(size_t)k * kSlice + l];
          const long long src = tp.src[(size_t)k * kSlice + l];
          if (src < 0) {
            EXPECT(c == -1, "padding slot with column %d", c);
            pad = true;
            continue;
          }
          EXPECT(!pad && i >= 0, "slot after padding (slice %d lane %d)", s, l);
          EXPECT((unsigned)src == q && q < qend, "slice %d lane %d slot %d: entry %lld, expected %u", s, l, k, src, q);
          if (q >= qend) break;
          const int j = (int)fcol[q];
          if (in_part) {
            EXPECT(c >= 0 && c < tp.nt, "in-phase column %d", c);
            if (c >= 0 && c < tp.nt) {
              EXPECT(tp.rows[(size_t)t * tp.nt + c] == j, "tile position %d holds %d, not %d", c,
                     tp.rows[(size_t)t * tp.nt + c], j);
              EXPECT(tp.rlev[(size_t)t * tp.nt + c] < lev, "dependency level not below the row's");
            }
            EXPECT(phase_of[j] == phase_of[i], "in-phase slot of another phase");
          } else {
            EXPECT(c == j, "other-phase column %d, entry column %d", c, j);
            EXPECT(upper ? phase_of[j] > phase_of[i] : phase_of[j] < phase_of[i], "column %d not done before", j);
          }
          q++;
        }
      }
      EXPECT(q == qend, "slice %d lane %d: %u of the row's entries unplaced", s, l, qend - q);
    }
  }
  // overflow LDS: each row's slots beyond the register budgets, packed row
  // after row inside the largest tile's area
  EXPECT(tp.rov.size() == tp.rows.size() * 4, "overflow places");
  if (g_fail) return;
  for (int t = 0; t < tp.ntile; t++) {
    int oi = 0, oo = 0;
    for (int r = 0; r < tp.nt; r++) {
      const size_t rix = (size_t)t * tp.nt + r;
      const int i = tp.rows[rix];
      if (i < 0) continue;
      int nin = 0, nout = 0;
      const unsigned q0 = upper ? fdia[i] : frow[i], q1 = upper ? frow[i + 1] : fdia[i];
      for (unsigned q = q0; q < q1; q++) (phase_of[fcol[q]] == phase_of[i] ? nin : nout)++;
      const int ci = std::max(0, nin - kTileInSlots), co = upper ? std::max(0, nout - kTileOutSlots) : 0;
      EXPECT(tp.rov[rix * 4] == oi && tp.rov[rix * 4 + 1] == ci, "tile %d row %d in-phase overflow", t, r);
      EXPECT(tp.rov[rix * 4 + 2] == oo && tp.rov[rix * 4 + 3] == co, "tile %d row %d later-phase overflow", t, r);
      oi += ci;
      oo += co;
    }
    EXPECT(oi <= tp.ovf_in && oo <= tp.ovf_out, "tile %d overflow beyond the launch's area", t);
  }
}

// TileSplit: every live lane (tile row t, block row r < b) holds the row's
// first-part then second-part slots of the TilePlan, in order, with
// src = q*b + r; idle and padding lanes hold nothing.
static void check_split(const TilePlan& tp, int b) {
  TileSplit sp;
  build_tile_split(tp, b, sp);
  const int g = sp.g, nt = tp.nt, rpw = kSlice / g, wpt = nt * g / kSlice;
  EXPECT(g == (b == 2 ? 2 : 4) && sp.nslice == tp.ntile * wpt, "split shape");
  EXPECT((int)sp.slice_ptr.size() == sp.nslice + 1 && (int)sp.split.size() == sp.nslice, "split slices");
  EXPECT(sp.col.size() == (size_t)sp.nk * kSlice && sp.src.size() == sp.col.size(), "split slots");
  if (g_fail) return;
  for (int s = 0; s < sp.nslice; s++) {
    const int t = s / wpt, w = s % wpt;
    for (int l = 0; l < kSlice; l++) {
      const int j = l / g, r = l % g;
      const int trow = w * rpw + j;
      const size_t rix = (size_t)t * nt + trow;
      const int row = tp.rows[rix];
      const bool live = row >= 0 && r < b;
      EXPECT(sp.rows[(size_t)s * kSlice + l] == (live ? row * b + r : -1), "split rows slice %d lane %d", s, l);
      // the TilePlan's slots of this row
      const int s0 = (int)(rix / kSlice), l0 = (int)(rix % kSlice);
      for (int part = 0; part < 2; part++) {
        std::vector<std::pair<int, long long>> want;
        const int a0 = part ? tp.split[s0] : tp.slice_ptr[s0], a1 = part ? tp.slice_ptr[s0 + 1] : tp.split[s0];
        for (int k = a0; k < a1; k++)
          if (tp.src[(size_t)k * kSlice + l0] >= 0)
            want.push_back({tp.col[(size_t)k * kSlice + l0], tp.src[(size_t)k * kSlice + l0]});
        const int k0 = part ? sp.split[s] : sp.slice_ptr[s], k1 = part ? sp.slice_ptr[s + 1] : sp.split[s];
        size_t n = 0;
        for (int k = k0; k < k1; k++) {
          const long long q = sp.src[(size_t)k * kSlice + l];
          const int c = sp.col[(size_t)k * kSlice + l];
          if (q < 0) {
            EXPECT(c == -1, "split padding with a column");
            continue;
          }
          EXPECT(live && n < want.size() && c == want[n].first && q == want[n].second * b + r,
                 "split slice %d lane %d slot %d", s, l, k);
          n++;
        }
        EXPECT(n == (live ? want.size() : 0), "split slice %d lane %d: %zu of %zu slots", s, l, n, want.size());
      }
    }
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: host_check <in> <out>\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[8];
  if (fread(hdr, sizeof(int), 8, f) != 8) return 2;
  const unsigned nnode = hdr[0], nelem = hdr[1], npe = hdr[2];
  const int lev = hdr[3], ordering = hdr[4], block_nodes = hdr[5], pack_rows = hdr[6], bsize = hdr[7];
  std::vector<unsigned> conn((size_t)nelem * npe);
  if (fread(conn.data(), sizeof(unsigned), conn.size(), f) != conn.size()) return 2;
  fclose(f);

  std::vector<unsigned> rowptr, col;
  fem_pattern(nnode, nelem, npe, conn.data(), rowptr, col);
  std::vector<int> order, phase_ptr;
  int ncolour = 0;
  if (ordering == DFM_ORDER_MULTICOLOUR) {
    std::vector<int> colour;
    ncolour = greedy_colour(nnode, rowptr, col, colour);
    colour_order(nnode, colour, ncolour, order);
    phase_ptr.assign(ncolour + 1, 0);
    for (unsigned i = 0; i < nnode; i++) phase_ptr[colour[i] + 1]++;
    for (int c = 0; c < ncolour; c++) phase_ptr[c + 1] += phase_ptr[c];
  } else if (ordering == DFM_ORDER_BLOCK_MULTICOLOUR) {
    int nblock = 0;
    ncolour = bmc_order(nnode, rowptr, col, (unsigned)block_nodes, order, nblock, &phase_ptr);
  } else {
    order.resize(nnode);
    for (unsigned i = 0; i < nnode; i++) order[i] = (int)i;
  }
  std::vector<unsigned> frow, fcol, fdia;
  std::vector<int> flev;
  const int rc = iluk_symbolic(nnode, rowptr, col, lev, ordering ? &order : nullptr, frow, fcol, flev, fdia);
  if (rc != DFM_OK) return 3;
  std::vector<int> levL, levU;
  const int nlL = level_sets(nnode, frow, fcol, fdia, true, levL);
  const int nlU = level_sets(nnode, frow, fcol, fdia, false, levU);
  for (unsigned i = 0; i < nnode; i++) {
    for (unsigned q = frow[i]; q < fdia[i]; q++) EXPECT(levL[fcol[q]] < levL[i], "lower level order");
    for (unsigned q = fdia[i]; q < frow[i + 1]; q++) EXPECT(levU[fcol[q]] < levU[i], "upper level order");
  }
  // sliced-ELL plan of the lower sweep (rows by level)
  {
    std::vector<int> gptr(nlL + 1, 0), rows(nnode);
    for (unsigned i = 0; i < nnode; i++) gptr[levL[i] + 1]++;
    for (int l = 0; l < nlL; l++) gptr[l + 1] += gptr[l];
    std::vector<int> fill(gptr.begin(), gptr.end() - 1);
    for (unsigned i = 0; i < nnode; i++) rows[fill[levL[i]]++] = (int)i;
    std::vector<unsigned> beg(nnode), end(nnode);
    for (unsigned i = 0; i < nnode; i++) {
      beg[i] = frow[i];
      end[i] = fdia[i];
    }
    SellPlan plan;
    build_sell_plan(rows, gptr, beg, end, fcol, nullptr, plan);
    EXPECT(plan.col.size() == (size_t)plan.nk * kSlice, "sell plan size");
  }
  int ntL = 0, ntU = 0, tiL = 0, tiU = 0, nphase = 0;
  if (ordering == DFM_ORDER_NATURAL && discover_phases(nnode, frow, fcol, fdia, std::max(nlL, nlU), phase_ptr)) {
    nphase = (int)phase_ptr.size() - 1;
    EXPECT(nphase >= 1 && phase_ptr[0] == 0 && phase_ptr[nphase] == (int)nnode, "phase bounds");
    for (int p = 0; p < nphase; p++) EXPECT(phase_ptr[p] < phase_ptr[p + 1], "empty phase %d", p);
  }
  if (phase_ptr.size() >= 2) {
    TilePlan tL, tU;
    if (build_tile_plan(nnode, frow, fcol, fdia, phase_ptr, false, pack_rows, bsize, tL)) {
      check_tiles(tL, nnode, frow, fcol, fdia, phase_ptr, false);
      if (bsize == 2 || bsize == 3) check_split(tL, bsize);
      ntL = tL.nt;
      tiL = tL.ntile;
    }
    if (build_tile_plan(nnode, frow, fcol, fdia, phase_ptr, true, pack_rows, bsize, tU)) {
      check_tiles(tU, nnode, frow, fcol, fdia, phase_ptr, true);
      if (bsize == 2 || bsize == 3) check_split(tU, bsize);
      ntU = tU.nt;
      tiU = tU.ntile;
    }
  }
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  const int out_hdr[11] = {g_fail ? 1 : 0, (int)col.size(), (int)fcol.size(), ncolour, nlL, nlU, ntL, ntU, tiL, tiU,
                           nphase};
  fwrite(out_hdr, sizeof(int), 11, o);
  fwrite(rowptr.data(), sizeof(unsigned), rowptr.size(), o);
  fwrite(col.data(), sizeof(unsigned), col.size(), o);
  fwrite(order.data(), sizeof(int), order.size(), o);
  fwrite(frow.data(), sizeof(unsigned), frow.size(), o);
  fwrite(fcol.data(), sizeof(unsigned), fcol.size(), o);
  fwrite(flev.data(), sizeof(int), flev.size(), o);
  if (ncolour || nphase) fwrite(phase_ptr.data(), sizeof(int), phase_ptr.size(), o);
  fclose(o);
  if (g_fail) fprintf(stderr, "%d checks failed\n", g_fail);
  return g_fail ? 1 : 0;
}
