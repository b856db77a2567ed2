// dfm_host.cpp -- integer setup work done once per pattern on the host:
// FE adjacency pattern, ILU(k) symbolic factorisation, greedy colouring,
// level sets of the factor DAG and the sliced-ELL layout plans.
#include <algorithm>
#include <cstring>
#include <functional>

#include "dfm_internal.h"

namespace dfm {

// CElemAry::MakePattern_FEM (src/femfield/elem_ary.cpp:225-245, 675-892):
// neighbours of each node through shared elements, self excluded, ascending.
long fem_pattern(unsigned nnode, unsigned nelem, unsigned npe, const unsigned* conn,
                 std::vector<unsigned>& rowptr, std::vector<unsigned>& col) {
  std::vector<unsigned> eptr(nnode + 1, 0), elist((size_t)nelem * npe);
  for (size_t t = 0; t < (size_t)nelem * npe; t++) eptr[conn[t] + 1]++;
  for (unsigned i = 0; i < nnode; i++) eptr[i + 1] += eptr[i];
  {
    std::vector<unsigned> fill(eptr.begin(), eptr.end() - 1);
    for (unsigned e = 0; e < nelem; e++)
      for (unsigned k = 0; k < npe; k++) elist[fill[conn[(size_t)e * npe + k]]++] = e;
  }
  std::vector<int> mark(nnode, -1);
  rowptr.assign(nnode + 1, 0);
  col.clear();
  col.reserve((size_t)nnode * 8);
  for (unsigned i = 0; i < nnode; i++) {
    const size_t r0 = col.size();
    mark[i] = (int)i;
    for (unsigned q = eptr[i]; q < eptr[i + 1]; q++) {
      const unsigned* no = conn + (size_t)elist[q] * npe;
      for (unsigned k = 0; k < npe; k++) {
        const unsigned j = no[k];
        if (mark[j] != (int)i) {
          mark[j] = (int)i;
          col.push_back(j);
        }
      }
    }
    std::sort(col.begin() + r0, col.end());
    rowptr[i + 1] = (unsigned)col.size();
  }
  return (long)col.size();
}

// ILU(k) symbolic with DelFEM's level rule (src/matvec/matdiafrac_blkcrs.cpp:1249-1377):
// new fill gets max(lev_ik, lev_kj)+1, an entry keeps the first level it was
// given, k is taken from the growing sorted row in ascending order.
// With fill_blk (AddFracPtn(lev, aBlkFill), :1379-1516) the rows and columns
// listed there take fill of any level: fill (i, j) of level l is kept when
// l <= lev for both i and j, or when i or j is listed; operands are not
// filtered by their own levels.
int iluk_symbolic(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
                  int lev_fill, const std::vector<int>* order, std::vector<unsigned>& frow,
                  std::vector<unsigned>& fcol, std::vector<int>& flev, std::vector<unsigned>& fdiaind,
                  const std::vector<unsigned>* fill_blk) {
  const bool blk = fill_blk && !fill_blk->empty();
  std::vector<int> rowlev;  // per row: its fill limit, -1 unlimited
  if (blk) {
    rowlev.assign(nb, lev_fill);
    for (unsigned r : *fill_blk) {
      if (r >= nb) {
        set_error("fill block %u out of range (%u rows)", r, nb);
        return DFM_BAD_ARG;
      }
      rowlev[r] = -1;
    }
  }
  const unsigned ncrs0 = rowptr[nb];
  std::vector<unsigned> prow(nb + 1), pcol(ncrs0);
  if (order) {  // src/matvec/mat_blkcrs.cpp:493-553
    std::vector<int> inv(nb);
    for (unsigned i = 0; i < nb; i++) inv[(*order)[i]] = (int)i;
    prow[0] = 0;
    for (unsigned inew = 0; inew < nb; inew++) {
      const unsigned iold = (unsigned)(*order)[inew];
      const unsigned n0 = prow[inew], cnt = rowptr[iold + 1] - rowptr[iold];
      for (unsigned t = 0; t < cnt; t++) pcol[n0 + t] = (unsigned)inv[col[rowptr[iold] + t]];
      std::sort(pcol.begin() + n0, pcol.begin() + n0 + cnt);
      prow[inew + 1] = n0 + cnt;
    }
  } else {
    prow = rowptr;
    pcol = col;
    for (unsigned i = 0; i < nb; i++)
      for (unsigned q = prow[i] + 1; q < prow[i + 1]; q++)
        if (pcol[q] <= pcol[q - 1]) {
          set_error("ILU pattern row %u is not strictly ascending", i);
          return DFM_UNSORTED;
        }
  }
  const bool unchanged = ncrs0 == 0 || (unsigned long long)ncrs0 == (unsigned long long)nb * (nb - 1) ||
                         (lev_fill == 0 && !blk);
  fdiaind.assign(nb, 0);
  if (unchanged) {
    frow = prow;
    fcol = pcol;
    flev.assign(ncrs0, 0);
    for (unsigned i = 0; i < nb; i++) {
      fdiaind[i] = prow[i + 1];
      for (unsigned q = prow[i]; q < prow[i + 1]; q++)
        if (pcol[q] > i) {
          fdiaind[i] = q;
          break;
        }
    }
    return DFM_OK;
  }
  struct Node {
    unsigned c;
    int lev, next;
  };
  std::vector<Node> list;
  list.reserve(256);
  frow.assign(nb + 1, 0);
  fcol.clear();
  flev.clear();
  fcol.reserve((size_t)ncrs0 * 2);
  flev.reserve((size_t)ncrs0 * 2);
  for (unsigned i = 0; i < nb; i++) {
    list.clear();
    int head = -1;
    for (unsigned q = prow[i]; q < prow[i + 1]; q++) list.push_back({pcol[q], 0, (int)list.size() + 1});
    if (!list.empty()) {
      list.back().next = -1;
      head = 0;
    }
    for (int kc = head; kc != -1; kc = list[kc].next) {
      const unsigned k = list[kc].c;
      const int lik = list[kc].lev;
      if (!blk && lev_fill != -1 && lik + 1 > lev_fill) continue;
      if (k >= i) break;
      int cur = kc;
      for (unsigned q = fdiaind[k]; q < frow[k + 1]; q++) {
        const int lkj = flev[q];
        if (!blk && lev_fill != -1 && lkj + 1 > lev_fill) continue;
        const unsigned j = fcol[q];
        if (j == i) continue;
        if (blk) {
          const int l = std::max(lik, lkj) + 1, li = rowlev[i], lj = rowlev[j];
          if (li != -1 && lj != -1 && (l > li || l > lj)) continue;
        }
        for (;;) {
          const int nx = list[cur].next;
          if (nx == -1 || list[nx].c >= j) break;
          cur = nx;
        }
        const int nx = list[cur].next;
        if (nx != -1 && list[nx].c == j) continue;
        list.push_back({j, std::max(lik, lkj) + 1, nx});
        list[cur].next = (int)list.size() - 1;
        cur = (int)list.size() - 1;
      }
    }
    const unsigned n0 = (unsigned)fcol.size();
    fdiaind[i] = n0 + (unsigned)list.size();
    for (int t = head; t != -1; t = list[t].next) {
      if (list[t].c > i && fdiaind[i] == n0 + (unsigned)list.size()) fdiaind[i] = (unsigned)fcol.size();
      fcol.push_back(list[t].c);
      flev.push_back(list[t].lev);
    }
    frow[i + 1] = (unsigned)fcol.size();
  }
  return DFM_OK;
}

// Greedy colouring in smallest-last order (Matula-Beck): vertices are removed
// one at a time, always the one of least remaining degree (ties: lowest
// index), then coloured in reverse removal order with the smallest colour no
// already coloured neighbour uses.  On the Kuhn tetrahedral grid this finds
// the 4-colouring (natural-order greedy needs 8).
int greedy_colour(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
                  std::vector<int>& colour) {
  std::vector<int> deg(nb);
  int maxdeg = 0;
  for (unsigned i = 0; i < nb; i++) {
    deg[i] = (int)(rowptr[i + 1] - rowptr[i]);
    maxdeg = std::max(maxdeg, deg[i]);
  }
  // per-degree min-heaps of vertex ids with lazy deletion
  std::vector<std::vector<unsigned>> bucket(maxdeg + 1);
  auto push = [&](int d, unsigned v) {
    bucket[d].push_back(v);
    std::push_heap(bucket[d].begin(), bucket[d].end(), std::greater<unsigned>());
  };
  for (unsigned i = 0; i < nb; i++) push(deg[i], i);
  std::vector<char> removed(nb, 0);
  std::vector<unsigned> seq;
  seq.reserve(nb);
  int cur = 0;
  while (seq.size() < nb) {
    while (cur <= maxdeg) {
      auto& h = bucket[cur];
      while (!h.empty() && (removed[h.front()] || deg[h.front()] != cur)) {
        std::pop_heap(h.begin(), h.end(), std::greater<unsigned>());
        h.pop_back();
      }
      if (!h.empty()) break;
      cur++;
    }
    auto& h = bucket[cur];
    const unsigned v = h.front();
    std::pop_heap(h.begin(), h.end(), std::greater<unsigned>());
    h.pop_back();
    removed[v] = 1;
    seq.push_back(v);
    for (unsigned q = rowptr[v]; q < rowptr[v + 1]; q++) {
      const unsigned w = col[q];
      if (removed[w]) continue;
      deg[w]--;
      push(deg[w], w);
    }
    cur = std::max(0, cur - 1);
  }
  colour.assign(nb, -1);
  std::vector<int> stamp(nb + 2, -1);
  int ncol = 0;
  for (long t = (long)nb - 1; t >= 0; t--) {
    const unsigned i = seq[t];
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const int c = colour[col[q]];
      if (c >= 0) stamp[c] = (int)i;
    }
    int c = 0;
    while (stamp[c] == (int)i) c++;
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void colour_order(unsigned nb, const std::vector<int>& colour, int ncolour, std::vector<int>& order) {
  std::vector<unsigned> cnt(ncolour + 1, 0);
  for (unsigned i = 0; i < nb; i++) cnt[colour[i] + 1]++;
  for (int c = 0; c < ncolour; c++) cnt[c + 1] += cnt[c];
  order.assign(nb, 0);
  for (unsigned i = 0; i < nb; i++) order[cnt[colour[i]]++] = (int)i;
}

// Block multicolour ordering: nodes are aggregated into connected blocks of
// at most block_nodes nodes (breadth-first from the lowest unassigned node,
// neighbours in ascending order), the block adjacency graph is coloured with
// greedy_colour, and nodes are ordered by (block colour, block, node).  Rows
// inside a block keep their natural relative order, so the factor keeps most
// of the natural-order ILU quality while blocks of one colour are independent.
int bmc_order(unsigned nb, const std::vector<unsigned>& rowptr, const std::vector<unsigned>& col,
              unsigned block_nodes, std::vector<int>& order, int& nblock, std::vector<int>* colour_ptr) {
  std::vector<int> blk(nb, -1);
  std::vector<unsigned> queue;
  queue.reserve(block_nodes);
  int nbk = 0;
  for (unsigned seed = 0; seed < nb; seed++) {
    if (blk[seed] >= 0) continue;
    queue.clear();
    queue.push_back(seed);
    blk[seed] = nbk;
    for (size_t h = 0; h < queue.size() && queue.size() < block_nodes; h++) {
      const unsigned u = queue[h];
      for (unsigned q = rowptr[u]; q < rowptr[u + 1] && queue.size() < block_nodes; q++) {
        const unsigned v = col[q];
        if (blk[v] >= 0) continue;
        blk[v] = nbk;
        queue.push_back(v);
      }
    }
    nbk++;
  }
  // block adjacency graph, rows ascending, self excluded
  std::vector<std::vector<unsigned>> adj(nbk);
  for (unsigned u = 0; u < nb; u++)
    for (unsigned q = rowptr[u]; q < rowptr[u + 1]; q++)
      if (blk[col[q]] != blk[u]) adj[blk[u]].push_back((unsigned)blk[col[q]]);
  std::vector<unsigned> qrow(nbk + 1, 0), qcol;
  for (int i = 0; i < nbk; i++) {
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
    qrow[i + 1] = qrow[i] + (unsigned)adj[i].size();
  }
  qcol.reserve(qrow[nbk]);
  for (int i = 0; i < nbk; i++) qcol.insert(qcol.end(), adj[i].begin(), adj[i].end());
  std::vector<int> bcol;
  const int ncol = greedy_colour((unsigned)nbk, qrow, qcol, bcol);
  // counting sort by colour, then block, then node
  std::vector<unsigned> bptr(nbk + 1, 0);
  for (unsigned u = 0; u < nb; u++) bptr[blk[u] + 1]++;
  for (int i = 0; i < nbk; i++) bptr[i + 1] += bptr[i];
  std::vector<int> bynode(nb);
  {
    std::vector<unsigned> fill(bptr.begin(), bptr.end() - 1);
    for (unsigned u = 0; u < nb; u++) bynode[fill[blk[u]]++] = (int)u;
  }
  std::vector<int> blocks_by_colour;
  colour_order((unsigned)nbk, bcol, ncol, blocks_by_colour);
  order.clear();
  order.reserve(nb);
  for (int k : blocks_by_colour)
    for (unsigned t = bptr[k]; t < bptr[k + 1]; t++) order.push_back(bynode[t]);
  nblock = nbk;
  if (colour_ptr) {
    colour_ptr->assign(ncol + 1, 0);
    for (int k = 0; k < nbk; k++) (*colour_ptr)[bcol[k] + 1] += (int)(bptr[k + 1] - bptr[k]);
    for (int c = 0; c < ncol; c++) (*colour_ptr)[c + 1] += (*colour_ptr)[c];
  }
  return ncol;
}

// Level sets of the factor DAG: lower -> longest chain through L(i),
// upper -> longest chain through U(i) taken from the last row.
int level_sets(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
               const std::vector<unsigned>& fdiaind, bool lower, std::vector<int>& level) {
  level.assign(nb, 0);
  int nlev = 0;
  if (lower) {
    for (unsigned i = 0; i < nb; i++) {
      int l = 0;
      for (unsigned q = frow[i]; q < fdiaind[i]; q++) l = std::max(l, level[fcol[q]] + 1);
      level[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (long ii = (long)nb - 1; ii >= 0; ii--) {
      int l = 0;
      for (unsigned q = fdiaind[ii]; q < frow[ii + 1]; q++) l = std::max(l, level[fcol[q]] + 1);
      level[ii] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  return nlev;
}

void build_sell_plan(const std::vector<int>& rows_by_group, const std::vector<int>& group_ptr,
                     const std::vector<unsigned>& beg, const std::vector<unsigned>& end,
                     const std::vector<unsigned>& colv, const std::vector<long long>* srcv,
                     SellPlan& plan) {
  const int ngroup = (int)group_ptr.size() - 1;
  plan.group_slice_ptr.assign(ngroup + 1, 0);
  plan.slice_ptr.assign(1, 0);
  plan.rows.clear();
  int nslice = 0;
  long nk = 0;
  for (int g = 0; g < ngroup; g++) {
    for (int r0 = group_ptr[g]; r0 < group_ptr[g + 1]; r0 += kSlice) {
      int w = 0;
      for (int l = 0; l < kSlice; l++) {
        const int t = r0 + l;
        if (t < group_ptr[g + 1]) {
          const int row = rows_by_group[t];
          plan.rows.push_back(row);
          w = std::max(w, (int)(end[row] - beg[row]));
        } else {
          plan.rows.push_back(-1);
        }
      }
      nk += w;
      plan.slice_ptr.push_back((int)nk);
      nslice++;
    }
    plan.group_slice_ptr[g + 1] = nslice;
  }
  plan.nslice = nslice;
  plan.nk = nk;
  plan.col.assign((size_t)nk * kSlice, -1);
  plan.src.assign((size_t)nk * kSlice, -1);
  for (int s = 0; s < nslice; s++) {
    for (int l = 0; l < kSlice; l++) {
      const int row = plan.rows[(size_t)s * kSlice + l];
      if (row < 0) continue;
      long k = plan.slice_ptr[s];
      for (unsigned q = beg[row]; q < end[row]; q++, k++) {
        plan.col[(size_t)k * kSlice + l] = (int)colv[q];
        plan.src[(size_t)k * kSlice + l] = srcv ? (*srcv)[q] : (long long)q;
      }
    }
  }
}

// Phase-tiled sweep plan (TilePlan).  Components by union-find over the
// in-phase entries, local levels in sweep order, components packed into tiles
// in the order of their first rows, rows of a tile in factor order.
bool build_tile_plan(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
                     const std::vector<unsigned>& fdiaind, const std::vector<int>& phase_ptr, bool upper,
                     int pack_rows, int b, TilePlan& tp) {
  tp = TilePlan();
  const int np = (int)phase_ptr.size() - 1;
  if (np < 1 || phase_ptr[0] != 0 || phase_ptr[np] != (int)nb) return false;
  for (int p = 0; p < np; p++)
    if (phase_ptr[p + 1] < phase_ptr[p]) return false;
  auto beg = [&](unsigned i) { return upper ? fdiaind[i] : frow[i]; };
  auto end = [&](unsigned i) { return upper ? frow[i + 1] : fdiaind[i]; };
  std::vector<int> phase_of(nb);
  for (int p = 0; p < np; p++)
    for (int i = phase_ptr[p]; i < phase_ptr[p + 1]; i++) phase_of[i] = p;
  // components and local levels, phase by phase
  std::vector<int> parent(nb), level(nb, 0), comp_size(nb, 0);
  std::function<int(int)> find = [&](int x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (unsigned i = 0; i < nb; i++) parent[i] = (int)i;
  int maxcomp = 0;
  for (int p = 0; p < np; p++) {
    const int a = phase_ptr[p], b = phase_ptr[p + 1];
    for (int i = a; i < b; i++)
      for (unsigned q = beg(i); q < end(i); q++) {
        const int j = (int)fcol[q];
        if (j < a || j >= b) continue;
        const int ri = find(i), rj = find(j);
        if (ri != rj) parent[std::max(ri, rj)] = std::min(ri, rj);  // root = smallest row
      }
    for (int ii = 0; ii < b - a; ii++) {
      const int i = upper ? b - 1 - ii : a + ii;
      int l = 0;
      for (unsigned q = beg(i); q < end(i); q++) {
        const int j = (int)fcol[q];
        if (j >= a && j < b) l = std::max(l, level[j] + 1);
      }
      level[i] = l;
    }
    for (int i = a; i < b; i++) maxcomp = std::max(maxcomp, ++comp_size[find(i)]);
  }
  if (maxcomp > kTileMaxRows) return false;
  const int nt = maxcomp <= 64 ? 64 : (maxcomp <= 128 ? 128 : 256);
  const int cap = std::min(nt, std::max(1, pack_rows));
  tp.nt = nt;
  // tiles, phases in launch order
  std::vector<std::vector<int>> members(nb);
  std::vector<int> tile_rows;  // concatenated rows of all tiles, nt per tile
  tp.phase_tile.assign(1, 0);
  std::vector<int> cur;
  auto flush = [&]() {
    if (cur.empty()) return;
    std::sort(cur.begin(), cur.end());
    for (int t = 0; t < nt; t++) tile_rows.push_back(t < (int)cur.size() ? cur[t] : -1);
    tp.ntile++;
    cur.clear();
  };
  for (int pp = 0; pp < np; pp++) {
    const int p = upper ? np - 1 - pp : pp;
    const int a = phase_ptr[p], b = phase_ptr[p + 1];
    for (int i = a; i < b; i++) members[find(i)].push_back(i);
    for (int i = a; i < b; i++) {
      if (members[i].empty()) continue;  // not a root
      const int sz = (int)members[i].size();
      if (!cur.empty() && (int)cur.size() + sz > cap) flush();
      cur.insert(cur.end(), members[i].begin(), members[i].end());
      if ((int)cur.size() >= cap) flush();
      std::vector<int>().swap(members[i]);
    }
    flush();
    tp.phase_tile.push_back(tp.ntile);
  }
  // per tile row: level; per slice: parts and widths
  const int nslice = tp.ntile * nt / kSlice;
  tp.rows = tile_rows;
  tp.rlev.assign(tile_rows.size(), -1);
  tp.depth.assign(tp.ntile, 0);
  tp.slice_ptr.assign(nslice + 1, 0);
  tp.split.assign(nslice, 0);
  std::vector<int> nfirst(tile_rows.size(), 0), nsecond(tile_rows.size(), 0);
  // in-phase columns: lower rows -> the columns at or after the phase start;
  // upper rows -> the columns before the phase end
  auto in_phase = [&](int i, int j) { return phase_of[j] == phase_of[i]; };
  for (size_t r = 0; r < tile_rows.size(); r++) {
    const int i = tile_rows[r];
    if (i < 0) continue;
    tp.rlev[r] = level[i];
    int nin = 0, nout = 0;
    for (unsigned q = beg(i); q < end(i); q++) (in_phase(i, (int)fcol[q]) ? nin : nout)++;
    nfirst[r] = upper ? nin : nout;
    nsecond[r] = upper ? nout : nin;
    const int t = (int)(r / nt);
    tp.depth[t] = std::max(tp.depth[t], level[i] + 1);
  }
  for (int s = 0; s < nslice; s++) {
    int wa = 0, wb = 0;
    for (int l = 0; l < kSlice; l++) {
      wa = std::max(wa, nfirst[(size_t)s * kSlice + l]);
      wb = std::max(wb, nsecond[(size_t)s * kSlice + l]);
    }
    tp.split[s] = tp.slice_ptr[s] + wa;
    tp.slice_ptr[s + 1] = tp.split[s] + wb;
  }
  tp.nk = tp.slice_ptr[nslice];
  // overflow LDS: per tile row, its in-phase slots beyond kTileInSlots and
  // (upper) its later-phase slots beyond kTileOutSlots, packed row after row
  tp.rov.assign(tile_rows.size() * 4, 0);
  for (int t = 0; t < tp.ntile; t++) {
    int oi = 0, oo = 0;
    for (int r = 0; r < nt; r++) {
      const size_t rix = (size_t)t * nt + r;
      if (tile_rows[rix] < 0) continue;
      const int nin = upper ? nfirst[rix] : nsecond[rix], nout = upper ? nsecond[rix] : 0;
      const int ci = std::max(0, nin - kTileInSlots), co = std::max(0, nout - kTileOutSlots);
      tp.rov[rix * 4 + 0] = oi;
      tp.rov[rix * 4 + 1] = ci;
      tp.rov[rix * 4 + 2] = oo;
      tp.rov[rix * 4 + 3] = co;
      oi += ci;
      oo += co;
    }
    tp.ovf_in = std::max(tp.ovf_in, oi);
    tp.ovf_out = std::max(tp.ovf_out, oo);
  }
  const long bytes = (long)tp.ovf_in * (b * b * 8 + 4) + (long)tp.ovf_out * b * 8;
  if (bytes > kTileOverflowBytes) {
    tp = TilePlan();
    return false;
  }
  tp.col.assign((size_t)tp.nk * kSlice, -1);
  tp.src.assign((size_t)tp.nk * kSlice, -1);
  std::vector<int> local(nb, -1);
  for (int t = 0; t < tp.ntile; t++) {
    for (int r = 0; r < nt; r++) {
      const int i = tile_rows[(size_t)t * nt + r];
      if (i >= 0) local[i] = r;
    }
    for (int r = 0; r < nt; r++) {
      const int i = tile_rows[(size_t)t * nt + r];
      if (i < 0) continue;
      const int s = (t * nt + r) / kSlice, l = r % kSlice;
      int ka = tp.slice_ptr[s], kb = tp.split[s];
      for (unsigned q = beg(i); q < end(i); q++) {
        const int j = (int)fcol[q];
        const bool in = in_phase(i, j);
        const bool first = upper ? in : !in;
        int& k = first ? ka : kb;
        tp.col[(size_t)k * kSlice + l] = in ? local[j] : j;
        tp.src[(size_t)k * kSlice + l] = (long long)q;
        k++;
      }
    }
  }
  return true;
}

// TileSplit from a TilePlan: each tile row's slots, read back in order from
// the row's lane of the TilePlan layout, laid out again for g lanes per row.
void build_tile_split(const TilePlan& tp, int b, TileSplit& ts) {
  ts = TileSplit();
  const int g = b <= 2 ? 2 : 4;
  const int nt = tp.nt, rpw = kSlice / g;  // tile rows per wave
  ts.g = g;
  ts.b = b;
  ts.nslice = tp.ntile * nt * g / kSlice;
  // per tile row: its first- and second-part slots (col, factor entry), in order
  const size_t nrow = tp.rows.size();
  std::vector<std::vector<std::pair<int, long long>>> first(nrow), second(nrow);
  const int nslice0 = tp.ntile * nt / kSlice;
  for (int s = 0; s < nslice0; s++)
    for (int l = 0; l < kSlice; l++) {
      const size_t rix = (size_t)s * kSlice + l;
      if (tp.rows[rix] < 0) continue;
      for (int k = tp.slice_ptr[s]; k < tp.slice_ptr[s + 1]; k++) {
        const long long q = tp.src[(size_t)k * kSlice + l];
        if (q < 0) continue;
        (k < tp.split[s] ? first : second)[rix].push_back({tp.col[(size_t)k * kSlice + l], q});
      }
    }
  ts.slice_ptr.assign(ts.nslice + 1, 0);
  ts.split.assign(ts.nslice, 0);
  for (int s = 0; s < ts.nslice; s++) {
    // rows of this wave: tile t, tile rows [w*rpw, (w+1)*rpw)
    const int waves_per_tile = nt * g / kSlice;
    const int t = s / waves_per_tile, w = s % waves_per_tile;
    int wa = 0, wb = 0;
    for (int j = 0; j < rpw; j++) {
      const size_t rix = (size_t)t * nt + w * rpw + j;
      wa = std::max(wa, (int)first[rix].size());
      wb = std::max(wb, (int)second[rix].size());
    }
    ts.split[s] = ts.slice_ptr[s] + wa;
    ts.slice_ptr[s + 1] = ts.split[s] + wb;
  }
  ts.nk = ts.slice_ptr[ts.nslice];
  ts.col.assign((size_t)ts.nk * kSlice, -1);
  ts.src.assign((size_t)ts.nk * kSlice, -1);
  ts.rows.assign((size_t)ts.nslice * kSlice, -1);
  for (int s = 0; s < ts.nslice; s++) {
    const int waves_per_tile = nt * g / kSlice;
    const int t = s / waves_per_tile, w = s % waves_per_tile;
    for (int l = 0; l < kSlice; l++) {
      const int j = l / g, r = l % g;
      const size_t rix = (size_t)t * nt + w * rpw + j;
      const int row = tp.rows[rix];
      if (row < 0 || r >= b) continue;
      ts.rows[(size_t)s * kSlice + l] = row * b + r;
      for (int part = 0; part < 2; part++) {
        const auto& v = part ? second[rix] : first[rix];
        int k = part ? ts.split[s] : ts.slice_ptr[s];
        for (const auto& cq : v) {
          ts.col[(size_t)k * kSlice + l] = cq.first;
          ts.src[(size_t)k * kSlice + l] = cq.second * b + r;
          k++;
        }
      }
    }
  }
}


// Phases for the tiled sweeps of a factor without a colour ordering.  Rows are
// scanned in factor order, each joined by union-find to the rows of the
// current phase it depends on (lower entries of the row and upper entries of
// earlier rows that name it, so that one partition serves both sweeps); a new
// phase starts at the row that would make a component exceed kTileMaxRows
// rows or whose local level would reach the depth cap.  Caps are tried from
// none down to 16 and the cheapest estimate kept: per phase a launch's fixed
// cost and its deepest local level (from the measured tiled sweeps), against
// one launch per level set.  Coloured and tile-numbered meshes give a few
// phases; lexicographic numberings give long chains and keep the level sets.
bool discover_phases(unsigned nb, const std::vector<unsigned>& frow, const std::vector<unsigned>& fcol,
                     const std::vector<unsigned>& fdiaind, int nlev, std::vector<int>& phase_ptr) {
  constexpr double kPhaseUs = 3.0, kTileLevelUs = 0.45, kLevelSetUs = 5.0;
  phase_ptr.clear();
  if (nb < 2 || nlev < 8) return false;
  // earlier rows each row depends on or is depended on by
  std::vector<unsigned> aptr(nb + 1, 0);
  for (unsigned i = 0; i < nb; i++) {
    aptr[i + 1] += fdiaind[i] - frow[i];
    for (unsigned q = fdiaind[i]; q < frow[i + 1]; q++) aptr[fcol[q] + 1]++;
  }
  for (unsigned i = 0; i < nb; i++) aptr[i + 1] += aptr[i];
  std::vector<unsigned> adj(aptr[nb]), fillp(aptr.begin(), aptr.end() - 1);
  for (unsigned i = 0; i < nb; i++) {
    for (unsigned q = frow[i]; q < fdiaind[i]; q++) adj[fillp[i]++] = fcol[q];
    for (unsigned q = fdiaind[i]; q < frow[i + 1]; q++) adj[fillp[fcol[q]]++] = i;
  }
  std::vector<int> parent(nb), size(nb), lvl(nb), mark(nb);
  auto find = [&](int x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  double best = 0.5 * nlev * kLevelSetUs;  // tiles must at least halve the estimate
  bool found = false;
  std::vector<int> ptr;
  for (int dcap : {0, 32, 24, 16}) {
    for (unsigned i = 0; i < nb; i++) {
      parent[i] = (int)i;
      size[i] = 1;
      mark[i] = -1;
    }
    ptr.assign(1, 0);
    unsigned p0 = 0;
    int depth = 0;
    double cost = 0.0;
    for (unsigned i = 0; i < nb && cost < best; i++) {
      int tot = 1, l = 0;
      for (unsigned q = aptr[i]; q < aptr[i + 1]; q++) {
        const unsigned j = adj[q];
        if (j < p0) continue;
        l = std::max(l, lvl[j] + 1);
        const int r = find((int)j);
        if (mark[r] != (int)i) {
          mark[r] = (int)i;
          tot += size[r];
        }
      }
      if (tot > kTileMaxRows || (dcap > 0 && l >= dcap)) {
        cost += kPhaseUs + kTileLevelUs * depth;
        ptr.push_back((int)i);
        p0 = i;
        depth = 0;
        l = 0;
      } else {
        for (unsigned q = aptr[i]; q < aptr[i + 1]; q++) {
          const unsigned j = adj[q];
          if (j < p0) continue;
          const int a = find((int)i), r = find((int)j);
          if (a == r) continue;
          const int big = size[a] >= size[r] ? a : r, small = big == a ? r : a;
          parent[small] = big;
          size[big] += size[small];
        }
      }
      lvl[i] = l;
      depth = std::max(depth, l + 1);
    }
    cost += kPhaseUs + kTileLevelUs * depth;
    if (cost < best) {
      best = cost;
      ptr.push_back((int)nb);
      phase_ptr = ptr;
      found = true;
    }
  }
  return found;
}

}  // namespace dfm
