// Drives the C++ LsSol classes (include/delfem_amd/ls.hpp) with several
// segments the way DelFEM's split-field equation systems do
// (src/femeqn/eqnsys_fluid.cpp:411-436 with src/ls/linearsystem.cpp):
// AddLinSysSeg per segment, AddMat_Dia / AddMat_NonDia patterns,
// InitializeMarge, one Mearge per element and block, residual and Dirichlet
// flags, FinalizeMarge, CPreconditioner_ILU(ls, lev) across the segments,
// then LsSol::Solve_PCG or Solve_PBiCGSTAB.  Input and output are raw binary
// files written / read by tests/test_gpu_shim.py.
//
// input : i32 nnode nelem npe nseg kind maxit lev | f64 tol | i32 len[nseg] |
//         u32 conn[nelem*npe] | u32 rowptr[nnode+1] col[] (diagonal blocks) |
//         u32 rowptr_nd[nnode+1] col_nd[] (off-diagonal blocks) |
//         f64 emat(i,j)[nelem*npe*npe*len_i*len_j] for (i, j) row by row |
//         f64 res[sum nnode*len_i] | i32 flags[sum nnode*len_i]
//         kind: 0 Solve_PCG, 1 Solve_PBiCGSTAB
// output: f64 norm_res conv | i32 ok iteration | f64 update[sum nnode*len_i] |
//         per block (i, j) row by row: f64 values (diagonal blocks: CRS values then diagonal)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "delfem_amd/ls.hpp"

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}
template <typename T>
static void wr(FILE* f, const T* p, size_t n) {
  if (n && fwrite(p, sizeof(T), n, f) != n) exit(3);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: msshim_test IN OUT\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[7];
  rd(f, hdr, 7);
  const unsigned nnode = hdr[0], nelem = hdr[1], npe = hdr[2], nseg = hdr[3];
  const int kind = hdr[4], lev = hdr[6];
  unsigned maxit = (unsigned)hdr[5];
  if (nseg == 0 || nseg > 8) return 2;
  double tol;
  rd(f, &tol, 1);
  std::vector<int> len(nseg);
  rd(f, len.data(), nseg);
  std::vector<unsigned> conn((size_t)nelem * npe), rp(nnode + 1), rp2(nnode + 1);
  rd(f, conn.data(), conn.size());
  rd(f, rp.data(), rp.size());
  std::vector<unsigned> cl(rp[nnode]);
  rd(f, cl.data(), cl.size());
  rd(f, rp2.data(), rp2.size());
  std::vector<unsigned> cl2(rp2[nnode]);
  rd(f, cl2.data(), cl2.size());
  std::vector<std::vector<double>> emat(nseg * nseg);
  size_t total = 0;
  for (unsigned i = 0; i < nseg; i++) total += (size_t)nnode * len[i];
  for (unsigned i = 0; i < nseg; i++)
    for (unsigned j = 0; j < nseg; j++) {
      emat[i * nseg + j].resize((size_t)nelem * npe * npe * len[i] * len[j]);
      rd(f, emat[i * nseg + j].data(), emat[i * nseg + j].size());
    }
  std::vector<double> res(total);
  std::vector<int> flags(total);
  rd(f, res.data(), total);
  rd(f, flags.data(), total);
  fclose(f);
  if (dfm_init(0) != DFM_OK) {
    fprintf(stderr, "dfm_init: %s\n", dfm_last_error());
    return 4;
  }
  try {
    LsSol::CLinearSystem ls;
    for (unsigned i = 0; i < nseg; i++) ls.AddLinSysSeg(nnode, len[i]);
    for (unsigned i = 0; i < nseg; i++)
      for (unsigned j = 0; j < nseg; j++) {
        const bool ok = i == j ? ls.AddMat_Dia(i, rp, cl) : ls.AddMat_NonDia(i, j, rp2, cl2);
        if (!ok) return 5;
      }
    ls.InitializeMarge();
    for (unsigned i = 0; i < nseg; i++)
      for (unsigned j = 0; j < nseg; j++) {
        const size_t bs = (size_t)len[i] * len[j], esz = (size_t)npe * npe * bs;
        for (unsigned e = 0; e < nelem; e++) {
          const unsigned* no = conn.data() + (size_t)e * npe;
          const double* em = emat[i * nseg + j].data() + e * esz;
          const bool ok = i == j ? ls.GetMatrix(i).Mearge(npe, no, npe, no, (unsigned)bs, em)
                                 : ls.GetMatrix(i, j).Mearge(npe, no, npe, no, (unsigned)bs, em);
          if (!ok) return 6;
        }
      }
    size_t off = 0;
    for (unsigned i = 0; i < nseg; i++) {
      ls.GetResidual(i).SetValues(res.data() + off);
      for (unsigned k = 0; k < nnode; k++)
        for (int d = 0; d < len[i]; d++)
          if (flags[off + (size_t)k * len[i] + d]) ls.GetBCFlag(i).SetBC(k, d);
      off += (size_t)nnode * len[i];
    }
    const double norm = ls.FinalizeMarge();
    // lev -2: a per-segment request, level 1 for the last segment's block
    // row and column, as a caller of SetFillInLevel(lev, ilss) writes it
    LsSol::CPreconditioner_ILU prec;
    if (lev == -2) prec.SetFillInLevel(1, (int)nseg - 1);
    else prec.SetFillInLevel(lev);
    prec.SetLinearSystem(ls);
    if (!prec.SetValue(ls)) return 7;
    LsSol::CLinearSystemPreconditioner lsp(ls, prec);
    double conv = tol;
    const bool ok = kind == 0 ? LsSol::Solve_PCG(conv, maxit, lsp) : LsSol::Solve_PBiCGSTAB(conv, maxit, lsp);
    const std::vector<double> u = ls.vec(-2).GetValues();
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    const double d2[2] = {norm, conv};
    const int i2[2] = {ok ? 1 : 0, (int)maxit};
    wr(o, d2, 2);
    wr(o, i2, 2);
    wr(o, u.data(), u.size());
    for (unsigned i = 0; i < nseg; i++)
      for (unsigned j = 0; j < nseg; j++) {
        std::vector<double> val, dia;
        if (i == j) {
          ls.GetMatrix(i).GetValues(val, dia);
          wr(o, val.data(), val.size());
          wr(o, dia.data(), dia.size());
        } else {
          ls.GetMatrix(i, j).GetValues(val);
          wr(o, val.data(), val.size());
        }
      }
    fclose(o);
  } catch (const delfem_amd::Error& e) {
    fprintf(stderr, "error: %s (%d): %s\n", e.what(), e.code, dfm_last_error());
    return 8;
  }
  return 0;
}
