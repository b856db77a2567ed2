// Drives the C++ drop-in classes (include/delfem_amd/{matvec,ls}.hpp) the way
// DelFEM's equation systems do: pattern from the element list, InitializeMarge,
// one Mearge per element, residual, Dirichlet flags, FinalizeMarge, ILU(0),
// Solve_PCG or Solve_PBiCGSTAB.  Input and output are raw binary files
// written / read by tests/test_gpu_shim.py.
//
// input : i32 nnode nelem len npe kind(0 PCG, 1 BiCGSTAB) maxit | f64 tol |
//         u32 conn[nelem*npe] | f64 emat[nelem*npe*npe*len*len] |
//         f64 eres[nelem*npe*len] | i32 flags[nnode*len]
// output: f64 norm_res conv | i32 ok iteration | f64 val[ncrs*len*len]
//         dia[nnode*len*len] res[ndof] x[ndof]
#include <cstdio>
#include <cstdlib>
#include <set>
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
    fprintf(stderr, "usage: shim_test IN OUT\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[6];
  rd(f, hdr, 6);
  const unsigned nnode = hdr[0], nelem = hdr[1], len = hdr[2], npe = hdr[3];
  const int kind = hdr[4];
  unsigned maxit = (unsigned)hdr[5];
  double tol;
  rd(f, &tol, 1);
  std::vector<unsigned> conn((size_t)nelem * npe);
  std::vector<double> emat((size_t)nelem * npe * npe * len * len), eres((size_t)nelem * npe * len);
  std::vector<int> flags((size_t)nnode * len);
  rd(f, conn.data(), conn.size());
  rd(f, emat.data(), emat.size());
  rd(f, eres.data(), eres.size());
  rd(f, flags.data(), flags.size());
  fclose(f);
  if (dfm_init(0) != DFM_OK) {
    fprintf(stderr, "dfm_init: %s\n", dfm_last_error());
    return 1;
  }
  try {
    // node adjacency (CElemAry::MakePattern_FEM)
    std::vector<std::set<unsigned>> adj(nnode);
    for (unsigned e = 0; e < nelem; e++)
      for (unsigned a = 0; a < npe; a++)
        for (unsigned b = 0; b < npe; b++)
          if (a != b) adj[conn[e * npe + a]].insert(conn[e * npe + b]);
    std::vector<unsigned> index(nnode + 1, 0), array;
    for (unsigned i = 0; i < nnode; i++) {
      array.insert(array.end(), adj[i].begin(), adj[i].end());
      index[i + 1] = (unsigned)array.size();
    }
    LsSol::CLinearSystem ls;
    const int ilss = ls.AddLinSysSeg(nnode, len);
    if (ilss != 0 || !ls.AddMat_Dia(0, index, array)) return 1;
    ls.InitializeMarge();
    std::vector<double> res((size_t)nnode * len, 0.0);
    MatVec::CMatDia_BlkCrs& A = ls.GetMatrix(0);
    const size_t esz = (size_t)npe * npe * len * len;
    for (unsigned e = 0; e < nelem; e++) {
      const unsigned* no = &conn[(size_t)e * npe];
      if (!A.Mearge(npe, no, npe, no, len * len, &emat[e * esz])) return 1;
      for (unsigned a = 0; a < npe; a++)
        for (unsigned r = 0; r < len; r++) res[(size_t)no[a] * len + r] += eres[((size_t)e * npe + a) * len + r];
    }
    ls.GetResidual(0).SetValues(res.data());
    MatVec::CBCFlag& bc = ls.GetBCFlag(0);
    for (unsigned i = 0; i < nnode; i++)
      for (unsigned r = 0; r < len; r++)
        if (flags[(size_t)i * len + r]) bc.SetBC(i, r);
    const double norm_res = ls.FinalizeMarge();
    std::vector<double> val, dia;
    A.GetValues(val, dia);
    const std::vector<double> rfin = ls.GetResidual(0).GetValues();
    LsSol::CPreconditioner_ILU prec(ls, 0);
    if (!prec.SetValue(ls)) return 1;
    LsSol::CLinearSystemPreconditioner lsp(ls, prec);
    double conv = tol;
    const bool ok = kind == 0 ? LsSol::Solve_PCG(conv, maxit, lsp) : LsSol::Solve_PBiCGSTAB(conv, maxit, lsp);
    const std::vector<double> x = ls.GetUpdate(0).GetValues();
    // the API's own check: ||b - A x|| with MATVEC on a scratch vector
    ls.ReSizeTmpVecSolver(1);
    ls.GetTmpVectorArySize();
    ls.vec(0).SetValues(rfin.data());
    ls.MATVEC(-1.0, -2, 1.0, 0);
    const double rr = ls.DOT(0, 0);
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    const double d2[2] = {norm_res, conv};
    const int i2[2] = {ok ? 1 : 0, (int)maxit};
    wr(o, d2, 2);
    wr(o, i2, 2);
    wr(o, val.data(), val.size());
    wr(o, dia.data(), dia.size());
    wr(o, rfin.data(), rfin.size());
    wr(o, x.data(), x.size());
    wr(o, &rr, 1);
    fclose(o);
    printf("shim_test: %s ok=%d iteration=%u conv=%.3e |b-Ax|=%.3e\n", kind == 0 ? "PCG" : "BiCGSTAB", ok ? 1 : 0,
           maxit, conv, rr > 0 ? __builtin_sqrt(rr) : 0.0);
  } catch (const delfem_amd::Error& e) {
    fprintf(stderr, "delfem_amd error: %s\n", e.what());
    return 1;
  }
  return 0;
}
