// Drives the C++ drop-in classes (include/delfem_amd/{matvec,ls}.hpp) the way
// DelFEM's equation systems do: pattern from the element list, InitializeMarge,
// one Mearge per element, residual, Dirichlet flags, FinalizeMarge, ILU(0),
// then one solver: LsSol::Solve_PCG / Solve_PBiCGSTAB, or the matrix-layer
// MatVec::Sol::Solve_PCG / Solve_PBiCGSTAB / Solve_CG / Solve_BiCGSTAB on the
// assembled matrix (with the ILU(0) factor itself, which takes the device
// solver, or wrapped in another CPrecond_Blk, which takes the generic loop).
// Input and output are raw binary files written / read by tests/test_gpu_shim.py.
//
// input : i32 nnode nelem len npe kind maxit | f64 tol |
//         u32 conn[nelem*npe] | f64 emat[nelem*npe*npe*len*len] |
//         f64 eres[nelem*npe*len] | i32 flags[nnode*len]
//         kind: 0 LsSol PCG, 1 LsSol BiCGSTAB, 2 Sol::Solve_PCG, 3 Sol::Solve_PBiCGSTAB,
//         4 / 5 the same through a wrapping CPrecond_Blk, 6 Sol::Solve_CG, 7 Sol::Solve_BiCGSTAB
// output: f64 norm_res conv | i32 ok iteration | f64 val[ncrs*len*len]
//         dia[nnode*len*len] res[ndof] x[ndof] |b-Ax|^2 r_final[ndof]
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "delfem_amd/ls.hpp"

// A preconditioner that is not the factor class itself: the matrix-layer
// solvers then run their generic loop.
struct WrappedPrecond : MatVec::CPrecond_Blk {
  explicit WrappedPrecond(const MatVec::CMatDiaFrac_BlkCrs& f) : f(f) {}
  bool SolvePrecond(const MatVec::CMatDia_BlkCrs& mat, MatVec::CVector_Blk& v) const override {
    return f.SolvePrecond(mat, v);
  }
  const MatVec::CMatDiaFrac_BlkCrs& f;
};

static const char* kind_name(int kind) {
  static const char* names[] = {"LsSol::Solve_PCG",          "LsSol::Solve_PBiCGSTAB",
                                "Sol::Solve_PCG",            "Sol::Solve_PBiCGSTAB",
                                "Sol::Solve_PCG (wrapped)",  "Sol::Solve_PBiCGSTAB (wrapped)",
                                "Sol::Solve_CG",             "Sol::Solve_BiCGSTAB"};
  return kind >= 0 && kind < 8 ? names[kind] : "?";
}

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
    double conv = tol;
    bool ok = false;
    std::vector<double> x, rout;
    if (kind <= 1) {
      LsSol::CPreconditioner_ILU prec(ls, 0);
      if (!prec.SetValue(ls)) return 1;
      LsSol::CLinearSystemPreconditioner lsp(ls, prec);
      ok = kind == 0 ? LsSol::Solve_PCG(conv, maxit, lsp) : LsSol::Solve_PBiCGSTAB(conv, maxit, lsp);
      x = ls.GetUpdate(0).GetValues();
      rout = ls.GetResidual(0).GetValues();
    } else {
      MatVec::CVector_Blk r(nnode, len), u(nnode, len);
      r.SetValues(rfin.data());
      MatVec::CMatDiaFrac_BlkCrs fac(0, A);
      if (!fac.SetValue(A)) return 1;
      WrappedPrecond wrapped(fac);
      switch (kind) {
        case 2: ok = MatVec::Sol::Solve_PCG(conv, maxit, A, r, u, fac, A); break;
        case 3: ok = MatVec::Sol::Solve_PBiCGSTAB(conv, maxit, A, r, u, fac, A); break;
        case 4: ok = MatVec::Sol::Solve_PCG(conv, maxit, A, r, u, wrapped, A); break;
        case 5: ok = MatVec::Sol::Solve_PBiCGSTAB(conv, maxit, A, r, u, wrapped, A); break;
        case 6: ok = MatVec::Sol::Solve_CG(conv, maxit, A, r, u); break;
        case 7: ok = MatVec::Sol::Solve_BiCGSTAB(conv, maxit, A, r, u); break;
        default: return 2;
      }
      x = u.GetValues();
      rout = r.GetValues();
    }
    // the API's own check: ||b - A x|| with MATVEC on a scratch vector
    ls.GetUpdate(0).SetValues(x.data());
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
    wr(o, rout.data(), rout.size());
    fclose(o);
    printf("shim_test: %s ok=%d iteration=%u conv=%.3e |b-Ax|=%.3e\n", kind_name(kind), ok ? 1 : 0, maxit, conv,
           rr > 0 ? __builtin_sqrt(rr) : 0.0);
  } catch (const delfem_amd::Error& e) {
    fprintf(stderr, "delfem_amd error: %s\n", e.what());
    return 1;
  }
  return 0;
}
