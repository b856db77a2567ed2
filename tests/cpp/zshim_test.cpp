// Drives the complex drop-in classes (include/delfem_amd/zmatvec.hpp) the way
// DelFEM's complex Helmholtz code does (src/femls/zsolver_ls_iter.cpp,
// include/delfem/femls/zpreconditioner.h): pattern from the triangles, one
// Mearge per element (triangles, then Sommerfeld lines), Dirichlet flags,
// ILU(lev) of the matrix, then one MatVec::Sol solver.  Input and output are
// raw binary files written / read by tests/test_gpu_shim.py.
//
// input : i32 nnode ntri nline kind maxit lev | f64 tol | u32 tri[ntri*3] |
//         u32 line[nline*2] | c128 emat_tri[ntri*9] | c128 emat_line[nline*4] |
//         c128 rhs[nnode] | i32 flags[nnode]
//         kind: 0 Solve_COCG, 1 Solve_PCOCG, 2 Solve_BiCGSTAB, 3 Solve_BiCGStabP,
//               4 Solve_CGNR, 5 Solve_CG and 6 Solve_PCG (Hermitian matrices)
// output: f64 conv | i32 ok iteration | c128 val[ncrs] dia[nnode] u[nnode] r[nnode]
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "delfem_amd/zmatvec.hpp"

typedef std::complex<double> cplx;

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
    fprintf(stderr, "usage: zshim_test IN OUT\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[6];
  rd(f, hdr, 6);
  const unsigned nnode = hdr[0], ntri = hdr[1], nline = hdr[2];
  const int kind = hdr[3], lev = hdr[5];
  unsigned maxit = (unsigned)hdr[4];
  double tol;
  rd(f, &tol, 1);
  std::vector<unsigned> tri((size_t)ntri * 3), line((size_t)nline * 2);
  std::vector<cplx> et((size_t)ntri * 9), el((size_t)nline * 4), rhs(nnode);
  std::vector<int> flags(nnode);
  rd(f, tri.data(), tri.size());
  rd(f, line.data(), line.size());
  rd(f, et.data(), et.size());
  rd(f, el.data(), el.size());
  rd(f, rhs.data(), rhs.size());
  rd(f, flags.data(), flags.size());
  fclose(f);
  if (dfm_init(0) != DFM_OK) {
    fprintf(stderr, "dfm_init: %s\n", dfm_last_error());
    return 1;
  }
  try {
    std::vector<std::set<unsigned>> adj(nnode);
    for (unsigned e = 0; e < ntri; e++)
      for (unsigned a = 0; a < 3; a++)
        for (unsigned b = 0; b < 3; b++)
          if (a != b) adj[tri[e * 3 + a]].insert(tri[e * 3 + b]);
    std::vector<unsigned> index(nnode + 1, 0), array;
    for (unsigned i = 0; i < nnode; i++) {
      array.insert(array.end(), adj[i].begin(), adj[i].end());
      index[i + 1] = (unsigned)array.size();
    }
    MatVec::CZMatDia_BlkCrs A(nnode, 1);
    A.AddPattern(index, array);
    A.SetZero();
    for (unsigned e = 0; e < ntri; e++)
      if (!A.Mearge(3, &tri[e * 3], 3, &tri[e * 3], 1, &et[(size_t)e * 9])) return 1;
    for (unsigned e = 0; e < nline; e++)
      if (!A.Mearge(2, &line[e * 2], 2, &line[e * 2], 1, &el[(size_t)e * 4])) return 1;
    MatVec::CBCFlag bc(nnode, 1);
    for (unsigned i = 0; i < nnode; i++)
      if (flags[i]) bc.SetBC(i, 0);
    A.SetBoundaryCondition(bc);
    std::vector<cplx> val, dia;
    A.GetValues(val, dia);
    MatVec::CZVector_Blk r(nnode, 1), u(nnode, 1);
    r.SetValues(rhs.data());
    MatVec::CZMatDiaFrac_BlkCrs prec(lev, A);
    prec.SetValue_Initialize(A);
    prec.DoILUDecomp();
    double conv = tol;
    bool ok = false;
    switch (kind) {
      case 0: ok = MatVec::Sol::Solve_COCG(conv, maxit, A, r, u); break;
      case 1: ok = MatVec::Sol::Solve_PCOCG(conv, maxit, A, r, u, prec, A); break;
      case 2: ok = MatVec::Sol::Solve_BiCGSTAB(conv, maxit, A, r, u); break;
      case 3: ok = MatVec::Sol::Solve_BiCGStabP(conv, maxit, A, r, u, prec, A); break;
      case 4: ok = MatVec::Sol::Solve_CGNR(conv, maxit, A, r, u); break;
      case 5: ok = MatVec::Sol::Solve_CG(conv, maxit, A, r, u); break;
      case 6: ok = MatVec::Sol::Solve_PCG(conv, maxit, A, r, u, prec, A); break;
      default: return 2;
    }
    const std::vector<cplx> uh = u.GetValues(), rh = r.GetValues();
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    const int i2[2] = {ok ? 1 : 0, (int)maxit};
    wr(o, &conv, 1);
    wr(o, i2, 2);
    wr(o, val.data(), val.size());
    wr(o, dia.data(), dia.size());
    wr(o, uh.data(), uh.size());
    wr(o, rh.data(), rh.size());
    fclose(o);
    printf("zshim_test: kind %d ok=%d iteration=%u conv=%.3e\n", kind, ok ? 1 : 0, maxit, conv);
  } catch (const delfem_amd::Error& e) {
    fprintf(stderr, "delfem_amd error: %s\n", e.what());
    return 1;
  }
  return 0;
}
