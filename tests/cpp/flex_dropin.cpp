// A caller of the drop-in with blocks of different sizes, written the way
// DelFEM's rigid-body system uses them (Ls::CLinearSystem_RigidField2::
// SetRigidSystem, src/rigid/linearsystem_rigidfield.cpp:496-562:
// AddLinSysSeg(nblk, aBlkSize) and AddMat_Dia(ilss, crs) with a
// Com::CIndexedArray pattern; AddResidual / UpdateValueOfRigidSystem,
// :563-620: per block-dof AddValue / GetValue on the segment's vectors):
// the blocks merged one at a time through GetMatrix(0).Mearge(1, &i, 1, &j,
// ...), FinalizeMarge, then one of the solvers.  Also checks that CBCFlag
// refuses flags of a segment of per-block lengths, as the reference's does
// (bcflag_blk.h:153-158).  Input and output are raw binary files written /
// read by tests/test_gpu_flex_dropin.py.
//
// input : i32 nblk nentry kind maxit lev | f64 tol |
//         u32 alen[nblk] index[nblk+1] array[index[nblk]] |
//         u32 iblk[nentry] jblk[nentry] | f64 emat[sum of len_i*len_j] |
//         f64 rhs[ndof]
//   kind: 0 Solve_PCG + CPreconditioner_ILU (device loop), 1 Solve_PBiCGSTAB +
//   CPreconditioner_ILU, 2 Solve_CG, 3 Solve_BiCGSTAB, 4 Solve_PCG + a user
//   CPreconditioner over MatVec::CMatDiaFrac_BlkCrs (the interface loop),
//   5 MatVec::Sol::Solve_PCG with a CMatDiaFrac_BlkCrs as CPrecond_Blk (the
//   reference loop over device vectors)
// output: i32 ok iteration setbc_refused len_minus1 | f64 conv norm_res |
//         f64 val[nval] dia[ndia] x[ndof]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "delfem_amd/ls.hpp"

using namespace MatVec;

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

// a user preconditioner over the matrix-layer factor of the segment
class UserILU : public LsSol::CPreconditioner {
 public:
  explicit UserILU(int lev) : lev_(lev) {}
  void SetLinearSystem(const LsSol::CLinearSystem& ls) override {
    f_.reset(new CMatDiaFrac_BlkCrs(lev_, ls.GetMatrix(0)));
  }
  bool SetValue(const LsSol::CLinearSystem& ls) override {
    f_->SetValue(ls.GetMatrix(0));
    return f_->DoILUDecomp();
  }
  bool SolvePrecond(LsSol::CLinearSystem& ls, unsigned iv) override { return f_->Solve(ls.GetVector((int)iv, 0)); }

 private:
  int lev_;
  std::unique_ptr<CMatDiaFrac_BlkCrs> f_;
};

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: flex_dropin IN OUT\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[5];
  rd(f, hdr, 5);
  const unsigned nblk = (unsigned)hdr[0], nentry = (unsigned)hdr[1];
  const int kind = hdr[2];
  unsigned maxit = (unsigned)hdr[3];
  const int lev = hdr[4];
  double tol;
  rd(f, &tol, 1);
  std::vector<unsigned> aBlkSize(nblk);
  rd(f, aBlkSize.data(), nblk);
  Com::CIndexedArray crs;
  crs.InitializeSize(nblk);
  rd(f, crs.index.data(), nblk + 1);
  crs.array.resize(crs.index[nblk]);
  rd(f, crs.array.data(), crs.array.size());
  std::vector<unsigned> ib(nentry), jb(nentry);
  rd(f, ib.data(), nentry);
  rd(f, jb.data(), nentry);
  size_t nemat = 0;
  for (unsigned e = 0; e < nentry; e++) nemat += (size_t)aBlkSize[ib[e]] * aBlkSize[jb[e]];
  std::vector<double> emat(nemat);
  rd(f, emat.data(), nemat);
  unsigned ndof = 0;
  for (unsigned i = 0; i < nblk; i++) ndof += aBlkSize[i];
  std::vector<double> rhs(ndof);
  rd(f, rhs.data(), ndof);
  fclose(f);
  if (dfm_init(0) != DFM_OK) {
    fprintf(stderr, "dfm_init: %s\n", dfm_last_error());
    return 1;
  }
  try {
    LsSol::CLinearSystem ls;
    const int ilss = ls.AddLinSysSeg(nblk, aBlkSize);
    if (ilss != 0 || !ls.AddMat_Dia(ilss, crs)) return 5;
    ls.InitializeMarge();
    // element contributions, one block per Mearge call
    size_t off = 0;
    CMatDia_BlkCrs& mat = ls.GetMatrix(ilss);
    for (unsigned e = 0; e < nentry; e++) {
      const unsigned bs = aBlkSize[ib[e]] * aBlkSize[jb[e]];
      if (!mat.Mearge(1, &ib[e], 1, &jb[e], bs, emat.data() + off)) return 6;
      off += bs;
    }
    // the residual, block dof by block dof (AddResidual)
    CVector_Blk& res = ls.GetVector(-1, ilss);
    unsigned d = 0;
    for (unsigned i = 0; i < nblk; i++)
      for (unsigned k = 0; k < res.Len(i); k++) res.AddValue(i, k, rhs[d++]);
    const int setbc_refused = ls.GetBCFlag(ilss).SetBC(0, 0) ? 0 : 1;
    const int len_minus1 = (res.Len() == -1 && mat.LenBlkCol() == -1 && ls.GetBCFlag(ilss).LenBlk() == -1) ? 1 : 0;
    const double norm_res = ls.FinalizeMarge();
    std::vector<double> val, dia;
    mat.GetValues(val, dia);

    double conv = tol;
    bool ok = false;
    if (kind == 0 || kind == 1) {
      LsSol::CPreconditioner_ILU prec(ls, (unsigned)lev);
      if (!prec.SetValue(ls)) return 7;
      LsSol::CLinearSystemPreconditioner lsp(ls, prec);
      ok = kind == 1 ? LsSol::Solve_PBiCGSTAB(conv, maxit, lsp) : LsSol::Solve_PCG(conv, maxit, lsp);
    } else if (kind == 2) {
      ok = LsSol::Solve_CG(conv, maxit, ls);
    } else if (kind == 3) {
      ok = LsSol::Solve_BiCGSTAB(conv, maxit, ls);
    } else if (kind == 4) {
      UserILU prec(lev);
      prec.SetLinearSystem(ls);
      if (!prec.SetValue(ls)) return 7;
      LsSol::CLinearSystemPreconditioner lsp(ls, prec);
      ok = LsSol::Solve_PCG(conv, maxit, lsp);
    } else if (kind == 5) {
      CMatDiaFrac_BlkCrs fac(lev, mat);
      fac.SetValue(mat);
      CVector_Blk r(ls.GetVector(-1, ilss));
      CVector_Blk& u = ls.GetVector(-2, ilss);
      ok = MatVec::Sol::Solve_PCG(conv, maxit, mat, r, u, fac, mat);
    } else {
      return 2;
    }
    // the update, block dof by block dof (UpdateValueOfRigidSystem)
    const CVector_Blk& upd = ls.GetVector(-2, ilss);
    std::vector<double> x;
    for (unsigned i = 0; i < nblk; i++)
      for (unsigned k = 0; k < upd.Len(i); k++) x.push_back(upd.GetValue(i, k));

    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    const int i4[4] = {ok ? 1 : 0, (int)maxit, setbc_refused, len_minus1};
    const double dd[2] = {conv, norm_res};
    wr(o, i4, 4);
    wr(o, dd, 2);
    wr(o, val.data(), val.size());
    wr(o, dia.data(), dia.size());
    wr(o, x.data(), x.size());
    fclose(o);
    printf("flex_dropin: kind %d ok=%d iteration=%u conv=%.3e\n", kind, ok ? 1 : 0, maxit, conv);
  } catch (const delfem_amd::Error& e) {
    fprintf(stderr, "delfem_amd error: %s\n", e.what());
    return 1;
  }
  return 0;
}
