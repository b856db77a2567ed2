/*
 * dfm_oracle.h -- CPU restatement of DelFEM's linear-solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in delfem_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links or calls
 * it, and it is never the thing measured as the GPU result.
 *
 * Every function restates one reference routine, cited as file:line relative
 * to the DelFEM source tree.  Storage follows the reference exactly:
 *   rowptr[nb+1]  -- the reference's m_colInd_Blk (row pointer)
 *   col[ncrs]     -- the reference's m_rowPtr_Blk (column index), diagonal excluded
 *   val[ncrs*b*b] -- off-diagonal blocks, row-major b x b
 *   dia[nb*b*b]   -- diagonal blocks, row-major b x b
 * Orderings use the reference convention order[new] = old
 * (src/matvec/ordering_blk.cpp:46-60).
 *
 * Pinning status of each function is recorded in the header comment of
 * dfm_oracle.c.
 */
#ifndef DFM_ORACLE_H
#define DFM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* a1: node adjacency from element connectivity, self excluded, rows ascending.
 * Returns ncrs.  col may be NULL (count only); cap is the capacity of col. */
long orc_fem_pattern(unsigned nnode, unsigned nelem, unsigned npe, const unsigned *conn,
                     unsigned *rowptr, unsigned *col, long cap);

/* a5: Mearge of nelem element matrices in element order.
 * emat: nelem * (npe*npe) blocks of b*b, layout [(i*npe+j)*b*b + r*b + c]. */
void orc_assemble(unsigned nb, unsigned b, const unsigned *rowptr, const unsigned *col,
                  double *val, double *dia, unsigned nelem, unsigned npe,
                  const unsigned *conn, const double *emat);
/* residual merge: CVector_Blk::AddValue per element node in element order. */
void orc_assemble_vec(unsigned b, double *vec, unsigned nelem, unsigned npe,
                      const unsigned *conn, const double *eres);

/* a6: CMatDia_BlkCrs::SetBoundaryCondition and CBCFlag::SetZeroToBCDof. */
void orc_set_bc(unsigned nb, unsigned b, const unsigned *rowptr, const unsigned *col,
                double *val, double *dia, const int *flag);
void orc_zero_bc_dof(unsigned ndof, const int *flag, double *vec);

/* a10: y = beta*y + alpha*A*x (CMatDia_BlkCrs::MatVec). */
void orc_matvec(unsigned nb, unsigned b, const unsigned *rowptr, const unsigned *col,
                const double *val, const double *dia, double alpha, const double *x,
                double beta, double *y);

/* a7: ILU(k) symbolic pattern (natural or ordered).  lev_fill = -1 means
 * unlimited.  order may be NULL.  Returns ncrs of the factor, or -(needed)
 * when cap is too small, or -1 on bad input (unsorted row). */
long orc_iluk_symbolic(unsigned nb, const unsigned *rowptr, const unsigned *col,
                       int lev_fill, const int *order,
                       unsigned *frowptr, unsigned *fcol, int *flev, unsigned *fdiaind,
                       long cap);
/* the same with fill blocks (AddFracPtn(lev, aBlkFill), natural order) */
long orc_iluk_symbolic_blk(unsigned nb, const unsigned *rowptr, const unsigned *col, int lev_fill,
                           const unsigned *fill_blk, unsigned nfill,
                           unsigned *frowptr, unsigned *fcol, int *flev, unsigned *fdiaind, long cap);

/* a8: copy A (optionally reordered) into the factor pattern, then factor.
 * orc_ilu_decomp returns the number of singular pivots (0..10), or -1 when
 * more than 10 were met (the reference stops and returns false). */
void orc_ilu_setvalue(unsigned nb, unsigned b, const unsigned *rowptr, const unsigned *col,
                      const double *val, const double *dia, const int *order,
                      const unsigned *frowptr, const unsigned *fcol,
                      double *fval, double *fdia);
int orc_ilu_decomp(unsigned nb, unsigned b, const unsigned *frowptr, const unsigned *fcol,
                   const unsigned *fdiaind, double *fval, double *fdia);

/* a9: in-place substitutions. */
void orc_ilu_forward(unsigned nb, unsigned b, const unsigned *frowptr, const unsigned *fcol,
                     const unsigned *fdiaind, const double *fval, const double *fdia,
                     double *vec);
void orc_ilu_backward(unsigned nb, unsigned b, const unsigned *frowptr, const unsigned *fcol,
                      const unsigned *fdiaind, const double *fval, const double *fdia,
                      double *vec);

/* a12/a13: LsSol::Solve_PCG and LsSol::Solve_PBiCGSTAB on a one-segment
 * CLinearSystem with CPreconditioner_ILU (optionally ordered).
 * r: in = right-hand side (residual vector, id -1), out = final residual.
 * x: out = update vector (id -2).  conv_ratio and iteration in/out as in the
 * reference.  Returns 1 (true) / 0 (false) like the reference. */
int orc_solve_pcg(unsigned nb, unsigned b,
                  const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                  const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                  const double *fval, const double *fdia, const int *order,
                  double *r, double *x, double *conv_ratio, unsigned *iteration);
int orc_solve_pbicgstab(unsigned nb, unsigned b,
                        const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                        const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                        const double *fval, const double *fdia, const int *order,
                        double *r, double *x, double *conv_ratio, unsigned *iteration);
/* LsSol::Solve_CG (kind 0) / Solve_BiCGSTAB (kind 2), unpreconditioned
 * (src/ls/solver_ls_iter.cpp:31-97, 177-286) */
int orc_solve_ls(int kind, unsigned nb, unsigned b,
                 const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                 double *r, double *x, double *conv_ratio, unsigned *num_iter);

/* a17: MatVec::Sol::Solve_CG (kind 0) / Solve_PCG (1) / Solve_BiCGSTAB (2) /
 * Solve_PBiCGSTAB (3), src/matvec/solver_mat_iter.cpp; natural-order factor. */
int orc_solve_mat(int kind, unsigned nb, unsigned b,
                  const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                  const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                  const double *fval, const double *fdia,
                  double *r, double *u, double *conv_ratio, unsigned *iteration);

/* a4: element kernels (per element arrays). */
void orc_emat_poisson2d_p1(unsigned nelem, const double *coord, const unsigned *conn,
                           const double *alpha, double source, const double *value,
                           double *emat, double *eres);
void orc_emat_solid3d_p1(unsigned nelem, const double *coord, const unsigned *conn,
                         const double *lambda, const double *myu, double rho,
                         double gx, double gy, double gz, const double *disp,
                         double *emat, double *eres);
/* positive Helmholtz alpha*K + c*M with P1 consistent mass */
void orc_emat_helmholtz2d_p1(unsigned nelem, const double *coord, const unsigned *conn,
                             const double *alpha, const double *cmass, double source,
                             const double *value, double *emat, double *eres);
/* Oseen P1-P1 with PSPG, combined (u,v,p) 3x3 blocks */
void orc_emat_oseen2d_p1p1(unsigned nelem, const double *coord, const unsigned *conn,
                           double myu, double rho, const double *adv /* nnode*2 */,
                           const double *value /* nnode*3 */, double *emat, double *eres);

/* Orderings for the multicolour path (product algorithm, restated). */
int orc_greedy_colour(unsigned nb, const unsigned *rowptr, const unsigned *col, int *colour);
void orc_colour_order(unsigned nb, const int *colour, int ncolour, int *order);
/* block multicolour ordering; block_of receives each node's block; returns #colours */
int orc_bmc_order(unsigned nb, const unsigned *rowptr, const unsigned *col, unsigned block_nodes,
                  int *order, int *block_of);
int orc_level_sets(unsigned nb, const unsigned *frowptr, const unsigned *fcol,
                   const unsigned *fdiaind, int lower, int *level);

#ifdef __cplusplus
}
#endif

/* Complex twin, scalar (len 1) blocks; complex arrays are (re, im) double pairs. */
void orc_zassemble(unsigned nb, const unsigned *rowptr, const unsigned *col, unsigned nelem, unsigned npe,
                   const unsigned *conn, const double *emat, double *val, double *dia);
void orc_zassemble_vec(double *res, unsigned nelem, unsigned npe, const unsigned *conn, const double *eres);
void orc_zset_bc(unsigned nb, const unsigned *rowptr, const unsigned *col, double *val, double *dia,
                 const int *flag, double *res);
void orc_zmatvec(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                 double alpha, const double *x, double beta, double *y);
void orc_zilu_setvalue(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val,
                       const double *dia, const int *order, const unsigned *frowptr, const unsigned *fcol,
                       double *fval, double *fdia);
int orc_zilu_decomp(unsigned nb, const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                    double *fval, double *fdia);
void orc_zilu_forward(unsigned nb, const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                      const double *fval, const double *fdia, double *vec);
void orc_zilu_backward(unsigned nb, const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                       const double *fval, const double *fdia, double *vec);
unsigned orc_bmat_pattern_boundary(unsigned nb, unsigned b, const unsigned *rowptr, const unsigned *col,
                                   const int *flag_col, const int *flag_row, unsigned *brow, unsigned *bcol);
void orc_bmat_mearge(unsigned nb, unsigned b, const unsigned *brow, const unsigned *bcol, unsigned nelem,
                     unsigned npe, const unsigned *conn, const double *emat, double *val);
void orc_bmat_bc_inverse_col(unsigned nb, unsigned b, const unsigned *brow, const unsigned *bcol, const int *flag,
                             double *val);
void orc_bmat_bc_row(unsigned nb, unsigned b, const unsigned *brow, const unsigned *bcol, const int *flag, double *val);
void orc_bmat_matvec(unsigned nb, unsigned b, const unsigned *brow, const unsigned *bcol, const double *val,
                     double alpha, const double *x, double beta, double *y);
void orc_bmat2_matvec(unsigned nbc, unsigned bc, unsigned nbr, unsigned br, const unsigned *rowptr,
                      const unsigned *col, const double *val, double alpha, const double *x, double beta, double *y);
void orc_bmat2_mearge(unsigned nbc, unsigned bc, unsigned nbr, unsigned br, const unsigned *rowptr,
                      const unsigned *col, unsigned nelem, unsigned npe_col, const unsigned *conn_col,
                      unsigned npe_row, const unsigned *conn_row, const double *emat, double *val);
void orc_bmat2_bc(unsigned nbc, unsigned bc, unsigned nbr, unsigned br, const unsigned *rowptr, const unsigned *col,
                  const int *flag_row_dofs, const int *flag_col_dofs, double *val);
void *orc_ms_create(unsigned nseg, const unsigned *nb, const unsigned *b);
void orc_ms_free(void *p);
void orc_ms_set_dia(void *p, unsigned i, const unsigned *rowptr, const unsigned *col, const double *val,
                    const double *dia);
void orc_ms_set_nondia(void *p, unsigned i, unsigned j, const unsigned *rowptr, const unsigned *col, const double *val);
void orc_ms_matvec(void *p, double alpha, const double *x, double beta, double *y);
int orc_ms_ilu(void *p, int lev_fill);
int orc_ms_ilu_levels(void *p, const int *alev, const unsigned *fill_blk, unsigned nfill);
unsigned orc_ms_factor_ncrs(void *p, unsigned i, unsigned j);
void orc_ms_get_pattern(void *p, unsigned i, unsigned j, unsigned *rowptr, unsigned *col);
void orc_ms_get_factor(void *p, unsigned i, unsigned j, double *fval, double *fdia);
void orc_ms_precond(void *p, double *z);
int orc_ms_solve(void *p, int bicg, double *r, double *x, double *conv_ratio, unsigned *iteration);
void orc_zmatvec_hermitian(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val,
                           const double *dia, double alpha, const double *x, double beta, double *y);
int orc_zsolve_cgnr(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                    double *r, double *u, double *conv_ratio, unsigned *iteration);
int orc_zsolve_cg_herm(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                       int precond, const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind,
                       const double *fval, const double *fdia, const int *order, double *r, double *u,
                       double *conv_ratio, unsigned *iteration);
int orc_zsolve_pcocg(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                     const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind, const double *fval,
                     const double *fdia, const int *order, double *r, double *u, double *conv_ratio,
                     unsigned *iteration);
int orc_zsolve_cocg(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val, const double *dia,
                    double *r, double *u, double *conv_ratio, unsigned *iteration);
int orc_zsolve_bicgstab(unsigned nb, const unsigned *rowptr, const unsigned *col, const double *val,
                        const double *dia, int precond, const unsigned *frowptr, const unsigned *fcol,
                        const unsigned *fdiaind, const double *fval, const double *fdia, const int *order,
                        double *r, double *u, double *conv_ratio, unsigned *iteration);
void orc_zemat_helmholtz2d_p1(unsigned nelem, const double *coord, const unsigned *conn, double wave_length,
                              const double *value, double *emat, double *eres);
void orc_zemat_sommerfeld_b1(unsigned nline, const double *coord, const unsigned *conn, double wave_length,
                             const double *value, double *emat, double *eres);


/* Variable block sizes (flex): dofptr[nb+1] dof offsets, valptr[ncrs+1] and
 * diaptr[nb+1] value offsets of the CRS and diagonal blocks (row-major
 * len_i x len_j), as m_DofPtrCol / m_ValPtr / m_DiaValPtr of the reference. */
void orc_fx_layout(unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                   unsigned *valptr, unsigned *diaptr);
void orc_fx_matvec(unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                   const unsigned *valptr, const double *val, const unsigned *diaptr, const double *dia,
                   double alpha, const double *x, double beta, double *y);
void orc_fx_set_bc(unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                   const unsigned *valptr, double *val, const unsigned *diaptr, double *dia, const int *flag);
void orc_fx_mearge(unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                   const unsigned *valptr, double *val, const unsigned *diaptr, double *dia,
                   unsigned n, const unsigned *iblk, const unsigned *jblk, const double *emat);
void orc_fx_ilu_setvalue(unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                         const unsigned *valptr, const double *val, const unsigned *diaptr, const double *dia,
                         const unsigned *frowptr, const unsigned *fcol, const unsigned *fvalptr,
                         double *fval, double *fdia);
int orc_fx_ilu_decomp(unsigned nb, const unsigned *dofptr, const unsigned *frowptr, const unsigned *fcol,
                      const unsigned *fdiaind, const unsigned *fvalptr, double *fval, const unsigned *diaptr,
                      double *fdia);
void orc_fx_ilu_forward(unsigned nb, const unsigned *dofptr, const unsigned *frowptr, const unsigned *fcol,
                        const unsigned *fdiaind, const unsigned *fvalptr, const double *fval,
                        const unsigned *diaptr, const double *fdia, double *vec);
void orc_fx_ilu_backward(unsigned nb, const unsigned *dofptr, const unsigned *frowptr, const unsigned *fcol,
                         const unsigned *fdiaind, const unsigned *fvalptr, const double *fval, double *vec);
int orc_fx_solve(int kind, unsigned nb, const unsigned *dofptr, const unsigned *rowptr, const unsigned *col,
                 const unsigned *valptr, const double *val, const unsigned *diaptr, const double *dia,
                 const unsigned *frowptr, const unsigned *fcol, const unsigned *fdiaind, const unsigned *fvalptr,
                 const double *fval, const double *fdia, double *r, double *x, double *conv_ratio,
                 unsigned *iteration);

#endif
