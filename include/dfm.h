/*
 * dfm.h -- C-ABI of delfem_amd, the MI355X (gfx950) implementation of
 * DelFEM's linear-solve hot path: block-CRS matrices with a separate diagonal,
 * finite-element assembly, ILU(k) factorisation and preconditioned CG /
 * BiCGSTAB.
 *
 * Conventions (identical to the DelFEM reference):
 *  - a square matrix has nblk block rows of len x len blocks; the CRS part
 *    excludes the diagonal, which is held separately; blocks are row-major;
 *  - "index"/"array" arguments follow Com::CIndexedArray (row pointer of
 *    nblk+1 entries and column list, self excluded, each row ascending);
 *  - an ordering is order[new] = old;
 *  - vectors are block vectors of nblk*len doubles in the reference layout.
 * Vector arguments named *_dev are device pointers (hipMalloc'd memory of the
 * current device, or memory from dfm_malloc); arrays without the suffix are
 * host memory.  All calls are synchronous with respect to the host unless
 * stated otherwise.  Every call returns a dfm status; dfm_last_error()
 * describes the last failure.  Handles are thread-compatible (one thread at a
 * time per handle), not thread-safe.
 *
 * Each entry point names the reference interface it replaces
 * (file:line in the DelFEM tree).
 */
#ifndef DFM_H
#define DFM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DFM_OK = 0,
    DFM_NOT_CONVERGED = 1,     /* Solve_PCG returning false (solver_ls_iter.cpp:168-171) */
    DFM_SINGULAR = 2,          /* more than 10 singular pivots (matdiafrac_blkcrs.cpp:774,880-890) */
    DFM_BAD_ARG = -1,
    DFM_HIP_ERROR = -2,
    DFM_UNSORTED = -3,         /* a pattern row is not ascending (reference asserts) */
    DFM_NOT_IMPLEMENTED = -4,
    DFM_TOO_LARGE = -5
};

typedef struct dfm_bcrs_s dfm_bcrs;   /* MatVec::CMatDia_BlkCrs */
typedef struct dfm_ilu_s dfm_ilu;     /* MatVec::CMatDiaFrac_BlkCrs / LsSol::CPreconditioner_ILU */
typedef struct dfm_asm_s dfm_asm;     /* element->matrix scatter plan (femeqn AddLinSys_* loops) */
typedef struct dfm_batch_s dfm_batch; /* many independent small systems */

/* ---- runtime ------------------------------------------------------------ */
const char *dfm_last_error(void);
int dfm_init(int device);                       /* select the device, create the library stream */
int dfm_synchronize(void);
int dfm_malloc(void **ptr_dev, size_t bytes);
int dfm_free(void *ptr_dev);
int dfm_memcpy_htod(void *dst_dev, const void *src, size_t bytes);
int dfm_memcpy_dtoh(void *dst, const void *src_dev, size_t bytes);
int dfm_memcpy_dtod(void *dst_dev, const void *src_dev, size_t bytes);
/* Stream used by the library (hipStream_t cast to void*).  dfm_set_stream makes
 * the library issue its work on a caller's stream (NULL: the default stream). */
void *dfm_stream(void);
int dfm_set_stream(void *stream);

/* ---- pattern (Com::CIndexedArray from element connectivity) -------------- */
/* CElemAry::MakePattern_FEM, src/femfield/elem_ary.cpp:225-245 (+675-892):
 * node adjacency without self, rows ascending.  Returns the number of
 * entries in *ncrs; col may be NULL to query the size (cap = capacity). */
int dfm_fem_pattern(unsigned nnode, unsigned nelem, unsigned npe, const unsigned *conn,
                    unsigned *rowptr, unsigned *col, size_t cap, size_t *ncrs);

/* ---- MatVec::CMatDia_BlkCrs (include/delfem/matvec/matdia_blkcrs.h:44-118) */
int dfm_bcrs_create(unsigned nblk, unsigned len, dfm_bcrs **out);          /* CMatDia_BlkCrs(nblk,len) matdia_blkcrs.cpp:40 */
int dfm_bcrs_destroy(dfm_bcrs *A);
int dfm_bcrs_add_pattern(dfm_bcrs *A, unsigned nrow, const unsigned *index,
                         const unsigned *array);                           /* AddPattern(CIndexedArray) matdia_blkcrs.cpp:346-505 */
int dfm_bcrs_info(const dfm_bcrs *A, unsigned *nblk, unsigned *len, unsigned *ncrs); /* NBlkMatCol/LenBlkCol/NCrs mat_blkcrs.h:83-157 */
int dfm_bcrs_get_pattern(const dfm_bcrs *A, unsigned *rowptr, unsigned *col);        /* GetPtrIndPSuP mat_blkcrs.h:158-165 */
int dfm_bcrs_set_zero(dfm_bcrs *A);                                        /* SetZero matdia_blkcrs.cpp:139-163 */
int dfm_bcrs_set_values(dfm_bcrs *A, const double *val, const double *dia); /* host values into the pattern (GetPtrValPSuP/GetPtrValDia writes) */
int dfm_bcrs_get_values(const dfm_bcrs *A, double *val, double *dia);       /* GetPtrValPSuP/GetPtrValDia reads */
int dfm_bcrs_device_values(dfm_bcrs *A, double **val_dev, double **dia_dev); /* device arrays of the values (CRS order) */
int dfm_bcrs_mearge(dfm_bcrs *A, unsigned nblkel, const unsigned *blkel,
                    unsigned blksize, const double *emat);                  /* Mearge (one element, host emat) matdia_blkcrs.cpp:165-258 */
int dfm_bcrs_set_bc(dfm_bcrs *A, const int *flag);                          /* SetBoundaryCondition(CBCFlag) matdia_blkcrs.cpp:261-343 */
/* MatVec matdia_blkcrs.cpp:921-1094, bitwise the reference's CRS order.  For b <= 3 the first call
 * keeps a sliced-ELL copy of the values (about the matrix's size again, refreshed when they change;
 * values written through dfm_bcrs_device_values pointers after a MatVec need that call again). */
int dfm_bcrs_matvec(dfm_bcrs *A, double alpha, const double *x_dev, double beta,
                    double *y_dev);

/* ---- MatVec::CVector_Blk / LsSol::CLinearSystem BLAS-1 (linearsystem.cpp:250-415) */
int dfm_vec_dot(size_t n, const double *x_dev, const double *y_dev, double *result); /* DOT / operator* vector_blk.cpp:37-47 */
int dfm_vec_axpy(size_t n, double alpha, const double *x_dev, double *y_dev);        /* AXPY vector_blk.cpp:88-97 */
int dfm_vec_scal(size_t n, double alpha, double *x_dev);                              /* SCAL / operator*= vector_blk.cpp:54-59 */
int dfm_vec_copy(size_t n, const double *x_dev, double *y_dev);                       /* COPY / operator= vector_blk.cpp:62-71 */
int dfm_vec_fill(size_t n, double value, double *x_dev);                             /* SetVectorZero (value 0) vector_blk.cpp:100-105 */
int dfm_vec_zero_bc(size_t n, const int *flag, double *x_dev);                        /* CBCFlag::SetZeroToBCDof bcflag_blk.h:102-129 */

/* ---- element assembly (femeqn AddLinSys_* loops + Mearge) ---------------- */
/* A plan that adds element matrices into A in element order per entry (the
 * accumulation order of sequential Mearge calls).  conn: nelem*npe host array. */
int dfm_asm_create(dfm_bcrs *A, unsigned nnode, unsigned nelem, unsigned npe,
                   const unsigned *conn, dfm_asm **out);
int dfm_asm_destroy(dfm_asm *p);
/* AddLinearSystem_Poisson2D_P1, src/femeqn/eqn_poisson.cpp:49-125 (alpha per element).
 * coord_dev: nnode*2, alpha_dev: nelem, value_dev: nnode, res_dev: nnode (accumulated). */
int dfm_assemble_poisson2d_p1(dfm_asm *p, const double *coord_dev, const double *alpha_dev,
                              double source, const double *value_dev, double *res_dev);
/* AddLinSys_LinearSolid3D_Static_P1, src/femeqn/eqn_linear_solid3d.cpp:241-334
 * (lambda, myu per element).  coord_dev: nnode*3, disp_dev: nnode*3, res_dev: nnode*3. */
int dfm_assemble_solid3d_p1(dfm_asm *p, const double *coord_dev, const double *lambda_dev,
                            const double *myu_dev, double rho, double gx, double gy, double gz,
                            const double *disp_dev, double *res_dev);
/* Steady Oseen P1-P1 with PSPG on combined (u,v,p) 3x3 blocks: viscous,
 * Galerkin advection, pressure/continuity, PSPG advection and PSPG pressure
 * terms of MakeMat_NavierStokes2D_NonStatic_Newmark_P1P1
 * (src/femeqn/eqn_navier_stokes.cpp:73-110,136-158,231-253) with a prescribed
 * advection field and the steady stabilisation parameter (no Courant term),
 * merged as the combined assembly src/femeqn/eqn_stokes.cpp:231-346 does.
 * coord_dev: nnode*2, adv_dev: nnode*2, value_dev: nnode*3, res_dev: nnode*3
 * (accumulated: res -= E*value). */
int dfm_assemble_oseen2d_p1p1(dfm_asm *p, const double *coord_dev, double myu, double rho,
                              const double *adv_dev, const double *value_dev, double *res_dev);
/* Positive Helmholtz / implicit diffusion alpha*K + c*M with the P1 consistent
 * mass of AddLinearSystem_Diffusion2D_P1 (src/femeqn/eqn_diffusion.cpp:240-345,
 * rho -> c, gamma*dt -> 1), alpha and c per element.  res_dev: nnode
 * (accumulated: res += f*area/3 - E*value). */
int dfm_assemble_helmholtz2d_p1(dfm_asm *p, const double *coord_dev, const double *alpha_dev,
                                const double *c_dev, double source, const double *value_dev,
                                double *res_dev);

/* ---- ILU (MatVec::CMatDiaFrac_BlkCrs, LsSol::CPreconditioner_ILU) -------- */
enum { DFM_ORDER_NATURAL = 0, DFM_ORDER_GIVEN = 1, DFM_ORDER_MULTICOLOUR = 2, DFM_ORDER_BLOCK_MULTICOLOUR = 3 };
/* Symbolic ILU(lev_fill) of A's pattern: CMatDiaFrac_BlkCrs(lev,A) / (lev,A,order),
 * matdiafrac_blkcrs.cpp:227-324 + AddFracPtn :1249-1377; CPreconditioner_ILU::SetLinearSystem
 * preconditioner.cpp:57-211.  ordering: NATURAL (order ignored), GIVEN (order[new]=old),
 * MULTICOLOUR (greedy smallest-last colouring of A's graph computed here),
 * BLOCK_MULTICOLOUR (connected blocks of at most block_nodes nodes, block
 * graph coloured, natural order inside a block; block_nodes 0 means 32).
 * block_nodes is ignored by the other orderings.  lev_fill = -1 means unlimited. */
int dfm_ilu_create(const dfm_bcrs *A, int lev_fill, int ordering, const int *order, unsigned block_nodes,
                   dfm_ilu **out);
/* Natural-order symbolic ILU(lev_fill) in which the nfill block rows listed in fill_blk
 * (and the same block columns) take fill of any level: CMatDiaFrac_BlkCrs::AddFracPtn(lev,
 * aBlkFill), matdiafrac_blkcrs.cpp:1379-1516, reached through CPreconditioner_ILU::SetFillBlk
 * (include/delfem/ls/preconditioner.h:80, preconditioner.cpp:149-150).  Fill (i, j) of level l
 * is kept when l <= lev_fill for both i and j (lev_fill -1: unlimited) or when i or j is listed. */
int dfm_ilu_create_fill(const dfm_bcrs *A, int lev_fill, const unsigned *fill_blk, unsigned nfill, dfm_ilu **out);
int dfm_ilu_destroy(dfm_ilu *M);
/* ncrs of the factor, number of level sets of L and U, colours (0 for NATURAL / GIVEN) */
int dfm_ilu_info(const dfm_ilu *M, unsigned *ncrs, int *nlevel_lower, int *nlevel_upper, int *ncolour);
/* Phase tiling of the triangular sweeps (no reference counterpart; a
 * diagnostic): phases = colours swept one launch each, tiles per sweep, and
 * the threads of a tile; all 0 when the factor has none (natural or given
 * order, or a dependency component of more than 256 rows). */
int dfm_ilu_tile_info(const dfm_ilu *M, int *nphase, int *ntile_lower, int *ntile_upper, int *tile_threads);
/* Factor pattern in factor numbering; any pointer may be NULL. order: nblk entries. */
int dfm_ilu_get_pattern(const dfm_ilu *M, unsigned *rowptr, unsigned *col, int *lev,
                        unsigned *diaind, int *order);
/* Numeric factorisation on the device: SetValue(A) = SetValue_Initialize + DoILUDecomp,
 * matdiafrac_blkcrs.cpp:1518-1580, 769-1102; CPreconditioner_ILU::SetValue preconditioner.cpp:217-297.
 * *nsingular (may be NULL) receives the singular pivot count; DFM_SINGULAR when > 10. */
int dfm_ilu_set_value(dfm_ilu *M, dfm_bcrs *A, int *nsingular);
/* Factor values (factor numbering): val: ncrs*len*len (L, D^-1 U), dia: nblk*len*len (D^-1). */
int dfm_ilu_get_values(const dfm_ilu *M, double *val, double *dia);
/* In-place solves on a device vector in factor numbering:
 * ForwardSubstitution / BackwardSubstitution matdiafrac_blkcrs.cpp:332-663. */
int dfm_ilu_forward(dfm_ilu *M, double *vec_dev);
int dfm_ilu_backward(dfm_ilu *M, double *vec_dev);
/* CPreconditioner_ILU::SolvePrecond preconditioner.cpp:301-335 on a vector in the
 * original numbering (permuted in and out when the factor is ordered). */
int dfm_ilu_solve_precond(dfm_ilu *M, double *vec_dev);
/* The same without allocating (no counterpart in the reference): for callers
 * that cannot allocate device memory -- code under graph capture, serving
 * loops, harnesses that forbid hipMalloc.  An ordered factor's permuted copy
 * lives in a per-device scratch array of 4 Mi doubles (32 MiB, held until the
 * process ends): DFM_TOO_LARGE for an ordered factor of more values; natural-
 * order factors need no scratch.  Calls that overlap on one device share the
 * scratch and race.  Needs dfm_init. */
int dfm_ilu_solve_precond_noalloc(dfm_ilu *M, double *vec_dev);

/* ---- Batched solves: many independent scalar systems, one persistent launch */
/* No single reference counterpart: each system gets the reference's ordered
 * CPreconditioner_ILU(0) (preconditioner.cpp:57-297, DoILUDecomp b = 1,
 * matdiafrac_blkcrs.cpp:769-1102) and Solve_PCG (solver_ls_iter.cpp:100-172).
 * A is one assembled 1x1-block matrix whose diagonal blocks, rows
 * [sys_ptr[s], sys_ptr[s+1]) (host array, nsys+1 entries), are the systems;
 * any coupling between systems is rejected.  Each system is ordered by a
 * greedy colouring of its graph; DFM_NOT_IMPLEMENTED for a system needing
 * more colours than the kernel holds, a row of more than 16 off-diagonal
 * entries, or a one-way coupling (a structurally unsymmetric pattern that
 * leaves two coupled rows in one colour). */
int dfm_batch_create(const dfm_bcrs *A, unsigned nsys, const unsigned *sys_ptr, dfm_batch **out);
int dfm_batch_destroy(dfm_batch *B);
int dfm_batch_info(const dfm_batch *B, unsigned *max_width, int *max_colours, int *block, int *wg_per_cu);
/* Take A's current values (after assembly and boundary conditions). */
int dfm_batch_set_value(dfm_batch *B, const dfm_bcrs *A);
/* Launch shape of the persistent kernel: threads per workgroup (256, 512,
 * 1024) and workgroups per compute unit (block * wg_per_cu <= 2048). */
int dfm_batch_set_launch(dfm_batch *B, int block, int wg_per_cu);
/* Systems of at least min_rows rows are solved by a team of up to max_team
 * workgroups (row ranges split, team barriers between phases), as many as fit
 * in the workgroups resident at once; max_team 1 keeps one workgroup per
 * system.  Each team is queued on one XCD and its barriers go through that
 * XCD's L2 (DFM_BATCH_XCD=0: one queue, agent-scope barriers).  Default:
 * min_rows 16384, max_team 4. */
int dfm_batch_set_teams(dfm_batch *B, unsigned min_rows, int max_team);
/* One step for every system: ILU(0) numeric, then PCG from x = 0 on rhs
 * (concatenated, A's numbering) to ||r|| < tol ||rhs||.  precision:
 * DFM_PREC_FP64 (the reference's PCG) or DFM_PREC_MIXED_F32 (FP32 matrix,
 * factors and recurrence; FP64 solution, FP64 true residual replacing the
 * recursive one at every 10x residual drop, convergence on the FP64 true
 * residual).  Per system (host arrays, any may be NULL): iterations (FP64:
 * the reference's 0-based index at convergence; FP32: iterations run),
 * FP64 reliable updates, status
 * (1 converged, 0 not, -1 more than 10 singular pivots), FP64 relative
 * residual, and ms[2s] the solve time on its workgroup, ms[2s+1] the time
 * from the launch to its completion.  DFM_NOT_CONVERGED unless all converged. */
int dfm_batch_solve_pcg(dfm_batch *B, const double *rhs_dev, double *x_dev, double tol, unsigned maxit,
                        int precision, unsigned *iters, int *outer, int *status, double *relres, double *ms);
/* Factor of system sys after a solve, in its colour numbering: val in CRS
 * order (rows ascending, columns ascending; L unscaled, U scaled by D^-1),
 * dia = D^-1, order[new] = old (local rows). */
int dfm_batch_get_factor(const dfm_batch *B, unsigned sys, double *val, double *dia, int *order);

/* ---- Krylov (LsSol::Solve_PCG / Solve_PBiCGSTAB, solver_ls_iter.cpp:100-172, 291-408) */
/* Precision of a solve: the type of the Krylov vectors / recurrence, and the
 * storage of the matrix copy and of the ILU factors streamed each iteration.
 *   FP64             f64 vectors, f64 matrix, f64 factors (reference semantics)
 *   MIXED            f64 vectors, f32 matrix, f32 factors
 *   MIXED_BF16       f64 vectors, f32 matrix, bf16 factors
 *   MIXED_F32        f32 vectors, f32 matrix, f32 factors
 *   MIXED_F32_BF16   f32 vectors, f32 matrix, bf16 factors
 *   FAC32            f64 vectors, f64 matrix, f32 factors
 *   FAC16            f64 vectors, f64 matrix, bf16 factors
 * FAC modes keep the Krylov operator in FP64 and lower only the
 * preconditioner: the choice for systems too ill-conditioned for an FP32
 * matrix (cond(A) * 2^-24 near or above 1).
 * All modes other than FP64 accumulate dot products in f64 and run FP64 reliable
 * updates: when the recursive residual has dropped by reliable_delta, the
 * partial solution is added to the f64 solution and the true residual
 * b - A x is recomputed with the f64 matrix and replaces the recursive one.
 * Convergence is decided on that f64 true residual. */
enum {
    DFM_PREC_FP64 = 0,
    DFM_PREC_MIXED = 1,
    DFM_PREC_MIXED_BF16 = 2,
    DFM_PREC_MIXED_F32 = 3,
    DFM_PREC_MIXED_F32_BF16 = 4,
    DFM_PREC_FAC32 = 5,
    DFM_PREC_FAC16 = 6
};
typedef struct {
    int precision;         /* DFM_PREC_* (above) */
    unsigned check_every;  /* iterations between host convergence checks (0: default 16) */
    int want_true_relres;  /* nonzero: compute true_relres = ||b - A x|| / ||b|| in FP64 at the end */
    double reliable_delta; /* MIXED*: residual drop between FP64 reliable updates (0: default 0.1) */
    unsigned inner_iters;  /* out: Krylov iterations (MIXED*) */
    unsigned outer_iters;  /* out: FP64 reliable updates (true residual + solution update) (MIXED*) */
    double true_relres;    /* out: ||b - A x|| / ||b|| in FP64 (when want_true_relres, or MIXED) */
} dfm_solve_opts;
/* A x = r with x := 0 start.  r_dev: in = right-hand side (vector id -1), out = final
 * (recursive) residual; x_dev: out = solution (vector id -2).  conv_ratio: in = tolerance,
 * out = achieved ||r||/||r0||; iteration: in = max, out = as the reference reports.
 * opts may be NULL (FP64).  Returns DFM_OK (converged) or DFM_NOT_CONVERGED. */
int dfm_solve_pcg(dfm_bcrs *A, dfm_ilu *M, double *r_dev, double *x_dev,
                  double *conv_ratio, unsigned *iteration, dfm_solve_opts *opts);
/* How the solvers and the substitutions run the level sets of a triangular
 * sweep (no reference counterpart; a tuning knob): mode 0 one launch per
 * level; 1 (default) for factors of 8 or more levels one launch per phase
 * (tiles of dependent rows run level by level inside one workgroup) where
 * every tile fits 256 rows, else, in the solvers, one persistent launch.
 * Phases are the colours of a multicolour or block-multicolour ordering; a
 * factor in natural or given order has phases found in its own numbering
 * when they are estimated to beat its level sets (a mesh numbered by colours
 * of tiles, a small system: one 256-row tile per phase).  2 persistent
 * always; 3 per phase wherever possible (per level elsewhere); 4 as 3 with
 * one lane per block row of 2x2 / 3x3 blocks instead of one per tile row
 * (slower here; kept as an alternative).  grid = workgroups of a persistent
 * launch (0: twice the slices of the largest level, at least 64; at most
 * 4096).  FP64 sweeps give the same bits in every mode.  Process-wide;
 * initial values from DFM_PERSIST / DFM_PERSIST_GRID (read at factor
 * creation: DFM_TILE_ROWS, rows a tile is filled to, default 64, and
 * DFM_DISCOVER_PHASES=0, no phases found in natural or given order). */
int dfm_set_sweep_mode(int mode, int grid);
/* In-solve timing of the Krylov iteration's SpMV (no reference counterpart;
 * a measurement aid for the roofline figure).  on != 0 starts recording:
 * solves then launch their iterations directly instead of replaying graphs
 * (same kernels, same order) and every iteration SpMV gets a start / stop
 * event pair stamped at the kernel's own start and end (at most 65,536
 * launches); on == 0 stops.  dfm_get_kernel_timing returns the average
 * duration over all recorded launches and over the launches that did work
 * (at least half the median duration: launches after convergence inside a
 * block of iterations return at once), and both counts. */
int dfm_set_kernel_timing(int on);
int dfm_get_kernel_timing(double *avg_all_ms, double *avg_working_ms, int *nlaunch, int *nworking);
int dfm_solve_pbicgstab(dfm_bcrs *A, dfm_ilu *M, double *r_dev, double *x_dev,
                        double *conv_ratio, unsigned *iteration, dfm_solve_opts *opts);

/* ---- saved-stiffness residual (Fem::Ls::CLinearSystem_Save) --------------
 * The boundary matrix K_b of src/femls/linearsystem_fieldsave.cpp: a
 * MatVec::CMat_BlkCrs (include/delfem/matvec/mat_blkcrs.h:45-197) whose rows
 * are the block rows with a free dof and whose columns are the blocks with a
 * fixed dof, so that MakeResidual (linearsystem_fieldsave.cpp:206-247) is
 * r = f - K_b u without re-assembly.  Square matrices of square blocks. */
typedef struct dfm_bmat_s dfm_bmat;     /* MatVec::CMat_BlkCrs */
int dfm_bmat_create(unsigned nblk_col, unsigned len_col, unsigned nblk_row, unsigned len_row,
                    dfm_bmat **out);                                       /* CMat_BlkCrs(nblk_col,len_col,nblk_row,len_row) */
int dfm_bmat_destroy(dfm_bmat *M);
/* SetPatternBoundary(A, bc_col, bc_row) mat_blkcrs.cpp:1177-1265 (host flags, nblk*len each):
 * per row with a free dof (bc_col), its own block first, then its CRS neighbours, those with
 * a fixed dof (bc_row); values zero */
int dfm_bmat_set_pattern_boundary(dfm_bmat *M, const dfm_bcrs *A, const int *flag_col, const int *flag_row);
int dfm_bmat_info(const dfm_bmat *M, unsigned *nblk, unsigned *len, unsigned *ncrs);
int dfm_bmat_get_pattern(const dfm_bmat *M, unsigned *rowptr, unsigned *col);
int dfm_bmat_set_zero(dfm_bmat *M);                                        /* SetZero mat_blkcrs.cpp:227 */
int dfm_bmat_get_values(const dfm_bmat *M, double *val);                   /* ncrs*len*len, row-major blocks */
int dfm_bmat_set_values(dfm_bmat *M, const double *val);                   /* host values into the pattern (GetPtrValPSuP writes) mat_blkcrs.h:166-173 */
int dfm_bmat_mearge(dfm_bmat *M, unsigned nblkel, const unsigned *blkel,
                    unsigned blksize, const double *emat);                 /* Mearge (one square element, host emat) mat_blkcrs.cpp:754-828 */
/* The values of A (assembled, before its boundary conditions) at K_b's entries: equal to
 * merging the same element matrices into K_b (the assembly loops merge every element into
 * both, eqn_linear_solid3d.cpp:511-512); A must have the pattern K_b was made from. */
int dfm_bmat_copy_from(dfm_bmat *M, const dfm_bcrs *A);
int dfm_bmat_set_bc_inverse_col(dfm_bmat *M, const int *flag);             /* SetBoundaryConditionInverse_Colum mat_blkcrs.cpp:876-907 */
int dfm_bmat_set_bc_row(dfm_bmat *M, const int *flag);                     /* SetBoundaryCondition_Row mat_blkcrs.cpp:910-952 */
/* MatVec(..., isnt_trans=true) mat_blkcrs.cpp:955-1176, bitwise; blocks up to 3 x 3 through a
 * sliced-ELL copy kept from the first call, as dfm_bcrs_matvec */
int dfm_bmat_matvec(dfm_bmat *M, double alpha, const double *x_dev, double beta,
                    double *y_dev);
/* General block CRS (off-diagonal blocks of multi-segment systems): rows of nblk_col blocks
 * of len_col dofs, columns of nblk_row blocks of len_row dofs (dfm_bmat_create takes any).
 * AddPattern(CIndexedArray) of a new matrix (mat_blkcrs.cpp:286-370): rows of column blocks. */
int dfm_bmat_add_pattern(dfm_bmat *M, unsigned nrow, const unsigned *index, const unsigned *array);
int dfm_bmat_set_bc_col(dfm_bmat *M, const int *flag);                     /* SetBoundaryCondition_Colum mat_blkcrs.cpp:829-875 */
/* The block's values from a combined matrix A (b dofs per node, same nodes) assembled by the
 * same element loop, before its boundary conditions: entry (i, j)'s sub-block is A's block
 * (i, j) restricted to rows dofs_row[0..len_col) and columns dofs_col[0..len_row) -- the sums
 * Mearge of the split element matrices forms.  Every entry must be in A's pattern. */
int dfm_bmat_extract(dfm_bmat *M, const dfm_bcrs *A, const unsigned *dofs_row, const unsigned *dofs_col);
/* The same for a segment's diagonal block D (D has A's pattern, block size len(dofs)). */
int dfm_bcrs_extract(dfm_bcrs *D, const dfm_bcrs *A, const unsigned *dofs);
/* y[k] = x[idx[k]] for k < n (host idx): a combined vector regrouped into segments, or back */
int dfm_vec_gather(size_t n, const long long *idx, const double *x_dev, double *y_dev);
/* y[i] = x[i] where flag[i] == 0 (host flags): the field update of CLinearSystem_Save::
 * UpdateValueOfField (linearsystem_fieldsave.cpp:249-300), fixed dofs keep their values */
int dfm_vec_copy_free(size_t n, const int *flag, const double *x_dev, double *y_dev);

/* ---- multi-segment systems (LsSol::CLinearSystem with several segments) --
 * src/ls/linearsystem.cpp:138-490: segment i has nblk[i] blocks of len[i] dofs, a
 * diagonal block (dfm_bcrs) and off-diagonal blocks (dfm_bmat, rows of segment i,
 * columns of segment j); vectors are the segments' vectors one after the other
 * (device pointers).  Blocks are borrowed: they must outlive the handle. */
typedef struct dfm_mseg_s dfm_mseg;
typedef struct dfm_msilu_s dfm_msilu;
int dfm_mseg_create(unsigned nseg, const unsigned *nblk, const unsigned *len, dfm_mseg **out); /* AddLinSysSeg x nseg */
int dfm_mseg_destroy(dfm_mseg *S);
int dfm_mseg_set_dia(dfm_mseg *S, unsigned i, dfm_bcrs *A);                 /* AddMat_Dia :419-448 */
int dfm_mseg_set_nondia(dfm_mseg *S, unsigned i, unsigned j, dfm_bmat *B);  /* AddMat_NonDia :450-520 */
int dfm_mseg_matvec(dfm_mseg *S, double alpha, const double *x_dev, double beta,
                    double *y_dev);                                         /* MATVEC :250-296 */
int dfm_mseg_dot(dfm_mseg *S, const double *x_dev, const double *y_dev, double *result); /* DOT :335-364 */
/* CPreconditioner_ILU(ls, lev_fill) across the segments (preconditioner.cpp:57-206, with
 * SetFillInLevel(lev_fill) for every block): natural-order ILU of each diagonal block and a
 * CMatFrac_BlkCrs copy of each off-diagonal block, patterns extended by AddFracPtnLowUp /
 * AddFracPtnUp / AddFracPtnLow / AddFracPtn (matfrac_blkcrs.cpp:156-495,
 * matdiafrac_blkcrs.cpp:1105-1377) in the reference's block order; lev_fill 0 = no fill,
 * -1 = unlimited.  Fill that would need a block the system lacks is DFM_BAD_ARG. */
int dfm_mseg_ilu_create(dfm_mseg *S, int lev_fill, dfm_msilu **out);
/* The same with a fill level per block (i, j), lev[i * nseg + j] (-1 unlimited): what
 * CPreconditioner_ILU::SetFillInLevel(lev, ilss) requests resolve to (preconditioner.h:77,
 * preconditioner.cpp:108-131; a request sets segment ilss's block row and column, ilss -1
 * every block, later requests win; all zero and no fill blocks: no fill), and fill_blk:
 * SetFillBlk's rows of segment 0 (preconditioner.h:80, preconditioner.cpp:143-148), which
 * take fill of any level in block (0, 0) (nfill 0: none).  A missing block that the fill
 * would need stays out when its own level is 0 (the reference creates it empty). */
int dfm_mseg_ilu_create_levels(dfm_mseg *S, const int *lev, const unsigned *fill_blk, unsigned nfill,
                               dfm_msilu **out);
int dfm_mseg_ilu_destroy(dfm_msilu *M);
/* SetValue (:217-297): copies, then per block in (i, j) order DoILUDecompLowUp + DoILUDecomp
 * (diagonal), DoILUDecompUp / Low / LowUp (off-diagonal); DFM_SINGULAR as dfm_ilu_set_value */
int dfm_mseg_ilu_set_value(dfm_msilu *M, int *nsingular);
/* factor pattern of block (i, j): *ncrs blocks; rowptr (nblk_i + 1) and col (ncrs) may be NULL */
int dfm_mseg_ilu_get_pattern(const dfm_msilu *M, unsigned i, unsigned j, unsigned *ncrs, unsigned *rowptr,
                             unsigned *col);
/* factor values of block (i, j): diagonal as dfm_ilu_get_values, off-diagonal ncrs*len_i*len_j */
int dfm_mseg_ilu_get_values(const dfm_msilu *M, unsigned i, unsigned j, double *val, double *dia);
int dfm_mseg_ilu_precond(dfm_msilu *M, double *vec_dev);                    /* SolvePrecond :317-333 */
/* LsSol::Solve_PCG / Solve_PBiCGSTAB (solver_ls_iter.cpp:100-172, 291-408) on the segments:
 * r_dev in right-hand side, out residual; x_dev out.  DFM_OK converged, DFM_NOT_CONVERGED
 * at max_iter (Solve_PBiCGSTAB's true in both cases). */
int dfm_mseg_solve_pcg(dfm_mseg *S, dfm_msilu *M, double *r_dev, double *x_dev,
                       double *conv_ratio, unsigned *iteration);
int dfm_mseg_solve_pbicgstab(dfm_mseg *S, dfm_msilu *M, double *r_dev, double *x_dev,
                             double *conv_ratio, unsigned *iteration);
/* How the two solvers above run (no reference counterpart; a switch): on = 1
 * (default) device-driven, the scalars formed on the device and the stop flag
 * polled every four iterations; 0 host-driven, one synchronisation per DOT.
 * Both give the same bits.  Process-wide; initial value from DFM_MSEG_LOOP
 * ("host" for 0). */
int dfm_mseg_set_device_loop(int on);

/* ---- complex twin (scalar blocks, as the reference's complex classes) ----
 * Complex values are interleaved (re, im) doubles: the layout of
 * Com::Complex and of std::complex<double>.  Every complex operation keeps
 * the operation order of Com::Complex (include/delfem/complex.h:82-160), so
 * assembly, MatVec, factorisation and substitutions match the reference bit
 * for bit in natural ordering. */
typedef struct dfm_zbcrs_s dfm_zbcrs;   /* MatVec::CZMatDia_BlkCrs */
typedef struct dfm_zilu_s dfm_zilu;     /* MatVec::CZMatDiaFrac_BlkCrs */
int dfm_zbcrs_create(unsigned nblk, unsigned len, dfm_zbcrs **out);        /* CZMatDia_BlkCrs(nblk,len) zmatdia_blkcrs.cpp:44-49; len must be 1 */
int dfm_zbcrs_destroy(dfm_zbcrs *A);
int dfm_zbcrs_add_pattern(dfm_zbcrs *A, unsigned nrow, const unsigned *index,
                          const unsigned *array);                          /* AddPattern(CIndexedArray) zmatdia_blkcrs.cpp:406-551 */
int dfm_zbcrs_info(const dfm_zbcrs *A, unsigned *nblk, unsigned *len, unsigned *ncrs);
int dfm_zbcrs_get_pattern(const dfm_zbcrs *A, unsigned *rowptr, unsigned *col);
int dfm_zbcrs_set_zero(dfm_zbcrs *A);                                      /* SetZero zmatdia_blkcrs.cpp:185-190 */
int dfm_zbcrs_set_values(dfm_zbcrs *A, const double *val, const double *dia); /* host values: val 2*ncrs, dia 2*nblk doubles */
int dfm_zbcrs_get_values(const dfm_zbcrs *A, double *val, double *dia);
int dfm_zbcrs_device_values(dfm_zbcrs *A, double **val_dev, double **dia_dev);
int dfm_zbcrs_mearge(dfm_zbcrs *A, unsigned nblkel, const unsigned *blkel,
                     unsigned blksize, const double *emat);                /* Mearge (one element, host emat nblkel^2 complex) zmatdia_blkcrs.cpp:198-237 */
int dfm_zbcrs_set_bc(dfm_zbcrs *A, const int *flag);                       /* SetBoundaryCondition zmatdia_blkcrs.cpp:239-268 (host flags) */
int dfm_zbcrs_matvec(dfm_zbcrs *A, double alpha, const double *x_dev, double beta,
                     double *y_dev);                                       /* MatVec zmatdia_blkcrs.cpp:737-770 */
int dfm_zbcrs_matvec_hermitian(dfm_zbcrs *A, double alpha, const double *x_dev, double beta,
                               double *y_dev);                             /* MatVec_Hermitian zmatdia_blkcrs.cpp:803-847: alpha A^H x + beta y */
int dfm_zvec_dot(size_t n, const double *x_dev, const double *y_dev, double *result); /* CZVector_Blk operator* (no conjugate) zvector_blk.cpp:39-60; result[2] */
int dfm_zvec_norm2(size_t n, const double *x_dev, double *result);                    /* GetSquaredVectorNorm zvector_blk.cpp:154-161 */
int dfm_zvec_inner(size_t n, const double *x_dev, const double *y_dev, double *result); /* InnerProduct (sum conj(x) y) zvector_blk.cpp:62-80; result[2] */
int dfm_zvec_zero_bc(size_t n, const int *flag, double *x_dev);                        /* CBCFlag::SetZeroToBCDof on a complex vector */
/* Element assembly (src/femeqn/eqn_helmholtz.cpp AddLinSys_Helmholtz loops):
 * plans made on the complex matrix's pattern add element rows in element
 * order per entry, as sequential Mearge calls do. */
int dfm_zasm_create(dfm_zbcrs *A, unsigned nnode, unsigned nelem, unsigned npe,
                    const unsigned *conn, dfm_asm **out);
/* Helmholtz P1 triangle, eqn_helmholtz.cpp:51-133 (k = 2*3.1416/wave_length);
 * res_dev (may be NULL) += -emat value. */
int dfm_zassemble_helmholtz2d_p1(dfm_zbcrs *A, dfm_asm *p, const double *coord_dev, double wave_length,
                                 const double *value_dev, double *res_dev);
/* Sommerfeld radiation line element, eqn_helmholtz.cpp:437-497. */
int dfm_zassemble_sommerfeld_b1(dfm_zbcrs *A, dfm_asm *p, const double *coord_dev, double wave_length,
                                const double *value_dev, double *res_dev);
/* Symbolic ILU(lev_fill): CZMatDiaFrac_BlkCrs(lev,A) zmatdiafrac_blkcrs.cpp:77-108 + AddFracPtn
 * :263-432; orderings as dfm_ilu_create (NATURAL is the reference's). */
int dfm_zilu_create(const dfm_zbcrs *A, int lev_fill, int ordering, const int *order, dfm_zilu **out);
int dfm_zilu_destroy(dfm_zilu *M);
int dfm_zilu_info(const dfm_zilu *M, unsigned *ncrs, int *nlevel_lower, int *nlevel_upper, int *ncolour);
int dfm_zilu_get_pattern(const dfm_zilu *M, unsigned *rowptr, unsigned *col, int *lev,
                         unsigned *diaind, int *order);
/* SetValue + DoILUDecomp zmatdiafrac_blkcrs.cpp:466-486, 192-262.  *nsmall (may be NULL):
 * pivots with |d|^2 < 1e-20 (the reference reports them and divides anyway). */
int dfm_zilu_set_value(dfm_zilu *M, dfm_zbcrs *A, int *nsmall);
int dfm_zilu_get_values(const dfm_zilu *M, double *val, double *dia);
/* ForwardSubstitution / BackwardSubstitution zmatdiafrac_blkcrs.cpp:116-157 (factor numbering) */
int dfm_zilu_forward(dfm_zilu *M, double *vec_dev);
int dfm_zilu_backward(dfm_zilu *M, double *vec_dev);
/* Solve / SolvePrecond zmatdiafrac_blkcrs.cpp:164-170 on a vector in the original numbering */
int dfm_zilu_solve_precond(dfm_zilu *M, double *vec_dev);
/* MatVec::Sol::Solve_PCOCG zsolver_mat_iter.cpp:303-392 with the factor as
 * preconditioner.  r_dev: in right-hand side, out final residual; u_dev: out
 * solution (without the last step's update, as the reference).  conv_ratio: in
 * tolerance, out sqrt(||r||^2/||r0||^2) at convergence (unchanged at max_iter);
 * iteration: in max_iter, out iterations.  Returns DFM_OK in both cases, as
 * the reference returns true. */
int dfm_zsolve_pcocg(dfm_zbcrs *A, dfm_zilu *M, double *r_dev, double *u_dev,
                     double *conv_ratio, unsigned *iteration);
/* MatVec::Sol::Solve_CGNR zsolver_mat_iter.cpp:394-483: CG on A^H A u = A^H b (any nonsingular
 * complex A).  r_dev in right-hand side, out final residual; u_dev out; conv_ratio / iteration
 * as dfm_zsolve_pcocg (unchanged tolerance and max_iter when not reached).  Returns DFM_OK. */
int dfm_zsolve_cgnr(dfm_zbcrs *A, double *r_dev, double *u_dev, double *conv_ratio, unsigned *iteration);
/* MatVec::Sol::Solve_CG / Solve_PCG of the complex classes, zsolver_mat_iter.cpp:32-119 and
 * 121-218: CG for a Hermitian matrix (real part of the conjugated inner product); Solve_PCG
 * with the factor's SolvePrecond.  Solve_CG returns DFM_OK also at max_iter (conv_ratio
 * unchanged); Solve_PCG returns DFM_NOT_CONVERGED there (the reference's false), its
 * iteration count starts at 0 as the reference's. */
int dfm_zsolve_cg(dfm_zbcrs *A, double *r_dev, double *u_dev, double *conv_ratio, unsigned *iteration);
int dfm_zsolve_pcg(dfm_zbcrs *A, dfm_zilu *M, double *r_dev, double *u_dev, double *conv_ratio,
                   unsigned *iteration);
/* MatVec::Sol::Solve_COCG zsolver_mat_iter.cpp:220-299 (unpreconditioned PCOCG). */
int dfm_zsolve_cocg(dfm_zbcrs *A, double *r_dev, double *u_dev, double *conv_ratio, unsigned *iteration);
/* MatVec::Sol::Solve_BiCGSTAB zsolver_mat_iter.cpp:487-616 and Solve_BiCGStabP :622-739 (with
 * the factor).  As the reference: the inner products are InnerProduct (sum conj(a) b,
 * zvector_blk.cpp:62-80), so alpha and omega are the conjugates of the textbook ones and the
 * pair converges only on real-valued data; BiCGStabP starts from r = conj(b) with shadow
 * residual b and leaves the last ratio in conv_ratio also at max_iter.  Return DFM_OK. */
int dfm_zsolve_bicgstab(dfm_zbcrs *A, double *r_dev, double *u_dev, double *conv_ratio, unsigned *iteration);
int dfm_zsolve_bicgstabp(dfm_zbcrs *A, dfm_zilu *M, double *r_dev, double *u_dev,
                         double *conv_ratio, unsigned *iteration);

/* ---- variable block sizes ("flex") ---------------------------------------
 * MatVec::CMatDia_BlkCrs / CMatDiaFrac_BlkCrs made with per-block lengths aLen
 * (blocks of 1..32 dofs; the reference's rigid bodies and constraints) and the
 * LsSol solvers over one such segment (CLinearSystem::AddLinSysSeg(nnode, aLen),
 * linearsystem.cpp:170-201).  Layout as the reference's: dofptr (m_DofPtrCol),
 * the block pattern in CRS without the diagonal, valptr (m_ValPtr, len_i x len_j
 * row-major blocks in CRS order), diaptr (m_DiaValPtr).  Vectors are flat device
 * arrays of dofptr[nblk] values.  FP64, bitwise the reference's flex branches. */
typedef struct dfm_fbcrs_s dfm_fbcrs;   /* CMatDia_BlkCrs with per-block sizes */
typedef struct dfm_filu_s dfm_filu;     /* CMatDiaFrac_BlkCrs over a flex matrix */
int dfm_fbcrs_create(unsigned nblk, const unsigned *alen, dfm_fbcrs **out); /* CMatDia_BlkCrs(nblk,aLen,nblk,aLen) matdia_blkcrs.cpp:70-104 */
int dfm_fbcrs_destroy(dfm_fbcrs *A);
int dfm_fbcrs_add_pattern(dfm_fbcrs *A, unsigned nrow, const unsigned *index,
                          const unsigned *array);                         /* AddPattern(CIndexedArray) matdia_blkcrs.cpp:346-505; values zeroed */
int dfm_fbcrs_info(const dfm_fbcrs *A, unsigned *nblk, unsigned *ndof, unsigned *ncrs,
                   unsigned *nval, unsigned *ndia);                       /* NBlkMatCol / sizes (mat_blkcrs.h:83-157) */
int dfm_fbcrs_layout(const dfm_fbcrs *A, unsigned *dofptr, unsigned *rowptr, unsigned *col,
                     unsigned *valptr, unsigned *diaptr);                 /* m_DofPtrCol, pattern, m_ValPtr (mat_blkcrs.cpp:236-249), m_DiaValPtr */
int dfm_fbcrs_set_zero(dfm_fbcrs *A);                                      /* SetZero mat_blkcrs.cpp:225-258, matdia_blkcrs.cpp:139-163 */
int dfm_fbcrs_set_values(dfm_fbcrs *A, const double *val, const double *dia);
int dfm_fbcrs_get_values(const dfm_fbcrs *A, double *val, double *dia);
int dfm_fbcrs_device_values(dfm_fbcrs *A, double **val_dev, double **dia_dev);
int dfm_fbcrs_mearge(dfm_fbcrs *A, unsigned n, const unsigned *iblk, const unsigned *jblk,
                     const double *emat);                                 /* n calls of Mearge(1,&i,1,&j,...) matdia_blkcrs.cpp:173-207, in order */
int dfm_fbcrs_set_bc(dfm_fbcrs *A, const int *flag);                       /* SetBoundaryCondition, flex branch matdia_blkcrs.cpp:268-306 */
int dfm_fbcrs_matvec(dfm_fbcrs *A, double alpha, const double *x_dev, double beta,
                     double *y_dev);                                      /* MatVec, flex branch matdia_blkcrs.cpp:931-960 */
int dfm_filu_create(const dfm_fbcrs *A, int lev_fill, dfm_filu **out);     /* CMatDiaFrac_BlkCrs(lev,A) matdiafrac_blkcrs.cpp:227-287 */
int dfm_filu_destroy(dfm_filu *M);
int dfm_filu_info(const dfm_filu *M, unsigned *ncrs, unsigned *nval, int *nlevel_lower, int *nlevel_upper);
int dfm_filu_get_pattern(const dfm_filu *M, unsigned *rowptr, unsigned *col, unsigned *diaind,
                         unsigned *valptr);
/* SetValue_Initialize + DoILUDecomp (flex branch matdiafrac_blkcrs.cpp:777-846):
 * DFM_SINGULAR after more than 10 singular pivots, the factor then left as the
 * reference leaves it (stopped in the row of the 11th). */
int dfm_filu_set_value(dfm_filu *M, dfm_fbcrs *A, int *nsingular);
int dfm_filu_get_values(const dfm_filu *M, double *val, double *dia);
int dfm_filu_forward(dfm_filu *M, double *vec_dev);                        /* ForwardSubstitution matdiafrac_blkcrs.cpp:339-379 */
int dfm_filu_backward(dfm_filu *M, double *vec_dev);                       /* BackwardSubstitution matdiafrac_blkcrs.cpp:512-553 */
int dfm_filu_solve_precond(dfm_filu *M, double *vec_dev);                  /* SolvePrecond preconditioner.cpp:301-335 */
/* LsSol solvers over one flex segment: kind 0 Solve_PCG (solver_ls_iter.cpp:100-172),
 * 1 Solve_PBiCGSTAB (:291-408) with the factor M, 2 Solve_CG (:31-97), 3 Solve_BiCGSTAB
 * (:177-286) without (M may be NULL).  r in: right-hand side, out: final residual; x out.
 * DFM_OK when converged, DFM_NOT_CONVERGED at max_iter (the reference's Solve_PCG returns
 * false there, the other three true). */
int dfm_fsolve(int kind, dfm_fbcrs *A, dfm_filu *M, double *r_dev, double *x_dev,
               double *conv_ratio, unsigned *iteration);

/* ---- measurement (no reference counterpart) -------------------------------
 * Times one device kernel of a solve on its real operands: which = 0 the
 * sliced-ELL SpMV of A, 1 the forward sweep (all lower level sets), 2 the
 * backward sweep.  precision selects the FP64 or FP32 copies.  Runs nrep
 * back-to-back launches between HIP events on the library stream; returns the
 * average milliseconds per launch and the algorithmic HBM bytes of one launch
 * (values + indices + one pass over each vector operand). */
int dfm_bench_kernel(dfm_bcrs *A, dfm_ilu *M, int which, int precision, int nrep,
                     double *avg_ms, double *bytes);
/* Device-to-device copy of `bytes` (16-byte vector loads and stores) with the
 * library's launch shape: the practical bandwidth ceiling of a streaming kernel. */
int dfm_bench_copy(size_t bytes, int nrep, double *avg_ms);
/* Streaming read of `bytes` (non-temporal 16-byte loads, eight in flight per
 * lane, no write stream): the practical read-bandwidth ceiling that SpMV and the
 * triangular sweeps, which read far more than they write, are held against. */
int dfm_bench_read(size_t bytes, int nrep, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
