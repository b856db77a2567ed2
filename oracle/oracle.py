"""ctypes access to the checkers -- TEST INFRASTRUCTURE ONLY.

  Orc  -- liborc.so, the C restatement of the reference hot path (dfm_oracle.c)
  Ref  -- _ref/libdfmref.so, the DelFEM reference's own sources compiled by
          oracle/Makefile, driven through ref_driver.cpp (exists only where
          the DelFEM tree was available at build time)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.  The product (delfem_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libdfmref.so")
REF_TREE = os.environ.get("DELFEM_REF", "/root/reference")

vp = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build(ref=True):
    """Compile liborc.so (always) and the reference library (when the tree exists)."""
    targets = ["oracle"]
    if ref and os.path.isdir(os.path.join(REF_TREE, "src", "matvec")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, f"REF={REF_TREE}", "-j8", *targets], check=True)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Orc:
    """The C restatement.  Methods mirror dfm_oracle.h with numpy arrays."""

    def __init__(self, path=ORC_PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run oracle.build()")
        L = C.CDLL(path)
        L.orc_fem_pattern.restype = C.c_long
        L.orc_fem_pattern.argtypes = [C.c_uint, C.c_uint, C.c_uint, vp, vp, vp, C.c_long]
        L.orc_assemble.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, C.c_uint, C.c_uint, vp, vp]
        L.orc_assemble_vec.argtypes = [C.c_uint, vp, C.c_uint, C.c_uint, vp, vp]
        L.orc_set_bc.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp]
        L.orc_zero_bc_dof.argtypes = [C.c_uint, vp, vp]
        L.orc_matvec.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_iluk_symbolic.restype = C.c_long
        L.orc_iluk_symbolic.argtypes = [C.c_uint, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_long]
        L.orc_iluk_symbolic_blk.restype = C.c_long
        L.orc_iluk_symbolic_blk.argtypes = [C.c_uint, vp, vp, C.c_int, vp, C.c_uint, vp, vp, vp, vp, C.c_long]
        L.orc_ilu_setvalue.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_ilu_decomp.restype = C.c_int
        L.orc_ilu_decomp.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp]
        L.orc_ilu_forward.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp]
        L.orc_ilu_backward.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp]
        for f in (L.orc_solve_pcg, L.orc_solve_pbicgstab):
            f.restype = C.c_int
            f.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                          C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_solve_ls.restype = C.c_int
        L.orc_solve_ls.argtypes = [C.c_int, C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_uint)]
        L.orc_solve_mat.restype = C.c_int
        L.orc_solve_mat.argtypes = [C.c_int, C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zassemble.argtypes = [C.c_uint, vp, vp, C.c_uint, C.c_uint, vp, vp, vp, vp]
        L.orc_zassemble_vec.argtypes = [vp, C.c_uint, C.c_uint, vp, vp]
        L.orc_zset_bc.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp]
        L.orc_zmatvec.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_zilu_setvalue.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_zilu_decomp.restype = C.c_int
        L.orc_zilu_decomp.argtypes = [C.c_uint, vp, vp, vp, vp, vp]
        L.orc_zilu_forward.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp]
        L.orc_zilu_backward.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp]
        L.orc_zsolve_pcocg.restype = C.c_int
        L.orc_zsolve_pcocg.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_bmat_pattern_boundary.restype = C.c_uint
        L.orc_bmat_pattern_boundary.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp, vp, vp]
        L.orc_bmat_mearge.argtypes = [C.c_uint, C.c_uint, vp, vp, C.c_uint, C.c_uint, vp, vp, vp]
        L.orc_bmat_bc_inverse_col.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp]
        L.orc_bmat_bc_row.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, vp]
        L.orc_bmat_matvec.argtypes = [C.c_uint, C.c_uint, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_bmat2_matvec.argtypes = [C.c_uint] * 4 + [vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_bmat2_mearge.argtypes = [C.c_uint] * 4 + [vp, vp, C.c_uint, C.c_uint, vp, C.c_uint, vp, vp, vp]
        L.orc_bmat2_bc.argtypes = [C.c_uint] * 4 + [vp, vp, vp, vp, vp]
        L.orc_ms_create.restype = vp
        L.orc_ms_create.argtypes = [C.c_uint, vp, vp]
        L.orc_ms_free.argtypes = [vp]
        L.orc_ms_set_dia.argtypes = [vp, C.c_uint, vp, vp, vp, vp]
        L.orc_ms_set_nondia.argtypes = [vp, C.c_uint, C.c_uint, vp, vp, vp]
        L.orc_ms_matvec.argtypes = [vp, C.c_double, vp, C.c_double, vp]
        L.orc_ms_ilu.restype = C.c_int
        L.orc_ms_ilu.argtypes = [vp, C.c_int]
        L.orc_ms_ilu_levels.restype = C.c_int
        L.orc_ms_ilu_levels.argtypes = [vp, vp, vp, C.c_uint]
        L.orc_ms_factor_ncrs.restype = C.c_uint
        L.orc_ms_factor_ncrs.argtypes = [vp, C.c_uint, C.c_uint]
        L.orc_ms_get_pattern.argtypes = [vp, C.c_uint, C.c_uint, vp, vp]
        L.orc_ms_get_factor.argtypes = [vp, C.c_uint, C.c_uint, vp, vp]
        L.orc_ms_precond.argtypes = [vp, vp]
        L.orc_ms_solve.restype = C.c_int
        L.orc_ms_solve.argtypes = [vp, C.c_int, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zmatvec_hermitian.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_zsolve_cg_herm.restype = C.c_int
        L.orc_zsolve_cg_herm.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp,
                                         C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zsolve_cgnr.restype = C.c_int
        L.orc_zsolve_cgnr.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zsolve_cocg.restype = C.c_int
        L.orc_zsolve_cocg.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zsolve_bicgstab.restype = C.c_int
        L.orc_zsolve_bicgstab.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp,
                                          C.POINTER(C.c_double), C.POINTER(C.c_uint)]
        L.orc_zemat_helmholtz2d_p1.argtypes = [C.c_uint, vp, vp, C.c_double, vp, vp, vp]
        L.orc_zemat_sommerfeld_b1.argtypes = [C.c_uint, vp, vp, C.c_double, vp, vp, vp]
        L.orc_emat_poisson2d_p1.argtypes = [C.c_uint, vp, vp, vp, C.c_double, vp, vp, vp]
        L.orc_emat_solid3d_p1.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, vp, vp, vp]
        L.orc_emat_oseen2d_p1p1.argtypes = [C.c_uint, vp, vp, C.c_double, C.c_double, vp, vp, vp, vp]
        L.orc_emat_helmholtz2d_p1.argtypes = [C.c_uint, vp, vp, vp, vp, C.c_double, vp, vp, vp]
        L.orc_greedy_colour.restype = C.c_int
        L.orc_greedy_colour.argtypes = [C.c_uint, vp, vp, vp]
        L.orc_colour_order.argtypes = [C.c_uint, vp, C.c_int, vp]
        L.orc_bmc_order.restype = C.c_int
        L.orc_bmc_order.argtypes = [C.c_uint, vp, vp, C.c_uint, vp, vp]
        L.orc_level_sets.restype = C.c_int
        L.orc_level_sets.argtypes = [C.c_uint, vp, vp, vp, C.c_int, vp]
        L.orc_fx_layout.argtypes = [C.c_uint, vp, vp, vp, vp, vp]
        L.orc_fx_matvec.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_fx_set_bc.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fx_mearge.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, C.c_uint, vp, vp, vp]
        L.orc_fx_ilu_setvalue.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fx_ilu_decomp.restype = C.c_int
        L.orc_fx_ilu_decomp.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fx_ilu_forward.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fx_ilu_backward.argtypes = [C.c_uint, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fx_solve.restype = C.c_int
        L.orc_fx_solve.argtypes = [C.c_int, C.c_uint] + [vp] * 13 + [vp, vp, C.POINTER(C.c_double),
                                                                   C.POINTER(C.c_uint)]
        self.L = L

    # -- patterns ---------------------------------------------------------
    def fem_pattern(self, nnode, conn):
        conn = u32(conn)
        nelem, npe = conn.shape
        rowptr = np.zeros(nnode + 1, dtype=np.uint32)
        n = self.L.orc_fem_pattern(nnode, nelem, npe, _p(conn), _p(rowptr), None, 0)
        col = np.zeros(max(1, n), dtype=np.uint32)
        self.L.orc_fem_pattern(nnode, nelem, npe, _p(conn), _p(rowptr), _p(col), n)
        return rowptr, col[:n]

    def iluk_symbolic(self, rowptr, col, lev, order=None, fill_blk=None):
        """ILU(lev) pattern; fill_blk (natural order only): the block rows and
        columns that take fill of any level (AddFracPtn(lev, aBlkFill))."""
        rowptr, col = u32(rowptr), u32(col)
        nb = rowptr.size - 1
        order = None if order is None else i32(order)
        blk = None if fill_blk is None else u32(fill_blk)
        assert blk is None or order is None
        cap = max(16, col.size * 4)
        while True:
            frow = np.zeros(nb + 1, dtype=np.uint32)
            fcol = np.zeros(cap, dtype=np.uint32)
            flev = np.zeros(cap, dtype=np.int32)
            fdia = np.zeros(max(1, nb), dtype=np.uint32)
            if blk is None:
                n = self.L.orc_iluk_symbolic(nb, _p(rowptr), _p(col), lev, _p(order), _p(frow), _p(fcol), _p(flev),
                                             _p(fdia), cap)
            else:
                n = self.L.orc_iluk_symbolic_blk(nb, _p(rowptr), _p(col), lev, _p(blk), blk.size, _p(frow), _p(fcol),
                                                 _p(flev), _p(fdia), cap)
            if n == -1:
                raise ValueError("unsorted pattern row or fill block out of range")
            if n >= 0:
                return dict(rowptr=frow, col=fcol[:n], lev=flev[:n], diaind=fdia[:nb])
            cap = -n

    # -- assembly ---------------------------------------------------------
    def assemble(self, nb, b, rowptr, col, conn, emat, val=None, dia=None):
        rowptr, col, conn, emat = u32(rowptr), u32(col), u32(conn), f64(emat)
        val = np.zeros(col.size * b * b) if val is None else f64(val).copy()
        dia = np.zeros(nb * b * b) if dia is None else f64(dia).copy()
        self.L.orc_assemble(nb, b, _p(rowptr), _p(col), _p(val), _p(dia), conn.shape[0], conn.shape[1], _p(conn),
                            _p(emat))
        return val, dia

    def assemble_vec(self, nb, b, conn, eres, vec=None):
        conn, eres = u32(conn), f64(eres)
        vec = np.zeros(nb * b) if vec is None else f64(vec).copy()
        self.L.orc_assemble_vec(b, _p(vec), conn.shape[0], conn.shape[1], _p(conn), _p(eres))
        return vec

    def set_bc(self, nb, b, rowptr, col, val, dia, flag):
        val, dia = f64(val).copy(), f64(dia).copy()
        self.L.orc_set_bc(nb, b, _p(u32(rowptr)), _p(u32(col)), _p(val), _p(dia), _p(i32(flag)))
        return val, dia

    def zero_bc(self, flag, vec):
        vec = f64(vec).copy()
        self.L.orc_zero_bc_dof(vec.size, _p(i32(flag)), _p(vec))
        return vec

    def matvec(self, nb, b, rowptr, col, val, dia, alpha, x, beta, y):
        y = f64(y).copy()
        self.L.orc_matvec(nb, b, _p(u32(rowptr)), _p(u32(col)), _p(f64(val)), _p(f64(dia)), alpha, _p(f64(x)), beta,
                          _p(y))
        return y

    # -- ILU ----------------------------------------------------------------
    def ilu_values(self, nb, b, rowptr, col, val, dia, F, order=None):
        """SetValue_Initialize + DoILUDecomp; returns (fval, fdia, nsing)."""
        nf = F["col"].size
        fval = np.zeros(nf * b * b)
        fdia = np.zeros(nb * b * b)
        order = None if order is None else i32(order)
        self.L.orc_ilu_setvalue(nb, b, _p(u32(rowptr)), _p(u32(col)), _p(f64(val)), _p(f64(dia)), _p(order),
                                _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(fval), _p(fdia))
        ns = self.L.orc_ilu_decomp(nb, b, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])), _p(fval),
                                   _p(fdia))
        return fval, fdia, ns

    def forward(self, nb, b, F, fval, fdia, vec):
        vec = f64(vec).copy()
        self.L.orc_ilu_forward(nb, b, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])), _p(f64(fval)),
                               _p(f64(fdia)), _p(vec))
        return vec

    def backward(self, nb, b, F, fval, fdia, vec):
        vec = f64(vec).copy()
        self.L.orc_ilu_backward(nb, b, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])),
                                _p(f64(fval)), _p(f64(fdia)), _p(vec))
        return vec

    def solve(self, kind, nb, b, rowptr, col, val, dia, F, fval, fdia, rhs, tol, maxit, order=None):
        """Solve_PCG / Solve_PBiCGSTAB: returns (ok, conv, iter, x, r_final)."""
        r = f64(rhs).copy()
        x = np.zeros(nb * b)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        order = None if order is None else i32(order)
        f = self.L.orc_solve_pcg if kind == "pcg" else self.L.orc_solve_pbicgstab
        ok = f(nb, b, _p(u32(rowptr)), _p(u32(col)), _p(f64(val)), _p(f64(dia)), _p(u32(F["rowptr"])),
               _p(u32(F["col"])), _p(u32(F["diaind"])), _p(f64(fval)), _p(f64(fdia)), _p(order), _p(r), _p(x),
               C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r

    # -- variable block sizes (flex) -------------------------------------------
    def fx_layout(self, dofptr, rowptr, col):
        """(valptr, diaptr): value offsets of the CRS and diagonal blocks."""
        dofptr, rowptr, col = u32(dofptr), u32(rowptr), u32(col)
        nb = rowptr.size - 1
        valptr = np.zeros(col.size + 1, dtype=np.uint32)
        diaptr = np.zeros(nb + 1, dtype=np.uint32)
        self.L.orc_fx_layout(nb, _p(dofptr), _p(rowptr), _p(col), _p(valptr), _p(diaptr))
        return valptr, diaptr

    def fx_mearge(self, M, blocks, val=None, dia=None):
        """Mearge of blocks [(i, j, values)] in order into the flex matrix M
        (dict dofptr, rowptr, col, valptr, diaptr); returns (val, dia)."""
        val = np.zeros(int(M["valptr"][-1])) if val is None else f64(val).copy()
        dia = np.zeros(int(M["diaptr"][-1])) if dia is None else f64(dia).copy()
        ib = u32([b[0] for b in blocks])
        jb = u32([b[1] for b in blocks])
        em = f64(np.concatenate([np.ravel(b[2]) for b in blocks])) if blocks else np.zeros(1)
        self.L.orc_fx_mearge(M["rowptr"].size - 1, _p(u32(M["dofptr"])), _p(u32(M["rowptr"])),
                             _p(u32(M["col"])), _p(u32(M["valptr"])), _p(val), _p(u32(M["diaptr"])), _p(dia),
                             ib.size, _p(ib), _p(jb), _p(em))
        return val, dia

    def fx_set_bc(self, M, val, dia, flag):
        val, dia = f64(val).copy(), f64(dia).copy()
        self.L.orc_fx_set_bc(M["rowptr"].size - 1, _p(u32(M["dofptr"])), _p(u32(M["rowptr"])), _p(u32(M["col"])),
                             _p(u32(M["valptr"])), _p(val), _p(u32(M["diaptr"])), _p(dia), _p(i32(flag)))
        return val, dia

    def fx_matvec(self, M, val, dia, alpha, x, beta, y):
        y = f64(y).copy()
        self.L.orc_fx_matvec(M["rowptr"].size - 1, _p(u32(M["dofptr"])), _p(u32(M["rowptr"])), _p(u32(M["col"])),
                             _p(u32(M["valptr"])), _p(f64(val)), _p(u32(M["diaptr"])), _p(f64(dia)), alpha,
                             _p(f64(x)), beta, _p(y))
        return y

    def fx_factor(self, M, lev):
        """The ILU(lev) pattern of the flex matrix with its value offsets."""
        F = self.iluk_symbolic(M["rowptr"], M["col"], lev)
        F["dofptr"] = u32(M["dofptr"])
        F["valptr"], F["diaptr"] = self.fx_layout(M["dofptr"], F["rowptr"], F["col"])
        return F

    def fx_ilu_values(self, M, val, dia, F):
        """SetValue_Initialize + DoILUDecomp (flex): (fval, fdia, ok)."""
        nb = M["rowptr"].size - 1
        fval = np.zeros(max(1, int(F["valptr"][-1])))
        fdia = np.zeros(max(1, int(F["diaptr"][-1])))
        self.L.orc_fx_ilu_setvalue(nb, _p(u32(M["dofptr"])), _p(u32(M["rowptr"])), _p(u32(M["col"])),
                                   _p(u32(M["valptr"])), _p(f64(val)), _p(u32(M["diaptr"])), _p(f64(dia)),
                                   _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["valptr"])), _p(fval),
                                   _p(fdia))
        ok = self.L.orc_fx_ilu_decomp(nb, _p(u32(F["dofptr"])), _p(u32(F["rowptr"])), _p(u32(F["col"])),
                                      _p(u32(F["diaind"])), _p(u32(F["valptr"])), _p(fval), _p(u32(F["diaptr"])),
                                      _p(fdia))
        return fval, fdia, bool(ok)

    def fx_forward(self, F, fval, fdia, vec):
        vec = f64(vec).copy()
        self.L.orc_fx_ilu_forward(F["rowptr"].size - 1, _p(u32(F["dofptr"])), _p(u32(F["rowptr"])),
                                  _p(u32(F["col"])), _p(u32(F["diaind"])), _p(u32(F["valptr"])), _p(f64(fval)),
                                  _p(u32(F["diaptr"])), _p(f64(fdia)), _p(vec))
        return vec

    def fx_backward(self, F, fval, vec):
        vec = f64(vec).copy()
        self.L.orc_fx_ilu_backward(F["rowptr"].size - 1, _p(u32(F["dofptr"])), _p(u32(F["rowptr"])),
                                   _p(u32(F["col"])), _p(u32(F["diaind"])), _p(u32(F["valptr"])), _p(f64(fval)),
                                   _p(vec))
        return vec

    def fx_solve(self, kind, M, val, dia, F, fval, fdia, rhs, tol, maxit):
        """LsSol solvers on the flex system, kind 'pcg', 'pbicgstab', 'cg' or
        'bicgstab': (ok, conv, iter, x, r_final)."""
        k = {"pcg": 0, "pbicgstab": 1, "cg": 2, "bicgstab": 3}[kind]
        r = f64(rhs).copy()
        x = np.zeros(r.size)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        if F is None:
            fr = fc = fd = fv = None
        else:
            fr, fc, fd, fv = u32(F["rowptr"]), u32(F["col"]), u32(F["diaind"]), u32(F["valptr"])
        ok = self.L.orc_fx_solve(k, M["rowptr"].size - 1, _p(u32(M["dofptr"])), _p(u32(M["rowptr"])),
                                 _p(u32(M["col"])), _p(u32(M["valptr"])), _p(f64(val)), _p(u32(M["diaptr"])),
                                 _p(f64(dia)), _p(fr), _p(fc), _p(fd), _p(fv), _p(None if fval is None else f64(fval)),
                                 _p(None if fdia is None else f64(fdia)), _p(r), _p(x), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r

    # -- complex twin (len 1; complex128 arrays) ------------------------------
    def zemat_helmholtz2d(self, coord, conn, wave_length, value):
        conn = u32(conn)
        ne = conn.shape[0]
        emat = np.zeros(ne * 9, dtype=np.complex128)
        eres = np.zeros(ne * 3, dtype=np.complex128)
        self.L.orc_zemat_helmholtz2d_p1(ne, _p(f64(coord)), _p(conn), float(wave_length), _p(c128(value)), _p(emat),
                                        _p(eres))
        return emat, eres

    def zemat_sommerfeld(self, coord, lconn, wave_length, value):
        lconn = u32(lconn)
        ne = lconn.shape[0]
        emat = np.zeros(ne * 4, dtype=np.complex128)
        eres = np.zeros(ne * 2, dtype=np.complex128)
        self.L.orc_zemat_sommerfeld_b1(ne, _p(f64(coord)), _p(lconn), float(wave_length), _p(c128(value)), _p(emat),
                                       _p(eres))
        return emat, eres

    def zassemble(self, nb, rowptr, col, conn, emat, val=None, dia=None):
        conn = u32(conn)
        val = np.zeros(int(rowptr[-1]), dtype=np.complex128) if val is None else c128(val).copy()
        dia = np.zeros(nb, dtype=np.complex128) if dia is None else c128(dia).copy()
        self.L.orc_zassemble(nb, _p(u32(rowptr)), _p(u32(col)), conn.shape[0], conn.shape[1], _p(conn),
                             _p(c128(emat)), _p(val), _p(dia))
        return val, dia

    def zassemble_vec(self, res, conn, eres):
        conn = u32(conn)
        res = c128(res).copy()
        self.L.orc_zassemble_vec(_p(res), conn.shape[0], conn.shape[1], _p(conn), _p(c128(eres)))
        return res

    def zset_bc(self, nb, rowptr, col, val, dia, flags, res):
        val, dia, res = c128(val).copy(), c128(dia).copy(), c128(res).copy()
        self.L.orc_zset_bc(nb, _p(u32(rowptr)), _p(u32(col)), _p(val), _p(dia), _p(i32(flags)), _p(res))
        return val, dia, res

    def zmatvec(self, nb, rowptr, col, val, dia, alpha, x, beta, y):
        y = c128(y).copy()
        self.L.orc_zmatvec(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)), float(alpha),
                           _p(c128(x)), float(beta), _p(y))
        return y

    def zmatvec_hermitian(self, nb, rowptr, col, val, dia, alpha, x, beta, y):
        """CZMatDia_BlkCrs::MatVec_Hermitian: alpha A^H x + beta y."""
        y = c128(y).copy()
        self.L.orc_zmatvec_hermitian(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)),
                                     float(alpha), _p(c128(x)), float(beta), _p(y))
        return y

    def zsolve_cg_herm(self, nb, rowptr, col, val, dia, rhs, tol, maxit, F=None, fval=None, fdia=None, order=None):
        """MatVec::Sol::Solve_CG (F None) or Solve_PCG with the ILU factor, for a
        Hermitian complex matrix: (ok, conv, iter, u, r_final)."""
        r = c128(rhs).copy()
        u = np.zeros(nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        pre = F is not None
        order = None if order is None else i32(order)
        ok = self.L.orc_zsolve_cg_herm(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)), int(pre),
                                       _p(u32(F["rowptr"])) if pre else None, _p(u32(F["col"])) if pre else None,
                                       _p(u32(F["diaind"])) if pre else None, _p(c128(fval)) if pre else None,
                                       _p(c128(fdia)) if pre else None, _p(order), _p(r), _p(u), C.byref(conv),
                                       C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def zsolve_cgnr(self, nb, rowptr, col, val, dia, rhs, tol, maxit):
        """MatVec::Sol::Solve_CGNR: (ok, conv, iter, u, r_final)."""
        r = c128(rhs).copy()
        u = np.zeros(nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.L.orc_zsolve_cgnr(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)), _p(r), _p(u),
                                    C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def zilu_values(self, nb, rowptr, col, val, dia, F, order=None):
        """Complex ILU on the pattern F (orc_iluk_symbolic): (fval, fdia, nsmall)."""
        fval = np.zeros(int(F["rowptr"][-1]), dtype=np.complex128)
        fdia = np.zeros(nb, dtype=np.complex128)
        order = None if order is None else i32(order)
        self.L.orc_zilu_setvalue(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)), _p(order),
                                 _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(fval), _p(fdia))
        ns = self.L.orc_zilu_decomp(nb, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])), _p(fval),
                                    _p(fdia))
        return fval, fdia, ns

    def zilu_solve(self, nb, F, fval, fdia, vec):
        v = c128(vec).copy()
        args = (nb, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])), _p(c128(fval)), _p(c128(fdia)))
        self.L.orc_zilu_forward(*args, _p(v))
        self.L.orc_zilu_backward(*args, _p(v))
        return v

    def zsolve_pcocg(self, nb, rowptr, col, val, dia, F, fval, fdia, rhs, tol, maxit, order=None):
        """MatVec::Sol::Solve_PCOCG: (ok, conv, iter, u, r_final)."""
        r = c128(rhs).copy()
        u = np.zeros(nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        order = None if order is None else i32(order)
        ok = self.L.orc_zsolve_pcocg(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)),
                                     _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])), _p(c128(fval)),
                                     _p(c128(fdia)), _p(order), _p(r), _p(u), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def bmat_pattern_boundary(self, nb, b, rowptr, col, flag_col, flag_row):
        """CMat_BlkCrs::SetPatternBoundary: (brow, bcol)."""
        rowptr, col = u32(rowptr), u32(col)
        brow = np.zeros(nb + 1, dtype=np.uint32)
        bcol = np.zeros(int(rowptr[-1]) + nb + 1, dtype=np.uint32)
        n = self.L.orc_bmat_pattern_boundary(nb, b, _p(rowptr), _p(col), _p(i32(flag_col)), _p(i32(flag_row)),
                                             _p(brow), _p(bcol))
        return brow, bcol[:n].copy()

    def bmat_mearge(self, nb, b, brow, bcol, conn, emat, val=None):
        conn = u32(conn)
        val = np.zeros(int(brow[-1]) * b * b) if val is None else f64(val).copy()
        self.L.orc_bmat_mearge(nb, b, _p(u32(brow)), _p(u32(bcol)), conn.shape[0], conn.shape[1], _p(conn),
                               _p(f64(emat)), _p(val))
        return val

    def bmat_bc(self, nb, b, brow, bcol, flag, val):
        """SetBoundaryConditionInverse_Colum then SetBoundaryCondition_Row with the same flags."""
        val = f64(val).copy()
        self.L.orc_bmat_bc_inverse_col(nb, b, _p(u32(brow)), _p(u32(bcol)), _p(i32(flag)), _p(val))
        self.L.orc_bmat_bc_row(nb, b, _p(u32(brow)), _p(u32(bcol)), _p(i32(flag)), _p(val))
        return val

    def bmat_matvec(self, nb, b, brow, bcol, val, alpha, x, beta, y):
        y = f64(y).copy()
        self.L.orc_bmat_matvec(nb, b, _p(u32(brow)), _p(u32(bcol)), _p(f64(val)), float(alpha), _p(f64(x)),
                               float(beta), _p(y))
        return y

    def zsolve_cocg(self, nb, rowptr, col, val, dia, rhs, tol, maxit):
        """MatVec::Sol::Solve_COCG: (ok, conv, iter, u, r_final)."""
        r = c128(rhs).copy()
        u = np.zeros(nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.L.orc_zsolve_cocg(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)), _p(r), _p(u),
                                    C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def zsolve_bicgstab(self, nb, rowptr, col, val, dia, rhs, tol, maxit, F=None, fval=None, fdia=None,
                        order=None):
        """MatVec::Sol::Solve_BiCGSTAB (F None) or Solve_BiCGStabP with the
        factor: (ok, conv, iter, u, r_final)."""
        r = c128(rhs).copy()
        u = np.zeros(nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        order = None if order is None else i32(order)
        fa = (None,) * 5 if F is None else (u32(F["rowptr"]), u32(F["col"]), u32(F["diaind"]), c128(fval), c128(fdia))
        ok = self.L.orc_zsolve_bicgstab(nb, _p(u32(rowptr)), _p(u32(col)), _p(c128(val)), _p(c128(dia)),
                                        0 if F is None else 1, *[_p(a) for a in fa], _p(order), _p(r), _p(u),
                                        C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    MAT_KINDS = {"cg": 0, "pcg": 1, "bicgstab": 2, "pbicgstab": 3}

    def solve_ls(self, kind, nb, b, rowptr, col, val, dia, rhs, tol, maxit):
        """LsSol::Solve_CG ("cg") / Solve_BiCGSTAB ("bicgstab"), unpreconditioned:
        returns (ok, conv, iter, x, r_final)."""
        r = f64(rhs).copy()
        x = np.zeros(nb * b)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.L.orc_solve_ls(self.MAT_KINDS[kind], nb, b, _p(u32(rowptr)), _p(u32(col)), _p(f64(val)),
                                 _p(f64(dia)), _p(r), _p(x), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r

    def solve_mat(self, kind, nb, b, rowptr, col, val, dia, rhs, tol, maxit, F=None, fval=None, fdia=None):
        """MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
        (natural-order factor for the preconditioned kinds): returns
        (ok, conv, iter, u, r_final)."""
        r = f64(rhs).copy()
        u = np.zeros(nb * b)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        fp = (None, None, None, None, None) if F is None else (u32(F["rowptr"]), u32(F["col"]), u32(F["diaind"]),
                                                             f64(fval), f64(fdia))
        ok = self.L.orc_solve_mat(self.MAT_KINDS[kind], nb, b, _p(u32(rowptr)), _p(u32(col)), _p(f64(val)),
                                  _p(f64(dia)), *[_p(a) for a in fp], _p(r), _p(u), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    # -- elements ------------------------------------------------------------
    def emat_poisson2d(self, coord, conn, alpha, source, value):
        conn = u32(conn)
        ne = conn.shape[0]
        emat = np.zeros(ne * 9)
        eres = np.zeros(ne * 3)
        self.L.orc_emat_poisson2d_p1(ne, _p(f64(coord)), _p(conn), _p(f64(alpha)), source, _p(f64(value)), _p(emat),
                                     _p(eres))
        return emat, eres

    def emat_solid3d(self, coord, conn, lam, mu, rho, g, disp):
        conn = u32(conn)
        ne = conn.shape[0]
        emat = np.zeros(ne * 144)
        eres = np.zeros(ne * 12)
        self.L.orc_emat_solid3d_p1(ne, _p(f64(coord)), _p(conn), _p(f64(lam)), _p(f64(mu)), rho, g[0], g[1], g[2],
                                   _p(f64(disp)), _p(emat), _p(eres))
        return emat, eres

    def emat_oseen2d(self, coord, conn, myu, rho, adv, value):
        conn = u32(conn)
        ne = conn.shape[0]
        emat = np.zeros(ne * 81)
        eres = np.zeros(ne * 9)
        self.L.orc_emat_oseen2d_p1p1(ne, _p(f64(coord)), _p(conn), myu, rho, _p(f64(adv)), _p(f64(value)), _p(emat),
                                     _p(eres))
        return emat, eres

    def emat_helmholtz2d(self, coord, conn, alpha, cmass, source, value):
        conn = u32(conn)
        ne = conn.shape[0]
        emat = np.zeros(ne * 9)
        eres = np.zeros(ne * 3)
        self.L.orc_emat_helmholtz2d_p1(ne, _p(f64(coord)), _p(conn), _p(f64(alpha)), _p(f64(cmass)), source,
                                       _p(f64(value)), _p(emat), _p(eres))
        return emat, eres

    # -- orderings -------------------------------------------------------------
    def colour_order(self, rowptr, col):
        rowptr, col = u32(rowptr), u32(col)
        nb = rowptr.size - 1
        colour = np.zeros(nb, dtype=np.int32)
        nc = self.L.orc_greedy_colour(nb, _p(rowptr), _p(col), _p(colour))
        order = np.zeros(nb, dtype=np.int32)
        self.L.orc_colour_order(nb, _p(colour), nc, _p(order))
        return colour, nc, order

    def bmc_order(self, rowptr, col, block_nodes=32):
        rowptr, col = u32(rowptr), u32(col)
        nb = rowptr.size - 1
        order = np.zeros(nb, dtype=np.int32)
        blk = np.zeros(nb, dtype=np.int32)
        nc = self.L.orc_bmc_order(nb, _p(rowptr), _p(col), block_nodes, _p(order), _p(blk))
        return blk, nc, order

    def level_sets(self, F, lower):
        nb = F["rowptr"].size - 1
        lev = np.zeros(nb, dtype=np.int32)
        n = self.L.orc_level_sets(nb, _p(u32(F["rowptr"])), _p(u32(F["col"])), _p(u32(F["diaind"])),
                                  1 if lower else 0, _p(lev))
        return lev, n


class Ref:
    """The compiled DelFEM reference (ref_driver.cpp entry points)."""

    def __init__(self, path=REF_PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: the reference library is built where the DelFEM tree exists")
        L = C.CDLL(path)
        sig = {
            "ref_quiet": (None, [C.c_int]),
            "ref_sys_create": (vp, [C.c_uint, C.c_uint, vp, vp]),
            "ref_sys_free": (None, [vp]),
            "ref_sys_initialize_marge": (None, [vp]),
            "ref_sys_ncrs": (C.c_uint, [vp]),
            "ref_sys_merge": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_sys_add_residual": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_sys_set_bc": (None, [vp, vp]),
            "ref_sys_finalize": (C.c_double, [vp]),
            "ref_sys_get_matrix": (None, [vp, vp, vp, vp, vp]),
            "ref_sys_set_matrix": (None, [vp, vp, vp]),
            "ref_sys_resize_tmp": (None, [vp, C.c_uint]),
            "ref_sys_set_vector": (None, [vp, C.c_int, vp]),
            "ref_sys_get_vector": (None, [vp, C.c_int, vp]),
            "ref_sys_matvec": (None, [vp, C.c_double, C.c_int, C.c_double, C.c_int]),
            "ref_prec_create": (vp, [vp, C.c_int, vp]),
            "ref_prec_free": (None, [vp]),
            "ref_prec_set_value": (C.c_int, [vp, vp]),
            "ref_prec_solve": (None, [vp, vp, C.c_int]),
            "ref_solve_pcg": (C.c_int, [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_solve_pbicgstab": (C.c_int, [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_ls_solve": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_frac_create": (vp, [vp, C.c_int, vp]),
            "ref_frac_create_fill": (vp, [vp, C.c_int, vp, C.c_uint]),
            "ref_frac_free": (None, [vp]),
            "ref_frac_ncrs": (C.c_uint, [vp]),
            "ref_frac_get_pattern": (None, [vp, vp, vp, vp]),
            "ref_frac_set_value": (C.c_int, [vp, vp]),
            "ref_frac_get_values": (None, [vp, vp, vp]),
            "ref_frac_forward": (None, [vp, vp]),
            "ref_frac_backward": (None, [vp, vp]),
            "ref_frac_solve": (C.c_int, [vp, vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_mat_solve": (C.c_int, [vp, vp, C.c_int, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_zsys_create": (vp, [C.c_uint, vp, vp]),
            "ref_zsys_free": (None, [vp]),
            "ref_zsys_merge": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_zsys_set_bc": (None, [vp, vp]),
            "ref_zsys_get": (None, [vp, vp, vp]),
            "ref_zsys_matvec": (None, [vp, C.c_double, vp, C.c_double, vp]),
            "ref_zsys_matvec_hermitian": (None, [vp, C.c_double, vp, C.c_double, vp]),
            "ref_zsolve_cgnr": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_zsolve_cg_herm": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_zfrac_create": (vp, [vp, C.c_int]),
            "ref_zfrac_free": (None, [vp]),
            "ref_zfrac_ncrs": (C.c_uint, [vp]),
            "ref_zfrac_get_pattern": (None, [vp, C.c_uint, vp, vp]),
            "ref_zfrac_set_value": (C.c_int, [vp, vp]),
            "ref_zfrac_get_values": (None, [vp, C.c_uint, vp, vp]),
            "ref_zfrac_solve": (None, [vp, C.c_uint, vp]),
            "ref_zsolve_pcocg": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_zsolve_cocg": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_bmat_create_boundary": (vp, [vp, vp, vp]),
            "ref_ms_create": (vp, [C.c_uint, vp, vp]),
            "ref_ms_free": (None, [vp]),
            "ref_ms_add_pattern": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_ms_initialize": (None, [vp]),
            "ref_ms_merge": (None, [vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp]),
            "ref_ms_set_bc": (None, [vp, C.c_uint, vp]),
            "ref_ms_set_vector": (None, [vp, C.c_int, vp]),
            "ref_ms_get_vector": (None, [vp, C.c_int, vp]),
            "ref_ms_finalize": (C.c_double, [vp]),
            "ref_ms_get_block": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_ms_matvec": (None, [vp, C.c_double, vp, C.c_double, vp]),
            "ref_ms_ilu": (C.c_int, [vp, C.c_int]),
            "ref_ms_ilu_inputs": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int]),
            "ref_ms_precond": (None, [vp, vp]),
            "ref_ms_solve": (C.c_int, [vp, C.c_int, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_bmat_free": (None, [vp]),
            "ref_bmat_ncrs": (C.c_uint, [vp]),
            "ref_bmat_get": (None, [vp, vp, vp, vp]),
            "ref_bmat_merge": (None, [vp, C.c_uint, C.c_uint, vp, vp]),
            "ref_bmat_bc": (None, [vp, vp, C.c_int]),
            "ref_bmat_matvec": (None, [vp, C.c_double, vp, C.c_double, vp]),
            "ref_zsolve_bicgstab": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_fx_create": (vp, [C.c_uint, vp, vp, vp]),
            "ref_fx_free": (None, [vp]),
            "ref_fx_merge": (None, [vp, C.c_uint, vp, vp, vp]),
            "ref_fx_set_bc": (None, [vp, vp]),
            "ref_fx_finalize": (C.c_double, [vp]),
            "ref_fx_ncrs": (C.c_uint, [vp]),
            "ref_fx_get_matrix": (None, [vp, vp, vp, vp, vp, vp, vp]),
            "ref_fx_set_vector": (None, [vp, C.c_int, vp]),
            "ref_fx_get_vector": (None, [vp, C.c_int, vp]),
            "ref_fx_resize_tmp": (None, [vp, C.c_uint]),
            "ref_fx_matvec": (None, [vp, C.c_double, C.c_int, C.c_double, C.c_int]),
            "ref_fx_frac_create": (vp, [vp, C.c_int]),
            "ref_fx_frac_free": (None, [vp]),
            "ref_fx_frac_ncrs": (C.c_uint, [vp]),
            "ref_fx_frac_set_value": (C.c_int, [vp, vp]),
            "ref_fx_frac_get": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_fx_frac_forward": (None, [vp, vp, vp]),
            "ref_fx_frac_backward": (None, [vp, vp, vp]),
            "ref_fx_solve": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
            "ref_emat_poisson2d_p1": (None, [C.c_uint, vp, vp, vp, C.c_double, vp, vp, vp]),
            "ref_emat_solid3d_p1": (None, [C.c_uint, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                           C.c_double, vp, vp, vp]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        L.ref_quiet(1)
        self.L = L


class RefSystem:
    """One LsSol::CLinearSystem of the reference with a single segment."""

    def __init__(self, ref, nb, b, index, array):
        self.ref, self.nb, self.b = ref, nb, b
        self._index, self._array = u32(index), u32(array)
        if self._array.size == 0:
            self._array = np.zeros(1, dtype=np.uint32)
        self.h = ref.L.ref_sys_create(nb, b, _p(self._index), _p(self._array))
        ref.L.ref_sys_initialize_marge(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.L.ref_sys_free(self.h)
            self.h = None

    def merge(self, conn, emat, eres=None):
        conn = u32(conn)
        self.ref.L.ref_sys_merge(self.h, conn.shape[0], conn.shape[1], _p(conn), _p(f64(emat)))
        if eres is not None:
            self.ref.L.ref_sys_add_residual(self.h, conn.shape[0], conn.shape[1], _p(conn), _p(f64(eres)))

    def set_bc(self, flags):
        self.ref.L.ref_sys_set_bc(self.h, _p(i32(flags)))

    def finalize(self):
        return self.ref.L.ref_sys_finalize(self.h)

    def matrix(self):
        n = self.ref.L.ref_sys_ncrs(self.h)
        bb = self.b * self.b
        rowptr = np.zeros(self.nb + 1, dtype=np.uint32)
        col = np.zeros(max(1, n), dtype=np.uint32)
        val = np.zeros(max(1, n * bb))
        dia = np.zeros(self.nb * bb)
        self.ref.L.ref_sys_get_matrix(self.h, _p(rowptr), _p(col), _p(val), _p(dia))
        return rowptr, col[:n], val[:n * bb], dia

    def set_matrix(self, val, dia):
        self.ref.L.ref_sys_set_matrix(self.h, _p(f64(val)), _p(f64(dia)))

    def set_vector(self, iv, v):
        self.ref.L.ref_sys_set_vector(self.h, iv, _p(f64(v)))

    def get_vector(self, iv):
        out = np.zeros(self.nb * self.b)
        self.ref.L.ref_sys_get_vector(self.h, iv, _p(out))
        return out

    def matvec(self, alpha, iv1, beta, iv2):
        self.ref.L.ref_sys_matvec(self.h, alpha, iv1, beta, iv2)

    def resize_tmp(self, n):
        self.ref.L.ref_sys_resize_tmp(self.h, n)

    def solve_ls(self, kind, rhs, tol, maxit):
        """LsSol::Solve_CG ("cg") / Solve_BiCGSTAB ("bicgstab") on this system:
        returns (ok, conv, iter, x, r_final)."""
        self.set_vector(-1, rhs)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_ls_solve(self.h, Orc.MAT_KINDS[kind], C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, self.get_vector(-2), self.get_vector(-1)

    def solve_mat(self, kind, rhs, tol, maxit, frac=None):
        """MatVec::Sol::Solve_* on this system's matrix (frac: a natural-order
        RefFrac holding values, for the preconditioned kinds): returns
        (ok, conv, iter, u, r_final)."""
        r = f64(rhs).copy()
        u = np.zeros(self.nb * self.b)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_mat_solve(self.h, None if frac is None else frac.h, Orc.MAT_KINDS[kind], _p(r), _p(u),
                                      C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r


class RefILU:
    """LsSol::CPreconditioner_ILU of the reference."""

    def __init__(self, sysobj, lev, order=None):
        self.s = sysobj
        self._order = None if order is None else i32(order)
        self.h = sysobj.ref.L.ref_prec_create(sysobj.h, lev, _p(self._order))

    def __del__(self):
        if getattr(self, "h", None):
            self.s.ref.L.ref_prec_free(self.h)
            self.h = None

    def set_value(self):
        return bool(self.s.ref.L.ref_prec_set_value(self.h, self.s.h))

    def solve(self, kind, tol, maxit):
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        f = self.s.ref.L.ref_solve_pcg if kind == "pcg" else self.s.ref.L.ref_solve_pbicgstab
        ok = f(self.s.h, self.h, C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value


class RefFrac:
    """MatVec::CMatDiaFrac_BlkCrs of the reference (natural or ordered ILU(k));
    with fill_blk: CMatDiaFrac_BlkCrs(0, A) + AddFracPtn(lev, aBlkFill) +
    MakePatternFinalize, as CPreconditioner_ILU builds it after SetFillBlk."""

    def __init__(self, sysobj, lev, order=None, fill_blk=None):
        self.s = sysobj
        self._order = None if order is None else i32(order)
        if fill_blk is None:
            self.h = sysobj.ref.L.ref_frac_create(sysobj.h, lev, _p(self._order))
        else:
            blk = u32(fill_blk)
            self.h = sysobj.ref.L.ref_frac_create_fill(sysobj.h, lev, _p(blk), blk.size)

    def __del__(self):
        if getattr(self, "h", None):
            self.s.ref.L.ref_frac_free(self.h)
            self.h = None

    def pattern(self):
        L = self.s.ref.L
        n = L.ref_frac_ncrs(self.h)
        nb = self.s.nb
        rowptr = np.zeros(nb + 1, dtype=np.uint32)
        col = np.zeros(max(1, n), dtype=np.uint32)
        dia = np.zeros(nb, dtype=np.uint32)
        L.ref_frac_get_pattern(self.h, _p(rowptr), _p(col), _p(dia))
        return dict(rowptr=rowptr, col=col[:n], diaind=dia)

    def set_value(self):
        return bool(self.s.ref.L.ref_frac_set_value(self.h, self.s.h))

    def values(self):
        L = self.s.ref.L
        n = L.ref_frac_ncrs(self.h)
        bb = self.s.b * self.s.b
        val = np.zeros(max(1, n * bb))
        dia = np.zeros(self.s.nb * bb)
        L.ref_frac_get_values(self.h, _p(val), _p(dia))
        return val[:n * bb], dia

    def forward(self, vec):
        v = f64(vec).copy()
        self.s.ref.L.ref_frac_forward(self.h, _p(v))
        return v

    def backward(self, vec):
        v = f64(vec).copy()
        self.s.ref.L.ref_frac_backward(self.h, _p(v))
        return v

    def solve(self, kind, tol, maxit):
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.s.ref.L.ref_frac_solve(self.h, self.s.h, 0 if kind == "pcg" else 1, C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value


def split_problem(orc, prob, groups):
    """A combined-unknown Problem (b dofs per node) split into segments of the
    given dof groups (e.g. [[0, 1], [2]]: velocity and pressure) the way the
    reference's split fields assemble it: per block (i, j) the element
    matrices restricted to the dofs of groups i and j; diagonal patterns are
    the node adjacency, off-diagonal ones the adjacency with the node itself.
    Returns dict(nb, b=[...], rowptr/col for 'dia' and 'nd', emat[(i, j)],
    res (concatenated, before the Dirichlet rows), flags[i])."""
    emat, eres = problem_emats(orc, prob)
    nb, B, npe, ne = prob.nnode, prob.b, prob.npe, prob.nelem
    em = emat.reshape(ne, npe, npe, B, B)
    er = eres.reshape(ne, npe, B)
    rowptr, col = orc.fem_pattern(nb, prob.conn)
    # off-diagonal pattern: adjacency plus the node itself, ascending
    rp2 = np.zeros(nb + 1, dtype=np.uint32)
    cols = []
    for i in range(nb):
        c = np.sort(np.append(col[rowptr[i]:rowptr[i + 1]], i)).astype(np.uint32)
        cols.append(c)
        rp2[i + 1] = rp2[i] + c.size
    col2 = np.concatenate(cols).astype(np.uint32)
    flags = np.ascontiguousarray(prob.flags, dtype=np.int32).reshape(nb, B)
    out = dict(nb=nb, b=[len(g) for g in groups], rowptr=rowptr, col=col, rowptr_nd=rp2, col_nd=col2, emat={},
               flags=[np.ascontiguousarray(flags[:, g]).reshape(-1) for g in groups])
    for i, gi in enumerate(groups):
        for j, gj in enumerate(groups):
            out["emat"][(i, j)] = np.ascontiguousarray(em[:, :, :, gi][:, :, :, :, gj]).reshape(-1)
    res = np.zeros(nb * B)
    res = orc.assemble_vec(nb, B, prob.conn, eres).reshape(nb, B)
    out["res"] = np.concatenate([np.ascontiguousarray(res[:, g]).reshape(-1) for g in groups])
    return out


def fill_levels(nseg, inputs):
    """The fill level of each block (i, j) after SetFillInLevel(lev, ilss)
    calls, in order (preconditioner.cpp:108-122): ilss -1 sets every block,
    otherwise the row and the column of segment ilss; later calls win."""
    alev = np.zeros((nseg, nseg), dtype=np.int32)
    for lev, ilss in inputs:
        if ilss == -1:
            alev[:, :] = lev
        else:
            alev[:, ilss] = lev
            alev[ilss, :] = lev
    return alev


class OrcMS:
    """The oracle's multi-segment system: blocks assembled (Mearge of the
    split element matrices), Dirichlet conditions as CLinearSystem::
    FinalizeMarge sets them, MATVEC, cross-segment ILU(0), SolvePrecond and
    the Krylov solvers."""

    def __init__(self, orc, sp, conn):
        self.orc, self.sp = orc, sp
        L = orc.L
        nseg = len(sp["b"])
        nb = sp["nb"]
        self.nseg, self.nb, self.b = nseg, nb, sp["b"]
        self.keep = []
        self.val, self.dia = {}, {}
        conn = u32(conn)
        for i in range(nseg):
            bi = self.b[i]
            val, dia = orc.assemble(nb, bi, sp["rowptr"], sp["col"], conn, sp["emat"][(i, i)])
            val, dia = orc.set_bc(nb, bi, sp["rowptr"], sp["col"], val, dia, sp["flags"][i])
            self.val[(i, i)], self.dia[i] = val, dia
            for j in range(nseg):
                if i == j:
                    continue
                bj = self.b[j]
                v = np.zeros(int(sp["rowptr_nd"][-1]) * bi * bj)
                L.orc_bmat2_mearge(nb, bi, nb, bj, _p(u32(sp["rowptr_nd"])), _p(u32(sp["col_nd"])), conn.shape[0],
                                   conn.shape[1], _p(conn), conn.shape[1], _p(conn), _p(f64(sp["emat"][(i, j)])), _p(v))
                L.orc_bmat2_bc(nb, bi, nb, bj, _p(u32(sp["rowptr_nd"])), _p(u32(sp["col_nd"])),
                               _p(i32(sp["flags"][i])), _p(i32(sp["flags"][j])), _p(v))
                self.val[(i, j)] = v
        self.nbs = np.array([nb] * nseg, dtype=np.uint32)
        self.bs = np.array(self.b, dtype=np.uint32)
        self.h = L.orc_ms_create(nseg, _p(self.nbs), _p(self.bs))
        rp, cl = u32(sp["rowptr"]), u32(sp["col"])
        rp2, cl2 = u32(sp["rowptr_nd"]), u32(sp["col_nd"])
        self.keep += [rp, cl, rp2, cl2]
        for i in range(nseg):
            L.orc_ms_set_dia(self.h, i, _p(rp), _p(cl), _p(self.val[(i, i)]), _p(self.dia[i]))
            for j in range(nseg):
                if i != j:
                    L.orc_ms_set_nondia(self.h, i, j, _p(rp2), _p(cl2), _p(self.val[(i, j)]))
        flags = np.concatenate(sp["flags"])
        self.res = np.where(flags != 0, 0.0, sp["res"])

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.L.orc_ms_free(self.h)

    def matvec(self, alpha, x, beta, y):
        y = f64(y).copy()
        self.orc.L.orc_ms_matvec(self.h, float(alpha), _p(f64(x)), float(beta), _p(y))
        return y

    def ilu(self, lev=0, inputs=None, fill_blk=None):
        """CPreconditioner_ILU(ls, lev) across the segments; True on success.
        inputs: SetFillInLevel(lev, ilss) calls [(lev, ilss), ...] instead;
        fill_blk: SetFillBlk's rows of segment 0."""
        if inputs is not None or fill_blk is not None:
            inputs = [(lev, -1)] if inputs is None else inputs
            alev = np.ascontiguousarray(fill_levels(self.nseg, inputs), dtype=np.int32)
            blk = u32(fill_blk if fill_blk is not None else [])
            return self.orc.L.orc_ms_ilu_levels(self.h, _p(alev), _p(blk) if blk.size else None, blk.size) == 0
        return self.orc.L.orc_ms_ilu(self.h, int(lev)) == 0

    def pattern(self, i, j):
        """Factor pattern of block (i, j) (rowptr, col) after ilu()."""
        n = self.orc.L.orc_ms_factor_ncrs(self.h, i, j)
        rp = np.zeros(self.nb + 1, dtype=np.uint32)
        cl = np.zeros(max(n, 1), dtype=np.uint32)
        self.orc.L.orc_ms_get_pattern(self.h, i, j, _p(rp), _p(cl))
        return rp, cl[:n]

    def factor(self, i, j):
        bb = self.b[i] * self.b[j]
        n = int(self.orc.L.orc_ms_factor_ncrs(self.h, i, j))
        if i == j:
            fv = np.zeros(n * bb)
            fd = np.zeros(self.nb * bb)
            self.orc.L.orc_ms_get_factor(self.h, i, j, _p(fv), _p(fd))
            return fv, fd
        fv = np.zeros(n * bb)
        self.orc.L.orc_ms_get_factor(self.h, i, j, _p(fv), None)
        return fv

    def precond(self, z):
        z = f64(z).copy()
        self.orc.L.orc_ms_precond(self.h, _p(z))
        return z

    def solve(self, bicg, rhs, tol, maxit):
        r = f64(rhs).copy()
        x = np.zeros_like(r)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.orc.L.orc_ms_solve(self.h, 1 if bicg else 0, _p(r), _p(x), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r


class RefMS:
    """The reference's multi-segment LsSol::CLinearSystem assembled from a
    split problem, with CPreconditioner_ILU(0) across the segments."""

    def __init__(self, ref, sp, conn):
        self.ref, self.sp = ref, sp
        L = ref.L
        nseg = len(sp["b"])
        self.nbs = np.array([sp["nb"]] * nseg, dtype=np.uint32)
        self.bs = np.array(sp["b"], dtype=np.uint32)
        self.h = L.ref_ms_create(nseg, _p(self.nbs), _p(self.bs))
        for i in range(nseg):
            for j in range(nseg):
                if i == j:
                    L.ref_ms_add_pattern(self.h, i, j, _p(u32(sp["rowptr"])), _p(u32(sp["col"])))
                else:
                    L.ref_ms_add_pattern(self.h, i, j, _p(u32(sp["rowptr_nd"])), _p(u32(sp["col_nd"])))
        L.ref_ms_initialize(self.h)
        conn = u32(conn)
        for i in range(nseg):
            for j in range(nseg):
                L.ref_ms_merge(self.h, i, j, conn.shape[0], conn.shape[1], _p(conn), _p(f64(sp["emat"][(i, j)])))
        L.ref_ms_set_vector(self.h, -1, _p(f64(sp["res"])))
        for i in range(nseg):
            L.ref_ms_set_bc(self.h, i, _p(i32(sp["flags"][i])))
        self.norm_res = L.ref_ms_finalize(self.h)
        self.nseg = nseg

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.L.ref_ms_free(self.h)

    def block(self, i, j):
        nb, bb = self.sp["nb"], self.sp["b"][i] * self.sp["b"][j]
        n = int((self.sp["rowptr"] if i == j else self.sp["rowptr_nd"])[-1])
        val = np.zeros(n * bb)
        dia = np.zeros(nb * bb) if i == j else None
        self.ref.L.ref_ms_get_block(self.h, i, j, _p(val), _p(dia))
        return (val, dia) if i == j else val

    def vector(self, iv):
        v = np.zeros(sum(self.sp["b"]) * self.sp["nb"])
        self.ref.L.ref_ms_get_vector(self.h, iv, _p(v))
        return v

    def matvec(self, alpha, x, beta, y):
        y = f64(y).copy()
        self.ref.L.ref_ms_matvec(self.h, float(alpha), _p(f64(x)), float(beta), _p(y))
        return y

    def ilu(self, lev=0, inputs=None, fill_blk=None):
        if inputs is not None or fill_blk is not None:
            inputs = [(lev, -1)] if inputs is None else inputs
            lv = np.ascontiguousarray([l for l, _ in inputs], dtype=np.int32)
            ss = np.ascontiguousarray([i for _, i in inputs], dtype=np.int32)
            blk = u32(fill_blk if fill_blk is not None else [])
            return bool(self.ref.L.ref_ms_ilu_inputs(self.h, len(inputs), _p(lv), _p(ss), _p(blk) if blk.size else None,
                                                     blk.size))
        return bool(self.ref.L.ref_ms_ilu(self.h, int(lev)))

    def precond(self, z):
        z = f64(z).copy()
        self.ref.L.ref_ms_precond(self.h, _p(z))
        return z

    def solve(self, bicg, rhs, tol, maxit):
        r = f64(rhs).copy()
        x = np.zeros_like(r)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_ms_solve(self.h, 1 if bicg else 0, _p(r), _p(x), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r


class RefBMat:
    """MatVec::CMat_BlkCrs of the reference, patterned by SetPatternBoundary
    on a RefSystem's matrix (the boundary matrix of CLinearSystem_Save)."""

    def __init__(self, rs, flag_col, flag_row):
        self.ref, self.nb, self.b = rs.ref, rs.nb, rs.b
        self.h = self.ref.L.ref_bmat_create_boundary(rs.h, _p(i32(flag_col)), _p(i32(flag_row)))

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.L.ref_bmat_free(self.h)

    def get(self):
        n = self.ref.L.ref_bmat_ncrs(self.h)
        brow = np.zeros(self.nb + 1, dtype=np.uint32)
        bcol = np.zeros(max(n, 1), dtype=np.uint32)
        val = np.zeros(max(n, 1) * self.b * self.b)
        self.ref.L.ref_bmat_get(self.h, _p(brow), _p(bcol), _p(val))
        return brow, bcol[:n], val[:n * self.b * self.b]

    def merge(self, conn, emat):
        conn = u32(conn)
        self.ref.L.ref_bmat_merge(self.h, conn.shape[0], conn.shape[1], _p(conn), _p(f64(emat)))

    def bc(self, flag):
        self.ref.L.ref_bmat_bc(self.h, _p(i32(flag)), 1)
        self.ref.L.ref_bmat_bc(self.h, _p(i32(flag)), 0)

    def matvec(self, alpha, x, beta, y):
        y = f64(y).copy()
        self.ref.L.ref_bmat_matvec(self.h, float(alpha), _p(f64(x)), float(beta), _p(y))
        return y


class RefZSystem:
    """MatVec::CZMatDia_BlkCrs (len 1) of the reference with its ILU(lev) factor."""

    def __init__(self, ref, nb, rowptr, col):
        self.ref, self.nb = ref, nb
        self._rp, self._cl = u32(rowptr), u32(col)
        self.h = ref.L.ref_zsys_create(nb, _p(self._rp), _p(self._cl))
        self.f = None

    def __del__(self):
        if getattr(self, "f", None):
            self.ref.L.ref_zfrac_free(self.f)
        if getattr(self, "h", None):
            self.ref.L.ref_zsys_free(self.h)

    def merge(self, conn, emat):
        conn = u32(conn)
        self.ref.L.ref_zsys_merge(self.h, conn.shape[0], conn.shape[1], _p(conn), _p(c128(emat)))

    def set_bc(self, flags):
        self.ref.L.ref_zsys_set_bc(self.h, _p(i32(flags)))

    def matrix(self):
        val = np.zeros(int(self._rp[-1]), dtype=np.complex128)
        dia = np.zeros(self.nb, dtype=np.complex128)
        self.ref.L.ref_zsys_get(self.h, _p(val), _p(dia))
        return val, dia

    def matvec(self, alpha, x, beta, y):
        y = c128(y).copy()
        self.ref.L.ref_zsys_matvec(self.h, float(alpha), _p(c128(x)), float(beta), _p(y))
        return y

    def factor(self, lev):
        if self.f:
            self.ref.L.ref_zfrac_free(self.f)
        self.f = self.ref.L.ref_zfrac_create(self.h, int(lev))
        n = self.ref.L.ref_zfrac_ncrs(self.f)
        rowptr = np.zeros(self.nb + 1, dtype=np.uint32)
        col = np.zeros(max(n, 1), dtype=np.uint32)
        self.ref.L.ref_zfrac_get_pattern(self.f, self.nb, _p(rowptr), _p(col))
        ok = bool(self.ref.L.ref_zfrac_set_value(self.f, self.h))
        fval = np.zeros(max(n, 1), dtype=np.complex128)
        fdia = np.zeros(self.nb, dtype=np.complex128)
        self.ref.L.ref_zfrac_get_values(self.f, self.nb, _p(fval), _p(fdia))
        return dict(rowptr=rowptr, col=col[:n]), fval[:n], fdia, ok

    def precond(self, vec):
        v = c128(vec).copy()
        self.ref.L.ref_zfrac_solve(self.f, self.nb, _p(v))
        return v

    def pcocg(self, rhs, tol, maxit):
        r = c128(rhs).copy()
        u = np.zeros(self.nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_zsolve_pcocg(self.h, self.f, _p(r), _p(u), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def cocg(self, rhs, tol, maxit):
        r = c128(rhs).copy()
        u = np.zeros(self.nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_zsolve_cocg(self.h, _p(r), _p(u), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def matvec_hermitian(self, alpha, x, beta, y):
        y = c128(y).copy()
        self.ref.L.ref_zsys_matvec_hermitian(self.h, float(alpha), _p(c128(x)), float(beta), _p(y))
        return y

    def cg_herm(self, rhs, tol, maxit, precond):
        """Solve_CG, or Solve_PCG with the factor when precond."""
        r = c128(rhs).copy()
        u = np.zeros(self.nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_zsolve_cg_herm(self.h, self.f if precond else None, _p(r), _p(u), C.byref(conv),
                                           C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def cgnr(self, rhs, tol, maxit):
        r = c128(rhs).copy()
        u = np.zeros(self.nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_zsolve_cgnr(self.h, _p(r), _p(u), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, u, r

    def bicgstab(self, rhs, tol, maxit, precond):
        """Solve_BiCGSTAB, or Solve_BiCGStabP with the factor when precond."""
        r = c128(rhs).copy()
        u = np.zeros(self.nb, dtype=np.complex128)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_zsolve_bicgstab(self.h, self.f if precond else None, _p(r), _p(u), C.byref(conv),
                                            C.byref(it))
        return bool(ok), conv.value, it.value, u, r


def assemble_zproblem(orc, prob):
    """Oracle pipeline of the complex Helmholtz problem: pattern (triangles),
    triangle and Sommerfeld line elements merged in that order, Dirichlet
    rows/columns and residual."""
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    et, rt = orc.zemat_helmholtz2d(prob.coord, prob.conn, prob.wave_length, prob.value)
    el, rl = orc.zemat_sommerfeld(prob.coord, prob.lconn, prob.wave_length, prob.value)
    val, dia = orc.zassemble(prob.nnode, rowptr, col, prob.conn, et)
    val, dia = orc.zassemble(prob.nnode, rowptr, col, prob.lconn, el, val, dia)
    res = np.zeros(prob.nnode, dtype=np.complex128)
    res = orc.zassemble_vec(res, prob.conn, rt)
    res = orc.zassemble_vec(res, prob.lconn, rl)
    val, dia, res = orc.zset_bc(prob.nnode, rowptr, col, val, dia, prob.flags, res)
    return dict(rowptr=rowptr, col=col, val=val, dia=dia, res=res, emat_tri=et, eres_tri=rt, emat_line=el,
                eres_line=rl)


def hermitian_zproblem(orc, prob, seed=0, skew=0.1):
    """A Hermitian positive definite complex system on the mesh of a complex
    Helmholtz problem (for the Hermitian Solve_CG / Solve_PCG): per triangle
    the real part of its Helmholtz element (a long wave length makes it
    positive definite) plus i*S with S real antisymmetric (random, scaled by
    skew against the element's diagonal), merged, Dirichlet rows/columns;
    right-hand side from the triangles' residuals."""
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    et, rt = orc.zemat_helmholtz2d(prob.coord, prob.conn, prob.wave_length, prob.value)
    ne = prob.conn.shape[0]
    E = et.reshape(ne, 3, 3).real.astype(np.complex128)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, (ne, 3))
    S = np.zeros((ne, 3, 3))
    S[:, 0, 1], S[:, 0, 2], S[:, 1, 2] = a[:, 0], a[:, 1], a[:, 2]
    S = S - S.transpose(0, 2, 1)
    scale = skew * np.abs(np.diagonal(E.real, axis1=1, axis2=2)).min(axis=1)
    E = E + 1j * S * scale[:, None, None]
    emat = np.ascontiguousarray(E.reshape(-1))
    val, dia = orc.zassemble(prob.nnode, rowptr, col, prob.conn, emat)
    res = orc.zassemble_vec(np.zeros(prob.nnode, dtype=np.complex128), prob.conn, rt)
    val, dia, res = orc.zset_bc(prob.nnode, rowptr, col, val, dia, prob.flags, res)
    return dict(rowptr=rowptr, col=col, val=val, dia=dia, res=res, emat_tri=emat)


def problem_emats(orc, prob, value=None):
    """Element matrices / residuals of a meshgen Problem via the C restatement."""
    if prob.kind == "poisson2d":
        v = np.zeros(prob.nnode) if value is None else value
        return orc.emat_poisson2d(prob.coord, prob.conn, prob.alpha, prob.source, v)
    if prob.kind == "helmholtz2d":
        v = np.zeros(prob.nnode) if value is None else value
        return orc.emat_helmholtz2d(prob.coord, prob.conn, prob.alpha, prob.cmass, prob.source, v)
    if prob.kind == "oseen2d":
        v = prob.value if value is None else value
        return orc.emat_oseen2d(prob.coord, prob.conn, prob.myu, prob.rho, prob.adv, v)
    v = np.zeros(prob.nnode * 3) if value is None else value
    return orc.emat_solid3d(prob.coord, prob.conn, prob.lam, prob.mu, prob.rho, prob.g, v)


def zero_block_rows(rowptr, col, val, dia, b, rows):
    """Copies of (val, dia) with whole block rows set to zero (off-diagonal
    blocks and diagonal block): the pivot of such a row is exactly zero after
    elimination, a singular pivot for DoILUDecomp whatever the ordering."""
    val, dia = f64(val).copy(), f64(dia).copy()
    bb = b * b
    for r in rows:
        r = int(r)
        val[int(rowptr[r]) * bb:int(rowptr[r + 1]) * bb] = 0.0
        dia[r * bb:(r + 1) * bb] = 0.0
    return val, dia


def assemble_problem(orc, prob):
    """Oracle pipeline: pattern, element loop, Mearge, BC, residual (CPU)."""
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    emat, eres = problem_emats(orc, prob)
    val, dia = orc.assemble(prob.nnode, prob.b, rowptr, col, prob.conn, emat)
    res = orc.assemble_vec(prob.nnode, prob.b, prob.conn, eres)
    val, dia = orc.set_bc(prob.nnode, prob.b, rowptr, col, val, dia, prob.flags)
    res = orc.zero_bc(prob.flags, res)
    return dict(rowptr=rowptr, col=col, val=val, dia=dia, res=res, emat=emat, eres=eres)


def flex_problem(nb=40, maxlen=6, seed=0, nfix=3, density=3, symmetric=True, lens=None, zero_rows=(), window=0):
    """A synthetic system of nb blocks of 1..maxlen dofs (rigid bodies and
    constraints in the reference's rigid-field system are such blocks): a
    random sparse symmetric block pattern (a path plus random links), block
    contributions listed for Mearge (each off-diagonal block merged in two
    parts and every diagonal block in two, so that Mearge accumulates),
    Dirichlet flags on the dofs of nfix random blocks, and a right-hand side.
    symmetric: off-diagonal block (j, i) = (i, j)^T and positive definite
    diagonal dominance (for PCG); otherwise unrelated random blocks.
    lens: the block lengths (instead of random ones); zero_rows: block rows
    that receive no contribution (an exactly zero pivot block); window > 0:
    random links only between blocks at most window apart (bounded ILU(k)
    fill for large nb)."""
    rng = np.random.default_rng(seed)
    alen = rng.integers(1, maxlen + 1, size=nb).astype(np.uint32)
    if lens is not None:
        alen = np.ascontiguousarray(lens, dtype=np.uint32)
    dofptr = np.zeros(nb + 1, dtype=np.uint32)
    dofptr[1:] = np.cumsum(alen)
    adj = [set() for _ in range(nb)]
    for i in range(nb - 1):
        adj[i].add(i + 1)
        adj[i + 1].add(i)
    for _ in range(density * nb):
        i, j = (int(v) for v in rng.integers(0, nb, size=2))
        if window > 0:
            j = min(nb - 1, max(0, i + int(rng.integers(-window, window + 1))))
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    rowptr = np.zeros(nb + 1, dtype=np.uint32)
    cols = []
    for i in range(nb):
        c = sorted(adj[i])
        cols.extend(c)
        rowptr[i + 1] = rowptr[i] + len(c)
    col = np.array(cols, dtype=np.uint32)
    blocks = []
    off = {}
    for i in range(nb):
        for j in sorted(adj[i]):
            if symmetric and j < i:
                continue
            off[(i, j)] = rng.uniform(-1.0, 1.0, size=(alen[i], alen[j]))
            if not symmetric or j > i:
                pass
    rowsum = np.zeros(int(dofptr[-1]))
    for (i, j), v in off.items():
        rowsum[dofptr[i]:dofptr[i + 1]] += np.abs(v).sum(axis=1)
        if symmetric:
            rowsum[dofptr[j]:dofptr[j + 1]] += np.abs(v).sum(axis=0)
    for i in range(nb):
        n = int(alen[i])
        m = rng.uniform(-0.5, 0.5, size=(n, n))
        d = 0.5 * (m + m.T) if symmetric else m
        d = d + np.diag(rowsum[dofptr[i]:dofptr[i + 1]] + n + 1.0)
        h = rng.uniform(0.2, 0.8)
        blocks.append((i, i, h * d))
        blocks.append((i, i, (1.0 - h) * d))
    for (i, j), v in sorted(off.items()):
        h = rng.uniform(0.2, 0.8)
        blocks.append((i, j, h * v))
        blocks.append((i, j, (1.0 - h) * v))
        if symmetric:
            blocks.append((j, i, v.T.copy()))
    order = rng.permutation(len(blocks))
    zr = set(int(v) for v in zero_rows)
    blocks = [blocks[k] for k in order if blocks[k][0] not in zr]
    flags = np.zeros(int(dofptr[-1]), dtype=np.int32)
    for i in rng.choice(nb, size=min(nfix, nb), replace=False):
        flags[dofptr[i]:dofptr[i + 1]] = 1
    rhs = rng.uniform(-1.0, 1.0, size=int(dofptr[-1]))
    rhs[flags != 0] = 0.0
    return dict(nb=nb, alen=alen, dofptr=dofptr, rowptr=rowptr, col=col, blocks=blocks, flags=flags, rhs=rhs)


class RefFlex:
    """One LsSol::CLinearSystem segment of per-block sizes in the reference."""

    def __init__(self, ref, P):
        self.ref, self.P = ref, P
        self._alen, self._index = u32(P["alen"]), u32(P["rowptr"])
        self._array = u32(P["col"]) if P["col"].size else np.zeros(1, dtype=np.uint32)
        self.h = ref.L.ref_fx_create(P["nb"], _p(self._alen), _p(self._index), _p(self._array))

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.L.ref_fx_free(self.h)
            self.h = None

    def merge(self, blocks):
        ib = u32([b[0] for b in blocks])
        jb = u32([b[1] for b in blocks])
        em = f64(np.concatenate([np.ravel(b[2]) for b in blocks]))
        self.ref.L.ref_fx_merge(self.h, ib.size, _p(ib), _p(jb), _p(em))

    def set_bc(self, flags):
        self.ref.L.ref_fx_set_bc(self.h, _p(i32(flags)))

    def finalize(self):
        return self.ref.L.ref_fx_finalize(self.h)

    def matrix(self):
        nb = self.P["nb"]
        ncrs = self.ref.L.ref_fx_ncrs(self.h)
        rowptr = np.zeros(nb + 1, dtype=np.uint32)
        col = np.zeros(max(1, ncrs), dtype=np.uint32)
        valptr = np.zeros(ncrs + 1, dtype=np.uint32)
        diaptr = np.zeros(nb + 1, dtype=np.uint32)
        self.ref.L.ref_fx_get_matrix(self.h, _p(rowptr), _p(col), _p(valptr), _p(diaptr), None, None)
        val = np.zeros(max(1, int(valptr[-1])))
        dia = np.zeros(max(1, int(diaptr[-1])))
        self.ref.L.ref_fx_get_matrix(self.h, None, None, None, None, _p(val), _p(dia))
        return dict(rowptr=rowptr, col=col[:ncrs], valptr=valptr, diaptr=diaptr, val=val[:int(valptr[-1])],
                    dia=dia[:int(diaptr[-1])])

    def set_vector(self, iv, v):
        self.ref.L.ref_fx_set_vector(self.h, iv, _p(f64(v)))

    def get_vector(self, iv):
        out = np.zeros(int(self.P["dofptr"][-1]))
        self.ref.L.ref_fx_get_vector(self.h, iv, _p(out))
        return out

    def matvec(self, alpha, x, beta, y):
        self.ref.L.ref_fx_resize_tmp(self.h, 2)
        self.set_vector(0, x)
        self.set_vector(1, y)
        self.ref.L.ref_fx_matvec(self.h, alpha, 0, beta, 1)
        return self.get_vector(1)

    def factor(self, lev):
        """(ok, F, fval, fdia) of CMatDiaFrac_BlkCrs(lev, A) after SetValue_Initialize + DoILUDecomp."""
        q = self.ref.L.ref_fx_frac_create(self.h, lev)
        try:
            ok = self.ref.L.ref_fx_frac_set_value(q, self.h)
            nb = self.P["nb"]
            ncrs = self.ref.L.ref_fx_frac_ncrs(q)
            F = dict(rowptr=np.zeros(nb + 1, dtype=np.uint32), col=np.zeros(max(1, ncrs), dtype=np.uint32),
                     diaind=np.zeros(max(1, nb), dtype=np.uint32), valptr=np.zeros(ncrs + 1, dtype=np.uint32),
                     diaptr=np.zeros(nb + 1, dtype=np.uint32))
            self.ref.L.ref_fx_frac_get(q, self.h, _p(F["rowptr"]), _p(F["col"]), _p(F["diaind"]), _p(F["valptr"]),
                                       _p(F["diaptr"]), None, None)
            fval = np.zeros(max(1, int(F["valptr"][-1])))
            fdia = np.zeros(max(1, int(F["diaptr"][-1])))
            self.ref.L.ref_fx_frac_get(q, self.h, None, None, None, None, None, _p(fval), _p(fdia))
            F["col"] = F["col"][:ncrs]
            F["dofptr"] = u32(self.P["dofptr"])
            return bool(ok), F, fval[:int(F["valptr"][-1])], fdia[:int(F["diaptr"][-1])]
        finally:
            self.ref.L.ref_fx_frac_free(q)

    def precond(self, lev, vec):
        """Forward then backward substitution with the ILU(lev) factor."""
        q = self.ref.L.ref_fx_frac_create(self.h, lev)
        try:
            self.ref.L.ref_fx_frac_set_value(q, self.h)
            v = f64(vec).copy()
            self.ref.L.ref_fx_frac_forward(q, self.h, _p(v))
            fwd = v.copy()
            self.ref.L.ref_fx_frac_backward(q, self.h, _p(v))
            return fwd, v
        finally:
            self.ref.L.ref_fx_frac_free(q)

    def solve(self, kind, lev, rhs, tol, maxit):
        """The update (-2) is zeroed first: Solve_PBiCGSTAB returns on a zero
        right-hand side before it zeroes the update, which would otherwise
        keep the previous solve's solution."""
        k = {"pcg": 0, "pbicgstab": 1, "cg": 2, "bicgstab": 3}[kind]
        self.set_vector(-2, np.zeros(int(self.P["dofptr"][-1])))
        r = f64(rhs).copy()
        x = np.zeros(r.size)
        conv = C.c_double(tol)
        it = C.c_uint(maxit)
        ok = self.ref.L.ref_fx_solve(self.h, k, lev, _p(r), _p(x), C.byref(conv), C.byref(it))
        return ok, conv.value, it.value, x, r


def flex_assemble(orc, P):
    """Oracle pipeline for a flex problem: layout, Mearge, Dirichlet rows and
    columns; returns the matrix dict M with val and dia."""
    M = dict(dofptr=u32(P["dofptr"]), rowptr=u32(P["rowptr"]), col=u32(P["col"]))
    M["valptr"], M["diaptr"] = orc.fx_layout(M["dofptr"], M["rowptr"], M["col"])
    val, dia = orc.fx_mearge(M, P["blocks"])
    val, dia = orc.fx_set_bc(M, val, dia, P["flags"])
    M["val"], M["dia"] = val, dia
    return M


def flex_dense(M, val, dia):
    """The flex matrix as a dense array (dofptr-indexed)."""
    d = M["dofptr"]
    n = int(d[-1])
    A = np.zeros((n, n))
    nb = M["rowptr"].size - 1
    for i in range(nb):
        li = int(d[i + 1] - d[i])
        A[d[i]:d[i + 1], d[i]:d[i + 1]] = np.asarray(dia[M["diaptr"][i]:M["diaptr"][i] + li * li]).reshape(li, li)
        for q in range(int(M["rowptr"][i]), int(M["rowptr"][i + 1])):
            j = int(M["col"][q])
            lj = int(d[j + 1] - d[j])
            A[d[i]:d[i + 1], d[j]:d[j + 1]] = np.asarray(val[M["valptr"][q]:M["valptr"][q] + li * lj]).reshape(li, lj)
    return A
