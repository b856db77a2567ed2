"""Python mirror of DelFEM's linear-solve API over the delfem_amd C-ABI.

Class and method names follow the reference (include/delfem/matvec/*.h,
include/delfem/ls/*.h) so tests and scripts read like DelFEM code:

    A = CMatDia_BlkCrs(nblk, len); A.AddPattern(index, array)
    ...assemble..., A.SetBoundaryCondition(flags)
    prec = CPreconditioner_ILU(A, lev=0, ordering="multicolour"); prec.SetValue(A)
    ok, conv, it = Solve_PCG(1e-10, 1000, A, prec, r, x)

Vectors are torch tensors on the GPU (torch is used only for device memory);
every kernel is the library's own HIP code.  Calls are synchronous.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call

_ORDER = {"natural": _lib.DFM_ORDER_NATURAL, "given": _lib.DFM_ORDER_GIVEN,
          "multicolour": _lib.DFM_ORDER_MULTICOLOUR, "block_multicolour": _lib.DFM_ORDER_BLOCK_MULTICOLOUR}


def _torch():
    import torch
    return torch


def init(device=0):
    """Select the device; the library then runs on its own stream."""
    call("dfm_init", device)


def _sync_torch():
    torch = _torch()
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def _ptr(a):
    """Host pointer of a contiguous numpy array."""
    return a.ctypes.data_as(C.c_void_p)


def _dptr(t):
    """Device pointer of a contiguous torch tensor."""
    assert t.is_cuda and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def device_vector(n, dtype=None, fill=0.0):
    torch = _torch()
    return torch.full((n,), fill, dtype=dtype or torch.float64, device="cuda")


def to_device(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def fem_pattern(nnode, conn):
    """CElemAry::MakePattern_FEM: (index, array) of the node adjacency."""
    conn = np.ascontiguousarray(conn, dtype=np.uint32)
    nelem, npe = conn.shape
    rowptr = np.zeros(nnode + 1, dtype=np.uint32)
    n = C.c_size_t(0)
    call("dfm_fem_pattern", nnode, nelem, npe, _ptr(conn), _ptr(rowptr), None, 0, C.byref(n))
    col = np.zeros(max(1, n.value), dtype=np.uint32)
    call("dfm_fem_pattern", nnode, nelem, npe, _ptr(conn), _ptr(rowptr), _ptr(col), col.size, C.byref(n))
    return rowptr, col[:n.value]


class CMatDia_BlkCrs:
    """MatVec::CMatDia_BlkCrs (include/delfem/matvec/matdia_blkcrs.h:44-118)."""

    def __init__(self, nblk, len_):
        h = C.c_void_p()
        call("dfm_bcrs_create", nblk, len_, C.byref(h))
        self.h = h
        self.nblk, self.len = nblk, len_

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_bcrs_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def NBlkMatCol(self):
        return self.nblk

    def LenBlkCol(self):
        return self.len

    def NCrs(self):
        n = C.c_uint()
        call("dfm_bcrs_info", self.h, None, None, C.byref(n))
        return n.value

    def AddPattern(self, index, array):
        index = np.ascontiguousarray(index, dtype=np.uint32)
        array = np.ascontiguousarray(array, dtype=np.uint32)
        if array.size == 0:
            array = np.zeros(1, dtype=np.uint32)
        call("dfm_bcrs_add_pattern", self.h, index.size - 1, _ptr(index), _ptr(array))
        return True

    def pattern(self):
        rowptr = np.zeros(self.nblk + 1, dtype=np.uint32)
        col = np.zeros(max(1, self.NCrs()), dtype=np.uint32)
        call("dfm_bcrs_get_pattern", self.h, _ptr(rowptr), _ptr(col))
        return rowptr, col[:self.NCrs()]

    def SetZero(self):
        call("dfm_bcrs_set_zero", self.h)
        return True

    def set_values(self, val, dia):
        bb = self.len * self.len
        val = np.ascontiguousarray(val, dtype=np.float64).reshape(-1)
        dia = np.ascontiguousarray(dia, dtype=np.float64).reshape(-1)
        assert val.size == self.NCrs() * bb and dia.size == self.nblk * bb
        call("dfm_bcrs_set_values", self.h, _ptr(val) if val.size else None, _ptr(dia))

    def values(self):
        bb = self.len * self.len
        val = np.zeros(max(1, self.NCrs() * bb))
        dia = np.zeros(max(1, self.nblk * bb))
        call("dfm_bcrs_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:self.NCrs() * bb], dia[:self.nblk * bb]

    def Mearge(self, nodes, emat):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        emat = np.ascontiguousarray(emat, dtype=np.float64).reshape(-1)
        call("dfm_bcrs_mearge", self.h, nodes.size, _ptr(nodes), self.len * self.len, _ptr(emat))
        return True

    def SetBoundaryCondition(self, flags):
        flags = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        call("dfm_bcrs_set_bc", self.h, _ptr(flags))
        return True

    def MatVec(self, alpha, x, beta, y):
        _sync_torch()
        call("dfm_bcrs_matvec", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))
        return True


class Assembler:
    """Element loops of femeqn (AddLinSys_*) gathered by row on the device."""

    def __init__(self, A, conn):
        conn = np.ascontiguousarray(conn, dtype=np.uint32)
        h = C.c_void_p()
        call("dfm_asm_create", A.h, A.nblk, conn.shape[0], conn.shape[1], _ptr(conn), C.byref(h))
        self.h, self.A = h, A

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_asm_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def poisson2d_p1(self, coord, alpha, source, value, res):
        _sync_torch()
        call("dfm_assemble_poisson2d_p1", self.h, _dptr(coord), _dptr(alpha), float(source), _dptr(value),
             _dptr(res) if res is not None else None)

    def solid3d_p1(self, coord, lam, mu, rho, g, disp, res):
        _sync_torch()
        call("dfm_assemble_solid3d_p1", self.h, _dptr(coord), _dptr(lam), _dptr(mu), float(rho), float(g[0]),
             float(g[1]), float(g[2]), _dptr(disp), _dptr(res) if res is not None else None)

    def oseen2d_p1p1(self, coord, myu, rho, adv, value, res):
        _sync_torch()
        call("dfm_assemble_oseen2d_p1p1", self.h, _dptr(coord), float(myu), float(rho), _dptr(adv), _dptr(value),
             _dptr(res) if res is not None else None)

    def helmholtz2d_p1(self, coord, alpha, cmass, source, value, res):
        _sync_torch()
        call("dfm_assemble_helmholtz2d_p1", self.h, _dptr(coord), _dptr(alpha), _dptr(cmass), float(source),
             _dptr(value), _dptr(res) if res is not None else None)


class CMatDiaFrac_BlkCrs:
    """MatVec::CMatDiaFrac_BlkCrs (include/delfem/matvec/matdiafrac_blkcrs.h:40-137):
    symbolic ILU(lev) at construction, numeric on SetValue."""

    def __init__(self, lev, A, order=None, ordering=None, block_nodes=0, fill_blk=None):
        """fill_blk (natural order only): block rows and columns that take fill
        of any level -- AddFracPtn(lev, aBlkFill) (matdiafrac_blkcrs.cpp:1379-1516),
        what CPreconditioner_ILU::SetFillBlk asks for."""
        if ordering is None:
            ordering = "given" if order is not None else "natural"
        h = C.c_void_p()
        ordarr = None
        if order is not None:
            ordarr = np.ascontiguousarray(order, dtype=np.int32)
        if fill_blk is not None:
            if ordering != "natural":
                raise ValueError("fill blocks need the natural order")
            blk = np.ascontiguousarray(fill_blk, dtype=np.uint32)
            call("dfm_ilu_create_fill", A.h, int(lev), _ptr(blk), blk.size, C.byref(h))
        else:
            call("dfm_ilu_create", A.h, int(lev), _ORDER[ordering], _ptr(ordarr) if ordarr is not None else None,
                 int(block_nodes), C.byref(h))
        self.h, self.nblk, self.len = h, A.nblk, A.len
        self.nsingular = 0

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_ilu_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def info(self):
        ncrs, nl, nu, nc = C.c_uint(), C.c_int(), C.c_int(), C.c_int()
        call("dfm_ilu_info", self.h, C.byref(ncrs), C.byref(nl), C.byref(nu), C.byref(nc))
        return dict(ncrs=ncrs.value, nlevel_lower=nl.value, nlevel_upper=nu.value, ncolour=nc.value)

    def tile_info(self):
        """Phase tiling of the sweeps: phases (colours), tiles per sweep, threads per tile."""
        v = [C.c_int() for _ in range(4)]
        call("dfm_ilu_tile_info", self.h, *(C.byref(x) for x in v))
        return dict(nphase=v[0].value, ntile_lower=v[1].value, ntile_upper=v[2].value, tile_threads=v[3].value)

    def NCrs(self):
        return self.info()["ncrs"]

    def pattern(self):
        n = self.NCrs()
        rowptr = np.zeros(self.nblk + 1, dtype=np.uint32)
        col = np.zeros(max(1, n), dtype=np.uint32)
        lev = np.zeros(max(1, n), dtype=np.int32)
        dia = np.zeros(self.nblk, dtype=np.uint32)
        order = np.zeros(self.nblk, dtype=np.int32)
        call("dfm_ilu_get_pattern", self.h, _ptr(rowptr), _ptr(col), _ptr(lev), _ptr(dia), _ptr(order))
        return dict(rowptr=rowptr, col=col[:n], lev=lev[:n], diaind=dia, order=order)

    def SetValue(self, A):
        ns = C.c_int()
        rc = call("dfm_ilu_set_value", self.h, A.h, C.byref(ns), ok=(_lib.DFM_OK, _lib.DFM_SINGULAR))
        self.nsingular = ns.value
        return rc == _lib.DFM_OK

    def values(self):
        bb = self.len * self.len
        n = self.NCrs()
        val = np.zeros(max(1, n * bb))
        dia = np.zeros(self.nblk * bb)
        call("dfm_ilu_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:n * bb], dia

    def ForwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_ilu_forward", self.h, _dptr(vec))
        return True

    def BackwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_ilu_backward", self.h, _dptr(vec))
        return True

    def SolvePrecond(self, vec):
        """CPreconditioner_ILU::SolvePrecond on a vector in the original numbering."""
        _sync_torch()
        call("dfm_ilu_solve_precond", self.h, _dptr(vec))
        return True


class CPreconditioner_ILU(CMatDiaFrac_BlkCrs):
    """LsSol::CPreconditioner_ILU (include/delfem/ls/preconditioner.h:59-108) on one segment."""

    def __init__(self, A, nlev=0, order=None, ordering=None, block_nodes=0):
        super().__init__(nlev, A, order=order, ordering=ordering, block_nodes=block_nodes)


def _solve(fn, conv, it, A, prec, r, x, precision, opts_kw):
    _sync_torch()
    opts = _lib.SolveOpts()
    opts.precision = _lib.PRECISIONS[precision]
    for k, v in (opts_kw or {}).items():
        setattr(opts, k, v)
    rc = call(fn, A.h, prec.h, _dptr(r), _dptr(x), C.byref(conv), C.byref(it), C.byref(opts),
              ok=(_lib.DFM_OK, _lib.DFM_NOT_CONVERGED))
    info = dict(inner_iters=opts.inner_iters, outer_iters=opts.outer_iters, true_relres=opts.true_relres,
                converged=rc == _lib.DFM_OK)
    return info


def Solve_PCG(conv_ratio, iteration, A, prec, r, x, precision="fp64", **opts):
    """LsSol::Solve_PCG (src/ls/solver_ls_iter.cpp:100-172).  Returns
    (ok, conv_ratio, iteration, info): ok as the reference's result, true when
    converged, false at max_iter; info["converged"] the same; r is
    overwritten by the residual."""
    conv, it = C.c_double(conv_ratio), C.c_uint(iteration)
    info = _solve("dfm_solve_pcg", conv, it, A, prec, r, x, precision, opts)
    return info["converged"], conv.value, it.value, info


def Solve_PBiCGSTAB(conv_ratio, iteration, A, prec, r, x, precision="fp64", **opts):
    """LsSol::Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:291-408).  As the
    reference it returns true also when stopped by max_iter (then iteration
    = max_iter and conv_ratio unchanged); info["converged"] tells them apart."""
    conv, it = C.c_double(conv_ratio), C.c_uint(iteration)
    info = _solve("dfm_solve_pbicgstab", conv, it, A, prec, r, x, precision, opts)
    return True, conv.value, it.value, info


def set_mseg_device_loop(on):
    """Multi-segment Solve_PCG / Solve_PBiCGSTAB device-driven (True, the
    default: scalars on the device, the stop flag polled every four
    iterations) or host-driven (one synchronisation per DOT); same bits."""
    call("dfm_mseg_set_device_loop", 1 if on else 0)


def set_sweep_mode(mode, grid=0):
    """Level sets of the triangular sweeps: mode 0 one launch per level; 1
    (default) for factors of 8 or more levels one launch per phase (dependent
    rows tiled into workgroups; phases are the colours of a colour ordering,
    or found in a natural-order factor's numbering), else in the solvers one
    persistent launch; 2 persistent always; 3 per phase wherever possible;
    4 as 3 with one lane per block row of 2x2 / 3x3 blocks (3: one per tile
    row); grid = workgroups of a persistent launch (0: automatic)."""
    call("dfm_set_sweep_mode", int(mode), int(grid))


def dot(x, y):
    _sync_torch()
    out = C.c_double()
    call("dfm_vec_dot", x.numel(), _dptr(x), _dptr(y), C.byref(out))
    return out.value


class FEMSystem:
    """A Problem (delfem_amd.meshgen) assembled on the device: matrix,
    residual, Dirichlet conditions -- the InitializeMarge / AddLinSys_* /
    FinalizeMarge sequence of src/ls/linearsystem.cpp:74-128."""

    def __init__(self, prob):
        torch = _torch()
        self.prob = prob
        rowptr, col = fem_pattern(prob.nnode, prob.conn)
        self.A = CMatDia_BlkCrs(prob.nnode, prob.b)
        self.A.AddPattern(rowptr, col)
        self.asm = Assembler(self.A, prob.conn)
        self.coord = to_device(prob.coord.ravel())
        if prob.kind == "poisson2d":
            self.coef = (to_device(prob.alpha),)
        elif prob.kind == "helmholtz2d":
            self.coef = (to_device(prob.alpha), to_device(prob.cmass))
        elif prob.kind == "oseen2d":
            self.coef = (to_device(prob.adv.ravel()),)
        else:
            self.coef = (to_device(prob.lam), to_device(prob.mu))
        self.res = torch.zeros(prob.ndof, dtype=torch.float64, device="cuda")
        if getattr(prob, "value", None) is not None:
            self.value = to_device(prob.value)
        else:
            self.value = torch.zeros(prob.ndof, dtype=torch.float64, device="cuda")
        self.flags = np.ascontiguousarray(prob.flags, dtype=np.int32)
        self.flags_dev = to_device(self.flags)

    def assemble_elements(self):
        """InitializeMarge + the element loop (matrix and residual, no
        boundary conditions yet)."""
        self.A.SetZero()
        self.res.zero_()
        p = self.prob
        if p.kind == "poisson2d":
            self.asm.poisson2d_p1(self.coord, self.coef[0], p.source, self.value, self.res)
        elif p.kind == "helmholtz2d":
            self.asm.helmholtz2d_p1(self.coord, self.coef[0], self.coef[1], p.source, self.value, self.res)
        elif p.kind == "oseen2d":
            self.asm.oseen2d_p1p1(self.coord, p.myu, p.rho, self.coef[0], self.value, self.res)
        else:
            self.asm.solid3d_p1(self.coord, self.coef[0], self.coef[1], p.rho, p.g, self.value, self.res)

    def assemble(self):
        """InitializeMarge + element loop + FinalizeMarge; returns ||r||."""
        self.assemble_elements()
        self.A.SetBoundaryCondition(self.flags)
        _sync_torch()
        call("dfm_vec_zero_bc", self.res.numel(), _ptr(self.flags), _dptr(self.res))
        return dot(self.res, self.res) ** 0.5

    def UpdateValueOfField(self, update):
        """CLinearSystem_Field::UpdateValueOfField (src/femls/linearsystem_field.cpp:715-793)
        for the node values: value += update (CNodeAry::AddValueToNodeSegment with
        factor 1.0; a combined (u, v, p) segment updates its components in block
        order).  The next assemble() forms the residual at the new values."""
        _sync_torch()
        if update.numel() != self.value.numel():
            raise ValueError("update and field sizes differ")
        call("dfm_vec_axpy", self.value.numel(), 1.0, _dptr(update), _dptr(self.value))


class CMat_BlkCrs:
    """MatVec::CMat_BlkCrs (include/delfem/matvec/mat_blkcrs.h:45-197) as the
    boundary matrix of the saved-stiffness system: square, square blocks."""

    def __init__(self, nblk_col, len_col, nblk_row, len_row):
        h = C.c_void_p()
        call("dfm_bmat_create", nblk_col, len_col, nblk_row, len_row, C.byref(h))
        self.h, self.nblk, self.len, self.len_row = h, nblk_col, len_col, len_row

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_bmat_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def NCrs(self):
        n = C.c_uint()
        call("dfm_bmat_info", self.h, None, None, C.byref(n))
        return n.value

    def SetPatternBoundary(self, A, flag_col, flag_row):
        fc = np.ascontiguousarray(flag_col, dtype=np.int32).reshape(-1)
        fr = np.ascontiguousarray(flag_row, dtype=np.int32).reshape(-1)
        call("dfm_bmat_set_pattern_boundary", self.h, A.h, _ptr(fc), _ptr(fr))
        return True

    def pattern(self):
        n = self.NCrs()
        rowptr = np.zeros(self.nblk + 1, dtype=np.uint32)
        col = np.zeros(max(1, n), dtype=np.uint32)
        call("dfm_bmat_get_pattern", self.h, _ptr(rowptr), _ptr(col))
        return rowptr, col[:n]

    def SetZero(self):
        call("dfm_bmat_set_zero", self.h)
        return True

    def values(self):
        n = self.NCrs() * self.len * self.len_row
        val = np.zeros(max(1, n))
        call("dfm_bmat_get_values", self.h, _ptr(val))
        return val[:n]

    def Mearge(self, nodes, emat):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        emat = np.ascontiguousarray(emat, dtype=np.float64).reshape(-1)
        call("dfm_bmat_mearge", self.h, nodes.size, _ptr(nodes), self.len * self.len, _ptr(emat))
        return True

    def CopyFrom(self, A):
        """The entries of the assembled diagonal-block matrix A (before its
        boundary conditions): the same sums as merging the elements here."""
        call("dfm_bmat_copy_from", self.h, A.h)
        return True

    def SetBoundaryConditionInverse_Colum(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        call("dfm_bmat_set_bc_inverse_col", self.h, _ptr(f))
        return True

    def SetBoundaryCondition_Row(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        call("dfm_bmat_set_bc_row", self.h, _ptr(f))
        return True

    def MatVec(self, alpha, x, beta, y):
        _sync_torch()
        call("dfm_bmat_matvec", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))
        return True


class FEMSystemSave(FEMSystem):
    """Fem::Ls::CLinearSystem_Save (src/femls/linearsystem_fieldsave.cpp) for
    a linear static Problem: the assembly also keeps the force f (element
    forces alone) and the boundary matrix K_b (the stiffness between free rows
    and fixed columns); MakeResidual forms r = f - K_b u for the current
    field u without re-assembly, the solve gives the free values directly and
    UpdateValueOfField sets them (fixed dofs keep their prescribed values)."""

    def __init__(self, prob):
        torch = _torch()
        super().__init__(prob)
        self.Kb = CMat_BlkCrs(prob.nnode, prob.b, prob.nnode, prob.b)
        self.Kb.SetPatternBoundary(self.A, self.flags, self.flags)
        self.force = torch.zeros(prob.ndof, dtype=torch.float64, device="cuda")

    def assemble(self):
        """InitializeMarge + element loop + FinalizeMarge of the saved system
        (linearsystem_fieldsave.cpp:166-204): the element loop runs at u = 0,
        so the assembled residual is the force."""
        torch = _torch()
        value = self.value
        self.value = torch.zeros_like(value)
        try:
            self.A.SetZero()
            self.res.zero_()
            p = self.prob
            if p.kind == "poisson2d":
                self.asm.poisson2d_p1(self.coord, self.coef[0], p.source, self.value, self.res)
            elif p.kind == "helmholtz2d":
                self.asm.helmholtz2d_p1(self.coord, self.coef[0], self.coef[1], p.source, self.value, self.res)
            elif p.kind == "solid3d":
                self.asm.solid3d_p1(self.coord, self.coef[0], self.coef[1], p.rho, p.g, self.value, self.res)
            else:
                raise ValueError(f"no saved-stiffness assembly for {p.kind}")
        finally:
            self.value = value
        self.force.copy_(self.res)
        self.Kb.CopyFrom(self.A)
        self.A.SetBoundaryCondition(self.flags)
        self.Kb.SetBoundaryConditionInverse_Colum(self.flags)
        self.Kb.SetBoundaryCondition_Row(self.flags)
        return 0.0

    def MakeResidual(self):
        """r = f - K_b u at the current field values (linearsystem_fieldsave.cpp:206-247); returns ||r||."""
        _sync_torch()
        call("dfm_vec_copy", self.res.numel(), _dptr(self.force), _dptr(self.res))
        self.Kb.MatVec(-1.0, self.value, 1.0, self.res)
        return dot(self.res, self.res) ** 0.5

    def UpdateValueOfField(self, update):
        """CLinearSystem_Save::UpdateValueOfField (linearsystem_fieldsave.cpp:249-300):
        the field takes the solution at free dofs; fixed dofs keep their values."""
        _sync_torch()
        if update.numel() != self.value.numel():
            raise ValueError("update and field sizes differ")
        call("dfm_vec_copy_free", self.value.numel(), _ptr(self.flags.reshape(-1)), _dptr(update), _dptr(self.value))


class MultiSegmentSystem:
    """LsSol::CLinearSystem with several segments (src/ls/linearsystem.cpp):
    a combined-unknown Problem split into the given dof groups (e.g.
    [[0, 1], [2]] for velocity and pressure), as the reference's split-field
    equation systems assemble it.  The combined matrix is assembled on the
    device and its sub-blocks taken into the segments' diagonal blocks
    (CMatDia_BlkCrs) and off-diagonal blocks (CMat_BlkCrs, pattern: the node
    adjacency with the node itself); FinalizeMarge sets the boundary
    conditions per block.  Vectors are the segments' vectors one after the
    other (torch tensors on the device)."""

    def __init__(self, prob, groups, ordering="natural", block_nodes=16):
        """ordering "multicolour": the mesh nodes are first renumbered by the
        greedy multicolouring of the node graph (the same numbering for every
        segment), so that each segment's natural-order ILU(0) is a colour-
        ordered one with one level set per colour; "block_multicolour": by
        colours of node blocks (about block_nodes each); "tile_multicolour":
        by colours of geometric tiles (meshgen.tile_colour_order), which keeps
        the level count low also with fill (ILU(1)).  Vectors are then in that
        numbering (self.node_order[new] = old node)."""
        torch = _torch()
        self.node_order = None
        if ordering == "tile_multicolour":
            from . import meshgen
            self.node_order = meshgen.tile_colour_order(prob, block_nodes)
            prob = meshgen.relabel(prob, self.node_order)
        elif ordering in ("multicolour", "block_multicolour"):
            rp0, cl0 = fem_pattern(prob.nnode, prob.conn)
            A0 = CMatDia_BlkCrs(prob.nnode, 1)
            A0.AddPattern(rp0, cl0)
            F0 = CMatDiaFrac_BlkCrs(0, A0, ordering=ordering, block_nodes=block_nodes)
            self.node_order = F0.pattern()["order"].copy()
            from . import meshgen
            prob = meshgen.relabel(prob, self.node_order)
        elif ordering != "natural":
            raise ValueError(f"unknown ordering {ordering}")
        self.prob, self.groups = prob, [list(g) for g in groups]
        self.fem = FEMSystem(prob)
        nb, B = prob.nnode, prob.b
        rowptr, col = fem_pattern(nb, prob.conn)
        rp2 = np.zeros(nb + 1, dtype=np.uint32)
        cols = []
        for i in range(nb):
            c = np.sort(np.append(col[rowptr[i]:rowptr[i + 1]], i)).astype(np.uint32)
            cols.append(c)
            rp2[i + 1] = rp2[i] + c.size
        col2 = np.concatenate(cols).astype(np.uint32)
        nseg = len(self.groups)
        self.lens = [len(g) for g in self.groups]
        self.dia = []
        for g in self.groups:
            D = CMatDia_BlkCrs(nb, len(g))
            D.AddPattern(rowptr, col)
            self.dia.append(D)
        self.nd = {}
        for i, gi in enumerate(self.groups):
            for j, gj in enumerate(self.groups):
                if i != j:
                    M = CMat_BlkCrs(nb, len(gi), nb, len(gj))
                    call("dfm_bmat_add_pattern", M.h, nb, _ptr(rp2), _ptr(col2))
                    self.nd[(i, j)] = M
        nbs = np.array([nb] * nseg, dtype=np.uint32)
        lens = np.array(self.lens, dtype=np.uint32)
        h = C.c_void_p()
        call("dfm_mseg_create", nseg, _ptr(nbs), _ptr(lens), C.byref(h))
        self.h = h
        for i in range(nseg):
            call("dfm_mseg_set_dia", h, i, self.dia[i].h)
        for (i, j), M in self.nd.items():
            call("dfm_mseg_set_nondia", h, i, j, M.h)
        flags = np.ascontiguousarray(prob.flags, dtype=np.int32).reshape(nb, B)
        self.flags = [np.ascontiguousarray(flags[:, g]).reshape(-1) for g in self.groups]
        self.flags_all = np.concatenate(self.flags)
        # combined dof of each segment-ordered dof
        self.perm = np.concatenate([(np.arange(nb)[:, None] * B + np.array(g)[None, :]).reshape(-1)
                                    for g in self.groups]).astype(np.int64)
        self.res = torch.zeros(nb * B, dtype=torch.float64, device="cuda")
        self.rowptr_nd, self.col_nd = rp2, col2

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_mseg_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def assemble(self):
        """InitializeMarge, the element loop of the combined problem split into
        the blocks, FinalizeMarge (diagonal SetBoundaryCondition, off-diagonal
        SetBoundaryCondition_Row then _Colum, residual zeroed at fixed dofs);
        returns ||r||."""
        self.fem.assemble_elements()
        A = self.fem.A
        for i, g in enumerate(self.groups):
            d = np.array(g, dtype=np.uint32)
            call("dfm_bcrs_extract", self.dia[i].h, A.h, _ptr(d))
        for (i, j), M in self.nd.items():
            di, dj = np.array(self.groups[i], dtype=np.uint32), np.array(self.groups[j], dtype=np.uint32)
            call("dfm_bmat_extract", M.h, A.h, _ptr(di), _ptr(dj))
        for i in range(len(self.groups)):
            self.dia[i].SetBoundaryCondition(self.flags[i])
        for (i, j), M in self.nd.items():
            M.SetBoundaryCondition_Row(self.flags[i])
        for (i, j), M in self.nd.items():
            call("dfm_bmat_set_bc_col", M.h, _ptr(self.flags[j]))
        _sync_torch()
        call("dfm_vec_gather", self.perm.size, _ptr(self.perm), _dptr(self.fem.res), _dptr(self.res))
        call("dfm_vec_zero_bc", self.res.numel(), _ptr(self.flags_all), _dptr(self.res))
        return self.dot(self.res, self.res) ** 0.5

    def dot(self, x, y):
        _sync_torch()
        out = C.c_double()
        call("dfm_mseg_dot", self.h, _dptr(x), _dptr(y), C.byref(out))
        return out.value

    def MATVEC(self, alpha, x, beta, y):
        _sync_torch()
        call("dfm_mseg_matvec", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))

    def preconditioner(self, lev=0, inputs=None, fill_blk=None):
        return MultiSegmentILU(self, lev, inputs, fill_blk)


def fill_levels(nseg, inputs):
    """The fill level of each block (i, j) that SetFillInLevel(lev, ilss)
    requests resolve to, in order (src/ls/preconditioner.cpp:108-122): ilss
    -1 sets every block, otherwise segment ilss's block row and column."""
    alev = np.zeros((nseg, nseg), dtype=np.int32)
    for lev, ilss in inputs:
        if ilss == -1:
            alev[:, :] = lev
        elif 0 <= ilss < nseg:
            alev[:, ilss] = lev
            alev[ilss, :] = lev
        else:
            raise ValueError(f"SetFillInLevel: segment {ilss} beyond the {nseg} segments")
    return alev


class MultiSegmentILU:
    """CPreconditioner_ILU(ls, lev) across the segments
    (src/ls/preconditioner.cpp:57-335): lev 0 no fill, -1 unlimited; inputs:
    SetFillInLevel(lev, ilss) requests [(lev, ilss), ...] instead of lev;
    fill_blk: SetFillBlk's rows of segment 0 (fill of any level there)."""

    def __init__(self, S, lev=0, inputs=None, fill_blk=None):
        h = C.c_void_p()
        if inputs is None and fill_blk is None:
            call("dfm_mseg_ilu_create", S.h, int(lev), C.byref(h))
        else:
            inputs = [(lev, -1)] if inputs is None else inputs
            alev = np.ascontiguousarray(fill_levels(len(S.groups), inputs), dtype=np.int32)
            blk = np.ascontiguousarray(fill_blk if fill_blk is not None else [], dtype=np.uint32)
            call("dfm_mseg_ilu_create_levels", S.h, _ptr(alev), _ptr(blk) if blk.size else None, blk.size,
                 C.byref(h))
        self.h, self.S, self.lev = h, S, lev

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_mseg_ilu_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def SetValue(self):
        ns = C.c_int()
        rc = call("dfm_mseg_ilu_set_value", self.h, C.byref(ns), ok=(_lib.DFM_OK, _lib.DFM_SINGULAR))
        return rc == _lib.DFM_OK

    def pattern(self, i, j):
        """Factor pattern (rowptr, col) of block (i, j), fill included."""
        n = C.c_uint()
        call("dfm_mseg_ilu_get_pattern", self.h, i, j, C.byref(n), None, None)
        rp = np.zeros(self.S.prob.nnode + 1, dtype=np.uint32)
        cl = np.zeros(max(1, n.value), dtype=np.uint32)
        call("dfm_mseg_ilu_get_pattern", self.h, i, j, C.byref(n), _ptr(rp), _ptr(cl))
        return rp, cl[:n.value]

    def values(self, i, j):
        S = self.S
        nb, bi, bj = S.prob.nnode, S.lens[i], S.lens[j]
        nc = C.c_uint()
        call("dfm_mseg_ilu_get_pattern", self.h, i, j, C.byref(nc), None, None)
        n = int(nc.value)
        if i == j:
            val, dia = np.zeros(max(1, n * bi * bi)), np.zeros(nb * bi * bi)
            call("dfm_mseg_ilu_get_values", self.h, i, j, _ptr(val), _ptr(dia))
            return val[:n * bi * bi], dia
        val = np.zeros(max(1, n * bi * bj))
        call("dfm_mseg_ilu_get_values", self.h, i, j, _ptr(val), None)
        return val[:n * bi * bj]

    def SolvePrecond(self, vec):
        _sync_torch()
        call("dfm_mseg_ilu_precond", self.h, _dptr(vec))

    def solve(self, kind, conv_ratio, iteration, r, x):
        """LsSol::Solve_PCG (kind "pcg") or Solve_PBiCGSTAB ("bicg") on the
        segments; returns (converged, conv_ratio, iteration)."""
        _sync_torch()
        conv = C.c_double(conv_ratio)
        it = C.c_uint(iteration)
        fn = "dfm_mseg_solve_pcg" if kind == "pcg" else "dfm_mseg_solve_pbicgstab"
        rc = call(fn, self.S.h, self.h, _dptr(r), _dptr(x), C.byref(conv), C.byref(it),
                  ok=(_lib.DFM_OK, _lib.DFM_NOT_CONVERGED))
        return rc == _lib.DFM_OK, conv.value, it.value


class BatchSystem:
    """Many independent scalar systems (the diagonal blocks of one assembled
    FEMSystem) solved together: dfm_batch_* -- per system the reference's
    ordered ILU(0) and Solve_PCG, all systems in one persistent launch."""

    SHAPES = ((256, 1), (256, 2), (256, 4), (512, 1), (512, 2), (512, 4), (1024, 1), (1024, 2))

    def __init__(self, fem, sys_ptr):
        self.fem = fem
        self.sys_ptr = np.ascontiguousarray(sys_ptr, dtype=np.uint32)
        self.nsys = self.sys_ptr.size - 1
        h = C.c_void_p()
        call("dfm_batch_create", fem.A.h, self.nsys, _ptr(self.sys_ptr), C.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_batch_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def info(self):
        w, nc, bl, wp = C.c_uint(), C.c_int(), C.c_int(), C.c_int()
        call("dfm_batch_info", self.h, C.byref(w), C.byref(nc), C.byref(bl), C.byref(wp))
        return dict(max_width=w.value, max_colours=nc.value, block=bl.value, wg_per_cu=wp.value)

    def SetValue(self):
        _sync_torch()
        call("dfm_batch_set_value", self.h, self.fem.A.h)

    def set_launch(self, block, wg_per_cu):
        call("dfm_batch_set_launch", self.h, int(block), int(wg_per_cu))

    def set_teams(self, min_rows, max_team):
        """Systems of at least min_rows rows get teams of up to max_team workgroups."""
        call("dfm_batch_set_teams", self.h, int(min_rows), int(max_team))

    def solve(self, rhs, x, tol, maxit, precision="mixed_f32"):
        """Returns (all_converged, dict of per-system arrays)."""
        _sync_torch()
        n = self.nsys
        it = np.zeros(n, dtype=np.uint32)
        outer = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        rel = np.zeros(n)
        ms = np.zeros(2 * n)
        rc = call("dfm_batch_solve_pcg", self.h, _dptr(rhs), _dptr(x), float(tol), int(maxit),
                  _lib.PRECISIONS[precision], _ptr(it), _ptr(outer), _ptr(st), _ptr(rel), _ptr(ms),
                  ok=(_lib.DFM_OK, _lib.DFM_NOT_CONVERGED))
        return rc == _lib.DFM_OK, dict(iters=it, outer=outer, status=st, relres=rel, solve_ms=ms[0::2],
                                       done_ms=ms[1::2])

    def factor(self, s):
        """(val, dia, order) of system s after a solve (colour numbering)."""
        n = int(self.sys_ptr[s + 1] - self.sys_ptr[s])
        rp, _ = self.fem.A.pattern()
        nnz = int(rp[self.sys_ptr[s + 1]] - rp[self.sys_ptr[s]])
        val = np.zeros(max(nnz, 1))
        dia = np.zeros(n)
        order = np.zeros(n, dtype=np.int32)
        call("dfm_batch_get_factor", self.h, int(s), _ptr(val), _ptr(dia), _ptr(order))
        return val[:nnz], dia, order

    # team axis: (rows from which a system gets a team, most workgroups per team)
    TEAMS = ((16384, 4), (16384, 8), (8192, 4), (8192, 3), (12288, 3))

    def autotune(self, rhs, x, tol, maxit, precision="mixed_f32", reps=2, shapes=None, teams=None):
        """Time each launch shape (threads per workgroup, workgroups per CU)
        with each team setting on this batch (graph replay, HIP events on the
        library stream) and keep the fastest.  Returns {(block, wg_per_cu,
        team_min_rows, team_max): ms}."""
        torch = _torch()
        stream = torch.cuda.ExternalStream(int(_lib.lib().dfm_stream()))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        times = {}
        for bl, wp in (shapes or self.SHAPES):
            for tr, tk in (teams or self.TEAMS):
                self.set_launch(bl, wp)
                self.set_teams(tr, tk)
                self.solve(rhs, x, tol, maxit, precision)
                best = float("inf")
                for _ in range(reps):
                    torch.cuda.synchronize()
                    ev[0].record(stream)
                    self.solve(rhs, x, tol, maxit, precision)
                    ev[1].record(stream)
                    ev[1].synchronize()
                    best = min(best, ev[0].elapsed_time(ev[1]))
                times[(bl, wp, tr, tk)] = best
        bl, wp, tr, tk = min(times, key=times.get)
        self.set_launch(bl, wp)
        self.set_teams(tr, tk)
        return times


# ---------------------------------------------------------------------------
# Complex twin.  Complex vectors are torch complex128 tensors on the GPU
# (interleaved re, im doubles, the layout of Com::Complex); host arrays are
# numpy complex128.

def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)


class CZMatDia_BlkCrs:
    """MatVec::CZMatDia_BlkCrs (include/delfem/matvec/zmatdia_blkcrs.h:39-100),
    scalar blocks (len 1) as the reference's complex solvers use it."""

    def __init__(self, nblk, len_=1):
        h = C.c_void_p()
        call("dfm_zbcrs_create", nblk, len_, C.byref(h))
        self.h = h
        self.nblk, self.len = nblk, len_

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_zbcrs_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def NBlkMatCol(self):
        return self.nblk

    def LenBlkCol(self):
        return self.len

    def NCrs(self):
        n = C.c_uint()
        call("dfm_zbcrs_info", self.h, None, None, C.byref(n))
        return n.value

    def AddPattern(self, index, array):
        index = np.ascontiguousarray(index, dtype=np.uint32)
        array = np.ascontiguousarray(array, dtype=np.uint32)
        if array.size == 0:
            array = np.zeros(1, dtype=np.uint32)
        call("dfm_zbcrs_add_pattern", self.h, index.size - 1, _ptr(index), _ptr(array))
        return True

    def pattern(self):
        rowptr = np.zeros(self.nblk + 1, dtype=np.uint32)
        col = np.zeros(max(1, self.NCrs()), dtype=np.uint32)
        call("dfm_zbcrs_get_pattern", self.h, _ptr(rowptr), _ptr(col))
        return rowptr, col[:self.NCrs()]

    def SetZero(self):
        call("dfm_zbcrs_set_zero", self.h)
        return True

    def set_values(self, val, dia):
        val, dia = _c128(val), _c128(dia)
        assert val.size == self.NCrs() and dia.size == self.nblk
        call("dfm_zbcrs_set_values", self.h, _ptr(val) if val.size else None, _ptr(dia))

    def values(self):
        n = self.NCrs()
        val = np.zeros(max(1, n), dtype=np.complex128)
        dia = np.zeros(max(1, self.nblk), dtype=np.complex128)
        call("dfm_zbcrs_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:n], dia[:self.nblk]

    def Mearge(self, nodes, emat):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        emat = _c128(emat)
        call("dfm_zbcrs_mearge", self.h, nodes.size, _ptr(nodes), 1, _ptr(emat))
        return True

    def SetBoundaryCondition(self, flags):
        flags = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        call("dfm_zbcrs_set_bc", self.h, _ptr(flags))
        return True

    def MatVec(self, alpha, x, beta, y):
        _sync_torch()
        call("dfm_zbcrs_matvec", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))
        return True

    def MatVec_Hermitian(self, alpha, x, beta, y):
        """y = alpha A^H x + beta y (src/matvec/zmatdia_blkcrs.cpp:803-847)."""
        _sync_torch()
        call("dfm_zbcrs_matvec_hermitian", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))
        return True


class ZAssembler:
    """Element loops of src/femeqn/eqn_helmholtz.cpp gathered by row on the device."""

    def __init__(self, A, conn):
        conn = np.ascontiguousarray(conn, dtype=np.uint32)
        h = C.c_void_p()
        call("dfm_zasm_create", A.h, A.nblk, conn.shape[0], conn.shape[1], _ptr(conn), C.byref(h))
        self.h, self.A = h, A

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_asm_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def helmholtz2d_p1(self, coord, wave_length, value, res):
        _sync_torch()
        call("dfm_zassemble_helmholtz2d_p1", self.A.h, self.h, _dptr(coord), float(wave_length), _dptr(value),
             _dptr(res) if res is not None else None)

    def sommerfeld_b1(self, coord, wave_length, value, res):
        _sync_torch()
        call("dfm_zassemble_sommerfeld_b1", self.A.h, self.h, _dptr(coord), float(wave_length), _dptr(value),
             _dptr(res) if res is not None else None)


class CZMatDiaFrac_BlkCrs:
    """MatVec::CZMatDiaFrac_BlkCrs (include/delfem/matvec/zmatdiafrac_blkcrs.h:44-96):
    symbolic ILU(lev) at construction, numeric on SetValue.  ordering
    "natural" is the reference's; "multicolour" / "given" reorder for the
    level-scheduled device sweeps."""

    def __init__(self, lev, A, order=None, ordering=None):
        if ordering is None:
            ordering = "given" if order is not None else "natural"
        ordarr = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        h = C.c_void_p()
        call("dfm_zilu_create", A.h, int(lev), _ORDER[ordering], _ptr(ordarr) if ordarr is not None else None,
             C.byref(h))
        self.h, self.nblk = h, A.nblk
        self.nsmall = 0

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_zilu_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def info(self):
        ncrs, nl, nu, nc = C.c_uint(), C.c_int(), C.c_int(), C.c_int()
        call("dfm_zilu_info", self.h, C.byref(ncrs), C.byref(nl), C.byref(nu), C.byref(nc))
        return dict(ncrs=ncrs.value, nlevel_lower=nl.value, nlevel_upper=nu.value, ncolour=nc.value)

    def NCrs(self):
        return self.info()["ncrs"]

    def pattern(self):
        n = self.NCrs()
        rowptr = np.zeros(self.nblk + 1, dtype=np.uint32)
        col = np.zeros(max(1, n), dtype=np.uint32)
        lev = np.zeros(max(1, n), dtype=np.int32)
        dia = np.zeros(self.nblk, dtype=np.uint32)
        order = np.zeros(self.nblk, dtype=np.int32)
        call("dfm_zilu_get_pattern", self.h, _ptr(rowptr), _ptr(col), _ptr(lev), _ptr(dia), _ptr(order))
        return dict(rowptr=rowptr, col=col[:n], lev=lev[:n], diaind=dia, order=order)

    def SetValue(self, A):
        """SetValue + DoILUDecomp; self.nsmall counts pivots with |d|^2 < 1e-20."""
        ns = C.c_int()
        call("dfm_zilu_set_value", self.h, A.h, C.byref(ns))
        self.nsmall = ns.value
        return True

    def values(self):
        n = self.NCrs()
        val = np.zeros(max(1, n), dtype=np.complex128)
        dia = np.zeros(self.nblk, dtype=np.complex128)
        call("dfm_zilu_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:n], dia

    def ForwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_zilu_forward", self.h, _dptr(vec))
        return True

    def BackwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_zilu_backward", self.h, _dptr(vec))
        return True

    def SolvePrecond(self, vec):
        """Solve (forward + backward) on a vector in the original numbering."""
        _sync_torch()
        call("dfm_zilu_solve_precond", self.h, _dptr(vec))
        return True


def Solve_PCOCG(conv_ratio, iteration, A, prec, r, u):
    """MatVec::Sol::Solve_PCOCG (src/matvec/zsolver_mat_iter.cpp:303-392).
    Returns (True, conv_ratio, iteration) as the reference (true also at
    max_iter); r is overwritten by the final residual, u by the solution."""
    _sync_torch()
    conv = C.c_double(conv_ratio)
    it = C.c_uint(iteration)
    call("dfm_zsolve_pcocg", A.h, prec.h, _dptr(r), _dptr(u), C.byref(conv), C.byref(it))
    return True, conv.value, it.value


def _zsolve(fn, conv_ratio, iteration, A, prec, r, u):
    _sync_torch()
    conv = C.c_double(conv_ratio)
    it = C.c_uint(iteration)
    args = (A.h,) + ((prec.h,) if prec is not None else ()) + (_dptr(r), _dptr(u), C.byref(conv), C.byref(it))
    call(fn, *args)
    return True, conv.value, it.value


def Solve_CG_Hermitian(conv_ratio, iteration, A, r, u):
    """MatVec::Sol::Solve_CG of the complex classes (src/matvec/zsolver_mat_iter.cpp:32-119):
    CG for a Hermitian complex matrix."""
    return _zsolve("dfm_zsolve_cg", conv_ratio, iteration, A, None, r, u)


def Solve_PCG_Hermitian(conv_ratio, iteration, A, prec, r, u):
    """MatVec::Sol::Solve_PCG of the complex classes (zsolver_mat_iter.cpp:121-218) with
    the complex ILU factor; returns (converged, conv_ratio, iteration) -- False at
    max_iter, as the reference."""
    _sync_torch()
    conv = C.c_double(conv_ratio)
    it = C.c_uint(iteration)
    rc = call("dfm_zsolve_pcg", A.h, prec.h, _dptr(r), _dptr(u), C.byref(conv), C.byref(it),
              ok=(_lib.DFM_OK, _lib.DFM_NOT_CONVERGED))
    return rc == _lib.DFM_OK, conv.value, it.value


def Solve_CGNR(conv_ratio, iteration, A, r, u):
    """MatVec::Sol::Solve_CGNR (src/matvec/zsolver_mat_iter.cpp:394-483): CG on
    the normal equations, for any nonsingular complex matrix."""
    return _zsolve("dfm_zsolve_cgnr", conv_ratio, iteration, A, None, r, u)


def Solve_COCG(conv_ratio, iteration, A, r, u):
    """MatVec::Sol::Solve_COCG (src/matvec/zsolver_mat_iter.cpp:220-299)."""
    return _zsolve("dfm_zsolve_cocg", conv_ratio, iteration, A, None, r, u)


def Solve_BiCGSTAB(conv_ratio, iteration, A, r, u):
    """MatVec::Sol::Solve_BiCGSTAB for the complex matrix (zsolver_mat_iter.cpp:487-616),
    with the reference's conjugated inner products (see include/dfm.h)."""
    return _zsolve("dfm_zsolve_bicgstab", conv_ratio, iteration, A, None, r, u)


def Solve_BiCGStabP(conv_ratio, iteration, A, prec, r, u):
    """MatVec::Sol::Solve_BiCGStabP (zsolver_mat_iter.cpp:622-739) with the complex ILU factor."""
    return _zsolve("dfm_zsolve_bicgstabp", conv_ratio, iteration, A, prec, r, u)


def zdot(x, y):
    """CZVector_Blk operator* (zvector_blk.cpp:39-60): sum x_i y_i, no conjugate."""
    _sync_torch()
    out = (C.c_double * 2)()
    call("dfm_zvec_dot", x.numel(), _dptr(x), _dptr(y), out)
    return complex(out[0], out[1])


def znorm2(x):
    """GetSquaredVectorNorm (zvector_blk.cpp:154-161)."""
    _sync_torch()
    out = C.c_double()
    call("dfm_zvec_norm2", x.numel(), _dptr(x), C.byref(out))
    return out.value


class ZHelmholtzSystem:
    """The complex Helmholtz problem (meshgen.zhelmholtz2d) assembled on the
    device: pattern from the triangles, triangle then Sommerfeld line
    elements, Dirichlet rows / columns and residual."""

    def __init__(self, prob):
        torch = _torch()
        self.prob = prob
        rowptr, col = fem_pattern(prob.nnode, prob.conn)
        self.A = CZMatDia_BlkCrs(prob.nnode, 1)
        self.A.AddPattern(rowptr, col)
        self.asm_tri = ZAssembler(self.A, prob.conn)
        self.asm_line = ZAssembler(self.A, prob.lconn)
        self.coord = to_device(prob.coord.ravel())
        self.value = to_device(_c128(prob.value))
        self.res = torch.zeros(prob.nnode, dtype=torch.complex128, device="cuda")
        self.flags = np.ascontiguousarray(prob.flags, dtype=np.int32)

    def assemble(self):
        """Returns ||r|| after the boundary conditions."""
        self.A.SetZero()
        self.res.zero_()
        p = self.prob
        self.asm_tri.helmholtz2d_p1(self.coord, p.wave_length, self.value, self.res)
        self.asm_line.sommerfeld_b1(self.coord, p.wave_length, self.value, self.res)
        self.A.SetBoundaryCondition(self.flags)
        _sync_torch()
        call("dfm_zvec_zero_bc", self.res.numel(), _ptr(self.flags), _dptr(self.res))
        return znorm2(self.res) ** 0.5


class CMatDiaFlex_BlkCrs:
    """MatVec::CMatDia_BlkCrs made with per-block lengths aLen
    (src/matvec/matdia_blkcrs.cpp:70-104): block i has alen[i] dofs.  Values
    are kept in the reference's layout (dofptr, valptr, diaptr)."""

    def __init__(self, nblk, alen):
        alen = np.ascontiguousarray(alen, dtype=np.uint32)
        assert alen.size == nblk
        h = C.c_void_p()
        call("dfm_fbcrs_create", nblk, _ptr(alen), C.byref(h))
        self.h, self.nblk, self.alen = h, nblk, alen
        self.dofptr = np.zeros(nblk + 1, dtype=np.uint32)
        self.dofptr[1:] = np.cumsum(alen)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_fbcrs_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def NBlkMatCol(self):
        return self.nblk

    def LenBlkCol(self, iblk=None):
        """-1 (flex), or the length of block iblk (mat_blkcrs.h:82-92)."""
        return -1 if iblk is None else int(self.alen[iblk])

    def ndof(self):
        return int(self.dofptr[-1])

    def info(self):
        v = [C.c_uint() for _ in range(5)]
        call("dfm_fbcrs_info", self.h, *(C.byref(x) for x in v))
        return dict(nblk=v[0].value, ndof=v[1].value, ncrs=v[2].value, nval=v[3].value, ndia=v[4].value)

    def NCrs(self):
        return self.info()["ncrs"]

    def AddPattern(self, index, array):
        index = np.ascontiguousarray(index, dtype=np.uint32)
        array = np.ascontiguousarray(array, dtype=np.uint32)
        if array.size == 0:
            array = np.zeros(1, dtype=np.uint32)
        call("dfm_fbcrs_add_pattern", self.h, index.size - 1, _ptr(index), _ptr(array))
        return True

    def layout(self):
        """dict(dofptr, rowptr, col, valptr, diaptr)."""
        i = self.info()
        out = dict(dofptr=np.zeros(self.nblk + 1, dtype=np.uint32), rowptr=np.zeros(self.nblk + 1, dtype=np.uint32),
                   col=np.zeros(max(1, i["ncrs"]), dtype=np.uint32), valptr=np.zeros(i["ncrs"] + 1, dtype=np.uint32),
                   diaptr=np.zeros(self.nblk + 1, dtype=np.uint32))
        call("dfm_fbcrs_layout", self.h, *(_ptr(out[k]) for k in ("dofptr", "rowptr", "col", "valptr", "diaptr")))
        out["col"] = out["col"][:i["ncrs"]]
        return out

    def SetZero(self):
        call("dfm_fbcrs_set_zero", self.h)
        return True

    def values(self):
        i = self.info()
        val, dia = np.zeros(max(1, i["nval"])), np.zeros(max(1, i["ndia"]))
        call("dfm_fbcrs_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:i["nval"]], dia[:i["ndia"]]

    def set_values(self, val, dia):
        val = np.ascontiguousarray(val, dtype=np.float64).reshape(-1)
        dia = np.ascontiguousarray(dia, dtype=np.float64).reshape(-1)
        call("dfm_fbcrs_set_values", self.h, _ptr(val) if val.size else None, _ptr(dia) if dia.size else None)

    def Mearge(self, blocks):
        """Mearge of single blocks [(iblk, jblk, len_i x len_j values)], in
        order (the flex branch takes one block per call, matdia_blkcrs.cpp:173-207)."""
        if not blocks:
            return True
        ib = np.ascontiguousarray([b[0] for b in blocks], dtype=np.uint32)
        jb = np.ascontiguousarray([b[1] for b in blocks], dtype=np.uint32)
        em = np.ascontiguousarray(np.concatenate([np.ravel(b[2]) for b in blocks]), dtype=np.float64)
        call("dfm_fbcrs_mearge", self.h, ib.size, _ptr(ib), _ptr(jb), _ptr(em))
        return True

    def SetBoundaryCondition(self, flags):
        flags = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        assert flags.size == self.ndof()
        call("dfm_fbcrs_set_bc", self.h, _ptr(flags))
        return True

    def MatVec(self, alpha, x, beta, y):
        _sync_torch()
        call("dfm_fbcrs_matvec", self.h, float(alpha), _dptr(x), float(beta), _dptr(y))
        return True


class CMatDiaFracFlex_BlkCrs:
    """MatVec::CMatDiaFrac_BlkCrs(lev, A) over a flex matrix
    (src/matvec/matdiafrac_blkcrs.cpp:227-287, numeric :777-846, substitutions
    :339-379 and :512-553), natural order; also the LsSol CPreconditioner_ILU
    of a single flex segment."""

    def __init__(self, lev, A):
        h = C.c_void_p()
        call("dfm_filu_create", A.h, int(lev), C.byref(h))
        self.h, self.A, self.nblk = h, A, A.nblk
        self.nsingular = 0

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.lib().dfm_filu_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def info(self):
        nc, nv, nl, nu = C.c_uint(), C.c_uint(), C.c_int(), C.c_int()
        call("dfm_filu_info", self.h, C.byref(nc), C.byref(nv), C.byref(nl), C.byref(nu))
        return dict(ncrs=nc.value, nval=nv.value, nlevel_lower=nl.value, nlevel_upper=nu.value)

    def NCrs(self):
        return self.info()["ncrs"]

    def pattern(self):
        i = self.info()
        out = dict(rowptr=np.zeros(self.nblk + 1, dtype=np.uint32), col=np.zeros(max(1, i["ncrs"]), dtype=np.uint32),
                   diaind=np.zeros(max(1, self.nblk), dtype=np.uint32),
                   valptr=np.zeros(i["ncrs"] + 1, dtype=np.uint32))
        call("dfm_filu_get_pattern", self.h, *(_ptr(out[k]) for k in ("rowptr", "col", "diaind", "valptr")))
        out["col"] = out["col"][:i["ncrs"]]
        out["diaind"] = out["diaind"][:self.nblk]
        return out

    def SetValue(self, A):
        ns = C.c_int()
        rc = call("dfm_filu_set_value", self.h, A.h, C.byref(ns), ok=(_lib.DFM_OK, _lib.DFM_SINGULAR))
        self.nsingular = ns.value
        return rc == _lib.DFM_OK

    def values(self):
        nv = self.info()["nval"]
        nd = self.A.info()["ndia"]
        val, dia = np.zeros(max(1, nv)), np.zeros(max(1, nd))
        call("dfm_filu_get_values", self.h, _ptr(val), _ptr(dia))
        return val[:nv], dia[:nd]

    def ForwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_filu_forward", self.h, _dptr(vec))
        return True

    def BackwardSubstitution(self, vec):
        _sync_torch()
        call("dfm_filu_backward", self.h, _dptr(vec))
        return True

    def SolvePrecond(self, vec):
        _sync_torch()
        call("dfm_filu_solve_precond", self.h, _dptr(vec))
        return True


_FX_KIND = {"pcg": 0, "pbicgstab": 1, "cg": 2, "bicgstab": 3}


def solve_flex(kind, conv_ratio, iteration, A, prec, r, x):
    """The LsSol solvers on one flex segment (src/ls/solver_ls_iter.cpp):
    kind "pcg" / "pbicgstab" with prec (a CMatDiaFracFlex_BlkCrs holding
    values), "cg" / "bicgstab" without (prec None).  r in: right-hand side,
    out: final residual; x out: solution.  Returns (ok, conv_ratio,
    iteration) with ok the reference's result (Solve_PCG false at max_iter,
    the others true)."""
    _sync_torch()
    conv = C.c_double(conv_ratio)
    it = C.c_uint(iteration)
    rc = call("dfm_fsolve", _FX_KIND[kind], A.h, prec.h if prec is not None else None, _dptr(r), _dptr(x),
              C.byref(conv), C.byref(it), ok=(_lib.DFM_OK, _lib.DFM_NOT_CONVERGED))
    ok = rc == _lib.DFM_OK or kind != "pcg"
    return ok, conv.value, it.value
