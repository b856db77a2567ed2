"""ctypes binding of libdelfem_amd.so (the C-ABI declared in include/dfm.h).

The shared library is built in-tree by delfem_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdelfem_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dfm.h")

DFM_OK = 0
DFM_NOT_CONVERGED = 1
DFM_SINGULAR = 2
DFM_BAD_ARG = -1
DFM_HIP_ERROR = -2
DFM_UNSORTED = -3
DFM_NOT_IMPLEMENTED = -4
DFM_TOO_LARGE = -5

DFM_ORDER_NATURAL = 0
DFM_ORDER_GIVEN = 1
DFM_ORDER_MULTICOLOUR = 2
DFM_ORDER_BLOCK_MULTICOLOUR = 3

DFM_PREC_FP64 = 0
DFM_PREC_MIXED = 1
DFM_PREC_MIXED_BF16 = 2
DFM_PREC_MIXED_F32 = 3
DFM_PREC_MIXED_F32_BF16 = 4
DFM_PREC_FAC32 = 5
DFM_PREC_FAC16 = 6

PRECISIONS = {"fp64": DFM_PREC_FP64, "mixed": DFM_PREC_MIXED, "mixed_bf16": DFM_PREC_MIXED_BF16,
              "mixed_f32": DFM_PREC_MIXED_F32, "mixed_f32_bf16": DFM_PREC_MIXED_F32_BF16,
              "fac32": DFM_PREC_FAC32, "fac16": DFM_PREC_FAC16}


class SolveOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("check_every", C.c_uint),
        ("want_true_relres", C.c_int),
        ("reliable_delta", C.c_double),
        ("inner_iters", C.c_uint),
        ("outer_iters", C.c_uint),
        ("true_relres", C.c_double),
    ]


u32p = C.POINTER(C.c_uint)
i32p = C.POINTER(C.c_int)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p
H = C.c_void_p  # opaque handle

_SIGS = {
    "dfm_last_error": (C.c_char_p, []),
    "dfm_init": (C.c_int, [C.c_int]),
    "dfm_synchronize": (C.c_int, []),
    "dfm_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "dfm_free": (C.c_int, [vp]),
    "dfm_memcpy_htod": (C.c_int, [vp, vp, C.c_size_t]),
    "dfm_memcpy_dtoh": (C.c_int, [vp, vp, C.c_size_t]),
    "dfm_memcpy_dtod": (C.c_int, [vp, vp, C.c_size_t]),
    "dfm_stream": (vp, []),
    "dfm_set_stream": (C.c_int, [vp]),
    "dfm_fem_pattern": (C.c_int, [C.c_uint, C.c_uint, C.c_uint, vp, vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "dfm_bcrs_create": (C.c_int, [C.c_uint, C.c_uint, C.POINTER(H)]),
    "dfm_bcrs_destroy": (C.c_int, [H]),
    "dfm_bcrs_add_pattern": (C.c_int, [H, C.c_uint, vp, vp]),
    "dfm_bcrs_info": (C.c_int, [H, u32p, u32p, u32p]),
    "dfm_bcrs_get_pattern": (C.c_int, [H, vp, vp]),
    "dfm_bcrs_set_zero": (C.c_int, [H]),
    "dfm_bcrs_set_values": (C.c_int, [H, vp, vp]),
    "dfm_bcrs_get_values": (C.c_int, [H, vp, vp]),
    "dfm_bcrs_device_values": (C.c_int, [H, C.POINTER(vp), C.POINTER(vp)]),
    "dfm_bcrs_mearge": (C.c_int, [H, C.c_uint, vp, C.c_uint, vp]),
    "dfm_bcrs_set_bc": (C.c_int, [H, vp]),
    "dfm_bcrs_matvec": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_vec_dot": (C.c_int, [C.c_size_t, vp, vp, f64p]),
    "dfm_vec_axpy": (C.c_int, [C.c_size_t, C.c_double, vp, vp]),
    "dfm_vec_scal": (C.c_int, [C.c_size_t, C.c_double, vp]),
    "dfm_vec_copy": (C.c_int, [C.c_size_t, vp, vp]),
    "dfm_vec_fill": (C.c_int, [C.c_size_t, C.c_double, vp]),
    "dfm_vec_zero_bc": (C.c_int, [C.c_size_t, vp, vp]),
    "dfm_asm_create": (C.c_int, [H, C.c_uint, C.c_uint, C.c_uint, vp, C.POINTER(H)]),
    "dfm_asm_destroy": (C.c_int, [H]),
    "dfm_assemble_poisson2d_p1": (C.c_int, [H, vp, vp, C.c_double, vp, vp]),
    "dfm_assemble_solid3d_p1": (C.c_int, [H, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double, vp, vp]),
    "dfm_batch_create": (C.c_int, [H, C.c_uint, vp, C.POINTER(H)]),
    "dfm_batch_destroy": (C.c_int, [H]),
    "dfm_batch_info": (C.c_int, [H, vp, vp, vp, vp]),
    "dfm_batch_set_value": (C.c_int, [H, H]),
    "dfm_batch_set_launch": (C.c_int, [H, C.c_int, C.c_int]),
    "dfm_batch_set_teams": (C.c_int, [H, C.c_uint, C.c_int]),
    "dfm_batch_solve_pcg": (C.c_int, [H, vp, vp, C.c_double, C.c_uint, C.c_int, vp, vp, vp, vp, vp]),
    "dfm_batch_get_factor": (C.c_int, [H, C.c_uint, vp, vp, vp]),
    "dfm_assemble_oseen2d_p1p1": (C.c_int, [H, vp, C.c_double, C.c_double, vp, vp, vp]),
    "dfm_assemble_helmholtz2d_p1": (C.c_int, [H, vp, vp, vp, C.c_double, vp, vp]),
    "dfm_ilu_create": (C.c_int, [H, C.c_int, C.c_int, vp, C.c_uint, C.POINTER(H)]),
    "dfm_ilu_create_fill": (C.c_int, [H, C.c_int, vp, C.c_uint, C.POINTER(H)]),
    "dfm_ilu_destroy": (C.c_int, [H]),
    "dfm_ilu_info": (C.c_int, [H, u32p, i32p, i32p, i32p]),
    "dfm_ilu_tile_info": (C.c_int, [H, i32p, i32p, i32p, i32p]),
    "dfm_ilu_get_pattern": (C.c_int, [H, vp, vp, vp, vp, vp]),
    "dfm_ilu_set_value": (C.c_int, [H, H, i32p]),
    "dfm_ilu_get_values": (C.c_int, [H, vp, vp]),
    "dfm_ilu_forward": (C.c_int, [H, vp]),
    "dfm_ilu_backward": (C.c_int, [H, vp]),
    "dfm_ilu_solve_precond_noalloc": (C.c_int, [H, vp]),
    "dfm_ilu_solve_precond": (C.c_int, [H, vp]),
    "dfm_set_sweep_mode": (C.c_int, [C.c_int, C.c_int]),
    "dfm_set_kernel_timing": (C.c_int, [C.c_int]),
    "dfm_get_kernel_timing": (C.c_int, [f64p, f64p, i32p, i32p]),
    "dfm_solve_pcg": (C.c_int, [H, H, vp, vp, f64p, u32p, C.POINTER(SolveOpts)]),
    "dfm_solve_pbicgstab": (C.c_int, [H, H, vp, vp, f64p, u32p, C.POINTER(SolveOpts)]),
    "dfm_bmat_create": (C.c_int, [C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.POINTER(H)]),
    "dfm_bmat_destroy": (C.c_int, [H]),
    "dfm_bmat_set_pattern_boundary": (C.c_int, [H, H, vp, vp]),
    "dfm_bmat_info": (C.c_int, [H, u32p, u32p, u32p]),
    "dfm_bmat_get_pattern": (C.c_int, [H, vp, vp]),
    "dfm_bmat_set_zero": (C.c_int, [H]),
    "dfm_bmat_get_values": (C.c_int, [H, vp]),
    "dfm_bmat_set_values": (C.c_int, [H, vp]),
    "dfm_bmat_mearge": (C.c_int, [H, C.c_uint, vp, C.c_uint, vp]),
    "dfm_bmat_copy_from": (C.c_int, [H, H]),
    "dfm_bmat_set_bc_inverse_col": (C.c_int, [H, vp]),
    "dfm_bmat_set_bc_row": (C.c_int, [H, vp]),
    "dfm_bmat_matvec": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_vec_copy_free": (C.c_int, [C.c_size_t, vp, vp, vp]),
    "dfm_bmat_add_pattern": (C.c_int, [H, C.c_uint, vp, vp]),
    "dfm_bmat_set_bc_col": (C.c_int, [H, vp]),
    "dfm_bmat_extract": (C.c_int, [H, H, vp, vp]),
    "dfm_bcrs_extract": (C.c_int, [H, H, vp]),
    "dfm_vec_gather": (C.c_int, [C.c_size_t, vp, vp, vp]),
    "dfm_mseg_create": (C.c_int, [C.c_uint, vp, vp, C.POINTER(H)]),
    "dfm_mseg_destroy": (C.c_int, [H]),
    "dfm_mseg_set_dia": (C.c_int, [H, C.c_uint, H]),
    "dfm_mseg_set_nondia": (C.c_int, [H, C.c_uint, C.c_uint, H]),
    "dfm_mseg_matvec": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_mseg_dot": (C.c_int, [H, vp, vp, f64p]),
    "dfm_mseg_ilu_create": (C.c_int, [H, C.c_int, C.POINTER(H)]),
    "dfm_mseg_ilu_create_levels": (C.c_int, [H, vp, vp, C.c_uint, C.POINTER(H)]),
    "dfm_mseg_ilu_get_pattern": (C.c_int, [H, C.c_uint, C.c_uint, u32p, vp, vp]),
    "dfm_mseg_ilu_destroy": (C.c_int, [H]),
    "dfm_mseg_ilu_set_value": (C.c_int, [H, i32p]),
    "dfm_mseg_ilu_get_values": (C.c_int, [H, C.c_uint, C.c_uint, vp, vp]),
    "dfm_mseg_ilu_precond": (C.c_int, [H, vp]),
    "dfm_mseg_solve_pcg": (C.c_int, [H, H, vp, vp, f64p, u32p]),
    "dfm_mseg_solve_pbicgstab": (C.c_int, [H, H, vp, vp, f64p, u32p]),
    "dfm_mseg_set_device_loop": (C.c_int, [C.c_int]),
    "dfm_zbcrs_create": (C.c_int, [C.c_uint, C.c_uint, C.POINTER(H)]),
    "dfm_zbcrs_destroy": (C.c_int, [H]),
    "dfm_zbcrs_add_pattern": (C.c_int, [H, C.c_uint, vp, vp]),
    "dfm_zbcrs_info": (C.c_int, [H, u32p, u32p, u32p]),
    "dfm_zbcrs_get_pattern": (C.c_int, [H, vp, vp]),
    "dfm_zbcrs_set_zero": (C.c_int, [H]),
    "dfm_zbcrs_set_values": (C.c_int, [H, vp, vp]),
    "dfm_zbcrs_get_values": (C.c_int, [H, vp, vp]),
    "dfm_zbcrs_device_values": (C.c_int, [H, C.POINTER(vp), C.POINTER(vp)]),
    "dfm_zbcrs_mearge": (C.c_int, [H, C.c_uint, vp, C.c_uint, vp]),
    "dfm_zbcrs_set_bc": (C.c_int, [H, vp]),
    "dfm_zbcrs_matvec": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_zbcrs_matvec_hermitian": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_zvec_dot": (C.c_int, [C.c_size_t, vp, vp, f64p]),
    "dfm_zvec_norm2": (C.c_int, [C.c_size_t, vp, f64p]),
    "dfm_zvec_zero_bc": (C.c_int, [C.c_size_t, vp, vp]),
    "dfm_zasm_create": (C.c_int, [H, C.c_uint, C.c_uint, C.c_uint, vp, C.POINTER(H)]),
    "dfm_zassemble_helmholtz2d_p1": (C.c_int, [H, H, vp, C.c_double, vp, vp]),
    "dfm_zassemble_sommerfeld_b1": (C.c_int, [H, H, vp, C.c_double, vp, vp]),
    "dfm_zilu_create": (C.c_int, [H, C.c_int, C.c_int, vp, C.POINTER(H)]),
    "dfm_zilu_destroy": (C.c_int, [H]),
    "dfm_zilu_info": (C.c_int, [H, u32p, i32p, i32p, i32p]),
    "dfm_zilu_get_pattern": (C.c_int, [H, vp, vp, vp, vp, vp]),
    "dfm_zilu_set_value": (C.c_int, [H, H, i32p]),
    "dfm_zilu_get_values": (C.c_int, [H, vp, vp]),
    "dfm_zilu_forward": (C.c_int, [H, vp]),
    "dfm_zilu_backward": (C.c_int, [H, vp]),
    "dfm_zilu_solve_precond": (C.c_int, [H, vp]),
    "dfm_zsolve_pcocg": (C.c_int, [H, H, vp, vp, f64p, u32p]),
    "dfm_zsolve_cocg": (C.c_int, [H, vp, vp, f64p, u32p]),
    "dfm_zsolve_cgnr": (C.c_int, [H, vp, vp, f64p, u32p]),
    "dfm_zsolve_cg": (C.c_int, [H, vp, vp, f64p, u32p]),
    "dfm_zsolve_pcg": (C.c_int, [H, H, vp, vp, f64p, u32p]),
    "dfm_zvec_inner": (C.c_int, [C.c_size_t, vp, vp, vp]),
    "dfm_zsolve_bicgstab": (C.c_int, [H, vp, vp, f64p, u32p]),
    "dfm_zsolve_bicgstabp": (C.c_int, [H, H, vp, vp, f64p, u32p]),
    "dfm_fbcrs_create": (C.c_int, [C.c_uint, vp, C.POINTER(H)]),
    "dfm_fbcrs_destroy": (C.c_int, [H]),
    "dfm_fbcrs_add_pattern": (C.c_int, [H, C.c_uint, vp, vp]),
    "dfm_fbcrs_info": (C.c_int, [H, u32p, u32p, u32p, u32p, u32p]),
    "dfm_fbcrs_layout": (C.c_int, [H, vp, vp, vp, vp, vp]),
    "dfm_fbcrs_set_zero": (C.c_int, [H]),
    "dfm_fbcrs_set_values": (C.c_int, [H, vp, vp]),
    "dfm_fbcrs_get_values": (C.c_int, [H, vp, vp]),
    "dfm_fbcrs_device_values": (C.c_int, [H, C.POINTER(vp), C.POINTER(vp)]),
    "dfm_fbcrs_mearge": (C.c_int, [H, C.c_uint, vp, vp, vp]),
    "dfm_fbcrs_set_bc": (C.c_int, [H, vp]),
    "dfm_fbcrs_matvec": (C.c_int, [H, C.c_double, vp, C.c_double, vp]),
    "dfm_filu_create": (C.c_int, [H, C.c_int, C.POINTER(H)]),
    "dfm_filu_destroy": (C.c_int, [H]),
    "dfm_filu_info": (C.c_int, [H, u32p, u32p, i32p, i32p]),
    "dfm_filu_get_pattern": (C.c_int, [H, vp, vp, vp, vp]),
    "dfm_filu_set_value": (C.c_int, [H, H, i32p]),
    "dfm_filu_get_values": (C.c_int, [H, vp, vp]),
    "dfm_filu_forward": (C.c_int, [H, vp]),
    "dfm_filu_backward": (C.c_int, [H, vp]),
    "dfm_filu_solve_precond": (C.c_int, [H, vp]),
    "dfm_fsolve": (C.c_int, [C.c_int, H, H, vp, vp, f64p, u32p]),
    "dfm_bench_kernel": (C.c_int, [H, H, C.c_int, C.c_int, C.c_int, f64p, f64p]),
    "dfm_bench_copy": (C.c_int, [C.c_size_t, C.c_int, f64p]),
    "dfm_bench_read": (C.c_int, [C.c_size_t, C.c_int, f64p]),
}

_lib = None


class DfmError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} returned {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m delfem_amd.build` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args, ok=(DFM_OK,)):
    """Call a dfm_* function and raise DfmError on a status outside `ok`."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc not in ok:
        raise DfmError(name, rc, L.dfm_last_error().decode(errors="replace"))
    return rc


def declared_symbols():
    """Function names declared in include/dfm.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dfm_[a-z0-9_]+)\s*\(", txt)))
