"""The persistent triangular sweep (one launch for all level sets of a sweep,
slices taken in level order, per-level completion counts) against the
one-launch-per-level sweep.  The slots of a row are subtracted in the same
order either way, and BiCGSTAB's reductions do not involve the sweeps, so the
whole FP64 and FAC32 BiCGSTAB runs must be the same bits in both modes and for
any grid -- one wave (which then takes every slice itself, level after level),
an odd handful, or the default.  PCG's first preconditioner application reduces
<r,z> over the persistent grid's partials, so PCG is held to one iteration and
1e-8 of the per-level run."""
import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import assemble_problem

pytestmark = pytest.mark.gpu

CASES = {
    "oseen2d_16_ilu1_bmc": (lambda: meshgen.oseen2d(16, seed=3), 1, "block_multicolour", "bicg"),
    "solid3d_4_ilu1_natural": (lambda: meshgen.solid3d(4, seed=3, sigma=0.5), 1, "natural", "pcg"),
    "poisson2d_20_ilu2_natural": (lambda: meshgen.poisson2d(20, seed=3), 2, "natural", "pcg"),
}


def _case(name):
    print(f"[case] {name}", flush=True)


@pytest.fixture(params=sorted(CASES))
def system(request, orc, gpu):
    _case(request.param)
    make, lev, ordering, _ = CASES[request.param]
    prob = make()
    s = gpu.FEMSystem(prob)
    s.assemble()
    M = gpu.CPreconditioner_ILU(s.A, lev, ordering=ordering, block_nodes=8)
    assert M.SetValue(s.A)
    assert M.info()["nlevel_lower"] >= 8, M.info()
    yield request.param, prob, s, M
    gpu.set_sweep_mode(1, 0)


def _solve(gpu, s, M, kind, precision, tol=1e-10):
    import torch
    r = s.res.clone()
    x = torch.zeros_like(r)
    fn = gpu.Solve_PCG if kind == "pcg" else gpu.Solve_PBiCGSTAB
    ok, conv, it, info = fn(tol, 5000, s.A, M, r, x, precision=precision)
    return info["converged"], conv, it, x.cpu().numpy(), r.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "fac32"])
def test_bicgstab_bitwise_across_modes(system, gpu, precision):
    name, prob, s, M = system
    gpu.set_sweep_mode(0, 0)
    base = _solve(gpu, s, M, "bicg", precision)
    assert base[0]
    for grid in (0, 1, 7):
        _case(f"{name} bicg {precision} persistent grid {grid}")
        gpu.set_sweep_mode(2, grid)
        got = _solve(gpu, s, M, "bicg", precision)
        assert got[:3] == base[:3], (grid, got[:3], base[:3])
        assert np.array_equal(got[3], base[3]) and np.array_equal(got[4], base[4]), grid


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_across_modes(system, gpu, precision):
    name, prob, s, M = system
    if CASES[name][3] != "pcg":
        pytest.skip("nonsymmetric system")
    gpu.set_sweep_mode(0, 0)
    ok0, _, it0, x0, _ = _solve(gpu, s, M, "pcg", precision)
    for grid in (0, 1):
        _case(f"{name} pcg {precision} persistent grid {grid}")
        gpu.set_sweep_mode(2, grid)
        ok, conv, it, x, _ = _solve(gpu, s, M, "pcg", precision)
        assert ok and ok0 and conv < 1e-10 and abs(it - it0) <= 1
        assert np.linalg.norm(x - x0) <= 1e-8 * np.linalg.norm(x0)


def test_kernel_timer_runs_persistent(system, gpu):
    import ctypes as C
    from delfem_amd import _lib
    name, prob, s, M = system
    gpu.set_sweep_mode(2, 0)
    for which in (1, 2):
        ms, nbytes = C.c_double(), C.c_double()
        _lib.call("dfm_bench_kernel", s.A.h, M.h, which, 0, 5, C.byref(ms), C.byref(nbytes))
        assert ms.value > 0 and nbytes.value > 0
