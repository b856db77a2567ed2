"""The phase-tiled triangular sweep (one launch per colour of a multicolour
or block-multicolour ordering; tiles of dependent rows run level by level
inside a workgroup) against the one-launch-per-level sweep, in both forms:
sweep mode 3 (one lane per tile row) and mode 4 (2x2 and 3x3 blocks: one
lane per block row, the D^-1 product through quad permutes).  A tile row
subtracts its slots in the factor's column order, products formed as in the
per-level kernels, so BiCGSTAB (whose reductions do not involve the sweeps)
must give the same bits in every mode for every storage type of the factors.
PCG's first preconditioner application reduces <r,z> over the tiles'
partials, so PCG is held to one iteration and 1e-8 of the per-level run.
Block sizes of the ordering select tiles of 64, 128 and 256 threads; a
factor whose components exceed 256 rows has no tiling and keeps the other
sweep forms."""
import numpy as np
import pytest

from delfem_amd import meshgen

pytestmark = pytest.mark.gpu

# name: (problem, ILU level, ordering, block nodes, solver)
CASES = {
    "oseen2d_16_ilu1_bmc8": (lambda: meshgen.oseen2d(16, seed=3), 1, "block_multicolour", 8, "bicg"),
    "oseen2d_24_ilu1_bmc48": (lambda: meshgen.oseen2d(24, seed=4), 1, "block_multicolour", 48, "bicg"),
    "oseen2d_32_ilu2_bmc40": (lambda: meshgen.oseen2d(32, seed=5), 2, "block_multicolour", 40, "bicg"),
    # a dependency component above 256 rows: no tiling, the modes fall back
    "oseen2d_32_ilu1_bmc100_untiled": (lambda: meshgen.oseen2d(32, seed=5), 1, "block_multicolour", 100, "bicg"),
    "solid3d_4_ilu1_mc": (lambda: meshgen.solid3d(4, seed=3, sigma=0.5), 1, "multicolour", 0, "pcg"),
    "poisson2d_24_ilu2_bmc16": (lambda: meshgen.poisson2d(24, seed=3), 2, "block_multicolour", 16, "pcg"),
    "poisson2d_24_ilu0_mc": (lambda: meshgen.poisson2d(24, seed=6), 0, "multicolour", 0, "pcg"),
}


def _case(name):
    print(f"[case] {name}", flush=True)


@pytest.fixture(params=sorted(CASES), scope="module")
def system(request, gpu):
    _case(request.param)
    make, lev, ordering, bn, _ = CASES[request.param]
    prob = make()
    s = gpu.FEMSystem(prob)
    s.assemble()
    M = gpu.CPreconditioner_ILU(s.A, lev, ordering=ordering, block_nodes=bn)
    assert M.SetValue(s.A)
    yield request.param, prob, s, M
    gpu.set_sweep_mode(1, 0)


def _solve(gpu, s, M, kind, precision, tol=1e-10):
    import torch
    r = s.res.clone()
    x = torch.zeros_like(r)
    fn = gpu.Solve_PCG if kind == "pcg" else gpu.Solve_PBiCGSTAB
    ok, conv, it, info = fn(tol, 5000, s.A, M, r, x, precision=precision)
    return info["converged"], conv, it, x.cpu().numpy(), r.cpu().numpy()


def test_tiling_present(system):
    name, prob, s, M = system
    ti, info = M.tile_info(), M.info()
    print(name, ti, info, flush=True)
    if name.endswith("untiled"):
        assert ti == dict(nphase=0, ntile_lower=0, ntile_upper=0, tile_threads=0)
        return
    assert ti["nphase"] == info["ncolour"] and ti["ntile_lower"] > 0 and ti["ntile_upper"] > 0
    assert ti["tile_threads"] in (64, 128, 256)
    if name.endswith("bmc40"):
        assert ti["tile_threads"] == 256


@pytest.mark.parametrize("precision", ["fp64", "fac32", "fac16", "mixed", "mixed_f32"])
def test_bicgstab_bitwise_tiled(system, gpu, precision):
    name, prob, s, M = system
    if CASES[name][4] != "bicg":
        pytest.skip("symmetric system: PCG")
    gpu.set_sweep_mode(0, 0)
    base = _solve(gpu, s, M, "bicg", precision)
    for mode in (3, 4, 1):
        _case(f"{name} bicg {precision} sweep mode {mode}")
        gpu.set_sweep_mode(mode, 0)
        got = _solve(gpu, s, M, "bicg", precision)
        assert got[:3] == base[:3], (mode, got[:3], base[:3])
        assert np.array_equal(got[3], base[3]) and np.array_equal(got[4], base[4]), mode
    assert base[0] or precision != "fp64"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_tiled(system, gpu, precision):
    name, prob, s, M = system
    if CASES[name][4] != "pcg":
        pytest.skip("nonsymmetric system: BiCGSTAB")
    gpu.set_sweep_mode(0, 0)
    ok0, _, it0, x0, _ = _solve(gpu, s, M, "pcg", precision)
    for mode in (3, 4):
        _case(f"{name} pcg {precision} tiled mode {mode}")
        gpu.set_sweep_mode(mode, 0)
        ok, conv, it, x, _ = _solve(gpu, s, M, "pcg", precision)
        assert ok and ok0 and conv < 1e-10 and abs(it - it0) <= 1
        assert np.linalg.norm(x - x0) <= 1e-8 * np.linalg.norm(x0)


def test_kernel_timer_runs_tiled(system, gpu):
    import ctypes as C
    from delfem_amd import _lib
    name, prob, s, M = system
    for mode in (3, 4):
        gpu.set_sweep_mode(mode, 0)
        for which in (1, 2):
            ms, nbytes = C.c_double(), C.c_double()
            _lib.call("dfm_bench_kernel", s.A.h, M.h, which, 0, 5, C.byref(ms), C.byref(nbytes))
            assert ms.value > 0 and nbytes.value > 0


# Natural-order factors whose phases are found in the factor's own numbering:
# a mesh numbered by colours of geometric tiles, and a small lexicographic one
# (whose phases are 256-row tiles, one workgroup each).
FOUND = {
    "oseen2d_32_tiles_ilu1": (lambda: meshgen.oseen2d(32, seed=3), True, 1),
    "poisson2d_40_tiles_ilu2": (lambda: meshgen.poisson2d(40, seed=2), True, 2),
    "solid3d_6_tiles_ilu0": (lambda: meshgen.solid3d(6, seed=2, sigma=0.5), True, 0),
    "poisson2d_20_lex_ilu2": (lambda: meshgen.poisson2d(20, seed=1), False, 2),
}


@pytest.mark.parametrize("numbering", ["mesh", "given"])
@pytest.mark.parametrize("name", sorted(FOUND))
def test_found_phases(orc, gpu, name, numbering):
    """The found phases' tile sweeps equal the oracle's substitutions bit for
    bit; BiCGSTAB gives the same bits with tiles and with one launch per
    level; PCG within one iteration and 1e-8 of the oracle.  numbering
    "mesh": the mesh is relabelled and the factor is in natural order;
    "given": the mesh keeps its numbering and the tile numbering is the
    factor's given order (phases found in factor numbering)."""
    import torch
    from oracle.oracle import assemble_problem
    _case(f"found phases {name} {numbering}")
    make, tiles, lev = FOUND[name]
    prob = make()
    order = None
    if tiles and numbering == "mesh":
        prob = meshgen.relabel(prob, meshgen.tile_colour_order(prob, 16))
    elif tiles:
        order = np.ascontiguousarray(meshgen.tile_colour_order(prob, 16), dtype=np.int32)
    elif numbering == "given":
        pytest.skip("a lexicographic mesh has no second numbering here")
    ref = assemble_problem(orc, prob)
    s = gpu.FEMSystem(prob)
    s.assemble()
    M = gpu.CPreconditioner_ILU(s.A, lev, order=order)
    ti, info = M.tile_info(), M.info()
    print(name, ti, info, flush=True)
    assert ti["nphase"] >= 1 and ti["ntile_lower"] > 0 and ti["ntile_upper"] > 0
    assert ti["nphase"] <= max(info["nlevel_lower"], info["nlevel_upper"]) // 4
    assert M.SetValue(s.A)
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], lev, order)
    ov, od, ns = orc.ilu_values(prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, order)
    fval, fdia = M.values()
    assert ns == 0 and np.array_equal(fval, ov) and np.array_equal(fdia, od)
    v = np.random.default_rng(5).standard_normal(prob.ndof)
    for mode in (1, 0):
        gpu.set_sweep_mode(mode, 0)
        vd = torch.from_numpy(v.copy()).cuda()
        M.ForwardSubstitution(vd)
        w = orc.forward(prob.nnode, prob.b, F, ov, od, v)
        assert np.array_equal(vd.cpu().numpy(), w), mode
        M.BackwardSubstitution(vd)
        w = orc.backward(prob.nnode, prob.b, F, ov, od, w)
        assert np.array_equal(vd.cpu().numpy(), w), mode
    gpu.set_sweep_mode(0, 0)
    base = _solve(gpu, s, M, "bicg", "fp64")
    gpu.set_sweep_mode(1, 0)
    got = _solve(gpu, s, M, "bicg", "fp64")
    assert got[:3] == base[:3] and np.array_equal(got[3], base[3]) and np.array_equal(got[4], base[4])
    if prob.kind != "oseen2d":
        ok_o, conv_o, it_o, x_o, _ = orc.solve("pcg", prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"],
                                               ref["dia"], F, ov, od, ref["res"], 1e-10, 5000, order)
        ok, conv, it, x, _ = _solve(gpu, s, M, "pcg", "fp64")
        print(f"  pcg {it} iterations, oracle {it_o}", flush=True)
        assert ok and ok_o and abs(it - it_o) <= 1
        assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)
