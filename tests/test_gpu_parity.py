"""Parity of the HIP path (through the C-ABI) with the oracle on the same inputs.

Integer work (patterns, orderings) must be identical.  FP64 kernels that the
device runs in the reference's operation order (assembly, Dirichlet rows,
MatVec, ILU numeric, substitutions) must match bitwise.  Krylov solves use
parallel reductions for the dot products, so iterates differ by rounding:
PCG iteration counts within 1 and solutions within 1e-8 relative error at
the 1e-10 residual tolerance; BiCGSTAB, whose iterates are sensitive to that
rounding, within 25 % of the iterations and 1e-6 relative error.
"""
import hashlib
import os

import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import assemble_problem

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cylinder_mesh():
    """The reference's unstructured tetrahedral mesh (its solid3d demo input)
    with the coefficients and Dirichlet flags stored in the fixture."""
    d = np.load(os.path.join(GOLDEN, "cylinder_tet.npz"), allow_pickle=False)
    nnode, nelem = int(d["nnode"]), d["conn"].shape[0]
    return meshgen.Problem(kind="solid3d", name="cylinder_tet", n=0, b=3, npe=4, coord=d["coord"],
                           conn=d["conn"], nnode=nnode, nelem=nelem, lam=d["lam"], mu=d["mu"],
                           rho=float(d["rho"]), g=tuple(d["g"]), flags=d["flags"], seed=0, sigma=0.0)


PROBLEMS = {
    "poisson2d_12": lambda: meshgen.poisson2d(12, seed=7),
    "solid3d_4": lambda: meshgen.solid3d(4, seed=7, sigma=0.5),
    "oseen2d_10": lambda: meshgen.oseen2d(10, seed=7),
    "helmholtz2d_9x5": lambda: meshgen.helmholtz2d(9, 5, 2.0, 1.0, seed=7, c=4.0),
    "cylinder_tet": cylinder_mesh,
}
ORDERINGS = ["natural", "multicolour", "block_multicolour"]
BLOCK_NODES = 8


def make_ilu(gpu, lev, A, ordering):
    return gpu.CMatDiaFrac_BlkCrs(lev, A, ordering=ordering, block_nodes=BLOCK_NODES)


def oracle_order(orc, ref, ordering):
    if ordering == "multicolour":
        return orc.colour_order(ref["rowptr"], ref["col"])[1:]
    if ordering == "block_multicolour":
        return orc.bmc_order(ref["rowptr"], ref["col"], BLOCK_NODES)[1:]
    return 0, None


def skip_unsuited(prob, kind):
    if kind == "pcg" and prob.kind == "oseen2d":
        pytest.skip("PCG needs a symmetric positive definite matrix; the Oseen system is not")


def _case(name):
    print(f"[case] {name}", flush=True)


def true_relres(s, x, rhs):
    """||rhs - A x|| / ||rhs|| with the API MatVec on the canonical CRS matrix."""
    import torch
    y = torch.from_numpy(rhs.copy()).cuda()
    s.A.MatVec(-1.0, x, 1.0, y)
    return float(torch.linalg.norm(y).item() / np.linalg.norm(rhs))


@pytest.fixture(scope="module", params=sorted(PROBLEMS))
def setup(request, orc, gpu):
    import torch
    _case(request.param)
    prob = PROBLEMS[request.param]()
    ref = assemble_problem(orc, prob)
    sys_ = gpu.FEMSystem(prob)
    nres = sys_.assemble()
    return dict(prob=prob, ref=ref, sys=sys_, nres=nres, torch=torch, gpu=gpu)


def test_pattern(setup):
    prob, ref, s = setup["prob"], setup["ref"], setup["sys"]
    rowptr, col = s.A.pattern()
    assert np.array_equal(rowptr, ref["rowptr"])
    assert np.array_equal(col, ref["col"])


def test_assembly_bitwise(setup):
    ref, s = setup["ref"], setup["sys"]
    val, dia = s.A.values()
    assert np.array_equal(val, ref["val"])
    assert np.array_equal(dia, ref["dia"])
    res = s.res.cpu().numpy()
    assert np.array_equal(res, ref["res"])
    assert setup["nres"] == pytest.approx(np.linalg.norm(ref["res"]), rel=1e-12)


def test_matvec_bitwise(setup, orc):
    prob, ref, s, torch = setup["prob"], setup["ref"], setup["sys"], setup["torch"]
    rng = np.random.default_rng(3)
    x = rng.standard_normal(prob.ndof)
    y0 = rng.standard_normal(prob.ndof)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        xd = torch.from_numpy(x).cuda()
        yd = torch.from_numpy(y0.copy()).cuda()
        s.A.MatVec(alpha, xd, beta, yd)
        want = orc.matvec(prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], alpha, x, beta, y0)
        assert np.array_equal(yd.cpu().numpy(), want)


@pytest.mark.parametrize("lev", [0, 1, 2])
@pytest.mark.parametrize("ordering", ORDERINGS)
def test_ilu_bitwise(setup, orc, lev, ordering):
    prob, ref, s, gpu, torch = setup["prob"], setup["ref"], setup["sys"], setup["gpu"], setup["torch"]
    _case(f"ilu lev={lev} {ordering}")
    M = make_ilu(gpu, lev, s.A, ordering)
    P = M.pattern()
    nc, order = oracle_order(orc, ref, ordering)
    if order is not None:
        assert np.array_equal(P["order"], order)
        assert M.info()["ncolour"] == nc
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], lev, order)
    for k in ("rowptr", "col", "lev", "diaind"):
        assert np.array_equal(P[k], F[k]), k
    levL, nL = orc.level_sets(F, True)
    levU, nU = orc.level_sets(F, False)
    info = M.info()
    assert info["nlevel_lower"] == nL and info["nlevel_upper"] == nU
    assert M.SetValue(s.A)
    fval, fdia = M.values()
    ov, od, ns = orc.ilu_values(prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, order)
    assert ns == 0
    assert np.array_equal(fval, ov)
    assert np.array_equal(fdia, od)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(prob.ndof)
    vd = torch.from_numpy(v).cuda()
    M.ForwardSubstitution(vd)
    w = orc.forward(prob.nnode, prob.b, F, ov, od, v)
    assert np.array_equal(vd.cpu().numpy(), w)
    M.BackwardSubstitution(vd)
    w = orc.backward(prob.nnode, prob.b, F, ov, od, w)
    assert np.array_equal(vd.cpu().numpy(), w)


@pytest.mark.parametrize("kind", ["pcg", "bicg"])
@pytest.mark.parametrize("ordering", ORDERINGS)
def test_krylov_fp64(setup, orc, kind, ordering):
    prob, ref, s, gpu, torch = setup["prob"], setup["ref"], setup["sys"], setup["gpu"], setup["torch"]
    skip_unsuited(prob, kind)
    _case(f"{kind} {ordering}")
    lev = 1 if prob.kind == "oseen2d" else 0
    M = make_ilu(gpu, lev, s.A, ordering)
    assert M.SetValue(s.A)
    order = M.pattern()["order"] if ordering != "natural" else None
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], lev, order)
    ov, od, _ = orc.ilu_values(prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, order)
    ok_o, conv_o, it_o, x_o, r_o = orc.solve(kind, prob.nnode, prob.b, ref["rowptr"], ref["col"], ref["val"],
                                             ref["dia"], F, ov, od, ref["res"], 1e-10, 1000, order)
    r = torch.from_numpy(ref["res"].copy()).cuda()
    x = torch.zeros_like(r)
    fn = gpu.Solve_PCG if kind == "pcg" else gpu.Solve_PBiCGSTAB
    ok, conv, it, info = fn(1e-10, 1000, s.A, M, r, x, want_true_relres=1, check_every=4)
    assert ok and ok_o
    # PCG's iterates are stable under the rounding of the parallel dot
    # products: same count within one, same solution to 1e-8.  BiCGSTAB's are
    # not (its residual is erratic, the count moves by 10 % and more): the bar
    # there is the solution's, 1e-6 relative error at a 1e-10 residual.
    pcg = kind == "pcg"
    assert abs(it - it_o) <= (1 if pcg else max(2, it_o // 4)), (it, it_o)
    assert conv < 1e-10
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - x_o) <= (1e-8 if pcg else 1e-6) * np.linalg.norm(x_o)
    assert info["true_relres"] < 1e-9
    assert true_relres(s, x, ref["res"]) < 1e-9


@pytest.mark.parametrize("precision", ["mixed", "mixed_bf16", "mixed_f32", "mixed_f32_bf16", "fac32", "fac16"])
@pytest.mark.parametrize("kind", ["pcg", "bicg"])
@pytest.mark.parametrize("ordering", ORDERINGS)
def test_krylov_mixed(setup, orc, ordering, kind, precision):
    """Low-precision recurrence + FP64 reliable updates reaches the FP64
    tolerance on the true residual (checked independently with the canonical
    FP64 MatVec).  Its error against a 1e-13 FP64 solve is bounded by
    cond(A) * relres; the stated bound for these small systems is 1e-5."""
    prob, ref, s, gpu, torch = setup["prob"], setup["ref"], setup["sys"], setup["gpu"], setup["torch"]
    skip_unsuited(prob, kind)
    _case(f"{kind} {precision} {ordering}")
    M = make_ilu(gpu, 1 if prob.kind == "oseen2d" else 0, s.A, ordering)
    assert M.SetValue(s.A)
    fn = gpu.Solve_PCG if kind == "pcg" else gpu.Solve_PBiCGSTAB

    def run(tol, prec):
        r = torch.from_numpy(ref["res"].copy()).cuda()
        x = torch.zeros_like(r)
        out = fn(tol, 4000, s.A, M, r, x, precision=prec, want_true_relres=1)
        return out, x.cpu().numpy(), true_relres(s, x, ref["res"])

    (ok, conv, it, info), xm, rel = run(1e-10, precision)
    assert ok and conv < 1e-10 and info["true_relres"] < 1e-10
    assert rel < 1.5e-10, rel
    assert info["outer_iters"] >= 1
    (okx, _, _, _), xx, _ = run(1e-13, "fp64")
    assert okx
    err_m = np.linalg.norm(xm - xx) / np.linalg.norm(xx)
    assert err_m <= 1e-5, err_m


def test_unstructured_mesh_vs_reference(gpu, orc):
    """The reference's own tetrahedral mesh through the device path against
    the reference's outputs stored in the fixture: assembled matrix and
    residual and the natural ILU(0) factor bit for bit, Solve_PCG within one
    iteration and 1e-8 of the reference's solution, Solve_PBiCGSTAB within
    25 % of its iterations and 1e-6."""
    import torch
    _case("cylinder_tet vs reference")
    d = np.load(os.path.join(GOLDEN, "cylinder_tet.npz"), allow_pickle=False)
    prob = cylinder_mesh()
    s = gpu.FEMSystem(prob)
    s.assemble()
    val, dia = s.A.values()
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(val) == str(d["sha_val"]) and sha(dia) == str(d["sha_dia"])
    assert np.array_equal(s.res.cpu().numpy(), d["res"])
    M = gpu.CMatDiaFrac_BlkCrs(0, s.A, ordering="natural")
    assert M.SetValue(s.A)
    fv, fd = M.values()
    assert sha(fv) == str(d["sha_fval_natural_k0"]) and sha(fd) == str(d["sha_fdiav_natural_k0"])
    for kind, solve in (("pcg", gpu.Solve_PCG), ("bicg", gpu.Solve_PBiCGSTAB)):
        r = s.res.clone()
        x = torch.zeros_like(r)
        ok, conv, it, _ = solve(1e-10, 3000, s.A, M, r, x)
        want = d[f"{kind}_natural_k0"]
        x_ref = d[f"{kind}_x_natural_k0"]
        assert ok and bool(want[0])
        pcg = kind == "pcg"
        assert abs(it - int(want[2])) <= (1 if pcg else max(2, int(want[2]) // 4)), (kind, it, want[2])
        assert np.linalg.norm(x.cpu().numpy() - x_ref) <= (1e-8 if pcg else 1e-6) * np.linalg.norm(x_ref)


def test_field_update_cycle(setup):
    """The femls loop around the solver: assemble, solve for the update in
    FP64, UpdateValueOfField (value += update), assemble again at the new
    values.  For these linear problems the new residual is the solver's
    final residual: below 1e-9 of the first, with the matrix unchanged."""
    prob, gpu, torch = setup["prob"], setup["gpu"], setup["torch"]
    _case(f"field update {prob.name}")
    s = gpu.FEMSystem(prob)
    n0 = s.assemble()
    val0, dia0 = s.A.values()
    M = gpu.CMatDiaFrac_BlkCrs(1 if prob.kind == "oseen2d" else 0, s.A)
    assert M.SetValue(s.A)
    fn = gpu.Solve_PBiCGSTAB if prob.kind == "oseen2d" else gpu.Solve_PCG
    r = s.res.clone()
    x = torch.zeros_like(r)
    ok, conv, it, _ = fn(1e-11, 5000, s.A, M, r, x)
    assert ok
    s.UpdateValueOfField(x)
    n1 = s.assemble()
    val1, dia1 = s.A.values()
    assert np.array_equal(val0, val1) and np.array_equal(dia0, dia1)
    assert n1 <= 1e-9 * n0, (n1, n0)
