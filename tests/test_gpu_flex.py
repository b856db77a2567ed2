"""Variable block sizes ("flex") on the device, against the oracle (itself
pinned bitwise against the compiled reference, tests/test_oracle_flex.py):
layout, Mearge, Dirichlet rows and columns, MatVec, the ILU(k) factor at
levels 0-3 with its stop at the 11th singular pivot and both substitutions
bit for bit; the four LsSol solvers within one iteration of the oracle and
1e-8 in the solution (their dot products are parallel reductions)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# small random graphs (dense fill at higher levels) and larger banded ones
# (random links at most 8 blocks apart, bounded fill)
CASES = [(30, 0, True), (30, 1, False), (2000, 2, True), (2000, 3, False)]


def _problem(nb, seed, sym, **kw):
    return O.flex_problem(nb=nb, seed=seed, symmetric=sym, window=8 if nb > 100 else 0, **kw)


def _case(name):
    print(f"[case] {name}", flush=True)


def _device(gpu, P):
    A = gpu.CMatDiaFlex_BlkCrs(P["nb"], P["alen"])
    A.AddPattern(P["rowptr"], P["col"])
    A.SetZero()
    A.Mearge(P["blocks"])
    A.SetBoundaryCondition(P["flags"])
    return A


@pytest.mark.parametrize("nb,seed,sym", CASES)
def test_flex_assembly_matvec(gpu, orc, nb, seed, sym):
    import torch
    _case(f"flex assembly nb={nb} seed={seed} sym={sym}")
    P = _problem(nb, seed, sym, nfix=max(3, nb // 20))
    M = O.flex_assemble(orc, P)
    A = _device(gpu, P)
    L = A.layout()
    for k in ("dofptr", "rowptr", "col", "valptr", "diaptr"):
        assert np.array_equal(L[k], M[k]), k
    val, dia = A.values()
    assert np.array_equal(val, M["val"]) and np.array_equal(dia, M["dia"])
    n = int(P["dofptr"][-1])
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    A.MatVec(0.7, xd, -0.3, yd)
    assert np.array_equal(yd.cpu().numpy(), orc.fx_matvec(M, M["val"], M["dia"], 0.7, x, -0.3, y))


@pytest.mark.parametrize("nb,seed,sym", CASES)
@pytest.mark.parametrize("lev", [0, 1, 2, 3])
def test_flex_factor_and_sweeps(gpu, orc, nb, seed, sym, lev):
    import torch
    _case(f"flex factor nb={nb} seed={seed} sym={sym} lev={lev}")
    P = _problem(nb, seed, sym, nfix=max(3, nb // 20))
    M = O.flex_assemble(orc, P)
    A = _device(gpu, P)
    Fd = gpu.CMatDiaFracFlex_BlkCrs(lev, A)
    F = orc.fx_factor(M, lev)
    pat = Fd.pattern()
    for k in ("rowptr", "col", "valptr"):
        assert np.array_equal(pat[k], F[k]), k
    assert np.array_equal(pat["diaind"], F["diaind"][:nb])
    assert Fd.SetValue(A)
    fv, fd, ok = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    gv, gd = Fd.values()
    assert ok and np.array_equal(gv, fv[:gv.size]) and np.array_equal(gd, fd[:gd.size])
    n = int(P["dofptr"][-1])
    v = np.random.default_rng(seed + 7).uniform(-1, 1, n)
    vd = torch.from_numpy(v).cuda()
    Fd.ForwardSubstitution(vd)
    f = orc.fx_forward(F, fv, fd, v)
    assert np.array_equal(vd.cpu().numpy(), f)
    Fd.BackwardSubstitution(vd)
    assert np.array_equal(vd.cpu().numpy(), orc.fx_backward(F, fv, f))


@pytest.mark.parametrize("nzero", [10, 11, 14])
def test_flex_singular_stop(gpu, orc, nzero):
    _case(f"flex singular nzero={nzero}")
    rows = list(range(3, 3 + 3 * nzero, 3))
    P = O.flex_problem(nb=60, seed=9, nfix=0, zero_rows=rows)
    M = O.flex_assemble(orc, P)
    A = _device(gpu, P)
    Fd = gpu.CMatDiaFracFlex_BlkCrs(1, A)
    okd = Fd.SetValue(A)
    F = orc.fx_factor(M, 1)
    fv, fd, ok = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    assert okd == ok == (nzero <= 10)
    assert Fd.nsingular >= min(nzero, 11)
    gv, gd = Fd.values()
    assert np.array_equal(gv, fv[:gv.size]) and np.array_equal(gd, fd[:gd.size])


@pytest.mark.parametrize("nb,seed,sym", CASES)
def test_flex_solvers(gpu, orc, nb, seed, sym):
    import torch
    _case(f"flex solvers nb={nb} seed={seed} sym={sym}")
    P = _problem(nb, seed, sym, nfix=max(3, nb // 20))
    M = O.flex_assemble(orc, P)
    A = _device(gpu, P)
    kinds = ("pcg", "cg") if sym else ("pbicgstab", "bicgstab")
    for lev in (0, 2):
        Fd = gpu.CMatDiaFracFlex_BlkCrs(lev, A)
        assert Fd.SetValue(A)
        F = orc.fx_factor(M, lev)
        fv, fd, _ = orc.fx_ilu_values(M, M["val"], M["dia"], F)
        for kind in kinds:
            pre = kind in ("pcg", "pbicgstab")
            r = torch.from_numpy(P["rhs"].copy()).cuda()
            x = torch.zeros_like(r)
            ok, conv, it = gpu.solve_flex(kind, 1e-10, 2000, A, Fd if pre else None, r, x)
            oo, oc, oi, ox, _ = orc.fx_solve(kind, M, M["val"], M["dia"], F if pre else None, fv if pre else None,
                                             fd if pre else None, P["rhs"], 1e-10, 2000)
            assert ok and oo, (kind, lev)
            assert abs(it - oi) <= 1, (kind, lev, it, oi)
            xs = x.cpu().numpy()
            assert np.linalg.norm(xs - ox) <= 1e-8 * np.linalg.norm(ox), (kind, lev)


def test_flex_solver_limits(gpu, orc):
    """max_iter and zero right-hand side behave as the reference's."""
    import torch
    _case("flex solver limits")
    P = O.flex_problem(nb=40, seed=21, nfix=2)
    M = O.flex_assemble(orc, P)
    A = _device(gpu, P)
    Fd = gpu.CMatDiaFracFlex_BlkCrs(0, A)
    Fd.SetValue(A)
    for kind in ("pcg", "pbicgstab", "cg", "bicgstab"):
        pre = kind in ("pcg", "pbicgstab")
        r = torch.from_numpy(P["rhs"].copy()).cuda()
        x = torch.zeros_like(r)
        ok, conv, it = gpu.solve_flex(kind, 1e-14, 3, A, Fd if pre else None, r, x)
        assert ok == (kind != "pcg") and it == 3, (kind, ok, it)
        z = torch.zeros_like(r)
        x = torch.zeros_like(r)
        ok, conv, it = gpu.solve_flex(kind, 1e-10, 50, A, Fd if pre else None, z, x)
        assert ok and it == 0 and conv == 0.0 and float(x.abs().max()) == 0.0, kind
