"""Variable block sizes ("flex"): the oracle's restatement of the reference's
flex branches (MatVec::CMatDia_BlkCrs / CMatDiaFrac_BlkCrs with aLen,
src/matvec/matdia_blkcrs.cpp:173-207, 268-306, 931-960,
src/matvec/matdiafrac_blkcrs.cpp:339-379, 512-553, 777-846; the LsSol
solvers over a segment made by AddLinSysSeg(nnode, aLen),
src/ls/linearsystem.cpp:170-201) pinned bitwise against the compiled
reference, and by what the mathematics fixes: the Dirichlet rows and
columns equal the fixed-size routine's when all blocks have one length, the
MatVec equals a dense product, an unlimited-fill factor solves the system
exactly, and the factorisation stops at the 11th singular pivot.

The reference's CBCFlag::SetBC refuses flex flags (bcflag_blk.h:153-158: it
asserts and returns false), so a flex segment of the reference never holds a
fixed dof: the live-reference cases carry no Dirichlet flags and the flex
SetBoundaryCondition is pinned by the fixed-size comparison.
"""
import numpy as np
import pytest

from oracle import oracle as O

CASES = [(seed, sym) for seed in range(3) for sym in (True, False)]


@pytest.mark.parametrize("seed,sym", CASES)
def test_flex_against_reference(orc, ref, seed, sym):
    P = O.flex_problem(nb=30, seed=seed, symmetric=sym, nfix=0)
    M = O.flex_assemble(orc, P)
    R = O.RefFlex(ref, P)
    R.merge(P["blocks"])
    R.finalize()
    RM = R.matrix()
    for k in ("rowptr", "col", "valptr", "diaptr", "val", "dia"):
        assert np.array_equal(RM[k], M[k]), k
    rng = np.random.default_rng(seed + 100)
    n = int(P["dofptr"][-1])
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert np.array_equal(R.matvec(0.7, x, -0.3, y), orc.fx_matvec(M, M["val"], M["dia"], 0.7, x, -0.3, y))
    kinds = ("pcg", "cg") if sym else ("pbicgstab", "bicgstab")
    for lev in (0, 1, 2, 3):
        ok, RF, rfv, rfd = R.factor(lev)
        F = orc.fx_factor(M, lev)
        for k in ("rowptr", "col", "valptr"):
            assert np.array_equal(RF[k], F[k]), (lev, k)
        assert np.array_equal(RF["diaind"], F["diaind"][:P["nb"]])
        fv, fd, ok2 = orc.fx_ilu_values(M, M["val"], M["dia"], F)
        assert ok and ok2
        assert np.array_equal(rfv, fv[:rfv.size]) and np.array_equal(rfd, fd[:rfd.size]), lev
        fwd, bwd = R.precond(lev, x)
        f2 = orc.fx_forward(F, fv, fd, x)
        assert np.array_equal(fwd, f2) and np.array_equal(bwd, orc.fx_backward(F, fv, f2)), lev
        for kind in kinds:
            r1 = R.solve(kind, lev, P["rhs"], 1e-10, 500)
            o1 = orc.fx_solve(kind, M, M["val"], M["dia"], F, fv, fd, P["rhs"], 1e-10, 500)
            assert bool(r1[0]) == o1[0] and r1[2] == o1[2] and r1[1] == o1[1], (kind, lev, r1[:3], o1[:3])
            assert np.array_equal(r1[3], o1[3]) and np.array_equal(r1[4], o1[4]), (kind, lev)


def test_flex_max_iter_and_zero_rhs(orc, ref):
    P = O.flex_problem(nb=20, seed=7, nfix=0)
    M = O.flex_assemble(orc, P)
    R = O.RefFlex(ref, P)
    R.merge(P["blocks"])
    R.finalize()
    F = orc.fx_factor(M, 0)
    fv, fd, _ = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    for kind in ("pcg", "pbicgstab", "cg", "bicgstab"):
        for rhs, it in ((P["rhs"], 3), (np.zeros_like(P["rhs"]), 50)):
            r1 = R.solve(kind, 0, rhs, 1e-12, it)
            o1 = orc.fx_solve(kind, M, M["val"], M["dia"], F, fv, fd, rhs, 1e-12, it)
            assert bool(r1[0]) == o1[0] and r1[1] == o1[1] and r1[2] == o1[2], (kind, it, r1[:3], o1[:3])
            assert np.array_equal(r1[3], o1[3]) and np.array_equal(r1[4], o1[4])


@pytest.mark.parametrize("nzero", [10, 11, 13])
def test_flex_singular_stop(orc, ref, nzero):
    """DoILUDecomp goes on past up to 10 singular pivots and stops at the
    11th, leaving later rows as copied (bitwise the reference's state)."""
    rows = list(range(2, 2 + 2 * nzero, 2))
    P = O.flex_problem(nb=40, seed=3, nfix=0, zero_rows=rows)
    M = O.flex_assemble(orc, P)
    R = O.RefFlex(ref, P)
    R.merge(P["blocks"])
    R.finalize()
    ok, RF, rfv, rfd = R.factor(1)
    F = orc.fx_factor(M, 1)
    fv, fd, ok2 = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    assert ok == ok2 == (nzero <= 10)
    assert np.array_equal(rfv, fv[:rfv.size]) and np.array_equal(rfd, fd[:rfd.size])


def test_flex_bc_equals_fixed_for_uniform_blocks(orc):
    """With every block of length b the flex Dirichlet treatment is the
    fixed-size SetBoundaryCondition (matdia_blkcrs.cpp:308-343), itself
    pinned against the reference."""
    for b in (1, 2, 3):
        P = O.flex_problem(nb=25, seed=b, lens=[b] * 25, nfix=5)
        M = dict(dofptr=P["dofptr"], rowptr=P["rowptr"], col=P["col"])
        M["valptr"], M["diaptr"] = orc.fx_layout(M["dofptr"], M["rowptr"], M["col"])
        val, dia = orc.fx_mearge(M, P["blocks"])
        v1, d1 = orc.fx_set_bc(M, val, dia, P["flags"])
        v2, d2 = orc.set_bc(25, b, P["rowptr"], P["col"], val, dia, P["flags"])
        assert np.array_equal(v1, v2) and np.array_equal(d1, d2), b
        assert not np.array_equal(v1, val)


def test_flex_matvec_dense_and_exact_lu(orc):
    P = O.flex_problem(nb=25, seed=11, symmetric=False, nfix=4)
    M = O.flex_assemble(orc, P)
    A = O.flex_dense(M, M["val"], M["dia"])
    n = A.shape[0]
    rng = np.random.default_rng(5)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    yy = orc.fx_matvec(M, M["val"], M["dia"], 1.3, x, 0.4, y)
    assert np.allclose(yy, 1.3 * A @ x + 0.4 * y, rtol=1e-13, atol=1e-12)
    # fixed dofs: unit rows and columns
    fix = np.nonzero(P["flags"])[0]
    assert np.all(A[fix][:, fix] == np.eye(fix.size)) and np.all(A[fix, :].sum(axis=1) == 1.0)
    # unlimited fill: the factor is an exact LU
    F = orc.fx_factor(M, 10 ** 6)
    fv, fd, ok = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    assert ok
    b = rng.uniform(-1, 1, n)
    z = orc.fx_backward(F, fv, orc.fx_forward(F, fv, fd, b))
    assert np.linalg.norm(z - np.linalg.solve(A, b)) < 1e-10 * np.linalg.norm(z)
    # and a textbook PCG on the symmetric case converges to the solution
    P = O.flex_problem(nb=25, seed=12, nfix=0)
    M = O.flex_assemble(orc, P)
    A = O.flex_dense(M, M["val"], M["dia"])
    F = orc.fx_factor(M, 0)
    fv, fd, _ = orc.fx_ilu_values(M, M["val"], M["dia"], F)
    ok, conv, it, xs, _ = orc.fx_solve("pcg", M, M["val"], M["dia"], F, fv, fd, P["rhs"], 1e-12, 300)
    assert ok and np.linalg.norm(A @ xs - P["rhs"]) < 1e-10 * np.linalg.norm(P["rhs"])
