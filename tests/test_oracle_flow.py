"""The oracle's Oseen P1-P1 PSPG element, Helmholtz element and block
multicolour ordering, pinned without the reference.

The Oseen element restates terms of MakeMat_NavierStokes2D_NonStatic_Newmark_P1P1
(src/femeqn/eqn_navier_stokes.cpp:73-110,136-158,231-253); that file does not
compile here, so its element is pinned by an independent quadrature of the weak
forms (basis functions from a 3x3 solve, mid-edge rule, exact for the quadratic
integrands), which fails on a transposed test/trial index, a dropped term or a
wrong sign.  The stabilisation parameter tau is a definition and is restated.
The block multicolour ordering has no reference counterpart and is pinned by
its defining properties."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from delfem_amd import meshgen
from oracle.oracle import assemble_problem, problem_emats

from test_oracle_math import bsr_of


def cross2(u, v):
    return u[0] * v[1] - u[1] * v[0]


def p1_basis(X):
    """phi_k(x) = c_k + g_k . x for the triangle X (3x2)."""
    M = np.column_stack([np.ones(3), X])
    Cf = np.linalg.inv(M)            # column k: coefficients of phi_k
    return Cf[0], Cf[1:].T          # constants (3,), gradients (3, 2)


def midedge_rule(X):
    area = 0.5 * abs(cross2(X[1] - X[0], X[2] - X[0]))
    pts = np.array([(X[0] + X[1]) / 2, (X[1] + X[2]) / 2, (X[2] + X[0]) / 2])
    return pts, np.full(3, area / 3.0), area


def tau_steady(abar, area, myu, rho):
    norm_v = np.hypot(*abar)
    h = np.sqrt(area / 3.14) * 2
    if norm_v * h * rho * 1e-30 > myu:
        return h * 0.5 / norm_v
    if norm_v * h * rho < myu * 1e-30:
        return h * h * rho * 0.5 / myu
    re_c = 0.5 * norm_v * h * rho / myu
    return (h * 0.5 / norm_v) / np.sqrt(1 + 1 / (re_c * re_c))


def oseen_element_quadrature(X, a, myu, rho):
    c0, g = p1_basis(X)
    pts, w, area = midedge_rule(X)
    phi = c0[None, :] + pts @ g.T                     # (q, k)
    aq = phi @ a                                       # advection at the points
    abar = a.mean(axis=0)
    tau = tau_steady(abar, area, myu, rho)
    E = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            gi, gj = g[i], g[j]
            for r in range(2):
                for c in range(2):
                    E[i, j, r, c] = myu * area * ((gi @ gj) * (r == c) + gi[c] * gj[r])
                conv = np.sum(w * phi[:, i] * (aq @ gj)) * rho
                E[i, j, r, r] += conv
                E[i, j, r, 2] = -np.sum(w * gi[r] * phi[:, j])            # -(p, div v)
                E[i, j, 2, r] = np.sum(w * phi[:, i] * gj[r])             # (q, div u)
                E[i, j, 2, r] += tau * area * gi[r] * (abar @ gj)         # PSPG advection
            E[i, j, 2, 2] = tau / rho * area * (gi @ gj)                  # PSPG pressure
    return E


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oseen_element_vs_quadrature(orc, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3, 2))
    if cross2(X[1] - X[0], X[2] - X[0]) < 0:
        X = X[[0, 2, 1]]
    a = rng.standard_normal((3, 2))
    u = rng.standard_normal(9)
    for myu in (0.01, 1.0):
        emat, eres = orc.emat_oseen2d(X, np.array([[0, 1, 2]]), myu, 1.3, a, u)
        want = oseen_element_quadrature(X, a, myu, 1.3)
        got = emat.reshape(3, 3, 3, 3)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
        # residual: -E u per node, in the [i][j][r][c] layout
        Ef = want.transpose(0, 2, 1, 3).reshape(9, 9)
        assert np.allclose(eres, -Ef @ u, rtol=1e-12, atol=1e-13 * np.abs(Ef).max())


def test_oseen_element_invariants(orc):
    p = meshgen.oseen2d(6, seed=4)
    emat, _ = orc.emat_oseen2d(p.coord, p.conn, p.myu, p.rho, p.adv, np.zeros(p.nnode * 3))
    E = emat.reshape(-1, 3, 3, 3, 3)
    scale = np.abs(E).max()
    # a uniform velocity is a null vector of the velocity and continuity rows
    # (sum_j grad phi_j = 0); a uniform pressure is a null vector of the PSPG rows
    assert np.abs(E[:, :, :, :, 0:2].sum(axis=2)).max() < 1e-13 * scale
    assert np.abs(E[:, :, :, 2, 2].sum(axis=2)).max() < 1e-13 * scale
    # advection makes the velocity block nonsymmetric; the PSPG pressure block is symmetric
    visc = E[:, :, :, 0:2, 0:2] - E[:, :, :, 0:2, 0:2].transpose(0, 2, 1, 4, 3)
    assert np.abs(visc).max() > 1e-6 * scale
    pp = E[:, :, :, 2, 2]
    assert np.allclose(pp, pp.transpose(0, 2, 1), atol=1e-15 * scale)
    assert np.all(pp[:, range(3), range(3)] > 0)


def test_helmholtz_element_vs_quadrature(orc):
    rng = np.random.default_rng(5)
    X = np.array([[0.1, 0.0], [1.2, 0.3], [0.4, 0.9]])
    alpha, cm, f = 1.7, 3.5, 0.8
    u = rng.standard_normal(3)
    emat, eres = orc.emat_helmholtz2d(X, np.array([[0, 1, 2]]), np.array([alpha]), np.array([cm]), f, u)
    c0, g = p1_basis(X)
    pts, w, area = midedge_rule(X)
    phi = c0[None, :] + pts @ g.T
    want = alpha * area * (g @ g.T) + cm * np.einsum("q,qi,qj->ij", w, phi, phi)
    assert np.allclose(emat.reshape(3, 3), want, rtol=1e-13)
    assert np.allclose(eres, f * area / 3.0 - want @ u, rtol=1e-12)


def test_oseen_assembly_and_solve_vs_direct(orc):
    p = meshgen.oseen2d(10, seed=7)
    A = assemble_problem(orc, p)
    nb, b = p.nnode, 3
    M = bsr_of(nb, b, A["rowptr"], A["col"], A["val"], A["dia"])
    # assembly = sum of element blocks (before boundary conditions)
    emat, eres = problem_emats(orc, p)
    val, dia = orc.assemble(nb, b, A["rowptr"], A["col"], p.conn, emat)
    raw = bsr_of(nb, b, A["rowptr"], A["col"], val, dia).toarray()
    E = emat.reshape(-1, 3, 3, 3, 3)
    want = np.zeros_like(raw)
    for e, el in enumerate(p.conn):
        for i in range(3):
            for j in range(3):
                want[el[i] * 3:el[i] * 3 + 3, el[j] * 3:el[j] * 3 + 3] += E[e, i, j]
    assert np.allclose(raw, want, rtol=1e-13, atol=1e-14 * np.abs(want).max())
    # residual = -K value on free dofs
    free = p.flags == 0
    assert np.allclose(A["res"][free], (-(want @ p.value))[free], atol=1e-13)
    xd = spla.spsolve(M.tocsc(), A["res"])
    _, _, order = orc.bmc_order(A["rowptr"], A["col"], 16)
    for ordr in (None, order):
        F = orc.iluk_symbolic(A["rowptr"], A["col"], 1, ordr)
        fv, fd, ns = orc.ilu_values(nb, b, A["rowptr"], A["col"], A["val"], A["dia"], F, ordr)
        assert ns == 0
        ok, conv, it, x, r = orc.solve("bicg", nb, b, A["rowptr"], A["col"], A["val"], A["dia"], F, fv, fd,
                                       A["res"], 1e-12, 2000, ordr)
        assert ok and conv < 1e-12
        assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    # the lid drives the flow: nonzero velocity inside, BC values kept
    assert np.abs(xd[free]).max() > 1e-3


@pytest.mark.parametrize("bs", [1, 5, 16, 1000])
def test_bmc_order_properties(orc, bs):
    for p in (meshgen.poisson2d(12, seed=1), meshgen.solid3d(3, seed=1), meshgen.oseen2d(9, seed=1)):
        rowptr, col = orc.fem_pattern(p.nnode, p.conn)
        n = p.nnode
        blk, nc, order = orc.bmc_order(rowptr, col, bs)
        assert np.array_equal(np.sort(order), np.arange(n))
        nbk = blk.max() + 1
        sizes = np.bincount(blk, minlength=nbk)
        assert sizes.min() >= 1 and sizes.max() <= bs
        # blocks are numbered by their seeds: the lowest node of block k rises with k,
        # and block 0 starts at node 0
        first = np.array([np.min(np.nonzero(blk == k)[0]) for k in range(nbk)])
        assert first[0] == 0 and np.all(np.diff(first) > 0)
        # every block is connected
        rows = np.repeat(np.arange(n), np.diff(rowptr.astype(np.int64)))
        for k in range(nbk):
            members = set(np.nonzero(blk == k)[0].tolist())
            seen, stack = {first[k]}, [first[k]]
            while stack:
                u = stack.pop()
                for v in col[rowptr[u]:rowptr[u + 1]]:
                    if v in members and v not in seen:
                        seen.add(int(v))
                        stack.append(int(v))
            assert seen == members
        # each block is contiguous in the order and keeps natural order inside
        seq = blk[order]
        change = np.concatenate([[True], seq[1:] != seq[:-1]])
        assert change.sum() == nbk
        for k in range(nbk):
            assert np.all(np.diff(order[seq == k]) > 0)
        if bs == 1:
            assert nbk == n
        # blocks come in nc runs of mutually non-adjacent blocks (the colour classes):
        # greedy split of the block sequence at the first adjacency gives at most nc runs
        cross = blk[rows] != blk[col]
        qa, qb = blk[rows][cross], blk[col][cross]
        adj = set(zip(qa.tolist(), qb.tolist()))
        run_of = np.zeros(nbk, dtype=np.int64)
        run, members = 0, []
        for k in seq[change]:
            if any((int(k), g) in adj for g in members):
                run += 1
                members = []
            members.append(int(k))
            run_of[k] = run
        assert run + 1 <= max(nc, 1)
        assert not np.any(run_of[qa] == run_of[qb])
