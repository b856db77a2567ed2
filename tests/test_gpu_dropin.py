"""A C++ caller written against the reference compiles unchanged against
include/delfem_amd/ and reproduces the oracle (tests/cpp/poisson_dropin.cpp).

Its element loop is AddLinearSystem_Poisson2D_P1's (src/femeqn/eqn_poisson.cpp:
99-122: per-element Mearge into LsSol::CLinearSystem's public m_Matrix_Dia,
res_c.AddValue into m_Residual), its pattern a Com::CIndexedArray.  Assembled
matrix and residual after FinalizeMarge: bit for bit.  Solvers through the
ILinearSystem_Sol / ILinearSystemPreconditioner_Sol interfaces:
Solve_PCG (device path and the interface loop with a user preconditioner)
within one iteration and 1e-8, Solve_CG within two and 1e-8, the BiCGSTAB pair
within 25 % and 1e-6; SetFillBlk's factor pattern and values bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import problem_emats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "poisson_dropin")


def _case(name):
    print(f"[case] {name}", flush=True)


def oracle_system(orc, prob, value):
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    emat, eres = problem_emats(orc, prob, value)
    val, dia = orc.assemble(prob.nnode, 1, rowptr, col, prob.conn, emat)
    res = orc.assemble_vec(prob.nnode, 1, prob.conn, eres)
    val, dia = orc.set_bc(prob.nnode, 1, rowptr, col, val, dia, prob.flags)
    res = orc.zero_bc(prob.flags, res)
    return dict(rowptr=rowptr, col=col, val=val, dia=dia, res=res)


def run(tmp_path, prob, value, kind, tol, maxit, lev=0, fill=()):
    fill = np.asarray(fill, dtype=np.uint32)
    inp, out = tmp_path / f"in{kind}.bin", tmp_path / f"out{kind}.bin"
    with open(inp, "wb") as f:
        f.write(np.array([prob.nnode, prob.nelem, kind, maxit, lev, fill.size], dtype=np.int32).tobytes())
        f.write(np.array([tol, prob.source], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(prob.coord, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(prob.conn, dtype=np.uint32).tobytes())
        f.write(np.ascontiguousarray(prob.alpha, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(value, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(prob.flags, dtype=np.int32).tobytes())
        f.write(fill.tobytes())
    r = subprocess.run([PROG, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    raw = open(out, "rb").read()
    ints = np.frombuffer(raw[:20], dtype=np.int32)
    rest = np.frombuffer(raw[20:], dtype=np.float64)
    conv, norm_res = rest[0], rest[1]
    return ints, conv, norm_res, rest[2:]


def split(rest, n, ncrs, nf=0):
    val, rest = rest[:ncrs], rest[ncrs:]
    dia, rest = rest[:n], rest[n:]
    res, rest = rest[:n], rest[n:]
    x, rest = rest[:n], rest[n:]
    fv, rest = rest[:nf], rest[nf:]
    return val, dia, res, x, fv, rest[:n]


@pytest.fixture(scope="module")
def poisson(orc):
    prob = meshgen.poisson2d(20, seed=31)
    value = np.random.default_rng(5).standard_normal(prob.nnode) * 0.1
    value[prob.flags == 1] = 0.0
    return prob, value, oracle_system(orc, prob, value)


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5])
def test_dropin_assembly_and_solvers(tmp_path, orc, poisson, kind):
    prob, value, ref = poisson
    _case(f"poisson_dropin kind {kind}")
    ncrs = ref["col"].size
    ints, conv, norm_res, rest = run(tmp_path, prob, value, kind, 1e-10, 2000)
    ok, it, _, stale, kept = (int(v) for v in ints)
    val, dia, res, x, _, _ = split(rest, prob.nnode, ncrs)
    assert np.array_equal(val, ref["val"]) and np.array_equal(dia, ref["dia"]), "assembled matrix differs"
    assert np.array_equal(res, ref["res"]), "assembled residual differs"
    assert norm_res == pytest.approx(np.linalg.norm(ref["res"]), rel=1e-12)
    assert kept == 1
    if kind in (0, 4):
        assert stale == 1
    if kind in (2, 3):
        ok_o, conv_o, it_o, x_o, _ = orc.solve_ls("cg" if kind == 2 else "bicgstab", prob.nnode, 1, ref["rowptr"],
                                                  ref["col"], ref["val"], ref["dia"], ref["res"], 1e-10, 2000)
    else:
        F = orc.iluk_symbolic(ref["rowptr"], ref["col"], 0)
        fv, fd, _ = orc.ilu_values(prob.nnode, 1, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F)
        ok_o, conv_o, it_o, x_o, _ = orc.solve("pcg" if kind in (0, 1) else "bicg", prob.nnode, 1, ref["rowptr"],
                                               ref["col"], ref["val"], ref["dia"], F, fv, fd, ref["res"], 1e-10, 2000)
    assert ok == 1 and ok_o and conv < 1e-10
    if kind in (0, 1, 2):
        assert abs(it - it_o) <= (1 if kind < 2 else 2), (it, it_o)
        assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)
    else:
        assert abs(it - it_o) <= max(2, it_o // 4), (it, it_o)
        assert np.linalg.norm(x - x_o) <= 1e-6 * np.linalg.norm(x_o)


@pytest.mark.parametrize("kind", [1, 2])
def test_dropin_max_iter(tmp_path, orc, poisson, kind):
    """The interface loops stopped by max_iter: Solve_PCG false with
    iteration unchanged, Solve_CG true with iteration = max_iter and
    conv_ratio unchanged (solver_ls_iter.cpp:68-97, 166-171)."""
    prob, value, ref = poisson
    _case(f"poisson_dropin max_iter kind {kind}")
    ints, conv, _, _ = run(tmp_path, prob, value, kind, 1e-14, 4)
    ok, it = int(ints[0]), int(ints[1])
    assert it == 4
    if kind == 1:
        assert ok == 0 and 1e-14 < conv < 1.0
    else:
        assert ok == 1 and conv == 1e-14


@pytest.mark.parametrize("lev", [0, 1, 2])
def test_dropin_fill_blocks(tmp_path, orc, poisson, lev):
    """CPreconditioner_ILU::SetFillBlk (preconditioner.h:80, preconditioner.cpp:
    149-150): the factor pattern and values of AddFracPtn(lev, aBlkFill)."""
    prob, value, ref = poisson
    _case(f"poisson_dropin fill blocks lev {lev}")
    n = prob.nnode
    fill = [5, n // 3, n // 2 + 7, n - 4]
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], lev, fill_blk=fill)
    nf = F["col"].size
    ints, conv, _, rest = run(tmp_path, prob, value, 6, 1e-10, 2000, lev=lev, fill=fill)
    ok, it, ncrs_f = int(ints[0]), int(ints[1]), int(ints[2])
    assert ncrs_f == nf
    _, _, _, x, fval, fdia = split(rest, n, ref["col"].size, nf)
    fv, fd, ns = orc.ilu_values(n, 1, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F)
    assert ns == 0 and np.array_equal(fval, fv) and np.array_equal(fdia, fd)
    ok_o, _, it_o, x_o, _ = orc.solve("pcg", n, 1, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, fv, fd,
                                      ref["res"], 1e-10, 2000)
    assert ok == 1 and ok_o and abs(it - it_o) <= 1
    assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)


@pytest.mark.parametrize("lev", [0, 1, -1])
def test_fill_blocks_blocked(orc, gpu, lev):
    """The same through the library's Python binding on 3x3 blocks."""
    from oracle.oracle import assemble_problem
    prob = meshgen.solid3d(4, seed=8, sigma=0.5)
    _case(f"fill blocks solid3d lev {lev}")
    ref = assemble_problem(orc, prob)
    s = gpu.FEMSystem(prob)
    s.assemble()
    fill = [2, 40, 77, prob.nnode - 1]
    M = gpu.CMatDiaFrac_BlkCrs(lev, s.A, fill_blk=fill)
    assert M.SetValue(s.A)
    P = M.pattern()
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], lev, fill_blk=fill)
    assert np.array_equal(P["rowptr"], F["rowptr"]) and np.array_equal(P["col"], F["col"])
    assert np.array_equal(P["lev"], F["lev"])
    fv, fd, ns = orc.ilu_values(prob.nnode, 3, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F)
    gv, gd = M.values()
    assert ns == 0 and np.array_equal(gv, fv) and np.array_equal(gd, fd)
