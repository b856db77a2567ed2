"""The C++ drop-in with blocks of different sizes: a program written the way
the reference's rigid-body system drives a segment of per-block lengths
(AddLinSysSeg(nblk, aBlkSize), AddMat_Dia with a Com::CIndexedArray,
one-block Mearge calls, per block-dof AddValue / GetValue; see
tests/cpp/flex_dropin.cpp) compiles unchanged against include/delfem_amd/,
assembles bitwise what the oracle assembles and solves within one iteration
of the oracle's LsSol solvers (1e-8 in the solution)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "flex_dropin")
KINDS = {0: "pcg", 1: "pbicgstab", 2: "cg", 3: "bicgstab", 4: "pcg", 5: "pcg"}


def _run(tmp_path, P, kind, lev, tol, maxit):
    blocks = P["blocks"]
    inp, out = tmp_path / f"in{kind}", tmp_path / f"out{kind}"
    with open(inp, "wb") as f:
        np.array([P["nb"], len(blocks), kind, maxit, lev], dtype=np.int32).tofile(f)
        np.array([tol], dtype=np.float64).tofile(f)
        np.asarray(P["alen"], dtype=np.uint32).tofile(f)
        np.asarray(P["rowptr"], dtype=np.uint32).tofile(f)
        np.asarray(P["col"], dtype=np.uint32).tofile(f)
        np.array([b[0] for b in blocks], dtype=np.uint32).tofile(f)
        np.array([b[1] for b in blocks], dtype=np.uint32).tofile(f)
        np.concatenate([np.ravel(b[2]) for b in blocks]).astype(np.float64).tofile(f)
        np.asarray(P["rhs"], dtype=np.float64).tofile(f)
    print(f"[case] flex_dropin kind {kind} lev {lev}", flush=True)
    r = subprocess.run([PROG, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    ints = np.frombuffer(raw[:16], dtype=np.int32)
    rest = np.frombuffer(raw[16:], dtype=np.float64)
    return ints, rest


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5])
def test_flex_dropin(tmp_path, orc, kind):
    sym = KINDS[kind] in ("pcg", "cg")
    P = O.flex_problem(nb=60, seed=30 + kind, symmetric=sym, nfix=0)
    M = O.flex_assemble(orc, P)
    lev = 1
    ints, rest = _run(tmp_path, P, kind, lev, 1e-10, 1000)
    ok, it, refused, len_m1 = (int(v) for v in ints)
    assert refused == 1 and len_m1 == 1
    nv, nd, n = M["val"].size, M["dia"].size, int(P["dofptr"][-1])
    conv, norm_res = rest[0], rest[1]
    val, dia, x = rest[2:2 + nv], rest[2 + nv:2 + nv + nd], rest[2 + nv + nd:2 + nv + nd + n]
    assert np.array_equal(val, M["val"]) and np.array_equal(dia, M["dia"])
    assert norm_res == pytest.approx(np.linalg.norm(P["rhs"]), rel=1e-14)
    A = O.flex_dense(M, M["val"], M["dia"])
    xs = np.linalg.solve(A, P["rhs"])
    assert ok == 1, (kind, it, conv)
    assert np.linalg.norm(x - xs) <= 1e-7 * np.linalg.norm(xs), kind
    if kind != 5:  # the LsSol loops: iterations against the oracle's
        pre = KINDS[kind] in ("pcg", "pbicgstab")
        F = orc.fx_factor(M, lev)
        fv, fd, _ = orc.fx_ilu_values(M, M["val"], M["dia"], F)
        oo, oc, oi, ox, _ = orc.fx_solve(KINDS[kind], M, M["val"], M["dia"], F if pre else None,
                                         fv if pre else None, fd if pre else None, P["rhs"], 1e-10, 1000)
        assert abs(it - oi) <= 1, (kind, it, oi)
        assert np.linalg.norm(x - ox) <= 1e-8 * np.linalg.norm(ox), kind
