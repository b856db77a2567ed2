"""The oracle against the live compiled reference (oracle/_ref/libdfmref.so),
at sizes beyond the fixtures: the full 2-D Poisson configuration (16,641
nodes) and random graphs for ILU(k) patterns.  Skipped where the reference
library was not built."""
import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import RefFrac, RefILU, RefSystem, assemble_problem


def test_poisson_128_pcg_trajectory(orc, ref):
    """Configuration 1 (scalar2d path) end to end: ILU(0)-PCG to 1e-8 with
    identical iteration count, conv_ratio and solution bits."""
    prob = meshgen.poisson2d(128, seed=12345)
    A = assemble_problem(orc, prob)
    rs = RefSystem(ref, prob.nnode, 1, A["rowptr"], A["col"])
    rs.merge(prob.conn, A["emat"], A["eres"])
    rs.set_bc(prob.flags)
    rs.finalize()
    rp, cl, val, dia = rs.matrix()
    assert np.array_equal(val, A["val"]) and np.array_equal(dia, A["dia"])
    pr = RefILU(rs, 0)
    assert pr.set_value()
    ok_r, conv_r, it_r = pr.solve("pcg", 1e-8, 1000)
    x_r = rs.get_vector(-2)
    F = orc.iluk_symbolic(A["rowptr"], A["col"], 0)
    fv, fd, _ = orc.ilu_values(prob.nnode, 1, A["rowptr"], A["col"], A["val"], A["dia"], F)
    ok, conv, it, x, _ = orc.solve("pcg", prob.nnode, 1, A["rowptr"], A["col"], A["val"], A["dia"], F, fv, fd,
                                   A["res"], 1e-8, 1000)
    assert (ok, conv, it) == (ok_r, conv_r, it_r)
    assert np.array_equal(x, x_r)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("lev", [1, 2, 3, -1])
def test_iluk_random_graphs(orc, ref, seed, lev):
    rng = np.random.default_rng(seed)
    n = 120
    adj = [set() for _ in range(n)]
    for i in range(n):
        for _ in range(2):
            j = int(rng.integers(n))
            if j != i:
                adj[i].add(j)
                adj[j].add(i)
    rowptr = np.concatenate([[0], np.cumsum([len(a) for a in adj])]).astype(np.uint32)
    col = np.array([j for a in adj for j in sorted(a)], dtype=np.uint32)
    rs = RefSystem(ref, n, 1, rowptr, col)
    for order in (None, rng.permutation(n).astype(np.int32)):
        P = RefFrac(rs, lev, order).pattern()
        F = orc.iluk_symbolic(rowptr, col, lev, order)
        for k in ("rowptr", "col", "diaind"):
            assert np.array_equal(P[k], F[k]), (k, lev)
