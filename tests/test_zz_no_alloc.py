"""dfm_ilu_solve_precond_noalloc (the allocation-free SolvePrecond) against
dfm_ilu_solve_precond through the C-ABI: bit for bit, on two calls in a row
(the second on a smaller system, so the shared scratch holds the first call's
values), for ordered and natural-order factors; an ordered factor larger than
the scratch is refused with DFM_TOO_LARGE."""
import ctypes as C

import numpy as np
import pytest

from delfem_amd import _lib, meshgen

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


def _factor(gpu, prob, lev, ordering):
    s = gpu.FEMSystem(prob)
    s.assemble()
    M = gpu.CPreconditioner_ILU(s.A, lev, ordering=ordering, block_nodes=8)
    assert M.SetValue(s.A)
    return s, M


def _both(M, n, seed):
    import torch
    v = torch.from_numpy(np.random.default_rng(seed).standard_normal(n)).cuda()
    a, b = v.clone(), v.clone()
    torch.cuda.synchronize()
    _lib.call("dfm_ilu_solve_precond", M.h, C.c_void_p(a.data_ptr()))
    _lib.call("dfm_ilu_solve_precond_noalloc", M.h, C.c_void_p(b.data_ptr()))
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("ordering", ["multicolour", "block_multicolour", "natural"])
def test_noalloc_matches_solve_precond(gpu, ordering):
    big = (meshgen.solid3d(6, seed=4, sigma=0.5), 1)
    small = (meshgen.oseen2d(7, seed=4), 1)
    for k, (prob, lev) in enumerate((big, small)):
        _case(f"noalloc {ordering} call {k} {prob.name}")
        _, M = _factor(gpu, prob, lev, ordering)
        want, got = _both(M, prob.nnode * prob.b, 10 + k)
        assert np.isfinite(want).all()
        assert np.array_equal(got, want)


def test_noalloc_refuses_too_large(gpu):
    _case("noalloc too large")
    import torch
    nb = (1 << 22) + 5                        # one more than the scratch holds, scalar blocks
    A = gpu.CMatDia_BlkCrs(nb, 1)
    A.AddPattern(np.zeros(nb + 1, dtype=np.uint32), np.zeros(0, dtype=np.uint32))
    A.set_values(np.zeros(0), np.full(nb, 2.0))
    M = gpu.CMatDiaFrac_BlkCrs(0, A, ordering="multicolour")
    assert M.SetValue(A)
    v = torch.ones(nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = _lib.lib().dfm_ilu_solve_precond_noalloc(M.h, C.c_void_p(v.data_ptr()))
    assert rc == _lib.DFM_TOO_LARGE
    assert b"allocation-free" in _lib.lib().dfm_last_error()
    assert torch.all(v == 1.0).item()        # untouched
