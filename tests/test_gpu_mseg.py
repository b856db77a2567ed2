"""Multi-segment systems on the device (through the C-ABI) against the
oracle's restatement, which the compiled reference pins bit for bit
(tests/test_oracle_ref.py): the split blocks, residual and MATVEC, the
cross-segment ILU(k) factor patterns (k = 0, 1, 2, unlimited), factor
values and SolvePrecond bit for bit; Solve_PCG within one iteration and
1e-8, Solve_PBiCGSTAB within 25 % and 1e-6 (the dot products are device
reductions)."""
import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import OrcMS, split_problem

pytestmark = pytest.mark.gpu

CASES = {
    "solid3d_uv_w": (lambda: meshgen.solid3d(4, seed=7, sigma=0.5), [[0, 1], [2]], "pcg"),
    "solid3d_u_v_w": (lambda: meshgen.solid3d(4, seed=7, sigma=0.5), [[0], [1], [2]], "pcg"),
    "oseen2d_uv_p": (lambda: meshgen.oseen2d(12, seed=7), [[0, 1], [2]], "bicg"),
}


def setup(orc, gpu, name):
    make, groups, kind = CASES[name]
    prob = make()
    print(f"[case] {name}", flush=True)
    S = gpu.MultiSegmentSystem(prob, groups)
    nres = S.assemble()
    sp = split_problem(orc, prob, groups)
    O = OrcMS(orc, sp, prob.conn)
    return prob, groups, kind, S, nres, O


@pytest.mark.parametrize("name", sorted(CASES))
def test_blocks_residual_matvec(orc, gpu, name):
    import torch
    prob, groups, kind, S, nres, O = setup(orc, gpu, name)
    for i in range(len(groups)):
        val, dia = S.dia[i].values()
        assert np.array_equal(val, O.val[(i, i)]) and np.array_equal(dia, O.dia[i]), i
        for j in range(len(groups)):
            if i != j:
                assert np.array_equal(S.nd[(i, j)].values(), O.val[(i, j)]), (i, j)
    assert np.array_equal(S.res.cpu().numpy(), O.res)
    assert nres == pytest.approx(np.linalg.norm(O.res), rel=1e-13)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(O.res.size), rng.standard_normal(O.res.size)
    yd = torch.from_numpy(y.copy()).cuda()
    S.MATVEC(-0.7, torch.from_numpy(x).cuda(), 0.3, yd)
    assert np.array_equal(yd.cpu().numpy(), O.matvec(-0.7, x, 0.3, y))


@pytest.mark.parametrize("name", sorted(CASES))
def test_cross_segment_ilu(orc, gpu, name):
    import torch
    prob, groups, kind, S, nres, O = setup(orc, gpu, name)
    M = S.preconditioner()
    assert M.SetValue() and O.ilu()
    for i in range(len(groups)):
        for j in range(len(groups)):
            if i == j:
                fv, fd = M.values(i, i)
                ofv, ofd = O.factor(i, i)
                assert np.array_equal(fv, ofv) and np.array_equal(fd, ofd), (i, j)
            else:
                assert np.array_equal(M.values(i, j), O.factor(i, j)), (i, j)
    x = np.random.default_rng(2).standard_normal(O.res.size)
    xd = torch.from_numpy(x.copy()).cuda()
    M.SolvePrecond(xd)
    assert np.array_equal(xd.cpu().numpy(), O.precond(x))


@pytest.mark.parametrize("name", sorted(CASES))
def test_segment_solvers(orc, gpu, name):
    import torch
    prob, groups, kind, S, nres, O = setup(orc, gpu, name)
    M = S.preconditioner()
    M.SetValue()
    O.ilu()
    r = S.res.clone()
    x = torch.zeros_like(r)
    ok, conv, it = M.solve(kind, 1e-10, 3000, r, x)
    ok_o, conv_o, it_o, x_o, r_o = O.solve(kind == "bicg", O.res, 1e-10, 3000)
    xh = x.cpu().numpy()
    print(f"  device it={it} conv={conv:.3e}  oracle it={it_o} conv={conv_o:.3e}  "
          f"|x-x_o|/|x_o|={np.linalg.norm(xh - x_o) / np.linalg.norm(x_o):.2e}", flush=True)
    assert ok and ok_o and conv < 1e-10
    if kind == "pcg":
        assert abs(it - it_o) <= 1
        assert np.linalg.norm(xh - x_o) <= 1e-8 * np.linalg.norm(x_o)
    else:
        assert abs(it - it_o) <= max(2, int(np.ceil(0.25 * it_o)))
        assert np.linalg.norm(xh - x_o) <= 1e-6 * np.linalg.norm(x_o)


@pytest.mark.parametrize("ordering,lev", [("multicolour", 0), ("tile_multicolour", 1)])
def test_multicolour_numbering(orc, gpu, ordering, lev):
    """Nodes renumbered (multicolouring, or coloured tiles with ILU(1)): the
    relabelled problem's blocks and cross-segment factor equal the oracle's
    on the same relabelled problem, and the PCG solution mapped back equals
    the natural-numbering one within 1e-8."""
    import torch
    from delfem_amd import meshgen
    prob = meshgen.solid3d(4, seed=7, sigma=0.5)
    print(f"[case] solid3d_4 {ordering} ILU({lev})", flush=True)
    groups = [[0, 1], [2]]
    S = gpu.MultiSegmentSystem(prob, groups, ordering=ordering)
    S.assemble()
    order = S.node_order
    sp = split_problem(orc, meshgen.relabel(prob, order), groups)
    O = OrcMS(orc, sp, S.prob.conn)
    assert np.array_equal(S.res.cpu().numpy(), O.res)
    M = S.preconditioner(lev)
    assert M.SetValue() and O.ilu(lev)
    for i in range(2):
        for j in range(2):
            got = M.values(i, j)
            want = O.factor(i, j)
            if i == j:
                assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
            else:
                assert np.array_equal(got, want)
    r = S.res.clone()
    x = torch.zeros_like(r)
    ok, conv, it = M.solve("pcg", 1e-11, 3000, r, x)
    Sn = gpu.MultiSegmentSystem(prob, groups)
    Sn.assemble()
    Mn = Sn.preconditioner()
    Mn.SetValue()
    rn = Sn.res.clone()
    xn = torch.zeros_like(rn)
    okn, _, itn = Mn.solve("pcg", 1e-11, 3000, rn, xn)
    assert ok and okn
    nb = prob.nnode
    xs = x.cpu().numpy()
    xn = xn.cpu().numpy()
    # segment-ordered vectors: segment 0 (2 dofs per node), then segment 1
    back = np.empty_like(xs)
    off = 0
    for g in groups:
        blk = xs[off:off + nb * len(g)].reshape(nb, len(g))
        out = np.empty_like(blk)
        out[order] = blk
        back[off:off + nb * len(g)] = out.reshape(-1)
        off += nb * len(g)
    print(f"  {ordering} {it} iterations, natural {itn}", flush=True)
    assert np.linalg.norm(back - xn) <= 1e-8 * np.linalg.norm(xn)


FILL_CASES = {
    "solid3d_uv_w": (lambda: meshgen.solid3d(3, seed=7, sigma=0.5), [[0, 1], [2]], "pcg"),
    "solid3d_u_v_w": (lambda: meshgen.solid3d(3, seed=7, sigma=0.5), [[0], [1], [2]], "pcg"),
    "oseen2d_uv_p": (lambda: meshgen.oseen2d(8, seed=7), [[0, 1], [2]], "bicg"),
}


@pytest.mark.parametrize("lev", [1, 2, -1])
@pytest.mark.parametrize("name", sorted(FILL_CASES))
def test_cross_segment_fill(orc, gpu, name, lev):
    """CPreconditioner_ILU(ls, lev) across segments: every block's fill
    pattern, factor values and SolvePrecond equal the oracle's bit for bit;
    the solvers as for ILU(0)."""
    import torch
    make, groups, kind = FILL_CASES[name]
    prob = make()
    print(f"[case] {name} lev {lev}", flush=True)
    S = gpu.MultiSegmentSystem(prob, groups)
    S.assemble()
    O = OrcMS(orc, split_problem(orc, prob, groups), prob.conn)
    M = S.preconditioner(lev)
    assert M.SetValue() and O.ilu(lev)
    n = len(groups)
    for i in range(n):
        for j in range(n):
            rp, cl = M.pattern(i, j)
            orp, ocl = O.pattern(i, j)
            assert np.array_equal(rp, orp) and np.array_equal(cl, ocl), (i, j)
            if i == j:
                fv, fd = M.values(i, i)
                ofv, ofd = O.factor(i, i)
                assert np.array_equal(fv, ofv) and np.array_equal(fd, ofd), (i, j)
            else:
                assert np.array_equal(M.values(i, j), O.factor(i, j)), (i, j)
    x = np.random.default_rng(2).standard_normal(O.res.size)
    xd = torch.from_numpy(x.copy()).cuda()
    M.SolvePrecond(xd)
    assert np.array_equal(xd.cpu().numpy(), O.precond(x))
    r = S.res.clone()
    xs = torch.zeros_like(r)
    ok, conv, it = M.solve(kind, 1e-10, 3000, r, xs)
    ok_o, conv_o, it_o, x_o, _ = O.solve(kind == "bicg", O.res, 1e-10, 3000)
    xh = xs.cpu().numpy()
    print(f"  device it={it} oracle it={it_o} |x-x_o|/|x_o|={np.linalg.norm(xh - x_o) / np.linalg.norm(x_o):.2e}",
          flush=True)
    assert ok and ok_o
    if kind == "pcg":
        assert abs(it - it_o) <= 1
        assert np.linalg.norm(xh - x_o) <= 1e-8 * np.linalg.norm(x_o)
    else:
        assert abs(it - it_o) <= max(2, int(np.ceil(0.25 * it_o)))
        assert np.linalg.norm(xh - x_o) <= 1e-6 * np.linalg.norm(x_o)


SEG_LEVEL_INPUTS = [[(1, 0)], [(2, -1), (0, 1)], [(1, 1), (2, 0)], [(-1, 1)], [(0, -1), (1, 2)]]


@pytest.mark.parametrize("inputs", SEG_LEVEL_INPUTS, ids=[str(i) for i in SEG_LEVEL_INPUTS])
@pytest.mark.parametrize("name", sorted(FILL_CASES))
def test_segment_fill_levels(orc, gpu, name, inputs):
    """SetFillInLevel(lev, ilss) per segment (preconditioner.h:77,
    preconditioner.cpp:108-131): every block's fill pattern, factor values
    and SolvePrecond equal the oracle's bit for bit (the oracle's level
    resolution and symbolic are pinned to the compiled reference in
    tests/test_oracle_ref.py)."""
    import torch
    make, groups, kind = FILL_CASES[name]
    if any(i >= len(groups) for _, i in inputs):
        pytest.skip("segment index beyond this split")
    prob = make()
    print(f"[case] {name} levels {inputs}", flush=True)
    S = gpu.MultiSegmentSystem(prob, groups)
    S.assemble()
    O = OrcMS(orc, split_problem(orc, prob, groups), prob.conn)
    M = S.preconditioner(inputs=inputs)
    assert M.SetValue() and O.ilu(inputs=inputs)
    n = len(groups)
    for i in range(n):
        for j in range(n):
            rp, cl = M.pattern(i, j)
            orp, ocl = O.pattern(i, j)
            assert np.array_equal(rp, orp) and np.array_equal(cl, ocl), (i, j)
            if i == j:
                fv, fd = M.values(i, i)
                ofv, ofd = O.factor(i, i)
                assert np.array_equal(fv, ofv) and np.array_equal(fd, ofd), (i, j)
            else:
                assert np.array_equal(M.values(i, j), O.factor(i, j)), (i, j)
    x = np.random.default_rng(3).standard_normal(O.res.size)
    xd = torch.from_numpy(x.copy()).cuda()
    M.SolvePrecond(xd)
    assert np.array_equal(xd.cpu().numpy(), O.precond(x))


@pytest.mark.parametrize("inputs", [None, [(1, 1)], [(2, -1)]], ids=["lev0", "lev1seg1", "lev2"])
@pytest.mark.parametrize("name", sorted(FILL_CASES))
def test_segment_fill_blocks(orc, gpu, name, inputs):
    """SetFillBlk with several segments (preconditioner.h:80,
    preconditioner.cpp:143-148): listed rows of segment 0 take fill of any
    level in block (0, 0); every block's pattern, factor values and
    SolvePrecond equal the oracle's bit for bit (the oracle is pinned to the
    compiled reference in tests/test_oracle_ref.py)."""
    import torch
    make, groups, kind = FILL_CASES[name]
    prob = make()
    print(f"[case] {name} fill blocks, levels {inputs}", flush=True)
    S = gpu.MultiSegmentSystem(prob, groups)
    S.assemble()
    sp = split_problem(orc, prob, groups)
    nb = sp["nb"]
    blk = [1, nb // 3, nb // 2 + 1, nb - 2]
    O = OrcMS(orc, sp, prob.conn)
    kw = dict(inputs=inputs) if inputs is not None else dict(lev=0)
    M = S.preconditioner(fill_blk=blk, **kw)
    assert M.SetValue() and O.ilu(fill_blk=blk, **kw)
    n = len(groups)
    for i in range(n):
        for j in range(n):
            rp, cl = M.pattern(i, j)
            orp, ocl = O.pattern(i, j)
            assert np.array_equal(rp, orp) and np.array_equal(cl, ocl), (i, j)
            if i == j:
                fv, fd = M.values(i, i)
                ofv, ofd = O.factor(i, i)
                assert np.array_equal(fv, ofv) and np.array_equal(fd, ofd), (i, j)
            else:
                assert np.array_equal(M.values(i, j), O.factor(i, j)), (i, j)
    x = np.random.default_rng(4).standard_normal(O.res.size)
    xd = torch.from_numpy(x.copy()).cuda()
    M.SolvePrecond(xd)
    assert np.array_equal(xd.cpu().numpy(), O.precond(x))


LOOP_CASES = {
    "solid3d_uv_w_pcg": (lambda: meshgen.solid3d(4, seed=7, sigma=0.5), [[0, 1], [2]], "pcg", 0, "natural"),
    "solid3d_u_v_w_pcg_mc": (lambda: meshgen.solid3d(5, seed=3, sigma=0.5), [[0], [1], [2]], "pcg", 0,
                             "multicolour"),
    "oseen2d_uv_p_bicg": (lambda: meshgen.oseen2d(10, seed=7), [[0, 1], [2]], "bicg", 1, "natural"),
    "oseen2d_uv_p_bicg_tiles": (lambda: meshgen.oseen2d(24, seed=5), [[0, 1], [2]], "bicg", 1, "tile_multicolour"),
}


@pytest.mark.parametrize("maxit", [3000, 9, 1, 0])
@pytest.mark.parametrize("name", sorted(LOOP_CASES))
def test_device_loop_bitwise(gpu, name, maxit):
    """The device-driven loops (scalars on the device, the stop flag polled
    every four iterations, updates skipped after it) return the host-driven
    loops' results bit for bit: status, conv_ratio, iteration, x and r, also
    when stopped by max_iter (iteration counts that are no multiple of the
    polling interval) and for a zero right-hand side."""
    import torch
    make, groups, kind, lev, ordering = LOOP_CASES[name]
    print(f"[case] {name} maxit {maxit}", flush=True)
    S = gpu.MultiSegmentSystem(make(), groups, ordering=ordering)
    S.assemble()
    M = S.preconditioner(lev)
    assert M.SetValue()
    tol = 1e-10
    for rhs_kind in ("res", "zero"):
        rhs = S.res.clone() if rhs_kind == "res" else torch.zeros_like(S.res)
        out = {}
        try:
            for dev in (False, True):
                gpu.set_mseg_device_loop(dev)
                r = rhs.clone()
                x = torch.full_like(r, 7.0)
                ok, conv, it = M.solve(kind, tol, maxit, r, x)
                out[dev] = (ok, conv, it, x.cpu().numpy(), r.cpu().numpy())
        finally:
            gpu.set_mseg_device_loop(True)
        print(f"  {rhs_kind}: ok {out[True][0]} conv {out[True][1]:.3e} it {out[True][2]}", flush=True)
        assert out[True][:3] == out[False][:3]
        assert np.array_equal(out[True][3], out[False][3]) and np.array_equal(out[True][4], out[False][4])
        if rhs_kind == "res" and maxit == 3000:
            assert out[True][0] and out[True][1] < tol


def test_graph_cache_many_operands(gpu):
    """More operand sets than the graph cache keeps (eight): every solve,
    replayed or newly captured after an eviction, gives the same bits."""
    import torch
    print("[case] oseen2d_10 12 operand sets", flush=True)
    S = gpu.MultiSegmentSystem(meshgen.oseen2d(10, seed=7), [[0, 1], [2]])
    S.assemble()
    M = S.preconditioner(1)
    assert M.SetValue()
    keep, first = [], None
    for k in range(12):
        r = S.res.clone()
        x = torch.zeros_like(r)
        keep.append((r, x))  # distinct addresses for every solve
        ok, conv, it = M.solve("bicg", 1e-10, 3000, r, x)
        got = (ok, conv, it, x.cpu().numpy())
        if first is None:
            first = got
        assert got[:3] == first[:3] and np.array_equal(got[3], first[3]), k
    r, x = keep[3]
    r.copy_(S.res)
    x.zero_()
    assert M.solve("bicg", 1e-10, 3000, r, x)[:3] == first[:3]
    assert np.array_equal(x.cpu().numpy(), first[3])
