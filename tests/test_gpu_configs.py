"""Every benchmark workload at its full stated size, on the device.

* 2-D Poisson 128^2 (16,641 nodes) and 3-D elasticity 48^3 (352,947 DOF):
  against the oracle -- assembly, residual and the natural-order ILU(0)
  factor bit for bit, natural-order ILU(0)-PCG within one iteration of the
  reference loop and 1e-8 relative error in the solution (the reference's own
  configuration); the benchmark's multicolour ordering in FP64 and in the
  mixed mode on the FP64 true residual (API MatVec on the canonical matrix)
  and against the natural-order solution.
* 3-D elasticity 96^3 (2,738,019 DOF), lognormal(0, 1) modulus: the
  multicolour ILU(0) factor bit for bit against the oracle's ordered
  factorisation of the device-assembled matrix (assembly itself is pinned
  bitwise at 48^3); FP64 and mixed (FP32 matrix and vectors, BF16 factors,
  FP64 refinement) PCG to 1e-10 checked on the FP64 true residual and against
  a 1e-13 FP64 solve.
* Oseen P1-P1 PSPG 256^2: assembly, block-multicolour order and ILU(1) factor
  bit for bit; FP64 and FAC32 BiCGSTAB to 1e-9 on the true residual and
  against a 1e-12 FP64 solve.
* 512 Helmholtz systems in one persistent launch: assembly bit for bit; every
  system's FP64 true residual recomputed on the host for the FP64 and the
  mixed batch; colour orders, factors and FP64 iterations of sampled systems
  (smallest, largest and four more) against the oracle.

Error bounds against the tight FP64 solves: 1e-6 (the FP64 parity bound) for
the symmetric problems; 1e-5 for the Oseen system, whose condition number is
about 3e8, so that a 1e-9 residual fixes the solution only to that order.
"""
import numpy as np
import pytest

from delfem_amd import meshgen
from oracle.oracle import assemble_problem

pytestmark = pytest.mark.gpu

SEED = 12345


def _case(name):
    print(f"[case] {name}", flush=True)


def true_relres(A, x, rhs_np):
    import torch
    y = torch.from_numpy(rhs_np.copy()).cuda()
    A.MatVec(-1.0, x, 1.0, y)
    return float(torch.linalg.norm(y).item() / np.linalg.norm(rhs_np))


def relerr(x, x_ref):
    x = x.cpu().numpy() if hasattr(x, "cpu") else x
    x_ref = x_ref.cpu().numpy() if hasattr(x_ref, "cpu") else x_ref
    return float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))


def device_solve(gpu, s, M, rhs_np, kind, tol, maxit, precision="fp64"):
    import torch
    r = torch.from_numpy(rhs_np.copy()).cuda()
    x = torch.zeros_like(r)
    fn = gpu.Solve_PCG if kind == "pcg" else gpu.Solve_PBiCGSTAB
    ok, conv, it, info = fn(tol, maxit, s.A, M, r, x, precision=precision)
    return ok, conv, it, info, x


def check_assembly(s, ref):
    val, dia = s.A.values()
    assert np.array_equal(val, ref["val"]), "assembled off-diagonal blocks differ"
    assert np.array_equal(dia, ref["dia"]), "assembled diagonal blocks differ"
    assert np.array_equal(s.res.cpu().numpy(), ref["res"]), "assembled residual differs"


@pytest.mark.parametrize("name", ["cfg1", "cfg2"])
def test_full_size_against_oracle(orc, gpu, name):
    _case(f"{name} full size")
    if name == "cfg1":
        prob, tol = meshgen.poisson2d(128, seed=SEED, sigma=0.5), 1e-8
    else:
        prob, tol = meshgen.solid3d(48, seed=SEED, sigma=0.0), 1e-10
    b, nb = prob.b, prob.nnode
    ref = assemble_problem(orc, prob)
    s = gpu.FEMSystem(prob)
    s.assemble()
    check_assembly(s, ref)
    # the reference's configuration: natural-order ILU(0)-PCG
    Mn = gpu.CPreconditioner_ILU(s.A, 0)
    assert Mn.SetValue(s.A)
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], 0)
    fv, fd, ns = orc.ilu_values(nb, b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F)
    gv, gd = Mn.values()
    assert ns == 0 and np.array_equal(gv, fv) and np.array_equal(gd, fd), "natural ILU(0) factor differs"
    ok_o, conv_o, it_o, x_o, _ = orc.solve("pcg", nb, b, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, fv, fd,
                                           ref["res"], tol, 5000)
    ok, conv, it, _, x = device_solve(gpu, s, Mn, ref["res"], "pcg", tol, 5000)
    assert ok and ok_o and abs(it - it_o) <= 1, (it, it_o)
    assert relerr(x, x_o) <= 1e-8
    assert true_relres(s.A, x, ref["res"]) <= 2 * tol
    # the benchmark's configuration: multicolour ordering, FP64 and mixed
    Mc = gpu.CPreconditioner_ILU(s.A, 0, ordering="multicolour")
    assert Mc.SetValue(s.A)
    for prec in ("fp64", "mixed_f32_bf16"):
        okc, _, _, info, xc = device_solve(gpu, s, Mc, ref["res"], "pcg", tol, 20000, prec)
        rr = true_relres(s.A, xc, ref["res"])
        assert okc and rr <= (2 if prec == "fp64" else 1.01) * tol, (prec, rr)
        assert relerr(xc, x_o) <= 1e-6, prec


@pytest.fixture(scope="module")
def elasticity96(gpu):
    _case("cfg3 setup")
    prob = meshgen.solid3d(96, seed=SEED, sigma=1.0)
    s = gpu.FEMSystem(prob)
    s.assemble()
    M = gpu.CPreconditioner_ILU(s.A, 0, ordering="multicolour")
    assert M.SetValue(s.A)
    return prob, s, M


def test_cfg3_factor_bitwise(orc, elasticity96):
    """The full multicolour ILU(0) factor of the 2.7 M DOF system."""
    _case("cfg3 factor")
    prob, s, M = elasticity96
    rowptr, col = s.A.pattern()
    val, dia = s.A.values()
    P = M.pattern()
    _, nc, order = orc.colour_order(rowptr, col)
    assert nc == 4 and np.array_equal(P["order"], order)
    F = orc.iluk_symbolic(rowptr, col, 0, order)
    assert np.array_equal(P["rowptr"], F["rowptr"]) and np.array_equal(P["col"], F["col"])
    fv, fd, ns = orc.ilu_values(prob.nnode, 3, rowptr, col, val, dia, F, order)
    gv, gd = M.values()
    assert ns == 0
    assert np.array_equal(gd, fd), "diagonal factor blocks differ"
    assert np.array_equal(gv, fv), "off-diagonal factor blocks differ"


def test_cfg3_solves(gpu, elasticity96):
    _case("cfg3 solves")
    prob, s, M = elasticity96
    rhs = s.res.cpu().numpy()
    ok_t, _, _, _, x_t = device_solve(gpu, s, M, rhs, "pcg", 1e-13, 20000)
    assert ok_t
    ok, conv, it, _, x = device_solve(gpu, s, M, rhs, "pcg", 1e-10, 20000)
    assert ok and conv < 1e-10
    assert true_relres(s.A, x, rhs) <= 2e-10
    assert relerr(x, x_t) <= 1e-6
    okm, _, _, info, xm = device_solve(gpu, s, M, rhs, "pcg", 1e-10, 20000, "mixed_f32_bf16")
    assert okm and info["outer_iters"] >= 1
    rr = true_relres(s.A, xm, rhs)
    assert rr <= 1.01e-10 and abs(rr - info["true_relres"]) <= 1e-2 * rr
    assert relerr(xm, x_t) <= 1e-6


def test_cfg4_oseen_full_size(orc, gpu):
    _case("cfg4 full size")
    prob = meshgen.oseen2d(256, seed=SEED)
    ref = assemble_problem(orc, prob)
    s = gpu.FEMSystem(prob)
    s.assemble()
    check_assembly(s, ref)
    M = gpu.CPreconditioner_ILU(s.A, 1, ordering="block_multicolour", block_nodes=16)
    assert M.SetValue(s.A)
    P = M.pattern()
    _, order = orc.bmc_order(ref["rowptr"], ref["col"], 16)[1:]
    assert np.array_equal(P["order"], order)
    F = orc.iluk_symbolic(ref["rowptr"], ref["col"], 1, order)
    assert np.array_equal(P["rowptr"], F["rowptr"]) and np.array_equal(P["col"], F["col"])
    fv, fd, ns = orc.ilu_values(prob.nnode, 3, ref["rowptr"], ref["col"], ref["val"], ref["dia"], F, order)
    gv, gd = M.values()
    assert ns == 0 and np.array_equal(gv, fv) and np.array_equal(gd, fd), "ILU(1) factor differs"
    rhs = ref["res"]
    _, _, _, info_t, x_t = device_solve(gpu, s, M, rhs, "bicg", 1e-12, 20000)
    assert info_t["converged"] and true_relres(s.A, x_t, rhs) <= 1e-10
    for prec in ("fp64", "fac32"):
        ok, conv, it, info, x = device_solve(gpu, s, M, rhs, "bicg", 1e-9, 20000, prec)
        rr = true_relres(s.A, x, rhs)
        assert ok and info["converged"] and rr <= 2e-9, (prec, rr)
        assert relerr(x, x_t) <= 1e-5, prec


def test_cfg5_batch_full_size(orc, gpu):
    import torch
    _case("cfg5 full size")
    parts = meshgen.helmholtz_batch(512, seed=SEED)
    prob = meshgen.concat(parts)
    ref = assemble_problem(orc, prob)
    fem = gpu.FEMSystem(prob)
    fem.assemble()
    check_assembly(fem, ref)
    B = gpu.BatchSystem(fem, prob.sys_ptr)
    B.SetValue()
    rhs = fem.res.clone()
    sp = prob.sys_ptr.astype(np.int64)
    rp = ref["rowptr"].astype(np.int64)
    out = {}
    for prec in ("fp64", "mixed_f32"):
        x = torch.zeros_like(rhs)
        ok, o = B.solve(rhs, x, 1e-8, 20000, precision=prec)
        assert ok and np.all(o["status"] == 1), prec
        out[prec] = (x.cpu().numpy(), o)
    y = orc.matvec(prob.nnode, 1, ref["rowptr"], ref["col"], ref["val"], ref["dia"], -1.0, out["mixed_f32"][0], 1.0,
                   ref["res"])
    yd = orc.matvec(prob.nnode, 1, ref["rowptr"], ref["col"], ref["val"], ref["dia"], -1.0, out["fp64"][0], 1.0,
                    ref["res"])
    for s in range(len(parts)):
        r0, r1 = sp[s], sp[s + 1]
        nr = np.linalg.norm(ref["res"][r0:r1])
        assert np.linalg.norm(y[r0:r1]) <= 1e-8 * nr, ("mixed", s)
        assert np.linalg.norm(yd[r0:r1]) <= 2e-8 * nr, ("fp64", s)
    nn = np.diff(sp)
    sample = sorted({int(np.argmin(nn)), int(np.argmax(nn))} | set(np.random.default_rng(1).choice(len(parts), 4,
                                                                                              replace=False).tolist()))
    xd = out["fp64"][0]
    for s in sample:
        r0, r1 = int(sp[s]), int(sp[s + 1])
        q0, q1 = rp[r0], rp[r1]
        S = dict(rowptr=(rp[r0:r1 + 1] - q0).astype(np.uint32), col=(ref["col"][q0:q1] - r0).astype(np.uint32),
                 val=ref["val"][q0:q1], dia=ref["dia"][r0:r1], res=ref["res"][r0:r1])
        n = r1 - r0
        _, _, order = orc.colour_order(S["rowptr"], S["col"])
        fval, fdia, gorder = B.factor(s)
        assert np.array_equal(gorder, order), s
        F = orc.iluk_symbolic(S["rowptr"], S["col"], 0, order)
        ov, od, ns = orc.ilu_values(n, 1, S["rowptr"], S["col"], S["val"], S["dia"], F, order)
        assert ns == 0 and np.array_equal(fval, ov) and np.array_equal(fdia, od), s
        ok_o, _, it_o, x_o, _ = orc.solve("pcg", n, 1, S["rowptr"], S["col"], S["val"], S["dia"], F, ov, od, S["res"],
                                          1e-8, 20000, order)
        assert ok_o and abs(int(out["fp64"][1]["iters"][s]) - it_o) <= 1, (s, out["fp64"][1]["iters"][s], it_o)
        assert np.linalg.norm(xd[r0:r1] - x_o) <= 1e-8 * np.linalg.norm(x_o), s
