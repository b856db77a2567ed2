"""The oracle (C restatement) against the golden fixtures written by the
compiled DelFEM reference (tests/golden/make_golden.py).  Everything is
compared bitwise: integer patterns, FP64 element matrices, assembled values,
ILU factors, substitutions, MatVec and whole PCG / BiCGSTAB runs (iteration
count, conv_ratio, solution and final residual)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = ["poisson2d_8", "solid3d_3", "graph60_b2"]


def load(name):
    return np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)


def orders_of(d):
    out = {"natural": None}
    for k in d.files:
        if k.startswith("order_"):
            out[k[len("order_"):]] = d[k]
    return out


def system(d):
    return int(d["nnode"]), int(d["b"]), d["rowptr"], d["col"], d["val"], d["dia"]


@pytest.mark.parametrize("name", ["poisson2d_8", "solid3d_3"])
def test_element_kernels(orc, name):
    d = load(name)
    ne = int(d["nemat"])
    conn = d["conn"][:ne]
    if str(d["kind"]) == "poisson2d":
        emat, eres = orc.emat_poisson2d(d["coord"], conn, d["alpha"][:ne], float(d["source"]), np.zeros(int(d["nnode"])))
    else:
        emat, eres = orc.emat_solid3d(d["coord"], conn, d["lam"][:ne], d["mu"][:ne], float(d["rho"]), tuple(d["g"]),
                                      np.zeros(int(d["nnode"]) * 3))
    assert np.array_equal(emat, d["emat"])
    assert np.array_equal(eres, d["eres"])


@pytest.mark.parametrize("name", ["poisson2d_8", "solid3d_3"])
def test_pattern_assembly_bc(orc, name):
    d = load(name)
    nb, b = int(d["nnode"]), int(d["b"])
    rowptr, col = orc.fem_pattern(nb, d["conn"])
    assert np.array_equal(rowptr, d["rowptr"]) and np.array_equal(col, d["col"])
    if str(d["kind"]) == "poisson2d":
        emat, eres = orc.emat_poisson2d(d["coord"], d["conn"], d["alpha"], float(d["source"]), np.zeros(nb))
    else:
        emat, eres = orc.emat_solid3d(d["coord"], d["conn"], d["lam"], d["mu"], float(d["rho"]), tuple(d["g"]),
                                      np.zeros(nb * 3))
    val, dia = orc.assemble(nb, b, rowptr, col, d["conn"], emat)
    res = orc.assemble_vec(nb, b, d["conn"], eres)
    val, dia = orc.set_bc(nb, b, rowptr, col, val, dia, d["flags"])
    res = orc.zero_bc(d["flags"], res)
    assert np.array_equal(val, d["val"])
    assert np.array_equal(dia, d["dia"])
    assert np.array_equal(res, d["res"])
    assert float(np.sqrt(np.sum(res * res))) == pytest.approx(float(d["norm_res"]), rel=1e-15)


@pytest.mark.parametrize("name", FIXTURES)
def test_matvec(orc, name):
    d = load(name)
    nb, b, rowptr, col, val, dia = system(d)
    y = orc.matvec(nb, b, rowptr, col, val, dia, float(d["mv_alpha"]), d["mv_x"], float(d["mv_beta"]), d["mv_y"])
    assert np.array_equal(y, d["mv_out"])


@pytest.mark.parametrize("name", FIXTURES)
def test_iluk_patterns(orc, name):
    d = load(name)
    nb, b, rowptr, col, val, dia = system(d)
    n_checked = 0
    for oname, order in orders_of(d).items():
        for lev in range(4):
            tag = f"{oname}_k{lev}"
            if f"frow_{tag}" not in d.files:
                continue
            F = orc.iluk_symbolic(rowptr, col, lev, order)
            assert np.array_equal(F["rowptr"], d[f"frow_{tag}"]), tag
            assert np.array_equal(F["col"], d[f"fcol_{tag}"]), tag
            assert np.array_equal(F["diaind"], d[f"fdia_{tag}"]), tag
            n_checked += 1
    assert n_checked >= 8


@pytest.mark.parametrize("name", FIXTURES)
def test_ilu_values_and_substitutions(orc, name):
    d = load(name)
    nb, b, rowptr, col, val, dia = system(d)
    n_checked = 0
    for oname, order in orders_of(d).items():
        for lev in range(2):
            tag = f"{oname}_k{lev}"
            if f"fval_{tag}" not in d.files:
                continue
            F = orc.iluk_symbolic(rowptr, col, lev, order)
            fv, fd, ns = orc.ilu_values(nb, b, rowptr, col, val, dia, F, order)
            assert bool(d[f"fok_{tag}"]) == (ns >= 0)
            assert np.array_equal(fv, d[f"fval_{tag}"]), tag
            assert np.array_equal(fd, d[f"fdiav_{tag}"]), tag
            assert np.array_equal(orc.forward(nb, b, F, fv, fd, d["vin"]), d[f"fwd_{tag}"]), tag
            assert np.array_equal(orc.backward(nb, b, F, fv, fd, d["vin"]), d[f"bwd_{tag}"]), tag
            n_checked += 1
    assert n_checked >= 3


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("kind", ["pcg", "bicg"])
def test_solvers(orc, name, kind):
    d = load(name)
    nb, b, rowptr, col, val, dia = system(d)
    n_checked = 0
    for oname, order in orders_of(d).items():
        for lev in range(2):
            tag = f"{oname}_k{lev}"
            if f"{kind}_{tag}" not in d.files:
                continue
            F = orc.iluk_symbolic(rowptr, col, lev, order)
            fv, fd, _ = orc.ilu_values(nb, b, rowptr, col, val, dia, F, order)
            ok, conv, it, x, r = orc.solve(kind, nb, b, rowptr, col, val, dia, F, fv, fd, d["res"], 1e-10, 2000,
                                           order)
            want = d[f"{kind}_{tag}"]
            assert (float(ok), conv, float(it)) == (want[0], want[1], want[2]), tag
            assert np.array_equal(x, d[f"{kind}_x_{tag}"]), tag
            assert np.array_equal(r, d[f"{kind}_r_{tag}"]), tag
            n_checked += 1
    assert n_checked >= 3


@pytest.mark.parametrize("name", ["poisson2d_8", "solid3d_3"])
@pytest.mark.parametrize("kind", ["cg", "pcg", "bicgstab", "pbicgstab"])
def test_matrix_layer_solvers(orc, name, kind):
    """MatVec::Sol::Solve_* with the natural ILU(0) factor: flag, iterations,
    conv_ratio, solution and final residual bit for bit."""
    d = load(name)
    nb, b, rowptr, col, val, dia = system(d)
    F = orc.iluk_symbolic(rowptr, col, 0)
    fv, fd, _ = orc.ilu_values(nb, b, rowptr, col, val, dia, F)
    ok, conv, it, u, r = orc.solve_mat(kind, nb, b, rowptr, col, val, dia, d["res"], 1e-10, 2000, F, fv, fd)
    want = d[f"mat_{kind}"]
    assert (float(ok), conv, float(it)) == (want[0], want[1], want[2])
    assert np.array_equal(u, d[f"mat_{kind}_u"]) and np.array_equal(r, d[f"mat_{kind}_r"])


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_unstructured_mesh(orc):
    """The reference's own tetrahedral mesh (747 nodes, 3,114 elements,
    E = 200, nu = 0.43): element matrices, pattern, assembled system, natural
    ILU(0) factor and the PCG / BiCGSTAB runs bit for bit (large arrays by
    SHA-256 digest)."""
    d = load("cylinder_tet")
    nb = int(d["nnode"])
    conn = d["conn"]
    emat, eres = orc.emat_solid3d(d["coord"], conn, d["lam"], d["mu"], float(d["rho"]), tuple(d["g"]),
                                  np.zeros(nb * 3))
    ne = int(d["nemat"])
    assert sha(emat[:ne * 144]) == str(d["sha_emat"]) and sha(eres[:ne * 12]) == str(d["sha_eres"])
    rowptr, col = orc.fem_pattern(nb, conn)
    assert sha(rowptr) == str(d["sha_rowptr"]) and sha(col) == str(d["sha_col"])
    val, dia = orc.assemble(nb, 3, rowptr, col, conn, emat)
    res = orc.assemble_vec(nb, 3, conn, eres)
    val, dia = orc.set_bc(nb, 3, rowptr, col, val, dia, d["flags"])
    res = orc.zero_bc(d["flags"], res)
    assert sha(val) == str(d["sha_val"]) and sha(dia) == str(d["sha_dia"])
    assert np.array_equal(res, d["res"])
    F = orc.iluk_symbolic(rowptr, col, 0)
    fv, fd, _ = orc.ilu_values(nb, 3, rowptr, col, val, dia, F)
    assert sha(fv) == str(d["sha_fval_natural_k0"]) and sha(fd) == str(d["sha_fdiav_natural_k0"])
    for kind in ("pcg", "bicg"):
        ok, conv, it, x, r = orc.solve(kind, nb, 3, rowptr, col, val, dia, F, fv, fd, res, 1e-10, 3000)
        want = d[f"{kind}_natural_k0"]
        assert (float(ok), conv, float(it)) == (want[0], want[1], want[2]), kind
        assert np.array_equal(x, d[f"{kind}_x_natural_k0"]) and sha(r) == str(d[f"sha_{kind}_r_natural_k0"])


def test_complex_twin(orc):
    """Complex twin against the reference's outputs: assembly from the stored
    element matrices' problem, MatVec, ILU(0)/ILU(1) pattern and factor, one
    preconditioner apply and Solve_PCOCG, bit for bit."""
    from delfem_amd import meshgen
    from oracle.oracle import assemble_zproblem
    d = load("zhelmholtz2d_10x8")
    prob = meshgen.zhelmholtz2d(10, 8, 1.0, 0.8, wave_length=float(d["wave_length"]))
    assert np.array_equal(prob.coord, d["coord"]) and np.array_equal(prob.lconn, d["lconn"])
    Z = assemble_zproblem(orc, prob)
    assert np.array_equal(Z["emat_tri"][:9 * 16], d["emat_tri"]) and np.array_equal(Z["emat_line"][:32], d["emat_line"])
    assert np.array_equal(Z["val"], d["zval"]) and np.array_equal(Z["dia"], d["zdia"])
    assert np.array_equal(Z["res"], d["zres"])
    nb = int(d["nnode"])
    assert np.array_equal(orc.zmatvec(nb, Z["rowptr"], Z["col"], Z["val"], Z["dia"], 1.5, d["zmv_x"], -0.25,
                                      d["zmv_x"]), d["zmv_out"])
    for lev in (0, 1):
        F = orc.iluk_symbolic(Z["rowptr"], Z["col"], lev)
        assert np.array_equal(F["rowptr"], d[f"zfrow_k{lev}"]) and np.array_equal(F["col"], d[f"zfcol_k{lev}"])
        fv, fd, _ = orc.zilu_values(nb, Z["rowptr"], Z["col"], Z["val"], Z["dia"], F)
        assert np.array_equal(fv, d[f"zfval_k{lev}"]) and np.array_equal(fd, d[f"zfdia_k{lev}"])
        assert np.array_equal(orc.zilu_solve(nb, F, fv, fd, d["zmv_x"]), d[f"zprec_k{lev}"])
        ok, conv, it, u, r = orc.zsolve_pcocg(nb, Z["rowptr"], Z["col"], Z["val"], Z["dia"], F, fv, fd, Z["res"],
                                              1e-10, 2000)
        want = d[f"zpcocg_k{lev}"]
        assert (float(ok), conv, float(it)) == (want[0], want[1], want[2])
        assert np.array_equal(u, d[f"zpcocg_u_k{lev}"]) and np.array_equal(r, d[f"zpcocg_r_k{lev}"])


def test_complex_solvers(orc):
    """MatVec::Sol::Solve_COCG, Solve_BiCGSTAB and Solve_BiCGStabP against the
    reference's runs, bit for bit: on the complex Helmholtz system (COCG to
    convergence, the BiCGSTAB pair for their first iterations, which is all
    they do there: the reference conjugates alpha and omega) and on a
    real-valued system in complex storage, where all four solvers converge."""
    from delfem_amd import meshgen
    from oracle.oracle import assemble_zproblem
    d = load("zhelmholtz2d_10x8")
    prob = meshgen.zhelmholtz2d(10, 8, 1.0, 0.8, wave_length=float(d["wave_length"]))
    Z = assemble_zproblem(orc, prob)
    nb = int(d["nnode"])
    args = (nb, Z["rowptr"], Z["col"], Z["val"], Z["dia"])

    def same(got, key, ukey=None, rkey=None):
        want = d[key]
        assert (float(got[0]), got[1], float(got[2])) == (want[0], want[1], want[2]), key
        assert np.array_equal(got[3], d[ukey or key + "_u"]) and np.array_equal(got[4], d[rkey or key + "_r"]), key

    same(orc.zsolve_cocg(*args, Z["res"], 1e-10, 2000), "zcocg")
    same(orc.zsolve_bicgstab(*args, Z["res"], 1e-10, 8), "zbicg")
    for lev in (0, 1):
        F = orc.iluk_symbolic(Z["rowptr"], Z["col"], lev)
        fv, fd, _ = orc.zilu_values(*args, F)
        same(orc.zsolve_bicgstab(*args, Z["res"], 1e-10, 8, F, fv, fd), f"zbicgp_k{lev}", f"zbicgp_u_k{lev}",
             f"zbicgp_r_k{lev}")
    rargs = (nb, Z["rowptr"], Z["col"], d["rval"], d["rdia"])
    F = orc.iluk_symbolic(Z["rowptr"], Z["col"], 0)
    fv, fd, _ = orc.zilu_values(*rargs, F)
    same(orc.zsolve_pcocg(*rargs, F, fv, fd, d["rres"], 1e-10, 2000), "rpcocg")
    same(orc.zsolve_cocg(*rargs, d["rres"], 1e-10, 2000), "rcocg")
    same(orc.zsolve_bicgstab(*rargs, d["rres"], 1e-10, 2000), "rbicg")
    same(orc.zsolve_bicgstab(*rargs, d["rres"], 1e-10, 2000, F, fv, fd), "rbicgp")
