"""Writes the golden fixtures of tests/golden/*.npz from the compiled DelFEM
reference (oracle/_ref/libdfmref.so, built by oracle/Makefile from the DelFEM
tree).  Run in a container where that tree exists:

    python tests/golden/make_golden.py

Each fixture holds inputs (mesh or graph, coefficients, vectors, orderings)
and the reference's outputs for them: element matrices, assembled block-CRS
matrix and residual after Dirichlet conditions, ILU(k) patterns (natural and
ordered, k = 0..3), ILU factor values, one MatVec, one forward/backward
substitution, the Solve_PCG / Solve_PBiCGSTAB results (flag, iterations,
conv_ratio, solution, final residual), and the same for the matrix-layer
MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB with the
natural ILU(0) factor; the reference's unstructured tetrahedral mesh; and the
complex twin (CZMatDia_BlkCrs, CZMatDiaFrac_BlkCrs, Solve_PCOCG).  Floating-point arrays are stored as
float64 and compared bitwise by the tests.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from delfem_amd import meshgen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def fem_fixture(ref, prob, lev_list=(0, 1, 2, 3), tol=1e-10):
    out = {}
    # element matrices through the reference's element kernels
    import ctypes as C
    conn = O.u32(prob.conn)
    ne = conn.shape[0]
    if prob.kind == "poisson2d":
        emat = np.zeros(ne * 9)
        eres = np.zeros(ne * 3)
        ref.L.ref_emat_poisson2d_p1(ne, O._p(O.f64(prob.coord)), O._p(conn), O._p(O.f64(prob.alpha)), prob.source,
                                    O._p(np.zeros(prob.nnode)), O._p(emat), O._p(eres))
        out.update(alpha=prob.alpha, source=np.array(prob.source))
    else:
        emat = np.zeros(ne * 144)
        eres = np.zeros(ne * 12)
        ref.L.ref_emat_solid3d_p1(ne, O._p(O.f64(prob.coord)), O._p(conn), O._p(O.f64(prob.lam)),
                                  O._p(O.f64(prob.mu)), prob.rho, prob.g[0], prob.g[1], prob.g[2],
                                  O._p(np.zeros(prob.nnode * 3)), O._p(emat), O._p(eres))
        out.update(lam=prob.lam, mu=prob.mu, rho=np.array(prob.rho), g=np.array(prob.g))
    # element matrices of the first elements pin the element kernels; the
    # assembled matrix below depends on all of them
    nkeep = min(ne, 24)
    per = emat.size // ne
    out.update(kind=np.array(prob.kind), b=np.array(prob.b), nnode=np.array(prob.nnode), coord=prob.coord,
               conn=prob.conn, flags=prob.flags, emat=emat[:nkeep * per], eres=eres[:nkeep * (eres.size // ne)],
               nemat=np.array(nkeep))
    # pattern: brute force (the reference builds it inside femfield, which is not compiled)
    orc = O.Orc()
    rowptr, col = orc.fem_pattern(prob.nnode, prob.conn)
    rs = O.RefSystem(ref, prob.nnode, prob.b, rowptr, col)
    rs.merge(prob.conn, emat, eres)
    rs.set_bc(prob.flags)
    nres = rs.finalize()
    rp, cl, val, dia = rs.matrix()
    res = rs.get_vector(-1)
    out.update(rowptr=rp, col=cl, val=val, dia=dia, res=res, norm_res=np.array(nres))
    return out, rs, rp, cl, val, dia, res


def ilu_fixture(out, rs, nb, b, res, orders, lev_list, tol, solve=True, max_value_lev=1):
    rng = np.random.default_rng(99)
    v = rng.standard_normal(nb * b)
    out["vin"] = v
    for oname, order in orders.items():
        if order is not None:
            out[f"order_{oname}"] = order
        for lev in lev_list:
            rf = O.RefFrac(rs, lev, order)
            P = rf.pattern()
            tag = f"{oname}_k{lev}"
            out[f"frow_{tag}"] = P["rowptr"]
            out[f"fcol_{tag}"] = P["col"]
            out[f"fdia_{tag}"] = P["diaind"]
            if lev <= (0 if oname == "random" else max_value_lev):
                ok = rf.set_value()
                fv, fd = rf.values()
                out[f"fval_{tag}"] = fv
                out[f"fdiav_{tag}"] = fd
                out[f"fok_{tag}"] = np.array(ok)
                out[f"fwd_{tag}"] = rf.forward(v)
                out[f"bwd_{tag}"] = rf.backward(v)
                if solve:
                    for kind in ("pcg", "bicg"):
                        rs.set_vector(-1, res)
                        ok2, conv, it = rf.solve(kind, tol, 2000)
                        out[f"{kind}_{tag}"] = np.array([float(ok2), conv, float(it)])
                        out[f"{kind}_x_{tag}"] = rs.get_vector(-2)
                        out[f"{kind}_r_{tag}"] = rs.get_vector(-1)
    rs.set_vector(-1, res)


def matvec_fixture(out, rs, nb, b):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(nb * b)
    y = rng.standard_normal(nb * b)
    rs.resize_tmp(2)
    rs.set_vector(0, x)
    rs.set_vector(1, y)
    rs.matvec(-0.5, 0, 2.0, 1)
    out.update(mv_x=x, mv_y=y, mv_alpha=np.array(-0.5), mv_beta=np.array(2.0), mv_out=rs.get_vector(1))


def mat_solver_fixture(out, rs, res, tol=1e-10, maxit=2000):
    """MatVec::Sol solvers on the assembled matrix and residual."""
    rf = O.RefFrac(rs, 0)
    rf.set_value()
    for kind in ("cg", "pcg", "bicgstab", "pbicgstab"):
        ok, conv, it, u, r = rs.solve_mat(kind, res, tol, maxit, rf if kind.startswith("p") else None)
        out[f"mat_{kind}"] = np.array([float(ok), conv, float(it)])
        out[f"mat_{kind}_u"] = u
        out[f"mat_{kind}_r"] = r


MSH = "/root/reference/test_glut/input_file/cylinder_tet.msh"


def gid_mesh_fixture(ref):
    """The reference's unstructured tetrahedral mesh (its solid3d demo input),
    elasticity with E = 200, nu = 0.43 as in that demo: mesh arrays, the
    reference's assembled system, natural ILU(0) and Solve_PCG results."""
    coord, conn = meshgen.read_gid_msh(MSH)
    prob = meshgen.solid3d_mesh(coord, conn, name="cylinder_tet")
    out, rs, rp, cl, val, dia, res = fem_fixture(ref, prob)
    rf = O.RefFrac(rs, 0)
    ok = rf.set_value()
    fv, fd = rf.values()
    out.update(fval_natural_k0=fv, fdiav_natural_k0=fd, fok_natural_k0=np.array(ok))
    for kind in ("pcg", "bicg"):
        rs.set_vector(-1, res)
        ok2, conv, it = rf.solve(kind, 1e-10, 3000)
        out[f"{kind}_natural_k0"] = np.array([float(ok2), conv, float(it)])
        out[f"{kind}_x_natural_k0"] = rs.get_vector(-2)
        out[f"{kind}_r_natural_k0"] = rs.get_vector(-1)
    rs.set_vector(-1, res)
    # the large outputs are kept as SHA-256 digests of their bytes (bitwise
    # comparison at a fraction of the size)
    for k in ("rowptr", "col", "val", "dia", "emat", "eres", "fval_natural_k0", "fdiav_natural_k0",
              "pcg_r_natural_k0", "bicg_r_natural_k0"):
        out[f"sha_{k}"] = np.array(digest(out.pop(k)))
    return out


def zhelmholtz_fixture(ref):
    """Complex twin: the reference's CZMatDia_BlkCrs assembled from the
    oracle's element matrices (eqn_helmholtz.cpp is not compiled: it needs the
    field layer), its ILU(0)/ILU(1) factors, one MatVec and one
    preconditioner apply, and MatVec::Sol::Solve_PCOCG."""
    orc = O.Orc()
    prob = meshgen.zhelmholtz2d(10, 8, 1.0, 0.8, wave_length=0.3)
    Z = O.assemble_zproblem(orc, prob)
    rz = O.RefZSystem(ref, prob.nnode, Z["rowptr"], Z["col"])
    rz.merge(prob.conn, Z["emat_tri"])
    rz.merge(prob.lconn, Z["emat_line"])
    rz.set_bc(prob.flags)
    val, dia = rz.matrix()
    out = dict(coord=prob.coord, conn=prob.conn, lconn=prob.lconn, wave_length=np.array(prob.wave_length),
               flags=prob.flags, value=prob.value, nnode=np.array(prob.nnode), rowptr=Z["rowptr"], col=Z["col"],
               emat_tri=Z["emat_tri"][:9 * 16], emat_line=Z["emat_line"][:4 * 8], zval=val, zdia=dia, zres=Z["res"])
    rng = np.random.default_rng(17)
    x = rng.standard_normal(prob.nnode) + 1j * rng.standard_normal(prob.nnode)
    out.update(zmv_x=x, zmv_out=rz.matvec(1.5, x, -0.25, x))
    for lev in (0, 1):
        P, fv, fd, ok = rz.factor(lev)
        out[f"zfrow_k{lev}"], out[f"zfcol_k{lev}"] = P["rowptr"], P["col"]
        out[f"zfval_k{lev}"], out[f"zfdia_k{lev}"] = fv, fd
        out[f"zprec_k{lev}"] = rz.precond(x)
        ok2, conv, it, u, r = rz.pcocg(Z["res"], 1e-10, 2000)
        out[f"zpcocg_k{lev}"] = np.array([float(ok2), conv, float(it)])
        out[f"zpcocg_u_k{lev}"], out[f"zpcocg_r_k{lev}"] = u, r
        # BiCGStabP with this factor: the first iterations (the reference's
        # complex BiCGSTAB conjugates alpha and omega and does not converge on
        # this complex system)
        ok2, conv, it, u, r = rz.bicgstab(Z["res"], 1e-10, 8, True)
        out[f"zbicgp_k{lev}"] = np.array([float(ok2), conv, float(it)])
        out[f"zbicgp_u_k{lev}"], out[f"zbicgp_r_k{lev}"] = u, r
    ok2, conv, it, u, r = rz.cocg(Z["res"], 1e-10, 2000)
    out["zcocg"] = np.array([float(ok2), conv, float(it)])
    out["zcocg_u"], out["zcocg_r"] = u, r
    ok2, conv, it, u, r = rz.bicgstab(Z["res"], 1e-10, 8, False)
    out["zbicg"] = np.array([float(ok2), conv, float(it)])
    out["zbicg_u"], out["zbicg_r"] = u, r
    # a real-valued system in complex storage (triangles only, long wave):
    # there all four solvers converge
    prob2 = meshgen.zhelmholtz2d(10, 8, 1.0, 0.8, wave_length=2.0)
    et, rt = orc.zemat_helmholtz2d(prob2.coord, prob2.conn, prob2.wave_length, prob2.value)
    rz2 = O.RefZSystem(ref, prob2.nnode, Z["rowptr"], Z["col"])
    rz2.merge(prob2.conn, et)
    rz2.set_bc(prob2.flags)
    res2 = orc.zassemble_vec(np.zeros(prob2.nnode, dtype=np.complex128), prob2.conn, rt)
    res2[prob2.flags == 1] = 0.0
    val2, dia2 = rz2.matrix()
    out.update(rval=val2, rdia=dia2, rres=res2)
    rz2.factor(0)
    for name, run in (("pcocg", lambda: rz2.pcocg(res2, 1e-10, 2000)), ("cocg", lambda: rz2.cocg(res2, 1e-10, 2000)),
                      ("bicg", lambda: rz2.bicgstab(res2, 1e-10, 2000, False)),
                      ("bicgp", lambda: rz2.bicgstab(res2, 1e-10, 2000, True))):
        ok2, conv, it, u, r = run()
        out[f"r{name}"] = np.array([float(ok2), conv, float(it)])
        out[f"r{name}_u"], out[f"r{name}_r"] = u, r
    return out


def digest(a):
    """SHA-256 of an array's bytes (little-endian, C order)."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def random_graph_fixture(ref, n=60, b=2, seed=7):
    """A random symmetric graph with 2x2 blocks: diagonally dominant values."""
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for i in range(n):
        for _ in range(3):
            j = int(rng.integers(n))
            if j != i:
                adj[i].add(j)
                adj[j].add(i)
    rowptr = np.zeros(n + 1, dtype=np.uint32)
    for i in range(n):
        rowptr[i + 1] = rowptr[i] + len(adj[i])
    col = np.array([j for i in range(n) for j in sorted(adj[i])], dtype=np.uint32)
    bb = b * b
    val = rng.standard_normal(col.size * bb) * 0.3
    dia = np.zeros(n * bb)
    for i in range(n):
        blk = rng.standard_normal((b, b)) * 0.1
        dia[i * bb:(i + 1) * bb] = (blk + np.eye(b) * (2.0 + len(adj[i]))).ravel()
    rs = O.RefSystem(ref, n, b, rowptr, col)
    rs.set_matrix(val, dia)
    res = rng.standard_normal(n * b)
    rs.set_vector(-1, res)
    out = dict(kind=np.array("graph"), b=np.array(b), nnode=np.array(n), rowptr=rowptr, col=col, val=val, dia=dia,
               res=res)
    return out, rs, res


def main():
    ref = O.Ref()
    orc = O.Orc()
    fixtures = {}
    for name, prob in (("poisson2d_8", meshgen.poisson2d(8, seed=3)),
                       ("solid3d_3", meshgen.solid3d(3, seed=3, sigma=0.5))):
        out, rs, rp, cl, val, dia, res = fem_fixture(ref, prob)
        _, _, order = orc.colour_order(rp, cl)
        rng = np.random.default_rng(11)
        orders = {"natural": None, "colour": order, "random": rng.permutation(prob.nnode).astype(np.int32)}
        matvec_fixture(out, rs, prob.nnode, prob.b)
        ilu_fixture(out, rs, prob.nnode, prob.b, res, orders, (0, 1, 2, 3), 1e-10,
                    max_value_lev=1 if prob.b == 1 else 0)
        mat_solver_fixture(out, rs, res)
        fixtures[name] = out
    fixtures["cylinder_tet"] = gid_mesh_fixture(ref)
    fixtures["zhelmholtz2d_10x8"] = zhelmholtz_fixture(ref)
    out, rs, res = random_graph_fixture(ref)
    rng = np.random.default_rng(13)
    orders = {"natural": None, "random": rng.permutation(60).astype(np.int32)}
    matvec_fixture(out, rs, 60, 2)
    ilu_fixture(out, rs, 60, 2, res, orders, (0, 1, 2, 3), 1e-10, solve=True)
    fixtures["graph60_b2"] = out
    for name, out in fixtures.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {os.path.getsize(path) / 1024:.1f} KiB, {len(out)} arrays")


if __name__ == "__main__":
    main()
