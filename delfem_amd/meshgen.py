"""Synthetic meshes and coefficient fields for the benchmark configurations.

Everything here is deterministic and platform independent: the random
numbers come from a counter-based splitmix64 stream (not numpy's generators),
normals from the Box-Muller cosine branch, so the CPU checkers and the device
path see identical inputs for a given seed.

Meshes
  tri_grid(nx, ny, lx, ly)  structured triangulation: cell (i, j) split along
                            the (i, j)-(i+1, j+1) diagonal, nodes row-major,
                            both triangles counter-clockwise (positive area).
  tet_grid(n)               unit cube of n^3 hexahedra, each split into the 6
                            Kuhn tetrahedra around the (0,0,0)-(1,1,1) diagonal,
                            vertex order fixed so every volume is positive.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """count outputs of the splitmix64 stream of `seed`, starting at `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, offset=0):
    """Uniform doubles in (0, 1): (top 53 bits + 0.5) * 2^-53."""
    z = splitmix64(seed, count, offset)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed, count, offset=0):
    """Standard normals, Box-Muller (cosine branch) on consecutive uniform pairs."""
    u = uniform(seed, 2 * count, 2 * offset)
    return np.sqrt(-2.0 * np.log(u[0::2])) * np.cos(2.0 * np.pi * u[1::2])


def lognormal(seed, count, mu, sigma, offset=0):
    return np.exp(mu + sigma * normal(seed, count, offset))


def tri_grid(nx, ny=None, lx=1.0, ly=1.0):
    ny = nx if ny is None else ny
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    coord = np.stack([i.ravel() * (lx / nx), j.ravel() * (ly / ny)], axis=1).astype(np.float64)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    a = (cj * (nx + 1) + ci).ravel()
    b, c, d = a + 1, a + nx + 2, a + nx + 1
    conn = np.empty((a.size * 2, 3), dtype=np.uint32)
    conn[0::2] = np.stack([a, b, c], axis=1)
    conn[1::2] = np.stack([a, c, d], axis=1)
    return np.ascontiguousarray(coord), np.ascontiguousarray(conn)


_KUHN = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def tet_grid(n):
    m = n + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    coord = np.stack([i.ravel() / n, j.ravel() / n, k.ravel() / n], axis=1).astype(np.float64)
    ck, cj, ci = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = np.stack([ci.ravel(), cj.ravel(), ck.ravel()], axis=1)
    ncell = base.shape[0]
    conn = np.empty((ncell, 6, 4), dtype=np.uint32)
    for t, perm in enumerate(_KUHN):
        c = base.copy()
        verts = [c[:, 0] + m * (c[:, 1] + m * c[:, 2])]
        for s in range(3):
            c[:, perm[s]] += 1
            verts.append(c[:, 0] + m * (c[:, 1] + m * c[:, 2]))
        v = np.stack(verts, axis=1)
        # an odd permutation of the axes gives a negatively oriented path: swap v1, v2
        odd = perm in ((0, 2, 1), (1, 0, 2), (2, 1, 0))
        if odd:
            v[:, [1, 2]] = v[:, [2, 1]]
        conn[:, t, :] = v
    return np.ascontiguousarray(coord), np.ascontiguousarray(conn.reshape(-1, 4))


def lame(E, nu):
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


class Problem:
    """A finite-element problem instance: mesh, coefficients, Dirichlet flags."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def ndof(self):
        return self.nnode * self.b


def poisson2d(n=128, seed=12345, sigma=0.5, source=1.0):
    """2-D Poisson (scalar2d path): unit square, n x n cells, lognormal diffusivity
    per element, f = source, u = 0 on x = 0."""
    coord, conn = tri_grid(n)
    nelem = conn.shape[0]
    alpha = lognormal(seed, nelem, 0.0, sigma) if sigma > 0 else np.ones(nelem)
    nnode = coord.shape[0]
    flags = np.zeros(nnode, dtype=np.int32)
    flags[np.isclose(coord[:, 0], 0.0)] = 1
    return Problem(kind="poisson2d", name=f"poisson2d_{n}", n=n, b=1, npe=3, coord=coord, conn=conn,
                   nnode=nnode, nelem=nelem, alpha=alpha, source=source, flags=flags, seed=seed)


def solid3d(n=48, seed=12345, sigma=0.0, E=1.0, nu=0.3, g=(0.0, 0.0, -1.0), rho=1.0):
    """3-D linear elasticity (solid3d path): unit cube, n^3 Kuhn-split hexahedra,
    fixed at z = 0, body force rho*g; E lognormal(0, sigma) per element when sigma > 0."""
    coord, conn = tet_grid(n)
    nelem = conn.shape[0]
    Ee = lognormal(seed, nelem, 0.0, sigma) if sigma > 0 else np.full(nelem, float(E))
    lam, mu = lame(Ee, nu)
    nnode = coord.shape[0]
    flags = np.zeros((nnode, 3), dtype=np.int32)
    flags[np.isclose(coord[:, 2], 0.0), :] = 1
    return Problem(kind="solid3d", name=f"solid3d_{n}", n=n, b=3, npe=4, coord=coord, conn=conn,
                   nnode=nnode, nelem=nelem, lam=np.ascontiguousarray(lam), mu=np.ascontiguousarray(mu),
                   rho=rho, g=tuple(g), flags=flags.ravel(), seed=seed, sigma=sigma)
