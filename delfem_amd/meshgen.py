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


NODAL_FIELDS = ("coord", "flags", "value", "adv")


def relabel(prob, order):
    """The same problem with its nodes renumbered, new node k = old node
    order[k]: nodal arrays permuted, connectivity mapped.  Element data and
    element order are unchanged, so every element matrix is the same."""
    order = np.asarray(order, dtype=np.int64)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    kw = dict(prob.__dict__)
    for k in NODAL_FIELDS:
        if k in kw and kw[k] is not None:
            a = np.asarray(kw[k])            # (nnode, ...) or raveled (nnode * b)
            kw[k] = np.ascontiguousarray(a.reshape(prob.nnode, -1)[order].reshape(a.shape))
    for k in ("conn", "lconn"):
        if k in kw and kw[k] is not None:
            kw[k] = np.ascontiguousarray(inv[kw[k]].astype(np.asarray(kw[k]).dtype))
    kw["name"] = prob.name + "_relabelled"
    return Problem(**kw)


def tile_colour_order(prob, tile_nodes=16):
    """A node numbering for level-scheduled ILU sweeps: the nodes are cut
    into square (cubic) tiles of about tile_nodes nodes by their
    coordinates, the tiles coloured greedily on the tile adjacency of the
    mesh, and numbered by (colour, tile, node).  Tiles of one colour are
    not adjacent, and for tiles of a few nodes a side not at distance two
    either, so ILU(0) and ILU(1) sweeps need about (colours x levels within
    a tile) level sets.  Returns order (order[new] = old node)."""
    coord = np.asarray(prob.coord, dtype=np.float64)
    nn, dim = coord.shape
    conn = np.asarray(prob.conn, dtype=np.int64)
    # mesh spacing: median edge length over the elements' first edges
    h = float(np.median(np.linalg.norm(coord[conn[:, 1]] - coord[conn[:, 0]], axis=1)))
    side = h * max(1.0, round(tile_nodes ** (1.0 / dim)))
    cell = np.floor((coord - coord.min(axis=0)) / side + 1e-9).astype(np.int64)
    dims = cell.max(axis=0) + 1
    tile = np.ravel_multi_index(tuple(cell.T), tuple(dims))
    tiles, tid = np.unique(tile, return_inverse=True)
    nt = tiles.size
    # tile adjacency from the element connectivity
    a = tid[conn]
    pairs = np.concatenate([np.stack([a[:, i], a[:, j]], axis=1)
                            for i in range(a.shape[1]) for j in range(a.shape[1]) if i != j])
    pairs = np.unique(pairs[pairs[:, 0] != pairs[:, 1]], axis=0)
    ptr = np.searchsorted(pairs[:, 0], np.arange(nt + 1))
    colour = -np.ones(nt, dtype=np.int64)
    for t in range(nt):
        used = set(colour[pairs[ptr[t]:ptr[t + 1], 1]].tolist())
        c = 0
        while c in used:
            c += 1
        colour[t] = c
    return np.lexsort((np.arange(nn), tid, colour[tid]))


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


def read_gid_msh(path):
    """Nodes and elements of a GiD ASCII mesh (one MESH block: a header line,
    `Coordinates` ... `end coordinates` with `id x y [z]` lines, `Elements` ...
    `end elements` with `id n1 .. nk [material]` lines; ids 1-based), the
    format Msh::CMesh3D::ReadFromFile_GiDMsh reads (src/msh/mesh3d.cpp:601-700).
    Returns (coord[nnode, dim], conn[nelem, k]) with 0-based node indices in
    the file's order."""
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    head = lines[0].split()
    dim = int(head[head.index("dimension") + 1])
    npe = int(head[head.index("Nnode") + 1])
    c0, c1 = lines.index("Coordinates"), lines.index("end coordinates")
    e0, e1 = lines.index("Elements"), lines.index("end elements")
    rows = [ln.split() for ln in lines[c0 + 1:c1] if ln]
    ids = np.array([int(r[0]) for r in rows])
    coord = np.array([[float(v) for v in r[1:1 + dim]] for r in rows])
    if not np.array_equal(ids, np.arange(1, ids.size + 1)):
        raise ValueError(f"{path}: node ids are not 1..n in order")
    conn = np.array([[int(v) for v in ln.split()[1:1 + npe]] for ln in lines[e0 + 1:e1] if ln], dtype=np.int64) - 1
    return np.ascontiguousarray(coord), np.ascontiguousarray(conn.astype(np.uint32))


def solid3d_mesh(coord, conn, E=200.0, nu=0.43, g=(0.0, 0.0, -1.0), rho=1.0, fix_fraction=0.01, name="mesh"):
    """Linear elasticity on a given (positively oriented) tetrahedral mesh: E,
    nu uniform, body force rho*g, displacement fixed at the nodes within
    fix_fraction of the mesh's length from one end of its principal axis."""
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.uint32)
    p = coord[conn.astype(np.int64)]
    vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    if np.any(vol <= 0.0):
        raise ValueError("tetrahedra must be positively oriented")
    c = coord - coord.mean(axis=0)
    axis = np.linalg.eigh(c.T @ c)[1][:, -1]
    t = c @ axis
    nnode, nelem = coord.shape[0], conn.shape[0]
    flags = np.zeros((nnode, 3), dtype=np.int32)
    flags[t <= t.min() + fix_fraction * (t.max() - t.min()), :] = 1
    lam, mu = lame(np.full(nelem, float(E)), nu)
    return Problem(kind="solid3d", name=name, n=0, b=3, npe=4, coord=coord, conn=conn, nnode=nnode, nelem=nelem,
                   lam=np.ascontiguousarray(lam), mu=np.ascontiguousarray(mu), rho=rho, g=tuple(g),
                   flags=flags.ravel(), seed=0, sigma=0.0)


def zhelmholtz2d(nx, ny=None, lx=1.0, ly=1.0, wave_length=0.25):
    """Complex Helmholtz (acoustic) problem of the reference's complex twin:
    -Lap u - k^2 u on an lx x ly rectangle of nx x ny cells (P1 triangles,
    src/femeqn/eqn_helmholtz.cpp:51-133) with Sommerfeld radiation line
    elements on the bottom, right and top edges (:437-497) and the incoming
    wave u = 1 prescribed on the x = 0 edge.  `value` holds that boundary
    value (complex); the linear system is solved for the update.  Scalar
    unknowns (b = 1), complex symmetric matrix."""
    ny = nx if ny is None else ny
    coord, conn = tri_grid(nx, ny, lx, ly)
    nnode = coord.shape[0]
    node = lambda i, j: j * (nx + 1) + i
    lines = [(node(i, 0), node(i + 1, 0)) for i in range(nx)]
    lines += [(node(nx, j), node(nx, j + 1)) for j in range(ny)]
    lines += [(node(i + 1, ny), node(i, ny)) for i in range(nx)]
    lconn = np.ascontiguousarray(np.array(lines, dtype=np.uint32))
    flags = np.zeros(nnode, dtype=np.int32)
    flags[np.isclose(coord[:, 0], 0.0)] = 1
    value = np.zeros(nnode, dtype=np.complex128)
    value[flags == 1] = 1.0
    return Problem(kind="zhelmholtz2d", name=f"zhelmholtz2d_{nx}x{ny}", n=nx, b=1, npe=3, coord=coord, conn=conn,
                   lconn=lconn, nnode=nnode, nelem=conn.shape[0], wave_length=float(wave_length), flags=flags,
                   value=value, seed=0)


def fourier_stream_velocity(coord, seed, modes=4):
    """Smooth divergence-free advection field on the unit square: the velocity
    (d psi/dy, -d psi/dx) of psi = sum_{k,l=1..modes} c_kl sin(k pi x) sin(l pi y),
    c_kl standard normal / (k*k + l*l), scaled so max_nodes |a| = 1.  The normal
    component vanishes on the walls."""
    c = normal(seed, modes * modes).reshape(modes, modes)
    x, y = coord[:, 0], coord[:, 1]
    ax = np.zeros(x.size)
    ay = np.zeros(x.size)
    for k in range(1, modes + 1):
        sx, cx = np.sin(k * np.pi * x), np.cos(k * np.pi * x)
        for l in range(1, modes + 1):
            sy, cy = np.sin(l * np.pi * y), np.cos(l * np.pi * y)
            w = c[k - 1, l - 1] / (k * k + l * l)
            ax += w * l * np.pi * sx * cy
            ay -= w * k * np.pi * cx * sy
    a = np.stack([ax, ay], axis=1)
    a /= np.max(np.sqrt(ax * ax + ay * ay))
    return np.ascontiguousarray(a)


def oseen2d(n=256, seed=12345, myu=0.01, rho=1.0, lid=1.0):
    """Linearised incompressible flow (fluid2d path): lid-driven cavity on the
    unit square, n x n cells, combined (u, v, p) unknowns per node, viscosity
    myu, prescribed Fourier advection field.  Velocity is fixed on the whole
    boundary (u = lid on the top edge without its corners, 0 elsewhere), and the
    pressure at node 0 is pinned to 0 (with Dirichlet velocity everywhere the
    pressure is otherwise defined up to a constant).  `value` holds the
    boundary values: the linear system is solved for the update."""
    coord, conn = tri_grid(n)
    nnode = coord.shape[0]
    adv = fourier_stream_velocity(coord, seed)
    x, y = coord[:, 0], coord[:, 1]
    wall = np.isclose(x, 0.0) | np.isclose(x, 1.0) | np.isclose(y, 0.0) | np.isclose(y, 1.0)
    flags = np.zeros((nnode, 3), dtype=np.int32)
    flags[wall, 0:2] = 1
    flags[0, 2] = 1
    value = np.zeros((nnode, 3))
    top = np.isclose(y, 1.0) & ~np.isclose(x, 0.0) & ~np.isclose(x, 1.0)
    value[top, 0] = lid
    return Problem(kind="oseen2d", name=f"oseen2d_{n}", n=n, b=3, npe=3, coord=coord, conn=conn, nnode=nnode,
                   nelem=conn.shape[0], myu=myu, rho=rho, adv=adv, flags=flags.ravel(),
                   value=np.ascontiguousarray(value.ravel()), seed=seed)


def helmholtz2d(nx, ny, lx, ly, seed=12345, sigma=0.5, c=1.0, source=1.0, offset=0):
    """Positive Helmholtz / implicit diffusion alpha*K + c*M on an lx x ly
    rectangle of nx x ny cells, alpha lognormal(0, sigma) per element drawn
    from the stream of `seed` at `offset`, f = source, u = 0 on the x = 0 edge."""
    coord, conn = tri_grid(nx, ny, lx, ly)
    nelem = conn.shape[0]
    alpha = lognormal(seed, nelem, 0.0, sigma, offset) if sigma > 0 else np.ones(nelem)
    nnode = coord.shape[0]
    flags = np.zeros(nnode, dtype=np.int32)
    flags[np.isclose(coord[:, 0], 0.0)] = 1
    return Problem(kind="helmholtz2d", name=f"helmholtz2d_{nx}x{ny}", b=1, npe=3, coord=coord, conn=conn,
                   nnode=nnode, nelem=nelem, alpha=alpha, cmass=np.full(nelem, float(c)), source=source,
                   flags=flags, seed=seed)


def helmholtz_batch(count=512, seed=12345, nmin=500, nmax=50000, sigma=0.5, cmax=10.0):
    """Batch of independent rectangles (batched path): node counts log-uniform
    in [nmin, nmax], aspect ratio lx/ly uniform in [1, 4] (ly = 1, near-square
    cells), mass coefficient c uniform in [0, cmax], alpha lognormal per
    element, u = 0 on the short edge x = 0."""
    u = uniform(seed, 3 * count).reshape(count, 3)
    out = []
    off = 3 * count
    for s in range(count):
        nn = nmin * (nmax / nmin) ** u[s, 0]
        ar = 1.0 + 3.0 * u[s, 1]
        ny = max(2, int(round(np.sqrt(nn / ar))) - 1)
        nx = max(2, int(round(ar * (ny + 1))) - 1)
        p = helmholtz2d(nx, ny, ar, 1.0, seed=seed, sigma=sigma, c=cmax * u[s, 2], offset=off)
        off += p.nelem
        out.append(p)
    return out


def concat(problems):
    """Disjoint union of scalar Helmholtz problems as one problem whose
    matrix has the systems as diagonal blocks; sys_ptr[s] is the first node
    of system s."""
    kinds = {p.kind for p in problems}
    assert kinds == {"helmholtz2d"}, kinds
    sys_ptr = np.zeros(len(problems) + 1, dtype=np.uint32)
    for s, p in enumerate(problems):
        sys_ptr[s + 1] = sys_ptr[s] + p.nnode
    conn = np.concatenate([p.conn + np.uint32(sys_ptr[s]) for s, p in enumerate(problems)]).astype(np.uint32)
    nnode = int(sys_ptr[-1])
    return Problem(kind="helmholtz2d", name=f"batch_{len(problems)}", b=1, npe=3,
                   coord=np.ascontiguousarray(np.concatenate([p.coord for p in problems])),
                   conn=np.ascontiguousarray(conn), nnode=nnode, nelem=conn.shape[0],
                   alpha=np.ascontiguousarray(np.concatenate([p.alpha for p in problems])),
                   cmass=np.ascontiguousarray(np.concatenate([p.cmass for p in problems])),
                   source=problems[0].source,
                   flags=np.ascontiguousarray(np.concatenate([p.flags for p in problems]).astype(np.int32)),
                   sys_ptr=sys_ptr, parts=problems)
