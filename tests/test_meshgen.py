"""Synthetic inputs: the splitmix64 stream, distributions and meshes."""
import numpy as np

from delfem_amd import meshgen


def test_splitmix64_known_answers():
    # published first outputs of splitmix64 seeded with 0 and with 1234567
    z = meshgen.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    z = meshgen.splitmix64(1234567, 2)
    assert [int(v) for v in z] == [6457827717110365317, 3203168211198807973]
    # counter-based: an offset continues the same stream
    assert np.array_equal(meshgen.splitmix64(9, 5)[2:], meshgen.splitmix64(9, 3, offset=2))


def test_distributions():
    u = meshgen.uniform(3, 200000)
    assert 0.0 < u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.005
    g = meshgen.normal(4, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    ln = meshgen.lognormal(5, 200000, 0.0, 0.5)
    assert abs(np.log(ln).std() - 0.5) < 0.01


def test_tri_grid():
    coord, conn = meshgen.tri_grid(4, 3, 2.0, 1.5)
    assert coord.shape == (20, 2) and conn.shape == (24, 3)
    p = coord[conn]
    area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    assert np.all(area > 0) and np.isclose(area.sum(), 3.0)


def test_tet_grid():
    coord, conn = meshgen.tet_grid(3)
    assert coord.shape == (64, 3) and conn.shape == (162, 4)
    p = coord[conn]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6.0
    assert np.all(vol > 0) and np.isclose(vol.sum(), 1.0)
    # Kuhn split: every interior node has 14 neighbours
    nbr = [set() for _ in range(64)]
    for el in conn:
        for a in el:
            nbr[a].update(int(b) for b in el if b != a)
    assert len(nbr[1 + 4 + 16]) == 14


def test_problems():
    p = meshgen.solid3d(4, sigma=1.0, seed=1)
    assert p.nnode == 125 and p.flags.sum() == 25 * 3
    assert np.all(p.lam > 0) and np.all(p.mu > 0)
    q = meshgen.poisson2d(8, seed=1)
    assert q.flags.sum() == 9 and q.alpha.size == 128


def test_relabel_keeps_element_matrices():
    """meshgen.relabel renumbers nodes only: every element matrix and residual
    is unchanged (the element data and order are kept), the nodal arrays
    (raveled or not) follow the permutation."""
    from oracle import oracle
    orc = oracle.Orc()
    for prob in (meshgen.solid3d(3, seed=2, sigma=0.5), meshgen.oseen2d(6, seed=2),
                 meshgen.zhelmholtz2d(6, 4, 1.0, 0.7, wave_length=0.5)):
        order = np.random.default_rng(1).permutation(prob.nnode)
        q = meshgen.relabel(prob, order)
        assert np.array_equal(q.coord, prob.coord[order])
        assert np.array_equal(np.asarray(q.flags).reshape(prob.nnode, -1),
                              np.asarray(prob.flags).reshape(prob.nnode, -1)[order])
        inv = np.empty_like(order)
        inv[order] = np.arange(order.size)
        assert np.array_equal(q.conn, inv[prob.conn])
        if prob.kind != "zhelmholtz2d":
            e1, r1 = oracle.problem_emats(orc, prob)
            e2, r2 = oracle.problem_emats(orc, q)
            assert np.array_equal(e1, e2) and np.array_equal(r1, r2)
        else:
            assert np.array_equal(q.lconn, inv[prob.lconn])


def test_tile_colour_order():
    """tile_colour_order is a permutation; each tile's nodes are numbered
    consecutively; and with tiles of one colour never adjacent, the natural
    ILU(0) sweep of the renumbered mesh has at most (colours x tile nodes)
    level sets (8 colours bound the greedy colouring of these tile grids)."""
    for prob, tn in ((meshgen.oseen2d(32, seed=1), 16), (meshgen.solid3d(8, seed=1), 8)):
        order = meshgen.tile_colour_order(prob, tn)
        assert np.array_equal(np.sort(order), np.arange(prob.nnode))
        # the tiles, recomputed as the numbering defines them
        coord = np.asarray(prob.coord)
        conn = np.asarray(prob.conn)
        h = float(np.median(np.linalg.norm(coord[conn[:, 1]] - coord[conn[:, 0]], axis=1)))
        side = h * max(1.0, round(tn ** (1.0 / coord.shape[1])))
        cell = np.floor((coord - coord.min(axis=0)) / side + 1e-9).astype(np.int64)
        tid = np.unique(np.ravel_multi_index(tuple(cell.T), tuple(cell.max(axis=0) + 1)), return_inverse=True)[1]
        pos = np.empty_like(order)
        pos[order] = np.arange(order.size)
        # nodes of one tile are numbered consecutively
        for t in np.unique(tid)[:50]:
            p = np.sort(pos[tid == t])
            assert p[-1] - p[0] == p.size - 1
        rp, cl = fem_pattern_numpy(prob.nnode, inv_conn(conn, order))
        lv = np.zeros(prob.nnode, dtype=np.int64)
        for r in range(prob.nnode):
            c = cl[rp[r]:rp[r + 1]]
            c = c[c < r]
            lv[r] = lv[c].max() + 1 if c.size else 0
        assert lv.max() + 1 <= 8 * tn


def inv_conn(conn, order):
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    return inv[conn]


def fem_pattern_numpy(nnode, conn):
    """Node adjacency (without the node itself) as CRS, ascending."""
    npe = conn.shape[1]
    a = np.concatenate([np.stack([conn[:, i], conn[:, j]], axis=1) for i in range(npe) for j in range(npe) if i != j])
    a = np.unique(a, axis=0)
    rp = np.searchsorted(a[:, 0], np.arange(nnode + 1))
    return rp, a[:, 1]
