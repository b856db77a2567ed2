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
