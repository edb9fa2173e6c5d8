"""The oracle's face neighbours and bordering rings (CCUpdateUsedCell and
compute_bordering, ConvexCell.cl:680-690, 754-807; voronoi.cpp:693-751,
1381-1393), pinned without their code:

- adjacency of power cells is symmetric: j is a face neighbour of i iff i is
  one of j (generic seeds);
- the neighbours with a face of more than 0.2 area (the Hessian, pinned in
  test_oracle_hessian.py) are face neighbours;
- rings nest: the cells kept with keep_distance d are kept with d + 1, and
  ring 1 is exactly the used cells and their face neighbours.

CPU only.
"""
import numpy as np

from conftest import oracle_grid


def _nbsets(nb):
    return [set(int(j) for j in row if j >= 0) for row in nb]


def test_face_neighbours_symmetric_and_cover_the_hessian(rpd, orc):
    n = 2000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=21, wmax=0.25 * h * h * 1e6)
    nb = _nbsets(orc.cell_neighbours(x, np.arange(n)))
    assert sum(len(s) for s in nb) > 10 * n
    for i in range(n):
        for j in nb[i]:
            assert i in nb[j], (i, j)
    _, st, _, _, ids, _ = orc.solve_hessian(x, None, H=64, HC=1, K=64, P=60, V=100)
    for i in range(n):
        if st[i] != 4:
            continue
        hs = set(int(j) for j in ids[i] if j >= 0)
        assert hs <= nb[i], i


def test_rings_nest(rpd, orc):
    n = 3000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=22)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.3))
    mg, st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=64, P=60, V=100, T=2000)
    keep = 1e-3 * 1e9 / n
    kept = {}
    for d in (0, 1, 2, 3):
        out, ids = orc.compact_used(x, mg, keep, d)
        kept[d] = set(ids.tolist())
        assert np.array_equal(out[:, :3], mg[ids, :3])
        flag = out[:, 3] == orc.NO_CELL_LIMITS
        assert (mg[ids[flag], 3] < keep).all()
    used = set(np.nonzero(mg[:, 3] >= keep)[0].tolist())
    assert kept[0] == used
    nb = _nbsets(orc.cell_neighbours(x, np.array(sorted(used))))
    assert kept[1] == used | set().union(*nb)
    assert kept[0] < kept[1] < kept[2] <= kept[3]
    all_out, all_ids = orc.compact_used(x, mg, keep, -1)
    assert all_ids.size == n
