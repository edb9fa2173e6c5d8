"""Pins of the CPU oracle (CPU only, no GPU).

Each test checks the oracle against something it does not share code with:
an O(N^2) numpy kNN with the canonical key, closed forms, invariants of the
power diagram, and brute-force polytope clipping (tests/polytope.py) of every
cell against every bisector and every kept voxel.
"""
import numpy as np
import pytest

from conftest import oracle_grid
import polytope as pt


def _brute_knn(xyzw, K, precision):
    """Canonical kNN: key (d^2, original index), d^2 = ((dx*dx + dy*dy) + dz*dz)
    unfused in the seed precision, self excluded (knearests.cl:12-78 with a
    total order in place of the heap's tie order)."""
    dt = np.float32 if precision == 0 else np.float64
    p = xyzw[:, :3].astype(dt)
    n = p.shape[0]
    out = np.full((n, K), -1, np.int64)
    idx = np.arange(n)
    for i in range(n):
        d = p - p[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        d2[i] = np.inf
        order = np.lexsort((idx, d2))[: min(K, n - 1)]
        out[i, : order.size] = order
    return out


@pytest.mark.parametrize("kind,n,K,prec", [("uniform", 3000, 30, 0), ("uniform", 3000, 30, 1),
                                           ("lattice", 1728, 26, 0), ("uniform", 300, 320, 1)])
def test_knn_grid_matches_brute_force(rpd, orc, kind, n, K, prec):
    if kind == "uniform":
        x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=11)
    else:
        x = rpd.generate_seeds(rpd.SEEDS_LATTICE, n, seed=0, p0=12, p1=0.0)  # exact ties
    perm, kn, counters = orc.knn(x, K, precision=prec, method=0)
    got = np.full((n, K), -1, np.int64)
    ok = kn >= 0
    got[perm] = np.where(ok, perm[np.where(ok, kn, 0)], -1)
    ref = _brute_knn(x, K, prec)
    assert np.array_equal(got, ref)
    # binning: perm is a bijection, counters an exclusive prefix of the cells
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert counters[0] == 0 and counters[-1] == n and np.all(np.diff(counters) >= 0)


def test_unrestricted_sums(rpd, orc):
    """No domain (CUBE_EPSILON 0): the cells tile [0,1000]^3, so the volumes
    sum to 1e9 and the first moments to 500 * 1e9 per axis."""
    n = 2000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=1, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    out, st, _, _ = orc.solve(x, None, K=40, P=40, V=60)
    assert np.all((st == 4) | (st == 6))
    assert abs(out[:, 3].sum() / 1e9 - 1) < 1e-12
    assert np.abs(out[:, :3].sum(0) / 5e11 - 1).max() < 1e-12
    assert np.all(out[st == 6] == 0)


def test_weight_shift_invariance(rpd, orc):
    """Adding one constant to every weight leaves the power diagram unchanged."""
    n = 1500
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=2, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    out, st, _, _ = orc.solve(x, None, K=40, P=40, V=60)
    x2 = x.copy()
    x2[:, 3] += 123.0
    out2, st2, _, _ = orc.solve(x2, None, K=40, P=40, V=60)
    v = 1e9 / n
    assert np.array_equal(st, st2)
    assert np.abs(out2[:, 3] - out[:, 3]).max() / v < 1e-12
    assert np.abs(out2[:, :3] - out[:, :3]).max() / (v * 1000) < 1e-12


def test_jittered_lattice_closed_form(rpd, orc):
    """Lattice (i+1/2)h with a tiny jitter: every cell is (up to the jitter)
    the axis cube of side h around its seed.  FP64 only: in FP32 a jitter of
    1e-7 h is below the coordinate resolution and the lattice is degenerate."""
    m = 8
    n = m ** 3
    x = rpd.generate_seeds(rpd.SEEDS_LATTICE, n, seed=5, p0=m, p1=1e-7)
    h = 1000.0 / m
    out, st, _, _ = orc.solve(x, None, K=30, P=40, V=60, precision=1)
    assert np.all(st == 4)
    assert np.abs(out[:, 3] - h ** 3).max() / h ** 3 < 1e-5
    c = out[:, :3] / out[:, 3:4]
    assert np.abs(c - x[:, :3]).max() / h < 1e-5


def test_equal_weights_is_voronoi(rpd, orc):
    """w = const gives the Voronoi diagram; the reference stop (weights
    ignored) and the exact power stop agree there."""
    n = 1200
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=4)
    a, sa, _, _ = orc.solve(x, None, K=40, P=40, V=60, security_mode=0)
    b, sb, _, _ = orc.solve(x, None, K=40, P=40, V=60, security_mode=1)
    assert np.array_equal(sa, sb)
    assert np.array_equal(a, b)


# the oracle's forms: FP64 in the reference's [0,1000]^3 frame, FP64 in each
# seed's frame, long double in the reference's frame
FORMS = [(1, 0), (1, 1), (2, 0)]


@pytest.mark.parametrize("prec,frame", FORMS)
@pytest.mark.parametrize("shape,g,params,n,seed", [
    ("sphere", 8, (0.5, 0.5, 0.5, 0.42), 40, 3),
    ("torus", 10, (0.5, 0.5, 0.5, 0.3, 0.15), 36, 7),
    ("cube_minus_sphere", 6, (0.5, 0.5, 0.5, 0.3), 30, 9),
    ("cube", 5, (), 30, 10),
])
def test_restricted_cells_brute_force(rpd, orc, shape, g, params, n, seed, prec, frame):
    """Every restricted cell (volume and first moments) against brute-force
    clipping of the cell by all N-1 bisectors and by every kept voxel."""
    code = {"sphere": rpd.SHAPE_SPHERE, "torus": rpd.SHAPE_TORUS, "cube": rpd.SHAPE_CUBE,
            "cube_minus_sphere": rpd.SHAPE_CUBE_MINUS_SPHERE}[shape]
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=seed, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    grid = rpd.Grid.shape(code, g, params)
    mask = rpd.voxel_mask(code, g, params)
    out, st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=n + 4, P=200, V=250, T=4000, precision=prec,
                              frame=frame)
    ref = np.array([pt.restricted_cell(pt.power_cell(x, i), mask, g) for i in range(n)])
    assert np.all((st == 4) | (st == 6))
    v = 1e9 / n
    assert np.abs(out[:, 3] - ref[:, 3]).max() / v < 1e-11
    assert np.abs(out[:, :3] - ref[:, :3]).max() / (v * 1000) < 1e-11


@pytest.mark.parametrize("prec,frame", FORMS)
def test_unrestricted_cells_brute_force(rpd, orc, prec, frame):
    n = 40
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=8, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    out, st, _, _ = orc.solve(x, None, K=n + 4, P=200, V=250, precision=prec, frame=frame)
    ref = np.array([pt.volume_moments(pt.power_cell(x, i)) for i in range(n)])
    v = 1e9 / n
    assert np.abs(out[:, 3] - ref[:, 3]).max() / v < 1e-12
    assert np.abs(out[:, :3] - ref[:, :3]).max() / (v * 1000) < 1e-12


def test_domain_sum_closed_form(rpd, orc):
    """Sum over cells of the restricted volume = kept voxels x voxel volume;
    sum of first moments = the voxel union's first moments."""
    n, g = 5000, 24
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=6, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    prm = (0.5, 0.5, 0.5, 0.3, 0.12)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, g, prm)
    mask = rpd.voxel_mask(rpd.SHAPE_TORUS, g, prm).reshape(-1).astype(bool)
    vs = 1000.0 / g
    vol = mask.sum() * vs ** 3
    idx = np.nonzero(mask)[0]
    cen = np.stack([idx % g, (idx // g) % g, idx // (g * g)], 1) * vs + vs / 2
    mom = cen.sum(0) * vs ** 3
    out, st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=40, P=40, V=60, T=200)
    assert np.all((st == 4) | (st == 6))
    assert abs(out[:, 3].sum() / vol - 1) < 1e-11
    assert np.abs(out[:, :3].sum(0) / mom - 1).max() < 1e-11


def test_forced_failures_converge(rpd, orc):
    """Tiny caps fail most cells in the first pass; the re-queue passes (caps
    x1.5 per overflow kind, voronoi.cpp:1294-1376) converge to the result of
    a run with ample caps."""
    n, g = 3000, 20
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=12, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    grid = oracle_grid(orc, rpd.Grid.shape(rpd.SHAPE_SPHERE, g, (0.5, 0.5, 0.5, 0.45)))
    ref, sr, _, _ = orc.solve(x, grid, K=60, P=80, V=120, T=600)
    got, sg, _, stats = orc.solve(x, grid, K=6, P=10, V=12, T=6, max_extra_steps=14)
    assert np.all((sr == 4) | (sr == 6))
    assert np.array_equal(sr, sg)
    assert stats["passes"] > 3
    v = 1e9 / n
    assert np.abs(got[:, 3] - ref[:, 3]).max() / v < 1e-12
    assert np.abs(got[:, :3] - ref[:, :3]).max() / (v * 1000) < 1e-12


def test_border_seeds_are_copied(rpd, orc):
    """A seed with w <= NO_CELL_LIMITS is copied to the output and acts as a
    wall (plane through it) for its neighbours (ConvexCell.cl:324-325, 873-876)."""
    n = 600
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=13)
    x[:5, 3] = rpd.NO_CELL_LIMITS
    out, st, _, _ = orc.solve(x, None, K=40, P=40, V=60)
    assert np.array_equal(out[:5], x[:5])
    assert np.all(st[5:] == 4)


@pytest.mark.parametrize("shape,prm", [(1, (0.5, 0.5, 0.5, 0.3, 0.12)), (0, (0.5, 0.5, 0.5, 0.2))])
def test_cull_is_exact(rpd, orc, shape, prm):
    """cull = 1 (the cover test before the kNN and the circumradius test in
    the neighbour loop) changes no cell: every moment equals the full clip's,
    and the cells it ends early have volume 0 when clipped.  Both tests must
    also be taken, by most seeds outside the domain."""
    n, g = 30000, 40
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=8, wmax=0.25 * h * h * 1e6)
    og = oracle_grid(orc, rpd.Grid.shape(shape, g, prm))
    full, st0, _, c0 = orc.solve(x, og, K=64, P=40, V=60, T=200, single_pass=1)
    got, st1, ku1, c1 = orc.solve(x, og, K=64, P=40, V=60, T=200, single_pass=1, cull=1)
    assert np.array_equal(full, got) or np.abs(full - got).max() <= 1e-12 * 1e9 / n
    ended = (st1 == 4) & (got == 0).all(1) & ((st0 != 4) | (full != 0).any(1))
    assert (np.abs(full[ended, 3]) <= 1e-12 * 1e9 / n).all()
    skipped = (st1 == 4) & (ku1 == 0)
    assert skipped.sum() > n // 3
    assert c1["n_pred"] < 0.6 * c0["n_pred"]


def test_cull_with_reference_stop_and_walls(rpd, orc):
    """The cover test needs the exact power stop and no border seeds: with
    either missing only the circumradius test runs, and results still match."""
    n, g = 8000, 32
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=9)
    og = oracle_grid(orc, rpd.Grid.shape(1, g, (0.5, 0.5, 0.5, 0.3, 0.12)))
    for sm, walls in ((1, False), (0, True)):
        y = x.copy()
        if walls:
            y[:40, 3] = -1e6
        full, _, _, _ = orc.solve(y, og, K=64, P=40, V=60, T=200, single_pass=1, security_mode=sm)
        got, st, ku, _ = orc.solve(y, og, K=64, P=40, V=60, T=200, single_pass=1, security_mode=sm, cull=1)
        assert np.abs(full - got).max() <= 1e-12 * 1e9 / n
        assert not ((st == 4) & (ku == 0) & (y[:, 3] > -1e5)).any()
