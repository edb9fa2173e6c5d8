"""Restricted power diagram on the GPU vs the FP64 oracle.

Bars: FP64 form within 1e-12, mixed form within 1e-6, of the mean cell volume
1e9/n (volume) and of mean cell volume x 1000 (first moments); per result, on
every cell above 1% of the mean cell volume, the relative volume error and
the barycentre displacement over the cell's size within 1e-12 (FP64, against
the oracle evaluated in each seed's frame) and 1e-6 (mixed).
"""
import numpy as np
import pytest

from conftest import cell_errors, moment_errors, oracle_grid

pytestmark = pytest.mark.gpu


def _run(rpd, orc, ctx, x, grid, prec, K, P, V, T=100, sm=0, cull=1):
    og = oracle_grid(orc, grid) if grid is not None else None
    ref, ref_st, _, _ = orc.solve(x, og, K=K, P=P, V=V, T=T, security_mode=sm, frame=1)
    dom = ctx.domain(grid) if grid is not None else None
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=K, P=P, V=V, T=T, precision=prec, security_mode=sm, cull=cull)
    got = ctx.moments()
    status = ctx.status()
    return ref, ref_st, got, status, st


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_unrestricted_voronoi(rpd, orc, ctx, prec, tol):
    n = 4000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=11)
    ref, ref_st, got, status, st = _run(rpd, orc, ctx, x, None, prec, 40, 40, 64)
    ev, em = moment_errors(got, ref, n)
    cv, cb, nc = cell_errors(got, ref, n)
    print("unrestricted", prec, st["status_hist"], ev, em, cv, cb, nc, flush=True)
    assert (status == 4).all()
    assert ev <= tol and em <= tol
    assert cv <= tol and cb <= tol and nc == n
    assert abs(got[:, 3].sum() - 1e9) / 1e9 < (1e-12 if prec == 0 else 1e-7)


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_sphere_domain(rpd, orc, ctx, prec, tol):
    n = 10000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=0)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.45))
    ref, ref_st, got, status, st = _run(rpd, orc, ctx, x, grid, prec, 50, 30, 40)
    ev, em = moment_errors(got, ref, n)
    cv, cb, nc = cell_errors(got, ref, n)
    print("sphere", prec, st, ev, em, cv, cb, nc, flush=True)
    assert (status == 4).all()
    assert ev <= tol and em <= tol
    assert cv <= tol and cb <= tol and nc > n // 3
    vol = grid.volume()
    assert abs(got[:, 3].sum() - vol) / vol < (1e-12 if prec == 0 else 1e-7)


@pytest.mark.parametrize("prec,tol,sm,cull", [(0, 1e-12, 0, 1), (1, 1e-6, 0, 1), (0, 1e-12, 1, 1), (1, 1e-6, 1, 1),
                                             (0, 1e-12, 0, 0), (2, 1e-4, 0, 1)])
def test_power_torus(rpd, orc, ctx, prec, tol, sm, cull):
    """Power diagram restricted to a torus; sm=1 is the reference stop test
    (weights ignored), checked against the oracle run with the same test.  In
    the mixed form a reference stop within the FP32 radius's margin is sent
    to the FP64 pass, so the mixed cells are the FP64 ones there too."""
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=2, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 40, (0.5, 0.5, 0.5, 0.3, 0.12))
    ref, ref_st, got, status, st = _run(rpd, orc, ctx, x, grid, prec, 64, 40, 60, sm=sm, cull=cull)
    ev, em = moment_errors(got, ref, n)
    cv, cb, nc = cell_errors(got, ref, n)
    print("torus power", prec, st, ev, em, cv, cb, nc, flush=True)
    if cull:
        # a culled cell ends as success with volume exactly 0 before it may
        # become empty; every other status must agree
        diff = status != ref_st
        assert (got[diff] == 0).all() and (ref[diff] == 0).all()
        assert set(zip(status[diff], ref_st[diff])) <= {(4, 6)}
    else:
        assert np.array_equal(status, ref_st)
    assert ev <= tol and em <= tol
    if prec != 2:  # the FP32 form has no per-cell bar
        assert cv <= tol and cb <= tol and nc > 1000


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_coarse_seeds_fine_grid(rpd, orc, ctx, prec, tol):
    """1000 power-weighted seeds in a sphere voxelized at 96^3: boundary cells
    meet up to a few thousand triangles, far beyond T = 100 (raised to 256 by
    the solver).  The re-queue pass takes T from the largest triangle union
    the finish kernel counted, so the solve ends in at most three passes; the
    oracle runs with T = 4000."""
    n = 1000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=5, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 96, (0.5, 0.5, 0.5, 0.45))
    ref, ref_st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=50, P=32, V=48, T=4000)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=50, P=32, V=48, T=100, precision=prec)
    got, status = ctx.moments(), ctx.status()
    dom.close()
    ev, em = moment_errors(got, ref, n)
    print("coarse seeds", prec, st["passes"], st["pass_cells"], st["pass_caps"], ev, em, flush=True)
    assert set(np.unique(ref_st)) <= {4, 6}
    diff = status != ref_st
    assert (got[diff] == 0).all() and (ref[diff] == 0).all()
    assert set(zip(status[diff], ref_st[diff])) <= {(4, 6)}
    assert st["passes"] <= 3
    assert ev <= tol and em <= tol


@pytest.mark.parametrize("prec", [0, 1])
def test_domain_cover_skip(rpd, ctx, prec):
    """Seeds proven outside the domain before the kNN (cull = 1) end with
    volume 0, and every cell agrees with the run that clips everything
    (cull = 0)."""
    n = 200000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=7, wmax=0.25 * h * h * 1e6)
    dom = ctx.domain(rpd.Grid.shape(rpd.SHAPE_TORUS, 64, (0.5, 0.5, 0.5, 0.3, 0.12)))
    ctx.set_seeds(x)
    st0 = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec, cull=0)
    full = ctx.moments()
    st1 = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec, cull=1)
    got = ctx.moments()
    dom.close()
    print("cover skip", prec, st1["skipped_cells"], st1["ms_cover"], st0["ms_total"], st1["ms_total"], flush=True)
    assert st0["skipped_cells"] == 0
    assert st1["skipped_cells"] > n // 2
    ev, em = moment_errors(got, full, n)
    assert ev <= 1e-12 and em <= 1e-12
    # the cells the cover test zeroed have (numerically) zero volume when clipped
    assert (np.abs(full[(got == 0).all(1), 3]) <= 1e-12 * 1e9 / n).all()


def test_knn_launch_sized_from_last_solve(rpd, ctx):
    """Pass 1's kNN is launched over the last solve's kept count (same seed
    count) plus a margin, before the new count is known: a larger count gets
    a second launch for the rest, a smaller one leaves empty waves.  Both
    orders give the bits of a fresh context, which launches over every
    seed."""
    n = 120000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=17, wmax=0.25 * h * h * 1e6)
    small = ctx.domain(rpd.Grid.shape(rpd.SHAPE_SPHERE, 48, (0.5, 0.5, 0.5, 0.12)))
    large = ctx.domain(rpd.Grid.shape(rpd.SHAPE_TORUS, 48, (0.5, 0.5, 0.5, 0.3, 0.12)))
    caps = dict(K=50, P=32, V=48, T=100, precision=rpd.PREC_MIXED)
    fresh = {}
    for name, d in (("small", small), ("large", large)):
        c2 = rpd.Context(0)
        d2 = c2.domain(d.grid)
        c2.set_seeds(x)
        st = c2.solve(d2, **caps)
        fresh[name] = (c2.moments(), st["skipped_cells"])
        d2.close()
        c2.close()
    kept = {k: n - v[1] for k, v in fresh.items()}
    # the large domain keeps more than the small one's count plus the margin
    assert kept["large"] > kept["small"] + kept["small"] // 16 + 4096
    ctx.set_seeds(x)
    for order in (("small", "large"), ("large", "small")):
        for name in order:
            d = small if name == "small" else large
            st = ctx.solve(d, **caps)
            print("kNN launch size", order, name, n - st["skipped_cells"], flush=True)
            assert np.array_equal(ctx.moments(), fresh[name][0]), (order, name)
    small.close()
    large.close()


@pytest.mark.parametrize("caps", [dict(K=64, P=34, V=60, T=100), dict(K=12, P=16, V=40, T=8)])
def test_caps_do_not_change_results(rpd, ctx, caps):
    """FP64 form: a solve whose small caps send cells through re-queue passes
    (recomputed, or resumed after their list ran out, and restricted by
    several waves per cell in the short queues) gives the same bits as one
    whose caps are ample (every cell done in pass 1, one wave per cell)."""
    n = 100000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=41, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 100, (0.5, 0.5, 0.5, 0.3, 0.12))
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st0 = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=rpd.PREC_FP64, cull=0)
    ref, ref_st = ctx.moments(), ctx.status()
    assert st0["queued_first"] > 32768
    st1 = ctx.solve(dom, precision=rpd.PREC_FP64, cull=0, max_extra_steps=12, **caps)
    got, status = ctx.moments(), ctx.status()
    print("caps", caps, st1["passes"], st1["pass_cells"], st1["status_hist"], flush=True)
    # a re-queue pass short enough to spread its cells over several waves
    assert st1["passes"] > 1 and any(0 < k <= 32768 for k in st1["pass_cells"][1:])
    assert np.array_equal(status, ref_st)
    assert np.array_equal(got, ref)
    dom.close()


def test_domain_outlives_its_context(rpd, ctx):
    """A domain made by one context serves another after the first is
    destroyed, and is destroyed last."""
    n = 5000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=5)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.45))
    ctx.set_seeds(x)
    dom0 = ctx.domain(grid)
    ctx.solve(dom0, K=50, P=30, V=40, precision=rpd.PREC_FP64)
    ref = ctx.moments()
    dom0.close()
    other = rpd.Context(0)
    dom = other.domain(grid)
    other.close()
    ctx.solve(dom, K=50, P=30, V=40, precision=rpd.PREC_FP64)
    assert np.array_equal(ctx.moments(), ref)
    dom.close()
