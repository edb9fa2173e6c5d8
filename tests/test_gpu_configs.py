"""The remaining workload configurations at their real shapes, on the GPU.

- 100,000 seeds in three distributions: i.i.d. uniform; a 47^3 lattice with
  +-0.25 spacing jitter, cropped to 100K; dart-throwing blue noise with
  minimum distance 0.8 (1/100000)^(1/3).  Torus R0.3 r0.12 at 100^3, K =
  20 / 50 / 70 respectively, P30 V40.  The FP64 form against the oracle's
  FP64 build in each seed's frame (1e-12 of the mean cell volume, and 1e-12
  per cell: relative volume and barycentre over cell size on every cell
  above 1% of the mean cell volume), the FP32 form against its FP32 build
  (1e-4 of the mean cell volume).
- 1,000,000 power-weighted seeds (w ~ U[0, 0.25 h^2]) on the torus at 128^3,
  K64 P40 V60: the mixed and FP64 forms on every cell, a 40,000-cell sample
  against the oracle per cell (FP64 1e-12, mixed 1e-6), total volume = the
  domain's.
- 1,000,000 seeds on a 100^3 lattice with +-1e-6 spacing jitter (nearly
  cospherical: every cell is a near-degenerate cube), unit cube at 100^3,
  weights 0 and, separately, U[0, 1e-4 h^2].  kNN lists bit-exact against
  the oracle's, mixed against FP64 within 1e-6, a 20,000-cell sample against
  the oracle (FP64 1e-12, mixed 1e-6), total volume = the cube's.  The FP64
  fall-back fraction of the mixed form is printed.
"""
import numpy as np
import pytest

from conftest import cell_errors, moment_errors, oracle_grid

pytestmark = pytest.mark.gpu

TORUS = (0.5, 0.5, 0.5, 0.3, 0.12)


def _seeds_100k(rpd, kind):
    n = 100000
    if kind == "uniform":
        return rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=21)
    if kind == "lattice":
        return rpd.generate_seeds(rpd.SEEDS_LATTICE, n, seed=22, p0=47, p1=0.25)
    return rpd.generate_seeds(rpd.SEEDS_BLUE_NOISE, n, seed=23, p0=0.8 * (1.0 / n) ** (1 / 3))


@pytest.mark.parametrize("kind,K", [("uniform", 20), ("lattice", 50), ("blue", 70)])
def test_100k_distributions_torus(rpd, orc, ctx, kind, K):
    x = _seeds_100k(rpd, kind)
    n = x.shape[0]
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 100, TORUS)
    og = oracle_grid(orc, grid)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    for prec, oprec, tol in ((rpd.PREC_FP64, 1, 1e-12), (rpd.PREC_FP32, 0, 1e-4)):
        ref, ref_st, _, _ = orc.solve(x, og, K=K, P=30, V=40, T=100, precision=oprec, vhs0=2, vhs1=2,
                                      frame=1 if oprec == 1 else 0)
        st = ctx.solve(dom, K=K, P=30, V=40, T=100, precision=prec)
        got = ctx.moments()
        ev, em = moment_errors(got, ref, n)
        cv, cb, nc = cell_errors(got, ref, n)
        print("100k", kind, K, prec, st["passes"], st["pass_cells"], st["status_hist"], ev, em, cv, cb, nc,
              flush=True)
        assert ev <= tol and em <= tol, (kind, prec, ev, em)
        if prec == rpd.PREC_FP64:
            assert cv <= 1e-12 and cb <= 1e-12 and nc > 5000, (kind, cv, cb, nc)
        vol = grid.volume()
        assert abs(got[:, 3].sum() - vol) / vol < (1e-11 if prec == rpd.PREC_FP64 else 1e-5)
    dom.close()


def _orig_lists(perm, kn):
    out = np.full(kn.shape, -1, np.int32)
    ok = kn >= 0
    out[perm] = np.where(ok, perm[np.where(ok, kn, 0)], -1)
    return out


@pytest.mark.parametrize("wscale", [0.0, 1e-4])
def test_1m_near_degenerate_lattice(rpd, orc, ctx, wscale):
    n, m = 1000000, 100
    h = 1.0 / m
    x = rpd.generate_seeds(rpd.SEEDS_LATTICE, n, seed=31, p0=m, p1=1e-6, wmax=wscale * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_CUBE, 100, ())
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    # kNN bit-exact with the oracle, both key precisions
    for prec in (0, 1):
        kn = ctx.knn(32, precision=prec)
        perm, okn, _ = orc.knn(x, 32, precision=prec)
        assert np.array_equal(_orig_lists(ctx.perm(), kn), _orig_lists(perm, okn)), prec
    res = {}
    for prec in (rpd.PREC_FP64, rpd.PREC_MIXED):
        st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec)
        res[prec] = (ctx.moments(), ctx.status(), st)
    f64, st64, s64 = res[rpd.PREC_FP64]
    mix, stm, smx = res[rpd.PREC_MIXED]
    ev, em = moment_errors(mix, f64, n)
    print("1M lattice w%.0e: fallback %.4f passes %d hist64 %s histmix %s mixed-vs-fp64 %.2e %.2e" %
          (wscale, smx["fallback_cells"] / n, smx["passes"], s64["status_hist"], smx["status_hist"], ev, em),
          flush=True)
    assert ev <= 1e-6 and em <= 1e-6
    assert abs(f64[:, 3].sum() / 1e9 - 1) < 1e-10
    # a sample of cells against the oracle
    ids = np.arange(0, n, n // 20000, dtype=np.int32)[:20000]
    ref, ref_st, _, _ = orc.solve_ids(x, ids, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, frame=1)
    e64 = moment_errors(f64[ids], ref, n)
    emx = moment_errors(mix[ids], ref, n)
    c64 = cell_errors(f64[ids], ref, n)
    cmx = cell_errors(mix[ids], ref, n)
    print("  sample vs oracle: fp64 %.2e %.2e mixed %.2e %.2e; per cell fp64 %.2e %.2e mixed %.2e %.2e" %
          (*e64, *emx, c64[0], c64[1], cmx[0], cmx[1]), flush=True)
    assert max(e64) <= 1e-12 and max(emx) <= 1e-6
    assert c64[0] <= 1e-12 and c64[1] <= 1e-12 and cmx[0] <= 1e-6 and cmx[1] <= 1e-6
    assert np.array_equal(st64[ids], ref_st)
    dom.close()


def test_1m_power_torus_per_cell(rpd, orc, ctx):
    n, g = 1000000, 128
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=0, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, g, TORUS)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    res = {}
    for prec in (rpd.PREC_FP64, rpd.PREC_MIXED):
        st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec)
        res[prec] = (ctx.moments(), ctx.status(), st)
    dom.close()
    f64, s64, st64 = res[rpd.PREC_FP64]
    mix, smx, stmx = res[rpd.PREC_MIXED]
    vol = grid.volume()
    assert abs(f64[:, 3].sum() - vol) / vol < 1e-11
    assert abs(mix[:, 3].sum() - vol) / vol < 1e-11
    assert set(np.unique(s64)) <= {4, 6} and np.array_equal(s64, smx)
    # the cells near the domain, where the volume is: every cell that has one
    # and a stride over the rest
    near = np.nonzero(f64[:, 3] != 0)[0]
    ids = np.union1d(near[:: max(1, near.size // 30000)], np.arange(0, n, 100)).astype(np.int32)
    ref, ref_st, _, _ = orc.solve_ids(x, ids, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, frame=1, cull=1)
    e64 = cell_errors(f64[ids], ref, n)
    emx = cell_errors(mix[ids], ref, n)
    m64 = moment_errors(f64[ids], ref, n)
    mmx = moment_errors(mix[ids], ref, n)
    print("1M torus: %d cells; FP64 per cell %.2e %.2e (%d), of mean %.2e %.2e; mixed per cell %.2e %.2e, "
          "of mean %.2e %.2e; fallback %d" % (ids.size, e64[0], e64[1], e64[2], *m64, emx[0], emx[1], *mmx,
                                               stmx["fallback_cells"]), flush=True)
    assert e64[2] > 20000
    assert e64[0] <= 1e-12 and e64[1] <= 1e-12 and max(m64) <= 1e-12
    assert emx[0] <= 1e-6 and emx[1] <= 1e-6 and max(mmx) <= 1e-6
