"""Packed batch (rpd_solve_batch): several diagrams of different sizes,
domains and weights in one launch of each kernel.  Every request's cells must
equal the same seeds solved alone with rpd_solve, bit for bit (each request
is binned and searched on its own, and a cell's computation reads only its
own request's seeds and domain), in either order of the requests and in
both the mixed and the FP64 forms; the batch's statuses are the requests'."""
import numpy as np
import pytest

from conftest import cell_errors, oracle_grid

pytestmark = pytest.mark.gpu

G = 32
SHAPES = [(0, (0.5, 0.5, 0.5, 0.45)), (1, (0.5, 0.5, 0.5, 0.3, 0.12)), (3, (0.5, 0.5, 0.5, 0.3))]


def _seeds(rpd, n, seed, walls=0.0):
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=seed, wmax=0.25 * h * h * 1e6)
    if walls > 0:
        rng = np.random.default_rng(seed)
        x[rng.choice(n, int(walls * n), replace=False), 3] = rpd.NO_CELL_LIMITS
    return x


@pytest.fixture(scope="module")
def doms(rpd, ctx):
    grids = [rpd.Grid.shape(code, G, prm) for code, prm in SHAPES]
    ds = [ctx.domain(g) for g in grids]
    yield grids, ds
    for d in ds:
        d.close()


@pytest.mark.parametrize("prec", [1, 0])
def test_batch_equals_alone(rpd, ctx, doms, prec):
    grids, ds = doms
    # sizes from tiny (fewer seeds than K) to large; a request without a
    # domain, one with walls (no cover test for it), the others with one
    reqs = [(ds[0], _seeds(rpd, 3000, 41)), (None, _seeds(rpd, 700, 42)), (ds[1], _seeds(rpd, 40000, 43)),
            (ds[2], _seeds(rpd, 30, 44)), (ds[1], _seeds(rpd, 9000, 45, walls=0.02)), (ds[0], _seeds(rpd, 15000, 46))]
    caps = dict(K=50, P=32, V=48, T=100, precision=prec)
    alone, alone_st = [], []
    for d, x in reqs:
        ctx.set_seeds(x)
        ctx.solve(d, **caps)
        alone.append(ctx.moments())
        alone_st.append(ctx.status())
    for order in (list(range(len(reqs))), list(reversed(range(len(reqs))))):
        st, outs = ctx.solve_batch([reqs[i] for i in order], **caps)
        status = ctx.status()
        print("batch", prec, order, "passes", st["passes"], "skipped", st["skipped_cells"], "hist",
              st["status_hist"], flush=True)
        assert st["n"] == sum(x.shape[0] for _, x in reqs)
        first = 0
        for k, i in enumerate(order):
            n = reqs[i][1].shape[0]
            assert outs[k].shape == (n, 4)
            assert np.array_equal(outs[k], alone[i]), "request %d differs from its solve alone" % i
            assert np.array_equal(status[first:first + n], alone_st[i]), "request %d statuses" % i
            first += n
        # moments() holds the requests back to back
        assert np.array_equal(ctx.moments(), np.concatenate(outs))
    # each restricted request's volume is its domain's
    for (d, x), o in zip(reqs, alone):
        if d is not None and (x[:, 3] > rpd.NO_CELL_LIMITS).all():
            vol = d.grid.volume()
            assert abs(o[:, 3].sum() - vol) <= 1e-9 * vol


def test_batch_matches_oracle_per_cell(rpd, orc, ctx, doms):
    """A batch's cells against the oracle directly (FP64 form, 1e-12 per
    cell; mixed 1e-6), on two requests of different domains."""
    grids, ds = doms
    reqs = [(1, _seeds(rpd, 6000, 51)), (2, _seeds(rpd, 2500, 52))]
    for prec, tol in ((0, 1e-12), (1, 1e-6)):
        _, outs = ctx.solve_batch([(ds[i], x) for i, x in reqs], K=50, P=32, V=48, T=100, precision=prec)
        for (i, x), got in zip(reqs, outs):
            ref, _, _, _ = orc.solve(x, oracle_grid(orc, grids[i]), K=50, P=32, V=48, T=100, frame=1)
            cv, cb, nc = cell_errors(got, ref, x.shape[0])
            print("batch vs oracle", prec, i, nc, cv, cb, flush=True)
            assert nc > 100 and cv <= tol and cb <= tol


def test_batch_refusals(rpd, ctx, doms):
    grids, ds = doms
    x = _seeds(rpd, 2000, 61)
    ctx.solve_batch([(ds[0], x), (ds[1], x)], K=40, P=32, V=48)
    # a context holding a batch refuses the single-diagram calls until new seeds
    with pytest.raises(rpd.RpdError):
        ctx.solve(ds[0], K=40, P=32, V=48)
    with pytest.raises(rpd.RpdError):
        ctx.knn(20)
    with pytest.raises(rpd.RpdError):
        ctx.compact_used(1.0, 1)
    out, perm = ctx.compact_used(1.0, 0)  # used cells only: fine
    assert perm.size > 0 and (perm < 4000).all()
    with pytest.raises(rpd.RpdError):
        ctx.solve_batch([(ds[0], x[:1])], K=40, P=32, V=48)
    ctx.set_seeds(x)
    ctx.solve(ds[0], K=40, P=32, V=48)
