"""Serving loop (rpd_server_*): requests of uneven size against resident
domains, run concurrently; every result must equal the same request solved
alone on a plain context, bit for bit (the kernels share no state between
requests and sum each cell in a fixed order)."""
import numpy as np
import pytest

from conftest import moment_errors, oracle_grid

pytestmark = pytest.mark.gpu

G = 32
DOMAINS = [(0, (0.5, 0.5, 0.5, 0.45)), (1, (0.5, 0.5, 0.5, 0.3, 0.12)), (2, ()), (3, (0.5, 0.5, 0.5, 0.3))]
CAPS = dict(K=50, P=32, V=48, T=100, precision=1)


def _requests(rpd, sizes, seed0=100):
    rng = np.random.default_rng(7)
    reqs = []
    base = {}
    for i, n in enumerate(sizes):
        h = (1.0 / n) ** (1.0 / 3.0)
        if n not in base:
            base[n] = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=seed0 + i, wmax=0.25 * h * h * 1e6)
        x = base[n].copy()
        # successive requests on one seed set: weights perturbed as in a transport iteration
        x[:, 3] = np.maximum(0.0, x[:, 3] + rng.normal(0.0, 0.01 * h * h * 1e6, n))
        reqs.append((i % len(DOMAINS), x))
    return reqs


@pytest.fixture(scope="module")
def setup(rpd, ctx):
    grids = [rpd.Grid.shape(code, G, prm) for code, prm in DOMAINS]
    doms = [ctx.domain(g) for g in grids]
    yield grids, doms
    for d in doms:
        d.close()


@pytest.mark.parametrize("inflight,budget", [(4, 0), (3, 64 << 20)])
def test_alone_equals_in_mix(rpd, ctx, setup, inflight, budget):
    grids, doms = setup
    sizes = [3000, 800, 12000, 1500, 5000, 2000, 9000, 1000]
    reqs = _requests(rpd, sizes)
    srv = rpd.Server(ctx, hbm_budget_bytes=budget, max_inflight=inflight)
    tickets = [srv.submit(doms[d], x, **CAPS) for d, x in reqs]
    got = {}
    lat = []
    while len(got) < len(tickets):
        r = srv.poll(wait=True)
        assert r is not None
        got[r[0]] = r[1]
        lat.append(r[2])
    srv.drain()
    assert srv.poll(wait=False) is None
    srv.close()
    assert all(v > 0 for v in lat)
    for t, (d, x) in zip(tickets, reqs):
        ctx.set_seeds(x)
        ctx.solve(doms[d], **CAPS)
        alone = ctx.moments()
        assert np.array_equal(got[t], alone), "ticket %d differs from its solve alone" % t


def test_served_result_matches_oracle(rpd, orc, ctx, setup):
    grids, doms = setup
    d, x = _requests(rpd, [2500])[0]
    srv = rpd.Server(ctx, max_inflight=2)
    srv.submit(doms[d], x, **CAPS)
    t, got, _ = srv.poll(wait=True)
    srv.close()
    ref, st, _, _ = orc.solve(x, oracle_grid(orc, grids[d]), K=50, P=32, V=48, T=100)
    ev, em = moment_errors(got, ref, x.shape[0])
    assert ev < 1e-6 and em < 1e-6


def test_full_size_domains_alone_equals_in_mix(rpd, orc, ctx):
    """The serving workload's domains at their real size (4 resident domains
    at 96^3) with a request of 250,000 seeds among small ones: each result is
    the same alone and in the mix, bit for bit, and a sample of the large
    request's cells matches the oracle (mixed form, 1e-6 per cell)."""
    from conftest import cell_errors
    grids = [rpd.Grid.shape(code, 96, prm) for code, prm in DOMAINS]
    doms = [ctx.domain(g) for g in grids]
    sizes = [250000, 1200, 30000, 4000, 2500, 60000, 1000, 9000]
    reqs = _requests(rpd, sizes, seed0=300)
    srv = rpd.Server(ctx, max_inflight=4)
    tickets = [srv.submit(doms[d], x, **CAPS) for d, x in reqs]
    got = {}
    while len(got) < len(tickets):
        r = srv.poll(wait=True)
        assert r is not None
        got[r[0]] = r[1]
    srv.close()
    for t, (d, x) in zip(tickets, reqs):
        ctx.set_seeds(x)
        ctx.solve(doms[d], **CAPS)
        assert np.array_equal(got[t], ctx.moments()), "ticket %d differs from its solve alone" % t
    d, x = reqs[0]
    n = x.shape[0]
    ids = np.arange(0, n, 25, dtype=np.int32)
    ref, _, _, _ = orc.solve_ids(x, ids, oracle_grid(orc, grids[d]), K=50, P=32, V=48, T=100, frame=1)
    cv, cb, nc = cell_errors(got[tickets[0]][ids], ref, n)
    ev, em = moment_errors(got[tickets[0]][ids], ref, n)
    print("96^3 serving: sample of %d cells, per cell %.2e %.2e, of mean %.2e %.2e" % (nc, cv, cb, ev, em),
          flush=True)
    assert nc > 1000 and cv <= 1e-6 and cb <= 1e-6 and ev <= 1e-6 and em <= 1e-6
    for dd in doms:
        dd.close()
