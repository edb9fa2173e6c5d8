"""kNN on the GPU vs the oracle: bit-exact neighbour lists and permutation.

The oracle (oracle/rpd_oracle.cpp, restating knearests.cpp/knearests.cl with a
stable binning and the canonical (d^2, original index) order) is itself
pinned against brute force in tests/test_oracle.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,K,kind", [(500, 20, "uniform"), (5000, 50, "uniform"), (20000, 64, "uniform"),
                                      (20000, 70, "uniform"), (4096, 100, "lattice"), (8000, 26, "lattice_exact"),
                                      (6000, 40, "clustered")])
def test_knn_bit_exact(rpd, orc, ctx, prec, n, K, kind):
    if kind == "uniform":
        x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=n + K)
    elif kind == "lattice":
        m = int(round(n ** (1 / 3)))
        x = rpd.generate_seeds(rpd.SEEDS_LATTICE, m ** 3, seed=3, p0=m, p1=1e-6)
    elif kind == "lattice_exact":  # massive distance ties, order decided by original index
        m = int(round(n ** (1 / 3)))
        x = rpd.generate_seeds(rpd.SEEDS_LATTICE, m ** 3, seed=3, p0=m, p1=0.0)
    else:  # two thirds of the seeds in a small ball: grid cells of far more than 32 seeds
        x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=7)
        k = 2 * n // 3
        x[:k, :3] = 500.0 + (x[:k, :3] - 500.0) * 0.02
    ctx.set_seeds(x)
    got = ctx.knn(K, prec)
    perm = ctx.perm()
    ref_perm, ref, _ = orc.knn(x, K, precision=prec, method=0)
    print("knn", kind, x.shape[0], K, prec, flush=True)
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(got, ref)


def test_knn_fewer_than_k(rpd, orc, ctx):
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, 40, seed=5)
    ctx.set_seeds(x)
    got = ctx.knn(64, 1)
    _, ref, _ = orc.knn(x, 64, precision=1, method=1)
    assert np.array_equal(got, ref)
    assert (got[:, 39:] == -1).all() and (got[:, :39] >= 0).all()


@pytest.mark.parametrize("prec", [0, 1])
def test_solve_binning_crowded_cells(rpd, orc, ctx, prec):
    """A solve queues the binning of new seeds without waiting for it; grid
    cells of more than 32 seeds take the sort path once the solve has
    synchronised its stream.  The permutation is then the oracle's, and the
    solve's cells and statuses are those of a context that binned the same
    seeds through the waiting path (ctx.knn) before solving."""
    n = 6000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=7)
    k = 2 * n // 3
    x[:k, :3] = 500.0 + (x[:k, :3] - 500.0) * 0.02
    caps = dict(K=40, P=40, V=64, T=100, precision=prec)
    key_prec = 1 if prec == rpd.PREC_FP64 else 0
    ctx.set_seeds(x)
    ctx.solve(None, **caps)
    ref_perm, _, _ = orc.knn(x, 40, precision=key_prec, method=0)
    print("crowded solve", prec, flush=True)
    assert np.array_equal(ctx.perm(), ref_perm)
    c2 = rpd.Context(0)
    c2.set_seeds(x)
    c2.knn(40, key_prec)
    c2.solve(None, **caps)
    assert np.array_equal(ctx.moments(), c2.moments())
    assert np.array_equal(ctx.status(), c2.status())
    c2.close()
