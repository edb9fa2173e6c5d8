"""Used cells and bordering rings on the GPU (rpd_compact_used) against the
oracle's rings (tests/test_oracle_bordering.py pins those), from the same
moments: the same kept cells in the same order, the same values and the
same NO_CELL_LIMITS flags, for keep_distance -1, 0, 1 and 2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_bordering_rings_vs_oracle(rpd, orc, ctx, prec):
    n = 5000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=23, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.3))
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    ctx.solve(dom, K=50, P=32, V=48, precision=prec)
    mg = ctx.moments()
    dom.close()
    keep = 1e-3 * 1e9 / n
    for d in (-1, 0, 1, 2):
        out, ids = ctx.compact_used(keep, d)
        ref_out, ref_ids = orc.compact_used(x, mg, keep, d)
        print("bordering", prec, d, ids.size, flush=True)
        assert np.array_equal(ids, ref_ids), d
        assert np.array_equal(out, ref_out), d
