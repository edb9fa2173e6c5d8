"""The drop-in boundary as a reference caller sees it.

- tests/voroalgo_flow (C++, compiled against include/rpd_voroalgo.hpp only)
  runs the flow of computeVoroDiagram (test_voronoi.cpp:30-62): with a grid
  and keepDistance >= 0 the output is the compacted used cells and rings
  (compressResult, voronoi.cpp:513-691), with their permutations and the
  renumbered Hessian; without a grid every seed.  Its outputs must equal
  the Python mirror's and the oracle's compaction, and its GTime["stat"]
  lines must carry the debug counters.
- the compaction on the device (rpd_compact_used_device) equals the host
  copy (rpd_compact_used) and the oracle's compact_used.
- the debug counters: the K-used histogram of pass 1 equals the oracle's,
  and each histogram sums to the cells it counts.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, oracle_grid

pytestmark = pytest.mark.gpu

FLOW = os.path.join(ROOT, "tests", "voroalgo_flow")


def _run_flow(tmp_path, x, grid_path, K, P, V, T, keep_dist, keep_vol, H, HC, counters):
    seeds = str(tmp_path / "seeds.bin")
    with open(seeds, "wb") as f:
        f.write(np.int64(x.shape[0]).tobytes())
        f.write(np.ascontiguousarray(x, np.float64).tobytes())
    out = str(tmp_path / "out.bin")
    r = subprocess.run([FLOW, seeds, grid_path or "-", str(K), str(P), str(V), str(T), str(keep_dist),
                        repr(float(keep_vol)), str(H), str(HC), str(int(counters)), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    b = open(out, "rb").read()
    o = 0

    def take(dtype, count):
        nonlocal o
        a = np.frombuffer(b, dtype, count, o).copy()
        o += a.nbytes
        return a

    count = int(take(np.int64, 1)[0])
    bary = take(np.float64, 4 * count).reshape(-1, 4)
    perm = take(np.int32, count)
    perm_pts = take(np.int32, int(take(np.int64, 1)[0]))
    hess = take(np.float64, int(take(np.int64, 1)[0]))
    hids = take(np.int32, int(take(np.int64, 1)[0]))
    return bary, perm, perm_pts, hess, hids, r.stdout


@pytest.mark.parametrize("with_grid", [True, False])
def test_voroalgo_flow_matches_mirror(rpd, orc, tmp_path, with_grid):
    n = 6000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=31, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.35))
    gpath = None
    if with_grid:
        gpath = str(tmp_path / "sphere.grid")
        grid.save(gpath)
        grid = rpd.Grid.load(gpath)  # the grid the C++ side reads
    K, P, V, T, H, HC = 50, 32, 48, 100, 24, 1
    keep_vol = 1e-3 * 1e9 / n
    bary, perm, perm_pts, hess, hids, stdout = _run_flow(tmp_path, x, gpath, K, P, V, T, 1, keep_vol, H, HC, True)
    # the Python mirror, same parameters (mixed form: the C++ default m_vhsAlgo {1, 2})
    algo = rpd.VoroAlgo(0)
    algo.m_K, algo.m_P, algo.m_V, algo.m_T = K, P, V, T
    algo.m_keepVolume, algo.m_keepDistance = keep_vol, 1
    algo.m_H, algo.m_HC = H, HC
    algo.precision = rpd.PREC_MIXED
    if with_grid:
        algo.setGrid(grid)
    algo.setInputPoints(x)
    algo.launch()
    print("flow", with_grid, "outputs", perm.size, "of", n, flush=True)
    assert perm.size == algo.getNumOutputPoints()
    assert np.array_equal(perm, algo.getPermutationToUser())
    assert np.array_equal(bary, algo.getBarycenters())
    assert np.array_equal(perm_pts, algo.getPermutationToInputPoints())
    assert np.array_equal(hess, algo.getHessians())
    assert np.array_equal(hids, algo.getHessianIds())
    if with_grid:
        # the compacted output: the oracle's used cells and ring, from the same moments
        full = algo._ctx.moments()
        ref_out, ref_ids = orc.compact_used(x, full, keep_vol, 1)
        assert np.array_equal(perm, ref_ids) and np.array_equal(bary, ref_out)
        assert np.array_equal(perm_pts, perm) and perm.size < n
        # every Hessian id is an output position of a cell with volume
        ids = hids.reshape(-1, H)
        ok = ids >= 0
        assert (bary[ids[ok], 3] >= keep_vol).all()
    else:
        assert perm.size == n and np.array_equal(perm, np.arange(n)) and perm_pts.size == 0
    # GTime["stat"] with the counters of every pass
    assert stdout.startswith('GTime["stat"] = [{"name" : "first"')
    for c in ("counter0", "counter1", "counter2", "counter3", "counter10"):
        assert '"%s":[' % c in stdout, c
    assert ('"counter4":[' in stdout) == with_grid


def test_device_compaction(rpd, orc, ctx):
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=33, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 48, (0.5, 0.5, 0.5, 0.3, 0.12))
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    ctx.solve(dom, K=64, P=40, V=60, precision=rpd.PREC_MIXED)
    full = ctx.moments()
    keep = 1e-3 * 1e9 / n
    for d in (-1, 0, 1, 2):
        cnt, po, pp = ctx.compact_used_device(keep, d)
        dev_out = rpd.read_device(po, np.float64, 4 * cnt).reshape(-1, 4)
        dev_perm = rpd.read_device(pp, np.int32, cnt)
        host_out, host_perm = ctx.compact_used(keep, d)
        ref_out, ref_ids = orc.compact_used(x, full, keep, d)
        print("device compaction", d, cnt, flush=True)
        assert np.array_equal(dev_out, host_out) and np.array_equal(dev_perm, host_perm)
        assert np.array_equal(dev_perm, ref_ids) and np.array_equal(dev_out, ref_out)
    # rpd_compact: the used cells only
    out, perm = ctx.compact(keep)
    sel = np.nonzero(full[:, 3] >= keep)[0]
    assert np.array_equal(perm, sel) and np.array_equal(out, full[sel])
    dom.close()


def test_debug_counters(rpd, orc, ctx):
    """Pass 1's K-used histogram equals the oracle's (every cell reaches its
    security radius with ample K, no domain), and the P / V / union
    histograms count every cell that ends the neighbour loop."""
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=35, wmax=0.25 * h * h * 1e6)
    ref, ref_st, ref_ku, _ = orc.solve(x, None, K=96, P=60, V=96, frame=1)
    # a power cell may be empty (empty_cell ends its loop early, uncounted)
    good = ref_st == 4
    assert set(np.unique(ref_st)) <= {4, 6} and good.sum() > n - 50
    ctx.set_seeds(x)
    st = ctx.solve(None, K=96, P=60, V=96, precision=rpd.PREC_FP64, debug_counters=1)
    dcs = [ctx.debug_counters(p) for p in range(st["passes"])]
    # a cell is counted in K by the pass where it reaches its radius (its
    # neighbours counted from the first pass when it was resumed): over the
    # passes, the oracle's histogram of K used
    hk = sum(d["K"] for d in dcs)
    want = np.bincount(np.minimum(ref_ku[good], 256), minlength=257)[:257]
    print("passes", st["passes"], "K histogram max", dcs[0]["max"], "K used", ref_ku[good].max(), flush=True)
    assert np.array_equal(hk, want)
    assert max(d["max"][0] for d in dcs) == min(ref_ku[good].max(), 256)
    # P and V: every cell that ends a pass's neighbour loop without failing
    assert sum(d["P"].sum() for d in dcs) >= good.sum() and dcs[0]["P"].sum() == dcs[0]["V"].sum()
    d0 = dcs[0]
    assert d0["max"][1] == np.nonzero(d0["P"])[0].max() and d0["max"][2] == np.nonzero(d0["V"])[0].max()
    # with a domain: the union histogram counts the queued cells of each pass
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.4))
    dom = ctx.domain(grid)
    st = ctx.solve(dom, K=64, P=40, V=60, precision=rpd.PREC_MIXED, debug_counters=1)
    tot = sum(int(ctx.debug_counters(p)["T"].sum()) for p in range(st["passes"]))
    print("union histogram total", tot, "queued first", st["queued_first"], flush=True)
    assert tot >= st["queued_first"] > 0
    dom.close()
    # no counters without the flag
    ctx.solve(None, K=96, P=60, V=96, precision=rpd.PREC_FP64)
    with pytest.raises(rpd.RpdError):
        ctx.debug_counters(0)
