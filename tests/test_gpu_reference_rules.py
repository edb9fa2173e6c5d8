"""The reference's own rules where they differ from the library's defaults,
through the HIP path, against the oracle run with the same rules.

- apex_mode 0: the cone apex at the first voxel corner of the cell's box whose
  flag is not 1, with the +1-layer fallback (ConvexCell.cl:950-1056).  On
  mesh domains it is sound: every cell succeeds and the results equal the
  oracle's (and apex_mode 1's).  On voxel domains a corner apex lies in the
  planes of the voxel faces it meets, so some cones are flat: the reference
  rule ends those cells with inconsistent_boundary or NaN moments, in the
  oracle as on the GPU.  The cells it completes must agree per cell.
- border seeds (w <= NO_CELL_LIMITS, ConvexCell.cl:324-325, 873-876): copied
  to the output and acting as walls (the plane through the border seed) for
  their neighbours, in both forms; also as the next solve of a compacted
  result, where the last bordering ring comes back as walls.
- the reference stop test (security_mode 1) in the mixed form is in
  test_gpu_solve.py (test_power_torus).
"""
import numpy as np
import pytest

from conftest import cell_errors, moment_errors, oracle_grid
from meshes import icosphere, torus

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", ["torus", "icosphere"])
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_reference_apex_on_mesh_domains(rpd, orc, ctx, mesh, prec, tol):
    P, T = torus(3.0, 1.1, 64, 32, rot_seed=12) if mesh == "torus" else icosphere(3, radius=2.0, rot_seed=5)
    grid = rpd.Grid.mesh(P, T, 32)
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=17, wmax=0.25 * h * h * 1e6)
    og = oracle_grid(orc, grid)
    ref, ref_st, _, _ = orc.solve(x, og, K=64, P=40, V=60, T=100, apex_mode=0, frame=1)
    ref1, _, _, _ = orc.solve(x, og, K=64, P=40, V=60, T=100, apex_mode=1, frame=1)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec, apex_mode=0)
    got, status = ctx.moments(), ctx.status()
    dom.close()
    ev, em = moment_errors(got, ref, n)
    cv, cb, nc = cell_errors(got, ref, n)
    print("apex 0 on", mesh, prec, st["status_hist"], ev, em, cv, cb, nc, flush=True)
    # the reference rule is sound on a mesh: it and apex_mode 1 agree
    assert set(np.unique(ref_st)) <= {4, 6}
    assert max(moment_errors(ref1, ref, n)) <= 1e-12
    diff = status != ref_st
    assert set(zip(status[diff], ref_st[diff])) <= {(4, 6)} and (got[diff] == 0).all()
    assert ev <= tol and em <= tol
    assert cv <= tol and cb <= tol and nc > 1000
    vol = grid.volume()
    assert abs(got[:, 3].sum() - vol) / vol < (1e-11 if prec == 0 else 1e-7)


@pytest.mark.parametrize("shape,g,prm", [(0, 32, (0.5, 0.5, 0.5, 0.45)), (1, 40, (0.5, 0.5, 0.5, 0.3, 0.12))])
def test_reference_apex_on_voxel_domains(rpd, orc, ctx, shape, g, prm):
    """FP64 form.  The reference rule fails on a few percent of the cells of
    a voxel domain, in the oracle and on the GPU; which cells fail depends
    on how each arithmetic resolves the flat cones.  The cells both
    complete (success, finite moments) agree per cell, and the cells the
    GPU completes equal the oracle's with apex_mode 1."""
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=2, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(shape, g, prm)
    ref, ref_st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, apex_mode=0, frame=1)
    ref1, ref1_st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, apex_mode=1, frame=1)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=rpd.PREC_FP64, apex_mode=0, cull=0)
    got, status = ctx.moments(), ctx.status()
    dom.close()
    ok_ref = (ref_st == 4) & np.isfinite(ref).all(1)
    ok_got = (status == 4) & np.isfinite(got).all(1)
    both = ok_ref & ok_got
    fail_ref, fail_got = int((~ok_ref).sum()), int((~ok_got).sum())
    only = int((ok_ref != ok_got).sum())
    print("apex 0 on voxels", shape, g, "oracle fails", fail_ref, "gpu fails", fail_got, "differ", only,
          st["status_hist"], flush=True)
    # the reference rule fails on a few percent of the cells of a voxel domain
    assert 0 < fail_ref < n // 10 and 0 < fail_got < n // 10
    cv, cb, nc = cell_errors(got[both], ref[both], n)
    ev, em = moment_errors(got[both], ref[both], n)
    print("  completed by both: %d, per cell %.2e %.2e, of mean %.2e %.2e" % (nc, cv, cb, ev, em), flush=True)
    assert ev <= 1e-12 and em <= 1e-12
    assert cv <= 1e-12 and cb <= 1e-12
    # what the GPU completes is the diagram (any valid apex gives it)
    assert (ref1_st == 4).all() or set(np.unique(ref1_st)) <= {4, 6}
    c1v, c1b, n1 = cell_errors(got[ok_got], ref1[ok_got], n)
    e1v, e1m = moment_errors(got[ok_got], ref1[ok_got], n)
    print("  GPU-completed vs apex_mode 1: %d, per cell %.2e %.2e, of mean %.2e %.2e" % (n1, c1v, c1b, e1v, e1m),
          flush=True)
    assert e1v <= 1e-12 and e1m <= 1e-12 and c1v <= 1e-12 and c1b <= 1e-12


def _walls(rpd, n, seed, frac):
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=seed, wmax=0.25 * h * h * 1e6)
    rng = np.random.default_rng(seed)
    w = rng.choice(n, int(frac * n), replace=False)
    x[w, 3] = rpd.NO_CELL_LIMITS
    return x, w


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_border_seeds_as_walls(rpd, orc, ctx, prec, tol):
    n = 20000
    x, w = _walls(rpd, n, 19, 0.03)
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 40, (0.5, 0.5, 0.5, 0.3, 0.12))
    ref, ref_st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, frame=1)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec)
    got, status = ctx.moments(), ctx.status()
    dom.close()
    ev, em = moment_errors(got, ref, n)
    cv, cb, nc = cell_errors(got, ref, n)
    print("walls", prec, st["status_hist"], st["skipped_cells"], ev, em, cv, cb, nc, flush=True)
    # border seeds are copied, the cover test is off with walls present
    assert np.array_equal(got[w], x[w])
    assert st["skipped_cells"] == 0
    diff = status != ref_st
    assert set(zip(status[diff], ref_st[diff])) <= {(4, 6)} and (got[diff] == 0).all()
    assert ev <= tol and em <= tol
    assert cv <= tol and cb <= tol and nc > 1000


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_solve_after_compaction_with_walls(rpd, orc, ctx, prec, tol):
    """A Lloyd-like step on a compacted result: the used cells and one ring
    of their face neighbours (rpd_compact_used, keep_distance 1) are solved
    again, the ring's empty cells as walls at their seeds."""
    n = 20000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=29, wmax=0.25 * h * h * 1e6)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 40, (0.5, 0.5, 0.5, 0.3))
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec)
    keep = 1e-3 * 1e9 / n
    out, ids = ctx.compact_used(keep, 1)
    walls = out[:, 3] == rpd.NO_CELL_LIMITS
    assert 0 < walls.sum() < ids.size
    y = x[ids].copy()
    y[walls, 3] = rpd.NO_CELL_LIMITS
    m = y.shape[0]
    ref, ref_st, _, _ = orc.solve(y, oracle_grid(orc, grid), K=64, P=40, V=60, T=100, frame=1)
    ctx.set_seeds(y)
    st = ctx.solve(dom, K=64, P=40, V=60, T=100, precision=prec)
    got = ctx.moments()
    dom.close()
    ev, em = moment_errors(got, ref, m)
    cv, cb, nc = cell_errors(got, ref, m)
    print("compacted + walls", prec, m, int(walls.sum()), st["status_hist"], ev, em, cv, cb, nc, flush=True)
    assert np.array_equal(got[walls], y[walls])
    assert ev <= tol and em <= tol
    assert cv <= tol and cb <= tol
    # walls only take space away from the cells of the used seeds
    assert got[~walls, 3].sum() <= grid.volume() * (1 + 1e-11)
