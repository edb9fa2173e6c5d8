"""Mesh-to-grid generator on the GPU (rpd_grid_from_mesh, the mesh path of
utils/createGrid/main.cpp:150-294) against the oracle:

- normalised points and per-voxel triangle lists bit-exact (the overlap test
  is single precision with every rounding as geometry.cpp:433-508 does it);
- corner flags: every 0 / 2 flag equals the oracle's brute-force class
  (distance over all triangles, winding number), every oracle 1 is a 1; the
  corners left undecided (1 where the oracle says 0 / 2) are counted and
  must be a small minority;
- the grid is valid, its volume is the mesh volume, it survives save / load
  unchanged, and a restricted diagram on it matches the oracle's (FP64
  1e-11, mixed 1e-6 of the mean cell volume) with cells that tile it.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, moment_errors, oracle_grid
from meshes import icosphere, signed_volume, torus, voxel_union, write_obj

pytestmark = pytest.mark.gpu


def _meshes(rpd):
    mask = rpd.voxel_mask(rpd.SHAPE_TORUS, 14, (0.5, 0.5, 0.5, 0.3, 0.12))
    return {
        "icosphere": icosphere(3, radius=2.0, centre=(1.0, -3.0, 0.5), rot_seed=11),
        "torus": torus(3.0, 1.1, 48, 24, rot_seed=12),
        "voxel_union": voxel_union(rpd, mask),
    }


def _check_against_oracle(rpd, orc, P, T, g, normalize=True):
    grid = rpd.Grid.mesh(P, T, g, normalize=normalize)
    xyz, tris = orc.mesh_points(P, T, normalize=normalize)
    assert np.array_equal(grid.points[:, :3], xyz)
    t3 = tris.reshape(-1, 3)
    if signed_volume(xyz, t3) > 0:  # re-oriented as Grid::load does
        t3 = t3[:, [1, 0, 2]]
    assert np.array_equal(grid.triangles.reshape(-1, 3), t3)
    offs, lst = orc.mesh_lists(xyz, tris, g)
    assert np.array_equal(grid.offsets, offs)
    assert np.array_equal(grid.tri_list, lst)
    ref = orc.mesh_corner_flags(xyz, tris, g)
    got = grid.in_domain
    decided = got != 1
    assert np.array_equal(got[decided], ref[decided])
    assert (got[ref == 1] == 1).all()
    undecided = int(((got == 1) & (ref != 1)).sum())
    return grid, undecided


@pytest.mark.parametrize("name", ["icosphere", "torus", "voxel_union"])
@pytest.mark.parametrize("g", [24, 41])
def test_mesh_grid_matches_oracle(rpd, orc, name, g):
    P, T = _meshes(rpd)[name]
    grid, undecided = _check_against_oracle(rpd, orc, P, T, g)
    print("mesh grid", name, g, "entries", grid.tri_list.size, "undecided", undecided, flush=True)
    assert undecided <= max(2, g ** 3 // 1000)
    assert grid.valid()
    xyz = grid.points[:, :3]
    assert abs(grid.volume() - abs(signed_volume(xyz, grid.triangles.reshape(-1, 3)))) <= 1e-9 * grid.volume()
    assert grid.volume() > 0


def test_lattice_aligned_mesh_without_normalisation(rpd, orc):
    # a voxel union kept in its own frame: mesh vertices on grid corners and
    # faces in lattice planes.  Corners on the surface are 1; the others take
    # the class of a voxel centre through a segment in general position, so
    # the degenerate lattice does not leave them undecided
    mask = rpd.voxel_mask(rpd.SHAPE_SPHERE, 10, (0.5, 0.5, 0.5, 0.4))
    P, T = voxel_union(rpd, mask)
    grid, undecided = _check_against_oracle(rpd, orc, P, T, 20, normalize=False)
    print("lattice-aligned: undecided", undecided, "of", 20 ** 3, flush=True)
    assert undecided <= 20 ** 3 // 1000
    assert grid.valid()


def test_mesh_grid_save_load(rpd, tmp_path):
    P, T = _meshes(rpd)["torus"]
    grid = rpd.Grid.mesh(P, T, 32)
    path = str(tmp_path / "torus.grid")
    grid.save(path)
    back = rpd.Grid.load(path)
    for f in ("in_domain", "points", "triangles", "tri_list", "offsets", "centre"):
        assert np.array_equal(getattr(back, f), getattr(grid, f)), f


# FP64 bound 1e-11 here, 1e-12 on voxel domains: mesh triangles of any
# direction are cut by bisector planes at grazing angles, where the GPU's
# fused multiply-adds and the oracle's separately rounded products place the
# cut vertices a few ulps apart (measured 3.4e-12 volume, 2.6e-12 moment)
@pytest.mark.parametrize("prec,tol", [(0, 1e-11), (1, 1e-6)])
def test_diagram_on_mesh_grid(rpd, orc, ctx, prec, tol):
    P, T = _meshes(rpd)["torus"]
    grid = rpd.Grid.mesh(P, T, 40)
    n = 8000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=21)
    ref, ref_st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=50, P=32, V=48, T=4000)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    st = ctx.solve(dom, K=50, P=32, V=48, T=100, precision=prec)
    got, status = ctx.moments(), ctx.status()
    dom.close()
    ev, em = moment_errors(got, ref, n)
    print("diagram on mesh grid", prec, st["passes"], st["status_hist"], ev, em, flush=True)
    assert set(np.unique(ref_st)) <= {4, 6}
    assert np.array_equal(status, ref_st) or set(zip(status[status != ref_st], ref_st[status != ref_st])) <= {(4, 6)}
    assert ev <= tol and em <= tol
    vol = grid.volume()
    assert abs(got[:, 3].sum() - vol) / vol < (1e-10 if prec == 0 else 1e-7)


def test_create_grid_cli(rpd, tmp_path):
    # a cube with quad faces next to an icosphere: the command line splits
    # the quads into fans as the reader of the package does
    P, T = icosphere(2, radius=1.0, rot_seed=13)
    c = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], float) * 0.8 + [1.5, -0.4, -0.4]
    quads = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2], [1, 3, 7, 5]]
    pts = np.vstack([P, c])
    quads = [[q + P.shape[0] for q in f] for f in quads]
    mesh = str(tmp_path / "two.obj")
    write_obj(mesh, pts, T, quads)
    out_grid, out_mesh = str(tmp_path / "two.grid"), str(tmp_path / "two_scaled.off")
    exe = os.path.join(ROOT, "gpu-restricted-power-diagrams_amd", "rpd_create_grid")
    r = subprocess.run([exe, mesh, "--division=30", "--grid=" + out_grid, "--output=" + out_mesh],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stderr.strip(), flush=True)
    got = rpd.Grid.load(out_grid)
    rp, rt = rpd.read_mesh(mesh)
    want = rpd.Grid.mesh(rp, rt, 30)
    for f in ("in_domain", "points", "triangles", "tri_list", "offsets"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    sp, st_ = rpd.read_mesh(out_mesh)
    assert np.allclose(sp, want.points[:, :3], rtol=0, atol=1e-12)
    assert np.array_equal(st_.reshape(-1), want.triangles)
