"""The oracle's mesh-to-grid generator (utils/createGrid/main.cpp:150-294),
pinned without its own formulas:

- triangle/box overlap (geometry.cpp:433-508): against polygon clipping of
  the triangle by the box grown and shrunk by a margin, on random cases that
  include ones only the nine edge axes can separate; closed boxes (a triangle
  touching a voxel face lies in both voxels, a flat one on the face only in
  the voxel the range floor(min)..ceil(max) reaches);
- point/triangle distance: against a dense barycentric sampling;
- winding number: against the half-space test of a convex polyhedron and
  the voxel membership of a non-convex voxel union;
- normalisation (main.cpp:164-175): unused points dropped, bounding box to
  [4.998, 4.998 + 990.222];
- lists: a triangle inside one voxel lists only it.

CPU only.
"""
import numpy as np
import pytest

from meshes import icosphere, voxel_union


def _clip_nonempty(tri, lo, hi):
    poly = [np.asarray(p, float) for p in tri]
    for d in range(3):
        for side, bound in ((1.0, hi[d]), (-1.0, lo[d])):
            out = []
            n = len(poly)
            for i in range(n):
                a, b = poly[i], poly[(i + 1) % n]
                fa, fb = side * (a[d] - bound), side * (b[d] - bound)
                if fa <= 0:
                    out.append(a)
                if (fa < 0 < fb) or (fb < 0 < fa):
                    out.append(a + (b - a) * (fa / (fa - fb)))
            poly = out
            if not poly:
                return False
    return True


def test_tri_box_overlap_against_clipping(orc):
    rng = np.random.default_rng(3)
    c = np.array([12.5, 37.5, 62.5], np.float32)
    h = np.float32(12.5)
    decided = edge_only = 0
    for _ in range(6000):
        s = rng.uniform(5, 40)
        tri = c.astype(float) + rng.uniform(-1, 1, size=(3, 3)) * s + rng.uniform(-20, 20, size=3)
        got = orc.tri_box_overlap(c, h, *tri)
        cd = c.astype(float)
        if _clip_nonempty(tri, cd - (float(h) - 1e-3), cd + (float(h) - 1e-3)):
            assert got
            decided += 1
        elif not _clip_nonempty(tri, cd - (float(h) + 1e-3), cd + (float(h) + 1e-3)):
            assert not got
            decided += 1
            # bounding boxes overlap and the plane cuts the box: only an edge
            # axis separates them
            bb = (tri.min(0) <= cd + float(h)).all() and (tri.max(0) >= cd - float(h)).all()
            n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
            corners = cd + float(h) * np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)])
            sd = (corners - tri[0]) @ n
            edge_only += bool(bb and sd.min() < 0 < sd.max())
    assert decided > 5900 and edge_only > 20, (decided, edge_only)


def test_closed_boxes_share_a_face_triangle(orc):
    g = 10
    # an edge on the face z = 500 between voxels k = 4 and 5: both list it
    offs, lst = orc.mesh_lists(np.array([[310.0, 420.0, 500.0], [380.0, 420.0, 500.0], [310.0, 480.0, 460.0]]),
                               np.array([0, 1, 2], np.int32), g)
    holders = sorted(int(v) for v in np.nonzero(np.diff(offs))[0])
    assert holders == [3 + 4 * g + 4 * g * g, 3 + 4 * g + 5 * g * g]
    # a flat triangle on that face: the range floor(min)..ceil(max) of
    # main.cpp:190-200 is k = 5 only, so only voxel 5 lists it
    offs, lst = orc.mesh_lists(np.array([[310.0, 420.0, 500.0], [380.0, 420.0, 500.0], [310.0, 480.0, 500.0]]),
                               np.array([0, 1, 2], np.int32), g)
    assert [int(v) for v in np.nonzero(np.diff(offs))[0]] == [3 + 4 * g + 5 * g * g]


def test_triangle_inside_one_voxel(orc):
    g = 10
    offs, lst = orc.mesh_lists(np.array([[310.0, 420.0, 530.0], [380.0, 430.0, 560.0], [340.0, 480.0, 590.0]]),
                               np.array([0, 1, 2], np.int32), g)
    assert list(np.nonzero(np.diff(offs))[0]) == [3 + 4 * g + 5 * g * g] and list(lst) == [0]


def test_point_triangle_distance_against_sampling(orc):
    rng = np.random.default_rng(5)
    m = 300
    s, t = np.meshgrid(np.linspace(0, 1, m + 1), np.linspace(0, 1, m + 1), indexing="ij")
    keep = (s + t) <= 1
    s, t = s[keep], t[keep]
    for _ in range(40):
        a, b, c = rng.uniform(-1, 1, size=(3, 3))
        p = rng.uniform(-2, 2, size=3)
        samples = a + s[:, None] * (b - a) + t[:, None] * (c - a)
        best = np.sqrt(((samples - p) ** 2).sum(1)).min()
        d = orc.point_triangle_distance(p, a, b, c)
        spacing = 2.0 * max(np.linalg.norm(b - a), np.linalg.norm(c - a)) / m
        assert best - spacing <= d <= best + 1e-12


def test_winding_number_convex(orc):
    P, T = icosphere(2, radius=1.0, rot_seed=1)
    rng = np.random.default_rng(7)
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    checked = 0
    for p in rng.uniform(-1.3, 1.3, size=(1500, 3)):
        sd = np.einsum("ij,ij->i", p - P[T[:, 0]], n)
        if np.abs(sd).min() < 1e-6:
            continue
        inside = (sd < 0).all()
        w = orc.winding_number(p, P, T)
        assert abs(w - (1.0 if inside else 0.0)) < 1e-9, (p, w, inside)
        checked += 1
    assert checked > 1400


def test_winding_number_voxel_union(rpd, orc):
    g0 = 10
    mask = rpd.voxel_mask(rpd.SHAPE_TORUS, g0, (0.5, 0.5, 0.5, 0.3, 0.12))
    P, T = voxel_union(rpd, mask)
    vs = 1000.0 / g0
    rng = np.random.default_rng(9)
    checked = 0
    for p in rng.uniform(0, 1000, size=(1500, 3)):
        f = p / vs
        if np.abs(f - np.round(f)).min() < 1e-6:
            continue
        i, j, k = (int(x) for x in np.floor(f))
        inside = bool(mask[i + j * g0 + k * g0 * g0])
        assert (abs(orc.winding_number(p, P, T)) > 0.5) == inside
        checked += 1
    assert checked > 1400


def test_normalisation_drops_unused_points(orc):
    P, T = icosphere(1, radius=3.0, centre=(5.0, -2.0, 1.0), rot_seed=2)
    P = np.vstack([P[:7], [[1e4, 1e4, 1e4]], P[7:]])
    T = np.where(T >= 7, T + 1, T)
    xyz, tris = orc.mesh_points(P, T)
    assert xyz.shape[0] == P.shape[0] - 1
    assert np.array_equal(tris.reshape(-1, 3), np.where(T > 7, T - 1, T))
    lo, hi = xyz.min(0), xyz.max(0)
    assert (lo == 4.998).all()
    assert abs(hi.max() - (990.222 + 4.998)) < 1e-9
    # the scale is one for all axes: extents keep their ratios
    ext = P[np.arange(P.shape[0]) != 7].max(0) - P[np.arange(P.shape[0]) != 7].min(0)
    assert np.allclose((hi - lo) / (hi - lo).max(), ext / ext.max(), rtol=1e-12)


def test_mesh_flags_convex(orc):
    # corners of a convex mesh: 2 exactly where the half-space test says
    # inside by more than the 1e-3 band (the distance to the surface of an
    # interior point is its smallest plane distance)
    P, T = icosphere(2, radius=1.0, rot_seed=4)
    xyz, tris = orc.mesh_points(P, T)
    T3 = tris.reshape(-1, 3)
    g = 12
    flags = orc.mesh_corner_flags(xyz, tris, g)
    n = np.cross(xyz[T3[:, 1]] - xyz[T3[:, 0]], xyz[T3[:, 2]] - xyz[T3[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    vs = 1000.0 / g
    idx = np.arange(g ** 3)
    C = np.stack([idx % g, (idx // g) % g, idx // (g * g)], 1) * vs
    sd = (C[:, None, :] - xyz[T3[:, 0]][None]) @ np.eye(3)
    sd = np.einsum("cti,ti->ct", sd, n)
    inside = (sd < 0).all(1)
    depth = -sd.max(1)
    assert np.array_equal(flags == 2, inside & (depth > 1e-3))
    assert ((flags == 2).sum() > 100) and ((flags == 0).sum() > 100)
