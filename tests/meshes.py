"""Closed, consistently oriented triangle meshes for the mesh-grid tests."""
import numpy as np


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def signed_volume(pts, tris):
    a, b, c = pts[tris[:, 0]], pts[tris[:, 1]], pts[tris[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def outward(pts, tris):
    """Triangles oriented so that the signed volume is positive (normals out)."""
    tris = tris.copy()
    if signed_volume(pts, tris) < 0:
        tris[:, [0, 1]] = tris[:, [1, 0]]
    return tris


def icosphere(subdiv, radius=1.0, centre=(0.0, 0.0, 0.0), rot_seed=None):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    pts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = pts[a] + pts[b]
                pts.append(m / np.linalg.norm(m))
                cache[key] = len(pts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    P = np.array(pts) * radius
    if rot_seed is not None:
        P = P @ _rotation(rot_seed).T
    P = P + np.asarray(centre, float)
    T = np.array(f, np.int32)
    return P, outward(P, T)


def torus(R, r, nu, nv, rot_seed=None):
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    P = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1).reshape(-1, 3)
    idx = lambda i, j: (i % nu) * nv + (j % nv)  # noqa: E731
    T = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            T += [(a, b, c), (a, c, d)]
    if rot_seed is not None:
        P = P @ _rotation(rot_seed).T
    T = np.array(T, np.int32)
    return P, outward(P, T)


def voxel_union(rpd, mask):
    """Boundary mesh of a voxel union (the library's voxel grid writer), in
    the [0,1000]^3 frame of the mask."""
    gr = rpd.Grid.voxels(mask)
    return gr.points[:, :3].copy(), gr.triangles.reshape(-1, 3).copy()


def write_obj(path, pts, tris, quads=None):
    with open(path, "w") as f:
        for p in pts:
            f.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in tris:
            f.write("f %d %d %d\n" % tuple(int(x) + 1 for x in t))
        for q in quads or ():
            f.write("f " + " ".join("%d/%d" % (int(x) + 1, int(x) + 1) for x in q) + "\n")
