"""Brute-force convex polytope geometry for small cases, independent of the
oracle's vertex representation: a polytope is a list of faces, each an
ordered polygon of 3-D points.  Cells are built by clipping a box by every
power bisector (all N-1 other seeds, no kNN, no stop test); restriction to a
voxel domain clips the cell by each kept voxel's box.

Used only to pin the oracle (restricted volumes and first moments,
ConvexCell.cl:918-1116 and :577-619) on tiny inputs.
"""
import numpy as np


def box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float64)
    faces = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [0, 4, 7, 3], [1, 2, 6, 5]]
    return [[v[i] for i in f] for f in faces]


def clip(faces, n, d):
    """Keep {x : n.x + d <= 0}."""
    n = np.asarray(n, dtype=np.float64)
    sv = [float(n @ p + d) for f in faces for p in f]
    if max(sv) <= 0:
        return faces      # nothing cut (a face may lie in the plane)
    if min(sv) >= 0:
        return []         # nothing strictly inside
    out_faces, cap = [], []
    for f in faces:
        m = len(f)
        s = [float(n @ p + d) for p in f]
        out = []
        for i in range(m):
            a, b = f[i], f[(i + 1) % m]
            sa, sb = s[i], s[(i + 1) % m]
            if sa <= 0:
                out.append(a)
                if sa == 0:
                    cap.append(a)
            if (sa < 0 < sb) or (sb < 0 < sa):
                p = a + (sa / (sa - sb)) * (b - a)
                out.append(p)
                cap.append(p)
        if len(out) >= 3:
            out_faces.append(out)
    if len(cap) >= 3 and out_faces:
        c = np.mean(cap, axis=0)
        u = cap[0] - c
        if np.linalg.norm(u) == 0:
            u = next((q - c for q in cap if np.linalg.norm(q - c) > 0), u)
        u = u / (np.linalg.norm(u) or 1.0)
        w = np.cross(n, u)
        ang = [np.arctan2(float((q - c) @ w), float((q - c) @ u)) for q in cap]
        order = np.argsort(ang)
        poly = []
        for i in order:
            if all(np.linalg.norm(cap[i] - q) > 1e-9 for q in poly):
                poly.append(cap[i])
        if len(poly) >= 3:
            out_faces.append(poly)
    return out_faces


def volume_moments(faces):
    """(mx, my, mz, volume) of a convex polytope."""
    if not faces:
        return np.zeros(4)
    pts = np.array([p for f in faces for p in f])
    c = pts.mean(axis=0)
    vol = 0.0
    mom = np.zeros(3)
    for f in faces:
        for i in range(1, len(f) - 1):
            a, b, e = f[0] - c, f[i] - c, f[i + 1] - c
            t = abs(float(np.linalg.det(np.array([a, b, e])))) / 6.0
            vol += t
            mom += t * (c + (a + b + e) / 4.0)  # tet centroid
    return np.array([mom[0], mom[1], mom[2], vol])


def power_cell(xyzw, i, lo=(0.0, 0.0, 0.0), hi=(1000.0, 1000.0, 1000.0)):
    """Power cell {x : |x-s_i|^2 + w_i <= |x-s_j|^2 + w_j for all j} in a box."""
    s = xyzw[i, :3]
    wi = xyzw[i, 3]
    faces = box(lo, hi)
    for j in range(xyzw.shape[0]):
        if j == i:
            continue
        t = xyzw[j, :3]
        # 2 (t - s) . x <= |t|^2 - |s|^2 + w_j - w_i
        n = 2.0 * (t - s)
        d = -(t @ t - s @ s + xyzw[j, 3] - wi)
        faces = clip(faces, n, d)
        if not faces:
            break
    return faces


def restricted_cell(faces, mask, g):
    """Volume and first moments of a convex polytope intersected with the
    union of the kept voxels (mask[z, y, x] or flat index x + y g + z g^2)."""
    if not faces:
        return np.zeros(4)
    mask = np.asarray(mask).reshape(-1)
    vs = 1000.0 / g
    pts = np.array([p for f in faces for p in f])
    lo = np.clip(np.floor(pts.min(axis=0) / vs).astype(int), 0, g - 1)
    hi = np.clip(np.floor(pts.max(axis=0) / vs).astype(int), 0, g - 1)
    acc = np.zeros(4)
    for z in range(lo[2], hi[2] + 1):
        for y in range(lo[1], hi[1] + 1):
            for x in range(lo[0], hi[0] + 1):
                if not mask[x + y * g + z * g * g]:
                    continue
                f = faces
                for ax in range(3):
                    e = np.zeros(3)
                    e[ax] = 1.0
                    c0 = (x, y, z)[ax] * vs
                    f = clip(f, -e, c0)          # x_ax >= c0
                    if not f:
                        break
                    f = clip(f, e, -(c0 + vs))   # x_ax <= c0 + vs
                    if not f:
                        break
                if f:
                    acc += volume_moments(f)
    return acc
