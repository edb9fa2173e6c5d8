"""ctypes binding of the CPU oracle (librpd_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, where it is the checker; the product
library never uses it.  The oracle restates the reference algorithm
(knearests.cpp/knearests.cl, ConvexCell.cl, voronoi.cpp pass logic) on the
CPU; see rpd_oracle.cpp for the file:line map.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librpd_oracle.so")

STATUS_NAMES = [
    "vertex_overflow", "plane_overflow", "inconsistent_boundary",
    "security_radius_not_reached", "success", "needs_exact_predicates",
    "empty_cell", "find_another_beginning_vertex", "triangle_overflow",
    "seed_on_border", "hessian_overflow",
]
SUCCESS = 4


class OrcGrid(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int),
        ("in_domain", ctypes.c_void_p),
        ("points", ctypes.c_void_p),
        ("num_points", ctypes.c_int),
        ("triangles", ctypes.c_void_p),
        ("num_triangles", ctypes.c_int),
        ("tri_list", ctypes.c_void_p),
        ("num_tri_list", ctypes.c_int),
        ("offsets", ctypes.c_void_p),
        ("centre", ctypes.c_void_p),
    ]


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "K", "P", "V", "T", "precision", "vhs0", "vhs1", "apex_mode",
        "fix_exact_requeue", "max_extra_steps", "threads", "single_pass", "security_mode", "cull", "H", "HC",
        "frame")]


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("n_pred", ctypes.c_longlong),
        ("n_radius_vertex", ctypes.c_longlong),
        ("n_planes", ctypes.c_longlong),
        ("n_bbox_vertex", ctypes.c_longlong),
        ("n_integrated_vertex", ctypes.c_longlong),
        ("n_clip", ctypes.c_longlong),
        ("status_hist", ctypes.c_int * 11),
        ("passes", ctypes.c_int),
        ("final_K", ctypes.c_int),
        ("final_P", ctypes.c_int),
        ("final_V", ctypes.c_int),
        ("final_T", ctypes.c_int),
        ("seconds_knn", ctypes.c_double),
        ("seconds_clip", ctypes.c_double),
        ("seconds_total", ctypes.c_double),
        ("n_pred_restrict", ctypes.c_longlong),
        ("n_planes_restrict", ctypes.c_longlong),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "status_hist"}
        d["status_hist"] = list(self.status_hist)
        return d


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_knn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_knn_dim.argtypes = [ctypes.c_int]
        L.orc_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcParams),
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(OrcStats)]
        L.orc_solve_hessian.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcGrid),
                                        ctypes.POINTER(OrcParams), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(OrcStats), ctypes.c_void_p, ctypes.c_void_p]
        L.orc_cell_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.orc_solve_ids.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcParams), ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(OrcStats)]
        L.orc_clip_cell.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(OrcGrid), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_grid_volume.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_grid_volume.restype = ctypes.c_double
        L.orc_grid_valid.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_grid_moments.argtypes = [ctypes.POINTER(OrcGrid), ctypes.c_void_p]
        L.orc_tri_box_overlap.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
        L.orc_point_triangle_distance.argtypes = [ctypes.c_void_p] * 4
        L.orc_point_triangle_distance.restype = ctypes.c_double
        L.orc_winding_number.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_winding_number.restype = ctypes.c_double
        L.orc_mesh_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]
        L.orc_mesh_lists.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_longlong]
        L.orc_mesh_lists.restype = ctypes.c_longlong
        L.orc_mesh_corner_flags.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Grid:
    """Numpy arrays of a domain grid (the content of the reference's Grid)."""

    def __init__(self, size, in_domain, points, triangles, tri_list, offsets, centre=None):
        self.size = int(size)
        self.in_domain = np.ascontiguousarray(in_domain, dtype=np.int8)
        self.points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 4)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1)
        self.tri_list = np.ascontiguousarray(tri_list, dtype=np.int32)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.centre = None if centre is None else np.ascontiguousarray(centre, dtype=np.int8)

    def orc(self):
        g = OrcGrid()
        g.size = self.size
        g.in_domain = _ptr(self.in_domain)
        g.points = _ptr(self.points)
        g.num_points = self.points.shape[0]
        g.triangles = _ptr(self.triangles)
        g.num_triangles = self.triangles.size // 3
        g.tri_list = _ptr(self.tri_list)
        g.num_tri_list = self.tri_list.size
        g.offsets = _ptr(self.offsets)
        g.centre = _ptr(self.centre)
        return g


def knn_dim(n):
    return lib().orc_knn_dim(int(n))


def knn(xyzw, K, precision=1, method=0):
    """Returns (perm, knn[n,K] stored ids, counters)."""
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    n = xyzw.shape[0]
    dim = knn_dim(n)
    perm = np.empty(n, np.int32)
    out = np.empty((n, K), np.int32)
    counters = np.empty(dim ** 3 + 1, np.int32)
    r = lib().orc_knn(_ptr(xyzw), n, K, precision, method, _ptr(perm), _ptr(out), _ptr(counters))
    if r < 0:
        raise RuntimeError("orc_knn failed %d" % r)
    return perm, out, counters


def params(K=25, P=35, V=40, T=100, precision=1, vhs0=2, vhs1=2, apex_mode=1, fix_exact_requeue=1,
           max_extra_steps=5, threads=0, single_pass=0, security_mode=0, cull=0, H=0, HC=1, frame=0):
    """Oracle parameters.  precision 0 float build, 1 double (USE_DOUBLE), 2
    long double (a higher-precision evaluation of the same algorithm); frame 0
    the reference's [0,1000]^3 frame, 1 each cell in its seed's frame."""
    p = OrcParams()
    for k, v in dict(K=K, P=P, V=V, T=T, precision=precision, vhs0=vhs0, vhs1=vhs1, apex_mode=apex_mode,
                     fix_exact_requeue=fix_exact_requeue, max_extra_steps=max_extra_steps, threads=threads,
                     single_pass=single_pass, security_mode=security_mode, cull=cull, H=H, HC=HC,
                     frame=frame).items():
        setattr(p, k, int(v))
    return p


def solve(xyzw, grid=None, **kw):
    """Restricted power diagram; returns (mg[n,4] per original seed, status[n], kused[n], stats dict)."""
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    n = xyzw.shape[0]
    out = np.zeros((n, 4), np.float64)
    status = np.zeros(n, np.int32)
    kused = np.zeros(n, np.int32)
    st = OrcStats()
    p = params(**kw)
    g = grid.orc() if grid is not None else None
    r = lib().orc_solve(_ptr(xyzw), n, ctypes.byref(g) if g is not None else None, ctypes.byref(p), _ptr(out),
                        _ptr(status), _ptr(kused), ctypes.byref(st))
    if r < 0:
        raise RuntimeError("orc_solve failed %d" % r)
    return out, status, kused, st.as_dict()


def solve_hessian(xyzw, grid=None, H=32, HC=1, **kw):
    """solve() plus the Hessian (ConvexCell.cl:621-674, 1133-1151): per original
    seed up to H entries, hess[n, H, HC] (HC = 1: face area; 4: area-weighted
    face centroid xyz, then area) and ids[n, H] (neighbour original index,
    -1 after the last).  Returns (mg, status, kused, hess, ids, stats)."""
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    n = xyzw.shape[0]
    out = np.zeros((n, 4), np.float64)
    status = np.zeros(n, np.int32)
    kused = np.zeros(n, np.int32)
    hess = np.zeros((n, H, HC), np.float64)
    ids = np.full((n, H), -1, np.int32)
    st = OrcStats()
    p = params(H=H, HC=HC, **kw)
    g = grid.orc() if grid is not None else None
    r = lib().orc_solve_hessian(_ptr(xyzw), n, ctypes.byref(g) if g is not None else None, ctypes.byref(p),
                                _ptr(out), _ptr(status), _ptr(kused), ctypes.byref(st), _ptr(hess), _ptr(ids))
    if r < 0:
        raise RuntimeError("orc_solve_hessian failed %d" % r)
    return out, status, kused, hess, ids, st.as_dict()


def cell_neighbours(xyzw, ids, K=64, P=60, V=100, security_mode=0, max_nb=128):
    """Face neighbours of the unrestricted power cells of the seeds `ids`
    (original indices, -1 padded): [len(ids), max_nb]."""
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.full((ids.size, max_nb), -1, np.int32)
    if ids.size == 0:
        return out
    r = lib().orc_cell_neighbours(_ptr(xyzw), xyzw.shape[0], _ptr(ids), ids.size, int(K), int(P), int(V),
                                  int(security_mode), _ptr(out), int(max_nb))
    if r < 0:
        raise RuntimeError("orc_cell_neighbours failed %d" % r)
    return out


NO_CELL_LIMITS = -100000.0


def compact_used(xyzw, mg, keep_volume, keep_distance, **kw):
    """Used cells and bordering rings (compressResult with m_keepDistance,
    voronoi.cpp:513-691, 693-751, 1381-1393): the cells with m >= keep_volume,
    then keep_distance rings of face neighbours; ring k expands the cells of
    ring k whose volume is below keep_volume.  A cell of the last ring without
    volume gets w = NO_CELL_LIMITS.  Returns (out[count, 4], ids[count]) in
    ascending original index."""
    n = mg.shape[0]
    if keep_distance < 0:
        return mg.copy(), np.arange(n, dtype=np.int32)
    used = (mg[:, 3] >= keep_volume).astype(np.int32)
    for k in range(keep_distance):
        if k == 0:
            sel = np.nonzero((used == 1) & (mg[:, 3] >= keep_volume))[0]
        else:
            sel = np.nonzero((used == k) & (mg[:, 3] < keep_volume))[0]
        if sel.size == 0:
            break
        nb = cell_neighbours(xyzw, sel, **kw)
        nb = np.unique(nb[nb >= 0])
        nb = nb[used[nb] == 0]
        used[nb] = k + 1
    keep = np.nonzero(used != 0)[0].astype(np.int32)
    out = mg[keep].copy()
    flag = (mg[keep, 3] < keep_volume) & (used[keep] == keep_distance)
    out[flag, 3] = NO_CELL_LIMITS
    return out, keep


def solve_ids(xyzw, ids, grid=None, **kw):
    """Cells of the seeds `ids` only (every seed still a neighbour)."""
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.zeros((ids.size, 4), np.float64)
    status = np.zeros(ids.size, np.int32)
    kused = np.zeros(ids.size, np.int32)
    st = OrcStats()
    p = params(**kw)
    g = grid.orc() if grid is not None else None
    r = lib().orc_solve_ids(_ptr(xyzw), xyzw.shape[0], _ptr(ids), ids.size, ctypes.byref(g) if g is not None else None,
                            ctypes.byref(p), _ptr(out), _ptr(status), _ptr(kused), ctypes.byref(st))
    if r < 0:
        raise RuntimeError("orc_solve_ids failed %d" % r)
    return out, status, kused, st.as_dict()


def clip_cell(xyzw, seed, neigs, K, P=64, V=128, T=0, precision=1, vhs=2, grid=None, apex_mode=1, out_domain=0,
              security_mode=1, wmin=0.0):
    xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
    neigs = np.ascontiguousarray(neigs, dtype=np.int32)
    out = np.zeros(4, np.float64)
    nv = ctypes.c_int(0)
    npl = ctypes.c_int(0)
    g = grid.orc() if grid is not None else None
    st = lib().orc_clip_cell(_ptr(xyzw), xyzw.shape[0], int(seed), _ptr(neigs), int(K), int(P), int(V), int(T),
                             int(precision), int(vhs), ctypes.byref(g) if g is not None else None, int(apex_mode),
                             int(out_domain), int(security_mode), ctypes.c_double(wmin), _ptr(out), ctypes.byref(nv),
                             ctypes.byref(npl))
    return st, out, nv.value, npl.value


def grid_volume(grid):
    g = grid.orc()
    return lib().orc_grid_volume(ctypes.byref(g))


def grid_valid(grid):
    g = grid.orc()
    return bool(lib().orc_grid_valid(ctypes.byref(g)))


def grid_moments(grid):
    g = grid.orc()
    out = np.zeros(4, np.float64)
    lib().orc_grid_moments(ctypes.byref(g), _ptr(out))
    return out


# ---- mesh-to-grid generator (utils/createGrid/main.cpp:150-294) ----------

def tri_box_overlap(centre, half, a, b, c):
    """Single-precision triangle/box overlap of geometry.cpp:433-508."""
    ctr = np.ascontiguousarray(centre, dtype=np.float32)
    v = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c)]
    return bool(lib().orc_tri_box_overlap(_ptr(ctr), ctypes.c_float(half), _ptr(v[0]), _ptr(v[1]), _ptr(v[2])))


def point_triangle_distance(p, a, b, c):
    v = [np.ascontiguousarray(x, dtype=np.float64) for x in (p, a, b, c)]
    return lib().orc_point_triangle_distance(*[_ptr(x) for x in v])


def winding_number(p, xyz, tris):
    p = np.ascontiguousarray(p, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    return lib().orc_winding_number(_ptr(p), _ptr(xyz), _ptr(tris), tris.size // 3)


def mesh_points(points, tris, normalize=True):
    """Kept points (renumbered) and triangles, normalised as main.cpp:164-175."""
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
    xyz = np.zeros_like(points)
    out_t = np.zeros_like(tris)
    m = lib().orc_mesh_points(_ptr(points), points.shape[0], _ptr(tris), tris.size // 3, int(bool(normalize)),
                              _ptr(xyz), _ptr(out_t))
    if m < 0:
        raise RuntimeError("orc_mesh_points failed %d" % m)
    return xyz[:m].copy(), out_t


def mesh_lists(xyz, tris, g):
    """Per-voxel triangle lists of main.cpp:178-219: (offsets[g^3+1], list)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    nt = tris.size // 3
    offs = np.zeros(g ** 3 + 1, np.int32)
    total = lib().orc_mesh_lists(_ptr(xyz), _ptr(tris), nt, int(g), _ptr(offs), None, 0)
    lst = np.zeros(max(total, 1), np.int32)
    lib().orc_mesh_lists(_ptr(xyz), _ptr(tris), nt, int(g), _ptr(offs), _ptr(lst), total)
    return offs, lst[:total].copy()


def mesh_corner_flags(xyz, tris, g, eps=1e-3):
    """Corner flags by brute force: 1 within eps of the surface, else 2
    inside / 0 outside by the winding number."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    out = np.zeros(g ** 3, np.int8)
    lib().orc_mesh_corner_flags(_ptr(xyz), _ptr(tris), tris.size // 3, int(g), float(eps), _ptr(out))
    return out
