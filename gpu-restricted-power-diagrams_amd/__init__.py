"""Python binding of librpd (the MI355X restricted power diagram hot path).

The product is the C-ABI shared library ``librpd.so`` built from ``csrc/``
(hand-written HIP kernels for gfx950 + host C++).  This module only loads it
with ctypes and mirrors the reference's ``VoroAlgo`` surface
(voronoi.h:105-256) so that callers and tests read like the reference's own
driver (test_voronoi.cpp:30-63).  Every compute call goes to librpd; when the
library or a GPU is missing the calls raise -- there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librpd.so")

# Status.h:4-16
STATUS_NAMES = [
    "vertex_overflow", "plane_overflow", "inconsistent_boundary",
    "security_radius_not_reached", "success", "needs_exact_predicates",
    "empty_cell", "find_another_beginning_vertex", "triangle_overflow",
    "seed_on_border", "hessian_overflow",
]
SUCCESS = 4
EMPTY_CELL = 6
NO_CELL_LIMITS = -100000.0
MAX_PASSES = 8  # RPD_MAX_PASSES

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP32 = 2

SHAPE_SPHERE = 0
SHAPE_TORUS = 1
SHAPE_CUBE = 2
SHAPE_CUBE_MINUS_SPHERE = 3

SEEDS_UNIFORM = 0
SEEDS_LATTICE = 1
SEEDS_BLUE_NOISE = 2


class RpdError(RuntimeError):
    pass


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int32),
        ("in_domain", ctypes.c_void_p),
        ("points", ctypes.c_void_p),
        ("num_points", ctypes.c_int32),
        ("triangles", ctypes.c_void_p),
        ("num_triangles", ctypes.c_int32),
        ("tri_list", ctypes.c_void_p),
        ("num_tri_list", ctypes.c_int32),
        ("offsets", ctypes.c_void_p),
        ("centre", ctypes.c_void_p),
    ]


# layout of rpd_read_debug_counters' output (rpd.h)
DCOUNT_MAX = 0
DCOUNT_HIST_K = 8
DCOUNT_HIST_P = DCOUNT_HIST_K + 257
DCOUNT_HIST_V = DCOUNT_HIST_P + 256
DCOUNT_HIST_T = DCOUNT_HIST_V + 257
DCOUNT_UNION_MAX = 4096
DCOUNT_INTS = DCOUNT_HIST_T + DCOUNT_UNION_MAX + 1


def read_device(ptr, dtype, count):
    """Copies count elements of dtype from a device pointer (as returned by
    the *_device entry points) to a new numpy array, through the HIP runtime
    librpd is linked with."""
    out = np.empty(int(count), dtype)
    if count:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [VP, VP, ctypes.c_size_t, ctypes.c_int]
        if hip.hipMemcpy(_p(out), VP(ptr), out.nbytes, 2) != 0:  # hipMemcpyDeviceToHost
            raise RpdError("hipMemcpy from the device failed")
    return out


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("K", "P", "V", "T", "precision", "max_extra_steps", "apex_mode", "security_mode", "cull", "H",
                 "HC", "debug_counters")]


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_bin", ctypes.c_double),
        ("ms_knn", ctypes.c_double),
        ("ms_clip", ctypes.c_double),
        ("ms_fallback", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_knn_first", ctypes.c_double),
        ("ms_clip_first", ctypes.c_double),
        ("ms_cells_first", ctypes.c_double),
        ("ms_finish_first", ctypes.c_double),
        ("status_hist", ctypes.c_int32 * 11),
        ("passes", ctypes.c_int32),
        ("fallback_cells", ctypes.c_int32),
        ("final_K", ctypes.c_int32),
        ("final_P", ctypes.c_int32),
        ("final_V", ctypes.c_int32),
        ("final_T", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("queued_first", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("pass_cells", ctypes.c_int32 * MAX_PASSES),
        ("pass_caps", (ctypes.c_int32 * 4) * MAX_PASSES),
        ("pass_ms_knn", ctypes.c_double * MAX_PASSES),
        ("pass_ms_clip", ctypes.c_double * MAX_PASSES),
        ("skipped_cells", ctypes.c_int64),
        ("ms_cover", ctypes.c_double),
    ]

    def as_dict(self):
        arrays = ("status_hist", "pass_cells", "pass_caps", "pass_ms_knn", "pass_ms_clip")
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in arrays}
        d["status_hist"] = list(self.status_hist)
        np_ = min(self.passes, MAX_PASSES)
        d["pass_cells"] = list(self.pass_cells)[:np_]
        d["pass_caps"] = [list(c) for c in self.pass_caps][:np_]
        d["pass_ms_knn"] = list(self.pass_ms_knn)[:np_]
        d["pass_ms_clip"] = list(self.pass_ms_clip)[:np_]
        return d


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
                ("max_batch", ctypes.c_int32)]


_lib = None
VP = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64


def lib():
    """Load librpd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RpdError("librpd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "rpd_voxel_mask": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, VP, VP]),
        "rpd_grid_from_voxels": (ctypes.c_int, [ctypes.c_int, VP, ctypes.POINTER(VP)]),
        "rpd_grid_from_mesh": (ctypes.c_int, [ctypes.c_int, VP, ctypes.c_int, VP, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(VP)]),
        "rpd_grid_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(VP)]),
        "rpd_grid_save": (ctypes.c_int, [VP, ctypes.c_char_p]),
        "rpd_grid_view": (ctypes.c_int, [VP, ctypes.POINTER(GridDesc)]),
        "rpd_grid_free": (None, [VP]),
        "rpd_grid_valid": (ctypes.c_int, [ctypes.POINTER(GridDesc)]),
        "rpd_grid_volume": (ctypes.c_double, [ctypes.POINTER(GridDesc)]),
        "rpd_grid_moments": (None, [ctypes.POINTER(GridDesc), VP]),
        "rpd_generate_seeds": (ctypes.c_int, [ctypes.c_int, I64, ctypes.c_uint64, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, VP]),
        "rpd_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
        "rpd_destroy": (None, [VP]),
        "rpd_last_error": (ctypes.c_char_p, [VP]),
        "rpd_default_params": (None, [ctypes.POINTER(Params)]),
        "rpd_domain_create": (ctypes.c_int, [VP, ctypes.POINTER(GridDesc), ctypes.POINTER(VP)]),
        "rpd_domain_destroy": (None, [VP]),
        "rpd_set_seeds": (ctypes.c_int, [VP, VP, I64]),
        "rpd_set_seeds_device": (ctypes.c_int, [VP, VP, I64]),
        "rpd_solve": (ctypes.c_int, [VP, VP, ctypes.POINTER(Params), ctypes.POINTER(Stats)]),
        "rpd_read_moments": (ctypes.c_int, [VP, VP]),
        "rpd_solve_batch": (ctypes.c_int, [VP, I32, VP, VP, VP, ctypes.POINTER(Params), ctypes.POINTER(Stats)]),
        "rpd_batch_range": (ctypes.c_int, [VP, I32, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "rpd_read_moments_range": (ctypes.c_int, [VP, I64, I64, VP]),
        "rpd_read_status": (ctypes.c_int, [VP, VP]),
        "rpd_read_perm": (ctypes.c_int, [VP, VP]),
        "rpd_moments_device": (ctypes.c_int, [VP, ctypes.POINTER(VP)]),
        "rpd_read_hessian": (ctypes.c_int, [VP, VP, VP]),
        "rpd_hessian_device": (ctypes.c_int, [VP, ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(I32),
                                              ctypes.POINTER(I32)]),
        "rpd_knn": (ctypes.c_int, [VP, I32, I32, VP]),
        "rpd_compact": (ctypes.c_int, [VP, ctypes.c_double, ctypes.POINTER(I64), VP, VP]),
        "rpd_compact_used": (ctypes.c_int, [VP, ctypes.c_double, I32, ctypes.POINTER(I64), VP, VP]),
        "rpd_compact_used_device": (ctypes.c_int, [VP, ctypes.c_double, I32, ctypes.POINTER(I64),
                                                   ctypes.POINTER(VP), ctypes.POINTER(VP)]),
        "rpd_read_debug_counters": (ctypes.c_int, [VP, I32, VP]),
        "rpd_server_create": (ctypes.c_int, [VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(VP)]),
        "rpd_server_destroy": (None, [VP]),
        "rpd_server_submit": (ctypes.c_int, [VP, VP, VP, I64, ctypes.POINTER(Params), ctypes.POINTER(I64)]),
        "rpd_server_poll": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(I64), VP, I64,
                                           ctypes.POINTER(ctypes.c_double)]),
        "rpd_server_drain": (ctypes.c_int, [VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            raise RpdError("librpd.so does not export %s; rebuild it" % name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = lib().rpd_last_error(ctx).decode() if ctx else ""
        raise RpdError("%s failed (%d) %s" % (what, rc, msg))


# ---------------------------------------------------------------------------
# host helpers: domains and seeds
# ---------------------------------------------------------------------------
def read_mesh(path):
    """Points [m, 3] and triangles [t, 3] of an ASCII .obj ("v x y z", "f a b
    c ..." with 1-based or negative ids, "a/b/c" forms) or .off file.  Polygons
    are split into fans about their first vertex, as geogram's
    facets.triangulate() (main.cpp:165) does."""
    pts, faces = [], []
    with open(path) as f:
        if path.lower().endswith(".off"):
            toks = [t for line in f for t in line.split("#")[0].split()]
            if not toks or toks[0] != "OFF":
                raise RpdError("%s: not an OFF file" % path)
            nv, nf = int(toks[1]), int(toks[2])
            k = 4
            for _ in range(nv):
                pts.append([float(x) for x in toks[k:k + 3]])
                k += 3
            for _ in range(nf):
                c = int(toks[k])
                faces.append([int(x) for x in toks[k + 1:k + 1 + c]])
                k += 1 + c
        else:
            for line in f:
                w = line.split()
                if not w:
                    continue
                if w[0] == "v":
                    pts.append([float(x) for x in w[1:4]])
                elif w[0] == "f":
                    ids = [int(x.split("/")[0]) for x in w[1:]]
                    faces.append([i - 1 if i > 0 else len(pts) + i for i in ids])
    tris = [[fc[0], fc[i], fc[i + 1]] for fc in faces for i in range(1, len(fc) - 1)]
    return np.asarray(pts, np.float64).reshape(-1, 3), np.asarray(tris, np.int32).reshape(-1, 3)


class Grid:
    """Host domain grid (the reference's Grid, Grid.h:14-70) owned by librpd."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        d = GridDesc()
        _check(lib().rpd_grid_view(self._h, ctypes.byref(d)), what="rpd_grid_view")
        self.desc = d
        g = d.size
        nv = g ** 3

        def arr(ptr, ctype, n, dtype):
            if n == 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (n,)).astype(dtype, copy=True)

        self.size = g
        self.in_domain = arr(d.in_domain, ctypes.c_int8, nv, np.int8)
        self.points = arr(d.points, ctypes.c_double, 4 * d.num_points, np.float64).reshape(-1, 4)
        self.triangles = arr(d.triangles, ctypes.c_int32, 3 * d.num_triangles, np.int32)
        self.tri_list = arr(d.tri_list, ctypes.c_int32, d.num_tri_list, np.int32)
        self.offsets = arr(d.offsets, ctypes.c_int32, nv + 1, np.int32)
        self.centre = arr(d.centre, ctypes.c_int8, nv, np.int8) if d.centre else None

    @classmethod
    def voxels(cls, mask):
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        g = int(round(mask.size ** (1.0 / 3.0)))
        assert g ** 3 == mask.size
        h = ctypes.c_void_p()
        _check(lib().rpd_grid_from_voxels(g, _p(mask), ctypes.byref(h)), what="rpd_grid_from_voxels")
        return cls(h.value)

    @classmethod
    def mesh(cls, points, triangles, g, normalize=True):
        """Grid of a closed triangle mesh built on the GPU (the mesh path of
        utils/createGrid/main.cpp:150-294): points [m, 3], triangles [t, 3]
        point ids; normalize scales into [4.998, 995.22]^3 as main.cpp:168-175."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        _check(lib().rpd_grid_from_mesh(int(g), _p(pts), pts.shape[0], _p(tri), tri.shape[0], int(bool(normalize)),
                                        ctypes.byref(h)), what="rpd_grid_from_mesh")
        return cls(h.value)

    @classmethod
    def shape(cls, shape, g, params=()):
        mask = voxel_mask(shape, g, params)
        return cls.voxels(mask)

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        _check(lib().rpd_grid_load(path.encode(), ctypes.byref(h)), what="rpd_grid_load")
        return cls(h.value)

    def save(self, path):
        _check(lib().rpd_grid_save(self._h, path.encode()), what="rpd_grid_save")

    def valid(self):
        return bool(lib().rpd_grid_valid(ctypes.byref(self.desc)))

    def volume(self):
        return lib().rpd_grid_volume(ctypes.byref(self.desc))

    def moments(self):
        out = np.zeros(4)
        lib().rpd_grid_moments(ctypes.byref(self.desc), _p(out))
        return out

    def __del__(self):
        try:
            if self._h:
                lib().rpd_grid_free(self._h)
                self._h = None
        except Exception:
            pass


def voxel_mask(shape, g, params=()):
    p = np.zeros(8, np.float64)
    p[: len(params)] = params
    out = np.zeros(g ** 3, np.uint8)
    _check(lib().rpd_voxel_mask(int(shape), int(g), _p(p), _p(out)), what="rpd_voxel_mask")
    return out


def generate_seeds(kind, n, seed=0, p0=0.0, p1=0.0, wmax=0.0):
    out = np.zeros((int(n), 4), np.float64)
    _check(lib().rpd_generate_seeds(int(kind), int(n), int(seed), float(p0), float(p1), float(wmax), _p(out)),
           what="rpd_generate_seeds")
    return out


def default_params():
    p = Params()
    lib().rpd_default_params(ctypes.byref(p))
    return p


# ---------------------------------------------------------------------------
# solver context
# ---------------------------------------------------------------------------
class Context:
    """One librpd context: one HIP device, one stream (rpd_create)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().rpd_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise RpdError("rpd_create(%d) failed (%d): no usable HIP device" % (device, rc))
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().rpd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what):
        _check(rc, self._h, what)

    def domain(self, grid):
        h = ctypes.c_void_p()
        self.check(lib().rpd_domain_create(self._h, ctypes.byref(grid.desc), ctypes.byref(h)), "rpd_domain_create")
        return Domain(self, h, grid)

    def set_seeds(self, xyzw):
        xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
        self.check(lib().rpd_set_seeds(self._h, _p(xyzw), xyzw.shape[0]), "rpd_set_seeds")
        self.n = xyzw.shape[0]

    def set_seeds_device(self, ptr, n):
        self.check(lib().rpd_set_seeds_device(self._h, ctypes.c_void_p(ptr), int(n)), "rpd_set_seeds_device")
        self.n = int(n)

    def solve(self, domain=None, K=25, P=35, V=40, T=100, precision=PREC_MIXED, max_extra_steps=5, apex_mode=1,
              security_mode=0, cull=1, H=0, HC=1, debug_counters=0):
        """rpd_solve; H > 0 also computes the Hessian (up to H neighbours per
        seed, HC = 1 face area or 4 face centroid times area and area);
        debug_counters keeps the per-pass counters (debug_counters())."""
        p = Params(int(K), int(P), int(V), int(T), int(precision), int(max_extra_steps), int(apex_mode),
                   int(security_mode), int(cull), int(H), int(HC), int(debug_counters))
        st = Stats()
        self.check(lib().rpd_solve(self._h, domain._h if domain is not None else None, ctypes.byref(p),
                                   ctypes.byref(st)), "rpd_solve")
        self._hess_shape = (int(H), int(HC)) if H > 0 else None
        return st.as_dict()

    def solve_batch(self, requests, K=25, P=35, V=40, T=100, precision=PREC_MIXED, max_extra_steps=5, apex_mode=1,
                    security_mode=0, cull=1, debug_counters=0):
        """rpd_solve_batch over requests [(domain or None, xyzw[n, 4]), ...]:
        one packed solve; returns (stats, [moments[n_r, 4] per request]).
        Afterwards moments() / status() hold every request's cells back to
        back."""
        xs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 4) for _, x in requests]
        nr = len(xs)
        doms = (VP * nr)(*[d._h.value if d is not None else None for d, _ in requests])
        ptrs = (VP * nr)(*[x.ctypes.data for x in xs])
        ns = (I64 * nr)(*[x.shape[0] for x in xs])
        p = Params(int(K), int(P), int(V), int(T), int(precision), int(max_extra_steps), int(apex_mode),
                   int(security_mode), int(cull), 0, 1, int(debug_counters))
        st = Stats()
        self.check(lib().rpd_solve_batch(self._h, nr, doms, ptrs, ns, ctypes.byref(p), ctypes.byref(st)),
                   "rpd_solve_batch")
        self.n = int(sum(x.shape[0] for x in xs))
        self._hess_shape = None
        outs = []
        for r in range(nr):
            first, count = I64(0), I64(0)
            self.check(lib().rpd_batch_range(self._h, r, ctypes.byref(first), ctypes.byref(count)), "rpd_batch_range")
            o = np.zeros((count.value, 4), np.float64)
            self.check(lib().rpd_read_moments_range(self._h, first.value, count.value, _p(o)),
                       "rpd_read_moments_range")
            outs.append(o)
        return st.as_dict(), outs

    def hessian(self):
        """(vals[n, H, HC], ids[n, H]) of the last solve with H > 0: per
        original seed the neighbours (original indices, -1 after the last)
        and their face values, in the cell's plane order."""
        H, HC = getattr(self, "_hess_shape", None) or (0, 1)
        if H <= 0:
            raise RpdError("hessian: the last solve had H = 0")
        vals = np.zeros((self.n, H, HC), np.float64)
        ids = np.zeros((self.n, H), np.int32)
        self.check(lib().rpd_read_hessian(self._h, _p(vals), _p(ids)), "rpd_read_hessian")
        return vals, ids

    def moments(self):
        out = np.zeros((self.n, 4), np.float64)
        self.check(lib().rpd_read_moments(self._h, _p(out)), "rpd_read_moments")
        return out

    def moments_device_ptr(self):
        p = ctypes.c_void_p()
        self.check(lib().rpd_moments_device(self._h, ctypes.byref(p)), "rpd_moments_device")
        return p.value

    def status(self):
        out = np.zeros(self.n, np.int32)
        self.check(lib().rpd_read_status(self._h, _p(out)), "rpd_read_status")
        return out

    def perm(self):
        out = np.zeros(self.n, np.int32)
        self.check(lib().rpd_read_perm(self._h, _p(out)), "rpd_read_perm")
        return out

    def knn(self, K, precision=1):
        out = np.zeros((self.n, int(K)), np.int32)
        self.check(lib().rpd_knn(self._h, int(K), int(precision), _p(out)), "rpd_knn")
        return out

    def compact(self, keep_volume):
        cnt = I64(0)
        self.check(lib().rpd_compact(self._h, float(keep_volume), ctypes.byref(cnt), None, None), "rpd_compact")
        out = np.zeros((cnt.value, 4), np.float64)
        perm = np.zeros(cnt.value, np.int32)
        self.check(lib().rpd_compact(self._h, float(keep_volume), ctypes.byref(cnt), _p(out), _p(perm)),
                   "rpd_compact")
        return out, perm

    def debug_counters(self, pass_index):
        """rpd_read_debug_counters: {"max": [K, P, V, T, H], "K": hist[257],
        "P": hist[256], "V": hist[257], "T": hist[4097]} of one pass."""
        out = np.zeros(DCOUNT_INTS, np.int32)
        self.check(lib().rpd_read_debug_counters(self._h, int(pass_index), _p(out)), "rpd_read_debug_counters")
        return {"max": out[DCOUNT_MAX:DCOUNT_MAX + 5].copy(), "K": out[DCOUNT_HIST_K:DCOUNT_HIST_P].copy(),
                "P": out[DCOUNT_HIST_P:DCOUNT_HIST_V].copy(), "V": out[DCOUNT_HIST_V:DCOUNT_HIST_T].copy(),
                "T": out[DCOUNT_HIST_T:].copy()}

    def compact_used_device(self, keep_volume, keep_distance):
        """rpd_compact_used_device: the compaction left in HBM.  Returns
        (count, device pointer of out[count, 4] doubles, device pointer of
        perm_to_user[count] int32), valid until the next compaction / solve."""
        cnt = I64(0)
        po, pp = VP(), VP()
        self.check(lib().rpd_compact_used_device(self._h, float(keep_volume), int(keep_distance), ctypes.byref(cnt),
                                                 ctypes.byref(po), ctypes.byref(pp)), "rpd_compact_used_device")
        return cnt.value, po.value, pp.value

    def compact_used(self, keep_volume, keep_distance):
        """rpd_compact_used: the used cells (m >= keep_volume) and keep_distance
        rings of face neighbours around them; last-ring cells without volume
        get w = NO_CELL_LIMITS.  Returns (out[count, 4], perm_to_user[count])."""
        cnt = I64(0)
        out = np.zeros((self.n, 4), np.float64)
        perm = np.zeros(self.n, np.int32)
        self.check(lib().rpd_compact_used(self._h, float(keep_volume), int(keep_distance), ctypes.byref(cnt), _p(out),
                                          _p(perm)), "rpd_compact_used")
        return out[:cnt.value].copy(), perm[:cnt.value].copy()


def params(K=25, P=35, V=40, T=100, precision=PREC_MIXED, max_extra_steps=5, apex_mode=1, security_mode=0, cull=1,
           H=0, HC=1):
    return Params(int(K), int(P), int(V), int(T), int(precision), int(max_extra_steps), int(apex_mode),
                  int(security_mode), int(cull), int(H), int(HC))


class Server:
    """Request loop for streams of solves (rpd_server_*): submit requests of
    any size against resident domains, poll finished ones with their latency.
    Queued requests run as packed batches (one launch per kernel for all of
    them), max_inflight batches at once, at most max_batch requests each."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_inflight=2, max_batch=256):
        h = ctypes.c_void_p()
        cfg = ServerConfig(int(hbm_budget_bytes), int(max_inflight), int(max_batch))
        _check(lib().rpd_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)), None, "rpd_server_create")
        self._h = h
        self._n = {}
        self._max_n = 0
        # submit() and poll() may run on different threads: a request's size
        # is recorded (and the output capacity raised) under this lock before
        # the request can finish
        self._lock = threading.Lock()

    def submit(self, domain, xyzw, **kw):
        xyzw = np.ascontiguousarray(xyzw, dtype=np.float64).reshape(-1, 4)
        p = params(**kw)
        t = I64(0)
        with self._lock:
            self._max_n = max(self._max_n, xyzw.shape[0])
            _check(lib().rpd_server_submit(self._h, domain._h if domain is not None else None, _p(xyzw),
                                           xyzw.shape[0], ctypes.byref(p), ctypes.byref(t)), None,
                   "rpd_server_submit")
            self._n[t.value] = xyzw.shape[0]
        return t.value

    def poll(self, wait=True):
        """(ticket, moments[n,4], latency_ms) of a finished request, or None."""
        with self._lock:
            cap = self._max_n
        out = np.zeros((max(cap, 1), 4), np.float64)
        t = I64(-1)
        lat = ctypes.c_double(0)
        rc = lib().rpd_server_poll(self._h, 1 if wait else 0, ctypes.byref(t), _p(out), cap, ctypes.byref(lat))
        if t.value < 0:
            _check(rc, None, "rpd_server_poll")
            return None
        with self._lock:
            n = self._n.pop(t.value)
        _check(rc, None, "rpd_server_poll (ticket %d)" % t.value)
        return t.value, out[:n].copy(), lat.value

    def drain(self):
        _check(lib().rpd_server_drain(self._h), None, "rpd_server_drain")

    def close(self):
        if getattr(self, "_h", None):
            lib().rpd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Domain:
    """A domain grid resident in HBM (rpd_domain_create)."""

    def __init__(self, ctx, handle, grid):
        self.ctx = ctx
        self._h = handle
        self.grid = grid

    def close(self):
        if getattr(self, "_h", None):
            lib().rpd_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VoroAlgo:
    """Mirror of the reference's VoroAlgo (voronoi.h:105-256) over librpd.

    setInputPoints / setGrid / launch / getBarycenters / getPermutationToUser
    keep the reference's meaning: launch computes the restricted power diagram
    and calls saveData(self) while the outputs are available.  Outputs are the
    used cells (volume >= m_keepVolume, ascending original index) in
    (m*x, m*y, m*z, m) form, and getPermutationToUser maps them to the user's
    seed indices (compressResult, voronoi.cpp:513-691).
    """

    def __init__(self, device=0, debug=False):
        self._ctx = Context(device)
        self._dom = None
        self.m_K, self.m_P, self.m_V, self.m_T = 25, 35, 40, 100
        self.m_keepVolume = 1.0
        self.m_keepDistance = 1  # voronoi.cpp:1052: one ring of bordering cells
        self.precision = PREC_MIXED
        self.max_extra_steps = 5
        self.apex_mode = 1
        self.security_mode = 0
        self.m_H, self.m_HC = 0, 1
        self.m_debug = debug
        self._bary = None
        self._perm = None
        self._input = None

    def setGrid(self, grid):
        self._dom = self._ctx.domain(grid)

    def setInputPoints(self, seeds):
        self._input = np.ascontiguousarray(seeds, dtype=np.float64).reshape(-1, 4)
        self._ctx.set_seeds(self._input)

    def getNumInputPoints(self):
        return 0 if self._input is None else self._input.shape[0]

    def launch(self, saveData=None, stats=None):
        st = self._ctx.solve(self._dom, self.m_K, self.m_P, self.m_V, self.m_T, self.precision,
                             self.max_extra_steps, self.apex_mode, self.security_mode, H=self.m_H, HC=self.m_HC)
        if stats is not None:
            stats.append(st)
        full = self._ctx.moments()
        self._compacted = not (self.m_keepDistance < 0 or self._dom is None)
        if not self._compacted:
            self._bary = full
            self._perm = np.arange(self._bary.shape[0], dtype=np.int32)
        else:
            self._bary, self._perm = self._ctx.compact_used(self.m_keepVolume, self.m_keepDistance)
        self._hess = self._hess_ids = None
        if self.m_H > 0:
            vals, ids = self._ctx.hessian()
            H = self.m_H
            if not self._compacted:
                self._hess, self._hess_ids = vals.reshape(-1), ids.reshape(-1)
            else:
                # compact_result (voronoi.cpp:594-631): the output cells'
                # entries, neighbours renumbered to output positions, dropped
                # when not output or without volume; a cell without volume has
                # no entries
                pos = np.full(self._ctx.n, -1, np.int64)
                pos[self._perm] = np.arange(self._perm.size)
                hv = np.zeros((self._perm.size, H, self.m_HC))
                hi = np.full((self._perm.size, H), -1, np.int32)
                for o, u in enumerate(self._perm):
                    if full[u, 3] < self.m_keepVolume:
                        continue
                    w = 0
                    for h in range(H):
                        j = ids[u, h]
                        if j < 0:
                            break
                        if pos[j] < 0 or full[j, 3] < self.m_keepVolume:
                            continue
                        hi[o, w] = pos[j]
                        hv[o, w] = vals[u, h]
                        w += 1
                self._hess, self._hess_ids = hv.reshape(-1), hi.reshape(-1)
        ok = True if saveData is None else bool(saveData(self))
        return ok

    def getNumOutputPoints(self):
        return 0 if self._bary is None else self._bary.shape[0]

    def getBarycenters(self):
        return self._bary

    def getPermutationToUser(self):
        return self._perm

    def getPermutationToInputPoints(self):
        """Output position -> index of the input points (the user's order);
        empty when it is the identity (voronoi.h:183-187)."""
        return self._perm if getattr(self, "_compacted", False) else np.zeros(0, np.int32)

    def getHessians(self):
        """m_HC * m_H values per output cell (voronoi.cpp:1179-1212); empty with m_H = 0."""
        return self._hess if getattr(self, "_hess", None) is not None else np.zeros(0)

    def getHessianIds(self):
        """m_H output indices per output cell, -1 after the last (voronoi.cpp:1214-1232)."""
        return self._hess_ids if getattr(self, "_hess_ids", None) is not None else np.zeros(0, np.int32)
