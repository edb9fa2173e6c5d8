// rpd_grid.cpp -- host side of the domain grid: voxel-union writer, the
// reference text format (Grid::load / Grid::save, Grid.cpp:186-344), the
// validity rules (Grid::valid, Grid.cpp:60-184) and the volume invariants
// (Grid.cpp:18-53).
//
// Frame: voxel (i,j,k) spans [i,i+1]x[j,j+1]x[k,k+1] * (1000/g); its
// "first corner" is (i,j,k)*(1000/g), whose flag is stored at i+j*g+k*g*g.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rpd.h"
#include "rpd_internal.h"

namespace {
void derive_centre_slices(rpd_grid *gr, int first, int stride);
}  // namespace

// Classes of the voxels' apex points for a grid read from a file: parity of
// the surface crossings of the ray from the point towards +x, row by row from
// the voxels' triangle lists (a triangle the ray meets lies in the list of a
// voxel of the row).  A row whose ray passes within rounding of a triangle
// edge, and a point within rounding of the surface, keep class 1 (unknown).
// Rows are independent: z-slices are shared out over host threads.
void rpd_derive_centre_classes(rpd_grid *gr) {
    const int g = gr->size;
    const size_t nv = gr->in_domain.size();
    gr->centre.assign(nv, 1);
    const int nthreads = std::max(1, std::min({int(std::thread::hardware_concurrency()), 16, g}));
    std::vector<std::thread> pool;
    for (int w = 1; w < nthreads; ++w) pool.emplace_back(derive_centre_slices, gr, w, nthreads);
    derive_centre_slices(gr, 0, nthreads);
    for (auto &t : pool) t.join();
}

namespace {
void derive_centre_slices(rpd_grid *gr, int first, int stride) {
    const int g = gr->size;
    const double vs = 1000.0 / double(g);
    std::vector<int> tris;
    std::vector<double> xs;
    for (int z = first; z < g; z += stride)
        for (int y = 0; y < g; ++y) {
            const double py = (y + kApexFracHost[1]) * vs, pz = (z + kApexFracHost[2]) * vs;
            tris.clear();
            for (int x = 0; x < g; ++x) {
                const size_t id = size_t(x) + size_t(y) * g + size_t(z) * g * g;
                for (int q = gr->offsets[id]; q < gr->offsets[id + 1]; ++q) tris.push_back(gr->tri_list[q]);
            }
            std::sort(tris.begin(), tris.end());
            tris.erase(std::unique(tris.begin(), tris.end()), tris.end());
            xs.clear();
            bool degenerate = false;
            for (int t : tris) {
                const double *A = &gr->points[4 * size_t(gr->triangles[3 * size_t(t)])];
                const double *B = &gr->points[4 * size_t(gr->triangles[3 * size_t(t) + 1])];
                const double *C = &gr->points[4 * size_t(gr->triangles[3 * size_t(t) + 2])];
                const double d = (B[1] - A[1]) * (C[2] - A[2]) - (B[2] - A[2]) * (C[1] - A[1]);
                const double scale = std::fabs(B[1] - A[1]) + std::fabs(C[1] - A[1]) + std::fabs(B[2] - A[2]) +
                                     std::fabs(C[2] - A[2]);
                if (std::fabs(d) <= 1e-14 * scale * scale) continue;  // parallel to the ray
                const double l1 = ((B[1] - py) * (C[2] - pz) - (B[2] - pz) * (C[1] - py)) / d;
                const double l2 = ((C[1] - py) * (A[2] - pz) - (C[2] - pz) * (A[1] - py)) / d;
                const double l3 = 1.0 - l1 - l2;
                const double eps = 1e-12;
                if (l1 < -eps || l2 < -eps || l3 < -eps) continue;
                if (l1 < eps || l2 < eps || l3 < eps) {
                    degenerate = true;
                    break;
                }
                xs.push_back(l1 * A[0] + l2 * B[0] + l3 * C[0]);
            }
            if (degenerate) continue;
            std::sort(xs.begin(), xs.end());
            for (int x = 0; x < g; ++x) {
                const double px = (x + kApexFracHost[0]) * vs;
                const auto it = std::upper_bound(xs.begin(), xs.end(), px);
                const bool near = (it != xs.end() && *it - px < 1e-9 * vs) ||
                                  (it != xs.begin() && px - *(it - 1) < 1e-9 * vs);
                if (near) continue;
                const size_t above = size_t(xs.end() - it);
                gr->centre[size_t(x) + size_t(y) * g + size_t(z) * g * g] = (above & 1) ? 2 : 0;
            }
        }
}
}  // namespace

namespace {

inline double det3(double a11, double a12, double a13, double a21, double a22, double a23, double a31, double a32,
                   double a33) {
    return a11 * (a22 * a33 - a23 * a32) - a21 * (a12 * a33 - a13 * a32) + a31 * (a12 * a23 - a13 * a22);
}

// Signed volume contribution of one triangle as in Grid.cpp:18-22.
inline double tet_volume(const double *a, const double *b, const double *c) {
    return -det3(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]) / 6;
}

}  // namespace

extern "C" int rpd_voxel_mask(int shape, int g, const double *p, uint8_t *out) {
    if (g <= 0 || !out) return RPD_E_ARG;
    if ((shape == RPD_SHAPE_SPHERE || shape == RPD_SHAPE_TORUS || shape == RPD_SHAPE_CUBE_MINUS_SPHERE) && !p)
        return RPD_E_ARG;
    for (int k = 0; k < g; ++k)
        for (int j = 0; j < g; ++j)
            for (int i = 0; i < g; ++i) {
                const double x = (i + 0.5) / g, y = (j + 0.5) / g, z = (k + 0.5) / g;
                bool in = false;
                switch (shape) {
                    case RPD_SHAPE_SPHERE: {
                        const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                        in = dx * dx + dy * dy + dz * dz <= p[3] * p[3];
                        break;
                    }
                    case RPD_SHAPE_TORUS: {
                        const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                        const double q = std::sqrt(dx * dx + dy * dy) - p[3];
                        in = q * q + dz * dz <= p[4] * p[4];
                        break;
                    }
                    case RPD_SHAPE_CUBE:
                        in = true;
                        break;
                    case RPD_SHAPE_CUBE_MINUS_SPHERE: {
                        const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                        in = dx * dx + dy * dy + dz * dz > p[3] * p[3];
                        break;
                    }
                    default:
                        return RPD_E_ARG;
                }
                out[size_t(i) + size_t(j) * g + size_t(k) * g * g] = in ? 1 : 0;
            }
    return RPD_OK;
}

extern "C" int rpd_grid_from_voxels(int g, const uint8_t *mask, rpd_grid **out) {
    if (g <= 0 || g > 1024 || !mask || !out) return RPD_E_ARG;
    auto *gr = new rpd_grid();
    gr->size = g;
    const size_t nv = size_t(g) * g * g;
    auto kept = [&](int i, int j, int k) -> bool {
        if (i < 0 || j < 0 || k < 0 || i >= g || j >= g || k >= g) return false;
        return mask[size_t(i) + size_t(j) * g + size_t(k) * g * g] != 0;
    };
    // corner flags, z-slices shared out over host threads
    gr->in_domain.resize(nv);
    auto flag_slices = [&](int first, int stride) {
        for (int k = first; k < g; k += stride)
            for (int j = 0; j < g; ++j)
                for (int i = 0; i < g; ++i) {
                    int c = 0;
                    for (int dz = -1; dz <= 0; ++dz)
                        for (int dy = -1; dy <= 0; ++dy)
                            for (int dx = -1; dx <= 0; ++dx) c += kept(i + dx, j + dy, k + dz);
                    gr->in_domain[size_t(i) + size_t(j) * g + size_t(k) * g * g] =
                        int8_t(c == 8 ? 2 : (c == 0 ? 0 : 1));
                }
    };
    {
        const int nthreads = std::max(1, std::min({int(std::thread::hardware_concurrency()), 16, g}));
        std::vector<std::thread> pool;
        for (int w = 1; w < nthreads; ++w) pool.emplace_back(flag_slices, w, nthreads);
        flag_slices(0, nthreads);
        for (auto &t : pool) t.join();
    }
    // lattice points, created on demand
    const double vs = 1000.0 / g;
    const int gp = g + 1;
    // dense index of the (g+1)^3 lattice points, -1 until first used: ids
    // follow the order in which the faces first touch the points
    std::vector<int32_t> pid(size_t(gp) * gp * gp, -1);
    auto point = [&](int i, int j, int k) -> int32_t {
        int32_t &id = pid[size_t(i) + size_t(j) * gp + size_t(k) * gp * gp];
        if (id >= 0) return id;
        id = int32_t(gr->points.size() / 4);
        gr->points.push_back(i * vs);
        gr->points.push_back(j * vs);
        gr->points.push_back(k * vs);
        gr->points.push_back(0.0);
        return id;
    };
    // boundary faces: for each axis a and each lattice plane position, the face
    // between voxel (.., c-1, ..) and (.., c, ..) along a is a boundary face iff
    // exactly one side is kept.  Each face = unit square in the two other axes,
    // split along its diagonal into two triangles.
    struct Tri {
        int a;        // normal axis
        int c;        // lattice plane index along a
        int u, v;     // square origin in the two other axes (a+1)%3, (a+2)%3
        int half;     // 0: corners (0,0),(1,0),(1,1); 1: (0,0),(1,1),(0,1)
    };
    std::vector<Tri> tris;
    int64_t nfaces = 0;  // boundary faces, counted first: ids are int32
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c <= g; ++c)
            for (int v = 0; v < g; ++v)
                for (int u = 0; u < g; ++u) {
                    int lo[3], hi[3];
                    lo[a] = c - 1;
                    hi[a] = c;
                    lo[(a + 1) % 3] = hi[(a + 1) % 3] = u;
                    lo[(a + 2) % 3] = hi[(a + 2) % 3] = v;
                    nfaces += kept(lo[0], lo[1], lo[2]) != kept(hi[0], hi[1], hi[2]);
                }
    if (6 * nfaces > int64_t(INT32_MAX)) {
        delete gr;
        return RPD_E_LIMIT;
    }
    tris.reserve(size_t(2 * nfaces));
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, cax = (a + 2) % 3;
        for (int c = 0; c <= g; ++c)
            for (int v = 0; v < g; ++v)
                for (int u = 0; u < g; ++u) {
                    int lo[3], hi[3];
                    lo[a] = c - 1;
                    hi[a] = c;
                    lo[b] = hi[b] = u;
                    lo[cax] = hi[cax] = v;
                    const bool klo = kept(lo[0], lo[1], lo[2]), khi = kept(hi[0], hi[1], hi[2]);
                    if (klo == khi) continue;
                    // square corners in (b, cax): q00=(u,v) q10=(u+1,v) q11=(u+1,v+1) q01=(u,v+1)
                    auto P = [&](int du, int dv) {
                        int x[3];
                        x[a] = c;
                        x[b] = u + du;
                        x[cax] = v + dv;
                        return point(x[0], x[1], x[2]);
                    };
                    const int32_t q00 = P(0, 0), q10 = P(1, 0), q11 = P(1, 1), q01 = P(0, 1);
                    // (b, cax, a) is a right-handed frame, so the loop
                    // q00->q10->q11 has its cross product along +a.  The
                    // reference convention (Grid.cpp:18-35, volume > 0 with the
                    // minus sign) wants cross products pointing into the
                    // domain: +a when the kept voxel is on the high side.
                    const bool inward_plus = khi;
                    const int32_t t0[3] = {q00, q10, q11}, t1[3] = {q00, q11, q01};
                    for (int h = 0; h < 2; ++h) {
                        const int32_t *t = h == 0 ? t0 : t1;
                        if (inward_plus) {
                            gr->triangles.insert(gr->triangles.end(), {t[0], t[1], t[2]});
                        } else {
                            gr->triangles.insert(gr->triangles.end(), {t[0], t[2], t[1]});
                        }
                        tris.push_back(Tri{a, c, u, v, h});
                    }
                }
    }
    // per-voxel lists by exact closed-box overlap.  The triangle lies in the
    // plane x_a = c, so candidate voxels have index c-1 or c along a; in the
    // (b, cax) plane the triangle is half of the unit square [u,u+1]x[v,v+1]
    // cut by its diagonal: half 0 keeps t_v <= t_u, half 1 keeps t_v >= t_u
    // (t = local coordinate in [0,1]).  A neighbour square overlaps iff some
    // corner of its intersection with the unit square satisfies that test.
    // The overlaps are visited twice, in increasing triangle order: counts,
    // then entries, so each list is sorted.
    auto visit = [&](auto &&emit) {
    for (size_t t = 0; t < tris.size(); ++t) {
        const Tri &T = tris[t];
        const int b = (T.a + 1) % 3, cax = (T.a + 2) % 3;
        for (int ia = T.c - 1; ia <= T.c; ++ia) {
            if (ia < 0 || ia >= g) continue;
            for (int jb = T.u - 1; jb <= T.u + 1; ++jb) {
                if (jb < 0 || jb >= g) continue;
                for (int kc = T.v - 1; kc <= T.v + 1; ++kc) {
                    if (kc < 0 || kc >= g) continue;
                    // intersection of [jb,jb+1]x[kc,kc+1] with [u,u+1]x[v,v+1] in local coords
                    const int u0 = std::max(jb, T.u) - T.u, u1 = std::min(jb + 1, T.u + 1) - T.u;
                    const int v0 = std::max(kc, T.v) - T.v, v1 = std::min(kc + 1, T.v + 1) - T.v;
                    bool hit = false;
                    for (int su = 0; su < 2 && !hit; ++su)
                        for (int sv = 0; sv < 2 && !hit; ++sv) {
                            const int tu = su ? u1 : u0, tv = sv ? v1 : v0;
                            hit = T.half == 0 ? (tv <= tu) : (tv >= tu);
                        }
                    if (!hit) continue;
                    int x[3];
                    x[T.a] = ia;
                    x[b] = jb;
                    x[cax] = kc;
                    emit(size_t(x[0]) + size_t(x[1]) * g + size_t(x[2]) * g * g, int32_t(t));
                }
            }
        }
    }
    };
    gr->centre.resize(nv);
    for (size_t i = 0; i < nv; ++i) gr->centre[i] = mask[i] ? 2 : 0;
    // list entries are counted in 64 bits: the grid's ids and offsets are
    // int32 (Grid.h), which a porous mask at g ~ 1000 can exceed
    std::vector<int64_t> cnt(nv + 1, 0);
    visit([&](size_t v, int32_t) { ++cnt[v + 1]; });
    for (size_t i = 0; i < nv; ++i) cnt[i + 1] += cnt[i];
    if (cnt[nv] > int64_t(INT32_MAX)) {
        delete gr;
        return RPD_E_LIMIT;
    }
    gr->offsets.assign(nv + 1, 0);
    for (size_t i = 0; i <= nv; ++i) gr->offsets[i] = int32_t(cnt[i]);
    std::vector<int64_t>().swap(cnt);
    gr->tri_list.resize(size_t(gr->offsets[nv]));
    std::vector<int32_t> cursor(gr->offsets.begin(), gr->offsets.end() - 1);
    visit([&](size_t v, int32_t t) { gr->tri_list[size_t(cursor[v]++)] = t; });
    *out = gr;
    return RPD_OK;
}

extern "C" int rpd_grid_view(const rpd_grid *gr, rpd_grid_desc *d) {
    if (!gr || !d) return RPD_E_ARG;
    d->size = gr->size;
    d->in_domain = gr->in_domain.data();
    d->points = gr->points.data();
    d->num_points = int32_t(gr->points.size() / 4);
    d->triangles = gr->triangles.data();
    d->num_triangles = int32_t(gr->triangles.size() / 3);
    d->tri_list = gr->tri_list.data();
    d->num_tri_list = int32_t(gr->tri_list.size());
    d->offsets = gr->offsets.data();
    d->centre = gr->centre.empty() ? nullptr : gr->centre.data();
    return RPD_OK;
}

extern "C" void rpd_grid_free(rpd_grid *gr) { delete gr; }

extern "C" double rpd_grid_volume(const rpd_grid_desc *g) {
    double vol = 0;
    for (int t = 0; t < g->num_triangles; ++t)
        vol += tet_volume(g->points + 4 * g->triangles[3 * t], g->points + 4 * g->triangles[3 * t + 1],
                          g->points + 4 * g->triangles[3 * t + 2]);
    return vol;
}

extern "C" void rpd_grid_moments(const rpd_grid_desc *g, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int t = 0; t < g->num_triangles; ++t) {
        const double *a = g->points + 4 * g->triangles[3 * t];
        const double *b = g->points + 4 * g->triangles[3 * t + 1];
        const double *c = g->points + 4 * g->triangles[3 * t + 2];
        const double v = tet_volume(a, b, c);
        for (int d = 0; d < 3; ++d) out[d] += v / 4 * (a[d] + b[d] + c[d]);
        out[3] += v;
    }
}

const char *rpd_grid_desc_check(const rpd_grid_desc *g) {
    if (!g || g->size <= 0) return "no grid";
    const int64_t nv = int64_t(g->size) * g->size * g->size;
    if (g->num_points < 0 || g->num_triangles < 0 || g->num_tri_list < 0) return "negative count";
    if (!g->in_domain || !g->offsets) return "missing flags or offsets";
    if ((g->num_points > 0 && !g->points) || (g->num_triangles > 0 && !g->triangles) ||
        (g->num_tri_list > 0 && !g->tri_list))
        return "missing array";
    for (int64_t i = 0; i < 3 * int64_t(g->num_triangles); ++i)
        if (g->triangles[i] < 0 || g->triangles[i] >= g->num_points) return "triangle point id out of range";
    for (int64_t i = 0; i < g->num_tri_list; ++i)
        if (g->tri_list[i] < 0 || g->tri_list[i] >= g->num_triangles) return "list entry out of range";
    if (g->offsets[0] != 0) return "offsets do not start at 0";
    for (int64_t v = 0; v < nv; ++v)
        if (g->offsets[v + 1] < g->offsets[v]) return "offsets decrease";
    if (g->offsets[nv] > g->num_tri_list) return "offsets end past the list";
    return nullptr;
}

extern "C" int rpd_grid_valid(const rpd_grid_desc *g) {
    if (rpd_grid_desc_check(g)) return 0;
    const long long nv = (long long)g->size * g->size * g->size;
    for (long long i = 0; i < nv; ++i)
        if (g->in_domain[i] & 0xfc) return 0;
    for (int i = 0; i < 3 * g->num_triangles; ++i)
        if (g->triangles[i] < 0 || g->triangles[i] >= g->num_points) return 0;
    for (int i = 0; i < g->num_tri_list; ++i)
        if (g->tri_list[i] < 0 || g->tri_list[i] >= g->num_triangles) return 0;
    for (long long i = 0; i <= nv; ++i)
        if (g->offsets[i] < 0 || g->offsets[i] > g->num_tri_list) return 0;
    if (rpd_grid_volume(g) <= 0) return 0;
    for (int t = 0; t < g->num_triangles; ++t) {
        const double *a = g->points + 4 * g->triangles[3 * t];
        const double *b = g->points + 4 * g->triangles[3 * t + 1];
        const double *c = g->points + 4 * g->triangles[3 * t + 2];
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        if ((cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 4 < 0.01 * 0.01) return 0;
    }
    // closed 2-manifold: every directed edge once, and its reverse present
    std::set<std::pair<int, int>> edges;
    for (int t = 0; t < g->num_triangles; ++t)
        for (int j = 0; j < 3; ++j)
            if (!edges.insert({g->triangles[3 * t + j], g->triangles[3 * t + (j + 1) % 3]}).second) return 0;
    for (const auto &e : edges)
        if (!edges.count({e.second, e.first})) return 0;
    // no voxel whose 8 corners are all on the border (flag 1): its corner
    // apex could not be placed.  The flag of voxel v is that of its first
    // corner, so voxel v's corners are the flags of v + {0,1}^3; a voxel on
    // a max face is skipped, its +1 corners taken to lie outside
    // (Grid.cpp:147-167, which means this test; its loop counts the 8 voxels
    // at the grid origin for every border voxel instead of those of v).
    const long long gs = g->size;
    for (long long z = 0; z + 1 < gs; ++z)
        for (long long y = 0; y + 1 < gs; ++y)
            for (long long x = 0; x + 1 < gs; ++x) {
                bool all = true;
                for (int c = 0; c < 8 && all; ++c)
                    all = g->in_domain[(x + (c & 1)) + (y + ((c >> 1) & 1)) * gs + (z + (c >> 2)) * gs * gs] == 1;
                if (all) return 0;
            }
    for (long long o = 0; o < nv; ++o) {
        if (g->offsets[o] > g->offsets[o + 1]) return 0;
        for (int t = g->offsets[o]; t + 1 < g->offsets[o + 1]; ++t)
            if (g->tri_list[t] >= g->tri_list[t + 1]) return 0;
    }
    return 1;
}

// Text format of Grid::save (Grid.cpp:290-344): size, the g^3 flags written
// as raw characters (values 0/1/2, each followed by a space), point count and
// xyz triples, triangle index count and indices, list count and ids, offset
// count and offsets.
extern "C" int rpd_grid_save(const rpd_grid *gr, const char *path) {
    if (!gr || !path) return RPD_E_ARG;
    std::ofstream f(path, std::ofstream::trunc | std::ofstream::binary);
    if (!f.is_open()) return RPD_E_IO;
    f.precision(17);
    f << gr->size << "\n";
    for (int8_t c : gr->in_domain) f << char(c) << " ";
    f << "\n" << gr->points.size() / 4 << "\n";
    for (size_t i = 0; i < gr->points.size() / 4; ++i)
        f << gr->points[4 * i] << " " << gr->points[4 * i + 1] << " " << gr->points[4 * i + 2] << " ";
    f << "\n" << gr->triangles.size() << "\n";
    for (int32_t t : gr->triangles) f << t << " ";
    f << "\n" << gr->tri_list.size() << "\n";
    for (int32_t t : gr->tri_list) f << t << " ";
    f << "\n" << gr->offsets.size() << "\n";
    for (int32_t t : gr->offsets) f << t << " ";
    f << "\n";
    return f.good() ? RPD_OK : RPD_E_IO;
}

extern "C" int rpd_grid_load(const char *path, rpd_grid **out) {
    if (!path || !out) return RPD_E_ARG;
    std::ifstream f(path, std::ifstream::binary);
    if (!f.is_open()) return RPD_E_IO;
    auto *gr = new rpd_grid();
    auto fail = [&]() {
        delete gr;
        return RPD_E_IO;
    };
    f >> gr->size;
    if (!f.good() || gr->size <= 0 || gr->size > 1024) return fail();
    const size_t nv = size_t(gr->size) * gr->size * gr->size;
    gr->in_domain.resize(nv);
    for (size_t i = 0; i < nv; ++i) {
        char c;
        f >> c;  // operator>> on char skips whitespace: flags are raw bytes
        if (c & 0xfc) return fail();
        gr->in_domain[i] = int8_t(c);
    }
    int np = 0;
    f >> np;
    if (!f.good() || np < 0) return fail();
    gr->points.resize(4 * size_t(np));
    for (int i = 0; i < np; ++i) {
        for (int j = 0; j < 3; ++j) f >> gr->points[4 * size_t(i) + j];
        gr->points[4 * size_t(i) + 3] = 0;
    }
    int nt = 0;
    f >> nt;
    if (!f.good() || nt < 0 || nt % 3) return fail();
    gr->triangles.resize(size_t(nt));
    for (auto &t : gr->triangles) {
        f >> t;
        if (t < 0 || t >= np) return fail();
    }
    int ni = 0;
    f >> ni;
    if (!f.good() || ni < 0) return fail();
    gr->tri_list.resize(size_t(ni));
    for (auto &t : gr->tri_list) {
        f >> t;
        if (t < 0 || t >= nt / 3) return fail();
    }
    int no = 0;
    f >> no;
    if (!f.good() || no <= int(nv)) return fail();
    gr->offsets.resize(size_t(no));
    int prev = 0;
    for (auto &o : gr->offsets) {
        f >> o;
        // offsets are a non-decreasing prefix into the list entries
        if (o < prev || o > ni) return fail();
        prev = o;
    }
    if (f.fail() || gr->offsets[0] != 0) return fail();
    // orientation normalisation of Grid::load (Grid.cpp:270-275)
    rpd_grid_desc d;
    rpd_grid_view(gr, &d);
    if (rpd_grid_volume(&d) < 0)
        for (size_t t = 0; t + 2 < gr->triangles.size(); t += 3) std::swap(gr->triangles[t], gr->triangles[t + 1]);
    rpd_derive_centre_classes(gr);
    *out = gr;
    return RPD_OK;
}
