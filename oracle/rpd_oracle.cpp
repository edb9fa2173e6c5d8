// rpd_oracle.cpp -- CPU oracle for the restricted power diagram hot path.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  The shipped library never
// links or calls this file.
//
// It restates, in plain C++ on the CPU, the algorithm of the reference:
//   * kNN binning and query    knearests.cpp:35-222, knearests.cl:12-78
//   * convex-cell clipping     ConvexCell.cl:278-563 (CCInit, CCNewPlane,
//                              CCClipByPlane, CCComputeBoundary, CCNewVertex,
//                              CCVertexIsInHalfSpace, CCIsSecurityRadiusReached)
//   * restriction to a grid    ConvexCell.cl:918-1116, merge :722-749,
//                              CCClipTetFromPointsInfinite :565-575
//   * integration              CCExportBaryVolume ConvexCell.cl:577-619
//   * pass policy              voronoi.cpp:1244-1376 (first / tr_overflow /
//                              extra passes, caps x1.5)
//   * per-precision constants  voronoi.cpp:915-947
// with these stated definitions where the reference is non-deterministic or
// reads out of bounds:
//   * kNN order is canonical: ascending (d^2, original index), self excluded,
//     d^2 = ((dx*dx + dy*dy) + dz*dz) unfused in the seed precision; the
//     search is exact (per-axis ring bounds, no row wrap-around, no ring cap).
//   * binning is stable: inside a grid cell, seeds keep original order.
//   * a kNN list shorter than K (fewer than K+1 seeds) is padded with -1 and a
//     cell that consumed every other seed is complete.
//   * GRID_VOXEL_SIZE is the exact 1000/g in the working precision (the
//     reference injects it printed with 6 significant digits).
// Build with -ffp-contract=off so every product and sum rounds separately.

#include "rpd_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Status numbering of Status.h:4-16.
enum {
    kVertexOverflow = 0,
    kPlaneOverflow = 1,
    kInconsistentBoundary = 2,
    kSecurityRadiusNotReached = 3,
    kSuccess = 4,
    kNeedsExactPredicates = 5,
    kEmptyCell = 6,
    kFindAnotherBeginningVertex = 7,
    kTriangleOverflow = 8,
    kSeedOnBorder = 9,
    kHessianOverflow = 10,
    kNumStatus = 11
};

constexpr double kNoCellLimits = -100000.0;  // voronoi.h:53
constexpr int kEndOfList = 255;              // ConvexCell.cl:132

template <class R>
struct R4 {
    R x, y, z, w;
};
template <class R>
inline R4<R> mk4(R x, R y, R z, R w) { return R4<R>{x, y, z, w}; }
template <class R>
inline R4<R> sub4(const R4<R> &a, const R4<R> &b) { return mk4<R>(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
template <class R>
inline R4<R> add4(const R4<R> &a, const R4<R> &b) { return mk4<R>(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
template <class R>
inline R4<R> scale4(R s, const R4<R> &a) { return mk4<R>(s * a.x, s * a.y, s * a.z, s * a.w); }
template <class R>
inline R dot3(const R4<R> &a, const R4<R> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R>
inline R dot4(const R4<R> &a, const R4<R> &b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// OpenCL cross() on 4-vectors: xyz cross product, w = 0.
template <class R>
inline R4<R> cross3(const R4<R> &a, const R4<R> &b) {
    return mk4<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, R(0));
}
template <class R>
inline R det2(R a11, R a12, R a21, R a22) { return a11 * a22 - a12 * a21; }
// ConvexCell.cl:72-77 -- rows A, B, C.
template <class R>
inline R det3(R a11, R a12, R a13, R a21, R a22, R a23, R a31, R a32, R a33) {
    return a11 * det2(a22, a23, a32, a33) - a21 * det2(a12, a13, a32, a33) + a31 * det2(a12, a13, a22, a23);
}
template <class R>
inline R det3r4(const R4<R> &A, const R4<R> &B, const R4<R> &C) {
    return det3<R>(A.x, A.y, A.z, B.x, B.y, B.z, C.x, C.y, C.z);
}
// ConvexCell.cl:79-97 (and det4x4D :104-122 for T = double).
template <class T>
inline T det4(T a11, T a12, T a13, T a14, T a21, T a22, T a23, T a24, T a31, T a32, T a33, T a34,
              T a41, T a42, T a43, T a44) {
    const T m12 = a21 * a12 - a11 * a22;
    const T m13 = a31 * a12 - a11 * a32;
    const T m14 = a41 * a12 - a11 * a42;
    const T m23 = a31 * a22 - a21 * a32;
    const T m24 = a41 * a22 - a21 * a42;
    const T m34 = a41 * a32 - a31 * a42;
    const T m123 = m23 * a13 - m13 * a23 + m12 * a33;
    const T m124 = m24 * a13 - m14 * a23 + m12 * a43;
    const T m134 = m34 * a13 - m14 * a33 + m13 * a43;
    const T m234 = m34 * a23 - m24 * a33 + m23 * a43;
    return (m234 * a14 - m134 * a24 + m124 * a34 - m123 * a44);
}
// ConvexCell.cl:125-129
template <class R>
inline R4<R> plane_from_points(const R4<R> &A, const R4<R> &B, const R4<R> &C) {
    R4<R> p = cross3(sub4(B, A), sub4(C, A));
    p.w = -dot3(p, A);
    return p;
}

// apex_mode 1 puts the cone apex at this fraction of a voxel whose interior
// has one known class (in or out).  Any interior point is valid; a generic
// one keeps the cone planes off the voxel lattice and off the CUBE_EPSILON
// box (the voxel centre lies on planes such as y + z = 1000 that contain box
// edges such as y = 1000.1, z = -0.1).
constexpr double kApexFrac[3] = {0.530901699437495, 0.520710678118655, 0.536602540378444};

struct Counters {
    long long n_pred = 0, n_radius_vertex = 0, n_planes = 0, n_bbox_vertex = 0, n_integrated_vertex = 0,
              n_clip = 0, n_pred_restrict = 0, n_planes_restrict = 0;
    bool restricting = false;  // the cell is in its restriction phase
    void add(const Counters &o) {
        n_pred_restrict += o.n_pred_restrict;
        n_planes_restrict += o.n_planes_restrict;
        n_pred += o.n_pred;
        n_radius_vertex += o.n_radius_vertex;
        n_planes += o.n_planes;
        n_bbox_vertex += o.n_bbox_vertex;
        n_integrated_vertex += o.n_integrated_vertex;
        n_clip += o.n_clip;
    }
};

// ---------------------------------------------------------------------------
// Grid view in the working precision (GridAccess, voronoi.cpp:42-144)
// ---------------------------------------------------------------------------
template <class R>
struct GridView {
    int g = 0;
    R vs = 0;  // voxel size 1000/g
    const signed char *flags = nullptr;
    std::vector<R4<R>> pts;
    const int *tris = nullptr;
    const int *list = nullptr;
    const int *offs = nullptr;
    const signed char *centre = nullptr;  // voxel-centre classes (nullable)
    bool has_border_vertices = false;
    int id(int x, int y, int z) const { return x + y * g + z * g * g; }
};

template <class R>
GridView<R> make_grid_view(const orc_grid *gr) {
    GridView<R> v;
    if (!gr) return v;
    v.g = gr->size;
    v.vs = R(1000) / R(gr->size);
    v.flags = gr->in_domain;
    v.pts.resize(size_t(gr->num_points));
    for (int i = 0; i < gr->num_points; ++i)
        v.pts[size_t(i)] = mk4<R>(R(gr->points[4 * i]), R(gr->points[4 * i + 1]), R(gr->points[4 * i + 2]), R(0));
    v.tris = gr->triangles;
    v.list = gr->tri_list;
    v.offs = gr->offsets;
    v.centre = gr->centre;
    size_t nv = size_t(v.g) * size_t(v.g) * size_t(v.g);
    for (size_t i = 0; i < nv; ++i)
        if (v.flags[i] == 1) {
            v.has_border_vertices = true;
            break;
        }
    return v;
}

// ---------------------------------------------------------------------------
// Convex cell (ConvexCell.cl:134-149), one private instance per seed.
// ---------------------------------------------------------------------------
template <class R>
struct Cell {
    using V3 = std::array<uint8_t, 3>;
    R4<R> seed;
    const R4<R> *pts = nullptr;
    // Origin of the frame the cell is built in: (0,0,0) for the reference's
    // [0,1000]^3 frame, the seed for the seed-local frame (w stays 0, so a
    // shifted point keeps its weight).
    R4<R> org = R4<R>{0, 0, 0, 0};
    int status = kSuccess;
    int P = 0, V = 0, vhs = 2;
    int security_mode = 1;  // 0 exact power stop, 1 reference stop
    double wmin = 0;        // minimum weight over non-border seeds
    int nb_vertex = 0, nb_plane = 0;
    R4<R> planes[256];
    int pnb[256];  // neighbour (stored id) of each plane, -1: cube face or cone plane
    V3 verts[256];
    uint8_t next[256];
    Counters *ctr = nullptr;

    // point vid in the cell's frame
    R4<R> at(int vid) const { return sub4(pts[vid], org); }

    // CCInit ConvexCell.cl:278-308; local: the frame's origin is the seed
    void init(const R4<R> *p_pts, int id, R cube_eps, bool local = false) {
        pts = p_pts;
        org = local ? mk4<R>(p_pts[id].x, p_pts[id].y, p_pts[id].z, R(0)) : mk4<R>(0, 0, 0, 0);
        seed = at(id);
        status = kSuccess;
        const R vmin = -cube_eps, vmax = R(1000) + cube_eps;
        planes[0] = mk4<R>(1, 0, 0, org.x - vmin);
        planes[1] = mk4<R>(-1, 0, 0, vmax - org.x);
        planes[2] = mk4<R>(0, 1, 0, org.y - vmin);
        planes[3] = mk4<R>(0, -1, 0, vmax - org.y);
        planes[4] = mk4<R>(0, 0, 1, org.z - vmin);
        planes[5] = mk4<R>(0, 0, -1, vmax - org.z);
        for (int i = 0; i < 6; ++i) pnb[i] = -1;
        nb_plane = 6;
        static const uint8_t cube[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2}, {5, 1, 3},
                                           {4, 2, 0}, {4, 0, 3}, {2, 4, 1}, {4, 3, 1}};
        for (int i = 0; i < 8; ++i) verts[i] = V3{cube[i][0], cube[i][1], cube[i][2]};
        nb_vertex = 8;
    }

    // CCNewPlane ConvexCell.cl:314-333
    int new_plane(int vid) {
        if (nb_plane >= P) {
            status = kPlaneOverflow;
            return -1;
        }
        ctr->n_planes++;
        const R4<R> B = at(vid);
        const R4<R> dir = sub4(seed, B);
        const R4<R> ave2 = add4(seed, B);
        R dt = dot3(ave2, dir) - (B.w - seed.w);
        if (double(B.w) <= kNoCellLimits) dt = R(2) * dot3(B, dir);
        const R nrm = std::sqrt(dot3(dir, dir));
        planes[nb_plane] = mk4<R>(dir.x / nrm, dir.y / nrm, dir.z / nrm, -dt / R(2) / nrm);
        pnb[nb_plane] = vid;  // CCPlaneNeighbour, ConvexCell.cl:329-331
        return nb_plane++;
    }

    void new_vertex(int i, int j, int k) {  // CCNewVertex :335-342
        if (nb_vertex >= V) {
            status = kVertexOverflow;
            return;
        }
        verts[nb_vertex++] = V3{uint8_t(i), uint8_t(j), uint8_t(k)};
    }

    // CCComputeVertexCoordinates :344-362
    R4<R> vertex_coords(const V3 &t, bool persp) const {
        const R4<R> p1 = planes[t[0]], p2 = planes[t[1]], p3 = planes[t[2]];
        const R m12 = p2.x * p3.y - p2.y * p3.x;
        const R m13 = p2.x * p3.z - p2.z * p3.x;
        const R m14 = p2.x * p3.w - p2.w * p3.x;
        const R m23 = p2.y * p3.z - p2.z * p3.y;
        const R m24 = p2.y * p3.w - p2.w * p3.y;
        const R m34 = p2.z * p3.w - p2.w * p3.z;
        R4<R> r;
        r.x = -p1.w * m23 - p1.y * m34 + p1.z * m24;
        r.y = p1.x * m34 + p1.w * m13 - p1.z * m14;
        r.z = -p1.x * m24 + p1.y * m14 - p1.w * m12;
        r.w = p1.x * m23 - p1.y * m13 + p1.z * m12;
        if (persp) return mk4<R>(r.x / r.w, r.y / r.w, r.z / r.w, R(1));
        return r;
    }

    // largest squared distance of a vertex from the seed
    R max_r2() const {
        R vd = 0;
        for (int i = 0; i < nb_vertex; ++i) {
            const R4<R> d = sub4(vertex_coords(verts[i], true), seed);
            vd = std::max(dot3(d, d), vd);
        }
        return vd;
    }

    // CCIsSecurityRadiusReached :364-382 (weights ignored, as the reference)
    bool radius_reached(const R4<R> &last) const {
        R vd = 0;
        for (int i = 0; i < nb_vertex; ++i) {
            const R4<R> pc = vertex_coords(verts[i], true);
            const R4<R> d = sub4(pc, seed);
            const R d2 = dot3(d, d);
            vd = std::max(d2, vd);
        }
        ctr->n_radius_vertex += nb_vertex;
        const R4<R> d = sub4(last, seed);
        const R d2 = dot3(d, d);
        const double c = double(seed.w) - wmin;
        if (security_mode == 1 || c <= 0) return d2 > R(4) * vd;
        // exact power stop: a seed B at distance d' >= dl has its bisector at
        // distance (d'^2 + w_B - w_s) / (2 d') >= (d'^2 - c) / (2 d') from the
        // seed, increasing in d'; no later bisector meets the ball of radius R
        // around the seed, which holds the cell, iff dl (dl - 2R) >= c.
        // Border seeds (plane through B, distance d' >= dl >= 2R) never do.
        const double dl = std::sqrt(double(d2)), Rm = std::sqrt(double(vd));
        return dl * (dl - 2.0 * Rm) >= c;
    }

    // CCVertexIsInHalfSpace :400-461
    bool in_half_space(const V3 &t, const R4<R> &e) {
        ctr->n_pred++;
        if (ctr->restricting) ctr->n_pred_restrict++;
        const R4<R> a = planes[t[0]], b = planes[t[1]], c = planes[t[2]];
        if (vhs == 2) {
            const double d = det4<double>(double(a.x), double(b.x), double(c.x), double(e.x), double(a.y),
                                          double(b.y), double(c.y), double(e.y), double(a.z), double(b.z),
                                          double(c.z), double(e.z), double(a.w), double(b.w), double(c.w),
                                          double(e.w));
            return d > 0.0;
        }
        if (vhs == 1) {
            // filtered real evaluation with the reference's magnitude bound
            auto fmaxabs = [](R u, R v) { return std::max(std::fabs(u), std::fabs(v)); };
            R tmp;
            R m12 = b.x * a.y; tmp = -a.x * b.y; R M12 = fmaxabs(m12, tmp); m12 += tmp;
            R m13 = c.x * a.y; tmp = -a.x * c.y; R M13 = fmaxabs(m13, tmp); m13 += tmp;
            R m14 = e.x * a.y; tmp = -a.x * e.y; R M14 = fmaxabs(m14, tmp); m14 += tmp;
            R m23 = c.x * b.y; tmp = -b.x * c.y; R M23 = fmaxabs(m23, tmp); m23 += tmp;
            R m24 = e.x * b.y; tmp = -b.x * e.y; R M24 = fmaxabs(m24, tmp); m24 += tmp;
            R m34 = e.x * c.y; tmp = -c.x * e.y; R M34 = fmaxabs(m34, tmp); m34 += tmp;
            const R m123 = m23 * a.z - m13 * b.z + m12 * c.z;
            const R M123 = std::max(std::max(std::fabs(M23 * a.z), std::fabs(M13 * b.z)), std::fabs(M12 * c.z));
            const R m124 = m24 * a.z - m14 * b.z + m12 * e.z;
            const R M124 = std::max(std::max(std::fabs(M24 * a.z), std::fabs(M14 * b.z)), std::fabs(M12 * e.z));
            const R m134 = m34 * a.z - m14 * c.z + m13 * e.z;
            const R M134 = std::max(std::max(std::fabs(M34 * a.z), std::fabs(M14 * c.z)), std::fabs(M13 * e.z));
            const R m234 = m34 * b.z - m24 * c.z + m23 * e.z;
            const R M234 = std::max(std::max(std::fabs(M34 * b.z), std::fabs(M24 * c.z)), std::fabs(M23 * e.z));
            const R res = m234 * a.w - m134 * b.w + m124 * c.w - m123 * e.w;
            const R M = std::max(std::max(std::fabs(M234 * a.w), std::fabs(M134 * b.w)),
                                 std::max(std::fabs(M124 * c.w), std::fabs(M123 * e.w)));
            if (std::fabs(res) < R(3.e-6) * M) {
                status = kNeedsExactPredicates;
                return res > R(-2.e-7) * M;
            }
            return res > R(0);
        }
        const R d = det4<R>(a.x, b.x, c.x, e.x, a.y, b.y, c.y, e.y, a.z, b.z, c.z, e.z, a.w, b.w, c.w, e.w);
        return d > R(0);
    }

    // CCComputeBoundary :463-515 -- grows the removed region triangle by
    // triangle, keeping a single simple boundary loop in next[].
    int compute_boundary(int nb_removed) {
        for (int i = 0; i < P; ++i) next[i] = kEndOfList;
        int first = kEndOfList;
        int nb_iter = 0;
        int t = nb_vertex;
        while (nb_removed > 0) {
            if (nb_iter++ > 1000) {
                status = kInconsistentBoundary;
                return first;
            }
            const V3 tri = verts[t];
            bool border[3], opp[3];
            for (int e = 0; e < 3; ++e) border[e] = next[tri[e]] != kEndOfList;
            for (int e = 0; e < 3; ++e) opp[e] = next[tri[(e + 1) % 3]] == tri[e];
            bool simple = true;
            for (int e = 0; e < 3; ++e)
                if (!opp[e] && !opp[(e + 1) % 3] && border[(e + 1) % 3]) simple = false;
            if (!opp[0] && !opp[1] && !opp[2]) {
                if (first == kEndOfList) {
                    for (int e = 0; e < 3; ++e) next[tri[e]] = tri[(e + 1) % 3];
                    first = tri[0];
                } else {
                    simple = false;
                }
            }
            if (!simple) {
                t++;
                if (t == nb_vertex + nb_removed) t = nb_vertex;
                continue;
            }
            for (int e = 0; e < 3; ++e)
                if (!opp[e]) next[tri[e]] = tri[(e + 1) % 3];
            for (int e = 0; e < 3; ++e)
                if (opp[e] && opp[(e + 1) % 3]) {
                    const int p = tri[(e + 1) % 3];
                    if (first == p) first = next[p];
                    next[p] = kEndOfList;
                }
            std::swap(verts[t], verts[nb_vertex + nb_removed - 1]);
            t = nb_vertex;
            nb_removed--;
        }
        return first;
    }

    // CCClipByPlane :517-563
    void clip_by_plane(int vid) {
        if (status == kPlaneOverflow) return;
        ctr->n_clip++;
        const R4<R> eqn = planes[vid];
        int nb_removed = 0;
        int i = 0;
        while (i < nb_vertex) {
            if (in_half_space(verts[i], eqn)) {
                nb_vertex--;
                std::swap(verts[i], verts[nb_vertex]);
                nb_removed++;
            } else {
                i++;
            }
        }
        if (nb_vertex < 1) {
            status = kEmptyCell;
            return;
        }
        if (status == kNeedsExactPredicates) return;
        if (nb_removed == 0) {
            --nb_plane;
            return;
        }
        const int first = compute_boundary(nb_removed);
        if (status != kSuccess) return;
        if (first == kEndOfList) return;
        int cir = first;
        do {
            const int nc = next[cir];
            new_vertex(vid, cir, nc);
            if (status != kSuccess) return;
            cir = nc;
        } while (cir != first);
    }

    // CCClipTetFromPointsInfinite :565-575
    void clip_tet(const R4<R> &A, const R4<R> &B, const R4<R> &C, const R4<R> &D) {
        const R4<R> *tri[4][3] = {{&A, &B, &C}, {&A, &D, &B}, {&A, &C, &D}, {&C, &D, &B}};
        for (int f = 0; f < 4; ++f) {
            if (nb_plane >= P) {
                status = kPlaneOverflow;
                return;
            }
            ctr->n_planes++;
            ctr->n_planes_restrict++;
            pnb[nb_plane] = -1;  // ConvexCell.cl:544-546
            planes[nb_plane++] = plane_from_points(*tri[f][0], *tri[f][1], *tri[f][2]);
            clip_by_plane(nb_plane - 1);
            if (status != kSuccess) return;
        }
    }

    // CCExportBaryVolume :577-619.  Accumulates factor/6 * sum into *out.
    // With surf (CCExportBaryVolumeAndArea :621-674): adds, per plane p of a
    // vertex, the signed area of each triangle (A, A + APlaneI, A +
    // ALineII1) times factor to surf[HC p + HC - 1], and for HC = 4 its
    // area-weighted centroid to surf[4p .. 4p + 2].
    void export_bary_volume(R4<R> *out, R factor, R vol_eps2, R *surf = nullptr, int HC = 1) {
        R4<R> s = mk4<R>(0, 0, 0, 0);
        const R4<R> C = seed;
        ctr->n_integrated_vertex += nb_vertex;
        for (int ve = 0; ve < nb_vertex; ++ve) {
            const R4<R> A = vertex_coords(verts[ve], true);
            const R4<R> AC = sub4(C, A);
            R4<R> pl[3], onPlane[3], onLine[3];
            for (int i = 0; i < 3; ++i) pl[i] = planes[verts[ve][i]];
            for (int i = 0; i < 3; ++i) {
                const R4<R> ni = pl[i];
                onPlane[i] = sub4(AC, scale4(dot3(AC, ni) / dot3(ni, ni), ni));
            }
            for (int i = 0; i < 3; ++i) {
                const R4<R> nn = cross3(pl[i], pl[(i + 1) % 3]);
                const R nn2 = dot4(nn, nn);
                if (nn2 < vol_eps2 * dot3(pl[i], pl[i]) * dot3(pl[(i + 1) % 3], pl[(i + 1) % 3]))
                    onLine[i] = mk4<R>(0, 0, 0, 0);
                else
                    onLine[i] = scale4(dot4(AC, nn) / nn2, nn);
            }
            for (int i = 0; i < 3; ++i) {
                for (int st = 0; st < 2; ++st) {
                    const int j = (i + 3 - st) % 3;
                    const R w = R(2 * st - 1) * det3r4(AC, onPlane[i], onLine[j]);
                    s.w += w;
                    const R4<R> sm = add4(add4(AC, onPlane[i]), onLine[j]);
                    s.x += w * (A.x + R(0.25) * sm.x);
                    s.y += w * (A.y + R(0.25) * sm.y);
                    s.z += w * (A.z + R(0.25) * sm.z);
                    if (surf) {
                        const int pid = verts[ve][i];
                        const R area = R(2 * st - 1) * factor * det3r4(onPlane[i], onLine[j], pl[i]) / R(2);
                        if (HC == 1) {
                            surf[pid] += area;
                        } else {
                            const R4<R> t = add4(onPlane[i], onLine[j]);
                            surf[4 * pid + 0] += area * (A.x + t.x / R(3));
                            surf[4 * pid + 1] += area * (A.y + t.y / R(3));
                            surf[4 * pid + 2] += area * (A.z + t.z / R(3));
                            surf[4 * pid + 3] += area;
                        }
                    }
                }
            }
        }
        const R f = factor / R(6);
        *out = add4(*out, scale4(f, s));
    }
};

template <class R>
struct CellResult {
    int status = kSuccess;
    int kused = 0;
    int maxP = 0, maxV = 0, ntri = 0;
};

// Sorted union of two sorted unique lists (merge, ConvexCell.cl:722-749).
// Returns false when the union would exceed cap entries.
inline bool merge_union(const std::vector<int> &a, const int *b, int nb, int cap, std::vector<int> &out) {
    out.clear();
    size_t i = 0;
    int j = 0;
    while (i < a.size() || j < nb) {
        int v;
        if (j >= nb || (i < a.size() && a[i] < b[j])) {
            v = a[i++];
        } else if (i >= a.size() || b[j] < a[i]) {
            v = b[j++];
        } else {
            v = a[i];
            ++i;
            ++j;
        }
        if (int(out.size()) >= cap) return false;
        out.push_back(v);
    }
    return true;
}

// compute_voro_cell ConvexCell.cl:812-1152, one seed.  pts are in the
// working precision; neigs holds K indices into pts (-1 = none).
template <class R>
CellResult<R> compute_cell(const R4<R> *pts, int seed, const int *neigs, int K, int P, int V, int T, int vhs,
                           const GridView<R> *grid, int apex_mode, bool out_domain, int security_mode, double wmin,
                           R4<R> *out, Counters *ctr, bool all_listed, double cull_r2 = 0, int H = 0, int HC = 1,
                           R *hess = nullptr, int *hids = nullptr, std::vector<int> *face_nb = nullptr,
                           bool local_frame = false) {
    CellResult<R> res;
    // Hessian (CHESSIAN): per plane the accumulated area (HC = 1) or
    // area-weighted face centroid and area (HC = 4) over the pieces of the
    // restricted cell; no entries unless the cell succeeds
    std::vector<R> surf;
    if (H > 0) {
        hids[0] = -1;
        surf.assign(size_t(HC) * 256, R(0));
    }
    R *sp = H > 0 ? surf.data() : nullptr;
    const R vol_eps2 = sizeof(R) == sizeof(float) ? R(1.e-6) : R(1.e-12);  // voronoi.cpp:919-922
    if (double(pts[seed].w) <= kNoCellLimits) {
        *out = pts[seed];
        return res;
    }
    *out = mk4<R>(0, 0, 0, 0);
    Cell<R> cc;
    cc.P = P;
    cc.V = V;
    cc.vhs = vhs;
    cc.security_mode = security_mode;
    cc.wmin = wmin;
    cc.ctr = ctr;
    cc.init(pts, seed, grid && grid->g > 0 ? R(0.1) : R(0), local_frame);
    const R4<R> org = cc.org;
    bool complete = false;  // every other seed consumed
    int i = 0;
    for (; i < K; ++i) {
        const int z = neigs[i];
        if (z < 0) {
            complete = true;
            break;
        }
        const int cur = cc.new_plane(z);
        cc.clip_by_plane(cur);
        if (cc.status != kSuccess) {
            res.status = cc.status;
            res.kused = i + 1;
            return res;
        }
        // cull = 1: a cell inside the ball of radius sqrt(cull_r2) about its
        // seed, which holds no domain point, has restricted volume 0
        if (cull_r2 > 0 && double(cc.max_r2()) * 1.001 < cull_r2) {
            res.kused = i + 1;
            return res;
        }
        if (cc.radius_reached(cc.at(z))) break;
    }
    res.kused = std::min(i + 1, K);
    if (complete) res.kused = i;
    // a list that holds every other seed leaves nothing to test: the cell is
    // exact once it is exhausted (the reference reads past the seeds there)
    if (!complete && i == K && all_listed) complete = true;
    res.maxP = cc.nb_plane;
    res.maxV = cc.nb_vertex;
    if (!complete && !cc.radius_reached(cc.at(neigs[K - 1]))) {
        res.status = kSecurityRadiusNotReached;
        return res;
    }
    if (face_nb) {
        // the seeds whose bisector carries a vertex of the (unrestricted)
        // cell: CCUpdateUsedCell, ConvexCell.cl:680-690
        for (int ve = 0; ve < cc.nb_vertex; ++ve)
            for (int i = 0; i < 3; ++i) {
                const int nb = cc.pnb[cc.verts[ve][i]];
                if (nb >= 0) face_nb->push_back(nb);
            }
        std::sort(face_nb->begin(), face_nb->end());
        face_nb->erase(std::unique(face_nb->begin(), face_nb->end()), face_nb->end());
    }
    bool add_volume = true;
    R4<R> Pt0 = mk4<R>(0, 0, 0, 1);
    std::vector<int> tr, tmp;
    if (grid && grid->g > 0) {
        const int g = grid->g;
        const R vs = grid->vs;
        R bmin[3] = {1000, 1000, 1000}, bmax[3] = {0, 0, 0};
        ctr->n_bbox_vertex += cc.nb_vertex;
        for (int ve = 0; ve < cc.nb_vertex; ++ve) {
            const R4<R> p = add4(cc.vertex_coords(cc.verts[ve], true), org);
            const R c[3] = {p.x, p.y, p.z};
            for (int d = 0; d < 3; ++d) {
                if (bmin[d] > c[d]) bmin[d] = c[d];
                if (bmax[d] < c[d]) bmax[d] = c[d];
            }
        }
        int vmin[3], vmax[3];
        for (int d = 0; d < 3; ++d) vmin[d] = std::max(0, int(bmin[d] / vs));
        for (int d = 0; d < 3; ++d) vmax[d] = std::min(g - 1, int(bmax[d] / vs));
        int in_domain = 1;
        // apex_mode 1: first voxel (z,y,x order) whose centre class is known;
        // without an explicit class array, a triangle-free voxel takes the
        // class of its first corner when that corner is not on the border.
        int free_flag = 1;
        R4<R> freePt = Pt0;
        for (int z = vmin[2]; z <= vmax[2]; z++) {
            for (int y = vmin[1]; y <= vmax[1]; y++) {
                for (int x = vmin[0]; x <= vmax[0]; x++) {
                    const int id = grid->id(x, y, z);
                    const signed char f = grid->flags[id];
                    if (in_domain == 1 && f != 1) {
                        in_domain = f;
                        Pt0 = mk4<R>(R(x) * vs, R(y) * vs, R(z) * vs, R(1));
                    }
                    const int o0 = grid->offs[id], o1 = grid->offs[id + 1];
                    if (free_flag == 1) {
                        const int cf = grid->centre ? grid->centre[id] : (o0 == o1 ? f : 1);
                        if (cf != 1) {
                            free_flag = cf;
                            freePt = mk4<R>((R(x) + R(kApexFrac[0])) * vs, (R(y) + R(kApexFrac[1])) * vs,
                                            (R(z) + R(kApexFrac[2])) * vs, R(1));
                        }
                    }
                    if (o0 == o1) continue;
                    if (T == 0) {
                        res.status = kTriangleOverflow;
                        return res;
                    }
                    if (!merge_union(tr, grid->list + o0, o1 - o0, T, tmp)) {
                        res.status = kTriangleOverflow;
                        return res;
                    }
                    tr.swap(tmp);
                }
            }
        }
        res.ntri = int(tr.size());
        if (apex_mode == 1 && free_flag != 1) {
            in_domain = free_flag;
            Pt0 = freePt;
        }
        if (in_domain == 1) {
            if (!out_domain) {
                res.status = kFindAnotherBeginningVertex;
                return res;
            }
            // GRID_OUT_DOMAIN fallback ConvexCell.cl:985-1051
            auto flag = [&](int x, int y, int z) { return grid->flags[grid->id(x, y, z)]; };
            auto corner = [&](int x, int y, int z) { return mk4<R>(R(x) * vs, R(y) * vs, R(z) * vs, R(1)); };
            if (vmax[2] + 1 < g) {
                const int z = vmax[2] + 1;
                for (int y = vmin[1]; y <= vmax[1] + 1; y++) {
                    if (y >= g) break;
                    for (int x = vmin[0]; x <= vmax[0] + 1; x++) {
                        if (x >= g) break;
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            Pt0 = corner(x, y, z);
                            break;
                        }
                    }
                    if (in_domain != 1) break;
                }
            } else {
                in_domain = 0;
                Pt0 = corner(vmin[0], vmin[1], vmax[2] + 1);
            }
            if (in_domain == 1 && vmax[1] + 1 < g) {
                const int y = vmax[1] + 1;
                for (int z = vmin[2]; z <= vmax[2]; z++) {
                    for (int x = vmin[0]; x <= vmax[0] + 1; x++) {
                        if (x >= g) break;
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            Pt0 = corner(x, y, z);
                            break;
                        }
                    }
                    if (in_domain != 1) break;
                }
            } else if (in_domain == 1) {
                in_domain = 0;
                Pt0 = corner(vmin[0], vmax[1] + 1, vmin[2]);
            }
            if (in_domain == 1 && vmax[0] + 1 < g) {
                const int x = vmax[0] + 1;
                for (int z = vmin[2]; z <= vmax[2]; z++) {
                    for (int y = vmin[1]; y <= vmax[1]; y++) {
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            Pt0 = corner(x, y, z);
                            break;
                        }
                    }
                    if (in_domain != 1) break;
                }
            } else if (in_domain == 1) {
                in_domain = 0;
                Pt0 = corner(vmax[0] + 1, vmin[1], vmin[2]);
            }
            if (in_domain == 1) {
                res.status = kFindAnotherBeginningVertex;
                return res;
            }
        }
        add_volume = in_domain == 2;
        Pt0 = sub4(Pt0, org);  // the apex in the cell's frame
    }
    if (add_volume) cc.export_bary_volume(out, R(1), vol_eps2, sp, HC);
    ctr->restricting = true;
    struct PhaseEnd {
        Counters *c;
        ~PhaseEnd() { c->restricting = false; }
    } phase_end{ctr};
    if (!tr.empty()) {
        // backup / restore ConvexCell.cl:692-715
        const int bk_nv = cc.nb_vertex, bk_np = cc.nb_plane, bk_st = cc.status;
        std::vector<typename Cell<R>::V3> bk(cc.verts, cc.verts + cc.nb_vertex);
        for (size_t t = 0; t < tr.size(); ++t) {
            const int tid = tr[t];
            R4<R> P1 = sub4(grid->pts[size_t(grid->tris[3 * tid + 0])], org);
            R4<R> P2 = sub4(grid->pts[size_t(grid->tris[3 * tid + 1])], org);
            R4<R> P3 = sub4(grid->pts[size_t(grid->tris[3 * tid + 2])], org);
            const bool go_out = det3r4(sub4(P1, Pt0), sub4(P2, Pt0), sub4(P3, Pt0)) > R(0);
            if (!go_out) std::swap(P2, P3);
            cc.clip_tet(Pt0, P1, P2, P3);
            res.maxP = std::max(res.maxP, cc.nb_plane);
            res.maxV = std::max(res.maxV, cc.nb_vertex);
            if (cc.status == kEmptyCell) {
                cc.nb_vertex = bk_nv;
                cc.nb_plane = bk_np;
                std::copy(bk.begin(), bk.end(), cc.verts);
                cc.status = kSuccess;
                continue;
            }
            if (cc.status != kSuccess) {
                // the reference zeroes only .w here (ConvexCell.cl:1104-1107);
                // the oracle clears all four moments of the failed cell
                *out = mk4<R>(0, 0, 0, 0);
                res.status = cc.status;
                return res;
            }
            cc.export_bary_volume(out, go_out ? R(1) : R(-1), vol_eps2, sp, HC);
            cc.nb_vertex = bk_nv;
            cc.nb_plane = bk_np;
            std::copy(bk.begin(), bk.end(), cc.verts);
            cc.status = bk_st;
        }
    }
    res.status = kSuccess;
    if (local_frame) {
        // moments about the seed -> moments in the [0,1000]^3 frame
        out->x += org.x * out->w;
        out->y += org.y * out->w;
        out->z += org.z * out->w;
        if (H > 0 && HC == 4)
            for (int p = 0; p < cc.nb_plane; ++p) {
                surf[size_t(4 * p + 0)] += org.x * surf[size_t(4 * p + 3)];
                surf[size_t(4 * p + 1)] += org.y * surf[size_t(4 * p + 3)];
                surf[size_t(4 * p + 2)] += org.z * surf[size_t(4 * p + 3)];
            }
    }
    if (H > 0) {
        // planes towards a seed whose face keeps more than 0.2 of area, in
        // plane order; more than H of them: hessian_overflow
        // (ConvexCell.cl:1133-1151)
        int cnt = 0;
        for (int p = 0; p < cc.nb_plane; ++p) {
            if (cc.pnb[p] < 0 || surf[size_t(HC) * p + HC - 1] <= R(0.2)) continue;
            if (cnt >= H) {
                res.status = kHessianOverflow;
                return res;
            }
            for (int c = 0; c < HC; ++c) hess[size_t(HC) * cnt + c] = surf[size_t(HC) * p + c];
            hids[cnt] = cc.pnb[p];
            ++cnt;
        }
        if (cnt < H) hids[cnt] = -1;
    }
    return res;
}

// ---------------------------------------------------------------------------
// kNN (knearests.cpp / knearests.cl)
// ---------------------------------------------------------------------------
int knn_dim(int n) {
    // knearests.cpp:63, evaluated in double
    return std::max(1, int(std::lround(std::cbrt(double(n) / 3.1))));
}

// cellFromPoint knearests.cpp:35-46: floor(p * (real)dim / 1000) per axis.
template <class R>
inline int cell_coord(R c, int dim) {
    int i = int(std::floor(c * R(dim) / R(1000.f)));
    return std::max(0, std::min(i, dim - 1));
}

template <class R>
struct KnnGrid {
    int dim = 0;
    std::vector<int> counters;  // dim^3+1 exclusive prefix
    std::vector<int> perm;      // stored -> original
    std::vector<R4<R>> stored;  // sorted seeds
    std::vector<int> cellOf;    // per stored point: cell index
};

template <class R>
KnnGrid<R> build_grid(const std::vector<R4<R>> &pts) {
    KnnGrid<R> kg;
    const int n = int(pts.size());
    const int dim = knn_dim(n);
    kg.dim = dim;
    const size_t nc = size_t(dim) * dim * dim;
    std::vector<int> cell(static_cast<size_t>(n));
    kg.counters.assign(nc + 1, 0);
    for (int i = 0; i < n; ++i) {
        const int c = cell_coord(pts[size_t(i)].x, dim) + cell_coord(pts[size_t(i)].y, dim) * dim +
                      cell_coord(pts[size_t(i)].z, dim) * dim * dim;
        cell[size_t(i)] = c;
        kg.counters[size_t(c)]++;
    }
    int acc = 0;
    for (auto &c : kg.counters) {
        const int v = c;
        c = acc;
        acc += v;
    }
    std::vector<int> cursor(kg.counters.begin(), kg.counters.end() - 1);
    kg.perm.resize(size_t(n));
    kg.stored.resize(size_t(n));
    kg.cellOf.resize(size_t(n));
    for (int i = 0; i < n; ++i) {  // stable: original order inside a cell
        const int pos = cursor[size_t(cell[size_t(i)])]++;
        kg.perm[size_t(pos)] = i;
        kg.stored[size_t(pos)] = pts[size_t(i)];
        kg.cellOf[size_t(pos)] = cell[size_t(i)];
    }
    return kg;
}

template <class R>
inline R dist2(const R4<R> &a, const R4<R> &b) {
    const R dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return (dx * dx + dy * dy) + dz * dz;
}

struct KnnKey {
    double d;
    int orig;
    int stored;
    bool operator<(const KnnKey &o) const { return d < o.d || (d == o.d && orig < o.orig); }
};

// Exact K nearest of stored point q by rings of Chebyshev cell distance.
template <class R>
void knn_query_grid(const KnnGrid<R> &kg, int q, int K, int *out, std::vector<KnnKey> &heap) {
    const int dim = kg.dim;
    const int n = int(kg.stored.size());
    const R4<R> p = kg.stored[size_t(q)];
    const int c = kg.cellOf[size_t(q)];
    const int ci = c % dim, cj = (c / dim) % dim, ck = c / (dim * dim);
    heap.clear();
    auto worse = [](const KnnKey &a, const KnnKey &b) { return a < b; };  // max-heap on key
    const double cs = 1000.0 / dim;
    auto visit_cell = [&](int x, int y, int z) {
        const int id = x + y * dim + z * dim * dim;
        for (int s = kg.counters[size_t(id)]; s < kg.counters[size_t(id) + 1]; ++s) {
            if (s == q) continue;
            const double d = double(dist2(p, kg.stored[size_t(s)]));
            const KnnKey key{d, kg.perm[size_t(s)], s};
            if (int(heap.size()) < K) {
                heap.push_back(key);
                std::push_heap(heap.begin(), heap.end(), worse);
            } else if (key < heap.front()) {
                std::pop_heap(heap.begin(), heap.end(), worse);
                heap.back() = key;
                std::push_heap(heap.begin(), heap.end(), worse);
            }
        }
    };
    const int rmax = dim;  // Chebyshev distance never exceeds dim-1
    for (int r = 0; r <= rmax; ++r) {
        if (r >= 1 && int(heap.size()) == K) {
            // every point of ring r is at least (r-1) cells away along one axis
            const double lb = double(r - 1) * cs * (1.0 - 1e-4);
            const double lb2 = lb * lb * (1.0 - 1e-6);
            if (heap.front().d < lb2) break;
        }
        if (r > 0 && ci - r < 0 && ci + r >= dim && cj - r < 0 && cj + r >= dim && ck - r < 0 && ck + r >= dim)
            break;  // the ring lies entirely outside the grid
        for (int dz = -r; dz <= r; ++dz) {
            const int z = ck + dz;
            if (z < 0 || z >= dim) continue;
            for (int dy = -r; dy <= r; ++dy) {
                const int y = cj + dy;
                if (y < 0 || y >= dim) continue;
                const bool full = std::abs(dz) == r || std::abs(dy) == r;
                if (full) {
                    for (int dx = -r; dx <= r; ++dx) {
                        const int x = ci + dx;
                        if (x < 0 || x >= dim) continue;
                        visit_cell(x, y, z);
                    }
                } else {
                    if (ci - r >= 0) visit_cell(ci - r, y, z);
                    if (r > 0 && ci + r < dim) visit_cell(ci + r, y, z);
                }
            }
        }
    }
    std::sort(heap.begin(), heap.end());
    for (int i = 0; i < K; ++i) out[i] = i < int(heap.size()) ? heap[size_t(i)].stored : -1;
    (void)n;
}

template <class R>
void knn_query_brute(const KnnGrid<R> &kg, int q, int K, int *out, std::vector<KnnKey> &all) {
    const int n = int(kg.stored.size());
    const R4<R> p = kg.stored[size_t(q)];
    all.clear();
    for (int s = 0; s < n; ++s) {
        if (s == q) continue;
        all.push_back(KnnKey{double(dist2(p, kg.stored[size_t(s)])), kg.perm[size_t(s)], s});
    }
    const int k = std::min(K, int(all.size()));
    std::partial_sort(all.begin(), all.begin() + k, all.end());
    for (int i = 0; i < K; ++i) out[i] = i < k ? all[size_t(i)].stored : -1;
}

template <class R>
std::vector<R4<R>> convert_pts(const double *xyzw, int n) {
    std::vector<R4<R>> p(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        p[size_t(i)] = mk4<R>(R(xyzw[4 * i]), R(xyzw[4 * i + 1]), R(xyzw[4 * i + 2]), R(xyzw[4 * i + 3]));
    return p;
}

void set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

template <class R>
int knn_all(const KnnGrid<R> &kg, int K, int method, const int *ids, int nids, int *knn_rows) {
    const long long nn = nids;
#pragma omp parallel
    {
        std::vector<KnnKey> buf;
#pragma omp for schedule(dynamic, 256)
        for (long long t = 0; t < nn; ++t) {
            const int q = ids ? ids[t] : int(t);
            if (method == 1)
                knn_query_brute(kg, q, K, knn_rows + t * K, buf);
            else
                knn_query_grid(kg, q, K, knn_rows + t * K, buf);
        }
    }
    return kg.dim;
}

// ---------------------------------------------------------------------------
// cull = 1: the library's tests for cells whose restriction is empty.  A
// domain voxel is one that holds triangles or whose centre class is not
// "outside"; dist is the Chebyshev voxel distance to the nearest one (BFS over
// the 26 neighbours, capped at 255).  A seed whose voxel lies d >= 2 voxels
// away is at least (d - 1) vs from every domain point.
// ---------------------------------------------------------------------------
constexpr double kCoverRel = 1e-9, kCoverAbs = 1e-6;

std::vector<uint8_t> domain_distance(const orc_grid *gr) {
    const int g = gr->size;
    const size_t nv = size_t(g) * g * g;
    std::vector<uint8_t> dist(nv, 255);
    std::vector<int> q;
    for (size_t v = 0; v < nv; ++v) {
        const bool tri = gr->offsets[v + 1] > gr->offsets[v];
        const bool in = gr->centre ? gr->centre[v] != 0 : gr->in_domain[v] != 0;
        if (tri || in) {
            dist[v] = 0;
            q.push_back(int(v));
        }
    }
    for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h], d = dist[size_t(v)];
        if (d >= 254) continue;
        const int x = v % g, y = (v / g) % g, z = v / (g * g);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int X = x + dx, Y = y + dy, Z = z + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= g || Y >= g || Z >= g) continue;
                    const size_t u = size_t(X) + size_t(Y) * g + size_t(Z) * g * g;
                    if (dist[u] == 255) {
                        dist[u] = uint8_t(d + 1);
                        q.push_back(int(u));
                    }
                }
    }
    return dist;
}

// Upper bound M of U(x) - w_min over the domain points x, U(x) the smallest
// power distance of x to a seed: for each domain voxel (centre c, half
// diagonal r) the smallest (|c - t| + r)^2 + w_t - w_min over the seeds t of
// the 3x3x3 kNN-grid cells around c; +inf when those cells are empty.
template <class R>
double cover_bound(const KnnGrid<R> &kg, const std::vector<uint8_t> &dist, int g, double wmin) {
    const double vs = 1000.0 / double(g), rv = 0.8660254037844387 * vs * (1.0 + 1e-12);
    const int dim = kg.dim;
    double M = 0;
    for (size_t v = 0; v < dist.size(); ++v) {
        if (dist[v] != 0) continue;
        const double cx = (double(v % g) + 0.5) * vs, cy = (double((v / g) % g) + 0.5) * vs,
                     cz = (double(v / (size_t(g) * g)) + 0.5) * vs;
        auto cc = [&](double c) { return std::max(0, std::min(dim - 1, int(c * double(dim) / 1000.0))); };
        const int ci = cc(cx), cj = cc(cy), ck = cc(cz);
        double best = std::numeric_limits<double>::infinity();
        for (int z = std::max(0, ck - 1); z <= std::min(dim - 1, ck + 1); ++z)
            for (int y = std::max(0, cj - 1); y <= std::min(dim - 1, cj + 1); ++y) {
                const int row = (y + z * dim) * dim;
                for (int j = kg.counters[size_t(row + std::max(0, ci - 1))],
                         je = kg.counters[size_t(row + std::min(dim - 1, ci + 1) + 1)];
                     j < je; ++j) {
                    const R4<R> t = kg.stored[size_t(j)];
                    const double dx = double(t.x) - cx, dy = double(t.y) - cy, dz = double(t.z) - cz;
                    const double dd = std::sqrt(dx * dx + dy * dy + dz * dz) + rv;
                    best = std::min(best, dd * dd + (double(t.w) - wmin));
                }
            }
        M = std::max(M, best * (1.0 + kCoverRel) + kCoverAbs);
    }
    return M;
}

// One clip pass over a list of stored ids (VoroAlgoComputeData::launch,
// voronoi.cpp:861-1036): kNN with K, then one cell per id.
template <class R>
void clip_pass(const KnnGrid<R> &kg, const std::vector<int> &ids, int K, int P, int V, int T, int vhs,
               const GridView<R> *grid, int apex_mode, bool out_domain, int security_mode, double wmin,
               std::vector<R4<R>> &out,
               std::vector<int> &status, std::vector<int> &kused, Counters &ctr, double &t_knn, double &t_clip,
               const std::vector<double> &cull_r2, int H = 0, int HC = 1, R *hess = nullptr, int *hids = nullptr,
               bool local_frame = false) {
    const long long nids = (long long)ids.size();
    if (nids == 0) return;
    std::vector<int> nb(size_t(nids) * size_t(K));
    auto t0 = std::chrono::steady_clock::now();
    knn_all(kg, K, 0, ids.data(), int(nids), nb.data());
    auto t1 = std::chrono::steady_clock::now();
    Counters total;
#pragma omp parallel
    {
        Counters local;
#pragma omp for schedule(dynamic, 256)
        for (long long t = 0; t < nids; ++t) {
            const int s = ids[size_t(t)];
            CellResult<R> r = compute_cell<R>(kg.stored.data(), s, nb.data() + t * K, K, P, V, T, vhs, grid,
                                              apex_mode, out_domain, security_mode, wmin, &out[size_t(s)], &local,
                                              K >= int(kg.stored.size()) - 1,
                                              cull_r2.empty() ? 0.0 : cull_r2[size_t(s)], H, HC,
                                              H > 0 ? hess + size_t(s) * H * HC : nullptr,
                                              H > 0 ? hids + size_t(s) * H : nullptr, nullptr, local_frame);
            status[size_t(s)] = r.status;
            kused[size_t(s)] = r.kused;
        }
#pragma omp critical
        total.add(local);
    }
    auto t2 = std::chrono::steady_clock::now();
    ctr.add(total);
    t_knn += std::chrono::duration<double>(t1 - t0).count();
    t_clip += std::chrono::duration<double>(t2 - t1).count();
}

template <class R>
int solve_impl(const double *xyzw, int n, const orc_grid *gr, const orc_params *prm, double *out_mg,
               int *out_status, int *out_kused, orc_stats *st, const int *ids_orig = nullptr, int n_ids = 0,
               double *out_hess = nullptr, int *out_hids = nullptr) {
    auto tstart = std::chrono::steady_clock::now();
    set_threads(prm->threads);
    const std::vector<R4<R>> pts = convert_pts<R>(xyzw, n);
    const KnnGrid<R> kg = build_grid(pts);
    GridView<R> gv = make_grid_view<R>(gr);
    const GridView<R> *grid = gr ? &gv : nullptr;
    // double and long double: the USE_DOUBLE pass policy; long double
    // evaluates the predicates in its own precision (det4 of VHSAlgo 0)
    const bool dbl = sizeof(R) >= sizeof(double);
    const bool local = prm->frame == 1;
    const int sm = prm->security_mode;
    double wmin = std::numeric_limits<double>::infinity();
    for (const auto &p : pts)
        if (double(p.w) > kNoCellLimits) wmin = std::min(wmin, double(p.w));
    if (!(wmin < std::numeric_limits<double>::infinity())) wmin = 0;
    std::vector<R4<R>> out(size_t(n), mk4<R>(0, 0, 0, 0));
    std::vector<int> status(size_t(n), kSuccess), kused(size_t(n), 0);
    // Hessian per stored seed (H entries of HC values, neighbour stored ids)
    const int H = out_hess && out_hids ? std::max(prm->H, 0) : 0, HC = prm->HC == 4 ? 4 : 1;
    std::vector<R> hess(size_t(n) * H * HC, R(0));
    std::vector<int> hids(size_t(n) * H, -1);
    Counters ctr;
    double t_knn = 0, t_clip = 0;
    int K = prm->K, P = prm->P, V = prm->V, T = prm->T;
    int vhs = sizeof(R) > sizeof(double) ? 0 : (dbl ? 2 : prm->vhs0);
    int passes = 0;
    std::vector<int> all;
    if (ids_orig) {
        std::vector<int> inv(static_cast<size_t>(n));
        for (int s = 0; s < n; ++s) inv[size_t(kg.perm[size_t(s)])] = s;
        all.resize(size_t(n_ids));
        for (int i = 0; i < n_ids; ++i) all[size_t(i)] = inv[size_t(ids_orig[i])];
    } else {
        all.resize(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) all[size_t(i)] = i;
    }

    // cull = 1 (the library's shortcuts, see rpd_cull.hip and the bisector
    // kernel): per stored seed the squared distance below which its cell
    // holds no domain point, and the seeds whose cover test proves volume 0
    std::vector<double> cull_r2;
    std::vector<int> todo = all;  // cells to clip (all keeps the output order)
    if (prm->cull && gr) {
        const int g = gr->size;
        const double vs = 1000.0 / double(g);
        const std::vector<uint8_t> dist = domain_distance(gr);
        auto vox = [&](double x) { return int(std::min(std::max(x / vs, 0.0), double(g - 1))); };
        auto dist_of = [&](int s) {
            const double *p = xyzw + 4 * size_t(kg.perm[size_t(s)]);
            return int(dist[size_t(vox(p[0])) + size_t(vox(p[1])) * g + size_t(vox(p[2])) * g * g]);
        };
        cull_r2.assign(size_t(n), 0.0);
        for (int s = 0; s < n; ++s) {
            const int d = dist_of(s);
            if (d >= 2) cull_r2[size_t(s)] = double(d - 1) * vs * double(d - 1) * vs;
        }
        bool border = false;
        for (const auto &p : pts) border = border || double(p.w) <= kNoCellLimits;
        if (sm == 0 && !border) {
            const double M = cover_bound(kg, dist, g, wmin);
            std::vector<int> kept;
            kept.reserve(all.size());
            for (int s : all) {
                const double D = double(std::max(0, dist_of(s) - 1)) * vs;
                const double w = xyzw[4 * size_t(kg.perm[size_t(s)]) + 3];
                if (D * D + (w - wmin) > M * (1.0 + kCoverRel) + kCoverAbs)
                    status[size_t(s)] = kSuccess;  // out stays 0
                else
                    kept.push_back(s);
            }
            todo.swap(kept);
        }
    }

    if (prm->single_pass) {
        clip_pass(kg, todo, K, P, V, T, vhs, grid, prm->apex_mode, grid && gv.has_border_vertices, sm, wmin, out,
                  status, kused, ctr, t_knn, t_clip, cull_r2, H, HC, hess.data(), hids.data(), local);
        passes = 1;
    } else {
        // first pass: T = 0 defers every boundary cell (voronoi.cpp:1294-1297)
        clip_pass(kg, todo, K, P, V, 0, vhs, grid, prm->apex_mode, false, sm, wmin, out, status, kused, ctr, t_knn,
                  t_clip, cull_r2, H, HC, hess.data(), hids.data(), local);
        passes = 1;
        auto collect = [&](const std::vector<int> &from, auto cond, std::vector<int> &ids0, std::vector<int> &ids1) {
            ids0.clear();
            ids1.clear();
            for (int s : from) {
                if (status[size_t(s)] == kSuccess) continue;
                ids0.push_back(s);
                if (cond(status[size_t(s)])) ids1.push_back(s);
            }
        };
        std::vector<int> ids0, ids1;
        collect(todo, [](int s) { return s == kTriangleOverflow; }, ids0, ids1);
        if (!ids1.empty()) {  // tr_overflow pass (voronoi.cpp:1304-1310)
            clip_pass(kg, ids1, K, P, V, T, vhs, grid, prm->apex_mode, false, sm, wmin, out, status, kused, ctr,
                      t_knn, t_clip, cull_r2, H, HC, hess.data(), hids.data(), local);
            passes++;
        }
        if (!ids0.empty()) {
            // extra passes (voronoi.cpp:1311-1376)
            vhs = sizeof(R) > sizeof(double) ? 0 : (dbl ? 2 : prm->vhs1);
            const bool out_domain = grid && gv.has_border_vertices;
            std::vector<int> cur = ids0;
            collect(cur, [](int) { return true; }, ids0, ids1);
            for (int step = 0; step < prm->max_extra_steps && !ids0.empty(); ++step) {
                if (!dbl && step > 0 && vhs <= 1) break;
                int hist[kNumStatus] = {0};
                for (int s : ids0) hist[status[size_t(s)]]++;
                bool relaunch = false;
                if (hist[kVertexOverflow]) { relaunch = true; V = int(V * 1.5); }
                if (hist[kPlaneOverflow]) { relaunch = true; P = int(P * 1.5); }
                if (hist[kSecurityRadiusNotReached]) { relaunch = true; K = int(K * 1.5); }
                if (hist[kTriangleOverflow]) { relaunch = true; T = int(T * 1.5); }
                if (prm->fix_exact_requeue && (hist[kNeedsExactPredicates] || hist[kFindAnotherBeginningVertex] ||
                                               hist[kInconsistentBoundary]))
                    relaunch = true;
                if (!relaunch) break;
                if (P >= 256) P = 255;  // uchar plane ids (voronoi.cpp:867-870)
                K = std::min(K, n - 1 > 0 ? n - 1 : 1);
                clip_pass(kg, ids1, K, P, V, T, vhs, grid, prm->apex_mode, out_domain, sm, wmin, out, status, kused,
                          ctr, t_knn, t_clip, cull_r2, H, HC, hess.data(), hids.data(), local);
                passes++;
                cur = ids0;
                collect(cur, [](int s) { return s != kEmptyCell; }, ids0, ids1);
            }
        }
    }
    for (size_t k = 0; k < all.size(); ++k) {
        const int s = all[k];
        const size_t o = ids_orig ? k : size_t(kg.perm[size_t(s)]);
        const R4<R> v = out[size_t(s)];
        out_mg[4 * o + 0] = double(v.x);
        out_mg[4 * o + 1] = double(v.y);
        out_mg[4 * o + 2] = double(v.z);
        out_mg[4 * o + 3] = double(v.w);
        if (out_status) out_status[o] = status[size_t(s)];
        if (out_kused) out_kused[o] = kused[size_t(s)];
        for (int h = 0; h < H; ++h) {
            const int nb = hids[size_t(s) * H + h];
            out_hids[o * H + h] = nb >= 0 ? kg.perm[size_t(nb)] : -1;
            for (int c = 0; c < HC; ++c) out_hess[(o * H + h) * HC + c] = double(hess[(size_t(s) * H + h) * HC + c]);
            if (nb < 0) break;
        }
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_pred = ctr.n_pred;
        st->n_radius_vertex = ctr.n_radius_vertex;
        st->n_planes = ctr.n_planes;
        st->n_bbox_vertex = ctr.n_bbox_vertex;
        st->n_integrated_vertex = ctr.n_integrated_vertex;
        st->n_clip = ctr.n_clip;
        st->n_pred_restrict = ctr.n_pred_restrict;
        st->n_planes_restrict = ctr.n_planes_restrict;
        for (int s : all) st->status_hist[status[size_t(s)]]++;
        st->passes = passes;
        st->final_K = K;
        st->final_P = P;
        st->final_V = V;
        st->final_T = T;
        st->seconds_knn = t_knn;
        st->seconds_clip = t_clip;
        st->seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
    }
    return 0;
}

}  // namespace

extern "C" {

int orc_knn_dim(int n) { return knn_dim(n); }

int orc_knn(const double *xyzw, int n, int K, int precision, int method, int *perm, int *knn, int *counters_out) {
    if (n <= 0 || K <= 0 || !xyzw) return -1;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        const auto pts = convert_pts<R>(xyzw, n);
        const KnnGrid<R> kg = build_grid(pts);
        if (perm) std::copy(kg.perm.begin(), kg.perm.end(), perm);
        if (counters_out) std::copy(kg.counters.begin(), kg.counters.end(), counters_out);
        if (knn) knn_all(kg, K, method, nullptr, n, knn);
        return kg.dim;
    };
    return precision == 0 ? run(float(0)) : run(double(0));
}

int orc_solve(const double *xyzw, int n, const orc_grid *grid, const orc_params *prm, double *out_mg,
              int *out_status, int *out_kused, orc_stats *st) {
    if (n <= 1 || !xyzw || !prm || !out_mg) return -1;
    if (prm->P >= 256 || prm->K <= 0 || prm->V <= 0) return -2;
    orc_params p = *prm;
    p.K = std::min(p.K, n - 1);
    if (p.precision == 2) return solve_impl<long double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st);
    return p.precision == 0 ? solve_impl<float>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st)
                            : solve_impl<double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st);
}

int orc_solve_hessian(const double *xyzw, int n, const orc_grid *grid, const orc_params *prm, double *out_mg,
                      int *out_status, int *out_kused, orc_stats *st, double *out_hess, int *out_hids) {
    if (n <= 1 || !xyzw || !prm || !out_mg) return -1;
    if (prm->P >= 256 || prm->K <= 0 || prm->V <= 0 || prm->H < 0 || (prm->HC != 1 && prm->HC != 4)) return -2;
    orc_params p = *prm;
    p.K = std::min(p.K, n - 1);
    if (p.precision == 2)
        return solve_impl<long double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, nullptr, 0, out_hess,
                                       out_hids);
    return p.precision == 0
               ? solve_impl<float>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, nullptr, 0, out_hess, out_hids)
               : solve_impl<double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, nullptr, 0, out_hess,
                                    out_hids);
}

// Face neighbours of the unrestricted power cells of the seeds ids_orig
// (FP64; caps grown x2 / x1.5 until the cell succeeds, at most 6 times):
// out_nb[i * max_nb ...] their original indices, ascending, -1 padded.
// Returns 0, or -1 (bad arguments), -2 (a cell failed), -3 (more than max_nb).
int orc_cell_neighbours(const double *xyzw, int n, const int *ids_orig, int n_ids, int K, int P, int V,
                        int security_mode, int *out_nb, int max_nb) {
    if (n <= 1 || !xyzw || !ids_orig || !out_nb || max_nb <= 0 || K <= 0) return -1;
    const std::vector<R4<double>> pts = convert_pts<double>(xyzw, n);
    const KnnGrid<double> kg = build_grid(pts);
    double wmin = std::numeric_limits<double>::infinity();
    for (const auto &p : pts)
        if (double(p.w) > kNoCellLimits) wmin = std::min(wmin, double(p.w));
    if (!(wmin < std::numeric_limits<double>::infinity())) wmin = 0;
    std::vector<int> inv(static_cast<size_t>(n));
    for (int s = 0; s < n; ++s) inv[size_t(kg.perm[size_t(s)])] = s;
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n_ids; ++i) {
        const int s = inv[size_t(ids_orig[i])];
        int k = std::min(K, n - 1), p = P, v = V;
        std::vector<int> nb, row, heap;
        Counters ctr;
        bool ok = false;
        for (int step = 0; step < 7 && !ok; ++step) {
            row.assign(size_t(k), -1);
            std::vector<KnnKey> hp;
            knn_query_grid(kg, s, k, row.data(), hp);
            R4<double> out;
            nb.clear();
            const CellResult<double> r =
                compute_cell<double>(kg.stored.data(), s, row.data(), k, p, v, 0, 2, nullptr, 1, false, security_mode,
                                     wmin, &out, &ctr, k >= n - 1, 0.0, 0, 1, nullptr, nullptr, &nb);
            if (r.status == kSuccess || r.status == kEmptyCell) {  // an empty power cell has no neighbour
                if (r.status == kEmptyCell) nb.clear();
                ok = true;
            } else {
                if (r.status == kSecurityRadiusNotReached) k = std::min(n - 1, 2 * k);
                if (r.status == kPlaneOverflow) p = std::min(255, p * 3 / 2);
                if (r.status == kVertexOverflow) v = std::min(255, v * 3 / 2);
                if (r.status != kSecurityRadiusNotReached && r.status != kPlaneOverflow &&
                    r.status != kVertexOverflow)
                    break;
            }
        }
        int *o = out_nb + size_t(i) * max_nb;
        for (int j = 0; j < max_nb; ++j) o[j] = -1;
        if (!ok) {
#pragma omp critical
            rc = -2;
            continue;
        }
        std::vector<int> orig;
        for (int z : nb) orig.push_back(kg.perm[size_t(z)]);
        std::sort(orig.begin(), orig.end());
        if (int(orig.size()) > max_nb) {
#pragma omp critical
            rc = -3;
            continue;
        }
        for (size_t j = 0; j < orig.size(); ++j) o[j] = orig[j];
    }
    return rc;
}

int orc_solve_ids(const double *xyzw, int n, const int *ids_orig, int n_ids, const orc_grid *grid,
                  const orc_params *prm, double *out_mg, int *out_status, int *out_kused, orc_stats *st) {
    if (n <= 1 || !xyzw || !prm || !out_mg || !ids_orig || n_ids <= 0) return -1;
    for (int i = 0; i < n_ids; ++i)
        if (ids_orig[i] < 0 || ids_orig[i] >= n) return -1;
    if (prm->P >= 256 || prm->K <= 0 || prm->V <= 0) return -2;
    orc_params p = *prm;
    p.K = std::min(p.K, n - 1);
    if (p.precision == 2)
        return solve_impl<long double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, ids_orig, n_ids);
    return p.precision == 0
               ? solve_impl<float>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, ids_orig, n_ids)
               : solve_impl<double>(xyzw, n, grid, &p, out_mg, out_status, out_kused, st, ids_orig, n_ids);
}

int orc_clip_cell(const double *xyzw, int n, int seed, const int *neigs, int K, int P, int V, int T, int precision,
                  int vhs, const orc_grid *grid, int apex_mode, int out_domain, int security_mode, double wmin,
                  double *out4, int *nb_vertex, int *nb_plane) {
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        const auto pts = convert_pts<R>(xyzw, n);
        GridView<R> gv = make_grid_view<R>(grid);
        Counters ctr;
        R4<R> out = mk4<R>(0, 0, 0, 0);
        CellResult<R> r = compute_cell<R>(pts.data(), seed, neigs, K, P, V, T, sizeof(R) == 8 ? 2 : vhs,
                                          grid ? &gv : nullptr, apex_mode, out_domain != 0, security_mode, wmin,
                                          &out, &ctr, K >= n - 1);
        out4[0] = double(out.x);
        out4[1] = double(out.y);
        out4[2] = double(out.z);
        out4[3] = double(out.w);
        if (nb_vertex) *nb_vertex = r.maxV;
        if (nb_plane) *nb_plane = r.maxP;
        return r.status;
    };
    return precision == 0 ? run(float(0)) : run(double(0));
}

// Grid.cpp:18-35 (signed tet volumes against the origin, minus sign included)
double orc_grid_volume(const orc_grid *g) {
    double vol = 0;
    for (int t = 0; t < g->num_triangles; ++t) {
        const double *a = g->points + 4 * g->triangles[3 * t];
        const double *b = g->points + 4 * g->triangles[3 * t + 1];
        const double *c = g->points + 4 * g->triangles[3 * t + 2];
        vol += -det3<double>(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]) / 6;
    }
    return vol;
}

void orc_grid_moments(const orc_grid *g, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int t = 0; t < g->num_triangles; ++t) {
        const double *a = g->points + 4 * g->triangles[3 * t];
        const double *b = g->points + 4 * g->triangles[3 * t + 1];
        const double *c = g->points + 4 * g->triangles[3 * t + 2];
        const double v = -det3<double>(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]) / 6;
        const double v4 = v / 4;
        for (int d = 0; d < 3; ++d) out[d] += v4 * (a[d] + b[d] + c[d]);
        out[3] += v;
    }
}

// Grid::valid Grid.cpp:60-184
int orc_grid_valid(const orc_grid *g) {
    const long long nv = (long long)g->size * g->size * g->size;
    for (long long i = 0; i < nv; ++i)
        if (g->in_domain[i] & 0xfc) return 0;
    for (int i = 0; i < 3 * g->num_triangles; ++i)
        if (g->triangles[i] < 0 || g->triangles[i] >= g->num_points) return 0;
    for (int i = 0; i < g->num_tri_list; ++i)
        if (g->tri_list[i] < 0 || g->tri_list[i] >= g->num_triangles) return 0;
    for (long long i = 0; i <= nv; ++i)
        if (g->offsets[i] < 0 || g->offsets[i] > g->num_tri_list) return 0;
    if (orc_grid_volume(g) <= 0) return 0;
    for (int t = 0; t < g->num_triangles; ++t) {
        const double *a = g->points + 4 * g->triangles[3 * t];
        const double *b = g->points + 4 * g->triangles[3 * t + 1];
        const double *c = g->points + 4 * g->triangles[3 * t + 2];
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double area2 = (cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 4;
        if (area2 < 0.01 * 0.01) return 0;
    }
    std::set<std::pair<int, int>> edges;
    for (int t = 0; t < g->num_triangles; ++t)
        for (int j = 0; j < 3; ++j) {
            auto e = std::make_pair(g->triangles[3 * t + j], g->triangles[3 * t + (j + 1) % 3]);
            if (!edges.insert(e).second) return 0;
        }
    for (const auto &e : edges)
        if (!edges.count(std::make_pair(e.second, e.first))) return 0;
    // "no cell has 8 border corners" (Grid.cpp:147-168).  As written, the
    // reference counts the border flags of the corners of voxel 0 (the probe
    // index x+y*g+z*g*g has no p in it) and then rejects every border corner
    // p off the upper faces when that count is 8; restated as such.
    const int s = g->size;
    if (s >= 2) {
        int found = 0;
        for (int dz = 0; dz <= 1; ++dz)
            for (int dy = 0; dy <= 1; ++dy)
                for (int dx = 0; dx <= 1; ++dx)
                    if (g->in_domain[dx + dy * s + (long long)dz * s * s] == 1) ++found;
        if (found == 8)
            for (long long p = 0; p < nv; ++p) {
                if (g->in_domain[p] != 1) continue;
                const int x = int(p % s), y = int((p / s) % s), z = int(p / s / s);
                if (x == s - 1 || y == s - 1 || z == s - 1) continue;
                return 0;
            }
    }
    for (long long o = 0; o < nv; ++o) {
        if (g->offsets[o] > g->offsets[o + 1]) return 0;
        for (int t = g->offsets[o]; t + 1 < g->offsets[o + 1]; ++t)
            if (g->tri_list[t] >= g->tri_list[t + 1]) return 0;
    }
    return 1;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Mesh-to-grid generator (utils/createGrid/main.cpp:150-294).
// ---------------------------------------------------------------------------
namespace {

// Single-precision triangle/box overlap, Akenine-Moller's separating axis
// test as geometry.cpp:433-508 evaluates it: v = float(vertex) - centre,
// edges e0 = v1-v0, e1 = v2-v1, e2 = v0-v2; nine edge axes with the
// vertex pairs of the reference's AXISTEST macros, the three box axes, then
// the triangle plane.  Built with -ffp-contract=off: every product and sum
// rounds on its own, as in the reference's x86-64 build.
bool orc_tri_box(const float ctr[3], float h, const double *T0, const double *T1, const double *T2) {
    float v0[3], v1[3], v2[3];
    for (int d = 0; d < 3; ++d) {
        v0[d] = float(T0[d]) - ctr[d];
        v1[d] = float(T1[d]) - ctr[d];
        v2[d] = float(T2[d]) - ctr[d];
    }
    float e0[3], e1[3], e2[3];
    for (int d = 0; d < 3; ++d) {
        e0[d] = v1[d] - v0[d];
        e1[d] = v2[d] - v1[d];
        e2[d] = v0[d] - v2[d];
    }
    auto sep = [&](float p, float q, float rad) { return std::min(p, q) > rad || std::max(p, q) < -rad; };
    // X tests: p = a*v[Y] - b*v[Z], a = e[Z], b = e[Y]
    auto tx = [&](const float *e, const float *va, const float *vb) {
        const float a = e[2], b = e[1];
        return sep(a * va[1] - b * va[2], a * vb[1] - b * vb[2], std::fabs(a) * h + std::fabs(b) * h);
    };
    // Y tests: p = -a*v[X] + b*v[Z], a = e[Z], b = e[X]
    auto ty = [&](const float *e, const float *va, const float *vb) {
        const float a = e[2], b = e[0];
        return sep(-a * va[0] + b * va[2], -a * vb[0] + b * vb[2], std::fabs(a) * h + std::fabs(b) * h);
    };
    // Z tests: p = a*v[X] - b*v[Y], a = e[Y], b = e[X]
    auto tz = [&](const float *e, const float *va, const float *vb) {
        const float a = e[1], b = e[0];
        return sep(a * va[0] - b * va[1], a * vb[0] - b * vb[1], std::fabs(a) * h + std::fabs(b) * h);
    };
    if (tx(e0, v0, v2) || ty(e0, v0, v2) || tz(e0, v1, v2)) return false;
    if (tx(e1, v0, v2) || ty(e1, v0, v2) || tz(e1, v0, v1)) return false;
    if (tx(e2, v0, v1) || ty(e2, v0, v1) || tz(e2, v1, v2)) return false;
    for (int d = 0; d < 3; ++d) {
        const float mn = std::min(v0[d], std::min(v1[d], v2[d]));
        const float mx = std::max(v0[d], std::max(v1[d], v2[d]));
        if (mn > h || mx < -h) return false;
    }
    const float nx = e0[1] * e1[2] - e0[2] * e1[1];
    const float ny = e0[2] * e1[0] - e0[0] * e1[2];
    const float nz = e0[0] * e1[1] - e0[1] * e1[0];
    const float dpl = -(nx * v0[0] + ny * v0[1] + nz * v0[2]);
    const float ax = nx > 0.0f ? h : -h, ay = ny > 0.0f ? h : -h, az = nz > 0.0f ? h : -h;
    if ((nx * -ax + ny * -ay + nz * -az) + dpl > 0.0f) return false;
    return (nx * ax + ny * ay + nz * az) + dpl >= 0.0f;
}

// Distance from p to triangle abc: the plane distance when the projection
// falls inside the triangle (barycentric signs), else the nearest of the
// three edge segments.
double orc_point_tri_dist(const double *p, const double *a, const double *b, const double *c) {
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double w[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const double wu = w[0] * u[0] + w[1] * u[1] + w[2] * u[2], wv = w[0] * v[0] + w[1] * v[1] + w[2] * v[2];
    const double den = uu * vv - uv * uv;
    const double s = (vv * wu - uv * wv) / den, t = (uu * wv - uv * wu) / den;
    if (s >= 0 && t >= 0 && s + t <= 1) {
        double q[3];
        for (int d = 0; d < 3; ++d) q[d] = a[d] + s * u[d] + t * v[d] - p[d];
        return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }
    auto seg = [&](const double *x, const double *y) {
        const double e[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        const double f[3] = {p[0] - x[0], p[1] - x[1], p[2] - x[2]};
        double l = (e[0] * f[0] + e[1] * f[1] + e[2] * f[2]) / (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        l = std::min(1.0, std::max(0.0, l));
        const double q[3] = {f[0] - l * e[0], f[1] - l * e[1], f[2] - l * e[2]};
        return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    };
    return std::min(seg(a, b), std::min(seg(b, c), seg(c, a)));
}

// Generalised winding number of a closed triangle mesh at p: the sum of the
// signed solid angles of its triangles (Van Oosterom & Strackee, IEEE TBME
// 30(2), 1983) over 4 pi; +-1 inside, 0 outside.
double orc_winding(const double *p, const double *pts, const int *tris, int nt) {
    double sum = 0;
    for (int t = 0; t < nt; ++t) {
        double r[3][3], l[3];
        for (int q = 0; q < 3; ++q) {
            for (int d = 0; d < 3; ++d) r[q][d] = pts[3 * size_t(tris[3 * t + q]) + d] - p[d];
            l[q] = std::sqrt(r[q][0] * r[q][0] + r[q][1] * r[q][1] + r[q][2] * r[q][2]);
        }
        const double det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
                           r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
                           r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
        auto dt = [&](int i, int j) { return r[i][0] * r[j][0] + r[i][1] * r[j][1] + r[i][2] * r[j][2]; };
        const double den = l[0] * l[1] * l[2] + dt(0, 1) * l[2] + dt(0, 2) * l[1] + dt(1, 2) * l[0];
        sum += 2.0 * std::atan2(det, den);
    }
    return sum / (4.0 * M_PI);
}

}  // namespace

extern "C" {

int orc_tri_box_overlap(const float centre[3], float half, const double *a, const double *b, const double *c) {
    return orc_tri_box(centre, half, a, b, c) ? 1 : 0;
}

double orc_point_triangle_distance(const double *p, const double *a, const double *b, const double *c) {
    return orc_point_tri_dist(p, a, b, c);
}

double orc_winding_number(const double *p, const double *pts, const int *tris, int nt) {
    return orc_winding(p, pts, tris, nt);
}

int orc_mesh_points(const double *points, int np, const int *tris, int nt, int normalize, double *out_xyz,
                    int *out_tris) {
    std::vector<int> remap(size_t(np), -1);
    for (size_t i = 0; i < 3 * size_t(nt); ++i) {
        if (tris[i] < 0 || tris[i] >= np) return -1;
        remap[size_t(tris[i])] = 0;
    }
    int m = 0;
    for (int i = 0; i < np; ++i)
        if (remap[size_t(i)] == 0) {
            remap[size_t(i)] = m;
            for (int d = 0; d < 3; ++d) out_xyz[3 * size_t(m) + d] = points[3 * size_t(i) + d];
            ++m;
        }
    for (size_t i = 0; i < 3 * size_t(nt); ++i) out_tris[i] = remap[size_t(tris[i])];
    if (normalize) {
        // main.cpp:168-175: bounding box of the kept points, scale 990.222 /
        // its largest side, then shift by 4.998
        double lo[3] = {out_xyz[0], out_xyz[1], out_xyz[2]}, hi[3] = {lo[0], lo[1], lo[2]};
        for (int i = 0; i < m; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], out_xyz[3 * size_t(i) + d]);
                hi[d] = std::max(hi[d], out_xyz[3 * size_t(i) + d]);
            }
        const double maxdim = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        if (!(maxdim > 0)) return -1;
        for (int i = 0; i < m; ++i)
            for (int d = 0; d < 3; ++d)
                out_xyz[3 * size_t(i) + d] = (990.222 / maxdim) * (out_xyz[3 * size_t(i) + d] - lo[d]) + 4.998;
    }
    return m;
}

long long orc_mesh_lists(const double *xyz, const int *tris, int nt, int g, int *out_offsets, int *out_list,
                         long long cap) {
    // main.cpp:178-219: per triangle, the voxels of the scaled bounding box
    // range that pass the overlap test; lists in increasing triangle order
    const long long nv = (long long)g * g * g;
    std::vector<std::vector<int>> lists(static_cast<size_t>(nv));
    const float vsf = float(1000. / double(g));
    const float half = 0.5f * vsf;
    for (int t = 0; t < nt; ++t) {
        const double *P[3] = {xyz + 3 * size_t(tris[3 * t]), xyz + 3 * size_t(tris[3 * t + 1]),
                              xyz + 3 * size_t(tris[3 * t + 2])};
        long long lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            const double mn = std::min(P[0][d], std::min(P[1][d], P[2][d]));
            const double mx = std::max(P[0][d], std::max(P[1][d], P[2][d]));
            lo[d] = (long long)std::floor(double(g) * mn / 1000.);
            hi[d] = (long long)std::ceil(double(g) * mx / 1000.);
        }
        for (long long i = lo[0]; i <= hi[0]; ++i) {
            if (i >= g) continue;
            for (long long j = lo[1]; j <= hi[1]; ++j) {
                if (j >= g) continue;
                for (long long k = lo[2]; k <= hi[2]; ++k) {
                    if (k >= g) continue;
                    const float c[3] = {float(double(vsf) * (0.5 + double(i))), float(double(vsf) * (0.5 + double(j))),
                                        float(double(vsf) * (0.5 + double(k)))};
                    if (orc_tri_box(c, half, P[0], P[1], P[2])) lists[size_t(k * g * g + j * g + i)].push_back(t);
                }
            }
        }
    }
    long long total = 0;
    for (const auto &l : lists) total += (long long)l.size();
    if (total > cap) return total;
    long long o = 0;
    for (long long v = 0; v < nv; ++v) {
        out_offsets[v] = int(o);
        for (int t : lists[size_t(v)]) out_list[o++] = t;
    }
    out_offsets[nv] = int(o);
    return total;
}

void orc_mesh_corner_flags(const double *xyz, const int *tris, int nt, int g, double eps, signed char *out_flags) {
    // the flag rule of the library's mesh grids: 1 within eps of the surface,
    // else 2 inside / 0 outside, by brute force over every triangle
    const long long nv = (long long)g * g * g;
    const double vs = 1000.0 / double(g);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long id = 0; id < nv; ++id) {
        const double p[3] = {double(id % g) * vs, double((id / g) % g) * vs, double(id / ((long long)g * g)) * vs};
        double dmin = 1e300;
        for (int t = 0; t < nt && dmin > eps; ++t)
            dmin = std::min(dmin, orc_point_tri_dist(p, xyz + 3 * size_t(tris[3 * t]), xyz + 3 * size_t(tris[3 * t + 1]),
                                                    xyz + 3 * size_t(tris[3 * t + 2])));
        if (dmin <= eps) {
            out_flags[id] = 1;
            continue;
        }
        out_flags[id] = std::fabs(orc_winding(p, xyz, tris, nt)) > 0.5 ? 2 : 0;
    }
}

}  // extern "C"
