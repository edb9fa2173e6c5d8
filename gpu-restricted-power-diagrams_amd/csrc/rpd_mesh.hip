// rpd_mesh.hip -- domain grid of a closed triangle mesh, the mesh path of
// utils/createGrid/main.cpp:150-294, built on the GPU:
//
//  1. points no triangle uses are dropped (main.cpp:164) and the others are
//     scaled into [4.998, 995.22]^3 (main.cpp:168-175);
//  2. per-voxel triangle lists: for each triangle, every voxel of its bounding
//     range (main.cpp:190-210) that passes the single-precision triangle/box
//     overlap test (geometry.cpp:433-508) -- one thread per triangle counts,
//     a scan places, a second pass emits (voxel, triangle) keys, a radix sort
//     orders them, so each list is ascending as the reference's;
//  3. corner flags of each voxel's first corner: 1 within 1e-3 of the
//     surface (the eps of main.cpp:276), else 2 inside / 0 outside.  The
//     reference decides inside/outside from a constrained tetrahedralization
//     (main.cpp:240-292, geogram/tetgen); here the class comes from a voxel
//     centre incident to the corner, whose class the grid already holds
//     (ray parity, rpd_derive_centre_classes), flipped by the parity of the
//     surface crossings of the segment corner -> centre, which lies inside
//     that voxel's box so only its triangle list can cross it.  Where every
//     incident centre is unknown or every segment passes within rounding of a
//     triangle edge, the flag is 1, which the restriction treats as "not a
//     usable apex" (always safe).  One thread per corner.

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rpd_internal.h"

namespace {

constexpr double kCornerEps = 1e-3;
__device__ constexpr double kCentreFrac[3] = {0.530901699437495, 0.520710678118655, 0.536602540378444};

// Triangle / axis-aligned cube overlap by the separating axis theorem in
// single precision (T. Akenine-Moller, "Fast 3D triangle-box overlap
// testing", J. Graphics Tools 6(1), 2001).  The operands and the order of
// every product and sum are those of geometry.cpp:433-508 (for each edge
// axis the same two vertices are projected), with no fused multiply-adds,
// so voxels that touch a triangle within rounding are decided as the
// reference decides them.  c: box centre, h: half size, a/b/c: vertices.
__device__ bool tri_box_overlap(const float c[3], float h, const float A[3], const float B[3], const float C[3]) {
#pragma clang fp contract(off)
    float v[3][3], e[3][3];
    for (int d = 0; d < 3; ++d) {
        v[0][d] = A[d] - c[d];
        v[1][d] = B[d] - c[d];
        v[2][d] = C[d] - c[d];
    }
    for (int d = 0; d < 3; ++d) {
        e[0][d] = v[1][d] - v[0][d];
        e[1][d] = v[2][d] - v[1][d];
        e[2][d] = v[0][d] - v[2][d];
    }
    // the two vertices each edge axis projects: (x, y) axes use (v0, v2) for
    // edges 0 and 1 and (v0, v1) for edge 2; the z axis uses (v1, v2), (v0,
    // v1), (v1, v2)
    const int px[3][2] = {{0, 2}, {0, 2}, {0, 1}};
    const int pz[3][2] = {{1, 2}, {0, 1}, {1, 2}};
    for (int k = 0; k < 3; ++k) {
        const float ex = e[k][0], ey = e[k][1], ez = e[k][2];
        const float fx = fabsf(ex), fy = fabsf(ey), fz = fabsf(ez);
        {  // axis (1,0,0) x edge
            const float p0 = ez * v[px[k][0]][1] - ey * v[px[k][0]][2];
            const float p1 = ez * v[px[k][1]][1] - ey * v[px[k][1]][2];
            const float rad = fz * h + fy * h;
            if (fminf(p0, p1) > rad || fmaxf(p0, p1) < -rad) return false;
        }
        {  // axis (0,1,0) x edge
            const float p0 = -ez * v[px[k][0]][0] + ex * v[px[k][0]][2];
            const float p1 = -ez * v[px[k][1]][0] + ex * v[px[k][1]][2];
            const float rad = fz * h + fx * h;
            if (fminf(p0, p1) > rad || fmaxf(p0, p1) < -rad) return false;
        }
        {  // axis (0,0,1) x edge
            const float p0 = ey * v[pz[k][0]][0] - ex * v[pz[k][0]][1];
            const float p1 = ey * v[pz[k][1]][0] - ex * v[pz[k][1]][1];
            const float rad = fy * h + fx * h;
            if (fminf(p0, p1) > rad || fmaxf(p0, p1) < -rad) return false;
        }
    }
    for (int d = 0; d < 3; ++d) {
        const float mn = fminf(fminf(v[0][d], v[1][d]), v[2][d]);
        const float mx = fmaxf(fmaxf(v[0][d], v[1][d]), v[2][d]);
        if (mn > h || mx < -h) return false;
    }
    // plane of the triangle against the box
    const float n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                        e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    const float dd = -(n[0] * v[0][0] + n[1] * v[0][1] + n[2] * v[0][2]);
    float lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = n[d] > 0.0f ? -h : h;
        hi[d] = n[d] > 0.0f ? h : -h;
    }
    if ((n[0] * lo[0] + n[1] * lo[1] + n[2] * lo[2]) + dd > 0.0f) return false;
    return (n[0] * hi[0] + n[1] * hi[1] + n[2] * hi[2]) + dd >= 0.0f;
}

// Voxel range of a triangle's bounding box (main.cpp:180-200): floor of the
// scaled minimum to ceil of the scaled maximum, inclusive, capped at g-1.
__device__ void tri_range(const double *P, int g, int lo[3], int hi[3]) {
    for (int d = 0; d < 3; ++d) {
        const double mn = fmin(fmin(P[d], P[3 + d]), P[6 + d]);
        const double mx = fmax(fmax(P[d], P[3 + d]), P[6 + d]);
        lo[d] = int(floor(double(g) * mn / 1000.0));
        hi[d] = min(int(ceil(double(g) * mx / 1000.0)), g - 1);
    }
}

// EMIT = false: count[t] = number of voxels the triangle overlaps.
// EMIT = true: write (voxel << 32 | t) at start[t] + k and count each voxel's
// entries in vcnt.
template <bool EMIT>
__global__ void __launch_bounds__(256) mesh_sat_kernel(const double *__restrict__ pts, const int *__restrict__ tris,
                                                       int nt, int g, int *__restrict__ count,
                                                       const long long *__restrict__ start,
                                                       uint64_t *__restrict__ keys,
                                                       int *__restrict__ vcnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double P[9];
    float F[3][3];
    for (int q = 0; q < 3; ++q) {
        const int id = tris[3 * t + q];
        for (int d = 0; d < 3; ++d) {
            P[3 * q + d] = pts[3 * size_t(id) + d];
            F[q][d] = float(P[3 * q + d]);
        }
    }
    int lo[3], hi[3];
    tri_range(P, g, lo, hi);
    const float vsf = float(1000.0 / double(g));
    const float h = 0.5f * vsf;
    int k = 0;
    const size_t base = EMIT ? size_t(start[t]) : 0;
    // slots the count pass reserved (the same test decides both passes;
    // the bound only keeps every write inside the triangle's range)
    const long long room = EMIT ? start[t + 1] - start[t] : 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const float c[3] = {float(double(vsf) * (0.5 + double(x))), float(double(vsf) * (0.5 + double(y))),
                                    float(double(vsf) * (0.5 + double(z)))};
                if (!tri_box_overlap(c, h, F[0], F[1], F[2])) continue;
                if (EMIT) {
                    if (k >= room) return;
                    const uint32_t vox = uint32_t(x + y * g + z * g * g);
                    keys[base + k] = (uint64_t(vox) << 32) | uint32_t(t);
                    atomicAdd(vcnt + vox, 1);
                }
                ++k;
            }
    if (!EMIT) count[t] = k;
}

__global__ void low_words_kernel(const uint64_t *__restrict__ keys, int m, int *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = int(uint32_t(keys[i]));
}

__device__ inline void sub3(const double *a, const double *b, double *r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void cross3(const double *a, const double *b, double *r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// Squared distance from p to triangle abc: the closest point found by the
// Voronoi-region walk of Ericson, Real-Time Collision Detection, 5.1.5.
__device__ double point_tri_dist2(const double *p, const double *a, const double *b, const double *c) {
    double ab[3], ac[3], ap[3], q[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    auto at = [&](double s, double t) {
        for (int d = 0; d < 3; ++d) q[d] = a[d] + s * ab[d] + t * ac[d];
        double r[3];
        sub3(p, q, r);
        return dot3(r, r);
    };
    if (d1 <= 0 && d2 <= 0) return at(0, 0);
    double bp[3];
    sub3(p, b, bp);
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) return at(1, 0);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return at(d1 / (d1 - d3), 0);
    double cp[3];
    sub3(p, c, cp);
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0 && d5 <= d6) return at(0, 1);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return at(0, d2 / (d2 - d6));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return at(1 - w, w);
    }
    const double den = 1.0 / (va + vb + vc);
    return at(vb * den, vc * den);
}

__device__ inline double orient(const double *a, const double *b, const double *c, const double *d) {
    double u[3], v[3], w[3], x[3];
    sub3(b, a, u);
    sub3(c, a, v);
    sub3(d, a, w);
    cross3(v, w, x);
    return dot3(u, x);
}

// Crossing of segment pq with triangle abc, p known to lie farther than
// kCornerEps from every triangle: 0 none, 1 one crossing, 2 within rounding
// of an edge or of q (undecided).
__device__ int segment_crossing(const double *p, const double *q, const double *a, const double *b,
                                const double *c) {
    double e[3];
    double L = 0.0;
    sub3(q, p, e);
    L = fmax(fmax(fabs(e[0]), fabs(e[1])), fabs(e[2]));
    double E = 0.0;
    for (int d = 0; d < 3; ++d) E = fmax(E, fmax(fabs(b[d] - a[d]), fmax(fabs(c[d] - b[d]), fabs(a[d] - c[d]))));
    const double s = L + E;
    const double tol = 1e-12 * s * s * s;
    const double s1 = orient(a, b, c, p), s2 = orient(a, b, c, q);
    if (fabs(s2) <= tol) return 2;
    // |s1| small: the plane meets the segment next to p, which is far from
    // the triangle, so there is no crossing
    if (fabs(s1) <= tol || (s1 > 0) == (s2 > 0)) return 0;
    const double u = orient(p, q, a, b), v = orient(p, q, b, c), w = orient(p, q, c, a);
    if (fabs(u) <= tol || fabs(v) <= tol || fabs(w) <= tol) return 2;
    return ((u > 0) == (v > 0) && (v > 0) == (w > 0)) ? 1 : 0;
}

__global__ void __launch_bounds__(256) corner_flags_kernel(const double *__restrict__ pts, const int *__restrict__ tris,
                                                           const int *__restrict__ tri_list,
                                                           const int *__restrict__ offs,
                                                           const int8_t *__restrict__ centre, int g,
                                                           int8_t *__restrict__ flags) {
    const size_t nv = size_t(g) * g * g;
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nv) return;
    const int i = int(id % g), j = int((id / g) % g), k = int(id / (size_t(g) * g));
    const double vs = 1000.0 / double(g);
    const double p[3] = {i * vs, j * vs, k * vs};
    auto vertex = [&](int t, int q, double *r) {
        const int pid = tris[3 * t + q];
        for (int d = 0; d < 3; ++d) r[d] = pts[3 * size_t(pid) + d];
    };
    // the voxels incident to the corner, the corner's own voxel first
    for (int n = 0; n < 8; ++n) {
        const int x = i - (n & 1), y = j - ((n >> 1) & 1), z = k - (n >> 2);
        if (x < 0 || y < 0 || z < 0) continue;
        const size_t v = size_t(x) + size_t(y) * g + size_t(z) * g * g;
        for (int q = offs[v]; q < offs[v + 1]; ++q) {
            double a[3], b[3], c[3];
            const int t = tri_list[q];
            vertex(t, 0, a);
            vertex(t, 1, b);
            vertex(t, 2, c);
            if (point_tri_dist2(p, a, b, c) <= kCornerEps * kCornerEps) {
                flags[id] = 1;
                return;
            }
        }
    }
    for (int n = 0; n < 8; ++n) {
        const int x = i - (n & 1), y = j - ((n >> 1) & 1), z = k - (n >> 2);
        if (x < 0 || y < 0 || z < 0) continue;
        const size_t v = size_t(x) + size_t(y) * g + size_t(z) * g * g;
        const int cls = centre[v];
        if (cls == 1) continue;
        const double qc[3] = {(x + kCentreFrac[0]) * vs, (y + kCentreFrac[1]) * vs, (z + kCentreFrac[2]) * vs};
        int parity = 0;
        bool undecided = false;
        for (int q = offs[v]; q < offs[v + 1] && !undecided; ++q) {
            double a[3], b[3], c[3];
            const int t = tri_list[q];
            vertex(t, 0, a);
            vertex(t, 1, b);
            vertex(t, 2, c);
            const int r = segment_crossing(p, qc, a, b, c);
            if (r == 2)
                undecided = true;
            else
                parity ^= r;
        }
        if (undecided) continue;
        flags[id] = int8_t(parity ? 2 - cls : cls);
        return;
    }
    flags[id] = 1;
}

inline int hip_ok(hipError_t e) { return e == hipSuccess ? RPD_OK : RPD_E_DEVICE; }

}  // namespace

extern "C" int rpd_grid_from_mesh(int g, const double *points, int num_points, const int32_t *triangles,
                                  int num_triangles, int normalize, rpd_grid **out) {
    if (g <= 0 || g > 1024 || !points || num_points <= 0 || !triangles || num_triangles <= 0 || !out)
        return RPD_E_ARG;
    for (size_t i = 0; i < 3 * size_t(num_triangles); ++i)
        if (triangles[i] < 0 || triangles[i] >= num_points) return RPD_E_ARG;
    for (size_t i = 0; i < 3 * size_t(num_points); ++i)
        if (!std::isfinite(points[i])) return RPD_E_ARG;
    // drop the points no triangle uses, keeping the order of the others
    std::vector<int32_t> remap(size_t(num_points), -1);
    for (size_t i = 0; i < 3 * size_t(num_triangles); ++i) remap[size_t(triangles[i])] = 0;
    int np = 0;
    for (auto &r : remap)
        if (r == 0) r = np++;
    std::vector<double> xyz(3 * size_t(np));
    for (int i = 0; i < num_points; ++i)
        if (remap[size_t(i)] >= 0)
            for (int d = 0; d < 3; ++d) xyz[3 * size_t(remap[size_t(i)]) + d] = points[3 * size_t(i) + d];
    std::vector<int32_t> tri(3 * size_t(num_triangles));
    for (size_t i = 0; i < tri.size(); ++i) tri[i] = remap[size_t(triangles[i])];
    if (normalize) {
        double mn[3], mx[3];
        for (int d = 0; d < 3; ++d) mn[d] = mx[d] = xyz[size_t(d)];
        for (int i = 0; i < np; ++i)
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], xyz[3 * size_t(i) + d]);
                mx[d] = std::max(mx[d], xyz[3 * size_t(i) + d]);
            }
        double maxdim = 0;
        for (int d = 0; d < 3; ++d) maxdim = std::max(maxdim, mx[d] - mn[d]);
        if (!(maxdim > 0) || !std::isfinite(maxdim)) return RPD_E_ARG;
        const double s = 990.222 / maxdim;
        for (int i = 0; i < np; ++i)
            for (int d = 0; d < 3; ++d) {
                double &c = xyz[3 * size_t(i) + d];
                c = s * (c - mn[d]) + 4.998;
            }
    } else {
        for (double c : xyz)
            if (!(c >= 0.0 && c <= 1000.0)) return RPD_E_ARG;
    }

    const size_t nv = size_t(g) * g * g;
    const int nt = num_triangles;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return RPD_E_DEVICE;
    DevBuf d_pts, d_tri, d_cnt, d_start, d_keys, d_keys2, d_vcnt, d_list, d_offs, d_tmp, d_centre, d_flags;
    auto *gr = new rpd_grid();
    int rc = RPD_OK;
    auto run = [&]() -> int {
        int r;
        if ((r = d_pts.ensure(xyz.size() * 8)) || (r = d_tri.ensure(tri.size() * 4)) ||
            (r = d_cnt.ensure((size_t(nt) + 1) * 4)) || (r = d_start.ensure((size_t(nt) + 1) * 8)) ||
            (r = d_vcnt.ensure((nv + 1) * 4)))
            return r;
        int *cnt = d_cnt.as<int>();
        long long *start = d_start.as<long long>();
        if ((r = hip_ok(hipMemcpyAsync(d_pts.ptr, xyz.data(), xyz.size() * 8, hipMemcpyHostToDevice, st))) ||
            (r = hip_ok(hipMemcpyAsync(d_tri.ptr, tri.data(), tri.size() * 4, hipMemcpyHostToDevice, st))) ||
            (r = hip_ok(hipMemsetAsync(cnt, 0, (size_t(nt) + 1) * 4, st))) ||
            (r = hip_ok(hipMemsetAsync(d_vcnt.ptr, 0, (nv + 1) * 4, st))))
            return r;
        const int tb = 256;
        mesh_sat_kernel<false><<<(nt + tb - 1) / tb, tb, 0, st>>>(d_pts.as<double>(), d_tri.as<int>(), nt, g, cnt,
                                                                   nullptr, nullptr, nullptr);
        size_t scan_bytes = 0;
        // 64-bit starts: the entry total is checked against the int32
        // offsets of the grid format below, not left to wrap
        if ((r = hip_ok(rocprim::exclusive_scan(nullptr, scan_bytes, cnt, start, 0LL, size_t(nt) + 1,
                                                rocprim::plus<long long>(), st))) ||
            (r = d_tmp.ensure(scan_bytes)) ||
            (r = hip_ok(rocprim::exclusive_scan(d_tmp.ptr, scan_bytes, cnt, start, 0LL, size_t(nt) + 1,
                                                rocprim::plus<long long>(), st))))
            return r;
        long long total = 0;
        if ((r = hip_ok(hipMemcpyAsync(&total, start + nt, 8, hipMemcpyDeviceToHost, st))) ||
            (r = hip_ok(hipStreamSynchronize(st))))
            return r;
        if (total < 0 || total > 0x7fffffffLL) return RPD_E_LIMIT;
        const size_t m = size_t(total);
        unsigned vbits = 1;
        while ((size_t(1) << vbits) < nv) ++vbits;
        // every temporary is sized before more work is queued: nothing is
        // reallocated under a kernel in flight
        size_t sort_bytes = 0, oscan = 0;
        if ((r = hip_ok(rocprim::radix_sort_keys(nullptr, sort_bytes, d_keys.as<uint64_t>(), d_keys2.as<uint64_t>(),
                                                 std::max<size_t>(m, 1), 0u, 32u + vbits, st))) ||
            (r = hip_ok(rocprim::exclusive_scan(nullptr, oscan, d_vcnt.as<int>(), d_offs.as<int>(), 0, nv + 1,
                                                rocprim::plus<int>(), st))) ||
            (r = d_tmp.ensure(std::max(std::max(sort_bytes, oscan), scan_bytes))) ||
            (r = d_keys.ensure(std::max<size_t>(m, 1) * 8)) || (r = d_keys2.ensure(std::max<size_t>(m, 1) * 8)) ||
            (r = d_list.ensure(std::max<size_t>(m, 1) * 4)) || (r = d_offs.ensure((nv + 1) * 4)) ||
            (r = d_centre.ensure(nv)) || (r = d_flags.ensure(nv)))
            return r;
        mesh_sat_kernel<true><<<(nt + tb - 1) / tb, tb, 0, st>>>(d_pts.as<double>(), d_tri.as<int>(), nt, g,
                                                                  nullptr, start, d_keys.as<uint64_t>(),
                                                                  d_vcnt.as<int>());
        if (m > 0) {
            if ((r = hip_ok(rocprim::radix_sort_keys(d_tmp.ptr, sort_bytes, d_keys.as<uint64_t>(),
                                                     d_keys2.as<uint64_t>(), m, 0u, 32u + vbits, st))))
                return r;
            low_words_kernel<<<int((m + tb - 1) / tb), tb, 0, st>>>(d_keys2.as<uint64_t>(), int(m),
                                                                    d_list.as<int>());
        }
        // offsets = exclusive scan of the per-voxel counts (vcnt[nv] = 0)
        int *offs = d_offs.as<int>();
        if ((r = hip_ok(rocprim::exclusive_scan(d_tmp.ptr, oscan, d_vcnt.as<int>(), offs, 0, nv + 1,
                                                rocprim::plus<int>(), st))))
            return r;
        gr->size = g;
        gr->tri_list.resize(m);
        gr->offsets.resize(nv + 1);
        if ((m && (r = hip_ok(hipMemcpyAsync(gr->tri_list.data(), d_list.ptr, m * 4, hipMemcpyDeviceToHost, st)))) ||
            (r = hip_ok(hipMemcpyAsync(gr->offsets.data(), offs, (nv + 1) * 4, hipMemcpyDeviceToHost, st))) ||
            (r = hip_ok(hipStreamSynchronize(st))))
            return r;
        // both passes must have found the same overlaps
        if (size_t(gr->offsets[nv]) != m) return RPD_E_DEVICE;
        gr->points.resize(4 * size_t(np));
        for (int i = 0; i < np; ++i) {
            for (int d = 0; d < 3; ++d) gr->points[4 * size_t(i) + d] = xyz[3 * size_t(i) + d];
            gr->points[4 * size_t(i) + 3] = 0.0;
        }
        gr->triangles = tri;
        // the centre classes take the voxel count from the flag array
        gr->in_domain.assign(nv, 1);
        rpd_derive_centre_classes(gr);
        if ((r = hip_ok(hipMemcpyAsync(d_centre.ptr, gr->centre.data(), nv, hipMemcpyHostToDevice, st)))) return r;
        corner_flags_kernel<<<int((nv + tb - 1) / tb), tb, 0, st>>>(
            d_pts.as<double>(), d_tri.as<int>(), d_list.as<int>(), offs, d_centre.as<int8_t>(), g,
            d_flags.as<int8_t>());
        if ((r = hip_ok(hipGetLastError())) ||
            (r = hip_ok(hipMemcpyAsync(gr->in_domain.data(), d_flags.ptr, nv, hipMemcpyDeviceToHost, st))) ||
            (r = hip_ok(hipStreamSynchronize(st))))
            return r;
        return RPD_OK;
    };
    rc = run();
    (void)hipStreamDestroy(st);
    if (rc != RPD_OK) {
        delete gr;
        return rc;
    }
    // orientation normalisation of Grid::load (Grid.cpp:270-275)
    rpd_grid_desc dsc;
    rpd_grid_view(gr, &dsc);
    if (rpd_grid_volume(&dsc) < 0)
        for (size_t t = 0; t + 2 < gr->triangles.size(); t += 3) std::swap(gr->triangles[t], gr->triangles[t + 1]);
    *out = gr;
    return RPD_OK;
}
