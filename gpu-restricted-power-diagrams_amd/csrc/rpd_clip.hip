// rpd_clip.hip -- per-seed power-cell clipping, restriction to the domain
// grid and fused volume / first-moment integration on gfx950.
//
// Replaces compute_voro_cell (ConvexCell.cl:812-1152) and the ConvexCell
// routines it drives (CCInit :278-308, CCNewPlane :314-333, CCClipByPlane
// :517-563, CCComputeBoundary :463-515, CCVertexIsInHalfSpace :400-461,
// CCIsSecurityRadiusReached :364-382, CCClipTetFromPointsInfinite :565-575,
// CCExportBaryVolume :577-619).
//
// Layout: ONE CELL PER WAVEFRONT.  Vertex j of the cell lives in lane j & 63
// (register slot j >> 6): its three plane ids and its homogeneous
// coordinates (X, Y, Z, W), i.e. the cofactors of CCComputeVertexCoordinates.
// Because det4x4(pi1, pi2, pi3, e) == e . (X, Y, Z, W), the plane-side test
// of CCVertexIsInHalfSpace is one 4-term dot product per lane.  Plane
// equations live in LDS.  A clip removes the vertices with e.v > 0, finds the
// cavity boundary as the edges of removed vertices whose reverse edge is not
// also removed (checked to be one simple cycle, else inconsistent_boundary),
// and places one new vertex (plane, a, b) per boundary edge a->b into free
// lanes -- the same vertex set CCComputeBoundary + CCNewVertex build.
//
// Everything is in the seed-local frame (origin at the seed, FP64 inputs):
// planes n.y + d with d = (|b|^2 + w_B - w_s) / (2|b|), b = B - s.
//
// Forms: FP64 (all double); mixed: the plane-side tests and the vertex
// registers are FP32 and each test also evaluates a certified bound on its
// FP32 error against the FP64 planes, 2^-18 * sum of |terms| of the det4x4
// expansion; a test inside the bound or an ill-conditioned new vertex ends the
// cell with needs_exact_predicates and the host re-runs it in FP64.  The
// planes are stored in FP64, and the cell geometry that produces the results
// (bbox, cone planes, volume and moments) is evaluated in FP64 from them, so
// an unflagged mixed cell has the FP64 combinatorics and FP64 geometry.  FP32
// form: everything in FP32, no filter.  Moments are accumulated in FP64.

#include <hip/hip_runtime.h>

#include "rpd_internal.h"
#include "rpd_wave.h"

namespace rpd {

template <typename R>
struct Pl {
    R x, y, z, w;
};

enum ClipResult : int { kNoClip = -1, kClipped = -2 };

struct ClipArgs {
    const double4 *seeds;
    const int *perm;
    const int *ids;
    const int *knn;
    int64_t n_ids;
    int K, P, V, T;
    int apex_mode;
    int g;
    double vs;
    const int8_t *flags;
    const int8_t *centre;
    const double4 *gpts;
    const int *tris;
    const int *tlist;
    const int *toffs;
    double4 *out;
    int *status;
    int *kused;
    int lds_stride;
    int tl_cap;
    int security_mode;    // 0 exact power stop, 1 reference stop
    double wmin;          // minimum weight over non-border seeds
    const uint8_t *dist;  // Chebyshev voxel distance to the nearest domain voxel (nullable)
};

__device__ __forceinline__ uint32_t pack3(uint32_t a, uint32_t b, uint32_t c) { return a | (b << 8) | (c << 16); }
__device__ __forceinline__ uint32_t t0(uint32_t t) { return t & 0xff; }
__device__ __forceinline__ uint32_t t1(uint32_t t) { return (t >> 8) & 0xff; }
__device__ __forceinline__ uint32_t t2(uint32_t t) { return (t >> 16) & 0xff; }

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R>
__device__ __forceinline__ R rabs(R v) {
    return v < R(0) ? -v : v;
}

template <typename R>
__device__ __forceinline__ Pl<R> rlane_pl(const Pl<R> &p, int l) {
    return Pl<R>{rlane(p.x, l), rlane(p.y, l), rlane(p.z, l), rlane(p.w, l)};
}

template <typename T, typename U>
__device__ __forceinline__ Pl<T> cvt(const Pl<U> &p) {
    return Pl<T>{T(p.x), T(p.y), T(p.z), T(p.w)};
}

// Homogeneous vertex of planes (p1, p2, p3): the cofactors of
// CCComputeVertexCoordinates (ConvexCell.cl:344-362), plus (FILTER) the sums
// of absolute values of their terms.
template <typename R, bool FILTER>
__device__ __forceinline__ void vertex_from_planes(const Pl<R> &p1, const Pl<R> &p2, const Pl<R> &p3, R &X, R &Y,
                                                   R &Z, R &W, float &MX, float &MY, float &MZ, float &MW) {
    const R m12 = p2.x * p3.y - p2.y * p3.x;
    const R m13 = p2.x * p3.z - p2.z * p3.x;
    const R m14 = p2.x * p3.w - p2.w * p3.x;
    const R m23 = p2.y * p3.z - p2.z * p3.y;
    const R m24 = p2.y * p3.w - p2.w * p3.y;
    const R m34 = p2.z * p3.w - p2.w * p3.z;
    X = -p1.w * m23 - p1.y * m34 + p1.z * m24;
    Y = p1.x * m34 + p1.w * m13 - p1.z * m14;
    Z = -p1.x * m24 + p1.y * m14 - p1.w * m12;
    W = p1.x * m23 - p1.y * m13 + p1.z * m12;
    if (FILTER) {
        const float a12 = fabsf(float(p2.x * p3.y)) + fabsf(float(p2.y * p3.x));
        const float a13 = fabsf(float(p2.x * p3.z)) + fabsf(float(p2.z * p3.x));
        const float a14 = fabsf(float(p2.x * p3.w)) + fabsf(float(p2.w * p3.x));
        const float a23 = fabsf(float(p2.y * p3.z)) + fabsf(float(p2.z * p3.y));
        const float a24 = fabsf(float(p2.y * p3.w)) + fabsf(float(p2.w * p3.y));
        const float a34 = fabsf(float(p2.z * p3.w)) + fabsf(float(p2.w * p3.z));
        const float ax = fabsf(float(p1.x)), ay = fabsf(float(p1.y)), az = fabsf(float(p1.z)),
                    aw = fabsf(float(p1.w));
        MX = aw * a23 + ay * a34 + az * a24;
        MY = ax * a34 + aw * a13 + az * a14;
        MZ = ax * a24 + ay * a14 + aw * a12;
        MW = ax * a23 + ay * a13 + az * a12;
    }
}

// Error bound factor of the mixed form: 64 units in the last place of FP32
// (first-order analysis gives 13: 4 from rounding the FP64 planes, 9 from the
// cofactor and dot-product evaluation).
constexpr float kFilterKappa = 0x1.0p-18f;
// vertices with |W| below this fraction of its term magnitude are
// ill-conditioned in FP32 and sent to FP64
constexpr float kFilterMinW = 1.0e-3f;

template <typename R, bool FILTER, int S>
struct CellState {
    uint32_t tri[S];
    R X[S], Y[S], Z[S], W[S];
    float MX[S], MY[S], MZ[S], MW[S];
    R r2[S];
    uint64_t valid[S];
};

// R: precision of the plane-side tests and of the vertex registers;
// RP: precision of the stored planes and of the cell geometry (bbox, cone
// planes, integration).  FP64 form <double, double>, mixed <float, double>,
// FP32 form <float, float>.
template <typename R, typename RP, bool FILTER, int S>
__global__ __launch_bounds__(256) void clip_kernel(ClipArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int lane = lane_id();
    const int wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int64_t t = int64_t(blockIdx.x) * wpb + wv;
    if (t >= a.n_ids) return;  // wave-uniform

    char *L = lds_raw + size_t(wv) * a.lds_stride;
    Pl<RP> *planes = reinterpret_cast<Pl<RP> *>(L);
    size_t off = (size_t(a.P) * sizeof(Pl<RP>) + 15) & ~size_t(15);
    uint32_t *rt = reinterpret_cast<uint32_t *>(L + off);
    off += 4 * 64 * S;
    uint32_t *nvs = reinterpret_cast<uint32_t *>(L + off);
    off += 4 * 64 * S;
    uint8_t *nxt = reinterpret_cast<uint8_t *>(L + off);
    off += 256;
    int *tl = reinterpret_cast<int *>(L + off);
    off += 4 * size_t(a.tl_cap);
    int *tu = reinterpret_cast<int *>(L + off);

    const int sid = a.ids ? a.ids[t] : int(t);
    const double4 S0 = a.seeds[sid];
    const int orig = a.perm[sid];
    const int K = a.K, P = a.P, V = a.V;

    auto finish = [&](int st, double m0, double m1, double m2, double m3, int ku) {
        if (lane == 0) {
            a.out[orig] = make_double4(m0, m1, m2, m3);
            a.status[orig] = st;
            if (a.kused) a.kused[orig] = ku;
        }
    };

    if (S0.w <= RPD_NO_CELL_LIMITS) {  // border seed: copied (ConvexCell.cl:873-876)
        finish(RPD_SUCCESS, S0.x, S0.y, S0.z, S0.w, 0);
        return;
    }

    for (int i = lane; i < 256; i += 64) nxt[i] = 0xff;

    // Domain cull: a seed whose voxel is d >= 2 voxels (Chebyshev) from every
    // voxel that holds triangles or lies inside has no point of the domain
    // within (d - 1) * vs; once the cell's circumradius about the seed is
    // below that, its restriction to the domain is empty (volume 0, as the
    // reference's apex-outside / no-triangle case, ConvexCell.cl:1067-1079).
    double cull_r2 = 0;
    if (a.g > 0 && a.dist) {
        const int g = a.g;
        auto vox = [&](double c) { return int(fmin(fmax(c / a.vs, 0.0), double(g - 1))); };
        const int d = a.dist[vox(S0.x) + vox(S0.y) * g + vox(S0.z) * g * g];
        if (d >= 2) cull_r2 = double(d - 1) * a.vs * double(d - 1) * a.vs;
    }
    const double wgap = S0.w - a.wmin;  // >= 0 except for the lightest seeds

    // ---- CCInit: bounding cube in the local frame ---------------------------
    const double eps = a.g > 0 ? 0.1 : 0.0;  // CUBE_EPSILON, voronoi.cpp:929
    const double vmin = -eps, vmax = 1000.0 + eps;
    if (lane < 6) {
        const int ax = lane >> 1;
        const double sc = ax == 0 ? S0.x : (ax == 1 ? S0.y : S0.z);
        const double sgn = (lane & 1) ? -1.0 : 1.0;
        const double d = (lane & 1) ? vmax - sc : sc - vmin;
        Pl<RP> p;
        p.x = ax == 0 ? RP(sgn) : RP(0);
        p.y = ax == 1 ? RP(sgn) : RP(0);
        p.z = ax == 2 ? RP(sgn) : RP(0);
        p.w = RP(d);
        planes[lane] = p;
    }
    wsync();
    int nb_plane = 6;

    CellState<R, FILTER, S> cs;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        cs.valid[s] = 0;
        cs.tri[s] = 0;
        cs.X[s] = cs.Y[s] = cs.Z[s] = R(0);
        cs.W[s] = R(1);
        cs.MX[s] = cs.MY[s] = cs.MZ[s] = cs.MW[s] = 0.f;
        cs.r2[s] = R(0);
    }
    {
        // CCInit vertex triples (ConvexCell.cl:299-306)
        const uint32_t cube_tri[8] = {pack3(2, 5, 0), pack3(5, 3, 0), pack3(1, 5, 2), pack3(5, 1, 3),
                                      pack3(4, 2, 0), pack3(4, 0, 3), pack3(2, 4, 1), pack3(4, 3, 1)};
        uint32_t tr = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (lane == i) tr = cube_tri[i];
        if (lane < 8) {
            cs.tri[0] = tr;
            vertex_from_planes<R, FILTER>(cvt<R>(planes[t0(tr)]), cvt<R>(planes[t1(tr)]), cvt<R>(planes[t2(tr)]),
                                          cs.X[0], cs.Y[0], cs.Z[0], cs.W[0], cs.MX[0], cs.MY[0], cs.MZ[0],
                                          cs.MW[0]);
            const R ix = cs.X[0] / cs.W[0], iy = cs.Y[0] / cs.W[0], iz = cs.Z[0] / cs.W[0];
            cs.r2[0] = ix * ix + iy * iy + iz * iz;
        }
        cs.valid[0] = 0xffull;
    }

    auto max_r2 = [&]() -> R {
        R m = R(0);
#pragma unroll
        for (int s = 0; s < S; ++s)
            if ((cs.valid[s] >> lane) & 1) m = cs.r2[s] > m ? cs.r2[s] : m;
        return wave_max(m);
    };
    R maxr2 = max_r2();

    // ---- one clip by plane e; returns kNoClip, kClipped or a status --------
    // e: the plane in the test precision (wave-uniform); the lane wl holds
    // the same plane in the storage precision in eP (wl < 0: every lane does)
    auto clip = [&](const Pl<R> &e, int wl, const Pl<RP> &eP) -> int {
        uint64_t rem[S];
        int nr = 0, nvalid = 0;
        uint64_t flag = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool v = (cs.valid[s] >> lane) & 1;
            const R sv = e.x * cs.X[s] + e.y * cs.Y[s] + e.z * cs.Z[s] + e.w * cs.W[s];
            rem[s] = ballot(v && sv > R(0));
            if (FILTER) {
                const float bound = kFilterKappa * (fabsf(float(e.x)) * cs.MX[s] + fabsf(float(e.y)) * cs.MY[s] +
                                                    fabsf(float(e.z)) * cs.MZ[s] + fabsf(float(e.w)) * cs.MW[s]);
                flag |= ballot(v && fabsf(float(sv)) <= bound);
            }
            nr += __popcll(rem[s]);
            nvalid += __popcll(cs.valid[s]);
        }
        if (FILTER && flag) return RPD_NEEDS_EXACT_PREDICATES;
        if (nr == 0) return kNoClip;
        if (nr == nvalid) return RPD_EMPTY_CELL;
        const int pid = nb_plane;
        if (lane == (wl < 0 ? 0 : wl)) planes[pid] = eP;
        // removed triangles, in slot-major order
        int carry = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if ((rem[s] >> lane) & 1) rt[carry + mbcnt(rem[s])] = cs.tri[s];
            carry += __popcll(rem[s]);
        }
        wsync();
        // boundary edges: an edge u->v of a removed triangle whose reverse
        // v->u belongs to no removed triangle
        uint32_t bm[S];
        int cnt_lane = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            bm[s] = 0;
            if ((rem[s] >> lane) & 1) {
                const uint32_t tr = cs.tri[s];
                const uint32_t a0 = t0(tr), a1 = t1(tr), a2 = t2(tr);
                const uint32_t r01 = (a1 << 8) | a0, r12 = (a2 << 8) | a1, r20 = (a0 << 8) | a2;
                bool h01 = false, h12 = false, h20 = false;
                for (int j = 0; j < nr; ++j) {
                    const uint32_t u = rt[j];
                    const uint32_t e0 = (t0(u) << 8) | t1(u), e1 = (t1(u) << 8) | t2(u), e2 = (t2(u) << 8) | t0(u);
                    h01 |= (r01 == e0) | (r01 == e1) | (r01 == e2);
                    h12 |= (r12 == e0) | (r12 == e1) | (r12 == e2);
                    h20 |= (r20 == e0) | (r20 == e1) | (r20 == e2);
                }
                bm[s] = (h01 ? 0u : 1u) | (h12 ? 0u : 2u) | (h20 ? 0u : 4u);
            }
            cnt_lane += __popc(bm[s]);
        }
        // slot-major exclusive offsets of the new vertices
        int nb = 0;
        int obase[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = __popc(bm[s]);
            const int o = wave_excl_scan(c);
            obase[s] = nb + o;
            nb += uni(rlane(o + c, 63));
        }
        int nkeep = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) nkeep += __popcll(cs.valid[s] & ~rem[s]);
        if (nkeep + nb > V) return RPD_VERTEX_OVERFLOW;
        // stage new triples (plane, a, b) and the boundary successor map
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (bm[s]) {
                const uint32_t tr = cs.tri[s];
                const uint32_t v3[3] = {t0(tr), t1(tr), t2(tr)};
                int o = obase[s];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (bm[s] & (1u << k)) {
                        const uint32_t ea = v3[k], eb = v3[(k + 1) % 3];
                        nvs[o++] = pack3(uint32_t(pid), ea, eb);
                        nxt[ea] = uint8_t(eb);
                    }
            }
        }
        wsync();
        // one simple cycle: every start plane written once, and the walk
        // from one start closes after exactly nb steps
        bool bad = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (bm[s]) {
                const uint32_t tr = cs.tri[s];
                const uint32_t v3[3] = {t0(tr), t1(tr), t2(tr)};
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (bm[s] & (1u << k)) bad |= nxt[v3[k]] != v3[(k + 1) % 3];
            }
        }
        bool inconsistent = ballot(bad) != 0;
        if (!inconsistent) {
            const uint32_t start = t1(uni(nvs[0]));
            uint32_t cur = start;
            int steps = 0;
            do {
                cur = uni(uint32_t(nxt[cur]));
                ++steps;
            } while (cur != start && cur != 0xffu && steps <= nb);
            inconsistent = !(cur == start && steps == nb);
        }
        wsync();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (bm[s]) {
                const uint32_t tr = cs.tri[s];
                if (bm[s] & 1u) nxt[t0(tr)] = 0xff;
                if (bm[s] & 2u) nxt[t1(tr)] = 0xff;
                if (bm[s] & 4u) nxt[t2(tr)] = 0xff;
            }
        }
        if (inconsistent) return RPD_INCONSISTENT_BOUNDARY;
        // new vertices fill the lowest free slots
        carry = 0;
        uint64_t newm[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            cs.valid[s] &= ~rem[s];
            const uint64_t fr = ~cs.valid[s];
            const int q = carry + mbcnt(fr);
            const bool take = ((fr >> lane) & 1) && q < nb;
            newm[s] = ballot(take);
            if (take) {
                const uint32_t tr = nvs[q];
                cs.tri[s] = tr;
                const Pl<R> pa = cvt<R>(planes[t1(tr)]), pb = cvt<R>(planes[t2(tr)]);
                vertex_from_planes<R, FILTER>(e, pa, pb, cs.X[s], cs.Y[s], cs.Z[s], cs.W[s], cs.MX[s], cs.MY[s],
                                              cs.MZ[s], cs.MW[s]);
                const R ix = cs.X[s] / cs.W[s], iy = cs.Y[s] / cs.W[s], iz = cs.Z[s] / cs.W[s];
                cs.r2[s] = ix * ix + iy * iy + iz * iz;
            }
            carry += __popcll(fr);
            cs.valid[s] |= newm[s];
        }
        nb_plane = pid + 1;
        if (FILTER) {
            uint64_t ill = 0;
#pragma unroll
            for (int s = 0; s < S; ++s)
                ill |= ballot(((newm[s] >> lane) & 1) && fabsf(float(cs.W[s])) < kFilterMinW * cs.MW[s]);
            if (ill) return RPD_NEEDS_EXACT_PREDICATES;
        }
        maxr2 = max_r2();
        return kClipped;
    };

    // ---- fused integration (CCExportBaryVolume) into FP64 accumulators -----
    // vertex coordinates in the geometry precision RP: the registers when
    // R == RP, else recomputed from the stored RP planes
    auto vertex_rp = [&](int s, RP &Ax, RP &Ay, RP &Az) {
        if (sizeof(R) == sizeof(RP)) {
            Ax = RP(cs.X[s] / cs.W[s]);
            Ay = RP(cs.Y[s] / cs.W[s]);
            Az = RP(cs.Z[s] / cs.W[s]);
        } else {
            const uint32_t tr = cs.tri[s];
            RP X, Y, Z, W;
            float d0, d1, d2, d3;
            vertex_from_planes<RP, false>(planes[t0(tr)], planes[t1(tr)], planes[t2(tr)], X, Y, Z, W, d0, d1, d2,
                                          d3);
            Ax = X / W;
            Ay = Y / W;
            Az = Z / W;
        }
    };
    double acc_m = 0, acc_x = 0, acc_y = 0, acc_z = 0;
    const RP vol_eps2 = sizeof(RP) == 8 ? RP(1e-12) : RP(1e-6);  // voronoi.cpp:919-922
    auto integrate = [&](double factor) -> bool {
        RP sw = 0, sx = 0, sy = 0, sz = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!((cs.valid[s] >> lane) & 1)) continue;
            const uint32_t tr = cs.tri[s];
            RP Ax, Ay, Az;
            vertex_rp(s, Ax, Ay, Az);
            const RP ACx = -Ax, ACy = -Ay, ACz = -Az;  // seed at the origin
            Pl<RP> pl[3] = {planes[t0(tr)], planes[t1(tr)], planes[t2(tr)]};
            RP opx[3], opy[3], opz[3], olx[3], oly[3], olz[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const RP nn = pl[i].x * pl[i].x + pl[i].y * pl[i].y + pl[i].z * pl[i].z;
                const RP f = (ACx * pl[i].x + ACy * pl[i].y + ACz * pl[i].z) / nn;
                opx[i] = ACx - f * pl[i].x;
                opy[i] = ACy - f * pl[i].y;
                opz[i] = ACz - f * pl[i].z;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const Pl<RP> &p = pl[i], &q = pl[(i + 1) % 3];
                const RP cx = p.y * q.z - p.z * q.y, cy = p.z * q.x - p.x * q.z, cz = p.x * q.y - p.y * q.x;
                const RP c2 = cx * cx + cy * cy + cz * cz;
                const RP pp = p.x * p.x + p.y * p.y + p.z * p.z, qq = q.x * q.x + q.y * q.y + q.z * q.z;
                if (c2 < vol_eps2 * pp * qq) {
                    olx[i] = oly[i] = olz[i] = RP(0);
                } else {
                    const RP f = (ACx * cx + ACy * cy + ACz * cz) / c2;
                    olx[i] = f * cx;
                    oly[i] = f * cy;
                    olz[i] = f * cz;
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
#pragma unroll
                for (int st = 0; st < 2; ++st) {
                    const int j = (i + 3 - st) % 3;
                    // det3 of rows AC, onPlane_i, onLine_j (ConvexCell.cl:72-77)
                    const RP d = ACx * (opy[i] * olz[j] - opz[i] * oly[j]) - opx[i] * (ACy * olz[j] - ACz * oly[j]) +
                                 olx[j] * (ACy * opz[i] - ACz * opy[i]);
                    const RP w = st ? d : -d;
                    sw += w;
                    sx += w * (Ax + RP(0.25) * (ACx + opx[i] + olx[j]));
                    sy += w * (Ay + RP(0.25) * (ACy + opy[i] + oly[j]));
                    sz += w * (Az + RP(0.25) * (ACz + opz[i] + olz[j]));
                }
            }
        }
        const double f = factor / 6.0;
        acc_m += f * wave_sum(double(sw));
        acc_x += f * wave_sum(double(sx));
        acc_y += f * wave_sum(double(sy));
        acc_z += f * wave_sum(double(sz));
        return true;
    };

    auto fail = [&](int st, int ku) { finish(st, 0.0, 0.0, 0.0, 0.0, ku); };

    // ---- bisector loop (ConvexCell.cl:882-913) --------------------------------
    // The 64 neighbours of a chunk are loaded one per lane with their plane,
    // distance |b| and plane offset t (signed distance of the bisector from
    // the seed).  A plane with t >= R, R the distance from the seed to the
    // farthest vertex, cannot cut the cell, so instead of visiting every
    // neighbour the wave ballots for the next neighbour whose plane may cut
    // (t < R) and for the first one where the stop test holds, and jumps to
    // the earlier of the two.  Neighbours in between would neither clip nor
    // stop, so the cell, the stop index and the plane-overflow rule (a full
    // plane list fails at the next neighbour, ConvexCell.cl:316-319) are
    // those of the one-by-one loop.
    const int *nrow = a.knn + t * int64_t(K);
    bool reached = false, complete = false;
    int kused = 0;
    Pl<RP> cpl;
    Pl<R> cplr;
    double cdl = 0, ct = 0;
    int cnb = -1;
    int loaded = -1;
    const double rmarg = sizeof(R) == 8 ? 1.0 + 1e-9 : 1.0 + 2e-4;  // FP32 R may be low by ~1e-4
    double Rcur = sqrt(double(maxr2));
    auto stop_at = [&](double dl, double Rm) -> bool {
        const double Rs = Rm * rmarg;
        if (a.security_mode == 1 || wgap <= 0) return dl * dl > 4.0 * Rs * Rs;
        return dl * (dl - 2.0 * Rs) >= wgap;
    };
    int i = 0;
    while (i < K) {
        const int base = i & ~63;
        if (base != loaded) {
            const int j = base + lane;
            cnb = j < K ? nrow[j] : -2;  // -1: list exhausted, -2: past K
            cpl = Pl<RP>{RP(0), RP(0), RP(0), RP(0)};
            cdl = 0;
            ct = 0;
            if (cnb >= 0) {
                const double4 B = a.seeds[cnb];
                const double bx = B.x - S0.x, by = B.y - S0.y, bz = B.z - S0.z;
                const double b2 = bx * bx + by * by + bz * bz;
                const double bl = sqrt(b2);
                double d;
                if (B.w <= RPD_NO_CELL_LIMITS)
                    d = bl;  // plane through B (ConvexCell.cl:325)
                else
                    d = (b2 + (B.w - S0.w)) / (2.0 * bl);
                cpl = Pl<RP>{RP(-bx / bl), RP(-by / bl), RP(-bz / bl), RP(d)};
                cdl = bl;
                ct = d;
            }
            cplr = cvt<R>(cpl);
            loaded = base;
        }
        const int l0 = i & 63;
        const bool in = lane >= l0 && cnb != -2;
        if (nb_plane >= P) {
            // the next neighbour's plane cannot be stored
            const uint64_t me = ballot(in && cnb == -1), mv = ballot(in && cnb >= 0);
            if (mv && (!me || lowest_bit(mv) < lowest_bit(me))) {
                fail(RPD_PLANE_OVERFLOW, base + lowest_bit(mv) + 1);
                return;
            }
            if (me) {
                complete = true;
                break;
            }
            i = base + 64;
            continue;
        }
        const double Rskip = Rcur * rmarg;
        const uint64_t m_end = ballot(in && cnb == -1);
        const uint64_t m_clip = ballot(in && cnb >= 0 && ct < Rskip);
        const uint64_t m_stop = ballot(in && cnb >= 0 && stop_at(cdl, Rcur));
        const int f_end = m_end ? lowest_bit(m_end) : 64;
        const int f_clip = m_clip ? lowest_bit(m_clip) : 64;
        const int f_stop = m_stop ? lowest_bit(m_stop) : 64;
        if (f_end < f_clip && f_end < f_stop) {
            complete = true;  // every other seed has been used
            kused = base + f_end;
            break;
        }
        if (f_stop < f_clip) {
            reached = true;
            kused = base + f_stop + 1;
            break;
        }
        if (f_clip == 64) {
            i = base + 64;
            kused = min(K, i);
            continue;
        }
        // clip by the plane of neighbour base + f_clip
        const Pl<R> e = rlane_pl(cplr, f_clip);
        const int res = clip(e, f_clip, cpl);
        kused = base + f_clip + 1;
        if (res >= 0) {
            fail(res, kused);
            return;
        }
        if (res == kClipped) {
            Rcur = sqrt(double(maxr2));
            if (double(maxr2) * 1.001 < cull_r2) {
                finish(RPD_SUCCESS, 0.0, 0.0, 0.0, 0.0, kused);
                return;
            }
        }
        if (stop_at(rlane(cdl, f_clip), Rcur)) {
            reached = true;
            break;
        }
        i = base + f_clip + 1;
    }
    if (!reached && !complete) {
        fail(RPD_SECURITY_RADIUS_NOT_REACHED, kused);
        return;
    }

    // ---- restriction to the domain grid (ConvexCell.cl:918-1116) -------------
    bool add_volume = true;
    int ntri = 0;
    double P0x = 0, P0y = 0, P0z = 0;  // apex, global frame
    if (a.g > 0) {
        const int g = a.g;
        const double vs = a.vs;
        double bmn[3] = {1000.0, 1000.0, 1000.0}, bmx[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!((cs.valid[s] >> lane) & 1)) continue;
            RP lx, ly, lz;
            vertex_rp(s, lx, ly, lz);
            const double gx = S0.x + double(lx), gy = S0.y + double(ly), gz = S0.z + double(lz);
            bmn[0] = fmin(bmn[0], gx);
            bmn[1] = fmin(bmn[1], gy);
            bmn[2] = fmin(bmn[2], gz);
            bmx[0] = fmax(bmx[0], gx);
            bmx[1] = fmax(bmx[1], gy);
            bmx[2] = fmax(bmx[2], gz);
        }
        int vmn[3], vmx[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // (int)(b / vs) clamped to [0, g-1] (ConvexCell.cl:939-940); the
            // clamp is applied in floating point first so that inf/NaN from a
            // degenerate vertex cannot produce an out-of-grid range
            const double lo = wave_min(bmn[d]) / vs, hi = wave_max(bmx[d]) / vs;
            vmn[d] = int(fmin(fmax(lo, 0.0), double(g - 1)));
            vmx[d] = int(fmin(fmax(hi, 0.0), double(g - 1)));
            if (vmx[d] < vmn[d]) vmx[d] = vmn[d];
        }
        const int nx = vmx[0] - vmn[0] + 1, ny = vmx[1] - vmn[1] + 1, nz = vmx[2] - vmn[2] + 1;
        const int nvox = nx * ny * nz;
        int first_ref = -1, first_c = -1, ref_flag = 1, c_flag = 1;
        int ncol = 0;
        bool overflow = false;
        for (int vb = 0; vb < nvox && !overflow; vb += 64) {
            const int v = vb + lane;
            const bool in = v < nvox;
            int f = 1, cf = 1, o0 = 0, cnt = 0;
            if (in) {
                const int x = vmn[0] + v % nx, y = vmn[1] + (v / nx) % ny, z = vmn[2] + v / (nx * ny);
                const int id = x + y * g + z * g * g;
                f = a.flags[id];
                o0 = a.toffs[id];
                cnt = a.toffs[id + 1] - o0;
                cf = a.centre ? a.centre[id] : (cnt == 0 ? f : 1);
            }
            const uint64_t mr = ballot(in && f != 1);
            if (first_ref < 0 && mr) {
                const int l = lowest_bit(mr);
                first_ref = vb + l;
                ref_flag = rlane(f, l);
            }
            const uint64_t mc = ballot(in && cf != 1);
            if (first_c < 0 && mc) {
                const int l = lowest_bit(mc);
                first_c = vb + l;
                c_flag = rlane(cf, l);
            }
            const int o = wave_excl_scan(cnt);
            const int tot = uni(rlane(o + cnt, 63));
            if (ncol + tot > a.tl_cap) {
                overflow = true;
                break;
            }
            for (int k = 0; k < cnt; ++k) tl[ncol + o + k] = a.tlist[o0 + k];
            ncol += tot;
        }
        if (overflow) {
            fail(RPD_TRIANGLE_OVERFLOW, kused);
            return;
        }
        wsync();
        // sorted-union semantics of merge (ConvexCell.cl:722-749): keep the
        // first occurrence of every triangle id
        int nu = 0;
        for (int ib = 0; ib < ncol; ib += 64) {
            const int i = ib + lane;
            const int val = i < ncol ? tl[i] : -1;
            bool dup = false;
            const int jend = min(ncol, ib + 64);
            for (int j = 0; j < jend; ++j) dup |= (j < i) && (tl[j] == val);
            const bool keep = i < ncol && !dup;
            const uint64_t m = ballot(keep);
            if (keep && nu + mbcnt(m) < a.T) tu[nu + mbcnt(m)] = val;
            nu += __popcll(m);
        }
        if (nu > a.T) {
            fail(RPD_TRIANGLE_OVERFLOW, kused);
            return;
        }
        ntri = nu;
        // apex
        int in_domain = 1;
        auto vox_xyz = [&](int v, int &x, int &y, int &z) {
            x = vmn[0] + v % nx;
            y = vmn[1] + (v / nx) % ny;
            z = vmn[2] + v / (nx * ny);
        };
        if (a.apex_mode == 1 && first_c >= 0) {
            int x, y, z;
            vox_xyz(first_c, x, y, z);
            P0x = (x + 0.5) * vs;
            P0y = (y + 0.5) * vs;
            P0z = (z + 0.5) * vs;
            in_domain = c_flag;
        } else if (first_ref >= 0) {
            int x, y, z;
            vox_xyz(first_ref, x, y, z);
            P0x = x * vs;
            P0y = y * vs;
            P0z = z * vs;
            in_domain = ref_flag;
        } else {
            // GRID_OUT_DOMAIN fallback (ConvexCell.cl:985-1051)
            auto flag = [&](int x, int y, int z) { return int(a.flags[x + y * g + z * g * g]); };
            auto setc = [&](int x, int y, int z) {
                P0x = x * vs;
                P0y = y * vs;
                P0z = z * vs;
            };
            if (vmx[2] + 1 < g) {
                const int z = vmx[2] + 1;
                for (int y = vmn[1]; y <= vmx[1] + 1 && y < g && in_domain == 1; y++)
                    for (int x = vmn[0]; x <= vmx[0] + 1 && x < g; x++)
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            setc(x, y, z);
                            break;
                        }
            } else {
                in_domain = 0;
                setc(vmn[0], vmn[1], vmx[2] + 1);
            }
            if (in_domain == 1 && vmx[1] + 1 < g) {
                const int y = vmx[1] + 1;
                for (int z = vmn[2]; z <= vmx[2] && in_domain == 1; z++)
                    for (int x = vmn[0]; x <= vmx[0] + 1 && x < g; x++)
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            setc(x, y, z);
                            break;
                        }
            } else if (in_domain == 1) {
                in_domain = 0;
                setc(vmn[0], vmx[1] + 1, vmn[2]);
            }
            if (in_domain == 1 && vmx[0] + 1 < g) {
                const int x = vmx[0] + 1;
                for (int z = vmn[2]; z <= vmx[2] && in_domain == 1; z++)
                    for (int y = vmn[1]; y <= vmx[1]; y++)
                        if (flag(x, y, z) != 1) {
                            in_domain = flag(x, y, z);
                            setc(x, y, z);
                            break;
                        }
            } else if (in_domain == 1) {
                in_domain = 0;
                setc(vmx[0] + 1, vmn[1], vmn[2]);
            }
            if (in_domain == 1) {
                fail(RPD_FIND_ANOTHER_BEGINNING_VERTEX, kused);
                return;
            }
        }
        add_volume = in_domain == 2;
    }
    if (add_volume) {
        if (!integrate(1.0)) {
            fail(RPD_NEEDS_EXACT_PREDICATES, kused);
            return;
        }
    }
    if (ntri > 0) {
        // backup (ConvexCell.cl:692-715): the vertex registers and the plane count
        CellState<R, FILTER, S> bk = cs;
        const int bk_np = nb_plane;
        const R bk_r2 = maxr2;
        const double lx0 = P0x - S0.x, ly0 = P0y - S0.y, lz0 = P0z - S0.z;
        for (int k = 0; k < ntri; ++k) {
            const int tid = tu[k];
            const int i1 = a.tris[3 * tid], i2 = a.tris[3 * tid + 1], i3 = a.tris[3 * tid + 2];
            const double4 G1 = a.gpts[i1];
            double4 G2 = a.gpts[i2], G3 = a.gpts[i3];
            // orientation test in the global frame, as the reference
            const double ux = G1.x - P0x, uy = G1.y - P0y, uz = G1.z - P0z;
            const double vx = G2.x - P0x, vy = G2.y - P0y, vz = G2.z - P0z;
            const double wx = G3.x - P0x, wy = G3.y - P0y, wz = G3.z - P0z;
            const double det = ux * (vy * wz - vz * wy) - vx * (uy * wz - uz * wy) + wx * (uy * vz - uz * vy);
            const bool go_out = det > 0.0;
            if (!go_out) {
                const double4 tmp = G2;
                G2 = G3;
                G3 = tmp;
            }
            // cone planes (A,B,C), (A,D,B), (A,C,D), (C,D,B) with A = apex
            const double Ax = lx0, Ay = ly0, Az = lz0;
            const double Bx = G1.x - S0.x, By = G1.y - S0.y, Bz = G1.z - S0.z;
            const double Cx = G2.x - S0.x, Cy = G2.y - S0.y, Cz = G2.z - S0.z;
            const double Dx = G3.x - S0.x, Dy = G3.y - S0.y, Dz = G3.z - S0.z;
            const double F[4][9] = {{Ax, Ay, Az, Bx, By, Bz, Cx, Cy, Cz},
                                    {Ax, Ay, Az, Dx, Dy, Dz, Bx, By, Bz},
                                    {Ax, Ay, Az, Cx, Cy, Cz, Dx, Dy, Dz},
                                    {Cx, Cy, Cz, Dx, Dy, Dz, Bx, By, Bz}};
            int res = kNoClip;
            for (int f = 0; f < 4; ++f) {
                if (nb_plane >= P) {
                    res = RPD_PLANE_OVERFLOW;
                    break;
                }
                const double e1x = F[f][3] - F[f][0], e1y = F[f][4] - F[f][1], e1z = F[f][5] - F[f][2];
                const double e2x = F[f][6] - F[f][0], e2y = F[f][7] - F[f][1], e2z = F[f][8] - F[f][2];
                double nx_ = e1y * e2z - e1z * e2y, ny_ = e1z * e2x - e1x * e2z, nz_ = e1x * e2y - e1y * e2x;
                double nd = -(nx_ * F[f][0] + ny_ * F[f][1] + nz_ * F[f][2]);
                const double nl = sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
                if (nl > 0) {
                    nx_ /= nl;
                    ny_ /= nl;
                    nz_ /= nl;
                    nd /= nl;
                }
                const Pl<RP> cp{RP(nx_), RP(ny_), RP(nz_), RP(nd)};
                res = clip(cvt<R>(cp), -1, cp);
                if (res >= 0) break;
            }
            if (res == RPD_EMPTY_CELL) {
                cs = bk;
                nb_plane = bk_np;
                maxr2 = bk_r2;
                continue;
            }
            if (res >= 0) {
                fail(res, kused);
                return;
            }
            if (!integrate(go_out ? 1.0 : -1.0)) {
                fail(RPD_NEEDS_EXACT_PREDICATES, kused);
                return;
            }
            cs = bk;
            nb_plane = bk_np;
            maxr2 = bk_r2;
        }
    }
    // local first moments -> global frame
    finish(RPD_SUCCESS, acc_x + acc_m * S0.x, acc_y + acc_m * S0.y, acc_z + acc_m * S0.z, acc_m, kused);
}

}  // namespace rpd

using namespace rpd;

int rpd_launch_clip(hipStream_t st, int form, const ClipLaunch &L) {
    if (L.n_ids <= 0) return RPD_OK;
    if (L.P < 7 || L.P > 255 || L.V < 8 || L.K < 1) return RPD_E_ARG;
    ClipArgs a{};
    a.seeds = L.seeds;
    a.perm = L.perm;
    a.ids = L.ids;
    a.knn = L.knn;
    a.n_ids = L.n_ids;
    a.K = L.K;
    a.P = L.P;
    a.V = L.V;
    a.T = std::max(L.T, 0);
    a.apex_mode = L.apex_mode;
    if (L.dom) {
        a.g = L.dom->g;
        a.vs = L.dom->vs;
        a.flags = L.dom->flags;
        a.centre = L.dom->centre;
        a.gpts = L.dom->pts;
        a.tris = L.dom->tris;
        a.tlist = L.dom->tlist;
        a.toffs = L.dom->toffs;
        a.dist = L.cull ? L.dom->dist : nullptr;
    }
    a.security_mode = L.security_mode;
    a.wmin = L.wmin;
    a.out = L.out;
    a.status = L.status;
    a.kused = L.kused;
    if (L.V > 256) return RPD_E_LIMIT;
    const int S = L.V <= 64 ? 1 : (L.V <= 128 ? 2 : 4);
    a.tl_cap = 4 * a.T + 64;
    const size_t pl_bytes = size_t(L.P) * (form == RPD_PREC_FP32 ? 16 : 32);
    size_t stride = ((pl_bytes + 15) & ~size_t(15)) + 2 * 4 * 64 * S + 256 + 4 * size_t(a.tl_cap + a.T);
    stride = (stride + 15) & ~size_t(15);
    a.lds_stride = int(stride);
    int wpb = 4;
    while (wpb > 1 && stride * wpb > 64 * 1024) wpb >>= 1;
    if (stride > 64 * 1024) return RPD_E_LIMIT;
    const unsigned blocks = unsigned((L.n_ids + wpb - 1) / wpb);
    const dim3 tb(64 * wpb);
    const size_t shm = stride * wpb;
#define RPD_CLIP_LAUNCH(R, RP, F)                                                          \
    do {                                                                                   \
        if (S == 1)                                                                        \
            clip_kernel<R, RP, F, 1><<<blocks, tb, shm, st>>>(a);                          \
        else if (S == 2)                                                                   \
            clip_kernel<R, RP, F, 2><<<blocks, tb, shm, st>>>(a);                          \
        else                                                                               \
            clip_kernel<R, RP, F, 4><<<blocks, tb, shm, st>>>(a);                          \
    } while (0)
    if (form == RPD_PREC_FP64)
        RPD_CLIP_LAUNCH(double, double, false);
    else if (form == RPD_PREC_MIXED)
        RPD_CLIP_LAUNCH(float, double, true);
    else
        RPD_CLIP_LAUNCH(float, float, false);
#undef RPD_CLIP_LAUNCH
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}
