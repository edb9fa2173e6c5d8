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
// Two kernels per pass:
//   clip_cells_kernel   the bisector loop for every cell (light, high
//                       occupancy).  A cell that provably misses the domain
//                       ends there with volume 0; a finished cell is appended
//                       to a queue as a compact record (FP64 planes + vertex
//                       triples);
//   finish_cells_kernel restriction to the domain (bbox, voxels, apex, cone
//                       clips per boundary triangle) and the integration, for
//                       the queued cells only.
//
// Everything is in the seed-local frame (origin at the seed, FP64 inputs):
// planes n.y + d with d = (|b|^2 + w_B - w_s) / (2|b|), b = B - s.
//
// Forms: FP64 (all double); mixed: the plane-side tests and the vertex
// registers are FP32 and each test also evaluates a certified bound on its
// FP32 error against the FP64 planes, 2^-18 * sum of |terms| of the det4x4
// expansion.  In the bisector kernel (FP32 planes in LDS) a test inside the
// bound or an ill-conditioned new vertex ends the cell with
// needs_exact_predicates and the host re-runs it in FP64; in the finish kernel
// (FP64 planes in LDS) a test inside the bound, and every test of an
// ill-conditioned vertex, is decided in FP64 from the vertex's planes, as the
// FP64 form decides it.  The
// planes are stored in FP64, and the cell geometry that produces the results
// (bbox, cone planes, volume and moments) is evaluated in FP64 from them, so
// an unflagged mixed cell has the FP64 combinatorics and FP64 geometry.  FP32
// form: everything in FP32, no filter.  Moments are accumulated in FP64.

#include <hip/hip_runtime.h>
#include <atomic>

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
    int64_t n_seeds;
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
    int lds_stride;   // LDS bytes of one wave
    int cell_stride;  // LDS bytes shared by the waves of one cell (finish kernel)
    const int *tvoffs;    // per triangle: offsets into tvox
    const uint32_t *tvox; // per triangle: sorted voxels whose list holds it, z<<20 | y<<10 | x
    const uint4 *tinfo;   // per triangle-list entry: its voxel, earlier near voxels holding it
    int security_mode;    // 0 exact power stop, 1 reference stop
    double wmin;          // minimum weight over non-border seeds
    const uint8_t *dist;  // Chebyshev voxel distance to the nearest domain voxel (nullable)
    char *rec;            // queue of finished cells (records of rec_stride bytes)
    int rec_stride;
    int *qcount;          // number of queued records
    // failed cells for the next pass (nullable): fl_cnt[0] exact-predicate
    // entries, stored at fl_eo[-1], fl_eo[-2], ..., [1] overflow entries at
    // fl_eo[0], fl_eo[1], ... (so that both kinds together are one range),
    // [2 + s] cells that ended with status s, [13] the largest triangle union
    // of a cell that overflowed T
    int *fl_cnt, *fl_eo;
    int exact_final;      // FP64 form: exact-predicate failures are final
    // A cell whose list runs out before its security radius is reached can be
    // resumed: its state goes to rs_out (records of rs_out_stride bytes, slot
    // from *rs_count below rs_cap, kept in rslot[sid]), the stored id of its
    // last neighbour to lastnb[sid], the cell to fl_sec (counter fl_cnt[14]).
    // rs_out null, or the buffer full: a plain overflow.  With rs_in set
    // (FP64 form), every cell resumes from record rs_in[rslot[sid]] (written
    // with rs_in_P planes) and its list holds only the neighbours after the
    // ones it has used.
    char *rs_out;
    int rs_out_stride, rs_cap;
    int *rs_count;
    const char *rs_in;
    int rs_in_stride, rs_in_P;
    int *rslot, *lastnb, *fl_sec;
    // Hessian (finish kernel, H > 0): per original seed up to H entries of
    // hc values in hess and the neighbours' original indices in hids
    double *hess;
    int *hids;
    int H, hc;
    int *dcnt;  // debug counters of this pass (kDcStride ints), nullable
    // packed batch (nullable): the fields of request req_of[sid] replace
    // n_seeds, wmin and the domain's
    const ReqDesc *reqs;
    const uint16_t *req_of;
    // finish kernel: work item i is the record at slot slots[i] (nullable:
    // slot i); fl_slot (nullable) receives the record slot of each failure it
    // lists, at the failure's index in fl_eo
    const int *slots;
    int *fl_slot;
    // layout of the records the finish kernel reads: planes and vertex
    // slots they were written with (the finish kernel's own P and S may be
    // larger)
    int rec_P, rec_S;
    // finish kernel (nullable): the queue length in device memory, read when
    // the launch covers an upper bound of it
    const int *nq_dev;
};

// The cell of stored id sid reads its request's seed count, weight minimum
// and domain (packed batch; wave-uniform, so the fields stay in SGPRs).  The
// kernels take it as a template parameter, so that a single diagram's
// kernels keep every field in the kernel arguments.
__device__ __forceinline__ void bind_request(ClipArgs &a, int sid) {
    const ReqDesc &d = a.reqs[uni(int(a.req_of[sid]))];
    a.n_seeds = d.n_seeds;
    a.wmin = d.wmin;
    a.g = d.g;
    a.vs = d.vs;
    a.flags = d.flags;
    a.centre = d.centre;
    a.gpts = d.gpts;
    a.tris = d.tris;
    a.tlist = d.tlist;
    a.toffs = d.toffs;
    a.tvoffs = d.tvoffs;
    a.tvox = d.tvox;
    a.tinfo = d.tinfo;
    a.dist = d.dist;
}

// Appends a failed cell (stored id sid) to the failure list: overflows
// (vertex, plane, K, triangle) always; needs_exact_predicates,
// inconsistent_boundary and find_another_beginning_vertex when the form was
// not FP64 (in FP64 a recomputation would fail the same way).  One lane.
__device__ __forceinline__ void note_failure(const ClipArgs &a, int sid, int st, bool resumable = false,
                                             int rslot = -1) {
    if (!a.fl_cnt || st == RPD_SUCCESS || st == RPD_EMPTY_CELL || st == RPD_HESSIAN_OVERFLOW) return;
    const bool exact = st == RPD_NEEDS_EXACT_PREDICATES || st == RPD_INCONSISTENT_BOUNDARY ||
                       st == RPD_FIND_ANOTHER_BEGINNING_VERTEX;
    if (exact && a.exact_final) return;
    atomicAdd(a.fl_cnt + 2 + st, 1);
    if (resumable) {
        a.fl_sec[atomicAdd(a.fl_cnt + 14, 1)] = sid;
        return;
    }
    const int k = atomicAdd(a.fl_cnt + (exact ? 0 : 1), 1);
    a.fl_eo[exact ? -1 - k : k] = sid;
    if (a.fl_slot) a.fl_slot[exact ? -1 - k : k] = rslot;
}

__device__ __forceinline__ uint32_t pack3(uint32_t a, uint32_t b, uint32_t c) { return a | (b << 8) | (c << 16); }
// no vertex (plane triples use 24 bits)
constexpr uint32_t kNoTri = 0xffffffffu;
__device__ __forceinline__ uint32_t t0(uint32_t t) { return t & 0xff; }
__device__ __forceinline__ uint32_t t1(uint32_t t) { return (t >> 8) & 0xff; }
__device__ __forceinline__ uint32_t t2(uint32_t t) { return (t >> 16) & 0xff; }

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
// of absolute values of their terms: MXYZ the largest of those of X, Y, Z,
// MW that of W.
template <typename R, bool FILTER>
__device__ __forceinline__ void vertex_from_planes(const Pl<R> &p1, const Pl<R> &p2, const Pl<R> &p3, R &X, R &Y,
                                                   R &Z, R &W, float &MXYZ, float &MW) {
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
        const float x2 = fabsf(float(p2.x)), y2 = fabsf(float(p2.y)), z2 = fabsf(float(p2.z)),
                    w2 = fabsf(float(p2.w));
        const float x3 = fabsf(float(p3.x)), y3 = fabsf(float(p3.y)), z3 = fabsf(float(p3.z)),
                    w3 = fabsf(float(p3.w));
        const float a12 = fmaf(x2, y3, y2 * x3), a13 = fmaf(x2, z3, z2 * x3), a14 = fmaf(x2, w3, w2 * x3);
        const float a23 = fmaf(y2, z3, z2 * y3), a24 = fmaf(y2, w3, w2 * y3), a34 = fmaf(z2, w3, w2 * z3);
        const float ax = fabsf(float(p1.x)), ay = fabsf(float(p1.y)), az = fabsf(float(p1.z)),
                    aw = fabsf(float(p1.w));
        const float MX = aw * a23 + ay * a34 + az * a24;
        const float MY = ax * a34 + aw * a13 + az * a14;
        const float MZ = ax * a24 + ay * a14 + aw * a12;
        MXYZ = fmaxf(MX, fmaxf(MY, MZ));
        MW = ax * a23 + ay * a13 + az * a12;
    }
}

// Whether the vertex of planes tr lies strictly outside plane e, in FP64:
// the vertex rebuilt from its FP64 planes (as the FP64 form builds it) and
// the FP64 form's test.  (Out of line, the call cost the finish kernel more
// spills than inlined.)
__device__ __forceinline__ bool outside64(const Pl<double> *planes, uint32_t tr, Pl<double> e) {
    double x, y, z, w;
    float mx, mw;
    vertex_from_planes<double, false>(planes[t0(tr)], planes[t1(tr)], planes[t2(tr)], x, y, z, w, mx, mw);
    return e.x * x + e.y * y + e.z * z + e.w * w > 0.0;
}

// Error bound factor of the mixed form: 64 units in the last place of FP32
// (first-order analysis gives 13: 4 from rounding the FP64 planes, 9 from the
// cofactor and dot-product evaluation).  The bound of a test e.v is
// kappa * (|e.x| MX + |e.y| MY + |e.z| MZ + |e.w| MW), taken here as
// kappa * ((|e.x| + |e.y| + |e.z|) MXYZ + |e.w| MW), which is not smaller.
constexpr float kFilterKappa = 0x1.0p-18f;
// vertices with |W| below this fraction of its term magnitude are
// ill-conditioned in FP32 and sent to FP64
constexpr float kFilterMinW = 1.0e-3f;

// apex_mode 1 puts the cone apex at this fraction of a voxel whose interior
// has one known class (in or out).  Any interior point is valid; a generic
// one keeps the cone planes off the voxel lattice and off the CUBE_EPSILON
// box (the voxel centre lies on planes such as y + z = 1000 that contain box
// edges such as y = 1000.1, z = -0.1).
__device__ constexpr double kApexFrac[3] = {0.530901699437495, 0.520710678118655, 0.536602540378444};

// CCInit vertex triples (ConvexCell.cl:299-306), one byte per vertex:
// vertex i has planes (kCube0 >> 8i, kCube1 >> 8i, kCube2 >> 8i) & 0xff.
constexpr uint64_t kCube0 = 0x0402040405010502ull;  // 2 5 1 5 4 4 2 4
constexpr uint64_t kCube1 = 0x0304000201050305ull;  // 5 3 5 1 2 0 4 3
constexpr uint64_t kCube2 = 0x0101030003020000ull;  // 0 0 2 3 0 3 1 1

// ---------------------------------------------------------------------------
// The cell of one wavefront: vertices in lanes, planes in LDS.
// ---------------------------------------------------------------------------
// The cell of one wavefront.  Planes are stored in LDS in precision PS: the
// bisector kernel keeps them in the test precision R, the finish kernel in
// the geometry precision RP.
template <typename R, typename RP, bool FILTER, int S, typename PS = RP>
struct WaveCell {
    Pl<PS> *planes;
    uint32_t *nvs;
    uint8_t *nxt;
    // bit matrix of directed edges, P rows of ceil(P / 32) words: bit v of
    // row u set while a removed vertex's triangle holds the edge u -> v
    // (all zero between clips)
    uint32_t *em;
    int P, V, nb_plane, lane;
    uint32_t tri[S];
    R X[S], Y[S], Z[S], W[S], r2[S];
    float MXYZ[S], MW[S];
    uint64_t valid[S];
    // (mixed finish kernel) vertices whose tests are all decided in FP64:
    // ill-conditioned in FP32 (|W| small against its terms)
    uint64_t x64[S];
    R maxr2;
    // FP64 planes in LDS under the mixed filter (the finish kernel): tests
    // the filter leaves open are decided in FP64 in place
    static constexpr bool kInline64 = FILTER && sizeof(PS) == 8;

    // R2: also the vertex's squared distance from the seed (the bisector
    // loop's stop test reads it through maxr2; the finish kernel does not)
    template <bool R2 = true>
    __device__ void set_vertex(int s, uint32_t tr, const Pl<R> &p1, const Pl<R> &p2, const Pl<R> &p3) {
        tri[s] = tr;
        vertex_from_planes<R, FILTER>(p1, p2, p3, X[s], Y[s], Z[s], W[s], MXYZ[s], MW[s]);
        if (R2) {
            // FP32: the hardware reciprocal (1 ulp); r2 is read only by the
            // stop and skip tests, whose radius margin (2e-4) is far larger
            R iw;
            if constexpr (sizeof(R) == 4)
                iw = __builtin_amdgcn_rcpf(W[s]);
            else
                iw = R(1) / W[s];
            const R ix = X[s] * iw, iy = Y[s] * iw, iz = Z[s] * iw;
            r2[s] = ix * ix + iy * iy + iz * iz;
        }
    }

    template <bool R2 = true>
    __device__ void set_vertex_lds(int s, uint32_t tr) {
        set_vertex<R2>(s, tr, cvt<R>(planes[t0(tr)]), cvt<R>(planes[t1(tr)]), cvt<R>(planes[t2(tr)]));
    }

    __device__ void refresh_max_r2() {
        R m = R(0);
#pragma unroll
        for (int s = 0; s < S; ++s)
            if ((valid[s] >> lane) & 1) m = r2[s] > m ? r2[s] : m;
        maxr2 = wave_max(m);
    }

    // True when plane e removes every vertex with a margin: e.v above the
    // certified FP32 error bound (mixed form) or above kSepTau times the sum
    // of |terms| of e.v (FP64 and FP32 forms).
    __device__ bool removes_all(const Pl<R> &e) const {
        bool keep = false;
        const float exyz = fabsf(float(e.x)) + fabsf(float(e.y)) + fabsf(float(e.z)), ew = fabsf(float(e.w));
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool v = (valid[s] >> lane) & 1;
            const R sv = e.x * X[s] + e.y * Y[s] + e.z * Z[s] + e.w * W[s];
            R m;
            if (FILTER)
                m = R(kFilterKappa * (exyz * MXYZ[s] + ew * MW[s]));
            else
                m = R(sizeof(R) == 8 ? 0x1.0p-33 : 0x1.0p-16) *
                    (fabs(e.x * X[s]) + fabs(e.y * Y[s]) + fabs(e.z * Z[s]) + fabs(e.w * W[s]));
            bool k = !(sv > m);
            if constexpr (kInline64) k = k || ((x64[s] >> lane) & 1);
            keep |= v && k;
        }
        return ballot(keep) == 0;
    }

    // One clip by plane e (wave-uniform, test precision).  Lane wl holds the
    // same plane in the storage precision in eP (wl < 0: every lane does).
    // Returns kNoClip, kClipped or a failure status.  R2: keep r2 and maxr2
    // up to date.
    template <bool R2 = true>
    __device__ int clip(const Pl<R> &e, int wl, const Pl<PS> &eP) {
        uint64_t rem[S], unsure[S];
        int nr = 0, nvalid = 0;
        uint64_t flag = 0;
        const float exyz = fabsf(float(e.x)) + fabsf(float(e.y)) + fabsf(float(e.z)), ew = fabsf(float(e.w));
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool v = (valid[s] >> lane) & 1;
            const R sv = e.x * X[s] + e.y * Y[s] + e.z * Z[s] + e.w * W[s];
            rem[s] = ballot(v && sv > R(0));
            unsure[s] = 0;
            if (FILTER) {
                const float bound = kFilterKappa * (exyz * MXYZ[s] + ew * MW[s]);
                unsure[s] = ballot(v && fabsf(float(sv)) <= bound);
                if constexpr (kInline64) unsure[s] |= x64[s] & valid[s];
                flag |= unsure[s];
            }
            nr += __popcll(rem[s]);
            nvalid += __popcll(valid[s]);
        }
        if (FILTER && flag) {
            if constexpr (sizeof(PS) == 8) {
                // FP64 planes at hand (the finish kernel): the vertices the
                // filter leaves open are rebuilt from their planes in FP64
                // and tested in FP64, the FP64 form's own evaluation, so the
                // clip decides as the FP64 form does
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!unsure[s]) continue;
                    const bool out = ((unsure[s] >> lane) & 1) && outside64(planes, tri[s], eP);
                    const uint64_t now = ballot(out);
                    nr += __popcll(now) - __popcll(rem[s] & unsure[s]);
                    rem[s] = (rem[s] & ~unsure[s]) | now;
                }
            } else {
                return RPD_NEEDS_EXACT_PREDICATES;
            }
        }
        if (nr == 0) return kNoClip;
        if (nr == nvalid) return RPD_EMPTY_CELL;
        const int pid = nb_plane;
        if (lane == (wl < 0 ? 0 : wl)) planes[pid] = eP;
        // boundary edges: an edge u->v of a removed triangle whose reverse
        // v->u belongs to no removed triangle.  Every removed triangle marks
        // its three directed edges in the bit matrix em (row u, bit v), reads
        // the three reverse edges back, then clears the words it set.
        uint32_t bm[S];
        const uint32_t rw = uint32_t(P + 31) >> 5;  // words per row
        // edge u -> v: word u * rw + v / 32, bit v % 32
        auto edge_word = [&](uint32_t u, uint32_t v) { return em + (u * rw + (v >> 5)); };
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if ((rem[s] >> lane) & 1) {
                const uint32_t tr = tri[s];
                const uint32_t a0 = t0(tr), a1 = t1(tr), a2 = t2(tr);
                atomicOr(edge_word(a0, a1), 1u << (a1 & 31));
                atomicOr(edge_word(a1, a2), 1u << (a2 & 31));
                atomicOr(edge_word(a2, a0), 1u << (a0 & 31));
            }
        }
        wsync();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            bm[s] = 0;
            if ((rem[s] >> lane) & 1) {
                const uint32_t tr = tri[s];
                const uint32_t a0 = t0(tr), a1 = t1(tr), a2 = t2(tr);
                const bool h01 = (*edge_word(a1, a0) >> (a0 & 31)) & 1u;  // a1 -> a0 removed too
                const bool h12 = (*edge_word(a2, a1) >> (a1 & 31)) & 1u;
                const bool h20 = (*edge_word(a0, a2) >> (a2 & 31)) & 1u;
                bm[s] = (h01 ? 0u : 1u) | (h12 ? 0u : 2u) | (h20 ? 0u : 4u);
            }
        }
        wsync();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if ((rem[s] >> lane) & 1) {
                const uint32_t tr = tri[s];
                const uint32_t a0 = t0(tr), a1 = t1(tr), a2 = t2(tr);
                *edge_word(a0, a1) = 0;
                *edge_word(a1, a2) = 0;
                *edge_word(a2, a0) = 0;
            }
        }
        // slot-major exclusive offsets of the new vertices
        int nb = 0;
        int obase[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = __popc(bm[s]);
            const int incl = wave_incl_scan_op(c, 0, [](int x, int y) { return x + y; });
            obase[s] = nb + incl - c;
            nb += uni(rlane(incl, 63));
        }
        int nkeep = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) nkeep += __popcll(valid[s] & ~rem[s]);
        if (nkeep + nb > V) return RPD_VERTEX_OVERFLOW;
        // stage new triples (plane, a, b) and (unfiltered forms) the
        // boundary successor map
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (bm[s]) {
                const uint32_t tr = tri[s];
                const uint32_t v0 = t0(tr), v1 = t1(tr), v2 = t2(tr);
                int o = obase[s];
                if (bm[s] & 1u) {
                    nvs[o++] = pack3(uint32_t(pid), v0, v1);
                    if (!FILTER) nxt[v0] = uint8_t(v1);
                }
                if (bm[s] & 2u) {
                    nvs[o++] = pack3(uint32_t(pid), v1, v2);
                    if (!FILTER) nxt[v1] = uint8_t(v2);
                }
                if (bm[s] & 4u) {
                    nvs[o++] = pack3(uint32_t(pid), v2, v0);
                    if (!FILTER) nxt[v2] = uint8_t(v0);
                }
            }
        }
        wsync();
        // The cavity must be a disk: its boundary one simple cycle.  With
        // certified signs (FILTER) it always is: the removed vertices are
        // those of a convex polytope strictly on one side of a plane, which,
        // like the kept ones, form a connected set, and every face meets the
        // plane's side in one contiguous arc, so no face is pinched.  The
        // unfiltered forms check it, as CCComputeBoundary does: every start
        // plane written once, and the walk from one start closes after
        // exactly nb steps.
        bool inconsistent = false;
        if (!FILTER) {
            bool bad = false;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (bm[s]) {
                    const uint32_t tr = tri[s];
                    const uint32_t v0 = t0(tr), v1 = t1(tr), v2 = t2(tr);
                    if (bm[s] & 1u) bad |= nxt[v0] != v1;
                    if (bm[s] & 2u) bad |= nxt[v1] != v2;
                    if (bm[s] & 4u) bad |= nxt[v2] != v0;
                }
            }
            inconsistent = ballot(bad) != 0;
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
                    const uint32_t tr = tri[s];
                    if (bm[s] & 1u) nxt[t0(tr)] = 0xff;
                    if (bm[s] & 2u) nxt[t1(tr)] = 0xff;
                    if (bm[s] & 4u) nxt[t2(tr)] = 0xff;
                }
            }
        }
        if (inconsistent) return RPD_INCONSISTENT_BOUNDARY;
        // new vertices fill the lowest free slots
        int carry = 0;
        uint64_t ill = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            valid[s] &= ~rem[s];
            const uint64_t fr = ~valid[s];
            const int q = carry + mbcnt(fr);
            const bool take = ((fr >> lane) & 1) && q < nb;
            const uint64_t newm = ballot(take);
            if (take) {
                const uint32_t tr = nvs[q];
                set_vertex<R2>(s, tr, e, cvt<R>(planes[t1(tr)]), cvt<R>(planes[t2(tr)]));
            }
            if (FILTER) {
                const uint64_t il = ballot(take && fabsf(float(W[s])) < kFilterMinW * MW[s]);
                if constexpr (kInline64)
                    x64[s] = (x64[s] & ~newm) | il;  // decided in FP64 from now on
                else
                    ill |= il;
            }
            carry += __popcll(fr);
            valid[s] |= newm;
        }
        nb_plane = pid + 1;
        if (FILTER && ill) return RPD_NEEDS_EXACT_PREDICATES;
        if (R2) refresh_max_r2();
        return kClipped;
    }
};

// LDS of one wave: planes[P] | plane sources src[P] (int32) | nvs[64S] |
// nxt[256] | em[P][ceil(P / 32)] | tu[T] (finish kernel) | chunk / cone planes cpl[64] |
// go_out flags[16] | triangle corners tcor[16][9] (finish kernel)
template <typename R, typename RP, bool FILTER, int S, typename PS>
__device__ __forceinline__ void bind_lds(char *L, int P, int T, WaveCell<R, RP, FILTER, S, PS> &c,
                                         int **src = nullptr, int **tu = nullptr, Pl<PS> **cpl = nullptr,
                                         int **goo = nullptr, double **tcor = nullptr) {
    c.planes = reinterpret_cast<Pl<PS> *>(L);
    size_t off = (size_t(P) * sizeof(Pl<PS>) + 15) & ~size_t(15);
    if (src) *src = reinterpret_cast<int *>(L + off);
    off += (4 * size_t(P) + 15) & ~size_t(15);
    c.nvs = reinterpret_cast<uint32_t *>(L + off);
    off += 4 * 64 * S;
    c.nxt = reinterpret_cast<uint8_t *>(L + off);
    off += 256;
    c.em = reinterpret_cast<uint32_t *>(L + off);
    off += 4 * size_t(P) * size_t((P + 31) >> 5);
    if (tu) *tu = reinterpret_cast<int *>(L + off);
    off += 4 * size_t(T);
    off = (off + 15) & ~size_t(15);
    if (cpl) *cpl = reinterpret_cast<Pl<PS> *>(L + off);
    off += 64 * sizeof(Pl<PS>);
    if (goo) *goo = reinterpret_cast<int *>(L + off);
    off += 64;
    if (tcor) *tcor = reinterpret_cast<double *>(L + off);
}

// A vertex of the finish kernel's cell as the lane-parallel skip tests read
// it: homogeneous coordinates in the test precision and, for the certified
// form, the magnitudes of their terms.
template <typename R>
struct VtxRec {
    R X, Y, Z, W;
    float MXYZ, MW;
};
constexpr size_t kVtxRecBytes = 40;  // sizeof(VtxRec<double>), the larger one

// LDS bytes of one wave (pl: bytes of one stored plane; corners: the finish
// kernel's triangle corners and vertex copy)
__host__ __device__ constexpr size_t finish_wave_lds(int P, int S, int T, size_t pl, bool corners = false) {
    return ((((((P * pl + 15) & ~size_t(15)) + ((4 * size_t(P) + 15) & ~size_t(15)) + 4 * 64 * size_t(S) +
               256 + 4 * size_t(P) * size_t((P + 31) >> 5) + 4 * size_t(T)) +
              15) & ~size_t(15)) + 64 * pl + 64 + (corners ? 16 * 9 * 8 + 64 * size_t(S) * kVtxRecBytes : 0) +
            15) & ~size_t(15);
}
// LDS the Hessian adds to a finish-kernel wave: per plane 4 sums, per vertex
// slot 4 values and a plane id
__host__ __device__ constexpr size_t hessian_wave_lds(int P, int S) {
    return 32 * size_t(P) + 64 * size_t(S) * (32 + 4);
}

// Power bisector of the seed S0 and neighbour B in the seed-local frame
// (CCNewPlane, ConvexCell.cl:314-333): unit normal -b/|b|, offset
// d = (|b|^2 + w_B - w_s) / (2|b|), b = B - S0; a border neighbour
// (w <= NO_CELL_LIMITS) gives the plane through B.  Evaluated without
// contraction so that every evaluation of one plane has the same bits.
__device__ __forceinline__ Pl<double> bisector_plane(const double4 &S0, const double4 &B, double &bl) {
#pragma clang fp contract(off)
    const double bx = B.x - S0.x, by = B.y - S0.y, bz = B.z - S0.z;
    const double b2 = bx * bx + by * by + bz * bz;
    bl = sqrt(b2);
    const double ib = 1.0 / bl;
    const double d = B.w <= RPD_NO_CELL_LIMITS ? bl : (b2 + (B.w - S0.w)) * (0.5 * ib);
    return Pl<double>{-bx * ib, -by * ib, -bz * ib, d};
}

// Face f of the bounding cube [-eps, 1000 + eps]^3 in the seed-local frame
// (CCInit, ConvexCell.cl:278-297): x >= -eps, x <= 1000 + eps, then y, z.
__device__ __forceinline__ Pl<double> cube_plane(int f, const double4 &S0, double eps) {
    const int ax = f >> 1;
    const double sc = ax == 0 ? S0.x : (ax == 1 ? S0.y : S0.z);
    const double sgn = (f & 1) ? -1.0 : 1.0;
    const double d = (f & 1) ? (1000.0 + eps) - sc : sc + eps;
    return Pl<double>{ax == 0 ? sgn : 0.0, ax == 1 ? sgn : 0.0, ax == 2 ? sgn : 0.0, d};
}

// Record of a finished cell in the queue: header (sid, nb_plane, nvalid,
// kused), planes[P] in the storage precision, vertex triples[64 * S].
struct RecHeader {
    int sid, nb_plane, nvalid, kused;
};

// ---------------------------------------------------------------------------
// Kernel A: the bisector loop (ConvexCell.cl:850-917)
// ---------------------------------------------------------------------------
// Minimum waves per SIMD the bisector kernel is compiled for (one vertex
// slot per lane): its run time is set by latency, so occupancy pays more than
// the few registers the compiler then spills (measured at 1M seeds: FP64 7,
// 72 VGPRs and 12 bytes of spill, 2.10 ms against 2.35 at 5 waves and 2.27 at
// 8; FP32 forms 8, which they reach without spills).  Wider cells (S > 1,
// re-queue passes) are left unconstrained.
template <typename R, int S>
constexpr int kClipWaves = S > 1 ? 1 : (sizeof(R) == 8 ? 7 : 8);

template <typename R, typename RP, bool FILTER, int S, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kClipWaves<R, S>)))
void clip_cells_kernel(ClipArgs a) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int lane = lane_id();
    const int wv = uni(int(threadIdx.x >> 6));  // wave-uniform: lets ids, seeds and pointers live in SGPRs
    const int wpb = blockDim.x >> 6;
    const int64_t t = int64_t(blockIdx.x) * wpb + wv;
    if (t >= a.n_ids) return;  // wave-uniform

    // planes in the test precision; src[p] records where plane p came from
    // (the neighbour's stored id, or -1 - f for cube face f), so that the
    // record of a queued cell carries the planes in the geometry precision
    // and each plane's neighbour (CCPlaneNeighbour, ConvexCell.cl:329-331)
    WaveCell<R, RP, FILTER, S, R> c;
    Pl<R> *chunk = nullptr;  // planes of the current 64 neighbours
    int *src = nullptr;
    bind_lds(lds_raw + size_t(wv) * a.lds_stride, a.P, 0, c, &src, nullptr, &chunk);
    c.P = a.P;
    c.V = a.V;
    c.lane = lane;

    const int sid = uni(a.ids ? a.ids[t] : int(t));
    if constexpr (BATCH) bind_request(a, sid);
    const double4 S0 = a.seeds[sid];
    const int orig = a.perm[sid];
    const int K = a.K, P = a.P;
    // resumed cells: record of the state, neighbours used before this pass
    const bool resume = sizeof(R) == 8 && a.rs_in != nullptr;
    const char *rs_rec = resume ? a.rs_in + size_t(a.rslot[sid]) * size_t(a.rs_in_stride) : nullptr;
    const int kbase = resume ? uni(reinterpret_cast<const RecHeader *>(rs_rec)->kused) : 0;

    auto finish = [&](int st, double m0, double m1, double m2, double m3, int ku, bool resumable = false) {
        if (lane == 0) {
            a.out[orig] = make_double4(m0, m1, m2, m3);
            a.status[orig] = st;
            if (a.kused) a.kused[orig] = kbase + ku;
            note_failure(a, sid, st, resumable);
        }
    };
    auto fail = [&](int st, int ku) { finish(st, 0.0, 0.0, 0.0, 0.0, ku); };

    if (S0.w <= RPD_NO_CELL_LIMITS) {  // border seed: copied (ConvexCell.cl:873-876)
        finish(RPD_SUCCESS, S0.x, S0.y, S0.z, S0.w, 0);
        return;
    }
    for (int i = lane; i < 256; i += 64) c.nxt[i] = 0xff;
    for (int i = lane; i < P * ((P + 31) >> 5); i += 64) c.em[i] = 0;

    // Domain cull: a seed whose voxel is d >= 2 voxels (Chebyshev) from every
    // voxel that holds triangles or lies inside has no point of the domain
    // within (d - 1) * vs; once the cell's circumradius about the seed is
    // below that, its restriction to the domain is empty (volume 0, as the
    // reference's apex-outside / no-triangle case, ConvexCell.cl:1067-1079).
    double cull_r2 = 0;
    if (a.g > 0 && a.dist) {
        const int g = a.g;
        auto vox = [&](double x) { return int(fmin(fmax(x / a.vs, 0.0), double(g - 1))); };
        const int d = a.dist[vox(S0.x) + vox(S0.y) * g + vox(S0.z) * g * g];
        if (d >= 2) cull_r2 = double(d - 1) * a.vs * double(d - 1) * a.vs;
    }
    const double wgap = S0.w - a.wmin;  // >= 0 except for the lightest seeds

    const double eps = a.g > 0 ? 0.1 : 0.0;  // CUBE_EPSILON, voronoi.cpp:929
#pragma unroll
    for (int s = 0; s < S; ++s) {
        c.valid[s] = 0;
        c.x64[s] = 0;
        c.tri[s] = 0;
        c.X[s] = c.Y[s] = c.Z[s] = R(0);
        c.W[s] = R(1);
        c.r2[s] = R(0);
        c.MXYZ[s] = c.MW[s] = 0.f;
    }
    if (resume) {
        // ---- the state a previous pass left: planes, then the vertices at
        // their lanes (nent = 64 x the S of the pass that left it) ----------
        const RecHeader hd = *reinterpret_cast<const RecHeader *>(rs_rec);
        const int nent = uni(hd.nvalid);
        c.nb_plane = uni(hd.nb_plane);
        const Pl<RP> *rpl = reinterpret_cast<const Pl<RP> *>(rs_rec + sizeof(RecHeader));
        const uint32_t *rtr =
            reinterpret_cast<const uint32_t *>(rs_rec + sizeof(RecHeader) + size_t(a.rs_in_P) * sizeof(Pl<RP>));
        const int *rnid = reinterpret_cast<const int *>(rtr + nent);
        for (int p = lane; p < c.nb_plane; p += 64) {
            c.planes[p] = cvt<R>(rpl[p]);
            src[p] = rnid[p];
        }
        wsync();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int j = s * 64 + lane;
            const uint32_t tr = j < nent ? rtr[j] : kNoTri;
            if (tr != kNoTri) c.set_vertex_lds(s, tr);
            c.valid[s] = ballot(tr != kNoTri);
        }
    } else {
        // ---- CCInit: bounding cube in the local frame ----------------------
        if (lane < 6) {
            c.planes[lane] = cvt<R>(cube_plane(lane, S0, eps));
            src[lane] = -1 - lane;
        }
        wsync();
        c.nb_plane = 6;
        if (lane < 8) {
            const uint32_t sh = 8 * uint32_t(lane);
            c.set_vertex_lds(0, pack3(uint32_t(kCube0 >> sh) & 0xff, uint32_t(kCube1 >> sh) & 0xff,
                                      uint32_t(kCube2 >> sh) & 0xff));
        }
        c.valid[0] = 0xffull;
    }
    c.refresh_max_r2();

    // ---- neighbour loop -----------------------------------------------------
    // The 64 neighbours of a chunk are loaded one per lane: the plane goes to
    // LDS, the distance |b| and plane offset t (signed distance of the
    // bisector from the seed) stay in registers.  A plane with t >= R, R the distance from the seed to the
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
    double cdl = 0;
    float ct = 0;
    int cnb = -1;
    int loaded = -1;
    const double rmarg = sizeof(R) == 8 ? 1.0 + 1e-9 : 1.0 + 2e-4;  // FP32 R may be low by ~1e-4
    // the cell's radius (FP32 forms: the hardware square root, 1 ulp, far
    // inside the radius margin)
    auto radius = [&]() -> double {
        if constexpr (sizeof(R) == 4)
            return double(__builtin_amdgcn_sqrtf(c.maxr2));
        else
            return sqrt(double(c.maxr2));
    };
    double Rcur = radius();
    auto stop_at = [&](double dl, double Rm) -> bool {
        // CCIsSecurityRadiusReached (reference stop, weights ignored) or the
        // exact power stop d (d - 2R) >= w_s - w_min
        const double Rs = Rm * rmarg;
        if (a.security_mode == 1 || wgap <= 0) return dl * dl > 4.0 * Rs * Rs;
        return dl * (dl - 2.0 * Rs) >= wgap;
    };
    // The exact stop only ever holds when no later bisector can cut, so the
    // FP32 radius's margin merely delays it.  The reference stop can hold
    // while a later bisector of a lighter seed still cuts; there a delayed
    // stop changes the cell.  So in the mixed form a reference-stop test
    // within the FP32 radius's margin (2R / rmarg < d <= 2R rmarg) that the
    // loop reaches ends the cell with needs_exact_predicates: the FP64 pass
    // decides it, as an uncertain plane-side test.
    const bool ref_stop = FILTER && a.security_mode == 1;
    auto stop_near = [&](double dl, double Rm) -> bool {
        const double Rl = Rm / rmarg;
        return dl * dl > 4.0 * Rl * Rl && !stop_at(dl, Rm);
    };
    int i = 0;
    while (i < K) {
        const int base = i & ~63;
        if (base != loaded) {
            const int j = base + lane;
            cnb = j < K ? nrow[j] : -2;  // -1: list exhausted, -2: past K
            Pl<R> cpl{R(0), R(0), R(0), R(0)};
            cdl = 0;
            ct = 0;
            if (cnb >= 0) {
                double bl;
                const Pl<double> q = bisector_plane(S0, a.seeds[cnb], bl);
                cpl = cvt<R>(q);
                cdl = bl;
                ct = float(q.w);
            }
            wsync();  // the previous chunk's planes are no longer read
            chunk[lane] = cpl;
            wsync();
            loaded = base;
        }
        const int l0 = i & 63;
        const bool in = lane >= l0 && cnb != -2;
        if (c.nb_plane >= P) {
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
        // skip test in FP32: t is rounded to FP32, so the radius gets a margin
        // well above FP32 rounding (a false candidate only costs a no-op clip)
        const float Rskip = float(Rcur * rmarg) * (1.0f + 1e-6f);
        const uint64_t m_end = ballot(in && cnb == -1);
        const uint64_t m_clip = ballot(in && cnb >= 0 && ct < Rskip);
        const uint64_t m_stop = ballot(in && cnb >= 0 && stop_at(cdl, Rcur));
        const int f_end = m_end ? lowest_bit(m_end) : 64;
        const int f_clip = m_clip ? lowest_bit(m_clip) : 64;
        const int f_stop = m_stop ? lowest_bit(m_stop) : 64;
        if (ref_stop) {
            const uint64_t m_near = ballot(in && cnb >= 0 && stop_near(cdl, Rcur));
            if (m_near && lowest_bit(m_near) < min(f_end, min(f_clip, f_stop))) {
                fail(RPD_NEEDS_EXACT_PREDICATES, base + lowest_bit(m_near) + 1);
                return;
            }
        }
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
        const Pl<R> e = chunk[f_clip];
        const int nbid = rlane(cnb, f_clip);
        if (lane == 0) src[c.nb_plane] = nbid;
        const int res = c.clip(e, -1, e);
        kused = base + f_clip + 1;
        if (res >= 0) {
            fail(res, kused);
            return;
        }
        if (res == kClipped) {
            Rcur = radius();
            if (double(c.maxr2) * 1.001 < cull_r2) {
                finish(RPD_SUCCESS, 0.0, 0.0, 0.0, 0.0, kused);
                return;
            }
        }
        {
            const double dl = rlane(cdl, f_clip);
            if (ref_stop && stop_near(dl, Rcur)) {
                fail(RPD_NEEDS_EXACT_PREDICATES, kused);
                return;
            }
            if (stop_at(dl, Rcur)) {
                reached = true;
                break;
            }
        }
        i = base + f_clip + 1;
    }
    // a list that holds every other seed leaves nothing to test: the cell is
    // exact once the list is used up (the reference reads past the seeds there)
    if (!reached && !complete && int64_t(kbase) + K >= a.n_seeds - 1) complete = true;

    // debug counters (ConvexCell.cl:887-911): K used where the security
    // radius was reached, P and V once the neighbour loop is done
    if (a.dcnt && lane == 0) {
        if (reached) {
            const int ku = min(kbase + kused, 256);
            atomicMax(a.dcnt + kDcMax + 0, ku);
            atomicAdd(a.dcnt + kDcHistK + ku, 1);
        }
        int nv = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) nv += __popcll(c.valid[s]);
        atomicMax(a.dcnt + kDcMax + 1, c.nb_plane);
        atomicAdd(a.dcnt + kDcHistP + min(c.nb_plane, 255), 1);
        atomicMax(a.dcnt + kDcMax + 2, nv);
        atomicAdd(a.dcnt + kDcHistV + min(nv, 256), 1);
    }

    // A finished cell is queued for restriction + integration.  A cell whose
    // list ran out keeps its state when there is room, so that the next pass
    // goes on from here with the neighbours after the last one used.
    const bool ran_out = !reached && !complete;
    int slot = -1;
    if (lane == 0) {
        if (!ran_out) {
            slot = atomicAdd(a.qcount, 1);
        } else if (a.rs_out) {
            slot = atomicAdd(a.rs_count, 1);
            if (slot >= a.rs_cap) slot = -1;
        }
    }
    slot = rlane(slot, 0);
    if (slot >= 0) {
        // the cell's state as a record: header, planes in the geometry
        // precision (rebuilt from their sources when the test precision is
        // lower), vertex triples; P planes' room.  (Written in line: as a
        // lambda the FP32-test kernel spills.)
        char *rec = ran_out ? a.rs_out + size_t(slot) * size_t(a.rs_out_stride)
                            : a.rec + size_t(slot) * size_t(a.rec_stride);
        Pl<RP> *rpl = reinterpret_cast<Pl<RP> *>(rec + sizeof(RecHeader));
        uint32_t *rtr = reinterpret_cast<uint32_t *>(rec + sizeof(RecHeader) + size_t(P) * sizeof(Pl<RP>));
        int *rnid = reinterpret_cast<int *>(rtr + 64 * S);
        // a queued cell's vertices are compacted; a kept state holds each
        // triple at its lane (kNoTri where none), so that the resumed cell
        // has the lanes, and so the sums, of one uninterrupted run
        int nv = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) nv += __popcll(c.valid[s]);
        if (ran_out) nv = 64 * S;
        if (lane == 0) *reinterpret_cast<RecHeader *>(rec) = RecHeader{sid, c.nb_plane, nv, kbase + kused};
        wsync();
        for (int p = lane; p < c.nb_plane; p += 64) {
            if (sizeof(R) == sizeof(RP)) {
                rpl[p] = cvt<RP>(c.planes[p]);
            } else {
                const int sj = src[p];
                double bl;
                rpl[p] = cvt<RP>(sj < 0 ? cube_plane(-1 - sj, S0, eps) : bisector_plane(S0, a.seeds[sj], bl));
            }
            rnid[p] = src[p];
        }
        int carry = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool v = (c.valid[s] >> lane) & 1;
            if (ran_out)
                rtr[64 * s + lane] = v ? c.tri[s] : kNoTri;
            else if (v)
                rtr[carry + mbcnt(c.valid[s])] = c.tri[s];
            carry += __popcll(c.valid[s]);
        }
    }
    if (ran_out) {
        if (slot >= 0 && lane == 0) {
            a.rslot[sid] = slot;
            a.lastnb[sid] = nrow[K - 1];
        }
        finish(RPD_SECURITY_RADIUS_NOT_REACHED, 0.0, 0.0, 0.0, 0.0, kused, slot >= 0);
    }
}

// tinfo[q] of a triangle-list entry (voxel v): x is v packed as
// z<<20 | y<<10 | x; (z << 32 | y) a 64-bit mask with bit
// (dz+2)*25 + (dy+2)*5 + (dx+2) set when the voxel v + (dx, dy, dz), within 2
// voxels of v per axis, precedes v in scan order and its list holds the same
// triangle (such offsets are the ones below the centre, bit 62, of the
// 5x5x5 block, so the bits are 0..61); kTinfoFar when a preceding voxel
// whose list holds it lies farther away.
constexpr uint64_t kTinfoFar = 1ull << 63;

// Mask of the preceding offsets of the 5x5x5 block (bits as in tinfo) that
// keep voxel (x, y, z) inside the box [mn, mx].
__device__ __forceinline__ uint64_t box_near(int x, int y, int z, const int *mn, const int *mx) {
    // per axis: the offsets d in [-2, 2] with mn <= v + d <= mx, bit d + 2
    auto axis = [](int v, int lo, int hi) -> uint32_t {
        const int a = max(lo - v, -2), b = min(hi - v, 2);
        return a > b ? 0u : ((1u << (b - a + 1)) - 1u) << (a + 2);
    };
    const uint32_t ax = axis(x, mn[0], mx[0]), ay = axis(y, mn[1], mx[1]), az = axis(z, mn[2], mx[2]);
    // the (dy, dx) plane: bit 5 (dy + 2) + dx + 2, i.e. ax repeated at 5 bit k
    // of ay (no carries: the copies are 5 bits apart)
    const uint32_t ry = (ay & 1u) | ((ay & 2u) << 4) | ((ay & 4u) << 8) | ((ay & 8u) << 12) | ((ay & 16u) << 16);
    const uint64_t pyx = uint64_t(ax) * uint64_t(ry);
    // planes dz = -2, -1, 0 at bits 0, 25, 50 (dz > 0 never precedes); bits
    // past 63 of the dz = 0 plane are offsets after the centre, never set
    const uint64_t rz = ((az & 1u) ? 1ull : 0ull) | ((az & 2u) ? 1ull << 25 : 0ull) | ((az & 4u) ? 1ull << 50 : 0ull);
    return pyx * rz;
}

// ---------------------------------------------------------------------------
// Kernel B: restriction to the domain (ConvexCell.cl:918-1116) and the
// integration (CCExportBaryVolume, ConvexCell.cl:577-619) of queued cells.
//
// W waves share one cell (W = 1 for large queues, up to 8 for the short
// re-queue passes whose run time is set by their heaviest cell): every wave
// rebuilds the cell and its triangle list, wave w takes the triangles
// k = w, w + W, ...  Each triangle starts from the same backed-up cell, so its
// contribution does not depend on which wave computes it.  The contributions
// are summed in increasing k, one after another: with W = 1 as they come (a
// wave-uniform running sum), with W > 1 from LDS, where the waves put them by
// triangle index, by wave 0.  Both make the same additions, so the result of
// a cell is the same bit for bit for every W and every batch it is in.
//
// Cone planes are built lane-parallel, 16 triangles x 4 planes per batch.  A
// triangle one of whose four planes removes every vertex of the cell with a
// margin contributes nothing: the reference clips it and ends with
// empty_cell at that plane at the latest (vertices created by earlier cone
// planes lie in the convex hull of the removed ones).  The shortcut is taken
// only when the caps guarantee that those earlier clips could not overflow
// either (at most 3 more planes, at most 2 more vertices per clip), so the
// outcome is the reference's.
// ---------------------------------------------------------------------------
template <typename R, typename RP, bool FILTER, int S, bool HESS, bool BATCH>
// 4 waves per SIMD for S = 1 (128 VGPRs): measured at 1M seeds against 5 and
// 6, whose spills cost more than the extra residency gives.  HESS: the
// Hessian variant (one wave per cell).
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(S == 1 && !HESS ? 4 : 1)))
void finish_cells_kernel(ClipArgs a, int nq, int W) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    const int lane = lane_id();
    const int wv = uni(int(threadIdx.x >> 6));  // wave-uniform: lets ids, seeds and pointers live in SGPRs
    const int wpb = blockDim.x >> 6;
    const int cpb = wpb / W;  // cells per block; W > 1 implies cpb == 1
    const int ci = wv / W, sub = wv - ci * W;
    const int item = int(blockIdx.x) * cpb + ci;
    if (item >= nq || (a.nq_dev && item >= uni(*a.nq_dev))) return;  // uniform over the waves of a cell
    const int slot = a.slots ? uni(a.slots[item]) : item;

    char *cell_l = lds_raw + size_t(ci) * a.cell_stride;
    int *fails = reinterpret_cast<int *>(cell_l);                  // [W][2]
    double4 *contrib = reinterpret_cast<double4 *>(cell_l + 64);  // [T], W > 1 only
    WaveCell<R, RP, FILTER, S> c;
    int *tu = nullptr, *goo = nullptr;
    Pl<RP> *cpl = nullptr;
    double *tcor = nullptr;
    bind_lds(lds_raw + size_t(cpb) * a.cell_stride + size_t(wv) * a.lds_stride, a.P, a.T, c, nullptr, &tu,
             &cpl, &goo, &tcor);
    c.P = a.P;
    c.V = a.V;
    c.lane = lane;
    const int P = a.P;

    const char *rec = a.rec + size_t(slot) * size_t(a.rec_stride);
    const RecHeader hd = *reinterpret_cast<const RecHeader *>(rec);
    const int sid = uni(hd.sid);
    if constexpr (BATCH) bind_request(a, sid);
    const int kused = uni(hd.kused);
    const int nvalid = uni(hd.nvalid);
    c.nb_plane = uni(hd.nb_plane);
    const Pl<RP> *rpl = reinterpret_cast<const Pl<RP> *>(rec + sizeof(RecHeader));
    const uint32_t *rtr =
        reinterpret_cast<const uint32_t *>(rec + sizeof(RecHeader) + size_t(a.rec_P) * sizeof(Pl<RP>));
    for (int p = lane; p < c.nb_plane; p += 64) c.planes[p] = rpl[p];
    for (int i = lane; i < 256; i += 64) c.nxt[i] = 0xff;
    for (int i = lane; i < P * ((P + 31) >> 5); i += 64) c.em[i] = 0;
    wsync();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int j = s * 64 + lane;
        c.tri[s] = 0;
        c.X[s] = c.Y[s] = c.Z[s] = R(0);
        c.W[s] = R(1);
        c.r2[s] = R(0);
        c.MXYZ[s] = c.MW[s] = 0.f;
        if (j < nvalid) c.template set_vertex_lds<false>(s, rtr[j]);
        c.valid[s] = ballot(j < nvalid);
        c.x64[s] = 0;
    }
    // (the vertices' distances from the seed are not kept here: the
    // restriction takes the cell's radius in FP64 below)

    const double4 S0 = a.seeds[sid];
    const int orig = a.perm[sid];
    auto finish = [&](int st, double m0, double m1, double m2, double m3) {
        if (sub == 0 && lane == 0) {
            a.out[orig] = make_double4(m0, m1, m2, m3);
            a.status[orig] = st;
            if (a.kused) a.kused[orig] = kused;
            note_failure(a, sid, st, false, slot);
        }
    };
    auto fail = [&](int st) { finish(st, 0.0, 0.0, 0.0, 0.0); };

    // vertex coordinates in the geometry precision RP: the registers when
    // R == RP, else recomputed from the stored RP planes
    auto vertex_rp = [&](int s, RP &Ax, RP &Ay, RP &Az) {
        if (sizeof(R) == sizeof(RP)) {
            const RP iw = RP(1) / RP(c.W[s]);
            Ax = RP(c.X[s]) * iw;
            Ay = RP(c.Y[s]) * iw;
            Az = RP(c.Z[s]) * iw;
        } else {
            const uint32_t tr = c.tri[s];
            RP X, Y, Z, W;
            float d0, d1;
            vertex_from_planes<RP, false>(c.planes[t0(tr)], c.planes[t1(tr)], c.planes[t2(tr)], X, Y, Z, W, d0, d1);
            const RP iw = RP(1) / W;
            Ax = X * iw;
            Ay = Y * iw;
            Az = Z * iw;
        }
    };

    // ---- integration (CCExportBaryVolume): factor/6 * (volume, moments) of
    // the current cell in the local frame, FP64 wave sums --------------------
    const RP vol_eps2 = sizeof(RP) == 8 ? RP(1e-12) : RP(1e-6);  // voronoi.cpp:919-922
    // Hessian sums: per record plane 4 doubles (area-weighted centroid in the
    // local frame, area), per vertex slot staging for their reduction
    double *hacc = nullptr, *hval = nullptr;
    int *hpid = nullptr;
    const int nbk = c.nb_plane;  // the record's planes; cone planes come after
    if (HESS) {
        hacc = reinterpret_cast<double *>(reinterpret_cast<char *>(tcor + 16 * 9) + 64 * size_t(S) * kVtxRecBytes);
        hval = hacc + 4 * size_t(P);
        hpid = reinterpret_cast<int *>(hval + 4 * 64);
        for (int i = lane; i < 4 * nbk; i += 64) hacc[i] = 0.0;
        wsync();
    }
    auto integrate = [&](double factor) -> double4 {
        RP sw = 0, sx = 0, sy = 0, sz = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool vok = (c.valid[s] >> lane) & 1;
            // per plane k of the vertex: area (CCExportBaryVolumeAndArea,
            // ConvexCell.cl:621-674) and area-weighted centroid
            double ha[3] = {0.0, 0.0, 0.0}, hx[3] = {0.0, 0.0, 0.0}, hy[3] = {0.0, 0.0, 0.0},
                   hz[3] = {0.0, 0.0, 0.0};
            uint32_t tr = c.tri[s];
            if (!HESS && !vok) continue;
            if (vok) {
                RP Ax, Ay, Az;
                vertex_rp(s, Ax, Ay, Az);
                const RP ACx = -Ax, ACy = -Ay, ACz = -Az;  // seed at the origin
                const Pl<RP> pl[3] = {c.planes[t0(tr)], c.planes[t1(tr)], c.planes[t2(tr)]};
                RP opx[3], opy[3], opz[3], olx[3], oly[3], olz[3];
    #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const RP nn = pl[k].x * pl[k].x + pl[k].y * pl[k].y + pl[k].z * pl[k].z;
                    const RP f = (ACx * pl[k].x + ACy * pl[k].y + ACz * pl[k].z) / nn;
                    opx[k] = ACx - f * pl[k].x;
                    opy[k] = ACy - f * pl[k].y;
                    opz[k] = ACz - f * pl[k].z;
                }
    #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const Pl<RP> &p = pl[k], &q = pl[(k + 1) % 3];
                    const RP cx = p.y * q.z - p.z * q.y, cy = p.z * q.x - p.x * q.z, cz = p.x * q.y - p.y * q.x;
                    const RP c2 = cx * cx + cy * cy + cz * cz;
                    const RP pp = p.x * p.x + p.y * p.y + p.z * p.z, qq = q.x * q.x + q.y * q.y + q.z * q.z;
                    if (c2 < vol_eps2 * pp * qq) {
                        olx[k] = oly[k] = olz[k] = RP(0);
                    } else {
                        const RP f = (ACx * cx + ACy * cy + ACz * cz) / c2;
                        olx[k] = f * cx;
                        oly[k] = f * cy;
                        olz[k] = f * cz;
                    }
                }
    #pragma unroll
                for (int k = 0; k < 3; ++k) {
    #pragma unroll
                    for (int st = 0; st < 2; ++st) {
                        const int j = (k + 3 - st) % 3;
                        // det3 of rows AC, onPlane_k, onLine_j (ConvexCell.cl:72-77)
                        const RP d = ACx * (opy[k] * olz[j] - opz[k] * oly[j]) - opx[k] * (ACy * olz[j] - ACz * oly[j]) +
                                     olx[j] * (ACy * opz[k] - ACz * opy[k]);
                        const RP w = st ? d : -d;
                        sw += w;
                        sx += w * (Ax + RP(0.25) * (ACx + opx[k] + olx[j]));
                        sy += w * (Ay + RP(0.25) * (ACy + opy[k] + oly[j]));
                        sz += w * (Az + RP(0.25) * (ACz + opz[k] + olz[j]));
                        if (HESS) {
                            // det3 of rows onPlane_k, onLine_j, n_k, halved
                            const double dt = double(opx[k]) * (double(oly[j]) * pl[k].z - double(olz[j]) * pl[k].y) -
                                              double(opy[k]) * (double(olx[j]) * pl[k].z - double(olz[j]) * pl[k].x) +
                                              double(opz[k]) * (double(olx[j]) * pl[k].y - double(oly[j]) * pl[k].x);
                            const double ar = (st ? 0.5 : -0.5) * factor * dt;
                            ha[k] += ar;
                            hx[k] += ar * (double(Ax) + (double(opx[k]) + double(olx[j])) / 3.0);
                            hy[k] += ar * (double(Ay) + (double(opy[k]) + double(oly[j])) / 3.0);
                            hz[k] += ar * (double(Az) + (double(opz[k]) + double(olz[j])) / 3.0);
                        }
                    }
                }
            }
            if (HESS) {
                // per plane, in a fixed order: plane slot k, then lane; the
                // lane that owns plane p (p = lane, lane + 64, ...) adds
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    hpid[lane] = vok ? int(k == 0 ? t0(tr) : (k == 1 ? t1(tr) : t2(tr))) : -1;
                    hval[4 * lane + 0] = hx[k];
                    hval[4 * lane + 1] = hy[k];
                    hval[4 * lane + 2] = hz[k];
                    hval[4 * lane + 3] = ha[k];
                    wsync();
                    for (int p = lane; p < nbk; p += 64) {
                        double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
                        for (int j = 0; j < 64; ++j) {
                            if (hpid[j] != p) continue;
                            e0 += hval[4 * j + 0];
                            e1 += hval[4 * j + 1];
                            e2 += hval[4 * j + 2];
                            e3 += hval[4 * j + 3];
                        }
                        hacc[4 * p + 0] += e0;
                        hacc[4 * p + 1] += e1;
                        hacc[4 * p + 2] += e2;
                        hacc[4 * p + 3] += e3;
                    }
                    wsync();
                }
            }
        }
        const double f = factor / 6.0;
        return make_double4(f * wave_sum(double(sx)), f * wave_sum(double(sy)), f * wave_sum(double(sz)),
                            f * wave_sum(double(sw)));
    };

    // ---- restriction to the domain grid --------------------------------------
    bool add_volume = true;
    int ntri = 0;
    double rcell = 0;  // largest distance of a vertex from the seed (FP64)
    double P0x = 0, P0y = 0, P0z = 0;  // apex, global frame
    if (a.g > 0) {
        const int g = a.g;
        const double vs = a.vs;
        double bmn[3] = {1000.0, 1000.0, 1000.0}, bmx[3] = {0.0, 0.0, 0.0}, vr2 = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!((c.valid[s] >> lane) & 1)) continue;
            RP lx, ly, lz;
            vertex_rp(s, lx, ly, lz);
            vr2 = fmax(vr2, double(lx) * double(lx) + double(ly) * double(ly) + double(lz) * double(lz));
            const double gx = S0.x + double(lx), gy = S0.y + double(ly), gz = S0.z + double(lz);
            bmn[0] = fmin(bmn[0], gx);
            bmn[1] = fmin(bmn[1], gy);
            bmn[2] = fmin(bmn[2], gz);
            bmx[0] = fmax(bmx[0], gx);
            bmx[1] = fmax(bmx[1], gy);
            bmx[2] = fmax(bmx[2], gz);
        }
        rcell = sqrt(wave_max(vr2));
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
        // First voxel of the box, in scan order (x fastest, then y, z),
        // whose flag is not 1 (the reference corner rule) or whose class is
        // known (apex_mode 1); usually the first voxel already.
        int first_ref = -1, first_c = -1, ref_flag = 1, c_flag = 1;
        for (int vb = 0; vb < nvox && (a.apex_mode == 1 ? first_c < 0 : first_ref < 0); vb += 64) {
            const int v = vb + lane;
            const bool in = v < nvox;
            int f = 1, cf = 1;
            if (in) {
                const int x = vmn[0] + v % nx, y = vmn[1] + (v / nx) % ny, z = vmn[2] + v / (nx * ny);
                const int id = x + y * g + z * g * g;
                f = a.flags[id];
                cf = a.centre ? a.centre[id] : (a.toffs[id + 1] == a.toffs[id] ? f : 1);
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
        }
        // Union of the voxels' lists (merge, ConvexCell.cl:722-749).  The
        // list entries of one row of the box are one range of tlist; the
        // rows' ranges are spread over the lanes.  An entry is kept in the
        // first voxel of the box, in scan order, whose list holds its
        // triangle: tinfo holds the entry's voxel and a mask of the voxels
        // within 2 of it per axis that precede it and hold the triangle too
        // (or kTinfoFar when one farther away does: then the triangle's own
        // voxel list is read).  Entries are kept in scan order.
        int nu = 0;
        const uint32_t pmn = (uint32_t(vmn[2]) << 20) | (uint32_t(vmn[1]) << 10) | uint32_t(vmn[0]);
        const uint32_t pmx = (uint32_t(vmx[2]) << 20) | (uint32_t(vmx[1]) << 10) | uint32_t(vmx[0]);
        int *seg_off = reinterpret_cast<int *>(cpl), *seg_a = seg_off + 64;
        const int nrows = ny * nz;
        for (int rb = 0; rb < nrows; rb += 64) {
            const int r = rb + lane;
            int sa = 0, len = 0;
            if (r < nrows) {
                const int id0 = vmn[0] + (vmn[1] + r % ny) * g + (vmn[2] + r / ny) * g * g;
                sa = a.toffs[id0];
                len = a.toffs[id0 + nx] - sa;
            }
            const int off = wave_excl_scan(len);
            const int total = uni(rlane(off + len, 63));
            if (total == 0) continue;
            seg_off[lane] = off;
            seg_a[lane] = sa;
            wsync();
            for (int base = 0; base < total; base += 64) {
                const int e = base + lane;
                bool keep = false;
                int t = -1;
                if (e < total) {
                    // last row whose offset is <= e (empty rows never win)
                    int lo = 0, hi = 63;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (seg_off[mid] <= e)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    const int q = seg_a[lo] + (e - seg_off[lo]);
                    t = a.tlist[q];
                    const uint4 inf = a.tinfo[q];
                    const uint32_t pv = inf.x;
                    const uint64_t near = (uint64_t(inf.z) << 32) | inf.y;
                    if (near & kTinfoFar) {
                        keep = true;
                        for (int u = a.tvoffs[t], ue = a.tvoffs[t + 1]; u < ue; ++u) {
                            const uint32_t w = a.tvox[u];
                            if (w >= pv) break;
                            const uint32_t ux = w & 1023u, uy = (w >> 10) & 1023u, uz = w >> 20;
                            if (ux >= (pmn & 1023u) && ux <= (pmx & 1023u) && uy >= ((pmn >> 10) & 1023u) &&
                                uy <= ((pmx >> 10) & 1023u) && uz >= (pmn >> 20) && uz <= (pmx >> 20)) {
                                keep = false;  // an earlier voxel of the box holds it
                                break;
                            }
                        }
                    } else {
                        const int x = int(pv & 1023u), y = int((pv >> 10) & 1023u), z = int(pv >> 20);
                        keep = (near & box_near(x, y, z, vmn, vmx)) == 0;
                    }
                }
                const uint64_t m = ballot(keep);
                if (keep && nu + mbcnt(m) < a.T) tu[nu + mbcnt(m)] = t;
                nu += __popcll(m);
            }
            wsync();  // the next rows overwrite seg_off / seg_a
        }
        if (nu > a.T) {
            // the cap a re-queue pass needs for this cell
            if (sub == 0 && lane == 0 && a.fl_cnt) atomicMax(a.fl_cnt + 13, nu);
            fail(RPD_TRIANGLE_OVERFLOW);
            return;
        }
        if (a.dcnt && sub == 0 && lane == 0) {  // ConvexCell.cl:981-982
            atomicMax(a.dcnt + kDcMax + 3, nu);
            atomicAdd(a.dcnt + kDcHistT + min(nu, kDcUnionMax), 1);
        }
        wsync();
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
            P0x = (x + kApexFrac[0]) * vs;
            P0y = (y + kApexFrac[1]) * vs;
            P0z = (z + kApexFrac[2]) * vs;
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
                fail(RPD_FIND_ANOTHER_BEGINNING_VERTEX);
                return;
            }
        }
        add_volume = in_domain == 2;
    }
    double4 cadd = make_double4(0.0, 0.0, 0.0, 0.0);
    if (add_volume && sub == 0) cadd = integrate(1.0);

    // The reference clips the triangles in increasing id order and stops at
    // the first failure: the cell's status is that of the smallest failing
    // id.  Triangles are visited here in list order, so after a failure only
    // triangles of smaller id are still clipped.
    int fail_t = 0x7fffffff, fail_st = 0;
    double sm = 0, sx = 0, sy = 0, sz = 0;  // running sums of the contributions (W = 1; wave-uniform)
    if (ntri > 0) {
        // backup (ConvexCell.cl:692-715): the vertex registers and the plane count
        const WaveCell<R, RP, FILTER, S> bk = c;
        const bool may_skip = c.nb_plane + 4 <= P && nvalid + 6 <= a.V;
        const double lx0 = P0x - S0.x, ly0 = P0y - S0.y, lz0 = P0z - S0.z;
        const double face_margin = sizeof(RP) == 8 ? 1e-7 : 1e-3;
        // The cell's vertices, compacted, for the lane-parallel tests: a cone
        // plane removes every vertex (with the margin of removes_all) when
        // each one is strictly on its removed side.
        VtxRec<R> *vtx = reinterpret_cast<VtxRec<R> *>(tcor + 16 * 9);
        int nvc = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if ((c.valid[s] >> lane) & 1)
                vtx[nvc + mbcnt(c.valid[s])] = VtxRec<R>{c.X[s], c.Y[s], c.Z[s], c.W[s], c.MXYZ[s], c.MW[s]};
            nvc += __popcll(c.valid[s]);
        }
        wsync();
        // A cone plane (unit normal n, offset d) with d < -rcell keeps no point
        // of the ball |y| <= rcell about the seed, which holds the cell: it
        // removes every vertex (kept side n.y + d >= 0), with a margin far
        // above the rounding of d and rcell.
        const double rsk = rcell * (1.0 + 1e-9) + 1e-9;
        for (int jb = sub; jb < ntri; jb += 16 * W) {
            // cone planes (A,B,C), (A,D,B), (A,C,D), (C,D,B), A = apex
            // (CCClipTetFromPointsInfinite): lane = 4 * triangle + plane
            const int i = lane >> 2, f = lane & 3;
            const int k = jb + W * i;
            Pl<RP> pl{RP(0), RP(0), RP(0), RP(0)};
            if (k < ntri) {
                const int tid = tu[k];
                const double4 G1 = a.gpts[a.tris[3 * tid]];
                double4 G2 = a.gpts[a.tris[3 * tid + 1]], G3 = a.gpts[a.tris[3 * tid + 2]];
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
                if (f == 0) goo[i] = go_out ? 1 : 0;
                // the three points of plane f in the local frame
                const double Bx = G1.x - S0.x, By = G1.y - S0.y, Bz = G1.z - S0.z;
                const double Cx = G2.x - S0.x, Cy = G2.y - S0.y, Cz = G2.z - S0.z;
                const double Dx = G3.x - S0.x, Dy = G3.y - S0.y, Dz = G3.z - S0.z;
                if (f < 3) {  // the corners in the local frame, for the face test
                    const double cx = f == 0 ? Bx : (f == 1 ? Cx : Dx);
                    const double cy = f == 0 ? By : (f == 1 ? Cy : Dy);
                    const double cz = f == 0 ? Bz : (f == 1 ? Cz : Dz);
                    tcor[9 * i + 3 * f] = cx;
                    tcor[9 * i + 3 * f + 1] = cy;
                    tcor[9 * i + 3 * f + 2] = cz;
                }
                double px, py, pz, qx, qy, qz, rx, ry, rz;
                if (f == 3) {
                    px = Cx, py = Cy, pz = Cz;
                    qx = Dx, qy = Dy, qz = Dz;
                    rx = Bx, ry = By, rz = Bz;
                } else {
                    px = lx0, py = ly0, pz = lz0;
                    if (f == 0) {
                        qx = Bx, qy = By, qz = Bz;
                        rx = Cx, ry = Cy, rz = Cz;
                    } else if (f == 1) {
                        qx = Dx, qy = Dy, qz = Dz;
                        rx = Bx, ry = By, rz = Bz;
                    } else {
                        qx = Cx, qy = Cy, qz = Cz;
                        rx = Dx, ry = Dy, rz = Dz;
                    }
                }
                const double e1x = qx - px, e1y = qy - py, e1z = qz - pz;
                const double e2x = rx - px, e2y = ry - py, e2z = rz - pz;
                double nx_ = e1y * e2z - e1z * e2y, ny_ = e1z * e2x - e1x * e2z, nz_ = e1x * e2y - e1y * e2x;
                double nd = -(nx_ * px + ny_ * py + nz_ * pz);
                const double nl = sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
                if (nl > 0) {
                    const double il = 1.0 / nl;
                    nx_ *= il;
                    ny_ *= il;
                    nz_ *= il;
                    nd *= il;
                }
                pl = Pl<RP>{RP(nx_), RP(ny_), RP(nz_), RP(nd)};
            }
            cpl[lane] = pl;
            // Skip tests, one lane per cone plane: the ball about the seed,
            // then every vertex (the test of removes_all).
            bool gone = false;
            if (may_skip && k < ntri) {
                gone = double(pl.w) < -rsk;
                if (!gone) {
                    const Pl<R> e = cvt<R>(pl);
                    const float exyz = fabsf(float(e.x)) + fabsf(float(e.y)) + fabsf(float(e.z)),
                                ew = fabsf(float(e.w));
                    gone = true;
                    // U vertices a step, their loads issued together (the
                    // last vertex stands in past the end: testing it twice
                    // changes nothing)
                    constexpr int U = sizeof(R) == 4 ? 4 : 2;
                    for (int j0 = 0; j0 < nvc && gone; j0 += U) {
                        VtxRec<R> v[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) v[u] = vtx[min(j0 + u, nvc - 1)];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const R sv = e.x * v[u].X + e.y * v[u].Y + e.z * v[u].Z + e.w * v[u].W;
                            R m;
                            if (FILTER)
                                m = R(kFilterKappa * (exyz * v[u].MXYZ + ew * v[u].MW));
                            else
                                m = R(sizeof(R) == 8 ? 0x1.0p-33 : 0x1.0p-16) *
                                    (fabs(e.x * v[u].X) + fabs(e.y * v[u].Y) + fabs(e.z * v[u].Z) +
                                     fabs(e.w * v[u].W));
                            gone = gone && sv > m;
                        }
                    }
                }
            }
            uint64_t skipm = ballot(gone);
            wsync();
            // The cone region of a triangle (apex A, beyond the triangle) lies
            // strictly outside one of the cell's planes when the triangle's
            // three corners do (with a margin) and the rays from A through
            // them do not turn back towards the plane.  The cell is inside
            // every plane of its list (each one cut it when added), so the two
            // are disjoint: the reference's clip ends empty_cell (same caps
            // guard as above).  Lane 4i + f tests planes f, f + 4, ... against
            // triangle i.
            {
                bool sep = false;
                if (may_skip && k < ntri && ((skipm >> (4 * i)) & 0xfull) == 0) {
                    double px[3], py[3], pz[3];
#pragma unroll
                    for (int v = 0; v < 3; ++v) {
                        px[v] = tcor[9 * i + 3 * v];
                        py[v] = tcor[9 * i + 3 * v + 1];
                        pz[v] = tcor[9 * i + 3 * v + 2];
                    }
                    // two planes a step, loaded together; past the end the
                    // last plane stands in (any plane of the cell may
                    // separate, and the triangle's other lanes test it too)
                    for (int j = f; j < c.nb_plane && !sep; j += 8) {
                        Pl<RP> q2[2];
#pragma unroll
                        for (int u = 0; u < 2; ++u) q2[u] = c.planes[min(j + 4 * u, c.nb_plane - 1)];
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const double nx = double(q2[u].x), ny = double(q2[u].y), nz = double(q2[u].z),
                                         nw = double(q2[u].w);
                            const double m = face_margin * (1.0 + fabs(nw));
                            bool out = true;
#pragma unroll
                            for (int v = 0; v < 3; ++v) {
                                out = out && nx * px[v] + ny * py[v] + nz * pz[v] + nw < -m;
                                out = out && nx * (px[v] - lx0) + ny * (py[v] - ly0) + nz * (pz[v] - lz0) <= 0.0;
                            }
                            sep = sep || out;
                        }
                    }
                }
                skipm |= ballot(sep);
            }
            // triangles of the batch that are clipped: none of their four
            // lanes skipped them
            uint32_t todo = 0;
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (jb + W * t < ntri && ((skipm >> (4 * t)) & 0xfull) == 0) todo |= 1u << t;
            if (W > 1 && lane < 16 && jb + W * lane < ntri) contrib[jb + W * lane] = make_double4(0.0, 0.0, 0.0, 0.0);
            // in increasing k: the running sums add the contributions in order
            // (a skipped triangle adds +0, which leaves them unchanged)
            while (todo) {
                const int t = __builtin_ctz(todo);
                todo &= todo - 1;
                const int kk = jb + W * t;
                const int tid = uni(tu[kk]);
                if (tid > fail_t) continue;
                double4 ck = make_double4(0.0, 0.0, 0.0, 0.0);
                int res = kNoClip;
                // each cone plane read from LDS just before its clip: the four
                // held at once would take 32 registers through the clips
#pragma unroll
                for (int f2 = 0; f2 < 4; ++f2) {
                    if (res < 0) {
                        const Pl<RP> q = cpl[4 * t + f2];
                        res = c.nb_plane >= P ? int(RPD_PLANE_OVERFLOW) : c.template clip<false>(cvt<R>(q), -1, q);
                    }
                }
                if (res >= 0 && res != RPD_EMPTY_CELL) {
                    fail_t = tid;
                    fail_st = res;
                } else if (res != RPD_EMPTY_CELL) {
                    ck = integrate(goo[t] ? 1.0 : -1.0);
                }
                c = bk;
                if (W == 1) {
                    sx = uni(sx + ck.x);
                    sy = uni(sy + ck.y);
                    sz = uni(sz + ck.z);
                    sm = uni(sm + ck.w);
                } else if (lane == 0) {
                    contrib[kk] = ck;
                }
            }
            wsync();  // the next batch overwrites cpl / goo
        }
    }
    if (lane == 0) {
        fails[2 * sub] = fail_t;
        fails[2 * sub + 1] = fail_st;
    }
    if (W > 1)
        __syncthreads();
    else
        wsync();
    if (sub != 0) return;
    // smallest failing triangle id over the waves
    int fk = 0x7fffffff, fs = 0;
    for (int w = 0; w < W; ++w) {
        const int k = fails[2 * w];
        if (k < fk) {
            fk = k;
            fs = fails[2 * w + 1];
        }
    }
    if (fk != 0x7fffffff) {
        fail(fs);
        return;
    }
    // fixed summation order: the contributions in increasing k, then the
    // whole-cell term
    if (W > 1) {
        for (int k = 0; k < ntri; ++k) {
            const double4 q = contrib[k];
            sx += q.x;
            sy += q.y;
            sz += q.z;
            sm += q.w;
        }
    }
    const double acc_m = cadd.w + sm, acc_x = cadd.x + sx, acc_y = cadd.y + sy, acc_z = cadd.z + sz;
    // Hessian entries (ConvexCell.cl:1133-1151): the record's planes towards
    // a seed whose face keeps more than 0.2 of area, in plane order, with the
    // neighbours' original indices; more than H: hessian_overflow, the first
    // H written
    int st_out = RPD_SUCCESS;
    if (HESS) {
        const int *rnid = reinterpret_cast<const int *>(rtr + 64 * a.rec_S);
        const int H = a.H;
        int cnt = 0;
        for (int pb = 0; pb < nbk; pb += 64) {
            const int p = pb + lane;
            int nb = -1;
            bool keep = false;
            if (p < nbk) {
                nb = rnid[p];
                keep = nb >= 0 && hacc[4 * p + 3] > 0.2;
            }
            const uint64_t m = ballot(keep);
            const int pos = cnt + mbcnt(m);
            if (keep && pos < H) {
                const size_t e = size_t(orig) * size_t(H) + size_t(pos);
                const double ar = hacc[4 * p + 3];
                a.hids[e] = a.perm[nb];
                if (a.hc == 1) {
                    a.hess[e] = ar;
                } else {
                    a.hess[4 * e + 0] = hacc[4 * p + 0] + ar * S0.x;
                    a.hess[4 * e + 1] = hacc[4 * p + 1] + ar * S0.y;
                    a.hess[4 * e + 2] = hacc[4 * p + 2] + ar * S0.z;
                    a.hess[4 * e + 3] = ar;
                }
            }
            cnt += __popcll(m);
        }
        if (cnt > H)
            st_out = RPD_HESSIAN_OVERFLOW;
        else if (cnt < H && lane == 0)
            a.hids[size_t(orig) * size_t(H) + size_t(cnt)] = -1;
        if (a.dcnt && lane == 0) atomicMax(a.dcnt + kDcMax + 4, min(cnt, H));  // ConvexCell.cl:1150
    }
    // local first moments -> global frame
    finish(st_out, acc_x + acc_m * S0.x, acc_y + acc_m * S0.y, acc_z + acc_m * S0.z, acc_m);
}

}  // namespace rpd

using namespace rpd;

// Bytes of one stored plane: the planes of the FP64 and mixed forms are FP64.
static size_t plane_bytes(int form) { return form == RPD_PREC_FP32 ? 16 : 32; }

// LDS bytes a workgroup may use on the current device.
// LDS a workgroup may use on the current device (asked once per device)
static size_t lds_limit() {
    static std::atomic<int> cached[64];
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (dev >= 0 && dev < 64 && (v = cached[dev].load(std::memory_order_relaxed)) > 0) return size_t(v);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 64 * 1024;
    if (dev >= 0 && dev < 64) cached[dev].store(v, std::memory_order_relaxed);
    return size_t(v);
}

// LDS of the cell part of a finish-kernel workgroup: fail slots, then, when
// W > 1 waves share the cell, its contributions (T).
static size_t finish_cell_stride(int T, int W) { return W > 1 ? (64 + 32 * size_t(T) + 15) & ~size_t(15) : 64; }

// LDS of one finish-kernel cell handled by a single wave.
static size_t finish_cell_lds(int form, int P, int V, int T, bool hess) {
    const int S = V <= 64 ? 1 : (V <= 128 ? 2 : 4);
    return finish_cell_stride(T, 1) + finish_wave_lds(P, S, T, plane_bytes(form), true) +
           (hess ? hessian_wave_lds(P, S) : 0);
}

int rpd_clip_max_T(int form, int P, int V, bool hess) {
    const size_t lim = lds_limit();
    int lo = 0, hi = 1 << 20;
    if (finish_cell_lds(form, P, V, 0, hess) > lim) return -1;
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (finish_cell_lds(form, P, V, mid, hess) <= lim)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

int rpd_clip_plane_room(int form, int P, int V, int T, int floor) {
    // LDS of a gfx950 CU, shared by the resident workgroups
    constexpr size_t kLdsPerCu = 160 * 1024;
    const int S = V <= 64 ? 1 : (V <= 128 ? 2 : 4);
    if (S != 1) return P;
    // cells per CU at the kernels' residency: 4 waves per SIMD for the
    // finish kernel, kClipWaves for the bisector kernel (one cell a wave)
    const size_t fin_cells = 4 * 4;
    const size_t bis_cells = 4 * size_t(form == RPD_PREC_FP64 ? kClipWaves<double, 1> : kClipWaves<float, 1>);
    const size_t bis_pl = form == RPD_PREC_FP64 ? 32 : 16;  // the bisector kernel's planes, in the test precision
    for (int p = std::min(floor, 255); p > P; --p)
        if (finish_cell_lds(form, p, V, T, false) * fin_cells <= kLdsPerCu &&
            finish_wave_lds(p, S, 0, bis_pl) * bis_cells <= kLdsPerCu)
            return p;
    return P;
}

// header | P planes | 64 S vertex triples | P plane neighbours (int)
size_t rpd_clip_record_bytes(int form, int P, int V) {
    const int S = V <= 64 ? 1 : (V <= 128 ? 2 : 4);
    const size_t pl = plane_bytes(form);
    return (sizeof(RecHeader) + size_t(P) * pl + 4 * 64 * size_t(S) + 4 * size_t(P) + 15) & ~size_t(15);
}

static int make_args(int form, const ClipLaunch &L, ClipArgs &a) {
    if (L.P < 7 || L.P > 255 || L.V < 8 || L.K < 1) return RPD_E_ARG;
    if (L.V > 256) return RPD_E_LIMIT;
    if (!L.rec || !L.qcount) return RPD_E_ARG;
    if (L.reqs && (!L.req_of || (L.hess && L.H > 0))) return RPD_E_ARG;
    a = ClipArgs{};
    a.seeds = L.seeds;
    a.perm = L.perm;
    a.ids = L.ids;
    a.knn = L.knn;
    a.n_ids = L.n_ids;
    a.n_seeds = L.n_seeds;
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
        a.tvoffs = L.dom->tvoffs;
        a.tvox = L.dom->tvox;
        a.tinfo = L.dom->tinfo;
        a.dist = L.cull ? L.dom->dist : nullptr;
    }
    a.security_mode = L.security_mode;
    a.wmin = L.wmin;
    a.out = L.out;
    a.status = L.status;
    a.kused = L.kused;
    a.rec = static_cast<char *>(L.rec);
    const int rec_P = L.rec_P > 0 ? L.rec_P : L.P, rec_V = L.rec_V > 0 ? L.rec_V : L.V;
    if (rec_P > L.P || rec_V > L.V) return RPD_E_ARG;
    a.rec_stride = int(rpd_clip_record_bytes(form, rec_P, rec_V));
    a.rec_P = rec_P;
    a.rec_S = rec_V <= 64 ? 1 : (rec_V <= 128 ? 2 : 4);
    a.qcount = L.qcount;
    a.fl_cnt = L.fl_cnt;
    a.fl_eo = L.fl_eo;
    a.fl_sec = L.fl_sec;
    a.rs_out = L.fl_sec ? L.rs_out : nullptr;
    a.rs_out_stride = L.rs_out_stride;
    a.rs_cap = L.rs_cap;
    a.rs_count = L.rs_count;
    a.rs_in = form == RPD_PREC_FP64 ? L.rs_in : nullptr;
    a.rs_in_stride = L.rs_in_stride;
    a.rs_in_P = L.rs_in_P;
    a.rslot = L.rslot;
    a.lastnb = L.lastnb;
    a.hess = L.hess;
    a.hids = L.hids;
    a.H = L.hess && L.hids ? L.H : 0;
    a.hc = L.hc == 4 ? 4 : 1;
    a.exact_final = form == RPD_PREC_FP64 ? 1 : 0;
    a.dcnt = L.dcnt;
    a.reqs = L.reqs;
    a.req_of = L.req_of;
    a.slots = L.slots;
    a.nq_dev = L.nq_dev;
    a.fl_slot = L.fl_slot;
    return RPD_OK;
}

int rpd_launch_cells(hipStream_t st, int form, const ClipLaunch &L, int *nq_out, bool wait) {
    *nq_out = 0;
    if (L.n_ids <= 0) return RPD_OK;
    ClipArgs a;
    const int r = make_args(form, L, a);
    if (r != RPD_OK) return r;
    const int S = L.V <= 64 ? 1 : (L.V <= 128 ? 2 : 4);
    if (!L.qcount_zeroed && rpd_zero_ints(st, L.qcount, 1) != RPD_OK) return RPD_E_DEVICE;

    // kernel A: the bisector loop of every cell
    {
        const size_t stride = finish_wave_lds(L.P, S, 0, form == RPD_PREC_FP64 ? 32 : 16);  // planes in R
        a.lds_stride = int(stride);
        // one wave per workgroup: each cell's wave starts and ends on its own
        const int wpb = 1;
        const unsigned blocks = unsigned((L.n_ids + wpb - 1) / wpb);
        const size_t shm = stride * wpb;
#define RPD_CLIP_A1(R, RP, F, B)                                                      \
    do {                                                                              \
        if (S == 1)                                                                   \
            clip_cells_kernel<R, RP, F, 1, B><<<blocks, dim3(64 * wpb), shm, st>>>(a);\
        else if (S == 2)                                                              \
            clip_cells_kernel<R, RP, F, 2, B><<<blocks, dim3(64 * wpb), shm, st>>>(a);\
        else                                                                          \
            clip_cells_kernel<R, RP, F, 4, B><<<blocks, dim3(64 * wpb), shm, st>>>(a);\
    } while (0)
#define RPD_CLIP_A(R, RP, F)              \
    do {                                  \
        if (a.reqs)                       \
            RPD_CLIP_A1(R, RP, F, true);  \
        else                              \
            RPD_CLIP_A1(R, RP, F, false); \
    } while (0)
        if (form == RPD_PREC_FP64)
            RPD_CLIP_A(double, double, false);
        else if (form == RPD_PREC_MIXED)
            RPD_CLIP_A(float, double, true);
        else
            RPD_CLIP_A(float, float, false);
#undef RPD_CLIP_A
#undef RPD_CLIP_A1
        if (hipGetLastError() != hipSuccess) return RPD_E_DEVICE;
    }
    if (L.ev_mid && hipEventRecord(L.ev_mid, st) != hipSuccess) return RPD_E_DEVICE;
    if (L.nq_dev_out) {
        const bool fl = L.fl_cnt && L.fl_cnt_host && L.fl_cnt_dev_out;
        if (rpd_pin_copy(st, L.qcount, L.nq_dev_out, 1, fl ? L.fl_cnt : nullptr, fl ? L.fl_cnt_dev_out : nullptr,
                         fl ? kFailCounters : 0) != RPD_OK)
            return RPD_E_DEVICE;
    } else {
        if (hipMemcpyAsync(nq_out, L.qcount, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
            return RPD_E_DEVICE;
        if (L.fl_cnt && L.fl_cnt_host &&
            hipMemcpyAsync(L.fl_cnt_host, L.fl_cnt, kFailCounters * sizeof(int), hipMemcpyDeviceToHost, st) !=
                hipSuccess)
            return RPD_E_DEVICE;
    }
    if (!wait) return RPD_OK;
    return hipStreamSynchronize(st) == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

namespace rpd {
__global__ __launch_bounds__(64) void pin_copy_kernel(const int *src0, int *dst0, int n0, const int *src1,
                                                      int *dst1, int n1) {
    for (int i = int(threadIdx.x); i < n0; i += 64) dst0[i] = src0[i];
    for (int i = int(threadIdx.x); i < n1; i += 64) dst1[i] = src1[i];
    __threadfence_system();  // the host reads them once the stream is done
}
}  // namespace rpd

namespace rpd {
__global__ __launch_bounds__(64) void zero_ints_kernel(int *p, int n) {
    for (int i = int(threadIdx.x); i < n; i += 64) p[i] = 0;
}
}  // namespace rpd

int rpd_zero_ints(hipStream_t st, int *p, int n) {
    zero_ints_kernel<<<1, 64, 0, st>>>(p, n);
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

int rpd_pin_copy(hipStream_t st, const int *src0, int *dst0, int n0, const int *src1, int *dst1, int n1) {
    pin_copy_kernel<<<1, 64, 0, st>>>(src0, dst0, n0, src1, dst1, n1);
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

int rpd_launch_finish(hipStream_t st, int form, const ClipLaunch &L, int nq, int beside) {
    if (nq <= 0) return RPD_OK;
    ClipArgs a;
    const int r = make_args(form, L, a);
    if (r != RPD_OK) return r;
    const int S = L.V <= 64 ? 1 : (L.V <= 128 ? 2 : 4);
    // kernel B: restriction + integration of the queued cells.  Short queues
    // (re-queue passes) spread each cell over W waves.
    {
        const size_t pl = plane_bytes(form);
        const bool hess = a.H > 0;
        const size_t wstride = finish_wave_lds(L.P, S, a.T, pl, true) + (hess ? hessian_wave_lds(L.P, S) : 0);
        const size_t kLdsMax = lds_limit();
        if (finish_cell_stride(a.T, 1) + wstride > kLdsMax) return RPD_E_LIMIT;
        // waves per cell from the cells that share the device: this queue's
        // and those of finish kernels running beside it (the Hessian variant:
        // one wave per cell)
        const int64_t load = int64_t(nq) + std::max(beside, 0);
        int W = hess || load > 32768 ? 1 : (load <= 2048 ? 8 : (load <= 8192 ? 4 : 2));
        while (W > 1 && finish_cell_stride(a.T, W) + W * wstride > kLdsMax) W >>= 1;
        const size_t cstride = finish_cell_stride(a.T, W);
        a.lds_stride = int(wstride);
        a.cell_stride = int(cstride);
        // One cell per workgroup: a workgroup keeps its LDS until its last
        // wave ends, and cell costs are skewed (interior cells are cheap), so
        // several cells per workgroup would hold the LDS of finished cells for
        // the slowest one.  LDS bounds the residency of this kernel.
        const int wpb = W;
        const int cpb = 1;
        const unsigned blocks = unsigned((nq + cpb - 1) / cpb);
        const size_t shm = size_t(cpb) * cstride + size_t(wpb) * wstride;
#define RPD_CLIP_B1(R, RP, F, HS, B)                                                                 \
    do {                                                                                             \
        if (S == 1)                                                                                  \
            finish_cells_kernel<R, RP, F, 1, HS, B><<<blocks, dim3(64 * wpb), shm, st>>>(a, nq, W);  \
        else if (S == 2)                                                                             \
            finish_cells_kernel<R, RP, F, 2, HS, B><<<blocks, dim3(64 * wpb), shm, st>>>(a, nq, W);  \
        else                                                                                         \
            finish_cells_kernel<R, RP, F, 4, HS, B><<<blocks, dim3(64 * wpb), shm, st>>>(a, nq, W);  \
    } while (0)
// a packed batch has no Hessian
#define RPD_CLIP_B(R, RP, F, HS)                  \
    do {                                          \
        if (!HS && a.reqs)                        \
            RPD_CLIP_B1(R, RP, F, false, true);   \
        else                                      \
            RPD_CLIP_B1(R, RP, F, HS, false);     \
    } while (0)
        if (form == RPD_PREC_FP64) {
            if (hess)
                RPD_CLIP_B(double, double, false, true);
            else
                RPD_CLIP_B(double, double, false, false);
        } else if (form == RPD_PREC_MIXED) {
            if (hess)
                RPD_CLIP_B(float, double, true, true);
            else
                RPD_CLIP_B(float, double, true, false);
        } else {
            if (hess)
                RPD_CLIP_B(float, float, false, true);
            else
                RPD_CLIP_B(float, float, false, false);
        }
#undef RPD_CLIP_B
#undef RPD_CLIP_B1
    }
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}
