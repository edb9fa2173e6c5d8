// rpd_internal.h -- declarations shared by the translation units of librpd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "rpd.h"

// Host domain grid behind the rpd_grid handle (the content of Grid,
// Grid.h:14-70, plus the voxel-centre classes).
struct rpd_grid {
    int32_t size = 0;
    std::vector<int8_t> in_domain;
    std::vector<double> points;  // xyzw
    std::vector<int32_t> triangles;
    std::vector<int32_t> tri_list;
    std::vector<int32_t> offsets;
    std::vector<int8_t> centre;  // voxel-centre classes 0 out / 2 in / 1 unknown
};

// Fractions of a voxel at which the restriction puts its cone apex (the
// kernels' apex_mode 1); the voxel-centre classes are those of these points.
constexpr double kApexFracHost[3] = {0.530901699437495, 0.520710678118655, 0.536602540378444};

// Voxel-centre classes of a grid from its triangle lists (rpd_grid.cpp).
void rpd_derive_centre_classes(rpd_grid *gr);
// Structural check of a grid description (rpd_grid.cpp): array pointers
// present, counts non-negative, triangles' point ids and list entries'
// triangle ids in range, offsets starting at 0, non-decreasing and ending at
// most at num_tri_list.  Returns nullptr when sound, else what is wrong.
const char *rpd_grid_desc_check(const rpd_grid_desc *g);

// Growable device buffer (hipMalloc'ed, freed with the owner).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes && ptr) return RPD_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (need == 0) return RPD_OK;
        if (hipMalloc(&ptr, need) != hipSuccess) {
            ptr = nullptr;
            return RPD_E_NOMEM;
        }
        bytes = need;
        return RPD_OK;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(ptr);
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// ---- kNN (rpd_knn.hip) ----------------------------------------------------
int rpd_knn_dim(int64_t n);
int rpd_bin_seeds(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch, int *perm,
                  int *counters, void *stored, double4 *stored64);
// The same binning in caller-given scratch (scratch_bytes of it; RPD_E_LIMIT
// when rpd_bin_scratch_bytes needs more); flag_host receives the
// crowded-cell flag (pinned host memory for an asynchronous copy).
int rpd_bin_seeds_in(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, void *scratch,
                     size_t scratch_bytes, int *flag_host, int *perm, int *counters, void *stored, double4 *stored64);
size_t rpd_bin_scratch_bytes(hipStream_t st, int64_t n, int dim);
// The binning queued without a wait: the crowded-cell flag goes to
// flag_pinned (pinned host memory).  Once the stream has been synchronised, a
// set flag calls for rpd_bin_seeds_redo (the crowded-cell path; it waits).
int rpd_bin_seeds_async(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch,
                        int *flag_pinned, int *perm, int *counters, void *stored, double4 *stored64);
int rpd_bin_seeds_redo(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch,
                       int *perm, void *stored, double4 *stored64);
// Several diagrams in one launch (a packed batch): request r owns the
// stored ids [base_r, base_r + n_r) of the shared arrays (seeds, sorted
// copies, permutation with original indices base_r + i, outputs), and its
// kNN grid is the counters (global stored ids) at its own offset; req_of[s]
// is the request of stored id s.  Per request, the kNN reads:
struct KnnReq {
    const int *counters;  // dim^3 + 1 entries, global stored ids
    int dim;
};
// K nearest of each query (stored ids ids[0..n_ids), nullptr: all), or, with
// lower set, the K after the neighbour lower[q] (stored-id indexed).  With
// kreqs set, each query searches the grid of its request (req_of) instead of
// (counters, dim).
// n_dev (nullable): the number of queries in device memory, at most n_ids
// (the launch covers n_ids; waves past the count exit).  narrow: one-wave
// workgroups, for launches beside a finish kernel (whose workgroups leave too
// little free LDS on a CU for wider ones).
int rpd_launch_knn(hipStream_t st, int prec, const void *stored, const int *perm, const int *counters, int dim,
                   const int *ids, int64_t n_ids, int K, int *out, const int *lower = nullptr,
                   const KnnReq *kreqs = nullptr, const uint16_t *req_of = nullptr, const int *n_dev = nullptr,
                   bool narrow = false);
// adds base to n ints (binning of one request of a packed batch)
int rpd_add_base(hipStream_t st, int *v, int64_t n, int base);

// ---- clip (rpd_clip.hip) --------------------------------------------------
struct DomainDev {
    int g = 0;
    double vs = 0;
    const int8_t *flags = nullptr;
    const int8_t *centre = nullptr;
    const double4 *pts = nullptr;
    const int *tris = nullptr;
    const int *tlist = nullptr;
    const int *toffs = nullptr;
    const int *tvoffs = nullptr;     // per triangle: offsets into tvox
    const uint32_t *tvox = nullptr;  // per triangle: sorted voxels (z<<20 | y<<10 | x) whose list holds it
    const uint4 *tinfo = nullptr;    // per triangle-list entry: its voxel, earlier near voxels holding it
    const uint8_t *dist = nullptr;   // Chebyshev voxel distance to domain voxels
    const int *dvox = nullptr;       // the domain voxels (distance 0), ascending
    int n_dvox = 0;
    bool has_border_vertices = false;
};

// Per-request fields of a packed batch that the clip kernels read in place
// of the launch's own (seed count, weight minimum, domain).
struct ReqDesc {
    int64_t n_seeds;
    double wmin;
    int g;
    double vs;
    const int8_t *flags;
    const int8_t *centre;
    const double4 *gpts;
    const int *tris;
    const int *tlist;
    const int *toffs;
    const int *tvoffs;
    const uint32_t *tvox;
    const uint4 *tinfo;
    const uint8_t *dist;  // nullptr: no cull
};

struct ClipLaunch {
    const double4 *seeds;  // stored order, FP64
    const int *perm;       // stored -> original
    const int *ids;        // stored ids to process (list position t)
    const int *knn;        // n_ids * K stored ids
    int64_t n_ids;
    int64_t n_seeds;       // seeds in the diagram
    int K, P, V, T;
    int apex_mode;
    int security_mode;
    double wmin;
    bool cull;
    const DomainDev *dom;  // nullptr: no domain
    double4 *out;          // original order
    int *status;           // original order
    int *kused;            // original order (nullable)
    void *rec;             // n_ids * rpd_clip_record_bytes(form, P, V) bytes
    int *qcount;           // one device int: queued cells
    bool qcount_zeroed;    // the caller has cleared qcount (rpd_launch_cells does not)
    hipEvent_t ev_mid;     // recorded after the bisector kernel (nullable)
    // failure list the kernels append to (nullable): kFailCounters counters
    // ([0] exact entries, [1] overflow entries, [2 + s] cells per status s,
    // [13] largest triangle union over T, [14] resumable entries), then the
    // stored ids: exact entries at fl_eo[-1 - k], overflows at fl_eo[k],
    // resumable cells at fl_sec[k]
    int *fl_cnt, *fl_eo, *fl_sec;
    int *fl_cnt_host;      // host copy of the counters, filled by rpd_launch_cells (nullable)
    // nullable: device addresses of nq_out / fl_cnt_host (mapped pinned
    // memory), written by a one-wave copy kernel instead of copy engine
    // blits, which wait for room beside a running finish kernel
    int *nq_dev_out, *fl_cnt_dev_out;
    // resumable cells (security radius not reached), see ClipArgs in rpd_clip.hip
    char *rs_out;          // records written by this pass (nullable)
    int rs_out_stride, rs_cap;
    int *rs_count;
    const char *rs_in;     // records this pass resumes from (nullable; FP64 form)
    int rs_in_stride, rs_in_P;
    int *rslot, *lastnb;   // per stored id
    // Hessian (H > 0): per original seed H entries of hc values, neighbours'
    // original indices (-1 after the last); finish kernel only
    double *hess;
    int *hids;
    int H, hc;
    int *dcnt;             // debug counters of the pass (kDcStride ints, nullable)
    // packed batch (nullable): per request fields, request of each stored id
    const ReqDesc *reqs;
    const uint16_t *req_of;
    // finish kernel: records at these slots instead of 0..nq-1 (nullable),
    // and where the record slot of each listed failure goes (nullable, laid
    // out as fl_eo)
    const int *slots;
    int *fl_slot;
    // finish kernel: the P and V the records were written with (0: P, V)
    int rec_P, rec_V;
    // finish kernel (nullable): the queue length in device memory, at most
    // the nq it is launched with (items past it exit)
    const int *nq_dev;
};

constexpr int kFailCounters = 16;

// Debug counters of one pass (the reference's debugCounter0..4,
// ConvexCell.cl:15-39, 887-911, 981-982, 1128-1150): [kDcMax + 0..4] largest
// K used, P and V after the neighbour loop, triangle union, Hessian entries;
// then histograms of K used (0..256), P (0..255), V (0..256) and the union
// size (0..kDcUnionMax, larger unions in the last bin).
constexpr int kDcMax = RPD_DCOUNT_MAX;
constexpr int kDcHistK = RPD_DCOUNT_HIST_K;
constexpr int kDcHistP = RPD_DCOUNT_HIST_P;
constexpr int kDcHistV = RPD_DCOUNT_HIST_V;
constexpr int kDcHistT = RPD_DCOUNT_HIST_T;
constexpr int kDcUnionMax = RPD_DCOUNT_UNION_MAX;
constexpr int kDcStride = RPD_DCOUNT_INTS;

// Cells whose restriction is provably empty, before the kNN (rpd_cull.hip).
// One per request: its cells are the stored ids [base, base + n).
struct CoverArgs {
    const double4 *stored64;  // sorted seeds, FP64 (every request)
    const int *perm;          // stored -> original (every request)
    const int *counters;      // the request's kNN grid: dim^3 + 1 exclusive prefix, global stored ids
    int dim;
    int64_t base, n;
    bool apply;               // false: every cell is kept
    const int *dvox;          // domain voxels
    int n_dvox;
    const uint8_t *dist;      // Chebyshev voxel distance to the domain voxels
    int g;
    double vs;
    double wmin;              // minimum weight (no border seeds)
    // nullable: the device pair (minimum weight over non-border seeds,
    // minimum over all) replacing wmin; a second one <= RPD_NO_CELL_LIMITS
    // (walls) turns the test off
    const double *wdev;
    double4 *out;             // original order: skipped cells get 0
    int *status;
    int *kused;
};
// Writes the outputs of the skipped cells and the stored ids of the others,
// ascending, to ids_out (n_total entries); *n_out = their number.  The
// requests' ranges lie in [0, n_total).  count_dev null: synchronizes st.
// Else nothing waits: *count_dev receives the device copy of the number and
// *n_out (pinned host memory) is written by a copy queued on st.
int rpd_cover_skip(hipStream_t st, const CoverArgs *reqs, int nreq, int64_t n_total, DevBuf &scratch, int *ids_out,
                   int *n_out, const int **count_dev = nullptr);

// Bytes of one queued-cell record.
size_t rpd_clip_record_bytes(int form, int P, int V);
// Largest T whose finish-kernel cell fits the current device's workgroup LDS
// (-1: none).
int rpd_clip_max_T(int form, int P, int V, bool hess = false);
// The plane cap P raised toward floor as far as the bisector and finish
// kernels keep their residency (LDS per CU) for one-slot cells (V <= 64)
// with T triangle entries; P itself otherwise.
int rpd_clip_plane_room(int form, int P, int V, int T, int floor);
// form: RPD_PREC_FP64 / RPD_PREC_MIXED / RPD_PREC_FP32.  The bisector-loop
// kernel of L.n_ids cells, then copies to the host the number of cells it
// Copies n0 ints src0 -> dst0 and n1 src1 -> dst1 with one one-wave kernel
// (dst: device addresses of mapped pinned host memory; n1 may be 0).  Small
// copies of counters to the host this way need one free wave slot, where a
// copy engine blit waits for room beside a running finish kernel.
int rpd_pin_copy(hipStream_t st, const int *src0, int *dst0, int n0, const int *src1 = nullptr, int *dst1 = nullptr,
                 int n1 = 0);
// Clears n device ints with one one-wave kernel (a fill blit, like a copy,
// waits for room beside a running finish kernel).
int rpd_zero_ints(hipStream_t st, int *p, int n);
// queued (and the failure counters when L.fl_cnt_host is set); waits for
// them unless wait is false (the caller synchronizes st before reading).
int rpd_launch_cells(hipStream_t st, int form, const ClipLaunch &L, int *nq, bool wait = true);
// The finish kernel over the nq records the bisector kernel queued (async);
// beside: cells of other finish kernels expected to run at the same time
// (they set how many waves each cell gets).
int rpd_launch_finish(hipStream_t st, int form, const ClipLaunch &L, int nq, int beside = 0);
