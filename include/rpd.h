/*
 * rpd.h -- C-ABI of librpd, the MI355X (gfx950) restricted power diagram
 * hot path: grid kNN -> per-seed power-cell clipping inside a voxel-gridded
 * domain -> fused per-cell volume and first-moment integration.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and
 * a negative RPD_E* code on failure; rpd_last_error() gives the message.
 *
 * Each declaration names the reference interface it replaces
 * (basselin7u/GPU-Restricted-Power-Diagrams, file:line).
 */
#ifndef RPD_H
#define RPD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define RPD_OK 0
#define RPD_E_ARG -1      /* invalid argument                               */
#define RPD_E_STATE -2    /* call out of order (e.g. solve before seeds)    */
#define RPD_E_DEVICE -3   /* HIP runtime error / no GPU                     */
#define RPD_E_NOMEM -4    /* device allocation failed                       */
#define RPD_E_IO -5       /* file could not be read or written             */
#define RPD_E_LIMIT -6    /* request exceeds a stated limit                 */

/* ---- per-cell status: identical numbering to Status.h:4-16 -------------- */
enum rpd_status {
    RPD_VERTEX_OVERFLOW = 0,
    RPD_PLANE_OVERFLOW = 1,
    RPD_INCONSISTENT_BOUNDARY = 2,
    RPD_SECURITY_RADIUS_NOT_REACHED = 3,
    RPD_SUCCESS = 4,
    RPD_NEEDS_EXACT_PREDICATES = 5,
    RPD_EMPTY_CELL = 6,
    RPD_FIND_ANOTHER_BEGINNING_VERTEX = 7,
    RPD_TRIANGLE_OVERFLOW = 8,
    RPD_SEED_ON_BORDER = 9,
    RPD_HESSIAN_OVERFLOW = 10,
    RPD_STATUS_NUM = 11
};

/* NO_CELL_LIMITS of voronoi.h:53: a seed with w <= this is a "border seed"
   (copied to the output, acts as a wall for its neighbours). */
#define RPD_NO_CELL_LIMITS (-100000.0)

/* ---- precision forms --------------------------------------------------- */
enum rpd_precision {
    RPD_PREC_FP64 = 0,   /* all-FP64 clip + integration (USE_DOUBLE build)    */
    RPD_PREC_MIXED = 1,  /* FP32 seed-local clip with a certified error filter;
                            flagged cells recomputed in FP64; FP64 moment
                            accumulation across domain pieces                */
    RPD_PREC_FP32 = 2    /* FP32 seed-local clip, no filter, no FP64 redo     */
};

/* ---- domain grid: content of Grid (Grid.h:14-70) ------------------------ */
typedef struct {
    int32_t size;                 /* voxels per axis g; voxel (i,j,k) id
                                     i + j*g + k*g*g, frame [0,1000]^3        */
    const int8_t *in_domain;      /* g^3 flags of each voxel's first corner:
                                     0 outside, 1 on border, 2 inside         */
    const double *points;         /* 4*num_points xyzw (w ignored)            */
    int32_t num_points;
    const int32_t *triangles;     /* 3*num_triangles point ids                */
    int32_t num_triangles;
    const int32_t *tri_list;      /* concatenated sorted per-voxel lists      */
    int32_t num_tri_list;
    const int32_t *offsets;       /* g^3+1 offsets into tri_list              */
    const int8_t *centre;         /* nullable: g^3 classes of voxel centres,
                                     0 outside / 2 inside / 1 unknown; used to
                                     put the cone apex of the restriction at a
                                     voxel centre, never on a lattice plane    */
} rpd_grid_desc;

/* Host-owned grid produced by the helpers below (free with rpd_grid_free). */
typedef struct rpd_grid rpd_grid;

/* Voxel domain shapes in the unit frame (scaled x1000). A voxel is kept iff
   its centre is inside the shape. */
enum rpd_shape {
    RPD_SHAPE_SPHERE = 0,        /* p0..p2 centre, p3 radius                   */
    RPD_SHAPE_TORUS = 1,         /* p0..p2 centre, p3 R, p4 r, axis z          */
    RPD_SHAPE_CUBE = 2,          /* whole grid                                 */
    RPD_SHAPE_CUBE_MINUS_SPHERE = 3 /* whole grid minus sphere p0..p2, p3     */
};

/* Voxel mask of an analytic shape: out_mask[g^3] = 1 kept, 0 not kept. */
int rpd_voxel_mask(int shape, int g, const double *params, uint8_t *out_mask);

/* Grid of a voxel union (replaces utils/createGrid/main.cpp:172-294 for
   voxel domains): boundary faces between kept and non-kept voxels (grid
   exterior counts as not kept), 2 triangles each, oriented so that
   Grid::getVolume() > 0; corner flags 2/0/1 = all/none/some of the 8 incident
   voxels kept; per-voxel sorted lists by exact closed-box overlap.
   RPD_E_LIMIT when the triangle point ids (3 per triangle) or the list
   entries exceed INT32_MAX (the int32 arrays of Grid). */
int rpd_grid_from_voxels(int g, const uint8_t *mask, rpd_grid **out);
/* Grid of a closed triangle mesh, built on the GPU (replaces the mesh path of
   utils/createGrid/main.cpp:150-294): points (3*num_points xyz) and
   triangles (3*num_triangles point ids); points no triangle uses are dropped
   and the rest renumbered in order (main.cpp:164).  normalize != 0 scales
   the mesh into [4.998, 995.22]^3 (main.cpp:168-175); normalize == 0
   requires every point in [0,1000]^3.  Per-voxel lists: the voxels of each
   triangle's bounding range (main.cpp:180-210) that pass the
   single-precision triangle/box overlap test (geometry.cpp:433-508), in
   ascending triangle order.  Corner flags: 1 within 1e-3 of the surface
   (main.cpp:276) or undecided, else 2 inside / 0 outside (the reference's
   tetrahedralization, main.cpp:240-292, is not used: its 0 and 2 corners
   get the same flags, some of its 1 corners get 0 or 2).  Triangles are
   re-oriented as Grid::load does when the volume is negative. */
int rpd_grid_from_mesh(int g, const double *points, int num_points, const int32_t *triangles, int num_triangles,
                       int normalize, rpd_grid **out);
/* Grid::load / Grid::save (Grid.cpp:186-279, 290-344), reference text format. */
int rpd_grid_load(const char *path, rpd_grid **out);
int rpd_grid_save(const rpd_grid *grid, const char *path);
/* Read-only view of a host grid (pointers stay valid until rpd_grid_free). */
int rpd_grid_view(const rpd_grid *grid, rpd_grid_desc *out);
void rpd_grid_free(rpd_grid *grid);
/* Grid::valid (Grid.cpp:60-184) -> 1 valid, 0 invalid.  Its rule that no
   voxel has 8 border corners (Grid.cpp:147-167) is applied per voxel, to the
   flags of v + {0,1}^3 (voxels on a max face skipped); the reference loop
   counts the corners of voxel 0 for every border voxel instead. */
int rpd_grid_valid(const rpd_grid_desc *grid);
/* Grid::getVolume (Grid.cpp:30-35) and first moments (Grid.cpp:37-53). */
double rpd_grid_volume(const rpd_grid_desc *grid);
void rpd_grid_moments(const rpd_grid_desc *grid, double out4[4]);

/* ---- synthetic seeds (SplitMix64, u = (x>>11) * 2^-53) ------------------- */
enum rpd_seed_kind {
    RPD_SEEDS_UNIFORM = 0,        /* i.i.d. uniform in [0,1]^3                  */
    RPD_SEEDS_LATTICE = 1,        /* m^3 lattice (i+1/2)/m + U(+-jitter/m),
                                     first n in (k,j,i) order                  */
    RPD_SEEDS_BLUE_NOISE = 2      /* dart throwing, min distance param        */
};
/* xyzw[4n] in the [0,1000]^3 frame; weights w_i ~ U[0, wmax] (already in
   frame units, i.e. scaled x1e6 from the unit frame); p0 = lattice m or
   dart min distance (unit frame), p1 = lattice jitter fraction. */
int rpd_generate_seeds(int kind, int64_t n, uint64_t seed, double p0, double p1, double wmax, double *xyzw);

/* ---- solver context ------------------------------------------------------ */
typedef struct rpd_ctx rpd_ctx;
typedef struct rpd_domain rpd_domain;

/* Solve parameters: the public caps of VoroAlgo (voronoi.h:199-230). */
typedef struct {
    int32_t K, P, V, T;           /* caps of the first pass.  A solve may
                                     raise caps where that costs the kernels
                                     no residency and changes no result: T to
                                     at least 256, V to its vertex slots'
                                     room (64/128/256) and, with
                                     max_extra_steps > 0, P toward 48.  A
                                     cell's result does not depend on the cap
                                     it succeeds with; the caps a solve used
                                     are in rpd_stats.final_* / pass_caps    */
    int32_t precision;            /* enum rpd_precision                       */
    int32_t max_extra_steps;      /* re-queue passes over failed cells with
                                     grown caps (ref: 5)                      */
    int32_t apex_mode;            /* 0 reference corner apex, 1 (default)
                                     centre of the first voxel of the cell's
                                     bbox whose centre class is known         */
    int32_t security_mode;        /* 0 (default) exact power-diagram stop:
                                     d*(d - 2R) >= w_s - w_min; 1 reference
                                     stop d^2 > 4R^2 (ConvexCell.cl:364-382),
                                     which ignores weights; both equal at w=0 */
    int32_t cull;                 /* 1 (default): end a cell as soon as its
                                     circumsphere provably misses the domain
                                     (restricted volume 0); 0: clip it fully  */
    int32_t H;                    /* Hessian (VoroAlgo::m_H, voronoi.h:214-215;
                                     CHESSIAN, ConvexCell.cl:621-678, 1133-1151):
                                     up to H neighbours per seed, 0 (default)
                                     none                                     */
    int32_t HC;                   /* values per Hessian entry (m_HC): 1 face
                                     area, 4 face centroid times area (x, y, z)
                                     and area                                 */
    int32_t debug_counters;       /* non-zero: keep the debug counters of each
                                     pass (VoroAlgo::m_enableDCounters*,
                                     voronoi.h:93-102), read with
                                     rpd_read_debug_counters; 0 (default): none */
} rpd_params;

/* Per-solve statistics (VoroAlgoStat, voronoi.h:60-103). */
#define RPD_MAX_PASSES 8
typedef struct {
    double ms_bin, ms_knn, ms_clip, ms_fallback, ms_total; /* device time   */
    double ms_knn_first, ms_clip_first; /* kNN and clip launch of pass 1     */
    double ms_cells_first, ms_finish_first; /* pass 1 clip split: bisector
                                       loop kernel / restriction+integration */
    int32_t status_hist[RPD_STATUS_NUM];  /* final status per cell           */
    int32_t passes;                 /* clip passes run                       */
    int32_t fallback_cells;         /* cells recomputed in FP64 (mixed form) */
    int32_t final_K, final_P, final_V, final_T;
    int32_t dim;                    /* kNN grid resolution                   */
    int32_t queued_first;           /* pass 1 cells that reached the finish
                                       kernel (not culled, not failed)      */
    int64_t n;                      /* seeds                                 */
    /* per pass (one VoroAlgoStat per step, voronoi.cpp:1269-1330); pass p >=
       RPD_MAX_PASSES-1 is accumulated into the last entry                   */
    int32_t pass_cells[RPD_MAX_PASSES];   /* cells entering the pass        */
    int32_t pass_caps[RPD_MAX_PASSES][4]; /* K, P, V, T of the pass          */
    double pass_ms_knn[RPD_MAX_PASSES], pass_ms_clip[RPD_MAX_PASSES];
    int64_t skipped_cells;          /* cells proven outside the domain before
                                       the kNN (cull = 1): volume 0          */
    double ms_cover;                /* device time of that test              */
} rpd_stats;

/* Replaces OpenCLContext::createContext (openCL.cpp:151-207) + VoroAlgo ctor
   (voronoi.cpp:1041-1073): one context = one HIP device + one stream. */
int rpd_create(int device, rpd_ctx **out);
void rpd_destroy(rpd_ctx *ctx);
const char *rpd_last_error(const rpd_ctx *ctx);
/* Default parameters (VoroAlgo ctor defaults K25 P35 V40 T100). */
void rpd_default_params(rpd_params *out);

/* Replaces VoroAlgo::setGrid / GridAccess::init (voronoi.cpp:1079-1082,
   78-144): copy the grid to HBM; it stays resident until rpd_domain_destroy.
   RPD_E_ARG for a malformed description: a missing array, a triangle point
   id or list entry out of range, offsets not starting at 0, decreasing or
   ending past num_tri_list. */
int rpd_domain_create(rpd_ctx *ctx, const rpd_grid_desc *grid, rpd_domain **out);
/* A domain may be destroyed before or after the context that created it; it
   may be used by any context of the same device. */
void rpd_domain_destroy(rpd_domain *dom);

/* Replaces VoroAlgo::setInputPoints(std::vector<real>) (voronoi.cpp:1098-1107)
   and the KNearests constructor's binning (knearests.cpp:48-222): copies n
   seeds xyzw (FP64, frame [0,1000]^3, w = power weight, sign as in
   ConvexCell.cl:324) to HBM and bins them.  Seeds can be released on return. */
int rpd_set_seeds(rpd_ctx *ctx, const double *xyzw, int64_t n);
/* Same, from a device pointer already in HBM (setInputPoints(cl_mem,int),
   voronoi.cpp:1088-1096). */
int rpd_set_seeds_device(rpd_ctx *ctx, const double *d_xyzw, int64_t n);

/* Replaces VoroAlgo::launch (voronoi.cpp:1244-1403): kNN + clip + restrict +
   integrate with the status re-queue.  dom may be NULL (domain [0,1000]^3). */
int rpd_solve(rpd_ctx *ctx, rpd_domain *dom, const rpd_params *prm, rpd_stats *stats);

/* Packed batch: nreq diagrams solved together, one launch of each kernel for
   all of their cells.  No counterpart in the reference, which runs one
   VoroAlgo::launch per diagram (voronoi.cpp:1244-1403) with the grid
   resident across launches (voronoi.h:126-131) and batches at most 100000
   cell ids of one diagram per kernel pair (voronoi.cpp:960-973).  Request r:
   ns[r] seeds xyzw[r] (host, copied) against doms[r] (NULL: [0,1000]^3;
   doms itself may be NULL).  Each request is binned and searched on its
   own, and each of its cells gives what rpd_solve of those seeds alone
   gives.  Afterwards the context holds the requests' cells back to back:
   rpd_read_moments / rpd_read_status return request 0's cells (original
   order), then request 1's, and so on (rpd_batch_range gives each range).
   rpd_solve and rpd_knn refuse such a context until rpd_set_seeds.  No
   Hessian (H = 0) and no bordering rings (rpd_compact_used with
   keep_distance > 0).  At most RPD_BATCH_MAX_REQUESTS requests. */
#define RPD_BATCH_MAX_REQUESTS 4096
int rpd_solve_batch(rpd_ctx *ctx, int32_t nreq, rpd_domain *const *doms, const double *const *xyzw, const int64_t *ns,
                    const rpd_params *prm, rpd_stats *stats);
/* Range [first, first + count) of request r's cells after rpd_solve_batch. */
int rpd_batch_range(const rpd_ctx *ctx, int32_t r, int64_t *first, int64_t *count);
/* count cells' moments (4 doubles each) from cell `first` of the last solve. */
int rpd_read_moments_range(rpd_ctx *ctx, int64_t first, int64_t count, double *out4);

/* Debug counters of pass `pass` (0 the first, k the k-th re-queue pass; the
   two re-queue chains add into the same k) of the last solve run with
   debug_counters set (the reference's debugCounter0..4, ConvexCell.cl:15-39,
   887-911, 981-982, 1128-1150), RPD_DCOUNT_INTS int32 in out:
   [RPD_DCOUNT_MAX + 0..4]  largest K used where the security radius was
                            reached, P and V after the neighbour loop,
                            triangle union, Hessian entries of a cell;
   [RPD_DCOUNT_HIST_K + k]  cells that reached their radius at neighbour k
                            (k <= 256; counting from the first pass of a
                            resumed cell);
   [RPD_DCOUNT_HIST_P + p]  cells with p planes after the neighbour loop;
   [RPD_DCOUNT_HIST_V + v]  cells with v vertices after it;
   [RPD_DCOUNT_HIST_T + t]  cells whose triangle union had t entries (the
                            last bin also counts larger unions).
   Cells the domain cull ends (cull = 1) count only where they reached the
   loop's end.  RPD_E_STATE without such a solve. */
#define RPD_DCOUNT_MAX 0
#define RPD_DCOUNT_HIST_K 8
#define RPD_DCOUNT_HIST_P (RPD_DCOUNT_HIST_K + 257)
#define RPD_DCOUNT_HIST_V (RPD_DCOUNT_HIST_P + 256)
#define RPD_DCOUNT_HIST_T (RPD_DCOUNT_HIST_V + 257)
#define RPD_DCOUNT_UNION_MAX 4096
#define RPD_DCOUNT_INTS (RPD_DCOUNT_HIST_T + RPD_DCOUNT_UNION_MAX + 1)
int rpd_read_debug_counters(rpd_ctx *ctx, int32_t pass, int32_t *out);

/* Outputs, valid until the next rpd_set_seeds/rpd_solve on this context. */
/* Replaces getBarycenters (voronoi.cpp:1173-1183) in the per-original-seed
   "mg" layout of test_voronoi.cpp:186-211: 4n doubles (m*x, m*y, m*z, m). */
int rpd_read_moments(rpd_ctx *ctx, double *out4n);
/* Final per-cell status, original seed order. */
int rpd_read_status(rpd_ctx *ctx, int32_t *out_n);
/* Binning permutation perm[stored] = original index (KNearests::getPermutation,
   knearests.h:101-103). */
int rpd_read_perm(rpd_ctx *ctx, int32_t *out_n);
/* Device pointer of the per-original-seed moments (getBarycentersCL). */
int rpd_moments_device(rpd_ctx *ctx, const double **out_ptr);
/* Replaces VoroAlgo::getHessians / getHessianIds (voronoi.cpp:1179-1232) for
   a solve with H > 0, per original seed: vals[n*H*HC] and ids[n*H], the
   neighbours (original indices) whose shared face keeps more than 0.2 of
   area in the frame [0,1000]^3, in the cell's plane order, -1 after the last
   (none for a seed that failed or has volume 0; H entries and status
   hessian_overflow when there are more).  The face values are restricted to
   the domain like the moments. */
int rpd_read_hessian(rpd_ctx *ctx, double *vals, int32_t *ids);
/* Device pointers of the same (valid until the next solve); H, HC of it. */
int rpd_hessian_device(rpd_ctx *ctx, const double **vals, const int32_t **ids, int32_t *H, int32_t *HC);

/* kNN only (KNearests::buildKnearests, knearests.cpp:238-315): K nearest of
   every seed, in stored order, ascending (d^2, original index); -1 padded.
   knn_precision 0: FP32 seed copy, 1: FP64. out: n*K int32, stored ids. */
int rpd_knn(rpd_ctx *ctx, int32_t K, int32_t knn_precision, int32_t *out_nk);

/* rpd_knn over device buffers, allocating nothing.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The same binning and kNN kernels as rpd_knn, in the same
   order: d_perm[n] receives the binning permutation (stored -> original,
   as rpd_read_perm), d_out_nk[n*K] the K nearest of every seed in stored
   order, ascending (d^2, original index), -1 padded.  d_xyzw: n seeds in
   HBM (FP64 xyzw, as rpd_set_seeds_device).  The working memory is a
   24 MiB device array of the library, one per GPU, held until the process
   ends: n <= 262144 and K <= 256, else RPD_E_LIMIT.  Every call on one GPU
   shares that GPU's array: calls that overlap on one GPU race.  Does not
   change the context's seeds or results; waits for its stream. */
int rpd_knn_device_noalloc(rpd_ctx *ctx, const double *d_xyzw, int64_t n, int32_t K, int32_t knn_precision,
                           int32_t *d_perm, int32_t *d_out_nk);

/* Compaction of used cells (compressResult, voronoi.cpp:513-691) in a
   deterministic order, on the device (a scan of the keep flags, then a
   scatter): cells with m >= keep_volume, ascending original index.
   out_count receives the number of used cells; out4 (4*count) and
   out_perm_to_user (count) may be NULL to query the count. */
int rpd_compact(rpd_ctx *ctx, double keep_volume, int64_t *out_count, double *out4, int32_t *out_perm_to_user);
/* Replaces compressResult with used cells and bordering rings (voronoi.cpp:
   513-691, 693-751, 1381-1393; CCUpdateUsedCell and compute_bordering,
   ConvexCell.cl:680-690, 754-807; VoroAlgo::m_keepDistance): the used cells
   (m >= keep_volume), the cells sharing a face of their unrestricted power
   cell (keep_distance >= 1), and so on for keep_distance rings; a cell of the
   last ring whose volume is below keep_volume is flagged with w =
   RPD_NO_CELL_LIMITS (a wall for the next solve).  keep_distance 0: the used
   cells only (as rpd_compact); < 0: every cell.  Ascending original index;
   out4 and out_perm_to_user may be NULL to query the count. */
int rpd_compact_used(rpd_ctx *ctx, double keep_volume, int32_t keep_distance, int64_t *out_count, double *out4,
                     int32_t *out_perm_to_user);
/* The same compaction left in HBM (getBarycentersCL and
   getPermutationToUserCL after compressResult, voronoi.h:150-160,
   voronoi.cpp:1140-1160): *out4 receives a device pointer to 4*count
   doubles, *out_perm_to_user one to count original indices; both stay valid
   until the next compaction, rpd_set_seeds or rpd_solve on this context. */
int rpd_compact_used_device(rpd_ctx *ctx, double keep_volume, int32_t keep_distance, int64_t *out_count,
                            const double **out4, const int32_t **out_perm_to_user);

/* ---- serving --------------------------------------------------------------
   A request loop for streams of solves with different seed counts and
   weights against resident domains (the inner step of Lloyd relaxation or
   semi-discrete transport).  The queued requests are packed into batches
   (rpd_solve_batch: one launch of each kernel over the cells of every
   request of the batch), in arrival order, as long as their parameters are
   equal and their footprints fit the HBM budget.  Up to max_inflight
   batches run at once, each on its own context and streams, so a batch is
   admitted while the previous one finishes its tail.  A request's result is
   the same alone and in any batch. */
typedef struct rpd_server rpd_server;
typedef struct {
    int64_t hbm_budget_bytes;     /* admission limit on the footprints in flight (default 64 GiB) */
    int32_t max_inflight;         /* batches running at once (default 2, at most 16)              */
    int32_t max_batch;            /* requests per batch (default 256, at most RPD_BATCH_MAX_REQUESTS) */
} rpd_server_config;

int rpd_server_create(rpd_ctx *ctx, const rpd_server_config *cfg, rpd_server **out);
void rpd_server_destroy(rpd_server *srv);
/* Enqueue a request; seeds are copied. Returns a ticket >= 0 in *ticket. */
int rpd_server_submit(rpd_server *srv, rpd_domain *dom, const double *xyzw, int64_t n, const rpd_params *prm,
                      int64_t *ticket);
/* Poll (wait=0) or block (wait=1) for a finished request; copies its 4n
   moments into out4n (caller sized from the submitted n) and returns the
   ticket in *done_ticket (-1 if none finished yet).  lat_ms: submit->done. */
int rpd_server_poll(rpd_server *srv, int wait, int64_t *done_ticket, double *out4n, int64_t cap_n, double *lat_ms);
/* Run the scheduler until every submitted request has finished. */
int rpd_server_drain(rpd_server *srv);

#ifdef __cplusplus
}
#endif
#endif
