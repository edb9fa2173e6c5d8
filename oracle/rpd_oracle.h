/*
 * rpd_oracle.h -- C interface of the CPU oracle for the restricted power
 * diagram hot path (grid kNN -> power-cell clipping -> restricted
 * volume/moment integration).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates, on the CPU, the algorithm
 * of the reference files knearests.cpp/knearests.cl, ConvexCell.cl and the
 * pass logic of voronoi.cpp.  It is the checker used by tests/, by the
 * smoke() entry point and by the CPU-baseline leg of bench.py.  The product
 * library (librpd) never links, loads or calls it.
 */
#ifndef RPD_ORACLE_H
#define RPD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Domain grid, same content as the reference's Grid (Grid.h:14-70). */
typedef struct {
    int size;                    /* voxels per axis g                          */
    const signed char *in_domain;/* g^3 corner flags 0 out / 1 border / 2 in   */
    const double *points;        /* 4*num_points, xyzw with w = 0              */
    int num_points;
    const int *triangles;        /* 3*num_triangles point ids                  */
    int num_triangles;
    const int *tri_list;         /* per-voxel sorted triangle ids              */
    int num_tri_list;
    const int *offsets;          /* g^3+1 offsets into tri_list                */
    const signed char *centre;   /* nullable: g^3 voxel-centre classes
                                    0 out / 2 in / 1 unknown                    */
} orc_grid;

typedef struct {
    int K, P, V, T;              /* caps of the first pass (voronoi.h:199-205) */
    int precision;               /* 0: float build, 1: double (USE_DOUBLE),
                                    2: long double (x87 80-bit, 64-bit
                                    significand) with the double pass policy
                                    and its own det4 predicate: a higher-
                                    precision evaluation of the same cells    */
    int vhs0, vhs1;              /* float build: half-space algo first/extra   */
    int apex_mode;               /* 0: reference corner rule (ConvexCell.cl:950-
                                    1056), 1: centre of the first voxel of the
                                    bbox range whose centre class is known,
                                    then the reference rule                    */
    int fix_exact_requeue;       /* 1: relaunch cells flagged only with
                                    needs_exact_predicates (reference drops them) */
    int max_extra_steps;         /* extra passes (reference: 5)                */
    int threads;                 /* OpenMP threads, <=0: all                   */
    int single_pass;             /* 1: no re-queue, one pass with T = T        */
    int security_mode;           /* 0: exact power stop d(d-2R) >= w_s - w_min;
                                    1: reference stop d^2 > 4R^2 (weights
                                    ignored, ConvexCell.cl:364-382)            */
    int cull;                    /* 1: the library's shortcuts for cells whose
                                    restriction is provably empty (the seed's
                                    cover test before the kNN, the circumradius
                                    test in the neighbour loop); 0: none, as
                                    the reference                              */
    int H, HC;                   /* Hessian (CHESSIAN): up to H neighbours per
                                    seed, HC = 1 face area, 4 face centroid x
                                    area and area; H = 0: none                 */
    int frame;                   /* 0: planes, vertices and moments in the
                                    [0,1000]^3 frame, as the reference; 1: in
                                    the frame of each cell's seed (the cell is
                                    translated by -seed, its moments back)     */
} orc_params;

typedef struct {
    long long n_pred;            /* plane-side predicate evaluations           */
    long long n_radius_vertex;   /* vertices visited by the security radius    */
    long long n_planes;          /* planes built (bisectors + cone planes)     */
    long long n_bbox_vertex;     /* vertices visited for the bbox              */
    long long n_integrated_vertex;/* vertices integrated                       */
    long long n_clip;            /* clip-by-plane calls                        */
    int status_hist[11];         /* final status histogram                     */
    int passes;                  /* number of clip passes run                  */
    int final_K, final_P, final_V, final_T; /* caps of the last pass           */
    double seconds_knn, seconds_clip, seconds_total;
    long long n_pred_restrict;   /* of n_pred: during the restriction          */
    long long n_planes_restrict; /* of n_planes: cone planes                   */
} orc_stats;

/* Grid kNN.  perm[stored] = original index; knn[stored*K + i] = stored index
   of the i-th neighbour, ascending by (d^2, original index), -1 padded.
   method 0: exact grid ring search, 1: brute force.  Returns dim (>0) or <0. */
int orc_knn(const double *xyzw, int n, int K, int precision, int method,
            int *perm, int *knn, int *counters_out /* nullable, dim^3+1 */);

/* kNN grid resolution used for n seeds. */
int orc_knn_dim(int n);

/* Full restricted power diagram with the reference pass policy.
   out_mg: 4*n (m*x, m*y, m*z, m) per ORIGINAL seed index.
   out_status: final status per original seed (nullable).
   out_kused: neighbours consumed per original seed (nullable). */
int orc_solve(const double *xyzw, int n, const orc_grid *grid,
              const orc_params *prm, double *out_mg, int *out_status,
              int *out_kused, orc_stats *st);
/* orc_solve plus the Hessian (prm->H > 0): out_hess[n*H*HC] and
   out_hids[n*H] per original seed, neighbours as original indices in plane
   order, -1 after the last. */
int orc_solve_hessian(const double *xyzw, int n, const orc_grid *grid,
                      const orc_params *prm, double *out_mg, int *out_status,
                      int *out_kused, orc_stats *st, double *out_hess,
                      int *out_hids);

/* Same as orc_solve for the cells of a subset of seeds only (ids_orig:
   original indices); every seed still acts as a neighbour.  out_mg, out_status
   and out_kused are indexed like ids_orig.  Used for bounded CPU timings. */
/* Face neighbours (original indices, ascending, -1 padded to max_nb) of the
   unrestricted FP64 power cells of the seeds ids_orig: the seeds whose
   bisector carries a vertex (CCUpdateUsedCell, ConvexCell.cl:680-690).
   0, or -1 bad arguments, -2 a cell failed, -3 more than max_nb. */
int orc_cell_neighbours(const double *xyzw, int n, const int *ids_orig, int n_ids, int K, int P, int V,
                        int security_mode, int *out_nb, int max_nb);
int orc_solve_ids(const double *xyzw, int n, const int *ids_orig, int n_ids,
                  const orc_grid *grid, const orc_params *prm, double *out_mg,
                  int *out_status, int *out_kused, orc_stats *st);

/* One cell, one pass, caller-given neighbour list (original indices of
   xyzw).  For unit tests of the clip.  Returns the status. */
int orc_clip_cell(const double *xyzw, int n, int seed, const int *neigs, int K,
                  int P, int V, int T, int precision, int vhs,
                  const orc_grid *grid, int apex_mode, int out_domain,
                  int security_mode, double wmin,
                  double *out4, int *nb_vertex, int *nb_plane);

/* Mesh-to-grid generator (utils/createGrid/main.cpp:150-294).
   orc_mesh_points: drops the points no triangle uses (main.cpp:164), then
   normalize != 0 scales into [4.998, 995.22]^3 (main.cpp:168-175); writes
   out_xyz[3*m] and out_tris[3*nt], returns m (<0 bad arguments).
   orc_mesh_lists: per-voxel lists of main.cpp:178-219 with the overlap test
   of geometry.cpp:433-508; returns the entry count and writes offsets (g^3+1)
   and list only when the count fits cap.
   orc_mesh_corner_flags: corner flags 1 within eps of the surface, else 2
   inside / 0 outside by the winding number (brute force). */
int orc_tri_box_overlap(const float centre[3], float half, const double *a, const double *b, const double *c);
double orc_point_triangle_distance(const double *p, const double *a, const double *b, const double *c);
double orc_winding_number(const double *p, const double *pts, const int *tris, int nt);
int orc_mesh_points(const double *points, int np, const int *tris, int nt, int normalize, double *out_xyz,
                    int *out_tris);
long long orc_mesh_lists(const double *xyz, const int *tris, int nt, int g, int *out_offsets, int *out_list,
                         long long cap);
void orc_mesh_corner_flags(const double *xyz, const int *tris, int nt, int g, double eps, signed char *out_flags);

/* Grid helpers restating Grid.cpp. */
double orc_grid_volume(const orc_grid *grid);              /* Grid.cpp:30-35 */
int orc_grid_valid(const orc_grid *grid);                  /* Grid.cpp:60-184 */
void orc_grid_moments(const orc_grid *grid, double out4[4]); /* Grid.cpp:37-53 */

#ifdef __cplusplus
}
#endif
#endif
