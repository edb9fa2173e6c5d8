// rpd_api.hip -- C-ABI of librpd: contexts, resident domains, seeds, the
// solve with its status re-queue, outputs.
//
// Pass policy (replaces VoroAlgo::launch, voronoi.cpp:1244-1376):
//   pass 1: every cell, caps K/P/V/T of rpd_params, in the requested form;
//   then up to max_extra_steps re-queue passes over the failed cells only,
//   which the kernels append to device lists:
//     * vertex/plane overflow -> that cap x1.5 (as the reference);
//     * security radius not reached -> K x2 (the reference: x1.5);
//     * triangle overflow -> T = the largest triangle union the finish
//       kernel counted, at least T x1.5 (the reference: x1.5), within the
//       LDS of one finish-kernel cell;
//     * needs_exact_predicates, inconsistent_boundary and
//       find_another_beginning_vertex from an FP32 form -> recomputed (the
//       reference drops cells flagged only for exact predicates,
//       voronoi.cpp:1349-1356; here they are always recomputed); from the
//       FP64 form they are final (a recomputation fails the same way);
//     * empty_cell is final (volume 0);
//   every re-queue pass runs in the FP64 form, in two chains that run at
//   once, each growing its own caps: the failures of pass 1's bisector
//   kernel (and theirs, and so on), whose first pass is queued before pass
//   1's finish kernel, and the failures of that finish kernel.  The host
//   drives both by polling their streams; small counters come back through
//   pinned memory, so that no copy waits for a busy stream.  A cell's result does not
//   depend on the caps it succeeds with.  A cell whose list ran out keeps its
//   state (planes, vertices, last neighbour) when there is room; the next
//   pass lists only the neighbours after that one and clips on from the
//   state, which is the clip sequence of one run with the longer list.
// Boundary cells are processed in the same pass as interior cells (the
// reference defers them with a T = 0 first pass and recomputes them).

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "rpd_internal.h"
#include "rpd_wave.h"

// A domain keeps no pointer to the context that made it: it may outlive that
// context and is shared by the contexts of one device.
struct rpd_domain {
    int device = 0;
    DomainDev dev;
    DevBuf flags, centre, pts, tris, tlist, toffs, tvoffs, tvox, tinfo, dist, dvox;
    int64_t bytes = 0;
};

// work buffers of one pass: neighbour lists, queued-cell records, queue length
struct PassBufs {
    DevBuf knn, rec, qcount;
};
// failed cells handed to the next pass: counters, then stored ids of the
// exact-predicate failures and the overflows (eo, 2n entries: exact ones
// down from the middle, overflows up from it) and of the resumable cells
struct FailBufs {
    DevBuf cnt, eo, sec;
};

// pinned host words of a context: queue lengths (one per stream), then the
// counters of each failure list, then the final status histogram
constexpr int kPinCounters = 16;
// the pinned word that receives a queued binning's crowded-cell flag
constexpr int kPinBinFlag = kPinCounters - 4;
// smallest triangle cap a solve uses
constexpr int kTriangleFloor = 256;
// the plane cap a solve raises a lower one to, where that costs no residency
constexpr int kPlaneFloor = 48;
constexpr int kPinHist = kPinCounters + 6 * kFailCounters;
constexpr int kPinWords = kPinHist + 16;
static_assert(RPD_STATUS_NUM <= 16, "status histogram room");
// blocks of the weight-minimum kernel
constexpr int kWminBlocks = 256;

struct rpd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;      // == streams[0]
    hipStream_t streams[4] = {nullptr, nullptr, nullptr, nullptr};  // logical streams (two may share one)
    int n_streams = 4;                                                 // streams created
    std::string err;
    int64_t n = 0;
    DevBuf seeds;  // original order, FP64 xyzw
    // binning (kNN precision bin_prec)
    int bin_prec = -1;
    int dim = 0;
    DevBuf perm, counters, stored, stored64, bin_scratch;
    // outputs, original order
    DevBuf out, status, kused;
    // Hessian of the last solve (H > 0): n*H*HC values, n*H neighbours
    DevBuf hess, hids;
    int hess_H = 0, hess_HC = 1;
    // used-cell marks (per stored id) and the scratch outputs of the
    // bordering passes; caps and stop test of the last solve
    DevBuf used, bout, bstatus;
    // compacted output (compressResult): per original index keep flag and
    // ring, their exclusive scan, the kept cells' values and original
    // indices, rocPRIM scan scratch
    DevBuf ckeep, cring, cpos, cout, cperm, ctmp;
    // debug counters of the last solve that asked for them: RPD_MAX_PASSES
    // blocks of kDcStride ints
    DevBuf dcnt;
    bool dcnt_ok = false;
    int64_t ccount = 0;  // cells in cout / cperm after the last compaction
    int last_K = 0, last_P = 0, last_V = 0, last_sm = 0, last_steps = 0;
    // cells the cover test kept in the last solve over kept_hint_n seeds
    // (-1: none): the size of the next pass-1 kNN launch, which is queued
    // before the count is known
    int64_t kept_hint = -1, kept_hint_n = -1;
    // packed batch set by rpd_solve_batch (empty: the seeds are one
    // diagram): per request its domain, stored-id range, weight minima, kNN
    // grid size and counters offset; on the device the per-request fields
    // of the clip kernels and of the kNN, and the request of each stored id
    struct BatchReq {
        rpd_domain *dom = nullptr;
        int64_t base = 0, n = 0;
        double wmin = 0, wall = 0;
        int dim = 0;
        int64_t coff = 0;
    };
    std::vector<BatchReq> batch;
    DevBuf req_desc, knn_req, req_of;
    // pass work buffers: up to three passes may run at once, one per stream
    PassBufs pb[4];
    // pass 1 writes lists 0 (bisector kernel) and 1 (finish kernel); the
    // re-queue chains 2, 3 and 4, 5
    FailBufs fl[6];
    DevBuf knn, ids1, cnt, wred, cull_scratch;
    // states of resumable cells (two record buffers per chain, their
    // counters) and, per stored id, the record slot and the last neighbour
    DevBuf rs[4], rs_cnt, rslot, lastnb;
    // record slots of pass 1's finish-kernel failures (laid out as their list)
    DevBuf fslot;
    bool solved = false;
    hipEvent_t ev[4][4] = {};  // per stream: before kNN, after kNN, after the bisector kernel, after the finish
    hipEvent_t dep[3] = {};    // per chain: orders work between its streams; [2]: pass 1's queue length copied
    hipEvent_t evb[2] = {};    // a solve's binning (and weight minima): before, after
    int *pin = nullptr;        // kPinWords pinned host words
    int *pin_dev = nullptr;    // their device address (nullptr: copies go through the copy engine)
    double *pin_red = nullptr;  // 2 x kWminBlocks pinned doubles: per-block weight minima
    // smallest weight of the non-border seeds and of all seeds, for the
    // current seeds (valid once wmin_ok)
    bool wmin_ok = false;
    double wmin = 0, wall = 0;
};

// ---- working memory of rpd_knn_device_noalloc -----------------------------
// A file-scope device array: every GPU the library runs on holds its own copy
// of it until the process ends.  Host code must never pass `scratch` itself
// to a kernel or a copy (in host code the name is a stand-in, not a device
// address): the address of each device's copy is resolved once, with that
// device current, by rpd_create, and kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(24) << 20;
constexpr int64_t kNoAllocMaxSeeds = int64_t(1) << 18;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

namespace {

// Up to 12 device int arrays cleared by one single-workgroup launch: the
// counters a solve starts from, without one fill launch each.
struct ZeroSet {
    int *p[12];
    int n[12];
    int k;
};
__global__ void zero_words_kernel(ZeroSet z) {
    for (int j = 0; j < z.k; ++j)
        for (int i = int(threadIdx.x); i < z.n[j]; i += int(blockDim.x)) z.p[j][i] = 0;
}

// Resolves the current device's copy of `scratch` once (device current).
int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return RPD_E_DEVICE;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return RPD_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return RPD_E_DEVICE;
    g_scratch_ptr[device] = p;
    return RPD_OK;
}

int set_err(rpd_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) return set_err(ctx, RPD_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_));  \
    } while (0)

#define RPDCHK(ctx, call, what)                                          \
    do {                                                                 \
        int r_ = (call);                                                 \
        if (r_ != RPD_OK) return set_err(ctx, r_, "%s failed (%d)", what, r_); \
    } while (0)

// per-block minimum of the weights of non-border seeds (out[b]) and of all
// weights (out[gridDim.x + b])
__global__ __launch_bounds__(256) void wmin_kernel(const double4 *__restrict__ seeds, int64_t n,
                                                   double *__restrict__ out) {
    __shared__ double red[2][256];
    double m = INFINITY, ma = INFINITY;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double w = seeds[i].w;
        if (w > RPD_NO_CELL_LIMITS) m = fmin(m, w);
        ma = fmin(ma, w);
    }
    red[0][threadIdx.x] = m;
    red[1][threadIdx.x] = ma;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[blockIdx.x] = red[0][0];
        out[gridDim.x + blockIdx.x] = red[1][0];
    }
}

// The two minima over the nb blocks' partials (non-border seeds, all seeds)
// to out[0], out[1]: for the kernels queued before the host has them.
__global__ __launch_bounds__(256) void wmin_final_kernel(const double *__restrict__ part, int nb,
                                                         double *__restrict__ out) {
    __shared__ double red[2][256];
    double m = INFINITY, ma = INFINITY;
    for (int i = threadIdx.x; i < nb; i += 256) {
        m = fmin(m, part[i]);
        ma = fmin(ma, part[nb + i]);
    }
    red[0][threadIdx.x] = m;
    red[1][threadIdx.x] = ma;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        out[1] = red[1][0];
    }
}

// Chebyshev (26-neighbour BFS) voxel distance to the voxels that hold
// triangles or whose centre is not known to be outside, capped at 255.
void domain_distance(const rpd_grid_desc *g, std::vector<uint8_t> &dist) {
    const int G = g->size;
    const size_t nv = size_t(G) * G * G;
    dist.assign(nv, 255);
    std::vector<int> q;
    q.reserve(nv / 8 + 16);
    for (size_t v = 0; v < nv; ++v) {
        const bool tri = g->offsets[v + 1] > g->offsets[v];
        const bool in = g->centre ? g->centre[v] != 0 : g->in_domain[v] != 0;
        if (tri || in) {
            dist[v] = 0;
            q.push_back(int(v));
        }
    }
    for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h];
        const int d = dist[size_t(v)];
        if (d >= 254) continue;
        const int x = v % G, y = (v / G) % G, z = v / (G * G);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int X = x + dx, Y = y + dy, Z = z + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= G || Y >= G || Z >= G) continue;
                    const size_t u = size_t(X) + size_t(Y) * G + size_t(Z) * G * G;
                    if (dist[u] != 255) continue;
                    dist[u] = uint8_t(d + 1);
                    q.push_back(int(u));
                }
    }
}

// The lanes of a wave that hold one status add their count once (a ballot
// per distinct status in the wave: nearly every cell ends with the same one),
// then each workgroup adds its counts to out.  (One LDS atomic per cell
// serialised on the one bin nearly every cell hits.)
__global__ __launch_bounds__(256) void status_hist_kernel(const int *__restrict__ status, int64_t n,
                                                          int *__restrict__ out) {
    __shared__ int hist[RPD_STATUS_NUM];
    if (threadIdx.x < RPD_STATUS_NUM) hist[threadIdx.x] = 0;
    __syncthreads();
    const int lane = rpd::lane_id();
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const int st = status[i];
        const bool ok = st >= 0 && st < RPD_STATUS_NUM;
        uint64_t todo = rpd::ballot(ok);
        while (todo) {
            const int l = rpd::lowest_bit(todo);
            const int v = __builtin_amdgcn_readlane(st, l);
            const uint64_t m = rpd::ballot(ok && st == v);
            if (lane == l) atomicAdd(&hist[v], __popcll(m));
            todo &= ~m;
        }
    }
    __syncthreads();
    if (threadIdx.x < RPD_STATUS_NUM && hist[threadIdx.x]) atomicAdd(&out[threadIdx.x], hist[threadIdx.x]);
}

// used[s] = 1 for the cells whose restricted volume reaches keep
// (ConvexCell.cl:1118-1121), else 0; s is the stored id
__global__ __launch_bounds__(256) void used_init_kernel(const double4 *__restrict__ out, const int *__restrict__ perm,
                                                        int64_t n, double keep, int *__restrict__ used) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s < n) used[s] = out[perm[s]].w >= keep ? 1 : 0;
}

// Compaction marks, per original index i: keep[i] = 1 when the cell is kept
// (ring[i] = its mark, used == nullptr: every cell, mark 1), 0 otherwise;
// keep[n] = 0 so that the scan's last entry is the count.
__global__ __launch_bounds__(256) void compact_mark_kernel(const int *__restrict__ used, const int *__restrict__ perm,
                                                           int64_t n, int *__restrict__ keep,
                                                           int *__restrict__ ring) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s < n) {
        const int u = used ? used[s] : 1;
        const int o = perm[s];
        keep[o] = u != 0 ? 1 : 0;
        ring[o] = u;
    } else if (s == n) {
        keep[n] = 0;
    }
}

// compact_result (voronoi.cpp:582-691) in ascending original index: the kept
// cells' moments and original indices at their scanned positions.  Marks are
// those of the bordering passes: 1 for the used cells and their face
// neighbours, d + 1 for the neighbours ring d adds; a cell marked last (the
// keep distance) whose volume is below keep gets w = NO_CELL_LIMITS (the
// reference's MAX_DIST test); last < 0: no flags.
__global__ __launch_bounds__(256) void compact_scatter_kernel(const double4 *__restrict__ out,
                                                              const int *__restrict__ keepf,
                                                              const int *__restrict__ ring,
                                                              const int *__restrict__ pos, int64_t n, double keep,
                                                              int last, double4 *__restrict__ cout,
                                                              int *__restrict__ cperm) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n || !keepf[i]) return;
    double4 v = out[i];
    if (last >= 0 && v.w < keep && ring[i] == last) v.w = RPD_NO_CELL_LIMITS;
    const int p = pos[i];
    cout[p] = v;
    cperm[p] = int(i);
}

// The cells ring k expands (find_bordering, voronoi.cpp:693-741): ring 0 the
// used cells, ring k >= 1 the cells marked k whose volume is below keep.
__global__ __launch_bounds__(256) void ring_select_kernel(const int *__restrict__ used,
                                                          const double4 *__restrict__ out, const int *__restrict__ perm,
                                                          int64_t n, double keep, int k, int *__restrict__ ids,
                                                          int *__restrict__ count) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= n) return;
    const double w = out[perm[s]].w;
    const bool sel = k == 0 ? (used[s] == 1 && w >= keep) : (used[s] == k && w < keep);
    if (sel) ids[atomicAdd(count, 1)] = int(s);
}

// One wave per queued cell record (FP64 planes): the neighbours of the
// planes of its vertices that carry no mark get d (CCUpdateUsedCell,
// ConvexCell.cl:680-690).  Every writer of a ring writes the same d.
__global__ __launch_bounds__(256) void mark_neighbours_kernel(const char *__restrict__ rec, int stride, int P,
                                                              int S, int nq, int *__restrict__ used, int d) {
    const int q = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
    const int lane = int(threadIdx.x & 63);
    if (q >= nq) return;
    const char *r = rec + size_t(q) * size_t(stride);
    const int nvalid = reinterpret_cast<const int *>(r)[2];  // RecHeader{sid, nb_plane, nvalid, kused}
    const uint32_t *tri = reinterpret_cast<const uint32_t *>(r + 16 + size_t(P) * 32);
    const int *nid = reinterpret_cast<const int *>(tri + 64 * S);
    for (int j = lane; j < nvalid; j += 64) {
        const uint32_t t = tri[j];
        for (int k = 0; k < 3; ++k) {
            const int nb = nid[(t >> (8 * k)) & 0xff];
            if (nb >= 0 && used[nb] == 0) used[nb] = d;
        }
    }
}

// Bins the seeds in key precision prec unless they are.  With pending set
// the binning is only queued: *pending tells whether it was, and the caller,
// once it has synchronised the stream, finishes it with finish_binning.
int ensure_binning(rpd_ctx *c, int prec, bool *pending = nullptr) {
    if (pending) *pending = false;
    if (c->bin_prec == prec) return RPD_OK;
    const int64_t n = c->n;
    const int dim = rpd_knn_dim(n);
    const int64_t nc = int64_t(dim) * dim * dim;
    if (nc + 1 > INT32_MAX) return set_err(c, RPD_E_LIMIT, "kNN grid too large");
    RPDCHK(c, c->perm.ensure(size_t(n) * 4), "alloc perm");
    RPDCHK(c, c->counters.ensure(size_t(nc + 1) * 4), "alloc counters");
    RPDCHK(c, c->stored.ensure(size_t(n) * (prec == 0 ? 16 : 32)), "alloc stored");
    RPDCHK(c, c->stored64.ensure(size_t(n) * 32), "alloc stored64");
    if (pending) {
        RPDCHK(c,
               rpd_bin_seeds_async(c->stream, c->seeds.as<double4>(), n, prec, dim, c->bin_scratch,
                                   c->pin + kPinBinFlag, c->perm.as<int>(), c->counters.as<int>(), c->stored.ptr,
                                   c->stored64.as<double4>()),
               "binning");
        *pending = true;
    } else {
        RPDCHK(c,
               rpd_bin_seeds(c->stream, c->seeds.as<double4>(), n, prec, dim, c->bin_scratch, c->perm.as<int>(),
                             c->counters.as<int>(), c->stored.ptr, c->stored64.as<double4>()),
               "binning");
    }
    c->bin_prec = prec;
    c->dim = dim;
    return RPD_OK;
}
// After the stream is synchronised: the crowded-cell path of a queued binning
// when its flag is set
int finish_binning(rpd_ctx *c) {
    if (!c->pin[kPinBinFlag]) return RPD_OK;
    RPDCHK(c,
           rpd_bin_seeds_redo(c->stream, c->seeds.as<double4>(), c->n, c->bin_prec, c->dim, c->bin_scratch,
                              c->perm.as<int>(), c->stored.ptr, c->stored64.as<double4>()),
           "binning (crowded cells)");
    return RPD_OK;
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

}  // namespace

// device of a context (used by the serving loop to open its slots)
int rpd_ctx_device(const rpd_ctx *c) { return c ? c->device : 0; }

// A context on n_streams HIP streams (4 or 2).  Its passes are laid out on 4
// logical streams: 0 pass 1 and the finish failures' chain, 1 and 2 the
// bisector failures' chain (resumed and recomputed cells), 3 the finish
// failures' recomputed cells.  With 2 streams, 0 and 3 share one, 1 and 2
// the other: pass 1's finish kernel still overlaps the bisector failures'
// chain, and a serving slot holds 2 of the device's hardware queues.
int rpd_create_streams(int device, int n_streams, rpd_ctx **out);

extern "C" {

int rpd_create(int device, rpd_ctx **out) { return rpd_create_streams(device, 4, out); }

}  // extern "C"

int rpd_create_streams(int device, int n_streams, rpd_ctx **out) {
    if (!out || (n_streams != 2 && n_streams != 4)) return RPD_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RPD_E_DEVICE;
    if (device < 0 || device >= ndev) return RPD_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return RPD_E_DEVICE;
    if (resolve_scratch(device) != RPD_OK) return RPD_E_DEVICE;
    auto *c = new rpd_ctx();
    c->device = device;
    for (int i = 0; i < n_streams; ++i)
        if (hipStreamCreateWithFlags(&c->streams[i], hipStreamNonBlocking) != hipSuccess) {
            rpd_destroy(c);
            return RPD_E_DEVICE;
        }
    if (n_streams == 2) {
        c->streams[2] = c->streams[1];
        c->streams[3] = c->streams[0];
    }
    c->n_streams = n_streams;
    c->stream = c->streams[0];
    for (auto &es : c->ev)
        for (auto &e : es)
            if (hipEventCreate(&e) != hipSuccess) {
                rpd_destroy(c);
                return RPD_E_DEVICE;
            }
    for (auto &e : c->dep)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            rpd_destroy(c);
            return RPD_E_DEVICE;
        }
    for (auto &e : c->evb)
        if (hipEventCreate(&e) != hipSuccess) {
            rpd_destroy(c);
            return RPD_E_DEVICE;
        }
    if (hipHostMalloc(reinterpret_cast<void **>(&c->pin), kPinWords * sizeof(int), hipHostMallocDefault) !=
        hipSuccess) {
        c->pin = nullptr;
        rpd_destroy(c);
        return RPD_E_DEVICE;
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&c->pin_dev), c->pin, 0) != hipSuccess) {
        (void)hipGetLastError();
        c->pin_dev = nullptr;
    }
    if (hipHostMalloc(reinterpret_cast<void **>(&c->pin_red), 2 * kWminBlocks * sizeof(double),
                      hipHostMallocDefault) != hipSuccess) {
        c->pin_red = nullptr;
        rpd_destroy(c);
        return RPD_E_DEVICE;
    }
    *out = c;
    return RPD_OK;
}

extern "C" {

void rpd_destroy(rpd_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    for (int i = 0; i < c->n_streams; ++i)
        if (c->streams[i]) hipStreamSynchronize(c->streams[i]);
    for (auto &es : c->ev)
        for (auto &e : es)
            if (e) hipEventDestroy(e);
    for (auto &e : c->dep)
        if (e) hipEventDestroy(e);
    for (auto &e : c->evb)
        if (e) hipEventDestroy(e);
    for (int i = 0; i < c->n_streams; ++i)
        if (c->streams[i]) hipStreamDestroy(c->streams[i]);
    if (c->pin) hipHostFree(c->pin);
    if (c->pin_red) hipHostFree(c->pin_red);
    delete c;
}

const char *rpd_last_error(const rpd_ctx *c) { return c ? c->err.c_str() : "null context"; }

void rpd_default_params(rpd_params *p) {
    if (!p) return;
    // VoroAlgo constructor defaults (voronoi.cpp:1041-1073)
    p->K = 25;
    p->P = 35;
    p->V = 40;
    p->T = 100;
    p->precision = RPD_PREC_MIXED;
    p->max_extra_steps = 5;
    p->apex_mode = 1;
    p->security_mode = 0;
    p->cull = 1;
    p->H = 0;  // no Hessian (VoroAlgo::m_H, voronoi.cpp:1048)
    p->HC = 1;
    p->debug_counters = 0;
}

int rpd_domain_create(rpd_ctx *c, const rpd_grid_desc *g, rpd_domain **out) {
    if (!c || !g || !out || g->size <= 0) return set_err(c, RPD_E_ARG, "rpd_domain_create: bad arguments");
    if (g->size > 1024) return set_err(c, RPD_E_LIMIT, "rpd_domain_create: at most 1024 voxels per axis");
    // the lists are read by index on the host and on the GPU: every range
    // and id must lie inside its array
    if (const char *why = rpd_grid_desc_check(g))
        return set_err(c, RPD_E_ARG, "rpd_domain_create: malformed grid: %s", why);
    hipSetDevice(c->device);
    auto *d = new rpd_domain();
    d->device = c->device;
    const size_t nv = size_t(g->size) * g->size * g->size;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
        if (bytes == 0) bytes = 4;
        if (b.ensure(bytes) != RPD_OK) return RPD_E_NOMEM;
        if (src && hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return RPD_E_DEVICE;
        d->bytes += int64_t(bytes);
        return RPD_OK;
    };
    std::vector<double> pts4(size_t(g->num_points) * 4);
    for (int i = 0; i < g->num_points; ++i)
        for (int k = 0; k < 4; ++k) pts4[size_t(i) * 4 + k] = k < 3 ? g->points[4 * size_t(i) + k] : 0.0;
    int r = up(d->flags, g->in_domain, nv);
    if (r == RPD_OK && g->centre) r = up(d->centre, g->centre, nv);
    if (r == RPD_OK) r = up(d->pts, pts4.data(), pts4.size() * 8);
    if (r == RPD_OK) r = up(d->tris, g->triangles, size_t(g->num_triangles) * 12);
    if (r == RPD_OK) r = up(d->tlist, g->tri_list, size_t(g->num_tri_list) * 4);
    if (r == RPD_OK) r = up(d->toffs, g->offsets, (nv + 1) * 4);
    // per triangle, the voxels whose list holds it, ascending (the finish
    // kernel keeps a list entry only in the first voxel of a cell's box)
    std::vector<int> tvoffs(size_t(g->num_triangles) + 1, 0);
    std::vector<uint32_t> tvox(size_t(std::max(g->num_tri_list, 0)));
    for (size_t v = 0; v < nv && r == RPD_OK; ++v)
        for (int q = g->offsets[v]; q < g->offsets[v + 1]; ++q) {
            const int t = g->tri_list[q];
            if (t < 0 || t >= g->num_triangles) {
                r = RPD_E_ARG;
                break;
            }
            tvoffs[size_t(t) + 1]++;
        }
    for (size_t t = 0; t < size_t(g->num_triangles); ++t) tvoffs[t + 1] += tvoffs[t];
    if (r == RPD_OK) {
        std::vector<int> cur(tvoffs.begin(), tvoffs.end() - 1);
        const size_t gg = size_t(g->size);
        for (size_t v = 0; v < nv; ++v)
            for (int q = g->offsets[v]; q < g->offsets[v + 1]; ++q)
                tvox[size_t(cur[size_t(g->tri_list[q])]++)] =
                    (uint32_t(v / (gg * gg)) << 20) | (uint32_t((v / gg) % gg) << 10) | uint32_t(v % gg);
    }
    // per list entry of voxel v: v, and which voxels within 2 of v per axis
    // that precede it in scan order also list the triangle (64-bit mask, bit
    // (dz+2)*25 + (dy+2)*5 + (dx+2), below 62 for a preceding voxel); bit 63
    // when a preceding voxel listing it is farther away
    std::vector<uint32_t> tinfo(4 * size_t(std::max(g->num_tri_list, 0)));
    if (r == RPD_OK) {
        const size_t gg = size_t(g->size);
        for (size_t v = 0; v < nv; ++v) {
            const int vx = int(v % gg), vy = int((v / gg) % gg), vz = int(v / (gg * gg));
            const uint32_t pv = (uint32_t(vz) << 20) | (uint32_t(vy) << 10) | uint32_t(vx);
            for (int q = g->offsets[v]; q < g->offsets[v + 1]; ++q) {
                const int t = g->tri_list[q];
                uint64_t m = 0;
                for (int u = tvoffs[size_t(t)]; u < tvoffs[size_t(t) + 1]; ++u) {
                    const uint32_t w = tvox[size_t(u)];
                    if (w >= pv) break;
                    const int dx = int(w & 1023u) - vx, dy = int((w >> 10) & 1023u) - vy, dz = int(w >> 20) - vz;
                    if (dx < -2 || dx > 2 || dy < -2 || dy > 2 || dz < -2 || dz > 2)
                        m |= 1ull << 63;
                    else
                        m |= 1ull << ((dz + 2) * 25 + (dy + 2) * 5 + (dx + 2));
                }
                tinfo[4 * size_t(q)] = pv;
                tinfo[4 * size_t(q) + 1] = uint32_t(m);
                tinfo[4 * size_t(q) + 2] = uint32_t(m >> 32);
                tinfo[4 * size_t(q) + 3] = 0;
            }
        }
    }
    if (r == RPD_OK) r = up(d->tvoffs, tvoffs.data(), tvoffs.size() * 4);
    if (r == RPD_OK) r = up(d->tvox, tvox.data(), tvox.size() * 4);
    if (r == RPD_OK) r = up(d->tinfo, tinfo.data(), tinfo.size() * 4);
    std::vector<uint8_t> dist;
    domain_distance(g, dist);
    if (r == RPD_OK) r = up(d->dist, dist.data(), nv);
    std::vector<int> dvox;
    for (size_t v = 0; v < nv; ++v)
        if (dist[v] == 0) dvox.push_back(int(v));
    if (r == RPD_OK) r = up(d->dvox, dvox.data(), dvox.size() * 4);
    if (r == RPD_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = RPD_E_DEVICE;
    if (r != RPD_OK) {
        delete d;
        return set_err(c, r, "rpd_domain_create: upload failed");
    }
    d->dev.g = g->size;
    d->dev.vs = 1000.0 / double(g->size);
    d->dev.flags = d->flags.as<int8_t>();
    d->dev.centre = g->centre ? d->centre.as<int8_t>() : nullptr;
    d->dev.pts = d->pts.as<double4>();
    d->dev.tris = d->tris.as<int>();
    d->dev.tlist = d->tlist.as<int>();
    d->dev.toffs = d->toffs.as<int>();
    d->dev.tvoffs = d->tvoffs.as<int>();
    d->dev.tvox = d->tvox.as<uint32_t>();
    d->dev.tinfo = d->tinfo.as<uint4>();
    d->dev.dist = d->dist.as<uint8_t>();
    d->dev.dvox = d->dvox.as<int>();
    d->dev.n_dvox = int(dvox.size());
    for (size_t i = 0; i < nv; ++i)
        if (g->in_domain[i] == 1) {
            d->dev.has_border_vertices = true;
            break;
        }
    *out = d;
    return RPD_OK;
}

void rpd_domain_destroy(rpd_domain *d) {
    if (!d) return;
    // wait for the work of any context that may still read the grid
    hipSetDevice(d->device);
    hipDeviceSynchronize();
    delete d;
}

static int set_seeds_common(rpd_ctx *c, const double *src, int64_t n, hipMemcpyKind kind) {
    if (!c || !src || n < 2) return set_err(c, RPD_E_ARG, "rpd_set_seeds: need at least 2 seeds");
    if (n > INT32_MAX / 2) return set_err(c, RPD_E_LIMIT, "rpd_set_seeds: too many seeds");
    hipSetDevice(c->device);
    RPDCHK(c, c->seeds.ensure(size_t(n) * 32), "alloc seeds");
    HIPCHK(c, hipMemcpyAsync(c->seeds.ptr, src, size_t(n) * 32, kind, c->stream));
    c->n = n;
    c->bin_prec = -1;
    c->wmin_ok = false;
    c->solved = false;
    c->batch.clear();
    return RPD_OK;
}

int rpd_set_seeds(rpd_ctx *c, const double *xyzw, int64_t n) {
    return set_seeds_common(c, xyzw, n, hipMemcpyHostToDevice);
}

int rpd_set_seeds_device(rpd_ctx *c, const double *d_xyzw, int64_t n) {
    return set_seeds_common(c, d_xyzw, n, hipMemcpyDeviceToDevice);
}

int rpd_knn(rpd_ctx *c, int32_t K, int32_t prec, int32_t *out_nk) {
    if (!c || c->n < 2 || K <= 0 || !out_nk || (prec != 0 && prec != 1)) return set_err(c, RPD_E_ARG, "rpd_knn");
    if (K > 256) return set_err(c, RPD_E_LIMIT, "rpd_knn: K > 256");
    if (!c->batch.empty()) return set_err(c, RPD_E_STATE, "rpd_knn: the seeds are a packed batch");
    hipSetDevice(c->device);
    RPDCHK(c, ensure_binning(c, prec), "binning");
    RPDCHK(c, c->knn.ensure(size_t(c->n) * K * 4), "alloc knn");
    RPDCHK(c,
           rpd_launch_knn(c->stream, prec, c->stored.ptr, c->perm.as<int>(), c->counters.as<int>(), c->dim, nullptr,
                          c->n, K, c->knn.as<int>()),
           "knn launch");
    HIPCHK(c, hipMemcpyAsync(out_nk, c->knn.ptr, size_t(c->n) * K * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_knn_device_noalloc(rpd_ctx *c, const double *d_xyzw, int64_t n, int32_t K, int32_t prec,
                           int32_t *d_perm, int32_t *d_out_nk) {
    if (!c || !d_xyzw || !d_perm || !d_out_nk || n < 2 || K <= 0 || (prec != 0 && prec != 1))
        return set_err(c, RPD_E_ARG, "rpd_knn_device_noalloc: bad arguments");
    if (K > 256) return set_err(c, RPD_E_LIMIT, "rpd_knn_device_noalloc: K > 256");
    if (n > kNoAllocMaxSeeds)
        return set_err(c, RPD_E_LIMIT, "rpd_knn_device_noalloc: at most %lld seeds", (long long)kNoAllocMaxSeeds);
    HIPCHK(c, hipSetDevice(c->device));
    char *base = static_cast<char *>(g_scratch_ptr[c->device]);
    if (!base) return set_err(c, RPD_E_STATE, "rpd_knn_device_noalloc: scratch not resolved");
    const int dim = rpd_knn_dim(n);
    const int64_t nc = int64_t(dim) * dim * dim;
    // scratch: counters (dim^3 + 1) | sorted seeds (FP32 or FP64) | binning
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t off_stored = up(size_t(nc + 1) * 4);
    const size_t off_bin = off_stored + up(size_t(n) * (prec == 0 ? 16 : 32));
    const size_t bin_bytes = rpd_bin_scratch_bytes(c->stream, n, dim);
    if (bin_bytes == 0) return set_err(c, RPD_E_DEVICE, "rpd_knn_device_noalloc: scratch size query");
    if (off_bin + bin_bytes > kNoAllocScratchBytes)
        return set_err(c, RPD_E_LIMIT, "rpd_knn_device_noalloc: %zu scratch bytes needed, %zu available",
                       off_bin + bin_bytes, kNoAllocScratchBytes);
    int *counters = reinterpret_cast<int *>(base);
    void *stored = base + off_stored;
    // the crowded-cell flag comes back through the context's pinned words
    RPDCHK(c,
           rpd_bin_seeds_in(c->stream, reinterpret_cast<const double4 *>(d_xyzw), n, prec, dim, base + off_bin,
                            kNoAllocScratchBytes - off_bin, c->pin + kPinCounters - 2, d_perm, counters, stored,
                            nullptr),
           "rpd_knn_device_noalloc: binning");
    RPDCHK(c, rpd_launch_knn(c->stream, prec, stored, d_perm, counters, dim, nullptr, n, K, d_out_nk),
           "rpd_knn_device_noalloc: knn launch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

// The solve of the context's seeds: one diagram (against dom), or the packed
// batch rpd_solve_batch has set up (dom null, each request's domain in its
// descriptor).
static int solve_impl(rpd_ctx *c, rpd_domain *dom, const rpd_params *prm, rpd_stats *stats) {
    if (!c || !prm) return set_err(c, RPD_E_ARG, "rpd_solve: bad arguments");
    const bool batch = !c->batch.empty();
    if (batch && (dom || prm->H > 0))
        return set_err(c, RPD_E_ARG, "rpd_solve_batch: no Hessian, domains per request");
    if (c->n < 2) return set_err(c, RPD_E_STATE, "rpd_solve: no seeds");
    if (prm->precision < 0 || prm->precision > 2) return set_err(c, RPD_E_ARG, "rpd_solve: bad precision");
    if (prm->K <= 0 || prm->K > 256 || prm->P < 7 || prm->P > 255 || prm->V < 8 || prm->V > 256 || prm->T < 0)
        return set_err(c, RPD_E_ARG, "rpd_solve: caps out of range (0<K<=256, 7<=P<=255, 8<=V<=256, T>=0)");
    if (prm->H < 0 || prm->H > 4096 || (prm->H > 0 && prm->HC != 1 && prm->HC != 4))
        return set_err(c, RPD_E_ARG, "rpd_solve: Hessian H in [0, 4096], HC 1 or 4");
    // a domain is read-only device memory: any context on its device may use it
    if (dom && dom->device != c->device) return set_err(c, RPD_E_ARG, "rpd_solve: domain on another device");
    hipSetDevice(c->device);
    const int64_t n = c->n;
    const int form0 = prm->precision;
    const int kprec = form0 == RPD_PREC_FP64 ? 1 : 0;
    int K = std::min<int64_t>(prm->K, n - 1), P = prm->P, V = prm->V, T = prm->T;
    // The triangle cap is raised to kTriangleFloor where it is lower: a cell
    // done by one wave keeps 4 bytes of LDS per union entry, which costs the
    // finish kernel no residency, and a cell's result does not depend on the
    // cap it succeeds with; a lower cap only sends cells to re-queue passes.
    const bool hess = prm->H > 0;
    // The plane cap likewise, toward kPlaneFloor as far as the bisector and
    // finish kernels keep their residency: the restriction's cone clips add
    // up to 4 planes to a cell, which pushed the cells closest to the cap
    // into an FP64 finish pass of their own (6 cells at P = 40 at the 1M
    // torus, which also grew the bisector failures' pass to P = 60).  Only
    // with re-queue passes allowed: without them an overflow is the answer.
    if (prm->max_extra_steps > 0 && !hess && P < kPlaneFloor)
        P = rpd_clip_plane_room(form0, P, V, std::max(T, kTriangleFloor), kPlaneFloor);
    T = std::max(T, std::min(kTriangleFloor, rpd_clip_max_T(form0, P, V, hess)));
    // Likewise the vertex cap is the room of the vertex slots it needs
    // (64 per slot): a cap below it saves nothing.
    V = V <= 64 ? 64 : (V <= 128 ? 128 : 256);
    rpd_stats st{};
    st.n = n;
    auto t_start = std::chrono::steady_clock::now();

    HIPCHK(c, hipEventRecord(c->evb[0], c->stream));
    // The weight minima of new seeds (over the non-border seeds, and over
    // all, whose sign tells whether there are walls) and the binning are
    // queued, and nothing waits for them here: the cover test reads the
    // minima on the device (wdev), and the host reads them, with the
    // binning's crowded-cell flag, after the solve's first wait (settle
    // below), which the kNN launch has already overlapped.
    const bool new_wmin = !batch && prm->security_mode == 0 && !c->wmin_ok;
    const double *wdev = nullptr;
    if (new_wmin) {
        RPDCHK(c, c->wred.ensure((2 * kWminBlocks + 2) * sizeof(double)), "alloc wred");
        double *wr = c->wred.as<double>();
        wmin_kernel<<<kWminBlocks, 256, 0, c->stream>>>(c->seeds.as<double4>(), n, wr);
        wmin_final_kernel<<<1, 256, 0, c->stream>>>(wr, kWminBlocks, wr + 2 * kWminBlocks);
        HIPCHK(c, hipMemcpyAsync(c->pin_red, wr + 2 * kWminBlocks, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        wdev = wr + 2 * kWminBlocks;
    }
    if (batch) {
        if (c->bin_prec != kprec) return set_err(c, RPD_E_STATE, "rpd_solve_batch: binned in another precision");
    }
    bool bin_pending = false;
    if (!batch) RPDCHK(c, ensure_binning(c, kprec, &bin_pending), "binning");
    HIPCHK(c, hipEventRecord(c->evb[1], c->stream));

    RPDCHK(c, c->out.ensure(size_t(n) * 32), "alloc out");
    RPDCHK(c, c->status.ensure(size_t(n) * 4), "alloc status");
    RPDCHK(c, c->kused.ensure(size_t(n) * 4), "alloc kused");
    for (auto &l : c->fl) {
        RPDCHK(c, l.cnt.ensure(kFailCounters * sizeof(int)), "alloc failure list");
        RPDCHK(c, l.eo.ensure(size_t(n) * 8), "alloc failure list");
        RPDCHK(c, l.sec.ensure(size_t(n) * 4), "alloc failure list");
    }
    RPDCHK(c, c->rs_cnt.ensure(4 * sizeof(int)), "alloc resume counters");
    RPDCHK(c, c->rslot.ensure(size_t(n) * 4), "alloc resume slots");
    RPDCHK(c, c->lastnb.ensure(size_t(n) * 4), "alloc resume slots");
    RPDCHK(c, c->fslot.ensure(size_t(n) * 8), "alloc failure slots");
    RPDCHK(c, c->cnt.ensure(64), "alloc cnt");
    for (auto &b : c->pb) RPDCHK(c, b.qcount.ensure(sizeof(int)), "alloc qcount");
    // The counters the solve starts from, cleared by one launch: the final
    // status histogram's, every failure list's, the resume buffers' and the
    // queue lengths.  A list, resume buffer or queue counter stays "fresh"
    // (known zero) until a pass writes it; clearing a fresh one is skipped.
    bool fresh_fl[6], fresh_rs[4], fresh_q[4];
    {
        ZeroSet z{};
        z.p[z.k] = c->cnt.as<int>();
        z.n[z.k++] = 16;
        for (int i = 0; i < 6; ++i) {
            z.p[z.k] = c->fl[i].cnt.as<int>();
            z.n[z.k++] = kFailCounters;
            fresh_fl[i] = true;
        }
        z.p[z.k] = c->rs_cnt.as<int>();
        z.n[z.k++] = 4;
        for (int i = 0; i < 4; ++i) {
            z.p[z.k] = c->pb[i].qcount.as<int>();
            z.n[z.k++] = 1;
            fresh_rs[i] = fresh_q[i] = true;
        }
        zero_words_kernel<<<1, 64, 0, c->stream>>>(z);
        HIPCHK(c, hipGetLastError());
    }
    c->hess_H = hess ? prm->H : 0;
    c->hess_HC = hess ? prm->HC : 1;
    if (hess) {
        RPDCHK(c, c->hess.ensure(size_t(n) * prm->H * prm->HC * sizeof(double)), "alloc hessian");
        RPDCHK(c, c->hids.ensure(size_t(n) * prm->H * sizeof(int)), "alloc hessian");
        // no entries (-1, values 0) until a cell writes its own
        HIPCHK(c, hipMemsetAsync(c->hids.ptr, 0xff, size_t(n) * prm->H * sizeof(int), c->stream));
        HIPCHK(c, hipMemsetAsync(c->hess.ptr, 0, size_t(n) * prm->H * prm->HC * sizeof(double), c->stream));
    }

    // debug counters: one zeroed block per pass (RPD_MAX_PASSES - 1 and
    // later passes share the last), read back by rpd_read_debug_counters
    c->dcnt_ok = false;
    if (prm->debug_counters) {
        RPDCHK(c, c->dcnt.ensure(size_t(RPD_MAX_PASSES) * kDcStride * sizeof(int)), "alloc debug counters");
        HIPCHK(c, hipMemsetAsync(c->dcnt.ptr, 0, size_t(RPD_MAX_PASSES) * kDcStride * sizeof(int), c->stream));
    }
    auto dcnt_of = [&](int index) -> int * {
        return prm->debug_counters ? c->dcnt.as<int>() + size_t(std::min(index, RPD_MAX_PASSES - 1)) * kDcStride
                                   : nullptr;
    };
    const DomainDev *dd = dom ? &dom->dev : nullptr;
    // the weight minimum the kernels subtract, and whether there are walls
    // (border seeds): known here unless the minima are new (then at settle)
    double wmin = 0;
    bool border_seeds = true;
    auto host_wmin = [&]() {
        if (prm->security_mode == 0) {
            wmin = c->wmin < INFINITY ? c->wmin : 0.0;
            border_seeds = c->wall <= RPD_NO_CELL_LIMITS;
        }
    };
    if (!new_wmin) host_wmin();
    // After the solve's first wait: the binning's time, its crowded-cell
    // path, the new weight minima.  *restart: the binning was redone (a
    // crowded cell), so what was queued on the old one is to be redone.
    auto settle = [&](bool *restart) -> int {
        *restart = false;
        if (hipEventSynchronize(c->evb[1]) != hipSuccess) return RPD_E_DEVICE;
        st.ms_bin = elapsed_ms(c->evb[0], c->evb[1]);
        if (new_wmin) {
            c->wmin = c->pin_red[0];
            c->wall = c->pin_red[1];
            c->wmin_ok = true;
            host_wmin();
        }
        if (bin_pending && c->pin[kPinBinFlag]) {
            const int r = finish_binning(c);
            if (r != RPD_OK) return r;
            *restart = true;
        }
        return RPD_OK;
    };
    // cells provably outside the domain skip pass 1 (exact stop, no walls);
    // in a batch, per request
    const int *ids1 = nullptr;
    int64_t n1 = n;
    std::vector<CoverArgs> cov;
    auto cover_of = [&](const DomainDev *d, const int *counters, int dim, int64_t base, int64_t cnt, double wm,
                        bool walls) {
        CoverArgs ca{};
        ca.stored64 = c->stored64.as<double4>();
        ca.perm = c->perm.as<int>();
        ca.counters = counters;
        ca.dim = dim;
        ca.base = base;
        ca.n = cnt;
        ca.apply = d && prm->cull && prm->security_mode == 0 && !walls;
        if (d) {
            ca.dvox = d->dvox;
            ca.n_dvox = d->n_dvox;
            ca.dist = d->dist;
            ca.g = d->g;
            ca.vs = d->vs;
        }
        ca.wmin = wm;
        ca.wdev = walls ? nullptr : wdev;
        ca.out = c->out.as<double4>();
        ca.status = c->status.as<int>();
        ca.kused = c->kused.as<int>();
        return ca;
    };
    if (batch) {
        for (const auto &q : c->batch)
            cov.push_back(cover_of(q.dom ? &q.dom->dev : nullptr, c->counters.as<int>() + q.coff, q.dim, q.base, q.n,
                                   prm->security_mode == 0 && q.wmin < INFINITY ? q.wmin : 0.0,
                                   prm->security_mode != 0 || q.wall <= RPD_NO_CELL_LIMITS));
    } else if (dd) {
        // new minima: the device tells walls (no cover test) from none
        cov.push_back(cover_of(dd, c->counters.as<int>(), c->dim, 0, n, wmin, new_wmin ? false : border_seeds));
    }
    bool any_cover = false;
    for (const auto &ca : cov) any_cover = any_cover || ca.apply;
    // The kept cells' count comes back through pinned memory without a wait:
    // pass 1's kNN is queued over n ids and reads the count on the device,
    // and the host waits for the count (event ev[1][1]) only after that
    // launch, so the wait overlaps the kNN.
    const int *cover_count = nullptr;
    int *cover_pin = c->pin + kPinCounters - 3;
    if (any_cover) {
        RPDCHK(c, c->ids1.ensure(size_t(n) * 4), "alloc ids");
        HIPCHK(c, hipEventRecord(c->ev[1][0], c->stream));
        RPDCHK(c,
               rpd_cover_skip(c->stream, cov.data(), int(cov.size()), n, c->cull_scratch, c->ids1.as<int>(), cover_pin,
                              &cover_count),
               "domain cover");
        HIPCHK(c, hipEventRecord(c->ev[1][1], c->stream));
        ids1 = c->ids1.as<int>();
    }
    const ReqDesc *reqs = batch ? c->req_desc.as<ReqDesc>() : nullptr;
    const KnnReq *kreqs = batch ? c->knn_req.as<KnnReq>() : nullptr;
    const uint16_t *req_of = batch ? c->req_of.as<uint16_t>() : nullptr;

    // One pass = kNN + bisector kernel (which waits for the queue length)
    // on a stream, then the finish kernel queued behind them.  Failed cells
    // go to failure list fl_out; fl_cnt_host receives the list's counters
    // as they stand once the bisector kernel is done.
    struct Pass {
        int index = 0;  // pass number (0 the first), for the debug counters
        int slot = 0;  // stream, events and buffers
        int form = 0;
        int64_t n_ids = 0;
        int caps[4] = {0, 0, 0, 0};
        int nq = 0;
        bool finish_only = false;  // a finish kernel over earlier records (events 1 and 3 only)
    };
    // resumable cells: records read (in, FP64 passes) and written (out)
    struct Resume {
        const char *in = nullptr;  // the cells of this pass resume from here
        int in_stride = 0, in_P = 0;
        int out = -1;              // rs buffer written, -1: none
        int out_stride = 0, out_cap = 0;
    };
    // device address of a pinned word (nullptr: no mapping)
    auto pin_d = [&](const int *host) -> int * { return c->pin_dev ? c->pin_dev + (host - c->pin) : nullptr; };
    // kNN + bisector kernel, the queue length copied to the host (not waited for)
    auto pass_front = [&](Pass &ps, const int *ids, int fl_out, int *fl_cnt_host, const Resume &rz,
                          ClipLaunch &L) -> int {
        const int sl = ps.slot;
        hipStream_t s = c->streams[sl];
        PassBufs &b = c->pb[sl];
        const int K_ = ps.caps[0], P_ = ps.caps[1], V_ = ps.caps[2], T_ = ps.caps[3];
        int r = b.knn.ensure(size_t(ps.n_ids) * K_ * 4);
        if (r == RPD_OK) r = b.rec.ensure(size_t(ps.n_ids) * rpd_clip_record_bytes(ps.form, P_, V_));
        if (r == RPD_OK) r = b.qcount.ensure(sizeof(int));
        if (r != RPD_OK) return r;
        if (hipEventRecord(c->ev[sl][0], s) != hipSuccess) return RPD_E_DEVICE;
        // (re-queue passes run beside a finish kernel: one-wave workgroups)
        r = rpd_launch_knn(s, kprec, c->stored.ptr, c->perm.as<int>(), c->counters.as<int>(), c->dim, ids, ps.n_ids,
                           K_, b.knn.as<int>(), rz.in ? c->lastnb.as<int>() : nullptr, kreqs, req_of, nullptr, true);
        if (r != RPD_OK) return r;
        if (hipEventRecord(c->ev[sl][1], s) != hipSuccess) return RPD_E_DEVICE;
        L = ClipLaunch{};
        L.seeds = c->stored64.as<double4>();
        L.perm = c->perm.as<int>();
        L.ids = ids;
        L.knn = b.knn.as<int>();
        L.n_ids = ps.n_ids;
        L.n_seeds = n;
        L.K = K_;
        L.P = P_;
        L.V = V_;
        L.T = T_;
        L.apex_mode = prm->apex_mode;
        L.security_mode = prm->security_mode;
        L.wmin = wmin;
        L.cull = prm->cull != 0;
        L.dom = dd;
        L.out = c->out.as<double4>();
        L.status = c->status.as<int>();
        L.kused = c->kused.as<int>();
        L.rec = b.rec.ptr;
        L.qcount = b.qcount.as<int>();
        L.ev_mid = c->ev[sl][2];
        L.fl_cnt = c->fl[fl_out].cnt.as<int>();
        L.fl_eo = c->fl[fl_out].eo.as<int>() + n;
        L.fl_sec = c->fl[fl_out].sec.as<int>();
        L.fl_cnt_host = fl_cnt_host;
        L.rs_in = rz.in;
        L.rs_in_stride = rz.in_stride;
        L.rs_in_P = rz.in_P;
        L.rs_out = rz.out >= 0 ? static_cast<char *>(c->rs[rz.out].ptr) : nullptr;
        L.rs_out_stride = rz.out_stride;
        L.rs_cap = rz.out_cap;
        L.rs_count = c->rs_cnt.as<int>() + std::max(rz.out, 0);
        L.rslot = c->rslot.as<int>();
        L.lastnb = c->lastnb.as<int>();
        if (hess) {
            L.hess = c->hess.as<double>();
            L.hids = c->hids.as<int>();
            L.H = prm->H;
            L.hc = prm->HC;
        }
        L.dcnt = dcnt_of(ps.index);
        L.reqs = reqs;
        L.req_of = req_of;
        L.qcount_zeroed = fresh_q[sl];
        fresh_q[sl] = false;
        L.nq_dev_out = pin_d(c->pin + sl);
        L.fl_cnt_dev_out = fl_cnt_host ? pin_d(fl_cnt_host) : nullptr;
        return rpd_launch_cells(s, ps.form, L, c->pin + sl, false);
    };
    // the finish kernel of a pass whose bisector kernel (and queue length) is done
    // The FP64 passes of a mixed solve finish in the mixed form: over the
    // same records (FP64 planes), its finish kernel decides in FP64, in
    // place, the tests its filter leaves open and those of ill-conditioned
    // vertices, so a cell gets the FP64 form's combinatorics for less work.
    const int finish_form_fp64 = form0 == RPD_PREC_MIXED && !hess ? RPD_PREC_MIXED : RPD_PREC_FP64;
    auto pass_back = [&](Pass &ps, const ClipLaunch &L, int beside) -> int {
        hipStream_t s = c->streams[ps.slot];
        ps.nq = c->pin[ps.slot];
        const int r =
            rpd_launch_finish(s, ps.form == RPD_PREC_FP64 ? finish_form_fp64 : ps.form, L, ps.nq, beside);
        if (r != RPD_OK) return r;
        return hipEventRecord(c->ev[ps.slot][3], s) == hipSuccess ? RPD_OK : RPD_E_DEVICE;
    };
    double ms_knn = 0, ms_fb_knn = 0, ms_fb = 0;
    auto account = [&](const Pass &ps, int index) {
        const int sl = ps.slot;
        const double mk = ps.finish_only ? 0.0 : elapsed_ms(c->ev[sl][0], c->ev[sl][1]),
                     mc = elapsed_ms(c->ev[sl][1], c->ev[sl][3]);
        const int pi = std::min(index, RPD_MAX_PASSES - 1);
        st.pass_cells[pi] += int32_t(ps.n_ids);
        for (int k = 0; k < 4; ++k) st.pass_caps[pi][k] = ps.caps[k];
        st.pass_ms_knn[pi] += mk;
        st.pass_ms_clip[pi] += mc;
        if (index == 0) {
            ms_knn = mk;
            st.ms_knn_first = mk;
            st.ms_clip_first = mc;
            st.ms_cells_first = elapsed_ms(c->ev[sl][1], c->ev[sl][2]);
            st.ms_finish_first = elapsed_ms(c->ev[sl][2], c->ev[sl][3]);
            st.queued_first = ps.nq;
        } else {
            ms_fb_knn += mk;
            ms_fb += mc;
        }
    };
    // Caps of the next pass from the counters of a failure list, grown from
    // cap: vertex/plane overflow x1.5, security radius x2 (a cell's result
    // does not depend on K once its radius is reached, and one wider kNN costs
    // less than another pass), triangle overflow to the largest union the
    // finish kernel met (within the LDS of one of its cells).
    struct Caps {
        int K, P, V, T;
    };
    auto grow = [&](const int *cnt, const Caps &cap) {
        const int *h = cnt + 2;
        Caps nc = cap;
        if (h[RPD_VERTEX_OVERFLOW]) nc.V = std::min(256, int(cap.V * 1.5));
        if (h[RPD_PLANE_OVERFLOW]) nc.P = std::min(255, int(cap.P * 1.5));
        // K doubles, by at most 256 a pass: a resumed cell lists only the
        // neighbours after the ones it used (at most 256, the kNN's limit),
        // so a cell may go on past 256 neighbours in all (a cell bounded by
        // walls reaches twice as far as its bisectors); a recomputed cell
        // lists at most 256
        if (h[RPD_SECURITY_RADIUS_NOT_REACHED])
            nc.K = int(std::min<int64_t>(n - 1, int64_t(cap.K) + std::min(cap.K, 256)));
        if (h[RPD_TRIANGLE_OVERFLOW]) nc.T = std::max(int(cap.T * 1.5) + 1, cnt[13]);
        nc.T = std::min(nc.T, rpd_clip_max_T(RPD_PREC_FP64, nc.P, nc.V, hess));
        return nc;
    };
    // (the list is written by the pass it is cleared for)
    auto clear_list = [&](int fl, hipStream_t s) {
        if (fresh_fl[fl]) {
            fresh_fl[fl] = false;
            return true;
        }
        return rpd_zero_ints(s, c->fl[fl].cnt.as<int>(), kFailCounters) == RPD_OK;
    };
    // Record buffer `which` for cap records of stride bytes, counter cleared.
    auto resume_buf = [&](int which, int64_t cap, int stride, hipStream_t s) -> int {
        const int r = c->rs[which].ensure(size_t(std::max<int64_t>(cap, 1)) * size_t(stride));
        if (r != RPD_OK) return r;
        if (fresh_rs[which]) {
            fresh_rs[which] = false;
            return RPD_OK;
        }
        return rpd_zero_ints(s, c->rs_cnt.as<int>() + which, 1);
    };
    // Where the previous pass of a chain left the states of its resumable
    // cells (rs buffer, stride, plane room), or which = -1.
    struct Left {
        int which = -1, stride = 0, P = 0;
    };
    // 1: done, 0: still running, -1: error
    auto done = [&](hipStream_t s) {
        const hipError_t e = hipStreamQuery(s);
        return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : -1);
    };
    // A chain of re-queue passes, driven by polling so that two chains and
    // pass 1's finish kernel run at once.  Each pass takes the failures of
    // the previous one (list cur, counters cnt): the cells whose state was
    // kept resume on stream sl (their kNN lists only the neighbours after the
    // last one used), the others are recomputed in FP64 on stream sl2, both
    // at once.  Pass k writes its failures to list outs[k % 2] and the states
    // of its resumable cells to rs buffer rsb[k % 2].
    struct Chain {
        int sl = 0, sl2 = 0, dep = 0;
        int outs[2] = {0, 0}, rsb[2] = {0, 0};
        int cur = 0;
        const int *cnt = nullptr;
        Caps cap{0, 0, 0, 0}, nc{0, 0, 0, 0};
        Left left;
        int k = 0, nxt = 0;
        int phase = 0;  // 0 idle, 1 bisector kernels queued, 2 finish kernels queued, 3 done, 4 waiting for pass 1
        int64_t n_res = 0, n_scr = 0;
        Pass pr, pz;
        ClipLaunch Lr, Lz;
        Resume rz;
        bool from_finish = false;  // the chain of pass 1's finish-kernel failures
        bool back_r = false, back_z = false;  // phase 1: the finish kernel of pr / pz queued
    };
    ClipLaunch L1{};  // pass 1's launch; its finish kernel's records serve chain B's first pass
    auto cnt_pin = [&](int list) { return c->pin + kPinCounters + kFailCounters * list; };
    // a failure list's counters to their pinned words, on stream s
    auto copy_counts = [&](int list, hipStream_t s) -> bool {
        if (c->pin_dev)
            return rpd_pin_copy(s, c->fl[list].cnt.as<int>(), pin_d(cnt_pin(list)), kFailCounters) == RPD_OK;
        return hipMemcpyAsync(cnt_pin(list), c->fl[list].cnt.ptr, kFailCounters * sizeof(int), hipMemcpyDeviceToHost,
                              s) == hipSuccess;
    };
    // plans the chain's next pass from the counters and queues its kNN and
    // bisector kernels (phase 1), or ends the chain (phase 3)
    auto chain_front = [&](Chain &ch) -> int {
        if (ch.k >= prm->max_extra_steps) {
            ch.phase = 3;
            return RPD_OK;
        }
        const int *cnt = ch.cnt;
        const int n_exact = cnt[0], n_over = cnt[1], n_sec = cnt[14];
        const Caps nc = grow(cnt, ch.cap);
        const bool caps_changed = nc.K != ch.cap.K || nc.P != ch.cap.P || nc.V != ch.cap.V || nc.T != ch.cap.T;
        // exact-predicate failures (only listed from the FP32 forms) always
        // go to FP64; overflows only when a cap grew; kept states when K grew
        // (else they would run out again: they end as they are)
        ch.n_res = n_sec > 0 && ch.left.which >= 0 && nc.K > ch.cap.K ? n_sec : 0;
        ch.n_scr = n_exact + (caps_changed ? n_over : 0);
        if (ch.n_res + ch.n_scr == 0) {
            ch.phase = 3;
            return RPD_OK;
        }
        st.fallback_cells += n_exact;
        hipStream_t s = c->streams[ch.sl], s2 = c->streams[ch.sl2];
        ch.nc = nc;
        ch.nxt = ch.outs[ch.k & 1];
        if (!clear_list(ch.nxt, s)) return RPD_E_DEVICE;
        // Chain B's first pass: pass 1's finish kernel failed these cells in
        // the restriction (a flagged test, or the triangle, vertex or plane
        // cap met by the cone clips).  Their bisector loop is done,
        // certified, and in pass 1's records, which are the FP64 form's
        // records too, so the FP64 finish kernel runs over those records
        // with the grown caps; a full FP64 recomputation would rebuild the
        // same records first.  The FP32 form's records hold FP32 planes (16
        // bytes, a different layout): its cells take the full pass.
        if (ch.from_finish && ch.k == 0 && ch.n_res == 0 && ch.n_scr > 0 && form0 != RPD_PREC_FP32 &&
            rpd_clip_record_bytes(form0, ch.cap.P, ch.cap.V) == rpd_clip_record_bytes(RPD_PREC_FP64, ch.cap.P, ch.cap.V)) {
            ch.rz = Resume{};
            ch.pr = Pass{};
            ch.pz = Pass{};
            ch.pz.slot = ch.sl;
            ch.pz.index = 1;
            ch.pz.form = RPD_PREC_FP64;
            ch.pz.n_ids = ch.n_scr;
            ch.pz.caps[0] = ch.cap.K;
            ch.pz.caps[1] = ch.cap.P;
            ch.pz.caps[2] = ch.cap.V;
            ch.pz.caps[3] = nc.T;
            ch.pz.caps[1] = nc.P;
            ch.pz.caps[2] = nc.V;
            ClipLaunch &L = ch.Lz;
            L = L1;
            L.rec_P = ch.cap.P;  // the records' layout: pass 1's caps
            L.rec_V = ch.cap.V;
            L.P = nc.P;
            L.V = nc.V;
            L.T = nc.T;
            L.slots = c->fslot.as<int>() + n - n_exact;  // the failures' range of the list, in place
            L.fl_slot = nullptr;
            L.fl_cnt = c->fl[ch.nxt].cnt.as<int>();
            L.fl_eo = c->fl[ch.nxt].eo.as<int>() + n;
            L.fl_sec = c->fl[ch.nxt].sec.as<int>();
            L.fl_cnt_host = nullptr;
            L.rs_out = nullptr;
            L.rs_in = nullptr;
            L.ev_mid = nullptr;
            L.dcnt = dcnt_of(1);
            ch.pz.finish_only = true;
            if (hipEventRecord(c->ev[ch.sl][1], s) != hipSuccess) return RPD_E_DEVICE;
            int r = rpd_launch_finish(s, RPD_PREC_FP64, L, int(ch.n_scr), 0);
            if (r != RPD_OK) return r;
            if (hipEventRecord(c->ev[ch.sl][3], s) != hipSuccess || !copy_counts(ch.nxt, s)) return RPD_E_DEVICE;
            ch.phase = 2;
            return RPD_OK;
        }
        ch.rz = Resume{};
        ch.rz.out = ch.rsb[ch.k & 1];
        ch.rz.out_stride = int(rpd_clip_record_bytes(RPD_PREC_FP64, nc.P, nc.V));
        ch.rz.out_cap = int(ch.n_res + ch.n_scr);
        int r = resume_buf(ch.rz.out, ch.rz.out_cap, ch.rz.out_stride, s);
        if (r != RPD_OK) return r;
        ch.pr = Pass{};
        ch.pz = Pass{};
        ch.pr.slot = ch.sl;
        ch.pz.slot = ch.sl2;
        ch.pr.index = ch.pz.index = 1 + ch.k;
        ch.pr.form = ch.pz.form = RPD_PREC_FP64;
        ch.pr.n_ids = ch.n_res;
        ch.pz.n_ids = ch.n_scr;
        ch.pr.caps[0] = nc.K - ch.cap.K;  // the neighbours after the ones used
        // recomputed cells list the grown K: a cell sent back for exact
        // predicates or an overflow starts from scratch, and with the old K
        // as many of them would run out of neighbours again as in the
        // previous pass (one more pass for them); the loop reads only the
        // neighbours it needs
        ch.pz.caps[0] = std::min(256, nc.K);
        ch.pr.caps[1] = ch.pz.caps[1] = nc.P;
        ch.pr.caps[2] = ch.pz.caps[2] = nc.V;
        ch.pr.caps[3] = ch.pz.caps[3] = nc.T;
        // the recomputed cells start once the list and the record counter are cleared
        if (ch.n_scr > 0 && (hipEventRecord(c->dep[ch.dep], s) != hipSuccess ||
                             hipStreamWaitEvent(s2, c->dep[ch.dep], 0) != hipSuccess))
            return RPD_E_DEVICE;
        if (ch.n_res > 0) {
            Resume rr = ch.rz;
            rr.in = static_cast<const char *>(c->rs[ch.left.which].ptr);
            rr.in_stride = ch.left.stride;
            rr.in_P = ch.left.P;
            r = pass_front(ch.pr, c->fl[ch.cur].sec.as<int>(), ch.nxt, nullptr, rr, ch.Lr);
            if (r != RPD_OK) return r;
        }
        if (ch.n_scr > 0) {
            // exact entries and (when recomputed) the overflows are one range
            r = pass_front(ch.pz, c->fl[ch.cur].eo.as<int>() + n - n_exact, ch.nxt, nullptr, ch.rz, ch.Lz);
            if (r != RPD_OK) return r;
        }
        ch.back_r = ch.n_res == 0;
        ch.back_z = ch.n_scr == 0;
        ch.phase = 1;
        return RPD_OK;
    };
    Pass p1;                 // pass 1 (stream 0)
    bool p1_running = true;  // its finish kernel is not known to be done
    // advances the chain as far as its streams allow, without waiting
    auto chain_poll = [&](Chain &ch, bool &progress) -> int {
        hipStream_t s = c->streams[ch.sl], s2 = c->streams[ch.sl2];
        if (ch.phase == 1) {
            // each finish kernel as soon as its own bisector kernel is done;
            // pass 1's finish kernel may still run beside them
            const int beside = p1_running ? p1.nq : 0;
            bool s_done = false;  // s found idle by this poll, before any launch on it
            if (!ch.back_r) {
                const int d = done(s);
                if (d < 0) return RPD_E_DEVICE;
                if (d) {
                    s_done = true;
                    if (pass_back(ch.pr, ch.Lr, beside) != RPD_OK) return RPD_E_DEVICE;
                    ch.back_r = true;
                    progress = true;
                }
            }
            if (!ch.back_z) {
                // one stream for both (two-stream contexts): pz's bisector
                // kernel was queued after pr's, so s idle covers it
                const int d = s2 == s && s_done ? 1 : done(s2);
                if (d < 0) return RPD_E_DEVICE;
                if (d) {
                    if (pass_back(ch.pz, ch.Lz, beside) != RPD_OK) return RPD_E_DEVICE;
                    ch.back_z = true;
                    progress = true;
                }
            }
            if (!ch.back_r || !ch.back_z) return RPD_OK;
            // the counters of the pass, once both finish kernels are done
            if (ch.n_scr > 0 && (hipEventRecord(c->dep[ch.dep], s2) != hipSuccess ||
                                 hipStreamWaitEvent(s, c->dep[ch.dep], 0) != hipSuccess))
                return RPD_E_DEVICE;
            if (!copy_counts(ch.nxt, s)) return RPD_E_DEVICE;
            ch.phase = 2;
            progress = true;
        }
        if (ch.phase == 2) {
            const int d = done(s);
            if (d < 0) return RPD_E_DEVICE;
            if (!d) return RPD_OK;
            if (ch.n_res > 0) account(ch.pr, 1 + ch.k);
            if (ch.n_scr > 0) account(ch.pz, 1 + ch.k);
            ch.cap = ch.nc;
            ch.left.which = ch.rz.out;
            ch.left.stride = ch.rz.out_stride;
            ch.left.P = ch.nc.P;
            ch.cur = ch.nxt;
            ch.cnt = cnt_pin(ch.nxt);
            ch.k += 1;
            ch.phase = 0;
            progress = true;
            return chain_front(ch);
        }
        return RPD_OK;
    };

    // Pass 1 on stream 0: the bisector kernel's failures go to list 0, the
    // finish kernel's to list 1.  Two chains of re-queue passes follow, at
    // once:
    //   * chain A, the bisector failures, known as soon as that kernel is
    //     done: its first pass is queued (streams 1, 2) while pass 1's finish
    //     kernel, queued behind the bisector kernel, runs;
    //   * chain B, the finish kernel's failures, once it is done (streams 0,
    //     3), with the caps of pass 1 grown only for them.
    if (!clear_list(0, c->stream) || !clear_list(1, c->stream)) return set_err(c, RPD_E_DEVICE, "clear lists");
    Left left1;
    {
        hipStream_t s = c->streams[0];
        PassBufs &b = c->pb[0];
        // pass 1's kNN: over the kept cells, their count read on the device
        // while the host does not know it yet.  The launch covers the last
        // solve's count over as many seeds, with a margin (n rows at most, or
        // all n without such a solve); rows past the count exit at once, and
        // rows past the launch, if the count turns out larger, get a second
        // launch once the host has it.
        RPDCHK(c, b.knn.ensure(size_t(std::max<int64_t>(n1, 1)) * K * 4), "alloc knn");
        int64_t rows = n1;
        if (cover_count && c->kept_hint >= 0 && c->kept_hint_n == n)
            rows = std::min<int64_t>(n1, c->kept_hint + c->kept_hint / 16 + 4096);
        HIPCHK(c, hipEventRecord(c->ev[0][0], s));
        RPDCHK(c,
               rpd_launch_knn(s, kprec, c->stored.ptr, c->perm.as<int>(), c->counters.as<int>(), c->dim, ids1, rows,
                              K, b.knn.as<int>(), nullptr, kreqs, req_of, cover_count),
               "pass 1 knn");
        HIPCHK(c, hipEventRecord(c->ev[0][1], s));
        if (cover_count) {
            HIPCHK(c, hipEventSynchronize(c->ev[1][1]));
            st.ms_cover = elapsed_ms(c->ev[1][0], c->ev[1][1]);
            n1 = *cover_pin;
            st.skipped_cells = n - n1;
            c->kept_hint = n1;
            c->kept_hint_n = n;
            if (n1 > rows) {
                RPDCHK(c,
                       rpd_launch_knn(s, kprec, c->stored.ptr, c->perm.as<int>(), c->counters.as<int>(), c->dim,
                                      ids1 + rows, n1 - rows, K, b.knn.as<int>() + size_t(rows) * size_t(K), nullptr,
                                      kreqs, req_of),
                       "pass 1 knn (rest)");
                HIPCHK(c, hipEventRecord(c->ev[0][1], s));
            }
        }
        {
            // (the cover test's wait, if any, has covered the binning)
            bool restart = false;
            RPDCHK(c, settle(&restart), "binning");
            // the seeds were rebinned: the cover test and the kNN queued
            // so far read the old order, so the solve starts over (once:
            // the seeds are binned now)
            if (restart) return solve_impl(c, dom, prm, stats);
        }
        p1.slot = 0;
        p1.form = form0;
        p1.n_ids = n1;
        p1.caps[0] = K;
        p1.caps[1] = P;
        p1.caps[2] = V;
        p1.caps[3] = T;
        // states of pass 1's cells whose K neighbours do not reach their
        // security radius are kept (FP64 planes: FP64 and mixed forms), for
        // up to 40% of the cells; the others are recomputed from scratch
        const int64_t cap_rs1 = std::min<int64_t>(n1, n1 * 2 / 5 + 1024);
        if (form0 != RPD_PREC_FP32 && prm->max_extra_steps > 0 && n1 > 0) {
            left1.which = 0;
            left1.stride = int(rpd_clip_record_bytes(form0, P, V));
            left1.P = P;
            RPDCHK(c, resume_buf(0, cap_rs1, left1.stride, c->streams[0]), "alloc resume records");
        }
        RPDCHK(c, b.rec.ensure(size_t(std::max<int64_t>(n1, 1)) * rpd_clip_record_bytes(form0, P, V)), "alloc records");
        RPDCHK(c, b.qcount.ensure(sizeof(int)), "alloc qcount");
        ClipLaunch &L = L1;
        L.seeds = c->stored64.as<double4>();
        L.perm = c->perm.as<int>();
        L.ids = ids1;
        L.knn = b.knn.as<int>();
        L.n_ids = n1;
        L.n_seeds = n;
        L.K = K;
        L.P = P;
        L.V = V;
        L.T = T;
        L.apex_mode = prm->apex_mode;
        L.security_mode = prm->security_mode;
        L.wmin = wmin;
        L.cull = prm->cull != 0;
        L.dom = dd;
        L.out = c->out.as<double4>();
        L.status = c->status.as<int>();
        L.kused = c->kused.as<int>();
        L.rec = b.rec.ptr;
        L.qcount = b.qcount.as<int>();
        L.ev_mid = c->ev[0][2];
        L.fl_cnt = c->fl[0].cnt.as<int>();
        L.fl_eo = c->fl[0].eo.as<int>() + n;
        L.fl_sec = c->fl[0].sec.as<int>();
        L.fl_cnt_host = cnt_pin(0);
        if (hess) {
            L.hess = c->hess.as<double>();
            L.hids = c->hids.as<int>();
            L.H = prm->H;
            L.hc = prm->HC;
        }
        if (left1.which == 0) {
            L.rs_out = static_cast<char *>(c->rs[0].ptr);
            L.rs_out_stride = left1.stride;
            L.rs_cap = int(cap_rs1);
            L.rs_count = c->rs_cnt.as<int>();
            L.rslot = c->rslot.as<int>();
            L.lastnb = c->lastnb.as<int>();
        }
        L.dcnt = dcnt_of(0);
        L.reqs = reqs;
        L.req_of = req_of;
        L.qcount_zeroed = fresh_q[0];
        fresh_q[0] = false;
        L.nq_dev_out = pin_d(c->pin + 0);
        L.fl_cnt_dev_out = pin_d(cnt_pin(0));
        RPDCHK(c, rpd_launch_cells(s, form0, L, c->pin + 0, false), "pass 1 cells");
        L.nq_dev_out = L.fl_cnt_dev_out = nullptr;
        L.qcount_zeroed = false;
        HIPCHK(c, hipEventRecord(c->dep[2], s));  // queue length and failure counters copied
        L.fl_cnt = c->fl[1].cnt.as<int>();
        L.fl_eo = c->fl[1].eo.as<int>() + n;
        L.fl_sec = c->fl[1].sec.as<int>();
        L.fl_cnt_host = nullptr;
        L.fl_slot = c->fslot.as<int>() + n;  // the finish kernel's failures keep their record slots
        // The finish kernel is queued at once, over the n1 cells the queue
        // can hold at most, reading its length on the device: no host round
        // trip between the two kernels.  Chain A's first pass follows once
        // the host has the bisector kernel's failure counters; its kNN uses
        // one-wave workgroups, which fit beside the finish kernel's.
        L.nq_dev = L.qcount;
        RPDCHK(c, rpd_launch_finish(s, form0, L, int(n1), 0), "pass 1 finish");
        HIPCHK(c, hipEventRecord(c->ev[0][3], s));
        if (!copy_counts(1, s)) return set_err(c, RPD_E_DEVICE, "pass 1 counters");
        HIPCHK(c, hipEventSynchronize(c->dep[2]));
        p1.nq = c->pin[0];
        // no cell to clip (every one skipped): the bisector kernel did not
        // run and copied no failure counters
        if (n1 <= 0) std::fill(cnt_pin(0), cnt_pin(0) + kFailCounters, 0);
        L.nq_dev = nullptr;
    }
    const Caps caps1{K, P, V, T};
    Chain ca, cb;
    ca.sl = 1, ca.sl2 = 2, ca.dep = 0;
    ca.outs[0] = 2, ca.outs[1] = 3;
    ca.rsb[0] = 1, ca.rsb[1] = 0;
    ca.cur = 0;
    ca.cnt = cnt_pin(0);
    ca.cap = caps1;
    ca.left = left1;
    cb.sl = 0, cb.sl2 = 3, cb.dep = 1;
    cb.outs[0] = 4, cb.outs[1] = 5;
    cb.rsb[0] = 2, cb.rsb[1] = 3;
    cb.cur = 1;
    cb.cnt = cnt_pin(1);
    cb.cap = caps1;
    cb.phase = 4;
    cb.from_finish = true;
    // The final status histogram: counts into cnt (cleared at the solve's
    // start), then to its pinned words, on stream s.  The counts of a
    // histogram queued before are cleared first; that one has finished (the
    // host has seen its stream idle since: its chain went on).
    const unsigned hist_blocks = unsigned(std::min<int64_t>(1024, (n + 255) / 256));
    bool hist_queued = false;
    auto queue_hist = [&](hipStream_t s) -> int {
        if (hist_queued && rpd_zero_ints(s, c->cnt.as<int>(), RPD_STATUS_NUM) != RPD_OK) return RPD_E_DEVICE;
        status_hist_kernel<<<hist_blocks, 256, 0, s>>>(c->status.as<int>(), n, c->cnt.as<int>());
        if (hipGetLastError() != hipSuccess) return RPD_E_DEVICE;
        hist_queued = true;
        // through pinned memory: a copy into pageable memory is staged and
        // starts late
        if (c->pin_dev) return rpd_pin_copy(s, c->cnt.as<int>(), c->pin_dev + kPinHist, RPD_STATUS_NUM);
        return hipMemcpyAsync(c->pin + kPinHist, c->cnt.ptr, sizeof(st.status_hist), hipMemcpyDeviceToHost, s) ==
                       hipSuccess
                   ? RPD_OK
                   : RPD_E_DEVICE;
    };
    // Once one chain is done, the other's queued pass may be the solve's
    // last: the histogram is queued behind that pass's counters on its
    // stream, and is ready when the host learns that the pass failed no
    // cell.  It holds if the chain then ends (its pass count grew by one
    // only).  A chain that is done stays done, so every such histogram is
    // queued for the same chain, on its stream.
    const Chain *hist_for = nullptr;
    int hist_k = -1;
    RPDCHK(c, chain_front(ca), "re-queue chain (bisector failures)");
    while (ca.phase != 3 || cb.phase != 3) {
        bool progress = false;
        if (cb.phase == 4) {
            const int d = done(c->streams[0]);
            if (d < 0) return set_err(c, RPD_E_DEVICE, "pass 1 finish");
            if (d) {
                account(p1, 0);
                p1_running = false;
                cb.phase = 0;
                progress = true;
                RPDCHK(c, chain_front(cb), "re-queue chain (finish failures)");
            }
        }
        RPDCHK(c, chain_poll(ca, progress), "re-queue chain (bisector failures)");
        RPDCHK(c, chain_poll(cb, progress), "re-queue chain (finish failures)");
        const Chain *last = ca.phase == 3 && cb.phase == 2 ? &cb : (cb.phase == 3 && ca.phase == 2 ? &ca : nullptr);
        if (last && !(hist_for == last && hist_k == last->k)) {
            RPDCHK(c, queue_hist(c->streams[last->sl]), "status histogram");
            hist_for = last;
            hist_k = last->k;
        }
        if (!progress) std::this_thread::yield();
    }
    const int steps_a = ca.k, steps_b = cb.k;
    const Caps cap_a = ca.cap, cap_b = cb.cap;
    st.passes = 1 + std::max(steps_a, steps_b);
    K = std::max(cap_a.K, cap_b.K);
    P = std::max(cap_a.P, cap_b.P);
    V = std::max(cap_a.V, cap_b.V);
    T = std::max(cap_a.T, cap_b.T);
    // final status histogram, unless the one queued early covers every pass
    hipStream_t hist_s = c->stream;
    if (hist_for && hist_for->k == hist_k + 1)
        hist_s = c->streams[hist_for->sl];
    else
        RPDCHK(c, queue_hist(c->stream), "status histogram");
    HIPCHK(c, hipStreamSynchronize(hist_s));
    std::memcpy(st.status_hist, c->pin + kPinHist, sizeof(st.status_hist));
    st.ms_knn = ms_knn + ms_fb_knn;
    st.ms_clip = st.ms_clip_first;
    st.ms_fallback = ms_fb;
    st.final_K = K;
    st.final_P = P;
    st.final_V = V;
    st.final_T = T;
    st.dim = c->dim;
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    c->last_K = K;
    c->last_P = P;
    c->last_V = V;
    c->last_sm = prm->security_mode;
    c->last_steps = prm->max_extra_steps;
    c->solved = true;
    c->dcnt_ok = prm->debug_counters != 0;
    if (stats) *stats = st;
    return RPD_OK;
}

int rpd_solve(rpd_ctx *c, rpd_domain *dom, const rpd_params *prm, rpd_stats *stats) {
    if (c && !c->batch.empty())
        return set_err(c, RPD_E_STATE, "rpd_solve: the seeds are a packed batch (set seeds, or rpd_solve_batch)");
    return solve_impl(c, dom, prm, stats);
}

int rpd_solve_batch(rpd_ctx *c, int32_t nreq, rpd_domain *const *doms, const double *const *xyzw, const int64_t *ns,
                    const rpd_params *prm, rpd_stats *stats) {
    if (!c || !xyzw || !ns || !prm || nreq <= 0) return set_err(c, RPD_E_ARG, "rpd_solve_batch: bad arguments");
    if (nreq > RPD_BATCH_MAX_REQUESTS)
        return set_err(c, RPD_E_LIMIT, "rpd_solve_batch: at most %d requests", RPD_BATCH_MAX_REQUESTS);
    if (prm->H > 0) return set_err(c, RPD_E_ARG, "rpd_solve_batch: no Hessian in a batch");
    if (prm->precision < 0 || prm->precision > 2) return set_err(c, RPD_E_ARG, "rpd_solve_batch: bad precision");
    int64_t total = 0;
    for (int r = 0; r < nreq; ++r) {
        if (!xyzw[r] || ns[r] < 2) return set_err(c, RPD_E_ARG, "rpd_solve_batch: request %d needs 2 seeds", r);
        if (doms && doms[r] && doms[r]->device != c->device)
            return set_err(c, RPD_E_ARG, "rpd_solve_batch: domain of request %d on another device", r);
        total += ns[r];
        if (total > INT32_MAX / 2) return set_err(c, RPD_E_LIMIT, "rpd_solve_batch: too many seeds");
    }
    hipSetDevice(c->device);
    const int kprec = prm->precision == RPD_PREC_FP64 ? 1 : 0;
    c->batch.clear();
    c->n = 0;
    c->bin_prec = -1;
    c->wmin_ok = false;
    c->solved = false;
    std::vector<rpd_ctx::BatchReq> reqs(static_cast<size_t>(nreq));
    int64_t base = 0, coff = 0;
    for (int r = 0; r < nreq; ++r) {
        auto &q = reqs[size_t(r)];
        q.dom = doms ? doms[r] : nullptr;
        q.base = base;
        q.n = ns[r];
        q.dim = rpd_knn_dim(q.n);
        q.coff = coff;
        // weight minima over the non-border seeds and over all, on the host
        q.wmin = q.wall = INFINITY;
        const double *x = xyzw[r];
        for (int64_t i = 0; i < q.n; ++i) {
            const double w = x[4 * i + 3];
            if (w > RPD_NO_CELL_LIMITS) q.wmin = std::min(q.wmin, w);
            q.wall = std::min(q.wall, w);
        }
        base += q.n;
        coff += int64_t(q.dim) * q.dim * q.dim + 1;
        if (coff > INT32_MAX) return set_err(c, RPD_E_LIMIT, "rpd_solve_batch: kNN grids too large");
    }
    const size_t vsz = kprec == 0 ? 16 : 32;
    RPDCHK(c, c->seeds.ensure(size_t(total) * 32), "alloc seeds");
    RPDCHK(c, c->perm.ensure(size_t(total) * 4), "alloc perm");
    RPDCHK(c, c->counters.ensure(size_t(coff) * 4), "alloc counters");
    RPDCHK(c, c->stored.ensure(size_t(total) * vsz), "alloc stored");
    RPDCHK(c, c->stored64.ensure(size_t(total) * 32), "alloc stored64");
    RPDCHK(c, c->req_of.ensure(size_t(total) * 2), "alloc request map");
    RPDCHK(c, c->req_desc.ensure(size_t(nreq) * sizeof(ReqDesc)), "alloc request descriptors");
    RPDCHK(c, c->knn_req.ensure(size_t(nreq) * sizeof(KnnReq)), "alloc request descriptors");
    hipStream_t s = c->stream;
    std::vector<ReqDesc> desc(static_cast<size_t>(nreq));
    std::vector<KnnReq> kdesc(static_cast<size_t>(nreq));
    for (int r = 0; r < nreq; ++r) {
        const auto &q = reqs[size_t(r)];
        double4 *sd = c->seeds.as<double4>() + q.base;
        HIPCHK(c, hipMemcpyAsync(sd, xyzw[r], size_t(q.n) * 32, hipMemcpyHostToDevice, s));
        // the request's own binning (as alone), then its ids made global
        int *perm = c->perm.as<int>() + q.base;
        int *cnt = c->counters.as<int>() + q.coff;
        const int64_t nc = int64_t(q.dim) * q.dim * q.dim;
        RPDCHK(c,
               rpd_bin_seeds(s, sd, q.n, kprec, q.dim, c->bin_scratch, perm, cnt,
                             static_cast<char *>(c->stored.ptr) + size_t(q.base) * vsz,
                             c->stored64.as<double4>() + q.base),
               "rpd_solve_batch: binning");
        RPDCHK(c, rpd_add_base(s, perm, q.n, int(q.base)), "rpd_solve_batch: binning");
        RPDCHK(c, rpd_add_base(s, cnt, nc + 1, int(q.base)), "rpd_solve_batch: binning");
        HIPCHK(c, hipMemsetD16Async(reinterpret_cast<hipDeviceptr_t>(c->req_of.as<uint16_t>() + q.base),
                                    uint16_t(r), size_t(q.n), s));
        ReqDesc &d = desc[size_t(r)];
        d = ReqDesc{};
        d.n_seeds = q.n;
        d.wmin = prm->security_mode == 0 && q.wmin < INFINITY ? q.wmin : 0.0;
        if (q.dom) {
            const DomainDev &dv = q.dom->dev;
            d.g = dv.g;
            d.vs = dv.vs;
            d.flags = dv.flags;
            d.centre = dv.centre;
            d.gpts = dv.pts;
            d.tris = dv.tris;
            d.tlist = dv.tlist;
            d.toffs = dv.toffs;
            d.tvoffs = dv.tvoffs;
            d.tvox = dv.tvox;
            d.tinfo = dv.tinfo;
            d.dist = prm->cull ? dv.dist : nullptr;
        }
        kdesc[size_t(r)] = KnnReq{cnt, q.dim};
    }
    HIPCHK(c, hipMemcpyAsync(c->req_desc.ptr, desc.data(), desc.size() * sizeof(ReqDesc), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->knn_req.ptr, kdesc.data(), kdesc.size() * sizeof(KnnReq), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipStreamSynchronize(s));  // the host vectors above are pageable
    c->n = total;
    c->bin_prec = kprec;
    c->dim = 0;
    c->batch = std::move(reqs);
    return solve_impl(c, nullptr, prm, stats);
}

int rpd_batch_range(const rpd_ctx *c, int32_t r, int64_t *first, int64_t *count) {
    if (!c || !first || !count || r < 0 || size_t(r) >= c->batch.size()) return RPD_E_ARG;
    *first = c->batch[size_t(r)].base;
    *count = c->batch[size_t(r)].n;
    return RPD_OK;
}

int rpd_read_moments_range(rpd_ctx *c, int64_t first, int64_t count, double *out4) {
    if (!c || !out4 || first < 0 || count < 0 || first + count > c->n)
        return set_err(c, RPD_E_ARG, "rpd_read_moments_range: bad range");
    if (!c->solved) return set_err(c, RPD_E_STATE, "rpd_read_moments_range: no solve");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(out4, c->out.as<double4>() + first, size_t(count) * 32, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_read_moments(rpd_ctx *c, double *out4n) {
    if (!c || !out4n) return set_err(c, RPD_E_ARG, "rpd_read_moments");
    if (!c->solved) return set_err(c, RPD_E_STATE, "rpd_read_moments: no solve");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(out4n, c->out.ptr, size_t(c->n) * 32, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_read_hessian(rpd_ctx *c, double *vals, int32_t *ids) {
    if (!c || !vals || !ids) return set_err(c, RPD_E_ARG, "rpd_read_hessian");
    if (!c->solved || c->hess_H <= 0) return set_err(c, RPD_E_STATE, "rpd_read_hessian: no solve with H > 0");
    hipSetDevice(c->device);
    const size_t m = size_t(c->n) * size_t(c->hess_H);
    HIPCHK(c, hipMemcpyAsync(vals, c->hess.ptr, m * c->hess_HC * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ids, c->hids.ptr, m * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_hessian_device(rpd_ctx *c, const double **vals, const int32_t **ids, int32_t *H, int32_t *HC) {
    if (!c || !vals || !ids || !H || !HC) return set_err(c, RPD_E_ARG, "rpd_hessian_device");
    if (!c->solved || c->hess_H <= 0) return set_err(c, RPD_E_STATE, "rpd_hessian_device: no solve with H > 0");
    *vals = c->hess.as<double>();
    *ids = c->hids.as<int32_t>();
    *H = c->hess_H;
    *HC = c->hess_HC;
    return RPD_OK;
}

int rpd_moments_device(rpd_ctx *c, const double **p) {
    if (!c || !p) return set_err(c, RPD_E_ARG, "rpd_moments_device");
    if (!c->solved) return set_err(c, RPD_E_STATE, "rpd_moments_device: no solve");
    *p = c->out.as<double>();
    return RPD_OK;
}

int rpd_read_status(rpd_ctx *c, int32_t *out) {
    if (!c || !out) return set_err(c, RPD_E_ARG, "rpd_read_status");
    if (!c->solved) return set_err(c, RPD_E_STATE, "rpd_read_status: no solve");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(out, c->status.ptr, size_t(c->n) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_read_perm(rpd_ctx *c, int32_t *out) {
    if (!c || !out || c->bin_prec < 0) return set_err(c, RPD_E_STATE, "rpd_read_perm: no binning yet");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(out, c->perm.ptr, size_t(c->n) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

}  // extern "C"

namespace {

// Used cells, bordering rings and the compacted output on the device
// (compressResult and findBordering, voronoi.cpp:513-751): fills c->cout,
// c->cperm and c->ccount.
int compact_device(rpd_ctx *c, double keep_volume, int32_t keep_distance) {
    if (!c->solved) return set_err(c, RPD_E_STATE, "rpd_compact_used: no solve");
    if (keep_distance > 254) return set_err(c, RPD_E_ARG, "rpd_compact_used: keep_distance > 254");
    hipSetDevice(c->device);
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    RPDCHK(c, c->ckeep.ensure(size_t(n + 1) * 4), "alloc compaction");
    RPDCHK(c, c->cring.ensure(size_t(n) * 4), "alloc compaction");
    RPDCHK(c, c->cpos.ensure(size_t(n + 1) * 4), "alloc compaction");
    RPDCHK(c, c->cout.ensure(size_t(n) * 32), "alloc compaction");
    RPDCHK(c, c->cperm.ensure(size_t(n) * 4), "alloc compaction");
    size_t scan_bytes = 0;
    HIPCHK(c, rocprim::exclusive_scan(nullptr, scan_bytes, c->ckeep.as<int>(), c->cpos.as<int>(), 0, size_t(n + 1),
                                      rocprim::plus<int>(), s));
    RPDCHK(c, c->ctmp.ensure(scan_bytes), "alloc compaction");
    const unsigned nb1 = unsigned((n + 256) / 256);  // n + 1 threads: the mark kernel also ends the array
    auto finish = [&](const int *used, int last) -> int {
        compact_mark_kernel<<<nb1, 256, 0, s>>>(used, c->perm.as<int>(), n, c->ckeep.as<int>(), c->cring.as<int>());
        HIPCHK(c, rocprim::exclusive_scan(c->ctmp.ptr, scan_bytes, c->ckeep.as<int>(), c->cpos.as<int>(), 0,
                                          size_t(n + 1), rocprim::plus<int>(), s));
        compact_scatter_kernel<<<nb1, 256, 0, s>>>(c->out.as<double4>(), c->ckeep.as<int>(), c->cring.as<int>(),
                                                  c->cpos.as<int>(), n, keep_volume, last, c->cout.as<double4>(),
                                                  c->cperm.as<int>());
        HIPCHK(c, hipMemcpyAsync(c->pin + kPinCounters - 1, c->cpos.as<int>() + n, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        c->ccount = c->pin[kPinCounters - 1];
        return RPD_OK;
    };
    if (keep_distance < 0) return finish(nullptr, -1);  // every cell
    if (keep_distance > 0 && !c->batch.empty())
        return set_err(c, RPD_E_STATE, "rpd_compact_used: no bordering rings for a packed batch");

    RPDCHK(c, c->used.ensure(size_t(n) * 4), "alloc used");
    RPDCHK(c, c->ids1.ensure(size_t(n) * 4), "alloc ids");
    const unsigned nb = unsigned((n + 255) / 256);
    used_init_kernel<<<nb, 256, 0, s>>>(c->out.as<double4>(), c->perm.as<int>(), n, keep_volume, c->used.as<int>());
    if (keep_distance > 0) {
        RPDCHK(c, c->bout.ensure(size_t(n) * 32), "alloc bordering outputs");
        RPDCHK(c, c->bstatus.ensure(size_t(n) * 4), "alloc bordering outputs");
    }
    for (int k = 0; k < keep_distance; ++k) {
        // the cells of ring k, then their unrestricted cells (FP64, no cull:
        // compute_bordering, ConvexCell.cl:754-807) mark their neighbours
        HIPCHK(c, hipMemsetAsync(c->cnt.ptr, 0, sizeof(int), s));
        ring_select_kernel<<<nb, 256, 0, s>>>(c->used.as<int>(), c->out.as<double4>(), c->perm.as<int>(), n,
                                              keep_volume, k, c->ids1.as<int>(), c->cnt.as<int>());
        int cnt = 0;
        HIPCHK(c, hipMemcpyAsync(&cnt, c->cnt.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (cnt <= 0) break;
        int K = std::min<int64_t>(std::max(c->last_K, 1), n - 1), P = c->last_P, V = c->last_V;
        for (int step = 0; cnt > 0 && step <= std::max(c->last_steps, 0); ++step) {
            PassBufs &b = c->pb[0];
            const size_t stride = rpd_clip_record_bytes(RPD_PREC_FP64, P, V);
            RPDCHK(c, b.knn.ensure(size_t(cnt) * K * 4), "alloc knn");
            RPDCHK(c, b.rec.ensure(size_t(cnt) * stride), "alloc records");
            RPDCHK(c, b.qcount.ensure(sizeof(int)), "alloc qcount");
            HIPCHK(c, hipMemsetAsync(c->fl[0].cnt.ptr, 0, kFailCounters * sizeof(int), s));
            RPDCHK(c,
                   rpd_launch_knn(s, c->bin_prec, c->stored.ptr, c->perm.as<int>(), c->counters.as<int>(), c->dim,
                                  c->ids1.as<int>(), cnt, K, b.knn.as<int>()),
                   "bordering knn");
            ClipLaunch L{};
            L.seeds = c->stored64.as<double4>();
            L.perm = c->perm.as<int>();
            L.ids = c->ids1.as<int>();
            L.knn = b.knn.as<int>();
            L.n_ids = cnt;
            L.n_seeds = n;
            L.K = K;
            L.P = P;
            L.V = V;
            L.T = 1;
            L.apex_mode = 1;
            L.security_mode = c->last_sm;
            L.wmin = c->wmin_ok && c->wmin < INFINITY ? c->wmin : 0.0;
            L.cull = false;
            L.dom = nullptr;
            L.out = c->bout.as<double4>();
            L.status = c->bstatus.as<int>();
            L.rec = b.rec.ptr;
            L.qcount = b.qcount.as<int>();
            L.fl_cnt = c->fl[0].cnt.as<int>();
            L.fl_eo = c->fl[0].eo.as<int>() + n;
            L.fl_sec = c->fl[0].sec.as<int>();
            L.fl_cnt_host = c->pin + kPinCounters;
            RPDCHK(c, rpd_launch_cells(s, RPD_PREC_FP64, L, c->pin + 0, true), "bordering cells");
            const int nq = c->pin[0];
            const int S = V <= 64 ? 1 : (V <= 128 ? 2 : 4);
            if (nq > 0)
                mark_neighbours_kernel<<<unsigned((nq + 3) / 4), 256, 0, s>>>(
                    static_cast<const char *>(b.rec.ptr), int(stride), P, S, nq, c->used.as<int>(), k + 1);
            // cells that overflowed or ran out of neighbours: again with grown caps
            const int *h = c->pin + kPinCounters + 2;
            const int n_over = c->pin[kPinCounters + 1];
            if (n_over <= 0) break;
            if (h[RPD_VERTEX_OVERFLOW]) V = std::min(256, int(V * 1.5));
            if (h[RPD_PLANE_OVERFLOW]) P = std::min(255, int(P * 1.5));
            if (h[RPD_SECURITY_RADIUS_NOT_REACHED]) K = int(std::min<int64_t>(n - 1, std::min(256, 2 * K)));
            HIPCHK(c, hipMemcpyAsync(c->ids1.ptr, c->fl[0].eo.as<int>() + n, size_t(n_over) * 4,
                                     hipMemcpyDeviceToDevice, s));
            cnt = n_over;
        }
        HIPCHK(c, hipStreamSynchronize(s));
    }
    // compaction (compute_permutation / compact_result, voronoi.cpp:513-691) in
    // ascending original index; a cell of the last ring without volume is
    // flagged with w = NO_CELL_LIMITS
    return finish(c->used.as<int>(), keep_distance);
}

}  // namespace

extern "C" {

int rpd_read_debug_counters(rpd_ctx *c, int32_t pass, int32_t *out) {
    if (!c || !out || pass < 0 || pass >= RPD_MAX_PASSES) return set_err(c, RPD_E_ARG, "rpd_read_debug_counters");
    if (!c->solved || !c->dcnt_ok) return set_err(c, RPD_E_STATE, "rpd_read_debug_counters: no solve with counters");
    hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(out, c->dcnt.as<int>() + size_t(pass) * kDcStride, kDcStride * sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RPD_OK;
}

int rpd_compact_used_device(rpd_ctx *c, double keep_volume, int32_t keep_distance, int64_t *out_count,
                            const double **out4, const int32_t **perm_to_user) {
    if (!c || !out_count) return set_err(c, RPD_E_ARG, "rpd_compact_used_device");
    const int r = compact_device(c, keep_volume, keep_distance);
    if (r != RPD_OK) return r;
    *out_count = c->ccount;
    if (out4) *out4 = c->cout.as<double>();
    if (perm_to_user) *perm_to_user = c->cperm.as<int32_t>();
    return RPD_OK;
}

int rpd_compact_used(rpd_ctx *c, double keep_volume, int32_t keep_distance, int64_t *out_count, double *out4,
                     int32_t *perm_to_user) {
    if (!c || !out_count) return set_err(c, RPD_E_ARG, "rpd_compact_used");
    const int r = compact_device(c, keep_volume, keep_distance);
    if (r != RPD_OK) return r;
    const int64_t k = c->ccount;
    if (k > 0 && out4)
        HIPCHK(c, hipMemcpyAsync(out4, c->cout.ptr, size_t(k) * 32, hipMemcpyDeviceToHost, c->stream));
    if (k > 0 && perm_to_user)
        HIPCHK(c, hipMemcpyAsync(perm_to_user, c->cperm.ptr, size_t(k) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out_count = k;
    return RPD_OK;
}

int rpd_compact(rpd_ctx *c, double keep_volume, int64_t *out_count, double *out4, int32_t *perm_to_user) {
    // the used cells alone: keep distance 0 (no ring, no flags)
    return rpd_compact_used(c, keep_volume, 0, out_count, out4, perm_to_user);
}

}  // extern "C"
