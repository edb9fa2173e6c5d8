// rpd_knn.hip -- seed binning and exact K-nearest-neighbour search on gfx950.
//
// Replaces the reference's KNearests (knearests.cpp:48-315) and
// build_knearests (knearests.cl:36-78):
//   * binning: cell = floor(p * dim / 1000) per axis (knearests.cpp:35-46) in
//     the seed precision; a stable radix sort of (cell, original index) gives
//     the permutation, so seeds keep original order inside a cell;
//   * query: one wavefront per query seed over the sorted array.  Candidate
//     cells are visited shell by shell (Chebyshev rings with per-axis bounds);
//     each row of a shell is one contiguous range of the sorted array, so the
//     64 lanes load consecutive seeds.  Candidates below the running K-th key
//     are kept in an LDS buffer; when it fills, and at the end of each shell,
//     a wave radix select finds the exact K-th key and the buffer keeps the K
//     keys up to it.  The search stops once the K-th distance is below the
//     lower bound of the next shell, and one bitonic sort orders the K kept
//     keys, so the result is the exact K nearest under the key (d^2, original
//     index), d^2 = ((dx*dx + dy*dy) + dz*dz) unfused.

#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rpd_internal.h"
#include "rpd_wave.h"

namespace rpd {

template <typename R>
struct Vec4T;
template <>
struct Vec4T<float> {
    using T = float4;
};
template <>
struct Vec4T<double> {
    using T = double4;
};

// floor(c * (R)dim / 1000) with the products and the quotient rounded once
// each, clamped to [0, dim-1] (knearests.cpp:35-46).
__device__ __forceinline__ int cell_coord(float c, int dim) {
    const float t = __fdiv_rn(__fmul_rn(c, float(dim)), 1000.0f);
    int i = int(floorf(t));
    return max(0, min(i, dim - 1));
}
__device__ __forceinline__ int cell_coord(double c, int dim) {
    const double t = __ddiv_rn(__dmul_rn(c, double(dim)), 1000.0);
    int i = int(floor(t));
    return max(0, min(i, dim - 1));
}

template <typename R>
__global__ void bin_cells_kernel(const double4 *__restrict__ seeds, int64_t n, int dim, int *__restrict__ cell,
                                 int *__restrict__ counts, int *__restrict__ rank) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double4 s = seeds[i];
    const R x = R(s.x), y = R(s.y), z = R(s.z);
    const int c = cell_coord(x, dim) + cell_coord(y, dim) * dim + cell_coord(z, dim) * dim * dim;
    cell[i] = c;
    rank[i] = atomicAdd(&counts[c], 1);  // a place in the cell, in no set order
}

// Seed i to its place in its cell's range of perm (order within a cell set
// by the atomics; sort_cells_kernel fixes it).
__global__ void scatter_cells_kernel(const int *__restrict__ cell, const int *__restrict__ rank, int64_t n,
                                     const int *__restrict__ counters, int *__restrict__ perm) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    perm[counters[cell[i]] + rank[i]] = int(i);
}

// Each cell's range of perm in increasing original index, so that perm is
// the stable sort of the seeds by cell.  Ranges longer than kSortRun are
// left to the caller (*big set): a clustered input goes to the radix sort.
constexpr int kSortRun = 32;
__global__ void sort_cells_kernel(const int *__restrict__ counters, int64_t nc, int *__restrict__ perm,
                                  int *__restrict__ big) {
    const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (c >= nc) return;
    const int lo = counters[c], hi = counters[c + 1];
    if (hi - lo > kSortRun) {
        *big = 1;
        return;
    }
    for (int j = lo + 1; j < hi; ++j) {
        const int v = perm[j];
        int k = j - 1;
        while (k >= lo && perm[k] > v) {
            perm[k + 1] = perm[k];
            --k;
        }
        perm[k + 1] = v;
    }
}

__global__ void iota_kernel(int *__restrict__ out, int64_t n) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < n) out[i] = int(i);
}

template <typename R>
__global__ void gather_seeds_kernel(const double4 *__restrict__ seeds, const int *__restrict__ perm, int64_t n,
                                    typename Vec4T<R>::T *__restrict__ stored, double4 *__restrict__ stored64) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double4 s = seeds[perm[i]];
    typename Vec4T<R>::T v;
    v.x = R(s.x);
    v.y = R(s.y);
    v.z = R(s.z);
    v.w = R(s.w);
    stored[i] = v;
    if (stored64) stored64[i] = s;
}

// ---------------------------------------------------------------------------
// top-K keys: (d^2 bits, original index) ordered, stored index as payload
// ---------------------------------------------------------------------------
template <typename R>
struct Key;

template <>
struct Key<float> {
    uint64_t k;  // (bits(d2) << 32) | original index  (d2 >= 0: bit order == value order)
    uint32_t s;  // stored index
    __device__ static Key inf() { return Key{~0ull, 0xffffffffu}; }
    __device__ static Key make(float d2, uint32_t orig, uint32_t st) {
        return Key{(uint64_t(__float_as_uint(d2)) << 32) | orig, st};
    }
    __device__ bool operator<(const Key &o) const { return k < o.k; }
    __device__ bool is_inf() const { return k == ~0ull; }
    __device__ double dist2() const { return double(__uint_as_float(uint32_t(k >> 32))); }
    __device__ Key xor_shfl(int m) const { return Key{xor_lane(k, m), xor_lane(s, m)}; }
    // key of lane - 1 (lane 0: inf)
    __device__ Key up_shfl() const { return Key{shr_lane1(k, ~0ull), shr_lane1(s, 0xffffffffu)}; }
    __device__ Key lane(int l) const { return Key{rlane(k, l), rlane(s, l)}; }
    // the ordering key as 32-bit words, most significant first (radix select)
    static constexpr int kWords = 2;
    __device__ uint32_t word(int w) const { return w == 0 ? uint32_t(k >> 32) : uint32_t(k); }
    // the largest key whose first word is w0 (no real key equals it: its
    // other bits are all ones)
    __device__ static Key top_of(uint32_t w0) { return Key{(uint64_t(w0) << 32) | 0xffffffffu, 0xffffffffu}; }
    // candidate buffer in LDS: ordering keys and stored indices
    struct Buf {
        uint64_t *k;
        uint32_t *s;
    };
    __device__ void put(const Buf &b, int i) const {
        b.k[i] = k;
        b.s[i] = s;
    }
    __device__ static Key get(const Buf &b, int i) { return Key{b.k[i], b.s[i]}; }
    static constexpr int kBufBytes = 12;  // per entry
};

template <>
struct Key<double> {
    uint64_t d;  // bits(d2)
    uint32_t o;  // original index
    uint32_t s;  // stored index
    __device__ static Key inf() { return Key{~0ull, 0xffffffffu, 0xffffffffu}; }
    __device__ static Key make(double d2, uint32_t orig, uint32_t st) {
        return Key{uint64_t(__double_as_longlong(d2)), orig, st};
    }
    __device__ bool operator<(const Key &b) const { return d < b.d || (d == b.d && o < b.o); }
    __device__ bool is_inf() const { return d == ~0ull && o == 0xffffffffu; }
    __device__ double dist2() const { return __longlong_as_double((long long)d); }
    __device__ Key xor_shfl(int m) const { return Key{xor_lane(d, m), xor_lane(o, m), xor_lane(s, m)}; }
    // key of lane - 1 (lane 0: inf)
    __device__ Key up_shfl() const {
        return Key{shr_lane1(d, ~0ull), shr_lane1(o, 0xffffffffu), shr_lane1(s, 0xffffffffu)};
    }
    __device__ Key lane(int l) const { return Key{rlane(d, l), rlane(o, l), rlane(s, l)}; }
    static constexpr int kWords = 3;
    __device__ uint32_t word(int w) const { return w == 0 ? uint32_t(d >> 32) : (w == 1 ? uint32_t(d) : o); }
    __device__ static Key top_of(uint32_t w0) {
        return Key{(uint64_t(w0) << 32) | 0xffffffffu, 0xffffffffu, 0xffffffffu};
    }
    struct Buf {
        uint64_t *d;
        uint32_t *o, *s;
    };
    __device__ void put(const Buf &b, int i) const {
        b.d[i] = d;
        b.o[i] = o;
        b.s[i] = s;
    }
    __device__ static Key get(const Buf &b, int i) { return Key{b.d[i], b.o[i], b.s[i]}; }
    static constexpr int kBufBytes = 16;
};

template <typename K>
__device__ __forceinline__ K kmin(const K &a, const K &b) { return b < a ? b : a; }
template <typename K>
__device__ __forceinline__ K kmax(const K &a, const K &b) { return b < a ? a : b; }

// Bitonic network over N = 64*S elements, element e = s*64 + lane.
// desc: sort descending instead of ascending.
template <int S, typename K>
__device__ __forceinline__ void bitonic_sort(K (&v)[S], bool desc) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= 64 * S; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int js = j >> 6;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s & js) continue;
                    const int e = s * 64 + lane;
                    const bool asc = ((e & k) == 0) != desc;
                    const K lo = v[s], hi = v[s | js];
                    v[s] = asc ? kmin(lo, hi) : kmax(lo, hi);
                    v[s | js] = asc ? kmax(lo, hi) : kmin(lo, hi);
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int e = s * 64 + lane;
                    const K o = v[s].xor_shfl(j);
                    const bool asc = ((e & k) == 0) != desc;
                    const bool keep_min = ((e & j) == 0) == asc;
                    v[s] = keep_min ? kmin(v[s], o) : kmax(v[s], o);
                }
            }
        }
    }
}

// Bitonic merge of a bitonic sequence into ascending order.
template <int S, typename K>
__device__ __forceinline__ void bitonic_merge_asc(K (&v)[S]) {
    const int lane = lane_id();
#pragma unroll
    for (int j = 32 * S; j > 0; j >>= 1) {
        if (j >= 64) {
            const int js = j >> 6;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (s & js) continue;
                const K lo = v[s], hi = v[s | js];
                v[s] = kmin(lo, hi);
                v[s | js] = kmax(lo, hi);
            }
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const K o = v[s].xor_shfl(j);
                const bool keep_min = (lane & j) == 0;
                v[s] = keep_min ? kmin(v[s], o) : kmax(v[s], o);
            }
        }
    }
}

template <typename R>
struct KnnArgs {
    const typename Vec4T<R>::T *pts;  // stored order
    const int *perm;                  // stored -> original
    const int *counters;              // dim^3 + 1 exclusive prefix
    const int *ids;                   // query stored ids (nullptr: 0..n_ids-1)
    int64_t n_ids;
    int dim;
    int K;
    int *out;  // n_ids * K stored ids, -1 padded
    // nullable: per stored id, a neighbour already known; only keys above
    // its key are listed (the K after it)
    const int *lower;
    // packed batch (nullable): the grid of request req_of[q] replaces
    // (counters, dim)
    const KnnReq *reqs;
    const uint16_t *req_of;
    // nullable: the number of queries, in device memory (<= n_ids)
    const int *n_dev;
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R>
__device__ __forceinline__ R dist2_unfused(const typename Vec4T<R>::T &a, const typename Vec4T<R>::T &b) {
#pragma clang fp contract(off)
    const R dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return (dx * dx + dy * dy) + dz * dz;
}

constexpr int kKnnWavesPerBlock = 4;
// candidate slots of one wave: the sorted-array indices of the seeds of up to
// 64 row segments, expanded in pieces of this many
constexpr int kCandBuf = 256;
// kept candidates of one wave (keys below the running K-th key): room for
// the K survivors of a selection plus one batch
template <int S>
constexpr int kKeep = S == 1 ? 256 : (S == 2 ? 384 : 512);

// Buffer index of the key of rank `rank` (0-based) among the M (distinct)
// keys of the buffer, by a radix select over 8-bit digits of the key words,
// most significant first: a 256-bin histogram of the digit of the keys still
// in play, its prefix sums across the lanes, the bin that holds the rank.
// Stops as soon as the chosen bin holds one key.  Lane l owns buffer entries
// l, l + 64, ... (SL of them); hist: 256 words of LDS.
template <typename KeyT, int SL>
__device__ int select_rank(const typename KeyT::Buf &buf, int M, int rank, uint32_t *hist) {
    const int lane = lane_id();
    uint32_t w[SL][KeyT::kWords];
    uint32_t alive = 0;
#pragma unroll
    for (int j = 0; j < SL; ++j) {
        const int e = lane + 64 * j;
        if (e < M) {
            const KeyT k = KeyT::get(buf, e);
#pragma unroll
            for (int q = 0; q < KeyT::kWords; ++q) w[j][q] = k.word(q);
            alive |= 1u << j;
        } else {
#pragma unroll
            for (int q = 0; q < KeyT::kWords; ++q) w[j][q] = 0;
        }
    }
    uint4 *h4 = reinterpret_cast<uint4 *>(hist);
    // The first word's digits start at the highest bit in which two keys in
    // play differ (the keys' distances span a few octaves, so the top bits
    // are mostly shared): windows of 8 bits from there down, the last one
    // at bit 0 (windows may overlap; bits every key shares make one bin).
    int sh0 = -1;  // -1: every key in play has the same first word
    {
        uint32_t mine = 0;
        bool has = false;
#pragma unroll
        for (int j = 0; j < SL; ++j)
            if (!has && ((alive >> j) & 1)) {
                mine = w[j][0];
                has = true;
            }
        const uint32_t ref = rlane(mine, lowest_bit(ballot(has)));
        uint32_t diff = 0;
#pragma unroll
        for (int j = 0; j < SL; ++j)
            if ((alive >> j) & 1) diff |= w[j][0] ^ ref;
        diff = uint32_t(rlane(wave_incl_scan_op(int(diff), 0, [](int a, int b) { return a | b; }), 63));
        if (diff) sh0 = max(31 - __clz(int(diff)) - 7, 0);
    }
    // words in an unrolled loop, digits of a word in a rolled one: the word
    // index stays a constant, so w[j][q] is a register, not a select chain
    bool one = false;
#pragma unroll
    for (int q = 0; q < KeyT::kWords; ++q) {
        int sh = q == 0 ? sh0 : 24;
#pragma unroll 1
        for (; sh >= 0 && !one; sh = sh > 0 && sh < 8 ? 0 : sh - 8) {
            h4[lane] = make_uint4(0u, 0u, 0u, 0u);
            wave_lds_sync();
#pragma unroll
            for (int j = 0; j < SL; ++j)
                if ((alive >> j) & 1) atomicAdd(&hist[(w[j][q] >> sh) & 255u], 1u);
            wave_lds_sync();
            const uint4 h = h4[lane];
            const int sum = int(h.x + h.y + h.z + h.w);
            const int incl = wave_incl_scan_op(sum, 0, [](int a, int b) { return a + b; });
            const int excl = incl - sum;
            const int L = lowest_bit(ballot(excl <= rank && rank < incl));
            // the bin of lane L's four that holds the rank
            int before = rlane(excl, L);
            const int c0 = int(rlane(h.x, L)), c1 = int(rlane(h.y, L)), c2 = int(rlane(h.z, L)),
                      c3 = int(rlane(h.w, L));
            int bin = 4 * L, cnt = c0;
            if (rank >= before + c0) {
                before += c0;
                bin += 1;
                cnt = c1;
                if (rank >= before + c1) {
                    before += c1;
                    bin += 1;
                    cnt = c2;
                    if (rank >= before + c2) {
                        before += c2;
                        bin += 1;
                        cnt = c3;
                    }
                }
            }
            rank -= before;
#pragma unroll
            for (int j = 0; j < SL; ++j)
                if (((w[j][q] >> sh) & 255u) != uint32_t(bin)) alive &= ~(1u << j);
            wave_lds_sync();  // the histogram is cleared again by the next pass
            one = cnt == 1;
        }
    }
    // the one key left (keys are distinct, so one is left after every digit)
    const int L = lowest_bit(ballot(alive != 0));
    const int jj = __ffs(int(rlane(int(alive), L))) - 1;
    return L + 64 * jj;
}

// An upper bound of the key of rank `rank` among the M keys of the buffer,
// from the first digit pass of select_rank alone: the largest key of the bin
// that holds the rank (the keys' first words from the highest bit in which two
// differ, 8 bits of them).  The bin's keys and those below it number at least
// rank + 1, so the key of that rank is at most the bound, and no key equals
// it.  false: every key in play has the same first word (no bound here).
template <typename KeyT, int SL>
__device__ bool bound_rank(const typename KeyT::Buf &buf, int M, int rank, uint32_t *hist, KeyT &ub) {
    const int lane = lane_id();
    uint32_t w[SL];
    uint32_t alive = 0;
#pragma unroll
    for (int j = 0; j < SL; ++j) {
        const int e = lane + 64 * j;
        w[j] = e < M ? KeyT::get(buf, e).word(0) : 0u;
        if (e < M) alive |= 1u << j;
    }
    uint32_t mine = 0;
    bool has = false;
#pragma unroll
    for (int j = 0; j < SL; ++j)
        if (!has && ((alive >> j) & 1)) {
            mine = w[j];
            has = true;
        }
    const uint32_t ref = rlane(mine, lowest_bit(ballot(has)));
    uint32_t diff = 0;
#pragma unroll
    for (int j = 0; j < SL; ++j)
        if ((alive >> j) & 1) diff |= w[j] ^ ref;
    diff = uint32_t(rlane(wave_incl_scan_op(int(diff), 0, [](int a, int b) { return a | b; }), 63));
    if (diff == 0) return false;
    const int sh = max(31 - __clz(int(diff)) - 7, 0);
    uint4 *h4 = reinterpret_cast<uint4 *>(hist);
    h4[lane] = make_uint4(0u, 0u, 0u, 0u);
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < SL; ++j)
        if ((alive >> j) & 1) atomicAdd(&hist[(w[j] >> sh) & 255u], 1u);
    wave_lds_sync();
    const uint4 h = h4[lane];
    const int sum = int(h.x + h.y + h.z + h.w);
    const int incl = wave_incl_scan_op(sum, 0, [](int a, int b) { return a + b; });
    const int excl = incl - sum;
    const int L = lowest_bit(ballot(excl <= rank && rank < incl));
    int before = rlane(excl, L);
    const int c0 = int(rlane(h.x, L)), c1 = int(rlane(h.y, L)), c2 = int(rlane(h.z, L));
    int bin = 4 * L;
    if (rank >= before + c0) {
        before += c0;
        bin += 1;
        if (rank >= before + c1) {
            before += c1;
            bin += 1;
            if (rank >= before + c2) bin += 1;
        }
    }
    wave_lds_sync();  // the histogram is cleared again by its next user
    // bits above the window are every key's; the bin's bits, then all ones
    const uint32_t low = sh + 8 >= 32 ? 0xffffffffu : (1u << (sh + 8)) - 1u;
    ub = KeyT::top_of((ref & ~low) | (uint32_t(bin) << sh) | ((1u << sh) - 1u));
    return true;
}

// The K nearest of query qi (its stored id ids[qi]): the wave's LDS pieces
// are cand (kCandBuf ints), hist (256 words) and the key buffer buf.
template <typename R, int S>
__device__ __forceinline__ void knn_query(const KnnArgs<R> &a, int64_t qi, int lane, int *cand, uint32_t *hist,
                                          const typename Key<R>::Buf &buf) {
    using KeyT = Key<R>;
    constexpr int KEEP = kKeep<S>;
    constexpr int SL = KEEP / 64;
    const int q = uni(a.ids ? a.ids[qi] : int(qi));
    const typename Vec4T<R>::T p = a.pts[q];
    int dim = a.dim;
    const int *counters = a.counters;
    if (a.reqs) {
        const KnnReq &rq = a.reqs[uni(int(a.req_of[q]))];
        dim = rq.dim;
        counters = rq.counters;
    }
    const int ci = cell_coord(R(p.x), dim), cj = cell_coord(R(p.y), dim), ck = cell_coord(R(p.z), dim);
    const int K = a.K;
    const double cs = 1000.0 / double(dim);

    // Candidates whose key is below the running K-th key are kept in the
    // buffer (M of them); once M reaches K, a selection finds the exact K-th
    // key of the buffer and the buffer keeps only the K keys up to it.
    int M = 0;
    KeyT kth = KeyT::inf();
    const bool has_low = a.lower != nullptr;
    KeyT klow = KeyT::inf();
    if (has_low) {
        const int L = a.lower[q];
        klow = KeyT::make(dist2_unfused<R>(p, a.pts[L]), uint32_t(a.perm[L]), uint32_t(L));
    }
    // the buffer cut to its keys <= kth
    auto cut = [&]() {
        KeyT v[SL];
        bool kp[SL];
#pragma unroll
        for (int j = 0; j < SL; ++j) {
            const int e = lane + 64 * j;
            kp[j] = e < M;
            v[j] = kp[j] ? KeyT::get(buf, e) : KeyT::inf();
            kp[j] = kp[j] && !(kth < v[j]);
        }
        wave_lds_sync();  // every entry is read before any is moved
        int base = 0;
#pragma unroll
        for (int j = 0; j < SL; ++j) {
            const uint64_t m = ballot(kp[j]);
            if (kp[j]) v[j].put(buf, base + mbcnt(m));
            base += __popcll(m);
        }
        M = base;
        wave_lds_sync();
    };
    // exact K-th key of the buffer, and the buffer cut to the K keys <= it
    auto shrink = [&]() {
        kth = KeyT::get(buf, select_rank<KeyT, SL>(buf, M, K - 1, hist));
        cut();
    };
    auto add_batch = [&](const KeyT &c) {
        const bool keep = !c.is_inf() && c < kth && (!has_low || klow < c);
        const uint64_t m = ballot(keep);
        if (m == 0) return;
        if (M + 64 > KEEP) shrink();  // M >= K here: KEEP >= K + 64
        const bool keep2 = keep && c < kth;
        const uint64_t m2 = ballot(keep2);
        if (keep2) c.put(buf, M + mbcnt(m2));
        M += __popcll(m2);
        wave_lds_sync();
    };

    // Cells whose box lies farther from the query than the current K-th
    // distance cannot hold a closer candidate, so from the second shell on,
    // rows and the x range of a row are cut to the ball of that radius (the
    // K-th key only decreases).  The bound is widened by a relative 1e-5 and
    // the boxes by kCullPad, far above the rounding of d^2, of the FP32 box
    // arithmetic (coordinates below 1000: ulp 6e-5) and of the cell index, so
    // no candidate at or below the K-th key is cut.
    constexpr float kCullPad = 0.01f;
    const float px = float(p.x), py = float(p.y), pz = float(p.z);
    const float csf = float(cs), inv_cs = float(1.0 / cs);
    auto box_gap = [&](int c, int cq, float v) -> float {
        if (c < cq) return fmaxf(0.0f, v - float(c + 1) * csf - kCullPad);
        if (c > cq) return fmaxf(0.0f, float(c) * csf - v - kCullPad);
        return 0.0f;
    };
    for (int r = 1;; ++r) {
        const int side = 2 * r + 1;
        const int nrows = side * side;
        const int nseg = 2 * nrows;
        const bool cull = r >= 2 && !kth.is_inf();
        const float k2 = cull ? float(kth.dist2() * (1.0 + 1e-5)) * (1.0f + 1e-6f) : 0.0f;
        // row t of the shell: (t % side, t / side); t < 2 side^2, and the
        // rounding of (t + 1/2) / side in FP32 stays far from an integer
        const float inv_side = 1.0f / float(side);
        for (int chunk = 0; chunk < nseg; chunk += 64) {
            const int u = chunk + lane;
            int sa = 0, sb = 0;
            if (u < nseg) {
                const int t = u >> 1, part = u & 1;
                const int tz = int((float(t) + 0.5f) * inv_side);
                const int dy = t - tz * side - r, dz = tz - r;
                const int y = cj + dy, z = ck + dz;
                if (y >= 0 && y < dim && z >= 0 && z < dim) {
                    // the first shell (r == 1) is the whole 3x3x3 block
                    const bool full = r == 1 || abs(dy) == r || abs(dz) == r;
                    int xlo, xhi;
                    if (full) {
                        xlo = part == 0 ? max(0, ci - r) : 1;
                        xhi = part == 0 ? min(dim - 1, ci + r) : 0;
                    } else {
                        xlo = xhi = part == 0 ? ci - r : ci + r;
                    }
                    if (cull) {
                        const float gy = box_gap(y, cj, py), gz = box_gap(z, ck, pz);
                        const float rem = k2 - gy * gy - gz * gz;
                        if (rem < 0.0f) {
                            xlo = 1;
                            xhi = 0;
                        } else {
                            const float w = sqrtf(rem) + kCullPad;
                            xlo = max(xlo, int(floorf((px - w) * inv_cs)));
                            xhi = min(xhi, int(floorf((px + w) * inv_cs)));
                        }
                    }
                    if (xlo >= 0 && xhi < dim && xlo <= xhi) {
                        const int row = y * dim + z * dim * dim;
                        sa = counters[row + xlo];
                        sb = counters[row + xhi + 1];
                    }
                }
            }
            const int len = sb - sa;
            const int off = wave_excl_scan(len);
            const int total = uni(rlane(off + len, 63));
            // The segments' seeds, in segment order, are candidates off ..
            // off + len - 1 of the chunk.  Each lane writes the indices of
            // its segment's part of the current piece into LDS, then every
            // batch of 64 reads one index per lane.
            for (int pbase = 0; pbase < total; pbase += kCandBuf) {
                const int lo = max(off, pbase), hi = min(off + len, pbase + kCandBuf);
                for (int gi = lo; gi < hi; ++gi) cand[gi - pbase] = sa + (gi - off);
                wave_lds_sync();
                const int nhere = min(total - pbase, kCandBuf);
                for (int base = 0; base < nhere; base += 64) {
                    const int g = base + lane;
                    KeyT c = KeyT::inf();
                    if (g < nhere) {
                        const int idx = cand[g];
                        if (idx != q) {
                            const typename Vec4T<R>::T cp = a.pts[idx];
                            const R d2 = dist2_unfused<R>(p, cp);
                            c = KeyT::make(d2, uint32_t(a.perm[idx]), uint32_t(idx));
                        }
                    }
                    add_batch(c);
                }
                wave_lds_sync();  // the next piece overwrites cand
            }
        }
        // The K-th key of the candidates so far: after the first shell, only
        // a bound of it (one digit pass instead of a full select), which
        // serves the next shell's cull and the stop test; the buffer keeps
        // the keys up to the bound.  After later shells, the exact key.
        if (r == 1 && M > K) {
            KeyT ub;
            if (bound_rank<KeyT, SL>(buf, M, K - 1, hist, ub)) {
                kth = ub;
                cut();
            } else {
                shrink();
            }
        } else if (M > K || (M == K && kth.is_inf())) {
            shrink();
        }
        // stop once the K-th key is closer than anything in the next shell
        if (!kth.is_inf()) {
            const double lb = double(r) * cs * (1.0 - 1e-4);
            if (kth.dist2() < lb * lb * (1.0 - 1e-6)) break;
        }
        const int rn = r + 1;
        if (ci - rn < 0 && ci + rn >= dim && cj - rn < 0 && cj + rn >= dim && ck - rn < 0 && ck + rn >= dim) break;
    }
    // the K kept keys exactly (a search that stops after the first shell
    // holds the keys up to its bound)
    if (M > K) shrink();
    int *out = a.out + qi * int64_t(K);
    if constexpr (S == 1 && sizeof(R) == 4) {
        // K <= 64, one-word keys: each kept key's place is the number of
        // kept keys below it (keys are distinct); places past M are padding.
        // (With FP64 keys, two words, the sort below is faster.)
        // The places come from the distances' bits (one 32-bit compare per
        // key); keys of equal distance get one place, which one LDS write
        // and read per lane detects, and then the full keys decide.
        const KeyT mine = lane < M ? KeyT::get(buf, lane) : KeyT::inf();
        const uint32_t mh = uint32_t(mine.k >> 32);
        const uint32_t *hi = reinterpret_cast<const uint32_t *>(buf.k) + 1;  // high words, every other one
        int rank = 0;
#pragma unroll 4
        for (int j = 0; j < M; ++j) rank += hi[2 * j] < mh ? 1 : 0;
        int *place = reinterpret_cast<int *>(hist);  // free after the last selection (K <= 256 words)
        if (lane < M) place[rank] = lane;
        wave_lds_sync();
        const bool clash = lane < M && place[rank] != lane;
        wave_lds_sync();
        if (ballot(clash)) {
            rank = 0;
            for (int j = 0; j < M; ++j) rank += KeyT::get(buf, j) < mine ? 1 : 0;
        }
        if (lane < M)
            out[rank] = int(mine.s);
        else if (lane < K)
            out[lane] = -1;
    } else {
        // the (at most K) kept keys in ascending order: one sort
        KeyT top[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * 64 + lane;
            top[s] = e < M ? KeyT::get(buf, e) : KeyT::inf();
        }
        bitonic_sort<S>(top, false);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = s * 64 + lane;
            if (e < K) out[e] = top[s].is_inf() ? -1 : int(top[s].s);
        }
    }
}

// One wave per query, WPB waves per workgroup.  With n_dev the launch covers
// an upper bound of the queries (n_ids) and the waves past the count exit at
// once.  (A grid-stride loop over fewer workgroups was measured slower: it
// holds more registers, 67 against 56 VGPRs for FP32 keys.)  Launches that
// run beside a finish kernel use one-wave workgroups: a CU that the finish
// kernel's workgroups fill keeps less free LDS than a four-wave workgroup
// needs, so such workgroups waited until the finish kernel drained.
template <typename R, int S, int WPB>
__global__ __launch_bounds__(64 * WPB) void knn_kernel(KnnArgs<R> a) {
    using KeyT = Key<R>;
    constexpr int KEEP = kKeep<S>;
    __shared__ int cand_all[WPB][kCandBuf];
    __shared__ __attribute__((aligned(16))) uint32_t hist_all[WPB][256];
    __shared__ __attribute__((aligned(16))) char keep_all[WPB][KEEP * KeyT::kBufBytes];
    const int lane = lane_id();
    const int wv = uni(int(threadIdx.x >> 6));  // wave-uniform: lets ids, seeds and pointers live in SGPRs
    const int64_t nq = a.n_dev ? min(a.n_ids, int64_t(uni(*a.n_dev))) : a.n_ids;
    typename KeyT::Buf buf;
    {
        char *kb = keep_all[wv];
        if constexpr (sizeof(R) == 4) {
            buf.k = reinterpret_cast<uint64_t *>(kb);
            buf.s = reinterpret_cast<uint32_t *>(kb + 8 * KEEP);
        } else {
            buf.d = reinterpret_cast<uint64_t *>(kb);
            buf.o = reinterpret_cast<uint32_t *>(kb + 8 * KEEP);
            buf.s = reinterpret_cast<uint32_t *>(kb + 12 * KEEP);
        }
    }
    const int64_t qi = int64_t(blockIdx.x) * WPB + wv;
    if (qi < nq) knn_query<R, S>(a, qi, lane, cand_all[wv], hist_all[wv], buf);
}

}  // namespace rpd

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
using namespace rpd;

int rpd_knn_dim(int64_t n) {
    // knearests.cpp:63 -- round((N/3.1)^(1/3)), at least 1
    const double d = std::cbrt(double(n) / 3.1);
    return std::max(1, int(std::lround(d)));
}

// Bins n FP64 seeds in precision prec (0 float, 1 double).  Outputs (device):
// perm[n], counters[dim^3+1], stored (float4 or double4 by prec) and stored64.
// perm is the stable sort of the seeds by cell (knearests.cpp:129-191 sorts
// by cell with a counting sort; here the order within a cell is fixed too):
// a counting sort whose cells are then sorted by index, or, when a cell holds
// more than kSortRun seeds, a radix sort of (cell, index).  Waits for the
// stream.
// Bytes of binning scratch for n seeds on a dim^3 grid (0 on a HIP error).
// Layout: cell_in, cell_out, rank / orig (n ints each), counts (dim^3 + 1),
// the crowded-cell flag, rocPRIM temporary storage.

// Layout of the binning's scratch: cell ids (in, out), ranks, per-cell
// counts, the crowded-cell flag, the library sort's / scan's temporary.
struct BinLayout {
    size_t off_cell_out, off_rank, off_counts, off_flag, off_tmp, sort_bytes, scan_bytes, need;
    unsigned end_bit;
};
static int bin_layout(hipStream_t st, int64_t n, int dim, BinLayout &b) {
    const int64_t nc = int64_t(dim) * dim * dim;
    const size_t bytes_cell = size_t(n) * 4;
    int *null_i = nullptr;
    b.end_bit = 1;
    while ((int64_t(1) << b.end_bit) < nc) ++b.end_bit;
    b.sort_bytes = b.scan_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, b.sort_bytes, null_i, null_i, null_i, null_i, n, 0u, b.end_bit, st) !=
        hipSuccess)
        return RPD_E_DEVICE;
    if (rocprim::exclusive_scan(nullptr, b.scan_bytes, null_i, null_i, 0, size_t(nc + 1), rocprim::plus<int>(), st) !=
        hipSuccess)
        return RPD_E_DEVICE;
    b.off_cell_out = bytes_cell;
    b.off_rank = 2 * bytes_cell;
    b.off_counts = 3 * bytes_cell;
    b.off_flag = (b.off_counts + size_t(nc + 1) * 4 + 255) & ~size_t(255);
    b.off_tmp = b.off_flag + 256;
    b.need = b.off_tmp + std::max(b.sort_bytes, b.scan_bytes) + 256;
    return RPD_OK;
}

size_t rpd_bin_scratch_bytes(hipStream_t st, int64_t n, int dim) {
    BinLayout b;
    return bin_layout(st, n, dim, b) == RPD_OK ? b.need : 0;
}

template <typename R>
static void gather_seeds(hipStream_t st, const double4 *seeds, const int *perm, int64_t n, void *stored,
                         double4 *stored64) {
    const int tb = 256;
    const unsigned nb = unsigned((n + tb - 1) / tb);
    gather_seeds_kernel<R><<<nb, tb, 0, st>>>(seeds, perm, n, static_cast<typename Vec4T<R>::T *>(stored), stored64);
}

// Queues the binning; its crowded-cell flag is copied to flag_host.
static int bin_queue(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, void *scratch_ptr,
                     size_t scratch_bytes, int *flag_host, int *perm, int *counters, void *stored,
                     double4 *stored64) {
    BinLayout b;
    if (bin_layout(st, n, dim, b) != RPD_OK) return RPD_E_DEVICE;
    if (b.need > scratch_bytes) return RPD_E_LIMIT;
    const int64_t nc = int64_t(dim) * dim * dim;
    char *base = static_cast<char *>(scratch_ptr);
    int *cell_in = reinterpret_cast<int *>(base);
    int *rank = reinterpret_cast<int *>(base + b.off_rank);
    int *counts = reinterpret_cast<int *>(base + b.off_counts);
    int *big = reinterpret_cast<int *>(base + b.off_flag);
    void *tmp = base + b.off_tmp;
    // the counts and the flag, which follows them, cleared at once
    if (hipMemsetAsync(counts, 0, b.off_flag + sizeof(int) - b.off_counts, st) != hipSuccess) return RPD_E_DEVICE;
    const int tb = 256;
    const unsigned nb = unsigned((n + tb - 1) / tb);
    if (prec == 0)
        bin_cells_kernel<float><<<nb, tb, 0, st>>>(seeds, n, dim, cell_in, counts, rank);
    else
        bin_cells_kernel<double><<<nb, tb, 0, st>>>(seeds, n, dim, cell_in, counts, rank);
    size_t scb = b.scan_bytes;
    if (rocprim::exclusive_scan(tmp, scb, counts, counters, 0, size_t(nc + 1), rocprim::plus<int>(), st) !=
        hipSuccess)
        return RPD_E_DEVICE;
    scatter_cells_kernel<<<nb, tb, 0, st>>>(cell_in, rank, n, counters, perm);
    sort_cells_kernel<<<unsigned((nc + tb - 1) / tb), tb, 0, st>>>(counters, nc, perm, big);
    // gathered at once; a crowded cell (rare) sorts and gathers again
    if (prec == 0)
        gather_seeds<float>(st, seeds, perm, n, stored, stored64);
    else
        gather_seeds<double>(st, seeds, perm, n, stored, stored64);
    if (hipMemcpyAsync(flag_host, big, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return RPD_E_DEVICE;
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

// The crowded-cell path: the permutation of a radix sort of (cell, index),
// then the gather again; waits for the stream.
static int bin_redo(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, void *scratch_ptr,
                    int *perm, void *stored, double4 *stored64) {
    BinLayout b;
    if (bin_layout(st, n, dim, b) != RPD_OK) return RPD_E_DEVICE;
    char *base = static_cast<char *>(scratch_ptr);
    int *cell_in = reinterpret_cast<int *>(base);
    int *cell_out = reinterpret_cast<int *>(base + b.off_cell_out);
    int *rank = reinterpret_cast<int *>(base + b.off_rank);
    void *tmp = base + b.off_tmp;
    const int tb = 256;
    const unsigned nb = unsigned((n + tb - 1) / tb);
    iota_kernel<<<nb, tb, 0, st>>>(rank, n);
    size_t sb = b.sort_bytes;
    if (rocprim::radix_sort_pairs(tmp, sb, cell_in, cell_out, rank, perm, n, 0u, b.end_bit, st) != hipSuccess)
        return RPD_E_DEVICE;
    if (prec == 0)
        gather_seeds<float>(st, seeds, perm, n, stored, stored64);
    else
        gather_seeds<double>(st, seeds, perm, n, stored, stored64);
    if (hipStreamSynchronize(st) != hipSuccess) return RPD_E_DEVICE;
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

int rpd_bin_seeds(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch, int *perm,
                  int *counters, void *stored, double4 *stored64) {
    const size_t need = rpd_bin_scratch_bytes(st, n, dim);
    if (need == 0) return RPD_E_DEVICE;
    if (scratch.ensure(need) != RPD_OK) return RPD_E_NOMEM;
    int flag_host = 0;
    return rpd_bin_seeds_in(st, seeds, n, prec, dim, scratch.ptr, need, &flag_host, perm, counters, stored, stored64);
}
int rpd_bin_seeds_in(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, void *scratch_ptr,
                     size_t scratch_bytes, int *flag_host, int *perm, int *counters, void *stored, double4 *stored64) {
    const int r = bin_queue(st, seeds, n, prec, dim, scratch_ptr, scratch_bytes, flag_host, perm, counters, stored,
                            stored64);
    if (r != RPD_OK) return r;
    if (hipStreamSynchronize(st) != hipSuccess) return RPD_E_DEVICE;
    if (*flag_host) return bin_redo(st, seeds, n, prec, dim, scratch_ptr, perm, stored, stored64);
    return RPD_OK;
}
int rpd_bin_seeds_async(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch,
                        int *flag_pinned, int *perm, int *counters, void *stored, double4 *stored64) {
    const size_t need = rpd_bin_scratch_bytes(st, n, dim);
    if (need == 0) return RPD_E_DEVICE;
    if (scratch.ensure(need) != RPD_OK) return RPD_E_NOMEM;
    return bin_queue(st, seeds, n, prec, dim, scratch.ptr, need, flag_pinned, perm, counters, stored, stored64);
}
int rpd_bin_seeds_redo(hipStream_t st, const double4 *seeds, int64_t n, int prec, int dim, DevBuf &scratch,
                       int *perm, void *stored, double4 *stored64) {
    return bin_redo(st, seeds, n, prec, dim, scratch.ptr, perm, stored, stored64);
}

namespace rpd {
__global__ void add_base_kernel(int *__restrict__ v, int64_t n, int base) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i < n) v[i] += base;
}
}  // namespace rpd

int rpd_add_base(hipStream_t st, int *v, int64_t n, int base) {
    if (n <= 0 || base == 0) return RPD_OK;
    add_base_kernel<<<unsigned((n + 255) / 256), 256, 0, st>>>(v, n, base);
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}

int rpd_launch_knn(hipStream_t st, int prec, const void *stored, const int *perm, const int *counters, int dim,
                   const int *ids, int64_t n_ids, int K, int *out, const int *lower, const KnnReq *kreqs,
                   const uint16_t *req_of, const int *n_dev, bool narrow) {
    if (n_ids <= 0) return RPD_OK;
    if (K <= 0 || K > 256) return RPD_E_LIMIT;
    const int wpb = narrow ? 1 : kKnnWavesPerBlock;
    const unsigned blocks = unsigned((n_ids + wpb - 1) / wpb);
    const dim3 tb(64 * wpb);
#define RPD_KNN_LAUNCH_W(R, S)                                 \
    do {                                                       \
        if (narrow)                                            \
            knn_kernel<R, S, 1><<<blocks, tb, 0, st>>>(a);     \
        else                                                   \
            knn_kernel<R, S, kKnnWavesPerBlock><<<blocks, tb, 0, st>>>(a); \
    } while (0)
#define RPD_KNN_LAUNCH(R, V4)                                                                       \
    do {                                                                                            \
        KnnArgs<R> a{static_cast<const V4 *>(stored), perm, counters, ids, n_ids, dim, K, out, lower, \
                     kreqs, req_of, n_dev};                                                         \
        if (K <= 64)                                                                                \
            RPD_KNN_LAUNCH_W(R, 1);                                                                 \
        else if (K <= 128)                                                                          \
            RPD_KNN_LAUNCH_W(R, 2);                                                                 \
        else                                                                                        \
            RPD_KNN_LAUNCH_W(R, 4);                                                                 \
    } while (0)
    if (prec == 0)
        RPD_KNN_LAUNCH(float, float4);
    else
        RPD_KNN_LAUNCH(double, double4);
#undef RPD_KNN_LAUNCH
#undef RPD_KNN_LAUNCH_W
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}
