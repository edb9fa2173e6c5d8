// rpd_cull.hip -- cells whose restriction to the domain is provably empty,
// found before the kNN search.
//
// A domain point x lies in some domain voxel v (a voxel holding triangles or
// whose centre is inside).  For any seed t, |x - t| <= |c_v - t| + r_v with
// c_v the voxel centre and r_v its half diagonal, so the smallest power
// distance U(x) = min_t |x - t|^2 + w_t of x satisfies
//     U(x) - w_min <= M = max_v min_t ((|c_v - t| + r_v)^2 + w_t - w_min).
// Here the min over t runs over the seeds of the kNN-grid cells around c_v
// (any seed gives an upper bound).  A seed s whose voxel lies d >= 2 voxels
// (Chebyshev) from every domain voxel is at least D = (d - 1) vs from every
// domain point.  When D^2 + w_s - w_min > M, every domain point has a strictly
// smaller power distance to another seed: the restricted cell of s has volume
// 0, which is what clipping it would give.  Such seeds skip the kNN search and
// the clip kernels.  Only for the exact power stop (the cell computed is the
// power cell) and without border seeds (walls are not power distances).

#include <hip/hip_runtime.h>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "rpd_internal.h"

namespace rpd {

// margin over the rounding of M and of the seed test (units of the [0,1000] frame)
constexpr double kCoverRel = 1e-9, kCoverAbs = 1e-6;

// One thread per domain voxel.  The distance to a seed is rounded up: the
// hardware FP32 square root (1 ulp) of d^2 rounded to FP32 is within 2^-22 of
// the exact one, relative, so scaling it by 1 + 2^-18 bounds it from above
// with room to spare (a d^2 that FP32 flushes to 0 is below 1e-37, far under
// kCoverAbs).  Any upper bound of M is a valid one; this one is looser by
// about 1e-5, relative, and takes a fraction of the VALU of the correctly
// rounded FP64 square root.  (Four lanes per voxel, each over some of the 9
// rows, was slower: the kernel is VALU bound and the rows do not split
// evenly.)
__global__ __launch_bounds__(256) void coverage_kernel(const double4 *__restrict__ stored64,
                                                       const int *__restrict__ counters, int dim,
                                                       const int *__restrict__ dvox, int n_dvox, int g, double vs,
                                                       double wmin, const double *__restrict__ wdev,
                                                       unsigned long long *__restrict__ mbits) {
    if (wdev) {
        if (wdev[1] <= RPD_NO_CELL_LIMITS) return;  // walls: no test (every block returns)
        wmin = wdev[0] < INFINITY ? wdev[0] : 0.0;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double best = INFINITY;
    if (i < n_dvox) {
        const int v = dvox[i];
        const double cx = (v % g + 0.5) * vs, cy = ((v / g) % g + 0.5) * vs, cz = (v / (g * g) + 0.5) * vs;
        const double rv = 0.8660254037844387 * vs * (1.0 + 1e-12);  // sqrt(3)/2, rounded up
        auto cc = [&](double c) { return max(0, min(dim - 1, int(c * double(dim) / 1000.0))); };
        const int ci = cc(cx), cj = cc(cy), ck = cc(cz);
        const int x0 = max(0, ci - 1), x1 = min(dim - 1, ci + 1);
        constexpr double kUp = 1.0 + 0x1p-18;
        for (int dz = -1; dz <= 1; ++dz) {
            const int z = ck + dz;
            if (z < 0 || z >= dim) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int y = cj + dy;
                if (y < 0 || y >= dim) continue;
                const int row = (y + z * dim) * dim;
                for (int j = counters[row + x0], je = counters[row + x1 + 1]; j < je; ++j) {
                    const double4 t = stored64[j];
                    const double ex = t.x - cx, ey = t.y - cy, ez = t.z - cz;
                    const double d2 = ex * ex + ey * ey + ez * ez;
                    const double dd = double(__builtin_amdgcn_sqrtf(float(d2))) * kUp + rv;
                    best = fmin(best, dd * dd + (t.w - wmin));
                }
            }
        }
        // no seed around this voxel: no bound
        best = best * (1.0 + kCoverRel) + kCoverAbs;
    }
    // block maximum, then one atomic per block (the values are >= 0, so their
    // bit patterns order as the values)
    __shared__ double red[256];
    red[threadIdx.x] = i < n_dvox ? best : 0.0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(mbits, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ __launch_bounds__(256) void skip_kernel(const double4 *__restrict__ stored64, const int *__restrict__ perm,
                                                   int64_t n, const uint8_t *__restrict__ dist, int g, double vs,
                                                   double wmin, const double *__restrict__ wdev,
                                                   const unsigned long long *__restrict__ mbits,
                                                   uint8_t *__restrict__ keep, double4 *__restrict__ out,
                                                   int *__restrict__ status, int *__restrict__ kused) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= n) return;
    if (wdev) {
        if (wdev[1] <= RPD_NO_CELL_LIMITS) {  // walls: every cell is kept
            keep[s] = 1;
            return;
        }
        wmin = wdev[0] < INFINITY ? wdev[0] : 0.0;
    }
    const double M = __longlong_as_double((long long)*mbits);
    const double4 p = stored64[s];
    auto vox = [&](double x) { return int(fmin(fmax(x / vs, 0.0), double(g - 1))); };
    const int d = dist[vox(p.x) + vox(p.y) * g + vox(p.z) * g * g];
    const double D = double(max(0, d - 1)) * vs;
    const bool skip = D * D + (p.w - wmin) > M * (1.0 + kCoverRel) + kCoverAbs;
    keep[s] = skip ? 0 : 1;
    if (skip) {
        const int o = perm[s];
        out[o] = make_double4(0.0, 0.0, 0.0, 0.0);
        status[o] = RPD_SUCCESS;
        if (kused) kused[o] = 0;
    }
}

}  // namespace rpd

using namespace rpd;

int rpd_cover_skip(hipStream_t st, const CoverArgs *reqs, int nreq, int64_t n_total, DevBuf &scratch, int *ids_out,
                   int *n_out, const int **count_dev) {
    if (count_dev) *count_dev = nullptr;
    if (!count_dev) *n_out = 0;
    if (n_total <= 0 || nreq <= 0) return RPD_E_ARG;
    const int64_t n = n_total;
    // scratch: M bits per request | keep flags | selected count | rocprim temp
    size_t sel_bytes = 0;
    const uint8_t *null_f = nullptr;
    int *null_i = nullptr;
    if (rocprim::select(nullptr, sel_bytes, rocprim::counting_iterator<int>(0), null_f, null_i, null_i, size_t(n), st) !=
        hipSuccess)
        return RPD_E_DEVICE;
    const size_t off_keep = (size_t(nreq) * 8 + 255) & ~size_t(255);
    const size_t off_cnt = (off_keep + size_t(n) + 255) & ~size_t(255), off_tmp = off_cnt + 256;
    if (scratch.ensure(off_tmp + sel_bytes + 256) != RPD_OK) return RPD_E_NOMEM;
    char *base = static_cast<char *>(scratch.ptr);
    auto *mbits = reinterpret_cast<unsigned long long *>(base);
    auto *keep = reinterpret_cast<uint8_t *>(base + off_keep);
    int *cnt = reinterpret_cast<int *>(base + off_cnt);
    if (hipMemsetAsync(mbits, 0, size_t(nreq) * 8, st) != hipSuccess) return RPD_E_DEVICE;
    for (int r = 0; r < nreq; ++r) {
        const CoverArgs &c = reqs[r];
        if (c.n <= 0) continue;
        if (c.base < 0 || c.base + c.n > n) return RPD_E_ARG;
        if (!c.apply) {  // every cell of this request is kept
            if (hipMemsetAsync(keep + c.base, 1, size_t(c.n), st) != hipSuccess) return RPD_E_DEVICE;
            continue;
        }
        if (c.n_dvox > 0)
            coverage_kernel<<<unsigned((c.n_dvox + 255) / 256), 256, 0, st>>>(
                c.stored64, c.counters, c.dim, c.dvox, c.n_dvox, c.g, c.vs, c.wmin, c.wdev, mbits + r);
        skip_kernel<<<unsigned((c.n + 255) / 256), 256, 0, st>>>(c.stored64 + c.base, c.perm + c.base, c.n, c.dist,
                                                                 c.g, c.vs, c.wmin, c.wdev, mbits + r, keep + c.base,
                                                                 c.out, c.status, c.kused);
    }
    size_t sb = sel_bytes;
    if (rocprim::select(base + off_tmp, sb, rocprim::counting_iterator<int>(0), keep, ids_out, cnt, size_t(n), st) !=
        hipSuccess)
        return RPD_E_DEVICE;
    if (hipMemcpyAsync(n_out, cnt, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return RPD_E_DEVICE;
    if (count_dev) {
        *count_dev = cnt;
    } else if (hipStreamSynchronize(st) != hipSuccess) {
        return RPD_E_DEVICE;
    }
    return hipGetLastError() == hipSuccess ? RPD_OK : RPD_E_DEVICE;
}
