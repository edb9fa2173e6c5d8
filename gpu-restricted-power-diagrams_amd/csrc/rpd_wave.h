// rpd_wave.h -- wave64 primitives for gfx950 (CDNA4): lane ids, ballots,
// lane reads of uniform values, scans and reductions across one wavefront.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

namespace rpd {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of m strictly below this lane
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }
__device__ __forceinline__ uint64_t uni(uint64_t v) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32))));
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ double uni(double v) {
    return __longlong_as_double((long long)uni(uint64_t(__double_as_longlong(v))));
}

// value of v in lane l (l wave-uniform)
__device__ __forceinline__ int rlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint32_t rlane(uint32_t v, int l) { return uint32_t(__builtin_amdgcn_readlane(int(v), l)); }
__device__ __forceinline__ float rlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), l));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(uint64_t(b) >> 32)), l));
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}
__device__ __forceinline__ uint64_t rlane(uint64_t v, int l) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), l));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), l));
    return (uint64_t(hi) << 32) | lo;
}

// ---- DPP (data-parallel primitives of the gfx9 VALU): row shifts inside a
// row of 16 lanes and row broadcasts of lanes 15 / 31 give a 64-lane
// inclusive scan in 6 steps without an LDS round trip.  Lanes a step does not
// write keep the identity element passed as `old`.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int CTRL, int ROW = 0xf, int BANK = 0xf>
__device__ __forceinline__ int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW, BANK, false);
}
template <int CTRL, int ROW = 0xf, int BANK = 0xf>
__device__ __forceinline__ float dpp_f(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, ROW, BANK, false));
}
template <int CTRL, int ROW = 0xf, int BANK = 0xf>
__device__ __forceinline__ double dpp_d(double old, double v) {
    const long long o = __double_as_longlong(old), b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(int(uint32_t(o)), int(uint32_t(b)), CTRL, ROW, BANK, false);
    const int hi = __builtin_amdgcn_update_dpp(int(uint32_t(uint64_t(o) >> 32)), int(uint32_t(uint64_t(b) >> 32)),
                                               CTRL, ROW, BANK, false);
    return __longlong_as_double((long long)((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}

// inclusive scan with operator OP over the 64 lanes; ID is OP's identity
template <typename T, typename OP>
__device__ __forceinline__ T wave_incl_scan_op(T x, T id, OP op) {
    if constexpr (sizeof(T) == 8) {
        x = op(x, dpp_d<kDppRowShr1>(id, x));
        x = op(x, dpp_d<kDppRowShr2>(id, x));
        x = op(x, dpp_d<kDppRowShr4>(id, x));
        x = op(x, dpp_d<kDppRowShr8>(id, x));
        x = op(x, dpp_d<kDppRowBcast15, 0xa>(id, x));
        x = op(x, dpp_d<kDppRowBcast31, 0xc>(id, x));
    } else if constexpr (std::is_same<T, float>::value) {
        x = op(x, dpp_f<kDppRowShr1>(id, x));
        x = op(x, dpp_f<kDppRowShr2>(id, x));
        x = op(x, dpp_f<kDppRowShr4>(id, x));
        x = op(x, dpp_f<kDppRowShr8>(id, x));
        x = op(x, dpp_f<kDppRowBcast15, 0xa>(id, x));
        x = op(x, dpp_f<kDppRowBcast31, 0xc>(id, x));
    } else {
        x = op(x, dpp_i<kDppRowShr1>(id, x));
        x = op(x, dpp_i<kDppRowShr2>(id, x));
        x = op(x, dpp_i<kDppRowShr4>(id, x));
        x = op(x, dpp_i<kDppRowShr8>(id, x));
        x = op(x, dpp_i<kDppRowBcast15, 0xa>(id, x));
        x = op(x, dpp_i<kDppRowBcast31, 0xc>(id, x));
    }
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    return rlane(wave_incl_scan_op(v, T(0), [](T a, T b) { return a + b; }), 63);
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    return rlane(wave_incl_scan_op(v, T(-INFINITY), [](T a, T b) { return a > b ? a : b; }), 63);
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    return rlane(wave_incl_scan_op(v, T(INFINITY), [](T a, T b) { return a < b ? a : b; }), 63);
}
// exclusive prefix sum across the wave
__device__ __forceinline__ int wave_excl_scan(int v) {
    return wave_incl_scan_op(v, 0, [](int a, int b) { return a + b; }) - v;
}

// Value of v in lane (lane ^ J), J a power of two below 64, on the VALU
// rather than the LDS crossbar: DPP quad permutes (J = 1, 2), row rotations
// by 4/12 and 8 within 16-lane rows (J = 4, 8), and the gfx950 cross-row
// swaps v_permlane16_swap / v_permlane32_swap (J = 16, 32).
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "xor_lane");
    if constexpr (J == 1) {
        return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xf, 0xf, false));  // quad_perm 1,0,3,2
    } else if constexpr (J == 2) {
        return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x4E, 0xf, 0xf, false));  // quad_perm 2,3,0,1
    } else if constexpr (J == 4) {
        const uint32_t up = uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x12C, 0xf, 0xf, false));  // row_ror 12
        const uint32_t dn = uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x124, 0xf, 0xf, false));  // row_ror 4
        return (lane_id() & 4) ? dn : up;
    } else if constexpr (J == 8) {
        return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x128, 0xf, 0xf, false));  // row_ror 8
    } else if constexpr (J == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane_id() & 16) ? r[0] : r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane_id() & 32) ? r[0] : r[1];
    }
}
// same, J known only after loop unrolling
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int j) {
    switch (j) {
        case 1: return xor_lane<1>(v);
        case 2: return xor_lane<2>(v);
        case 4: return xor_lane<4>(v);
        case 8: return xor_lane<8>(v);
        case 16: return xor_lane<16>(v);
        default: return xor_lane<32>(v);
    }
}
__device__ __forceinline__ uint64_t xor_lane(uint64_t v, int j) {
    return (uint64_t(xor_lane(uint32_t(v >> 32), j)) << 32) | xor_lane(uint32_t(v), j);
}
// Value of v in lane - 1 (DPP wave_shr:1); lane 0 gets `first`.
__device__ __forceinline__ uint32_t shr_lane1(uint32_t v, uint32_t first) {
    return uint32_t(__builtin_amdgcn_update_dpp(int(first), int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ uint64_t shr_lane1(uint64_t v, uint64_t first) {
    return (uint64_t(shr_lane1(uint32_t(v >> 32), uint32_t(first >> 32))) << 32) | shr_lane1(uint32_t(v), uint32_t(first));
}

// index of the lowest set bit (m != 0)
__device__ __forceinline__ int lowest_bit(uint64_t m) { return __ffsll((long long)m) - 1; }

}  // namespace rpd
