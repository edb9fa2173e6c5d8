// rpd_gen.cpp -- synthetic seed sets (SplitMix64, u = (x >> 11) * 2^-53),
// generated in FP64 in the unit frame and scaled x1000 into [0,1000]^3.

#include <cmath>
#include <cstdint>
#include <vector>

#include "rpd.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

}  // namespace

extern "C" int rpd_generate_seeds(int kind, int64_t n, uint64_t seed, double p0, double p1, double wmax,
                                  double *xyzw) {
    if (n <= 0 || !xyzw) return RPD_E_ARG;
    SplitMix64 rng(seed);
    switch (kind) {
        case RPD_SEEDS_UNIFORM:
            for (int64_t i = 0; i < n; ++i) {
                const double x = rng.uniform(), y = rng.uniform(), z = rng.uniform();
                xyzw[4 * i + 0] = 1000.0 * x;
                xyzw[4 * i + 1] = 1000.0 * y;
                xyzw[4 * i + 2] = 1000.0 * z;
                xyzw[4 * i + 3] = 0.0;
            }
            break;
        case RPD_SEEDS_LATTICE: {
            const int64_t m = int64_t(p0);
            if (m <= 0 || m * m * m < n) return RPD_E_ARG;
            int64_t c = 0;
            for (int64_t k = 0; k < m && c < n; ++k)
                for (int64_t j = 0; j < m && c < n; ++j)
                    for (int64_t i = 0; i < m && c < n; ++i, ++c) {
                        const double jx = p1 * (2.0 * rng.uniform() - 1.0);
                        const double jy = p1 * (2.0 * rng.uniform() - 1.0);
                        const double jz = p1 * (2.0 * rng.uniform() - 1.0);
                        xyzw[4 * c + 0] = 1000.0 * ((double(i) + 0.5 + jx) / double(m));
                        xyzw[4 * c + 1] = 1000.0 * ((double(j) + 0.5 + jy) / double(m));
                        xyzw[4 * c + 2] = 1000.0 * ((double(k) + 0.5 + jz) / double(m));
                        xyzw[4 * c + 3] = 0.0;
                    }
            break;
        }
        case RPD_SEEDS_BLUE_NOISE: {
            // dart throwing with a background grid of cell size r
            const double r = p0;
            if (!(r > 0) || r >= 0.5) return RPD_E_ARG;
            const int gd = std::max(1, int(std::floor(1.0 / r)));
            std::vector<std::vector<int64_t>> cells(size_t(gd) * gd * gd);
            auto cid = [&](double v) { return std::min(gd - 1, int(v * gd)); };
            const double r2 = r * r;
            int64_t accepted = 0;
            const int64_t max_darts = 2000 * n + 1000000;
            for (int64_t d = 0; d < max_darts && accepted < n; ++d) {
                const double x = rng.uniform(), y = rng.uniform(), z = rng.uniform();
                const int ci = cid(x), cj = cid(y), ck = cid(z);
                bool ok = true;
                for (int dk = -1; dk <= 1 && ok; ++dk)
                    for (int dj = -1; dj <= 1 && ok; ++dj)
                        for (int di = -1; di <= 1 && ok; ++di) {
                            const int a = ci + di, b = cj + dj, c = ck + dk;
                            if (a < 0 || b < 0 || c < 0 || a >= gd || b >= gd || c >= gd) continue;
                            for (int64_t q : cells[size_t(a) + size_t(b) * gd + size_t(c) * gd * gd]) {
                                const double dx = xyzw[4 * q] / 1000.0 - x, dy = xyzw[4 * q + 1] / 1000.0 - y,
                                             dz = xyzw[4 * q + 2] / 1000.0 - z;
                                if (dx * dx + dy * dy + dz * dz < r2) {
                                    ok = false;
                                    break;
                                }
                            }
                        }
                if (!ok) continue;
                xyzw[4 * accepted + 0] = 1000.0 * x;
                xyzw[4 * accepted + 1] = 1000.0 * y;
                xyzw[4 * accepted + 2] = 1000.0 * z;
                xyzw[4 * accepted + 3] = 0.0;
                cells[size_t(ci) + size_t(cj) * gd + size_t(ck) * gd * gd].push_back(accepted);
                ++accepted;
            }
            if (accepted < n) return RPD_E_LIMIT;
            break;
        }
        default:
            return RPD_E_ARG;
    }
    if (wmax > 0)
        for (int64_t i = 0; i < n; ++i) xyzw[4 * i + 3] = wmax * rng.uniform();
    return RPD_OK;
}
