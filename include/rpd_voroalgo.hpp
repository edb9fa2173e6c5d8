// rpd_voroalgo.hpp -- header-only C++ class with the surface of the
// reference's VoroAlgo (voronoi.h:105-256), implemented over the C-ABI of
// rpd.h.  A caller of the reference keeps its code:
//
//     rpd::VoroAlgo algo(device);              // VoroAlgo(ctx, debug)
//     algo.m_K = 64; algo.m_P = 40; ...        // public caps
//     algo.setInputPoints(xyzw);               // N x (x, y, z, w)
//     algo.setGrid(grid);                      // resident domain
//     algo.launch([&](rpd::VoroAlgo const &a) { bary = a.getBarycenters(); return true; }, stats);
//
// OpenCL-typed members of the reference map as follows:
//   setInputPoints(cl_mem, int)  -> setInputPointsDevice(const double *, int64_t)
//   getBarycentersCL()           -> getBarycentersDevice() (every seed), or
//                                   rpd_compact_used_device (the compacted output)
//   getContext/getCLQueue        -> no counterpart (one HIP stream per object)
// Output layout as the reference's: with a grid and m_keepDistance >= 0 the
// used cells and m_keepDistance rings of face neighbours (compressResult,
// voronoi.cpp:513-691, 1381-1394), in ascending user index (the reference's
// order comes from atomics), else every seed in the user's order.
// Outputs stay valid until the next launch (the reference frees them after
// saveData returns, voronoi.cpp:1398-1401).
#ifndef RPD_VOROALGO_HPP
#define RPD_VOROALGO_HPP

#include <algorithm>
#include <cstdint>
#include <functional>
#include <map>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rpd.h"

namespace rpd {

// VoroAlgoStat (voronoi.h:60-103): one entry per pass.
struct VoroAlgoStat {
    explicit VoroAlgoStat(std::string name) : m_name(std::move(name)) {}
    // one statistic as a python dict, the form test_voronoi prints in its
    // GTime["stat"] list (voronoi.h:69-86)
    void pythonPrint(std::ostream &output, std::string const &prefix = "") const {
        output << "{\"name\" : \"" << prefix << m_name << "\", ";
        const char *wh[] = {"knn", "CC", "time"};
        for (int i = 0; i < 3; ++i) output << "\"" << wh[i] << "\":" << m_times[i] << ", ";
        for (auto const &it : m_idToHistoMap) {
            output << "\"counter" << it.first << "\":[";
            for (size_t i = 0; i < it.second.size(); ++i) output << it.second[i] << (i + 1 != it.second.size() ? "," : "");
            output << "], ";
        }
        if (m_numDone) output << "\"nbdone\":" << m_numDone << ",";
        output << "\"nbcells\":" << m_num << "}";
    }
    std::string m_name;
    int m_num = 0;      // cells entering the pass
    int m_numDone = 0;  // of them, cells the pass finished
    double m_times[3] = {0, 0, 0};  // knn, CC, total (ms)
    // debug counters (voronoi.h:93-102), when enabled: 0 {maxK, maxP, maxV,
    // maxT, maxH}, 1 K histogram, 2 P histogram, 3 V histogram, 4 T
    // histogram; 10 the status histogram (Status.h numbering)
    std::map<int, std::vector<int>> m_idToHistoMap;
};

// Grid (Grid.h:14-70) as an owned host object.
class Grid {
public:
    Grid() = default;
    Grid(const Grid &) = delete;
    Grid &operator=(const Grid &) = delete;
    ~Grid() { rpd_grid_free(m_g); }
    bool load(const char *path) {
        rpd_grid_free(m_g);
        m_g = nullptr;
        return rpd_grid_load(path, &m_g) == RPD_OK;
    }
    bool save(const char *path) const { return m_g && rpd_grid_save(m_g, path) == RPD_OK; }
    bool fromVoxels(int g, const std::vector<uint8_t> &mask) {
        rpd_grid_free(m_g);
        m_g = nullptr;
        return int64_t(mask.size()) == int64_t(g) * g * g && rpd_grid_from_voxels(g, mask.data(), &m_g) == RPD_OK;
    }
    bool valid() const { return m_g && rpd_grid_valid(&desc()) == 1; }
    double getVolume() const { return m_g ? rpd_grid_volume(&desc()) : 0.0; }
    rpd_grid_desc const &desc() const {
        if (m_g) rpd_grid_view(m_g, &m_desc);
        return m_desc;
    }
    bool empty() const { return m_g == nullptr; }

private:
    rpd_grid *m_g = nullptr;
    mutable rpd_grid_desc m_desc{};
};

class VoroAlgo {
public:
    // VoroAlgo(ctx, debug) with the constructor defaults of voronoi.cpp:1041-1073
    explicit VoroAlgo(int device = 0, bool debug = false) : m_debug(debug) {
        if (rpd_create(device, &m_ctx) != RPD_OK) throw std::runtime_error("rpd_create: no usable HIP device");
        rpd_params p;
        rpd_default_params(&p);
        m_K = p.K;
        m_P = p.P;
        m_V = p.V;
        m_T = p.T;
        m_maxExtraSteps = p.max_extra_steps;
    }
    VoroAlgo(const VoroAlgo &) = delete;
    VoroAlgo &operator=(const VoroAlgo &) = delete;
    ~VoroAlgo() {
        if (m_dom) rpd_domain_destroy(m_dom);
        rpd_destroy(m_ctx);
    }

    // setGrid (voronoi.cpp:1079-1082): the grid is copied to HBM once
    void setGrid(Grid const &grid) {
        if (m_dom) rpd_domain_destroy(m_dom);
        m_dom = nullptr;
        check(rpd_domain_create(m_ctx, &grid.desc(), &m_dom), "setGrid");
    }
    // setInputPoints (voronoi.cpp:1098-1107): N x (x, y, z, w), copied
    void setInputPoints(std::vector<double> const &xyzw) {
        check(rpd_set_seeds(m_ctx, xyzw.data(), int64_t(xyzw.size() / 4)), "setInputPoints");
        m_n = int64_t(xyzw.size() / 4);
        m_input = xyzw;
    }
    void setInputPoints(std::vector<float> const &xyzw) {
        setInputPoints(std::vector<double>(xyzw.begin(), xyzw.end()));
    }
    void setInputPointsDevice(const double *d_xyzw, int64_t n) {
        check(rpd_set_seeds_device(m_ctx, d_xyzw, n), "setInputPoints");
        m_n = n;
        m_input.clear();
    }
    int getNumInputPoints() const { return int(m_n); }
    // the input points in the user's order (the reference returns them in its
    // kNN order; its getPermutationToInputPoints indexes that order, here
    // this one)
    std::vector<double> const &getInputPoints() const { return m_input; }

    // launch (voronoi.cpp:1244-1403): solve, compact the output as
    // compressResult does, then hand the results to saveData
    bool launch(std::function<bool(VoroAlgo const &)> saveData, std::vector<VoroAlgoStat> &stats) {
        if (m_n < 2) {
            m_error = "launch: input points are not initialised";
            return false;
        }
        rpd_params p;
        rpd_default_params(&p);
        p.K = m_K;
        p.P = m_P;
        p.V = m_V;
        p.T = m_T;
        p.precision = precision();
        p.max_extra_steps = m_maxExtraSteps;
        // voronoi.cpp:1276-1279: an unknown component count turns the Hessian off
        p.H = m_H > 0 && (m_HC == 1 || m_HC == 4) ? m_H : 0;
        p.HC = m_HC == 4 ? 4 : 1;
        const bool counters = m_enableDCountersMax || m_enableDCountersHistoK || m_enableDCountersHistoP ||
                              m_enableDCountersHistoV || m_enableDCountersHistoT;
        p.debug_counters = counters ? 1 : 0;
        rpd_stats st{};
        if (rpd_solve(m_ctx, m_dom, &p, &st) != RPD_OK) {
            m_error = rpd_last_error(m_ctx);
            return false;
        }
        const int64_t n = m_n;
        std::vector<double> full(size_t(n) * 4);
        if (rpd_read_moments(m_ctx, full.data()) != RPD_OK) return fail();
        // used cells and bordering rings (only with a grid and m_keepDistance
        // >= 0, voronoi.cpp:1291, 1381-1394); else every cell
        m_permToPoints.clear();
        if (m_dom && m_keepDistance >= 0) {
            int64_t cnt = 0;
            if (rpd_compact_used(m_ctx, m_keepVolume, m_keepDistance, &cnt, nullptr, nullptr) != RPD_OK) return fail();
            m_out.assign(size_t(cnt) * 4, 0.0);
            m_perm.assign(size_t(cnt), 0);
            if (rpd_compact_used(m_ctx, m_keepVolume, m_keepDistance, &cnt, m_out.data(), m_perm.data()) != RPD_OK)
                return fail();
            m_permToPoints = m_perm;  // into getInputPoints(), the user's order
        } else {
            m_out.swap(full);
            m_perm.resize(size_t(n));
            for (int64_t i = 0; i < n; ++i) m_perm[size_t(i)] = int(i);
        }
        const int64_t nout = int64_t(m_perm.size());
        m_hess.clear();
        m_hids.clear();
        if (p.H > 0) {
            const size_t H = size_t(p.H), HC = size_t(p.HC);
            std::vector<double> hv(size_t(n) * H * HC);
            std::vector<int32_t> hi(size_t(n) * H);
            if (rpd_read_hessian(m_ctx, hv.data(), hi.data()) != RPD_OK) return fail();
            if (m_permToPoints.empty()) {
                m_hess.swap(hv);
                m_hids.assign(hi.begin(), hi.end());
            } else {
                // compact_result's Hessian (voronoi.cpp:594-631): neighbours
                // renumbered to output positions, those not output or without
                // volume dropped; a cell without volume has no entries
                const std::vector<double> &mom = full;  // every seed's moments
                std::vector<int> pos(size_t(n), -1);
                for (int64_t o = 0; o < nout; ++o) pos[size_t(m_perm[size_t(o)])] = int(o);
                m_hess.assign(size_t(nout) * H * HC, 0.0);
                m_hids.assign(size_t(nout) * H, -1);
                for (int64_t o = 0; o < nout; ++o) {
                    const size_t u = size_t(m_perm[size_t(o)]);
                    if (mom[4 * u + 3] < m_keepVolume) continue;
                    size_t w = 0;
                    for (size_t h = 0; h < H; ++h) {
                        const int j = hi[u * H + h];
                        if (j < 0) break;
                        if (pos[size_t(j)] < 0 || mom[4 * size_t(j) + 3] < m_keepVolume) continue;
                        m_hids[size_t(o) * H + w] = pos[size_t(j)];
                        for (size_t c = 0; c < HC; ++c) m_hess[(size_t(o) * H + w) * HC + c] = hv[(u * H + h) * HC + c];
                        ++w;
                    }
                }
            }
        }
        // one statistic per pass (voronoi.cpp:1293-1376)
        const int np = st.passes < RPD_MAX_PASSES ? st.passes : RPD_MAX_PASSES;
        std::vector<int32_t> dc(counters ? size_t(RPD_DCOUNT_INTS) : 0);
        for (int i = 0; i < np; ++i) {
            VoroAlgoStat s(i == 0 ? "first" : "extra" + std::to_string(i));
            s.m_num = st.pass_cells[i];
            s.m_numDone = st.pass_cells[i] - (i + 1 < np ? st.pass_cells[i + 1] : 0);
            s.m_times[0] = st.pass_ms_knn[i];
            s.m_times[1] = st.pass_ms_clip[i];
            s.m_times[2] = st.pass_ms_knn[i] + st.pass_ms_clip[i];
            if (counters && rpd_read_debug_counters(m_ctx, i, dc.data()) == RPD_OK) {
                // the reference's sizes (initCounters, voronoi.cpp:1233-1242)
                const int K = st.pass_caps[i][0], P = st.pass_caps[i][1], V = st.pass_caps[i][2],
                          T = st.pass_caps[i][3];
                auto take = [&](int off, int len) {
                    return std::vector<int>(dc.begin() + off, dc.begin() + off + len);
                };
                if (m_enableDCountersMax) s.m_idToHistoMap[0] = take(RPD_DCOUNT_MAX, p.H > 0 ? 5 : (m_dom ? 4 : 3));
                if (m_enableDCountersHistoK) s.m_idToHistoMap[1] = take(RPD_DCOUNT_HIST_K, 1 + std::min(K, 256));
                if (m_enableDCountersHistoP) s.m_idToHistoMap[2] = take(RPD_DCOUNT_HIST_P, std::min(P, 256));
                if (m_enableDCountersHistoV) s.m_idToHistoMap[3] = take(RPD_DCOUNT_HIST_V, std::min(V + 1, 257));
                if (m_enableDCountersHistoT && m_dom)
                    s.m_idToHistoMap[4] = take(RPD_DCOUNT_HIST_T, std::min(T, RPD_DCOUNT_UNION_MAX) + 1);
            }
            stats.push_back(s);
        }
        if (m_enableDCountersStatus && !stats.empty())
            stats.back().m_idToHistoMap[10] = std::vector<int>(st.status_hist, st.status_hist + RPD_STATUS_NUM);
        m_last = st;
        return saveData ? saveData(*this) : true;
    }

    // number of output cells: the used cells and their rings after the
    // compaction, else every seed (voronoi.h:160-161)
    int getNumOutputPoints() const { return int(m_perm.size()); }
    // per output cell: (m x, m y, m z, m); border seeds are copied, a cell of
    // the last ring without volume has w = NO_CELL_LIMITS (voronoi.h:47-53)
    std::vector<double> const &getBarycenters() const { return m_out; }
    // the moments of every seed in the user's order, left in HBM
    // (getBarycentersCL before any compaction)
    const double *getBarycentersDevice() const {
        const double *p = nullptr;
        rpd_moments_device(m_ctx, &p);
        return p;
    }
    // output position -> the user's seed index (voronoi.h:188-190)
    std::vector<int> const &getPermutationToUser() const { return m_perm; }
    // output position -> index into getInputPoints(); empty when it is the
    // identity (voronoi.h:183-187)
    std::vector<int> getPermutationToInputPoints() const { return m_permToPoints; }
    // m_HC * m_H values and m_H neighbour ids (output positions, -1 after the
    // last) per output cell (getHessians / getHessianIds,
    // voronoi.cpp:1179-1232); empty with m_H = 0
    std::vector<double> const &getHessians() const { return m_hess; }
    std::vector<int> const &getHessianIds() const { return m_hids; }
    // final status per seed, in the user's order
    std::vector<int32_t> getStatus() const {
        std::vector<int32_t> s(static_cast<size_t>(m_n));
        rpd_read_status(m_ctx, s.data());
        return s;
    }
    rpd_stats const &lastStats() const { return m_last; }
    std::string const &lastError() const { return m_error; }
    rpd_ctx *handle() const { return m_ctx; }
    // the precision form a launch uses: m_precision when set (>= 0), else
    // the reference's half-space algorithm of the first step (m_vhsAlgo[0]:
    // 0 FP32 -> RPD_PREC_FP32, 1 FP32 with check -> RPD_PREC_MIXED, whose
    // flagged cells are recomputed in FP64 as m_vhsAlgo[1] = 2 asks, 2 FP64
    // -> RPD_PREC_FP64)
    int precision() const {
        if (m_precision >= 0) return m_precision;
        return m_vhsAlgo[0] == 0 ? RPD_PREC_FP32 : (m_vhsAlgo[0] == 1 ? RPD_PREC_MIXED : RPD_PREC_FP64);
    }

    // public members of voronoi.h:199-256 with their defaults
    int m_K, m_P, m_T, m_V;
    double m_PGlobalPercent = 0;  // accepted; the planes of a cell always fit the LDS
    int m_H = 0, m_HC = 1;
    int m_maxKComputed = 100000;  // accepted; the kNN runs as one launch over every query
    int m_vhsAlgo[2] = {1, 2};
    int m_keepDistance = 1;
    double m_keepVolume = 1;
    bool m_enableCLProfiling = false;  // accepted; the per-pass times are always measured
    bool m_enableDCountersHistoK = false, m_enableDCountersHistoP = false, m_enableDCountersHistoT = false,
         m_enableDCountersHistoV = false, m_enableDCountersMax = false, m_enableDCountersStatus = false;
    // this build's additions: the precision form (-1: from m_vhsAlgo) and the
    // number of re-queue passes
    int m_precision = -1;
    int m_maxExtraSteps;
    bool m_debug;

private:
    static void check(int rc, const char *what) {
        if (rc != RPD_OK) throw std::runtime_error(std::string(what) + " failed");
    }
    bool fail() {
        m_error = rpd_last_error(m_ctx);
        return false;
    }
    rpd_ctx *m_ctx = nullptr;
    rpd_domain *m_dom = nullptr;
    int64_t m_n = 0;
    std::vector<double> m_input, m_out, m_hess;
    std::vector<int> m_perm, m_permToPoints, m_hids;
    rpd_stats m_last{};
    std::string m_error;
};

}  // namespace rpd

#endif
