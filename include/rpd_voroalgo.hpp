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
//   getBarycentersCL()           -> getBarycentersDevice()
//   getContext/getCLQueue        -> no counterpart (one HIP stream per object)
// Outputs stay valid until the next launch (the reference frees them after
// saveData returns, voronoi.cpp:1398-1401).
#ifndef RPD_VOROALGO_HPP
#define RPD_VOROALGO_HPP

#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "rpd.h"

namespace rpd {

// VoroAlgoStat (voronoi.h:60-103): one entry per pass.
struct VoroAlgoStat {
    explicit VoroAlgoStat(std::string name) : m_name(std::move(name)) {}
    std::string m_name;
    int m_num = 0;
    int m_numDone = 0;
    double m_times[3] = {0, 0, 0};  // knn, CC, total (ms)
    std::map<int, std::vector<int>> m_idToHistoMap;  // 0: {maxK, maxP, maxV, maxT}, 10: status histogram
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
    explicit VoroAlgo(int device = 0, bool debug = false) : m_debug(debug) {
        if (rpd_create(device, &m_ctx) != RPD_OK) throw std::runtime_error("rpd_create: no usable HIP device");
        rpd_params p;
        rpd_default_params(&p);
        m_K = p.K;
        m_P = p.P;
        m_V = p.V;
        m_T = p.T;
        m_precision = p.precision;
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
    std::vector<double> const &getInputPoints() const { return m_input; }

    // launch (voronoi.cpp:1244-1403): solve, then hand the results to saveData
    bool launch(std::function<bool(VoroAlgo const &)> saveData, std::vector<VoroAlgoStat> &stats) {
        rpd_params p;
        rpd_default_params(&p);
        p.K = m_K;
        p.P = m_P;
        p.V = m_V;
        p.T = m_T;
        p.precision = m_precision;
        p.max_extra_steps = m_maxExtraSteps;
        p.H = m_H;
        p.HC = m_HC;
        rpd_stats st{};
        if (rpd_solve(m_ctx, m_dom, &p, &st) != RPD_OK) {
            m_error = rpd_last_error(m_ctx);
            return false;
        }
        m_out.assign(size_t(m_n) * 4, 0.0);
        if (rpd_read_moments(m_ctx, m_out.data()) != RPD_OK) return false;
        m_hess.clear();
        m_hids.clear();
        if (m_H > 0) {
            m_hess.assign(size_t(m_n) * m_H * m_HC, 0.0);
            m_hids.assign(size_t(m_n) * m_H, -1);
            if (rpd_read_hessian(m_ctx, m_hess.data(), m_hids.data()) != RPD_OK) return false;
        }
        m_perm.resize(size_t(m_n));
        for (int64_t i = 0; i < m_n; ++i) m_perm[size_t(i)] = int(i);  // per-original-seed layout
        const int np = st.passes < RPD_MAX_PASSES ? st.passes : RPD_MAX_PASSES;
        for (int i = 0; i < np; ++i) {
            VoroAlgoStat s(i == 0 ? "step0" : "extra" + std::to_string(i));
            s.m_num = st.pass_cells[i];
            s.m_times[0] = st.pass_ms_knn[i];
            s.m_times[1] = st.pass_ms_clip[i];
            s.m_times[2] = st.pass_ms_knn[i] + st.pass_ms_clip[i];
            s.m_idToHistoMap[0] = {st.pass_caps[i][0], st.pass_caps[i][1], st.pass_caps[i][2], st.pass_caps[i][3]};
            stats.push_back(s);
        }
        if (!stats.empty())
            stats.back().m_idToHistoMap[10] = std::vector<int>(st.status_hist, st.status_hist + RPD_STATUS_NUM);
        m_last = st;
        return saveData ? saveData(*this) : true;
    }

    int getNumOutputPoints() const { return int(m_n); }
    // per original seed: (mx, my, mz, m); border seeds are copied
    std::vector<double> const &getBarycenters() const { return m_out; }
    const double *getBarycentersDevice() const {
        const double *p = nullptr;
        rpd_moments_device(m_ctx, &p);
        return p;
    }
    std::vector<int> const &getPermutationToUser() const { return m_perm; }
    std::vector<int> getPermutationToInputPoints() const { return m_perm; }
    // m_HC * m_H values and m_H neighbour ids (-1 after the last) per seed
    // (getHessians / getHessianIds, voronoi.cpp:1179-1232); empty with m_H = 0
    std::vector<double> const &getHessians() const { return m_hess; }
    std::vector<int> const &getHessianIds() const { return m_hids; }
    std::vector<int32_t> getStatus() const {
        std::vector<int32_t> s(static_cast<size_t>(m_n));
        rpd_read_status(m_ctx, s.data());
        return s;
    }
    rpd_stats const &lastStats() const { return m_last; }
    std::string const &lastError() const { return m_error; }
    rpd_ctx *handle() const { return m_ctx; }

    // public caps (voronoi.h:199-230)
    int m_K, m_P, m_V, m_T;
    int m_precision;
    int m_maxExtraSteps;
    int m_H = 0, m_HC = 1;
    bool m_debug;

private:
    static void check(int rc, const char *what) {
        if (rc != RPD_OK) throw std::runtime_error(std::string(what) + " failed");
    }
    rpd_ctx *m_ctx = nullptr;
    rpd_domain *m_dom = nullptr;
    int64_t m_n = 0;
    std::vector<double> m_input, m_out, m_hess;
    std::vector<int> m_perm, m_hids;
    rpd_stats m_last{};
    std::string m_error;
};

}  // namespace rpd

#endif
