// voroalgo_flow.cpp -- test driver of the header-only rpd::VoroAlgo: the flow
// of the reference's computeVoroDiagram (test_voronoi.cpp:30-62) with
// Options::initVoroAlgoParameters (Options.cpp:253-280), compiled against
// include/rpd_voroalgo.hpp only.
//
//   voroalgo_flow seeds.bin grid.txt|- K P V T keepDistance keepVolume H HC counters out.bin
//
// seeds.bin: int64 n, then 4n doubles.  out.bin: int64 count, then
// 4*count doubles (getBarycenters), count int32 (getPermutationToUser),
// int64 m, m int32 (getPermutationToInputPoints), int64 h, h doubles
// (getHessians), int64 hi, hi int32 (getHessianIds).  The statistics are
// printed as GTime["stat"] with VoroAlgoStat::pythonPrint.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "rpd_voroalgo.hpp"

template <typename T>
static void put(std::ofstream &f, const std::vector<T> &v, bool with_count) {
    if (with_count) {
        const int64_t m = int64_t(v.size());
        f.write(reinterpret_cast<const char *>(&m), sizeof(m));
    }
    f.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

int main(int argc, char **argv) {
    if (argc != 13) {
        std::cerr << "usage: " << argv[0]
                  << " seeds.bin grid.txt|- K P V T keepDistance keepVolume H HC counters out.bin\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int64_t n = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(n));
    std::vector<double> seeds(size_t(n) * 4);
    in.read(reinterpret_cast<char *>(seeds.data()), std::streamsize(seeds.size() * sizeof(double)));
    if (!in || n < 2) {
        std::cerr << "bad seeds file\n";
        return 2;
    }
    rpd::Grid grid;
    const bool has_grid = std::string(argv[2]) != "-";
    if (has_grid && !grid.load(argv[2])) {
        std::cerr << "bad grid file\n";
        return 2;
    }
    std::vector<double> bary, hess;
    std::vector<int> perm, perm_pts, hids;
    std::vector<rpd::VoroAlgoStat> stats;
    try {
        rpd::VoroAlgo algo(0, false);
        algo.m_K = std::atoi(argv[3]);
        algo.m_P = std::atoi(argv[4]);
        algo.m_V = std::atoi(argv[5]);
        algo.m_T = std::atoi(argv[6]);
        algo.m_keepDistance = std::atoi(argv[7]);
        algo.m_keepVolume = std::atof(argv[8]);
        algo.m_H = std::atoi(argv[9]);
        algo.m_HC = std::atoi(argv[10]);
        const bool counters = std::atoi(argv[11]) != 0;
        algo.m_enableDCountersHistoK = algo.m_enableDCountersHistoP = algo.m_enableDCountersHistoT =
            algo.m_enableDCountersHistoV = counters;
        algo.m_enableDCountersMax = algo.m_enableDCountersStatus = counters;
        algo.setInputPoints(seeds);
        if (has_grid) algo.setGrid(grid);
        const bool ok = algo.launch(
            [&](rpd::VoroAlgo const &a) {
                bary = a.getBarycenters();
                perm = a.getPermutationToUser();
                perm_pts = a.getPermutationToInputPoints();
                if (a.m_H > 0) {
                    hess = a.getHessians();
                    hids = a.getHessianIds();
                }
                return int64_t(bary.size()) == 4 * int64_t(a.getNumOutputPoints());
            },
            stats);
        if (!ok) {
            std::cerr << "launch failed: " << algo.lastError() << "\n";
            return 1;
        }
    } catch (std::exception const &e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    std::cout << "GTime[\"stat\"] = [";
    for (size_t i = 0; i < stats.size(); ++i) {
        stats[i].pythonPrint(std::cout);
        if (i + 1 < stats.size()) std::cout << ",\n";
    }
    std::cout << "]\n";
    std::ofstream out(argv[12], std::ios::binary | std::ios::trunc);
    const int64_t count = int64_t(perm.size());
    out.write(reinterpret_cast<const char *>(&count), sizeof(count));
    put(out, bary, false);
    put(out, perm, false);
    put(out, perm_pts, true);
    put(out, hess, true);
    put(out, hids, true);
    return out ? 0 : 1;
}
