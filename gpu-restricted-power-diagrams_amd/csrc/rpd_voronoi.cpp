// rpd_voronoi.cpp -- command-line driver with the options and outputs of the
// reference's test_voronoi (test_voronoi.cpp:63-223, Options.cpp:20-283),
// running the restricted power diagram on the GPU through rpd::VoroAlgo.
//
//   rpd_voronoi [options] points.xyz
//
// Seeds: ascii "n" then "x y z" (or "x y z w" with --points4D) per line
// (Utils.cpp:36-79), or the binary format of --loadBinaryPoints (size_t n,
// then 4n reals; float or double, told apart by the file size).  Outputs
// (--output=...): xyz -> gridGPU.xyz, mg -> gridGPU_mg.xyzw, volume ->
// volume sums on stderr.  Additional option: --precision=fp64|mixed|fp32.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "rpd_voroalgo.hpp"

namespace {

struct Options {
    int K = 20, P = 25, V = 30, T = 100;  // Options.cpp:20-28
    double keepVolume = 0.001;            // USE_DOUBLE default (Options.cpp:36-40)
    int precision = -1;                   // -1: from --cVoroIsInHalfSpaceAlgo
    int vhs = 1;                          // half-space algorithm of the first step (Options.cpp:55, float build)
    int keepDistLast = 0;                 // --keepDist2 (Options.cpp:42): the final keep distance
    int maxK = 100000;                    // --maxK (accepted)
    bool counterMax = false, counterHisto = false, counterStatus = false;
    int H = 50, HC = 1;                   // Options.cpp:190-191
    bool hessian = false;                 // --hessian (Options.cpp:77, 269)
    bool help = false, points4D = false, loadBinary = false, checkGrid = false, invariants = false, debug = false;
    std::string grid, saveBinary;
    std::vector<std::string> outputs;
};

void usage(const char *prog) {
    std::cerr << "Usage: " << prog << " [options] points.xyz\n\n"
              << "\tOPTIONS:\n"
              << "\t--help|-h: show this help\n"
              << "\t--K=k: set the first K value[default 20]\n"
              << "\t--P=p: set the first P value[default 25]\n"
              << "\t--T=t: set the max number of grid's triangles[default 100]\n"
              << "\t--V=v: set the first V value[default 30]\n"
              << "\t--points4D: assume that the seeds' file is a 4D ascii file\n"
              << "\t--loadBinaryPoints: assume that the seeds' file is a 4D binary file\n"
              << "\t--saveBinaryPoints=file: save the seeds' in a 4D binary file (double)\n"
              << "\t--GPU: do the computation on the GPU (the only device)\n"
              << "\t--keepVolume=v: volume below which a cell is reported as unused[default 0.001]\n"
              << "\t--grid=FILE: set the grid input file to FILE\n"
              << "\t--checkGrid: checks if the grid if valid\n"
              << "\t--H=h: set the max number of Hessian values in a row[default 50]\n"
              << "\t--HC=C: set the number of Hessian components(1 or 4)[default 1]\n"
              << "\t--hessian: compute the hessian value\n"
              << "\t--precision=fp64|mixed|fp32: clip precision form[default: from --cVoroIsInHalfSpaceAlgo]\n"
              << "\t--cVoroIsInHalfSpaceAlgo=algo: 0 FP32, 1 FP32 with check (mixed), 2 FP64[default 1]\n"
              << "\t--keepDist=d, --keepDist2=d: rings of bordering cells kept in the output[default 0]\n"
              << "\t--maxK=k: accepted (the kNN is one launch)\n"
              << "\t--cVoroDebugCounterMax, --cVoroDebugCounterHisto, --cVoroDebugCounterStatus: per-pass\n"
              << "\t    counters of K, P, V, T (maxima, histograms) and statuses in GTime[\"stat\"]\n"
              << "\t--output=xyz|mg|geogram|volume: outputs (gridGPU.xyz, gridGPU_mg.xyzw,\n"
              << "\t    gridGPU.geogram_ascii, volumes on stderr)\n"
              << "\t--printInvariants: displays the final volume and barycenter\n"
              << "\t--debug: print per-pass statistics\n";
}

bool parse(int argc, char **argv, Options &o, std::string &file) {
    for (int i = 1; i < argc; ++i) {
        std::string a(argv[i]), v;
        const size_t eq = a.find('=');
        if (eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
        }
        auto val = [&]() -> std::string {
            if (!v.empty()) return v;
            if (i + 1 < argc) return std::string(argv[++i]);
            return std::string();
        };
        if (a == "-h" || a == "--help") o.help = true;
        else if (a == "--points4D") o.points4D = true;
        else if (a == "--loadBinaryPoints") o.loadBinary = true;
        else if (a == "--checkGrid") o.checkGrid = true;
        else if (a == "--printInvariants") o.invariants = true;
        else if (a == "--debug") o.debug = true;
        else if (a == "--GPU") {
        } else if (a == "--CPU") {
            std::cerr << argv[0] << ": there is no CPU path; the computation runs on the GPU\n";
        } else if (a == "--K") o.K = std::atoi(val().c_str());
        else if (a == "--P") o.P = std::atoi(val().c_str());
        else if (a == "--V") o.V = std::atoi(val().c_str());
        else if (a == "--T") o.T = std::atoi(val().c_str());
        else if (a == "--H") o.H = std::atoi(val().c_str());
        else if (a == "--HC") o.HC = std::atoi(val().c_str());
        else if (a == "--hessian") o.hessian = true;
        else if (a == "--keepVolume") {
            o.keepVolume = std::atof(val().c_str());
            if (o.keepVolume <= 0) o.keepVolume = 0.001;
        } else if (a == "--grid") o.grid = val();
        else if (a == "--saveBinaryPoints") o.saveBinary = val();
        else if (a == "--output") o.outputs.push_back(val());
        else if (a == "--cVoroIsInHalfSpaceAlgo") o.vhs = std::atoi(val().c_str());
        else if (a == "--keepDist" || a == "--keepDist2") o.keepDistLast = std::atoi(val().c_str());
        else if (a == "--maxK") o.maxK = std::atoi(val().c_str());
        else if (a == "--cVoroDebugCounterMax") o.counterMax = true;
        else if (a == "--cVoroDebugCounterHisto") o.counterHisto = true;
        else if (a == "--cVoroDebugCounterStatus") o.counterStatus = true;
        else if (a == "--enableProfiling") {
        } else if (a == "--precision") {
            const std::string p = val();
            o.precision = p == "fp64" ? RPD_PREC_FP64 : (p == "fp32" ? RPD_PREC_FP32 : RPD_PREC_MIXED);
        } else if (a.size() > 1 && a[0] == '-') {
            std::cerr << argv[0] << ": unknown option " << a << "\n";
            return false;
        } else if (file.empty()) file = argv[i];
        else return false;
    }
    return !file.empty();
}

// Utils::loadSeedsFile (Utils.cpp:36-79) without normalisation
bool load_ascii(const std::string &path, bool four, std::vector<double> &xyzw) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    long long n = -1;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream is(line);
        if (n < 0) {
            is >> n;
            xyzw.reserve(size_t(std::max(0LL, n)) * 4);
            continue;
        }
        double x = 0, y = 0, z = 0, w = 0;
        is >> x >> y >> z;
        if (four) is >> w;
        xyzw.insert(xyzw.end(), {x, y, z, w});
    }
    return n >= 0 && (long long)(xyzw.size() / 4) == n;
}

// Utils::readBinarySeedsFile (Utils.cpp:81-98): size_t n, then 4n reals
bool load_binary(const std::string &path, std::vector<double> &xyzw) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return false;
    const long long size = (long long)in.tellg();
    in.seekg(0);
    uint64_t n = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(n));
    if (!in || n == 0) return false;
    const long long body = size - 8;
    xyzw.resize(size_t(n) * 4);
    if (body == (long long)(n * 32)) {
        in.read(reinterpret_cast<char *>(xyzw.data()), std::streamsize(n * 32));
    } else if (body == (long long)(n * 16)) {
        std::vector<float> f(size_t(n) * 4);
        in.read(reinterpret_cast<char *>(f.data()), std::streamsize(n * 16));
        std::copy(f.begin(), f.end(), xyzw.begin());
    } else {
        return false;
    }
    return bool(in);
}

bool save_binary(const std::string &path, const std::vector<double> &xyzw) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return false;
    const uint64_t n = xyzw.size() / 4;
    out.write(reinterpret_cast<const char *>(&n), sizeof(n));
    out.write(reinterpret_cast<const char *>(xyzw.data()), std::streamsize(n * 32));
    return bool(out);
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    std::string file;
    if (!parse(argc, argv, o, file) || o.help) {
        usage(argv[0]);
        return 1;
    }
    std::vector<double> pts;
    if (!(o.loadBinary ? load_binary(file, pts) : load_ascii(file, o.points4D, pts))) {
        std::cerr << argv[0] << ": could not load file " << file << std::endl;
        return 1;
    }
    if (!o.saveBinary.empty()) save_binary(o.saveBinary, pts);

    rpd::Grid grid;
    const bool hasGrid = !o.grid.empty() && grid.load(o.grid.c_str());
    if (!o.grid.empty() && !hasGrid) std::cerr << argv[0] << ": could not load grid " << o.grid << std::endl;
    if (hasGrid && o.checkGrid && !grid.valid()) std::cerr << argv[0] << ": the grid is invalid." << std::endl;

    std::vector<double> bary;
    std::vector<int> permutation;
    std::vector<rpd::VoroAlgoStat> stats;
    try {
        // computeVoroDiagram (test_voronoi.cpp:30-62) with
        // Options::initVoroAlgoParameters (Options.cpp:253-280)
        rpd::VoroAlgo algo(0, o.debug);
        algo.m_K = o.K;
        algo.m_maxKComputed = o.maxK;
        algo.m_vhsAlgo[0] = o.vhs;
        algo.m_vhsAlgo[1] = 2;
        algo.m_precision = o.precision;
        algo.m_P = o.P;
        algo.m_V = o.V;
        algo.m_T = o.T;
        algo.m_H = o.hessian ? o.H : 0;
        algo.m_HC = o.HC;
        algo.m_keepDistance = o.keepDistLast;  // the final keep distance
        algo.m_keepVolume = o.keepVolume;
        algo.m_enableDCountersHistoK = algo.m_enableDCountersHistoP = algo.m_enableDCountersHistoT =
            algo.m_enableDCountersHistoV = o.counterHisto;
        algo.m_enableDCountersMax = o.counterMax;
        algo.m_enableDCountersStatus = o.counterStatus;
        algo.setInputPoints(pts);
        if (hasGrid) algo.setGrid(grid);
        if (!algo.launch([&](rpd::VoroAlgo const &a) {
                bary = a.getBarycenters();
                permutation = a.getPermutationToUser();
                return true;
            }, stats)) {
            std::cerr << argv[0] << ": solve failed: " << algo.lastError() << std::endl;
            return 1;
        }
    } catch (std::exception const &e) {
        std::cerr << argv[0] << ": " << e.what() << std::endl;
        return 1;
    }

    // GTime["stat"] (test_voronoi.cpp:52-60, VoroAlgoStat::pythonPrint)
    std::cout << "GTime[\"stat\"] = [";
    for (size_t i = 0; i < stats.size(); ++i) {
        stats[i].pythonPrint(std::cout);
        if (i + 1 < stats.size()) std::cout << ",\n";
    }
    std::cout << "]\n";

    const size_t n = bary.size() / 4;
    auto volume_report = [&]() {
        double vol = 0, minv = 1e20, maxv = -1e20;
        for (size_t i = 0; i < n; ++i) {
            const double m = bary[4 * i + 3];
            if (m <= o.keepVolume) continue;
            vol += m;
            minv = std::min(minv, m);
            maxv = std::max(maxv, m);
        }
        std::cerr << "Sum of volumes: " << std::setprecision(20) << vol << " minvol: " << minv << " maxvol: " << maxv
                  << std::endl;
        return vol;
    };
    if (o.invariants) {
        const double vol = volume_report();
        if (vol > 0) {
            double b[3] = {0, 0, 0};
            for (size_t i = 0; i < n; ++i)
                if (bary[4 * i + 3] > RPD_NO_CELL_LIMITS)
                    for (int j = 0; j < 3; ++j) b[j] += bary[4 * i + j];
            std::cerr << "Cells barycenter=" << std::setprecision(20) << b[0] / vol << "x" << b[1] / vol << "x"
                      << b[2] / vol << std::endl;
        }
    }
    for (auto const &s : o.outputs) {
        if (s == "xyz") {
            // Utils::dropXYZFile with pts4D (Utils.cpp:115-146)
            std::ofstream f("gridGPU.xyz");
            size_t k = 0;
            auto keep = [&](size_t i) { return bary[4 * i + 3] > RPD_NO_CELL_LIMITS && std::fabs(bary[4 * i + 3]) >= o.keepVolume; };
            for (size_t i = 0; i < n; ++i) k += keep(i);
            f << k << "\n";
            for (size_t i = 0; i < n; ++i)
                if (keep(i))
                    f << bary[4 * i] / bary[4 * i + 3] << "  " << bary[4 * i + 1] / bary[4 * i + 3] << "  "
                      << bary[4 * i + 2] / bary[4 * i + 3] << "\n";
        } else if (s == "geogram") {
            // Utils::dropXYZGeogram with every cell and its w (Utils.cpp:148-170,
            // test_voronoi.cpp:162-167): ascii geogram vertices, one number a line
            std::ofstream f("gridGPU.geogram_ascii");
            f << std::setprecision(17);
            f << "[HEAD]\n\"GEOGRAM\"\n\"1.0\"\n[ATTS]\n\"GEO::Mesh::vertices\"\n" << n << "\n";
            f << "[ATTR]\n\"GEO::Mesh::vertices\"\n\"point\"\n\"double\"\n8\n3\n";
            for (size_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) f << bary[4 * i + d] / bary[4 * i + 3] << "\n";
            f << "[ATTR]\n\"GEO::Mesh::vertices\"\n\"w\"\n\"double\"\n8\n1\n";
            for (size_t i = 0; i < n; ++i) f << bary[4 * i + 3] << "\n";
        } else if (s == "mg") {
            // per original seed, through the output permutation (test_voronoi.cpp:186-211)
            std::ofstream f("gridGPU_mg.xyzw");
            const size_t np = pts.size() / 4;
            std::vector<long long> pos(np, -1);
            for (size_t k = 0; k < permutation.size(); ++k) pos[size_t(permutation[k])] = (long long)k;
            f << np << "\n";
            for (size_t i = 0; i < np; ++i) {
                const long long q = pos[i];
                if (q < 0 || bary[4 * size_t(q) + 3] <= o.keepVolume) {
                    f << "0 0 0 0\n";
                    continue;
                }
                const double *b = &bary[4 * size_t(q)];
                f << std::setprecision(20) << b[0] << " " << b[1] << " " << b[2] << " " << b[3] << "\n";
            }
        } else if (s == "volume") {
            if (!o.invariants) volume_report();
        } else {
            std::cerr << argv[0] << ": ignored output : " << s << "\n";
        }
    }
    return 0;
}
