// rpd_diff_result -- compares two per-cell result files (the mg output of
// rpd_voronoi: a count line, then "mx my mz m" per cell), with the report of
// the reference's utils/diffResult (main.cpp:63-119): the summed volume and
// barycentre of each file and their difference, then the largest and the
// average per-cell differences of volume, barycentre and barycentre x volume
// over the cells with a positive volume in both files.  Host code only.

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

// count line, then four numbers per line (utils/diffResult/Utils.cpp:11-38)
bool load(const std::string &path, std::vector<double> &v) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    long long n = -1;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream is(line);
        if (n < 0) {
            is >> n;
            continue;
        }
        double a[4] = {0, 0, 0, 0};
        is >> a[0] >> a[1] >> a[2] >> a[3];
        v.insert(v.end(), a, a + 4);
    }
    return n >= 0 && (long long)(v.size() / 4) == n;
}

struct Report {
    double sum[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    double max_b[3] = {0, 0, 0}, max_m[3] = {0, 0, 0}, max_v = 0;
    double ave_b[3] = {0, 0, 0}, ave_m[3] = {0, 0, 0}, ave_v = 0;
    size_t compared = 0;
};

Report compare(const std::vector<double> &a, const std::vector<double> &b) {
    Report r;
    const size_t n = a.size() / 4;
    for (size_t i = 0; i < n; ++i)
        for (int j = 0; j < 4; ++j) {
            r.sum[0][j] += a[4 * i + j];
            r.sum[1][j] += b[4 * i + j];
        }
    for (size_t i = 0; i < n; ++i) {
        const double va = a[4 * i + 3], vb = b[4 * i + 3];
        r.max_v = std::max(r.max_v, std::fabs(vb - va));
        if (va <= 0 || vb <= 0) continue;  // barycentres only where both cells have volume
        ++r.compared;
        r.ave_v += std::fabs(vb - va);
        for (int j = 0; j < 3; ++j) {
            const double db = std::fabs(b[4 * i + j] / vb - a[4 * i + j] / va);
            const double dm = std::fabs(b[4 * i + j] - a[4 * i + j]);
            r.ave_b[j] += db;
            r.ave_m[j] += dm;
            r.max_b[j] = std::max(r.max_b[j], db);
            r.max_m[j] = std::max(r.max_m[j], dm);
        }
    }
    return r;
}

void print(const Report &r) {
    auto bc = [&](int f, int j) { return r.sum[f][j] / r.sum[f][3]; };
    const double k = r.compared ? 1.0 / double(r.compared) : 0.0;
    std::cout.precision(20);
    std::cout << "------------------- intrinsic (sum) --------------------------\n";
    std::cout << "Volume : " << r.sum[0][3] << "\t" << r.sum[1][3] << "\t" << (r.sum[1][3] - r.sum[0][3]) << "\n";
    std::cout << "Barycenter : " << bc(0, 0) << "x" << bc(0, 1) << "x" << bc(0, 2) << "\t" << bc(1, 0) << "x"
              << bc(1, 1) << "x" << bc(1, 2) << "\t" << (bc(1, 0) - bc(0, 0)) << "x" << (bc(1, 1) - bc(0, 1)) << "x"
              << (bc(1, 2) - bc(0, 2)) << "\n";
    std::cout << "------------------- maximal difference by cells--------------------------\n";
    std::cout << "Volume : " << r.max_v << "\n";
    std::cout << "Barycenter : " << r.max_b[0] << "x" << r.max_b[1] << "x" << r.max_b[2] << "\n";
    std::cout << "Barycenter x volume : " << r.max_m[0] << "x" << r.max_m[1] << "x" << r.max_m[2] << "\n";
    std::cout << "------------------- average difference by cells--------------------------\n";
    std::cout << "Volume : " << r.ave_v * k << "\n";
    std::cout << "Barycenter : " << r.ave_b[0] * k << "x" << r.ave_b[1] * k << "x" << r.ave_b[2] * k << "\n";
    std::cout << "Barycenter x volume : " << r.ave_m[0] * k << "x" << r.ave_m[1] * k << "x" << r.ave_m[2] * k << "\n";
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<double> a, b;
    if (argc != 3 || !load(argv[1], a) || !load(argv[2], b)) {
        std::cerr << "Error:" << argv[0] << " res1.xyzw res2.xyzw\n";
        return 1;
    }
    if (a.size() != b.size() || a.size() % 4 != 0) {
        std::cerr << "Error:" << argv[0] << " files " << argv[1] << " and " << argv[2]
                  << " do not contain the same numbers of data\n";
        return 1;
    }
    print(compare(a, b));
    return 0;
}
