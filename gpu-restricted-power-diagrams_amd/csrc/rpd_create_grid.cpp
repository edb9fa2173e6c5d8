// rpd_create_grid -- command line of utils/createGrid (main.cpp:112-300) over
// the C-ABI: reads an ASCII .obj or .off triangle mesh, scales it into
// [4.998, 995.22]^3 and writes the domain grid built on the GPU by
// rpd_grid_from_mesh.
//
//   rpd_create_grid mesh.obj [--division=75] [--grid=out.grid] [--output=scaled.obj]
//
// --division: voxels per axis (main.cpp:127); --grid: the grid file
// (Grid::save format); --output: the scaled mesh, as .obj or .off by its
// extension (main.cpp:177-180).  Polygons are split into fans about their
// first vertex, as geogram's facets.triangulate() does.  Host code only.

#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "rpd.h"

namespace {

bool ends_with(const std::string &s, const std::string &e) {
    if (s.size() < e.size()) return false;
    for (size_t i = 0; i < e.size(); ++i)
        if (std::tolower(static_cast<unsigned char>(s[s.size() - e.size() + i])) != e[i]) return false;
    return true;
}

void fan(const std::vector<int> &face, std::vector<int32_t> &tris) {
    for (size_t i = 1; i + 1 < face.size(); ++i) tris.insert(tris.end(), {face[0], face[i], face[i + 1]});
}

bool read_off(std::istream &in, std::vector<double> &pts, std::vector<int32_t> &tris) {
    std::vector<std::string> tok;
    std::string line;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string w;
        while (is >> w) tok.push_back(w);
    }
    if (tok.size() < 4 || tok[0] != "OFF") return false;
    const long nv = std::atol(tok[1].c_str()), nf = std::atol(tok[2].c_str());
    size_t k = 4;
    for (long i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            if (k >= tok.size()) return false;
            pts.push_back(std::atof(tok[k++].c_str()));
        }
    for (long f = 0; f < nf; ++f) {
        if (k >= tok.size()) return false;
        const long c = std::atol(tok[k++].c_str());
        std::vector<int> face;
        for (long i = 0; i < c; ++i) {
            if (k >= tok.size()) return false;
            face.push_back(std::atoi(tok[k++].c_str()));
        }
        fan(face, tris);
    }
    return true;
}

bool read_obj(std::istream &in, std::vector<double> &pts, std::vector<int32_t> &tris) {
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream is(line);
        std::string w;
        if (!(is >> w)) continue;
        if (w == "v") {
            double x = 0, y = 0, z = 0;
            if (!(is >> x >> y >> z)) return false;
            pts.insert(pts.end(), {x, y, z});
        } else if (w == "f") {
            std::vector<int> face;
            std::string v;
            while (is >> v) {
                const int id = std::atoi(v.substr(0, v.find('/')).c_str());
                const int np = int(pts.size() / 3);
                face.push_back(id > 0 ? id - 1 : np + id);
            }
            fan(face, tris);
        }
    }
    return true;
}

bool write_mesh(const std::string &path, const rpd_grid_desc &g) {
    std::ofstream f(path);
    if (!f) return false;
    f.precision(17);
    if (ends_with(path, ".off")) {
        f << "OFF\n" << g.num_points << " " << g.num_triangles << " 0\n";
        for (int i = 0; i < g.num_points; ++i)
            f << g.points[4 * i] << " " << g.points[4 * i + 1] << " " << g.points[4 * i + 2] << "\n";
        for (int t = 0; t < g.num_triangles; ++t)
            f << "3 " << g.triangles[3 * t] << " " << g.triangles[3 * t + 1] << " " << g.triangles[3 * t + 2] << "\n";
    } else {
        for (int i = 0; i < g.num_points; ++i)
            f << "v " << g.points[4 * i] << " " << g.points[4 * i + 1] << " " << g.points[4 * i + 2] << "\n";
        for (int t = 0; t < g.num_triangles; ++t)
            f << "f " << g.triangles[3 * t] + 1 << " " << g.triangles[3 * t + 1] + 1 << " "
              << g.triangles[3 * t + 2] + 1 << "\n";
    }
    return f.good();
}

}  // namespace

int main(int argc, char **argv) {
    int division = 75;
    std::string mesh, grid_path, out_path;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--division=", 0) == 0)
            division = std::atoi(a.c_str() + 11);
        else if (a.rfind("--grid=", 0) == 0)
            grid_path = a.substr(7);
        else if (a.rfind("--output=", 0) == 0)
            out_path = a.substr(9);
        else if (a.rfind("--", 0) == 0) {
            std::cerr << "unknown option " << a << "\n";
            return 1;
        } else
            mesh = a;
    }
    if (mesh.empty()) {
        std::cout << "Error:" << argv[0] << " meshfile [--division=75] [--grid=out.grid] [--output=scaled.obj]\n";
        return 1;
    }
    if (division <= 0) division = 75;
    std::ifstream in(mesh);
    std::vector<double> pts;
    std::vector<int32_t> tris;
    if (!in || !(ends_with(mesh, ".off") ? read_off(in, pts, tris) : read_obj(in, pts, tris)) || tris.empty()) {
        std::cerr << "can not read the mesh " << mesh << "\n";
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    rpd_grid *grid = nullptr;
    const int rc = rpd_grid_from_mesh(division, pts.data(), int(pts.size() / 3), tris.data(), int(tris.size() / 3), 1,
                                      &grid);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc != RPD_OK) {
        std::cerr << "rpd_grid_from_mesh failed: " << rc << "\n";
        return 1;
    }
    rpd_grid_desc d;
    rpd_grid_view(grid, &d);
    std::cerr << "grid " << division << "^3 of " << d.num_triangles << " triangles: " << d.num_tri_list
              << " list entries, volume " << rpd_grid_volume(&d) << ", valid " << rpd_grid_valid(&d) << ", " << ms
              << " ms\n";
    int status = 0;
    if (!out_path.empty() && !write_mesh(out_path, d)) {
        std::cerr << "can not write " << out_path << "\n";
        status = 1;
    }
    if (!grid_path.empty() && rpd_grid_save(grid, grid_path.c_str()) != RPD_OK) {
        std::cerr << "can not write " << grid_path << "\n";
        status = 1;
    }
    rpd_grid_free(grid);
    return status;
}
