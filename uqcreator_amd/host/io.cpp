// io.cpp -- the solver's files: mean/variance fields for visualisation,
// restart files, and the run log that carries the input for a restart.
//
//   write_fields     Mesh2D::Export / ExportShallowWater (mesh2d.cpp:399-615)
//                    and Mesh1D::Export (mesh1d.cpp:90-220): a legacy ASCII
//                    VTK unstructured grid with the E/Var arrays as cell data
//                    and their averages over the cells of each node as point
//                    data (vtkCellDataToPointData); 1D meshes become a strip
//                    of quadrilaterals one tenth of the domain high and also
//                    get the ExpectedValue / Variance text files.
//   write_log        the header and the input echo of main.cpp:160-180, each
//                    line prefixed "date time | " as the event log's sink.
//   read_log_config  MomentSolver::ImportPrevSettings (momentsolver.cpp:592-612).
//   write_restart_field / read_restart_field
//                    MomentSolver::Export and ImportPrevMoments /
//                    ImportPrevDuals (momentsolver.cpp:565-650): one line per
//                    cell, states outer, moments inner, 15 significant digits,
//                    each value followed by a comma.
// Array names are escaped as VTK's legacy writer escapes them (bytes outside
// '!'..'~' and '%' as %XX); every cell gets its own type (the reference
// declares all 2D cells triangles).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.h"
#include "io.h"

namespace uqk_host {

namespace {

std::string vtk_name(const std::string& s) {
    std::string out;
    for (unsigned char c : s) {
        if (c <= ' ' || c > '~' || c == '%') {
            char b[4];
            std::snprintf(b, sizeof b, "%%%02X", c);
            out += b;
        } else {
            out += (char)c;
        }
    }
    return out;
}

struct Field {
    std::string name;
    int comp;      // 1 or 3
    int rows[2];   // rows of the mean/variance matrix (x, y for vectors)
};

std::vector<Field> fields_of(const Settings& s) {
    if (s.problem == "Burgers1D") return {{"E(u)", 1, {0, 0}}, {"Var(u)", 1, {1, 1}}};
    if (s.problem == "ShallowWater2D")
        return {{"E(h)", 1, {0, 0}}, {"E(hU)", 3, {1, 2}}, {"Var(h)", 1, {3, 3}}, {"Var(hU)", 3, {4, 5}}};
    return {{"E(\xcf\x81)", 1, {0, 0}},     {"E(\xcf\x81U)", 3, {1, 2}},   {"E(\xcf\x81" "E)", 1, {3, 3}},
            {"Var(\xcf\x81)", 1, {4, 4}},   {"Var(\xcf\x81U)", 3, {5, 6}}, {"Var(\xcf\x81" "E)", 1, {7, 7}}};
}

void put_values(std::ostream& f, const std::vector<double>& v) {
    for (size_t i = 0; i < v.size(); ++i) f << v[i] << ((i % 9 == 8 || i + 1 == v.size()) ? "\n" : " ");
}

}  // namespace

void write_fields(const std::string& prefix, const Settings& s, const Mesh& m, const std::vector<double>& mv) {
    const long N = m.n_cells;
    const int S = s.n_states, F = 2 * S;
    if ((long)mv.size() != N * F) throw std::runtime_error("write_fields: mean/variance size");
    // points and cells
    std::vector<double> pts;  // x, y, z
    std::vector<std::vector<int64_t>> cells((size_t)N);
    std::vector<int> types((size_t)N);
    if (m.dim == 1) {
        const long P = (long)m.node_xy.size() / 2;
        const double height = (m.node_xy[2 * (P - 1)] - m.node_xy[0]) / 10.0;  // mesh1d.cpp:92
        pts.assign((size_t)6 * P, 0.0);
        for (long i = 0; i < P; ++i) {
            pts[3 * i] = m.node_xy[2 * i];
            pts[3 * (i + P)] = m.node_xy[2 * i];
            pts[3 * (i + P) + 1] = height;
        }
        for (long j = 0; j < N; ++j) {
            const int64_t a = m.cell_nodes[2 * j], b = m.cell_nodes[2 * j + 1];
            cells[j] = {a, a + P, b + P, b};
            types[j] = 9;
        }
    } else {
        const long P = (long)m.node_xy.size() / 2;
        pts.assign((size_t)3 * P, 0.0);
        for (long i = 0; i < P; ++i) {
            pts[3 * i] = m.node_xy[2 * i];
            pts[3 * i + 1] = m.node_xy[2 * i + 1];
        }
        const int K = m.cell_node_count;
        for (long j = 0; j < N; ++j) {
            cells[j].assign(m.cell_nodes.begin() + (size_t)j * K, m.cell_nodes.begin() + (size_t)(j + 1) * K);
            types[j] = K == 3 ? 5 : 9;
        }
        if (s.problem == "ShallowWater2D") {
            // ExportShallowWater: z = node average of E(h) over the largest E(h)
            std::vector<double> hsum((size_t)P, 0.0), cnt((size_t)P, 0.0);
            double hmax = 0.0;
            for (long j = 0; j < N; ++j)
                for (int64_t v : cells[j]) {
                    hmax = std::max(hmax, mv[(size_t)j * F]);
                    hsum[v] += mv[(size_t)j * F];
                    cnt[v] += 1.0;
                }
            for (long i = 0; i < P; ++i) pts[3 * i + 2] = cnt[i] > 0 ? hsum[i] / (cnt[i] * hmax) : 0.0;
        }
    }
    const long P = (long)pts.size() / 3;
    std::vector<double> users((size_t)P, 0.0);
    for (long j = 0; j < N; ++j)
        for (int64_t v : cells[j]) users[v] += 1.0;

    std::ofstream f(prefix + ".vtk");
    if (!f) throw std::runtime_error("cannot write " + prefix + ".vtk");
    f << std::setprecision(17);
    f << "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    f << "POINTS " << P << " double\n";
    put_values(f, pts);
    size_t total = 0;
    for (const auto& c : cells) total += c.size() + 1;
    f << "CELLS " << N << " " << total << "\n";
    for (const auto& c : cells) {
        f << c.size();
        for (int64_t v : c) f << " " << v;
        f << "\n";
    }
    f << "CELL_TYPES " << N << "\n";
    for (int t : types) f << t << "\n";
    const std::vector<Field> fl = fields_of(s);
    auto block = [&](bool points) {
        f << "FIELD FieldData " << fl.size() << "\n";
        for (const Field& fd : fl) {
            const long n = points ? P : N;
            std::vector<double> v((size_t)n * fd.comp, 0.0);
            for (long j = 0; j < N; ++j)
                for (int c = 0; c < std::min(fd.comp, 2); ++c) {
                    const double x = mv[(size_t)j * F + fd.rows[c]];
                    if (points) {
                        for (int64_t p : cells[j]) v[(size_t)p * fd.comp + c] += x / users[p];
                    } else {
                        v[(size_t)j * fd.comp + c] = x;
                    }
                }
            f << vtk_name(fd.name) << " " << fd.comp << " " << n << " double\n";
            put_values(f, v);
        }
    };
    f << "CELL_DATA " << N << "\n";
    block(false);
    f << "POINT_DATA " << P << "\n";
    block(true);
    if (!f) throw std::runtime_error("write failed: " + prefix + ".vtk");

    if (m.dim == 1) {  // mesh1d.cpp:198-218, default stream precision
        std::ofstream e(prefix + "ExpectedValue"), v(prefix + "Variance");
        for (long j = 0; j < N; ++j) {
            e << m.center[j];
            v << m.center[j];
            for (int q = 0; q < S; ++q) {
                e << " " << mv[(size_t)j * F + q];
                v << " " << mv[(size_t)j * F + S + q];
            }
            e << "\n";
            v << "\n";
        }
    }
}

void write_log(const std::string& path, const std::string& input_path, const std::string& input_text,
               const std::vector<std::string>& extra) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    auto stamp = []() {
        const auto now = std::chrono::system_clock::now();
        const std::time_t t = std::chrono::system_clock::to_time_t(now);
        const long us = (long)(std::chrono::duration_cast<std::chrono::microseconds>(now.time_since_epoch()).count() %
                               1000000);
        std::tm tm{};
        gmtime_r(&t, &tm);
        char b[64];
        std::strftime(b, sizeof b, "%Y-%m-%d %H:%M:%S", &tm);
        char c[80];
        std::snprintf(c, sizeof c, "%s.%06ld | ", b, us);
        return std::string(c);
    };
    const std::string bar(64, '=');
    f << stamp() << "UQCreator\n" << stamp() << bar << "\n";
    f << stamp() << "Git commit hash:\t-\n" << stamp() << "Config file:\t" << input_path << "\n";
    f << stamp() << "MPI Threads:\t1\n" << stamp() << "OMP Threads:\t1\n" << stamp() << bar << "\n";
    std::istringstream in(input_text);
    for (std::string line; std::getline(in, line);)
        if (line.empty() || line[0] != '#') f << stamp() << " " << line << "\n";
    f << stamp() << bar << "\n" << stamp() << "\n";
    for (const std::string& l : extra) f << stamp() << l << "\n";
}

std::string read_log_config(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("restart log not found: " + path);
    std::string line, out;
    bool section = false;
    while (std::getline(f, line)) {
        if (line.find("Config file") != std::string::npos) {
            section = true;
            for (int i = 0; i < 4; ++i) std::getline(f, line);  // thread counts and the bar
        } else if (section && line.find("==================================") != std::string::npos) {
            break;
        }
        if (section) {
            const size_t bar = line.find('|');
            out += (bar == std::string::npos ? line : line.substr(bar + 1)) + "\n";
        }
    }
    if (!section) throw std::runtime_error("restart log has no input section: " + path);
    return out;
}

void write_restart_field(const std::string& path, const std::vector<double>& x, long n_cells, int S, int M) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << std::setprecision(15);  // std::numeric_limits<double>::digits10
    for (long j = 0; j < n_cells; ++j) {
        for (int s = 0; s < S; ++s)
            for (int i = 0; i < M; ++i) f << x[((size_t)j * M + i) * S + s] << ",";
        f << "\n";
    }
}

std::vector<double> read_restart_field(const std::string& path, long n_cells, int S, int M) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("restart file not found: " + path);
    std::vector<double> x((size_t)n_cells * S * M, 0.0);
    std::string line, cell;
    for (long j = 0; j < n_cells; ++j) {
        if (!std::getline(f, line)) throw std::runtime_error(path + ": fewer lines than cells");
        std::stringstream ls(line);
        for (int s = 0; s < S; ++s)
            for (int i = 0; i < M; ++i) {
                if (!std::getline(ls, cell, ',')) throw std::runtime_error(path + ": short line " + std::to_string(j));
                x[((size_t)j * M + i) * S + s] = std::stod(cell);
            }
    }
    return x;
}

}  // namespace uqk_host
