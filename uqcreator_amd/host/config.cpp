// config.cpp -- Settings from TOML (settings.cpp:22-323 key names and
// semantics, plus the keys the configs add), meshes and initial conditions.
#include "config.h"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <fstream>
#include <stdexcept>

namespace uqk_host {

namespace {

const double kPi = 3.14159265358979323846;

std::vector<double> numbers(const TomlValue& v) {
    std::vector<double> out;
    if (v.kind == TomlValue::NUMBER) out.push_back(v.num);
    for (const auto& e : v.arr)
        if (e.kind == TomlValue::NUMBER) out.push_back(e.num);
    return out;
}

int closure_kind(const std::string& s) {
    if (s == "BoundedBarrier") return UQK_BOUNDED_BARRIER;
    if (s == "Euler") return UQK_EULER1D;
    if (s == "Euler2D") return UQK_EULER2D;
    if (s == "ShallowWater2D") return UQK_SHALLOW_WATER2D;
    if (s == "ShallowWater2DBarrier") return UQK_SHALLOW_WATER2D_BARRIER;
    throw std::runtime_error("[moment_system] 'closure' = '" + s + "' is not supported by the GPU path");
}

}  // namespace

Settings Settings::from_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream text;
    text << f.rdbuf();
    Settings s = from_toml(Toml::parse(text.str()));
    const size_t slash = path.find_last_of('/');
    s.input_dir = slash == std::string::npos ? "." : path.substr(0, slash);
    s.input_path = path;
    s.input_text = text.str();
    return s;
}

Settings Settings::from_toml(const Toml& t) {
    Settings s;
    s.problem = t.get_string("general", "problem", "");
    if (s.problem != "Burgers1D" && s.problem != "Euler2D" && s.problem != "ShallowWater2D")
        throw std::runtime_error("[general] 'problem' must be Burgers1D, Euler2D or ShallowWater2D");
    s.output_dir = t.get_string("general", "outputDir", ".");
    s.restart_file = t.get_string("general", "restartFile", "");  // settings.cpp:74-79
    s.import_dual_state = t.get_bool("general", "importDualState", false);
    s.dimension = (int)t.get_number("mesh", "dimension", s.problem == "Burgers1D" ? 1 : 2);
    s.mesh_type = t.get_string("mesh", "type", s.dimension == 1 ? "line" : "nacaOGrid");
    if (t.get_string("mesh", "format", "") == "SU2") {  // mesh2d.cpp:3-61
        s.mesh_type = "SU2";
        s.su2_file = t.get_string("mesh", "SU2MeshFile", "");
        if (s.su2_file.empty()) throw std::runtime_error("[mesh] 'SU2MeshFile' not set");
        const TomlValue* bcv = t.get("mesh", "SU2BC");
        if (!bcv || bcv->kind != TomlValue::ARRAY) throw std::runtime_error("[mesh] 'SU2BC' not set");
        for (const TomlValue& e : bcv->arr) {
            if (e.kind != TomlValue::ARRAY || e.arr.size() != 2)
                throw std::runtime_error("[mesh] 'SU2BC' entries are [tag, type]");
            s.su2_bc.push_back({e.arr[0].str, e.arr[1].str});
        }
    }
    s.n_cells_1d = (long)t.get_number("mesh", "numberOfCells", 0);
    s.a = t.get_number("mesh", "a", 0.0);
    s.b = t.get_number("mesh", "b", 0.0);
    s.ni = (int)t.get_number("mesh", "ni", 512);
    s.nj = (int)t.get_number("mesh", "nj", 128);
    s.radius = t.get_number("mesh", "radius", 20.0);
    s.first_height = t.get_number("mesh", "firstHeight", 2e-3);
    s.nx = (int)t.get_number("mesh", "nx", 0);
    s.ny = (int)t.get_number("mesh", "ny", 0);
    s.x_min = t.get_number("mesh", "xMin", -1.0);
    s.x_max = t.get_number("mesh", "xMax", 1.0);
    s.y_min = t.get_number("mesh", "yMin", -1.0);
    s.y_max = t.get_number("mesh", "yMax", 1.0);
    const std::string bnd = t.get_string("mesh", "boundary", "dirichlet");
    s.boundary = bnd == "swwall" ? UQK_BC_SWWALL : bnd == "noslip" ? UQK_BC_NOSLIP : UQK_BC_DIRICHLET;

    if (t.get_string("problem", "timestepping", "explicitEuler") != "explicitEuler")
        throw std::runtime_error("[problem] 'timestepping' must be explicitEuler");
    const TomlValue* dist = t.get("problem", "distribution");
    if (!dist || dist->kind != TomlValue::ARRAY || dist->arr.size() != 2)
        throw std::runtime_error("[problem] 'distribution' not set");
    for (const auto& d : dist->arr[0].arr)
        if (d.str != "Legendre") throw std::runtime_error("[problem] only Legendre distributions are supported");
    s.dim_xi = (int)dist->arr[0].arr.size();
    s.sigma = numbers(dist->arr[1]);
    if ((int)s.sigma.size() != s.dim_xi) throw std::runtime_error("[problem] 'distribution' sizes differ");
    if (!t.has("problem", "CFL")) throw std::runtime_error("[problem] 'CFL' not set");
    s.cfl = t.get_number("problem", "CFL", 0.5);
    if (!t.has("problem", "tEnd")) throw std::runtime_error("[problem] 'tEnd' not set");
    s.t_end = t.get_number("problem", "tEnd", 0.0);
    s.min_residual = t.get_number("problem", "residual", -1.0);
    if (const TomlValue* rt = t.get("problem", "retardationResiduals")) s.retardation = numbers(*rt);
    s.gamma = t.get_number("problem", "gamma", 1.4);
    const std::string flux = t.get_string("problem", "flux", "reference");
    s.flux = (flux == "laxFriedrichs" || flux == "rusanov") ? UQK_FLUX_LAX_FRIEDRICHS : UQK_FLUX_REFERENCE;
    s.max_steps = (long)t.get_number("problem", "maxSteps", -1);
    s.friction = t.get_number("problem", "friction", 0.0);
    s.friction_dim = (int)t.get_number("problem", "frictionDim", -1);
    s.ic = t.get_string("problem", "ic", s.problem == "Burgers1D" ? "ramp" : s.problem == "Euler2D" ? "naca" : "dambreak");

    s.closure = closure_kind(t.get_string("moment_system", "closure", ""));
    const TomlValue* mom = t.get("moment_system", "moments");
    if (!mom || mom->kind != TomlValue::ARRAY || mom->arr.size() != 2)
        throw std::runtime_error("[moment_system] 'moments' must be [[\"totalDegree\"|\"maxDegree\"], [degrees]]");
    const std::string dt = mom->arr[0].arr.empty() ? "" : mom->arr[0].arr[0].str;
    if (dt != "totalDegree" && dt != "maxDegree") throw std::runtime_error("[moment_system] 'moments' type is invalid");
    s.max_degree = dt == "maxDegree";
    for (double v : numbers(mom->arr[1])) s.degrees.push_back((int)v);
    if (s.degrees.empty()) throw std::runtime_error("[moment_system] 'moments' has no degree");
    // extension keys: the refinement thresholds the reference hard-codes
    s.refine_above = t.get_number("moment_system", "refineAbove", 0.02);
    s.coarsen_below = t.get_number("moment_system", "coarsenBelow", 0.004);
    const TomlValue* qp = t.get("moment_system", "quadPoints");
    if (!qp || qp->kind != TomlValue::ARRAY || qp->arr.size() != 2)
        throw std::runtime_error("[moment_system] 'quadPoints' must be [[\"tensorizedGrid\"], [n]]");
    const std::string gt = qp->arr[0].arr.empty() ? "" : qp->arr[0].arr[0].str;
    if (gt == "sparseGrid") throw std::runtime_error("[moment_system] sparse grids are not supported");
    std::vector<double> q = numbers(qp->arr[1]);
    if (q.empty()) throw std::runtime_error("[moment_system] 'quadPoints' has no size");
    s.quad_points = (int)q.back();  // settings.cpp:276
    s.max_iterations = (int)t.get_number("moment_system", "maxIterations", 1000);
    s.epsilon = t.get_number("moment_system", "epsilon", 5e-5);
    // bounded barrier defaults of boundedbarrier.cpp:7-14
    const bool two = s.dim_xi == 2;
    s.u_minus = t.get_number("moment_system", "uMinus", two ? 1.0 : 3.0);
    s.u_plus = t.get_number("moment_system", "uPlus", two ? 12.2 : 12.0);
    s.barrier = t.get_number("moment_system", "barrier", 0.01);
    s.precision = t.get_string("moment_system", "precision", "fp64") == "mixed" ? UQK_MIXED : UQK_FP64;
    s.du_exact = t.get_bool("moment_system", "duExact", true);
    s.safe_line_search = t.get_bool("moment_system", "safeLineSearch", false);
    s.n_states = s.closure == UQK_BOUNDED_BARRIER ? 1 : s.closure == UQK_EULER2D ? 4 : 3;
    if (s.problem == "Burgers1D" && s.closure != UQK_BOUNDED_BARRIER)
        throw std::runtime_error("Burgers1D needs the BoundedBarrier closure");
    if (s.problem == "Euler2D" && s.closure != UQK_EULER2D) throw std::runtime_error("Euler2D needs the Euler2D closure");
    if (s.problem == "ShallowWater2D" && s.closure != UQK_SHALLOW_WATER2D && s.closure != UQK_SHALLOW_WATER2D_BARRIER)
        throw std::runtime_error("ShallowWater2D needs a ShallowWater2D closure");
    return s;
}

// Mesh1D (mesh1d.cpp:5-67, line.cpp:3-44): cells 0 and N-1 Dirichlet, the
// cell centre is its left node, maxEdge = area^2 so dx_param = area.
Mesh Mesh::line(long n, double a, double b) {
    Mesh m;
    m.n_cells = n;
    m.n_edges = 2;
    m.dim = 1;
    m.nbr.resize(2 * n);
    m.bc.assign(n, UQK_BC_NONE);
    m.unit_normal.resize(2 * n);
    m.normal.resize(2 * n);
    m.area.resize(n);
    m.dx_param.resize(n);
    m.bnormal.assign(n, 0.0);
    m.center.resize(n);
    for (long j = 0; j < n; ++j) {
        const double x0 = a + j * (b - a) / n, x1 = a + (j + 1) * (b - a) / n;
        m.area[j] = std::fabs(x1 - x0);
        m.dx_param[j] = m.area[j] * m.area[j] / m.area[j];
        m.center[j] = x0;
        m.nbr[2 * j] = (j == 0 || j == n - 1) ? (int32_t)n : (int32_t)(j - 1);
        m.nbr[2 * j + 1] = (j == 0 || j == n - 1) ? (int32_t)n : (int32_t)(j + 1);
        m.unit_normal[2 * j] = m.normal[2 * j] = -1.0;
        m.unit_normal[2 * j + 1] = m.normal[2 * j + 1] = 1.0;
    }
    m.bc[0] = m.bc[n - 1] = UQK_BC_DIRICHLET;
    m.cell_node_count = 2;
    m.node_xy.assign(2 * (size_t)(n + 1), 0.0);
    m.cell_nodes.resize(2 * (size_t)n);
    for (long i = 0; i <= n; ++i) m.node_xy[2 * i] = a + i * (b - a) / n;
    for (long j = 0; j < n; ++j) {
        m.cell_nodes[2 * j] = j;
        m.cell_nodes[2 * j + 1] = j + 1;
    }
    return m;
}

// Quadrilaterals from node arrays X, Y (ni_nodes x nj_nodes, i fastest in
// the arrays' second index).  Cell (i, j) is j*ni + i, edge e joins node e and
// node e+1 of (i,j), (i+1,j), (i+1,j+1), (i,j+1); outward normals by the
// centroid test (quadrangle.cpp:72-83); shoelace area; dx_param = longest
// edge / area; the boundary normal is the edge whose neighbour id is largest
// (cell.cpp:57-68).
Mesh Mesh::quad(const std::vector<double>& X, const std::vector<double>& Y, int ni_nodes, int nj_nodes,
                bool periodic_i, int bc_j0, int bc_j1, int bc_i0, int bc_i1) {
    Mesh m;
    const int ni = periodic_i ? ni_nodes : ni_nodes - 1;
    const int nj = nj_nodes - 1;
    const long N = (long)ni * nj;
    m.n_cells = N;
    m.n_edges = 4;
    m.dim = 2;
    m.nbr.resize(4 * N);
    m.bc.assign(N, UQK_BC_NONE);
    m.unit_normal.resize(8 * N);
    m.normal.resize(8 * N);
    m.area.resize(N);
    m.dx_param.resize(N);
    m.bnormal.resize(2 * N);
    m.center.resize(2 * N);
    m.cell_node_count = 4;
    m.node_xy.resize(2 * (size_t)ni_nodes * nj_nodes);
    for (size_t k = 0; k < (size_t)ni_nodes * nj_nodes; ++k) {
        m.node_xy[2 * k] = X[k];
        m.node_xy[2 * k + 1] = Y[k];
    }
    m.cell_nodes.resize(4 * (size_t)N);
    auto node = [&](int i, int j, double& x, double& y) {
        x = X[(size_t)i * nj_nodes + j];
        y = Y[(size_t)i * nj_nodes + j];
    };
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const long c = (long)j * ni + i;
            const int ip = periodic_i ? (i + 1) % ni_nodes : i + 1;
            double px[4], py[4];
            node(i, j, px[0], py[0]);
            node(ip, j, px[1], py[1]);
            node(ip, j + 1, px[2], py[2]);
            node(i, j + 1, px[3], py[3]);
            m.cell_nodes[4 * c + 0] = (int64_t)i * nj_nodes + j;
            m.cell_nodes[4 * c + 1] = (int64_t)ip * nj_nodes + j;
            m.cell_nodes[4 * c + 2] = (int64_t)ip * nj_nodes + j + 1;
            m.cell_nodes[4 * c + 3] = (int64_t)i * nj_nodes + j + 1;
            const double cx = 0.25 * (px[0] + px[1] + px[2] + px[3]), cy = 0.25 * (py[0] + py[1] + py[2] + py[3]);
            double ar = 0.0, lmax = 0.0;
            for (int e = 0; e < 4; ++e) {
                const int f = (e + 1) % 4;
                ar += px[e] * py[f] - px[f] * py[e];
                const double dx = px[f] - px[e], dy = py[f] - py[e];
                const double len = std::hypot(dx, dy);
                lmax = std::max(lmax, len);
                double nx = dy, ny = -dx;
                if (nx * (cx - px[e]) + ny * (cy - py[e]) > 0) {
                    nx = -nx;
                    ny = -ny;
                }
                const double nn = std::hypot(nx, ny);
                m.unit_normal[8 * c + 2 * e] = nx / nn;
                m.unit_normal[8 * c + 2 * e + 1] = ny / nn;
                m.normal[8 * c + 2 * e] = nx / nn * len;
                m.normal[8 * c + 2 * e + 1] = ny / nn * len;
            }
            m.area[c] = 0.5 * std::fabs(ar);
            m.dx_param[c] = lmax / m.area[c];
            m.center[2 * c] = cx;
            m.center[2 * c + 1] = cy;
            auto id = [&](int ii, int jj) { return (int32_t)((long)jj * ni + ii); };
            m.nbr[4 * c + 0] = j > 0 ? id(i, j - 1) : (int32_t)N;
            m.nbr[4 * c + 2] = j < nj - 1 ? id(i, j + 1) : (int32_t)N;
            if (periodic_i) {
                m.nbr[4 * c + 1] = id((i + 1) % ni, j);
                m.nbr[4 * c + 3] = id((i - 1 + ni) % ni, j);
            } else {
                m.nbr[4 * c + 1] = i < ni - 1 ? id(i + 1, j) : (int32_t)N;
                m.nbr[4 * c + 3] = i > 0 ? id(i - 1, j) : (int32_t)N;
            }
            int bc = UQK_BC_NONE;
            if (j == 0) bc = bc_j0;
            if (j == nj - 1) bc = bc_j1;
            if (!periodic_i && i == 0) bc = bc_i0;
            if (!periodic_i && i == ni - 1) bc = bc_i1;
            m.bc[c] = bc;
            int best = 0;
            for (int e = 1; e < 4; ++e)
                if (m.nbr[4 * c + e] > m.nbr[4 * c + best]) best = e;
            m.bnormal[2 * c] = m.unit_normal[8 * c + 2 * best];
            m.bnormal[2 * c + 1] = m.unit_normal[8 * c + 2 * best + 1];
        }
    return m;
}

Mesh Mesh::quad_grid(int nx, int ny, double x0, double x1, double y0, double y1, int bc) {
    std::vector<double> X((size_t)(nx + 1) * (ny + 1)), Y(X.size());
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j) {
            X[(size_t)i * (ny + 1) + j] = x0 + (x1 - x0) * i / nx;
            Y[(size_t)i * (ny + 1) + j] = y0 + (y1 - y0) * j / ny;
        }
    return quad(X, Y, nx + 1, ny + 1, false, bc, bc, bc, bc);
}

static double naca0012(double x) {
    return 0.6 * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
}

// O-grid around an analytic NACA0012 of unit chord (half cosine, half uniform
// spacing along the chord, lower side first from the trailing edge), nj rings
// out to a circle of `radius` chords around mid chord, geometric stretching
// from `first_height`.  Profile ring: slip wall; outer ring: Dirichlet.
Mesh Mesh::naca_ogrid(int ni, int nj, double radius, double first_height) {
    const int half = ni / 2;
    std::vector<double> sx, sy;
    std::vector<double> xs(half + 1);
    for (int t = 0; t <= half; ++t) {
        const double th = kPi * t / half;
        xs[t] = 0.5 * (0.5 * (1 + std::cos(th)) + (1 - th / kPi));
    }
    for (int t = 0; t <= half; ++t) {
        sx.push_back(xs[t]);
        sy.push_back(-naca0012(xs[t]));
    }
    for (int t = half - 1; t >= 1; --t) {
        sx.push_back(xs[t]);
        sy.push_back(naca0012(xs[t]));
    }
    const int n = (int)sx.size();  // == ni
    std::vector<double> fx(n), fy(n);
    double dist = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ang = std::atan2(sy[i], sx[i] - 0.5);
        fx[i] = 0.5 + radius * std::cos(ang);
        fy[i] = radius * std::sin(ang);
        dist += std::hypot(fx[i] - sx[i], fy[i] - sy[i]);
    }
    dist /= n;
    const double h0 = first_height / dist;
    double lo = 1.0 + 1e-9, hi = 2.0, r = 1.0;
    for (int it = 0; it < 200; ++it) {
        r = 0.5 * (lo + hi);
        const double tot = h0 * (std::pow(r, nj) - 1) / (r - 1);
        if (tot < 1.0)
            lo = r;
        else
            hi = r;
    }
    std::vector<double> s(nj + 1, 0.0);
    for (int k = 1; k <= nj; ++k) s[k] = h0 * (std::pow(r, k) - 1) / (r - 1);
    for (int k = 0; k <= nj; ++k) s[k] /= s[nj];
    std::vector<double> X((size_t)n * (nj + 1)), Y(X.size());
    for (int i = 0; i < n; ++i)
        for (int k = 0; k <= nj; ++k) {
            X[(size_t)i * (nj + 1) + k] = sx[i] + (fx[i] - sx[i]) * s[k];
            Y[(size_t)i * (nj + 1) + k] = sy[i] + (fy[i] - sy[i]) * s[k];
        }
    return quad(X, Y, n, nj + 1, true, UQK_BC_NOSLIP, UQK_BC_DIRICHLET, 0, 0);
}

Mesh Mesh::from_settings(const Settings& s) {
    if (s.mesh_type == "line") {
        if (s.n_cells_1d <= 2 || (std::fabs(s.a) == 0.0 && std::fabs(s.b) == 0.0))
            throw std::runtime_error("[mesh] 'numberOfCells', 'a', 'b' must describe a line grid");
        return line(s.n_cells_1d, s.a, s.b);
    }
    if (s.mesh_type == "nacaOGrid") return naca_ogrid(s.ni, s.nj, s.radius, s.first_height);
    if (s.mesh_type == "SU2") {
        const std::string p = (!s.su2_file.empty() && s.su2_file[0] == '/') ? s.su2_file : s.input_dir + "/" + s.su2_file;
        return su2(p, s.su2_bc);
    }
    if (s.mesh_type == "quadGrid") {
        if (s.nx <= 0 || s.ny <= 0) throw std::runtime_error("[mesh] 'nx', 'ny' not set");
        return quad_grid(s.nx, s.ny, s.x_min, s.x_max, s.y_min, s.y_max, s.boundary);
    }
    throw std::runtime_error("[mesh] 'type' must be line, nacaOGrid or quadGrid, or format = \"SU2\"");
}

void initial_state(const Settings& s, const double* x, const double* xi, double* out) {
    const std::vector<double>& sg = s.sigma;
    if (s.problem == "Burgers1D") {
        const double x0 = 0.5, x1 = 1.5, uL = 12.0, uR = 3.0;
        const double sh = sg.empty() ? 0.0 : sg[0] * xi[0];
        if (s.ic == "step") {
            out[0] = x[0] < x0 + sh ? uL : uR;  // the config's shock at x0 + sigma xi
        } else {
            // burgers.cpp:51-67: ramp from x0 to x1
            if (x[0] < x0 + sh)
                out[0] = uL;
            else if (x[0] < x1 + sh)
                out[0] = uL + (uR - uL) * (x0 + sh - x[0]) / (x0 - x1);
            else
                out[0] = uR;
        }
        return;
    }
    if (s.problem == "Euler2D") {
        const double gamma = 1.4, R = 287.87, T = 273.15, p = 101325.0;
        const double a = std::sqrt(gamma * R * T);
        const double rho = p / (R * T);
        if (s.ic == "pipe") {
            // euler2d.cpp:139-180 (the branch the reference runs)
            const double s0 = sg.empty() ? 0.0 : sg[0];
            out[0] = rho;
            out[1] = 0.0;
            out[2] = 0.0;
            out[3] = 1.0;
            if (x[1] < 1.1 + s0 * xi[0]) {
                out[0] = 0.5 * rho + (s.dim_xi > 1 ? sg[1] * xi[1] : 0.0);
                out[3] = 0.3 + (s.dim_xi > 2 ? sg[2] * xi[2] : 0.0);
            }
            return;
        }
        // euler2d.cpp:182-206: uncertain angle of attack and Mach number
        const double ma = 0.8 + (s.dim_xi > 1 ? sg[1] * xi[1] : 0.0);
        const double ang = (1.25 + (sg.empty() ? 0.0 : sg[0] * xi[0])) * (2.0 * kPi) / 360.0;
        const double uf = ma * a * std::cos(ang), vf = ma * a * std::sin(ang);
        out[0] = rho;
        out[1] = rho * uf;
        out[2] = rho * vf;
        out[3] = 0.5 * rho * (uf * uf + vf * vf) + p / (gamma - 1);
        return;
    }
    // shallow water dam break: depth (2 + sigma0 xi0) left of x = sigma1 xi1, 1 right
    const double hl = 2.0 + (sg.empty() ? 0.0 : sg[0] * xi[0]);
    const double dam = s.dim_xi > 1 ? sg[1] * xi[1] : 0.0;
    out[0] = x[0] < dam ? hl : 1.0;
    out[1] = 0.0;
    out[2] = 0.0;
}

}  // namespace uqk_host
