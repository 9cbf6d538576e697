// config.h -- host classes that keep the reference's interfaces around the
// C ABI: Settings (TOML keys of settings.cpp:22-323), Mesh (1D line grid,
// structured quads), Problem (initial conditions), MomentSolver (the time loop
// of momentsolver.cpp:37-269 driving libuqk on the device).
#pragma once

#include <string>
#include <vector>

#include "toml.h"
#include "uqk.h"

namespace uqk_host {

struct Settings {
    // [general]
    std::string problem;          // Burgers1D | Euler2D | ShallowWater2D
    std::string output_dir = ".";
    std::string restart_file;        // restartFile: a run log with _moments / _duals beside it
    bool import_dual_state = false;  // importDualState
    // [mesh]
    int dimension = 1;
    std::string mesh_type;        // line | nacaOGrid | quadGrid | SU2 (format = "SU2")
    std::string su2_file;         // SU2MeshFile, relative to the input file's directory
    std::vector<std::pair<std::string, std::string>> su2_bc;  // SU2BC: marker tag -> type
    std::string input_dir = ".";
    std::string input_path, input_text;  // the file itself, for the run log
    long n_cells_1d = 0;
    double a = 0.0, b = 0.0;
    int ni = 512, nj = 128;
    double radius = 20.0, first_height = 2e-3;
    int nx = 0, ny = 0;
    double x_min = -1, x_max = 1, y_min = -1, y_max = 1;
    int boundary = UQK_BC_DIRICHLET;
    // [problem]
    std::vector<double> sigma;
    double cfl = 0.5, t_end = 0.0, min_residual = -1.0, gamma = 1.4;
    // refinement retardation (steady runs): the level cap rises by one each
    // time the residual falls below the next of these, descending
    std::vector<double> retardation;
    int flux = UQK_FLUX_REFERENCE;
    long max_steps = -1;
    double friction = 0.0;
    int friction_dim = -1;
    std::string ic;
    // [moment_system]
    int closure = UQK_BOUNDED_BARRIER;
    bool max_degree = false;
    std::vector<int> degrees;  // one per refinement level
    double refine_above = 0.02, coarsen_below = 0.004;  // momentsolver.cpp:108-111
    int quad_points = 0;
    int max_iterations = 1000;
    double epsilon = 5e-5;
    double u_minus = 3.0, u_plus = 12.0;
    double barrier = 0.01;
    int precision = UQK_FP64;
    bool du_exact = true;
    bool safe_line_search = false;
    // derived
    int dim_xi = 0, n_states = 0;

    static Settings from_toml(const Toml& t);
    // from_toml of the file, with input_dir its directory (settings.cpp:30)
    static Settings from_file(const std::string& path);
};

struct Mesh {
    long n_cells = 0;
    int n_edges = 0, dim = 0;
    std::vector<int32_t> nbr, bc;
    std::vector<double> unit_normal, normal, area, dx_param, bnormal, center;
    // geometry for output files: node coordinates (x, y) and the nodes of each
    // cell (cell_node_count per cell, counter-clockwise or as read)
    std::vector<double> node_xy;
    std::vector<int64_t> cell_nodes;
    int cell_node_count = 0;

    static Mesh line(long n, double a, double b);
    static Mesh quad(const std::vector<double>& X, const std::vector<double>& Y, int ni_nodes, int nj_nodes,
                     bool periodic_i, int bc_j0, int bc_j1, int bc_i0, int bc_i1);
    static Mesh quad_grid(int nx, int ny, double x0, double x1, double y0, double y1, int bc);
    static Mesh naca_ogrid(int ni, int nj, double radius, double first_height);
    // SU2 file (mesh2d.cpp:68-255): triangles or quadrilaterals, markers typed by bcs
    static Mesh su2(const std::string& path, const std::vector<std::pair<std::string, std::string>>& bcs);
    static Mesh from_settings(const Settings& s);
};

// IC(x, xi) -> S values (Problem::IC, burgers.cpp:48-86, euler2d.cpp:136-206,
// shallowwater2d.cpp:77-107 and the config dam break)
void initial_state(const Settings& s, const double* x, const double* xi, double* out);

}  // namespace uqk_host
