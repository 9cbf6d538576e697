// io.h -- output and restart files of a run (see io.cpp).
#pragma once

#include <string>
#include <vector>

#include "config.h"

namespace uqk_host {

// <prefix>.vtk with the E/Var fields (mv: n_cells x 2S, means then variances);
// 1D meshes also <prefix>ExpectedValue and <prefix>Variance.
void write_fields(const std::string& prefix, const Settings& s, const Mesh& m, const std::vector<double>& mv);
// run log: header, the input file's lines, then `extra` lines.
void write_log(const std::string& path, const std::string& input_path, const std::string& input_text,
               const std::vector<std::string>& extra);
// the input echoed in a run log, as TOML text
std::string read_log_config(const std::string& path);
// x: n_cells x M x S (state fastest), one text line per cell
void write_restart_field(const std::string& path, const std::vector<double>& x, long n_cells, int S, int M);
std::vector<double> read_restart_field(const std::string& path, long n_cells, int S, int M);

}  // namespace uqk_host
