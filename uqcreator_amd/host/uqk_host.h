// uqk_host.h -- host-side C++ classes and numerics of the moment solver.
#pragma once

#include <vector>

namespace uqk_host {

bool gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
double legendre_p(int n, double x);
std::vector<std::vector<int>> index_set(int d, const std::vector<int>& degrees, bool max_degree,
                                        std::vector<int>& ntotal_per_level);
bool basis_tables(int d, int q1, const std::vector<std::vector<int>>& idx, std::vector<double>& xi,
                  std::vector<double>& wq, std::vector<double>& phi, std::vector<double>& phiw);

}  // namespace uqk_host
