// quadrature.cpp -- host-side numerics substrate: Gauss-Legendre rules,
// Legendre basis, the total-degree index set and the Closure's basis tables.
//
//   Legendre::Compute       legendre.cpp:9-28: Golub-Welsch -- eigenvalues of
//                           the symmetric Jacobi matrix are the nodes and the
//                           squared first eigenvector components (times 2)
//                           the weights; nodes sorted ascending
//                           (polynomial.cpp:7-16), |x| < 1e-15 snapped to 0.
//   Settings index set      settings.cpp:242-266
//   TensorizedQuadrature    tensorizedquadrature.cpp:27-74 (dimension 0 fastest)
//   Closure tables          closure.cpp:45-66
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "uqk.h"
#include "uqk_host.h"

namespace uqk_host {

// Symmetric tridiagonal eigenproblem by implicit QL with Wilkinson shifts:
// the published tql2 method (Bowdler, Martin, Reinsch and Wilkinson, Numer.
// Math. 11 (1968) 293-306; EISPACK TQL2), the method mathtools.cpp:41-104
// also follows.
// d: diagonal (eigenvalues on return), e: sub-diagonal e[0..n-2],
// z: n x n row-major, returns eigenvectors in columns.
static bool tridiag_ql(std::vector<double>& d, std::vector<double>& e, std::vector<double>& z, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) z[i * n + j] = (i == j) ? 1.0 : 0.0;
    e.resize(n);
    e[n - 1] = 0.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) return false;
            // Wilkinson shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    double zk1 = z[k * n + i + 1];
                    z[k * n + i + 1] = s * z[k * n + i] + c * zk1;
                    z[k * n + i] = c * z[k * n + i] - s * zk1;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

bool gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    if (n <= 0) return false;
    if (n == 1) {
        w[0] = 2.0;
        return true;
    }
    std::vector<double> d(n, 0.0), e(n, 0.0), z((size_t)n * n);
    for (int i = 0; i < n - 1; ++i) {
        double k = i + 1.0;
        e[i] = std::sqrt(1.0 / (4.0 - 1.0 / (k * k)));
    }
    if (!tridiag_ql(d, e, z, n)) return false;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < n; ++i) {
        int j = order[i];
        x[i] = std::fabs(d[j]) < 1e-15 ? 0.0 : d[j];
        w[i] = 2.0 * z[0 * n + j] * z[0 * n + j];
    }
    return true;
}

double legendre_p(int n, double x) {
    if (n == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

std::vector<std::vector<int>> index_set(int d, const std::vector<int>& degrees, bool max_degree,
                                        std::vector<int>& ntotal_per_level) {
    std::vector<std::vector<int>> out;
    const int N = degrees.back();
    long span = 1;
    for (int l = 0; l < d; ++l) span *= (N + 1);
    ntotal_per_level.assign(degrees.size(), 0);
    int prev = -1;
    for (size_t level = 0; level < degrees.size(); ++level) {
        for (long i = 0; i < span; ++i) {
            std::vector<int> mi(d);
            long rem = i;
            int td = 0;
            for (int l = 0; l < d; ++l) {
                mi[l] = int(rem % (N + 1));
                rem /= (N + 1);
                td += mi[l];
            }
            if ((td <= degrees[level] && td > prev) || max_degree) out.push_back(mi);
        }
        ntotal_per_level[level] = int(out.size());
        if (max_degree) {
            for (size_t l2 = level + 1; l2 < degrees.size(); ++l2) ntotal_per_level[l2] = int(out.size());
            break;
        }
        prev = degrees[level];
    }
    return out;
}

bool basis_tables(int d, int q1, const std::vector<std::vector<int>>& idx, std::vector<double>& xi,
                  std::vector<double>& wq, std::vector<double>& phi, std::vector<double>& phiw) {
    std::vector<double> x1, w1;
    if (!gauss_legendre(q1, x1, w1)) return false;
    long Q = 1;
    for (int l = 0; l < d; ++l) Q *= q1;
    const int M = int(idx.size());
    xi.assign((size_t)Q * d, 0.0);
    wq.assign(Q, 1.0);
    phi.assign((size_t)Q * M, 1.0);
    phiw.assign((size_t)Q * M, 1.0);
    for (long k = 0; k < Q; ++k) {
        long rem = k;
        for (int l = 0; l < d; ++l) {
            int a = int(rem % q1);
            rem /= q1;
            xi[k * d + l] = x1[a];
            wq[k] *= w1[a];
        }
        for (int i = 0; i < M; ++i) {
            double p = 1.0, pw = 1.0;
            for (int l = 0; l < d; ++l) {
                int deg = idx[i][l];
                // P_n(x) / ||P_n|| with ||P_n||^2 = 1/(2n+1) (legendre.cpp:38), f_Xi = 1/2 (:36)
                double v = legendre_p(deg, xi[k * d + l]) / std::sqrt(1.0 / (2.0 * deg + 1.0));
                p *= v;
                pw *= v * 0.5;
            }
            phi[k * M + i] = p;
            phiw[k * M + i] = pw * wq[k];
        }
    }
    return true;
}

}  // namespace uqk_host

extern "C" int uqk_gauss_legendre(int32_t n, double* x, double* w) {
    std::vector<double> xv, wv;
    if (!uqk_host::gauss_legendre(n, xv, wv)) return -1;
    std::copy(xv.begin(), xv.end(), x);
    std::copy(wv.begin(), wv.end(), w);
    return 0;
}

extern "C" int32_t uqk_index_set(int32_t d, int32_t n_levels, const int32_t* degrees, int32_t max_degree,
                                 int32_t cap, int32_t* idx, int32_t* ntotal_per_level) {
    std::vector<int> deg(degrees, degrees + n_levels), ntl;
    auto set = uqk_host::index_set(d, deg, max_degree != 0, ntl);
    if ((int)set.size() > cap) return -1;
    for (size_t i = 0; i < set.size(); ++i)
        for (int l = 0; l < d; ++l) idx[i * d + l] = set[i][l];
    for (int l = 0; l < n_levels; ++l) ntotal_per_level[l] = ntl[l];
    return int32_t(set.size());
}

extern "C" int32_t uqk_basis_tables(int32_t d, int32_t q1, int32_t m, const int32_t* idx, double* xi, double* wq,
                                    double* phi, double* phiw) {
    std::vector<std::vector<int>> set(m, std::vector<int>(d));
    for (int i = 0; i < m; ++i)
        for (int l = 0; l < d; ++l) set[i][l] = idx[i * d + l];
    std::vector<double> x, w, p, pw;
    if (!uqk_host::basis_tables(d, q1, set, x, w, p, pw)) return -1;
    std::copy(x.begin(), x.end(), xi);
    std::copy(w.begin(), w.end(), wq);
    std::copy(p.begin(), p.end(), phi);
    std::copy(pw.begin(), pw.end(), phiw);
    return int32_t(w.size());
}
