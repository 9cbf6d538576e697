/*
 * mlp_oracle.cpp -- FP64 CPU reference of the learned warm start's network.
 *
 * TEST INFRASTRUCTURE ONLY (same rule as oracle.cpp): loaded by tests/ as the
 * reference the device forward/backward kernels are checked against.  The network has no
 * counterpart in UQCreator; it follows the textbook definitions:
 *   z_l = W_l a_{l-1} + b_l,  a_l = GELU(z_l) (l < L),  y = z_L,
 *   GELU(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))   (Hendrycks and
 *   Gimpel 2016, tanh form), loss = mean over rows and the first n_real
 *   outputs of (y - t)^2, and its gradient by reverse-mode differentiation.
 * Parameters are one flat array [W_1 (N_1 x K_1 row-major), b_1, W_2, b_2,
 * ...] with K_1 = dims[0], N_l = K_{l+1} = dims[l].
 *
 * With emulate_bf16 = 1 the values the device rounds to bfloat16 (round to
 * nearest even) are rounded the same way here: the layer inputs a_{l-1}, the
 * weights, the output gradient dy and the pre-activation gradients dz; sums
 * stay FP64, so what remains is the device's FP32 accumulation order.
 *
 * Pinning: the gradient is checked against central differences of the loss
 * (tests/test_mlp_oracle.py) and GELU' against the difference quotient.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

const double kC = 0.7978845608028654;  // sqrt(2/pi)

double gelu(double z) { return 0.5 * z * (1.0 + std::tanh(kC * (z + 0.044715 * z * z * z))); }

double gelu_prime(double z) {
    const double u = kC * (z + 0.044715 * z * z * z);
    const double th = std::tanh(u);
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * kC * (1.0 + 3.0 * 0.044715 * z * z);
}

long param_count(int nl, const int* dims) {
    long n = 0;
    for (int l = 0; l < nl; ++l) n += (long)dims[l + 1] * dims[l] + dims[l + 1];
    return n;
}

}  // namespace

extern "C" {

long orc_mlp_param_count(int nl, const int* dims) { return param_count(nl, dims); }

double orc_gelu(double z, int derivative) { return derivative ? gelu_prime(z) : gelu(z); }

double orc_bf16_round(double v) { return bf16_round(v); }

/* Forward pass of B rows: x (B x dims[0]) -> y (B x dims[nl]). */
void orc_mlp_forward(int nl, const int* dims, const double* params, long B, const double* x, double* y,
                     int emulate_bf16) {
    auto r = [&](double v) { return emulate_bf16 ? bf16_round(v) : v; };
    int wmax = 0;
    for (int l = 0; l <= nl; ++l) wmax = dims[l] > wmax ? dims[l] : wmax;
#pragma omp parallel for schedule(static)
    for (long b = 0; b < B; ++b) {
        std::vector<double> a(wmax), z(wmax);
        for (int k = 0; k < dims[0]; ++k) a[k] = x[b * dims[0] + k];
        const double* p = params;
        for (int l = 0; l < nl; ++l) {
            const int K = dims[l], N = dims[l + 1];
            const double* W = p;
            const double* bias = p + (long)N * K;
            for (int n = 0; n < N; ++n) {
                double s = 0.0;
                for (int k = 0; k < K; ++k) s += r(W[(long)n * K + k]) * r(a[k]);
                z[n] = s + bias[n];
            }
            for (int n = 0; n < N; ++n) a[n] = l + 1 < nl ? gelu(z[n]) : z[n];
            p += (long)N * K + N;
        }
        for (int n = 0; n < dims[nl]; ++n) y[b * dims[nl] + n] = a[n];
    }
}

/* Loss mean((y - t)^2) over B rows and the first n_real outputs (t is
 * B x dims[nl]); grad (same layout as params) receives dLoss/dparams. */
double orc_mlp_loss_grad(int nl, const int* dims, const double* params, long B, const double* x, const double* t,
                         int n_real, double* grad, int emulate_bf16) {
    auto r = [&](double v) { return emulate_bf16 ? bf16_round(v) : v; };
    const long np = param_count(nl, dims);
    std::memset(grad, 0, sizeof(double) * np);
    std::vector<long> off(nl + 1, 0);
    for (int l = 0; l < nl; ++l) off[l + 1] = off[l] + (long)dims[l + 1] * dims[l] + dims[l + 1];
    double loss = 0.0;
    const double scale = 1.0 / ((double)B * n_real);
    for (long b = 0; b < B; ++b) {
        std::vector<std::vector<double>> a(nl + 1), z(nl + 1);
        a[0].assign(x + b * dims[0], x + (b + 1) * dims[0]);
        for (int l = 0; l < nl; ++l) {
            const int K = dims[l], N = dims[l + 1];
            const double* W = params + off[l];
            const double* bias = W + (long)N * K;
            z[l + 1].resize(N);
            a[l + 1].resize(N);
            for (int n = 0; n < N; ++n) {
                double s = 0.0;
                for (int k = 0; k < K; ++k) s += r(W[(long)n * K + k]) * r(a[l][k]);
                z[l + 1][n] = s + bias[n];
                a[l + 1][n] = l + 1 < nl ? gelu(z[l + 1][n]) : z[l + 1][n];
            }
        }
        const int NO = dims[nl];
        std::vector<double> dz(NO, 0.0);
        for (int n = 0; n < n_real; ++n) {
            const double e = a[nl][n] - t[b * NO + n];
            loss += e * e * scale;
            dz[n] = 2.0 * e * scale;
        }
        for (int l = nl - 1; l >= 0; --l) {
            const int K = dims[l], N = dims[l + 1];
            const double* W = params + off[l];
            double* gW = grad + off[l];
            double* gb = gW + (long)N * K;
            std::vector<double> dzr(N);
            for (int n = 0; n < N; ++n) dzr[n] = r(dz[n]);
            for (int n = 0; n < N; ++n) {
                gb[n] += l + 1 < nl ? dzr[n] : dz[n];
                for (int k = 0; k < K; ++k) gW[(long)n * K + k] += dzr[n] * r(a[l][k]);
            }
            if (l == 0) break;
            std::vector<double> da(K, 0.0);
            for (int k = 0; k < K; ++k) {
                double s = 0.0;
                for (int n = 0; n < N; ++n) s += dzr[n] * r(W[(long)n * K + k]);
                da[k] = s;
            }
            dz.assign(K, 0.0);
            for (int k = 0; k < K; ++k) dz[k] = da[k] * gelu_prime(z[l][k]);
        }
    }
    return loss;
}

}  // extern "C"
