// FP64 restatement of the learned pruner's network (TEST INFRASTRUCTURE ONLY).
//
// The network has no counterpart in the reference; this is the textbook
// forward pass and backpropagation of an MLP 64 -> 128 -> 128 -> 1 with ReLU
// hidden layers and a logistic output, and the weighted binary cross-entropy
// sum_i w_i * (softplus(z_i) - y_i z_i) * inv_wsum.  Parameters are packed as
// W1[128][64] b1[128] W2[128][128] b2[128] w3[128] b3, given as float32 (the
// engine's bf16 weights widen to float32 exactly) and used in double.  The
// tests pin these routines against central finite differences and an
// independent float64 autograd, and check the engine's kernels against them.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {
constexpr int kIn = 64, kH = 128;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kParams = kB3 + 1;

void forward(const float *P, const float *x, double *h1, double *h2, double &z) {
    for (int j = 0; j < kH; ++j) {
        double a = P[kB1 + j];
        for (int i = 0; i < kIn; ++i) a += (double) P[kW1 + j * kIn + i] * (double) x[i];
        h1[j] = a > 0.0 ? a : 0.0;
    }
    for (int j = 0; j < kH; ++j) {
        double a = P[kB2 + j];
        for (int i = 0; i < kH; ++i) a += (double) P[kW2 + j * kH + i] * h1[i];
        h2[j] = a > 0.0 ? a : 0.0;
    }
    double a = P[kB3];
    for (int i = 0; i < kH; ++i) a += (double) P[kW3 + i] * h2[i];
    z = a;
}
}  // namespace

extern "C" {

int oracle_mlp_params(void) { return kParams; }

// logits[i] of samples X[i][0..64)
void oracle_mlp_forward(const float *P, const float *X, long long N, double *logits) {
    std::vector<double> h1(kH), h2(kH);
    for (long long s = 0; s < N; ++s) forward(P, X + s * kIn, h1.data(), h2.data(), logits[s]);
}

// grad[0..kParams) = gradient, grad[kParams] = value of the weighted loss
void oracle_mlp_grad(const float *P, const float *X, const float *Y, const float *WT, long long N, double inv_wsum,
                     double *grad) {
    for (int i = 0; i <= kParams; ++i) grad[i] = 0.0;
    std::vector<double> h1(kH), h2(kH), d1(kH), d2(kH);
    for (long long s = 0; s < N; ++s) {
        const float *x = X + s * kIn;
        double z;
        forward(P, x, h1.data(), h2.data(), z);
        const double w = (double) WT[s] * inv_wsum, y = Y[s];
        const double p = 1.0 / (1.0 + std::exp(-z));
        grad[kParams] += w * (std::fmax(z, 0.0) + std::log1p(std::exp(-std::fabs(z))) - y * z);
        const double dz = w * (p - y);
        grad[kB3] += dz;
        for (int j = 0; j < kH; ++j) {
            grad[kW3 + j] += dz * h2[j];
            d2[j] = h2[j] > 0.0 ? dz * (double) P[kW3 + j] : 0.0;
        }
        for (int j = 0; j < kH; ++j) {
            grad[kB2 + j] += d2[j];
            for (int i = 0; i < kH; ++i) grad[kW2 + j * kH + i] += d2[j] * h1[i];
        }
        for (int i = 0; i < kH; ++i) {
            double a = 0.0;
            for (int j = 0; j < kH; ++j) a += d2[j] * (double) P[kW2 + j * kH + i];
            d1[i] = h1[i] > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < kH; ++j) {
            grad[kB1 + j] += d1[j];
            for (int i = 0; i < kIn; ++i) grad[kW1 + j * kIn + i] += d1[j] * (double) x[i];
        }
    }
}

}  // extern "C"
