// CPU oracle of the path-guiding network (no reference counterpart: the
// reference has no path guiding; this restates the published methods the
// HIP kernels implement, in FP64).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP guiding kernels.
//
// Network: frequency-encoded position (3 -> 36: sin/cos(2^k pi x), k = 0..5)
// and normal (3 -> 12: k = 0..1) -> 4 hidden layers of 64 ReLU units -> 32
// outputs = 8 von Mises-Fisher lobes x (weight logit, concentration, two
// direction parameters).  Lobe i: pi_i = softmax(a)_i, kappa_i =
// softplus(b_i) + 0.01, mu_i = (r cos phi, z, r sin phi) with z = tanh(c_i),
// phi = pi tanh(d_i), r = sqrt(1 - z^2).  vMF density C(k) exp(k (mu.w - 1)),
// C(k) = k / (2 pi (1 - exp(-2k))); sampling after W. Jakob, "Numerically
// stable sampling of the von Mises-Fisher distribution on S^2" (2012).
// Training loss: mean over samples of -w * log q(w_dir) (Mueller et al.,
// "Neural Importance Sampling", ACM TOG 2019, KL divergence form); optimiser:
// Adam (Kingma & Ba 2015).
// Parameters in canonical order: W0[48x64], b0[64], W1[64x64], b1, W2, b2,
// W3, b3, W4[64x32], b4[32] (row-major, input-major), 17,696 values.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kIn = 48, kHid = 64, kOut = 32, kLayers = 5;
constexpr int kLobes = 8;
const int kDimIn[kLayers] = {kIn, kHid, kHid, kHid, kHid};
const int kDimOut[kLayers] = {kHid, kHid, kHid, kHid, kOut};
constexpr double kPiD = 3.14159265358979323846;

int layer_offset(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += kDimIn[i] * kDimOut[i] + kDimOut[i];
    return off;
}

// Round to bfloat16 (round to nearest even), returned as a double.
double bf16(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;   // inf / nan
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

void encode(const float* in6, double* feat) {
    // in6: position normalised to [0,1]^3, unit normal
    int o = 0;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 6; ++k) {
            const double a = std::ldexp(kPiD, k) * (double)in6[c];
            feat[o++] = std::sin(a);
            feat[o++] = std::cos(a);
        }
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 2; ++k) {
            const double a = std::ldexp(kPiD, k) * (double)in6[3 + c];
            feat[o++] = std::sin(a);
            feat[o++] = std::cos(a);
        }
}

struct Acts {
    double x[kLayers + 1][kHid];   // x[0] input features, x[l+1] output of layer l
};

void forward(const float* P, const float* in6, bool emulate, Acts& A) {
    encode(in6, A.x[0]);
    if (emulate)
        for (int k = 0; k < kIn; ++k) A.x[0][k] = bf16(A.x[0][k]);
    for (int l = 0; l < kLayers; ++l) {
        const float* W = P + layer_offset(l);
        const float* b = W + kDimIn[l] * kDimOut[l];
        for (int n = 0; n < kDimOut[l]; ++n) {
            double s = 0.0;
            for (int k = 0; k < kDimIn[l]; ++k) {
                const double w = emulate ? bf16(W[k * kDimOut[l] + n]) : (double)W[k * kDimOut[l] + n];
                s += A.x[l][k] * w;
            }
            s += (double)b[n];
            if (l < kLayers - 1) {
                s = s > 0.0 ? s : 0.0;
                if (emulate) s = bf16(s);
            }
            A.x[l + 1][n] = s;
        }
    }
}

struct Mixture {
    double pi[kLobes], kappa[kLobes], mu[kLobes][3];
    double z[kLobes], phi[kLobes], sig_b[kLobes], dtanh_d[kLobes];
};

double softplus(double x) { return x > 20.0 ? x : std::log1p(std::exp(x)); }

void mixture(const double* out, Mixture& m) {
    double mx = -1e300;
    for (int i = 0; i < kLobes; ++i) mx = std::max(mx, out[4 * i]);
    double sum = 0.0, e[kLobes];
    for (int i = 0; i < kLobes; ++i) { e[i] = std::exp(out[4 * i] - mx); sum += e[i]; }
    for (int i = 0; i < kLobes; ++i) {
        m.pi[i] = e[i] / sum;
        const double b = out[4 * i + 1];
        m.kappa[i] = std::min(softplus(b) + 0.01, 1.0e4);
        m.sig_b[i] = 1.0 / (1.0 + std::exp(-b));
        const double z = std::tanh(out[4 * i + 2]);
        const double td = std::tanh(out[4 * i + 3]);
        const double phi = kPiD * td;
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        m.z[i] = z;
        m.phi[i] = phi;
        m.dtanh_d[i] = 1.0 - td * td;
        m.mu[i][0] = r * std::cos(phi);
        m.mu[i][1] = z;
        m.mu[i][2] = r * std::sin(phi);
    }
}

double vmf_norm(double k) { return k / (2.0 * kPiD * (-std::expm1(-2.0 * k))); }

double lobe_pdf(const Mixture& m, int i, const double* w) {
    const double c = m.mu[i][0] * w[0] + m.mu[i][1] * w[1] + m.mu[i][2] * w[2];
    return vmf_norm(m.kappa[i]) * std::exp(m.kappa[i] * (c - 1.0));
}

double mixture_pdf(const Mixture& m, const double* w) {
    double q = 0.0;
    for (int i = 0; i < kLobes; ++i) q += m.pi[i] * lobe_pdf(m, i, w);
    return q;
}

// d(-weight * log q)/d outputs
void output_grad(const double* out, const double* w, double weight, double* g, double* loss) {
    Mixture m;
    mixture(out, m);
    double v[kLobes];
    double q = 0.0;
    for (int i = 0; i < kLobes; ++i) { v[i] = lobe_pdf(m, i, w); q += m.pi[i] * v[i]; }
    q = std::max(q, 1e-300);
    if (loss) *loss += -weight * std::log(q);
    for (int i = 0; i < kLobes; ++i) {
        const double k = m.kappa[i];
        const double cosw = m.mu[i][0] * w[0] + m.mu[i][1] * w[1] + m.mu[i][2] * w[2];
        // softmax logits
        g[4 * i] = -weight * m.pi[i] * (v[i] - q) / q;
        // concentration: dlogC/dk = 1/k - 2/(e^{2k} - 1)
        const double dlogc = 1.0 / k - 2.0 / std::expm1(2.0 * k);
        const double dv_dk = v[i] * (dlogc + cosw - 1.0);
        const bool clamped = softplus(out[4 * i + 1]) + 0.01 > 1.0e4;
        g[4 * i + 1] = clamped ? 0.0 : -weight * m.pi[i] * dv_dk / q * m.sig_b[i];
        // direction: dv/dmu = v k w
        const double s = -weight * m.pi[i] * v[i] * k / q;
        const double z = m.z[i], phi = m.phi[i];
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        // d mu / dc = d mu/dz * (1 - z^2); d mu/dz = (-z/r cos phi, 1, -z/r sin phi)
        const double dmu_dc[3] = {-z * r * std::cos(phi), 1.0 - z * z, -z * r * std::sin(phi)};
        const double dmu_dd[3] = {-r * std::sin(phi) * kPiD * m.dtanh_d[i], 0.0,
                                  r * std::cos(phi) * kPiD * m.dtanh_d[i]};
        g[4 * i + 2] = s * (dmu_dc[0] * w[0] + dmu_dc[1] * w[1] + dmu_dc[2] * w[2]);
        g[4 * i + 3] = s * (dmu_dd[0] * w[0] + dmu_dd[1] * w[1] + dmu_dd[2] * w[2]);
    }
}

}  // namespace

extern "C" {

int or_mlp_param_count(void) { return layer_offset(kLayers); }

void or_mlp_forward(const float* params, int n, const float* in6, int emulate_bf16, double* out32) {
    Acts A;
    for (int s = 0; s < n; ++s) {
        forward(params, in6 + 6 * s, emulate_bf16 != 0, A);
        for (int o = 0; o < kOut; ++o) out32[kOut * s + o] = A.x[kLayers][o];
    }
}

void or_mlp_encode(const float* in6, double* feat48) { encode(in6, feat48); }

double or_vmf_pdf(const double* out32, const float* dir) {
    Mixture m;
    mixture(out32, m);
    const double w[3] = {dir[0], dir[1], dir[2]};
    return mixture_pdf(m, w);
}

void or_vmf_sample(const double* out32, const float* u3, float* dir, double* pdf) {
    Mixture m;
    mixture(out32, m);
    // lobe by inverse CDF of pi
    int lobe = kLobes - 1;
    double acc = 0.0;
    for (int i = 0; i < kLobes; ++i) {
        acc += m.pi[i];
        if ((double)u3[0] < acc) { lobe = i; break; }
    }
    const double k = m.kappa[lobe];
    const double u1 = u3[1], u2 = u3[2];
    // clamped to -1 where exp(-2k) underflows and u1 = 0 (log of 0)
    const double wz = std::max(1.0 + std::log(u1 + (1.0 - u1) * std::exp(-2.0 * k)) / k, -1.0);
    const double sr = std::sqrt(std::max(0.0, 1.0 - wz * wz));
    const double lx = sr * std::cos(2.0 * kPiD * u2), ly = sr * std::sin(2.0 * kPiD * u2);
    // frame around mu (any orthonormal completion)
    const double* mu = m.mu[lobe];
    double t[3];
    if (std::fabs(mu[0]) > std::fabs(mu[1])) {
        const double inv = 1.0 / std::sqrt(mu[0] * mu[0] + mu[2] * mu[2]);
        t[0] = mu[2] * inv; t[1] = 0.0; t[2] = -mu[0] * inv;
    } else {
        const double inv = 1.0 / std::sqrt(mu[1] * mu[1] + mu[2] * mu[2]);
        t[0] = 0.0; t[1] = mu[2] * inv; t[2] = -mu[1] * inv;
    }
    const double b[3] = {t[1] * mu[2] - t[2] * mu[1], t[2] * mu[0] - t[0] * mu[2], t[0] * mu[1] - t[1] * mu[0]};
    double d[3];
    for (int c = 0; c < 3; ++c) d[c] = lx * t[c] + ly * b[c] + wz * mu[c];
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int c = 0; c < 3; ++c) dir[c] = (float)(d[c] / len);
    const double wd[3] = {dir[0], dir[1], dir[2]};
    *pdf = mixture_pdf(m, wd);
}

// Gradient of mean_s(-weight_s * log q_s(dir_s)) w.r.t. the canonical
// parameters; returns the loss.
double or_mlp_grad(const float* params, int n, const float* in6, const float* dirs,
                   const float* weights, int emulate_bf16, double* grad) {
    const int np = layer_offset(kLayers);
    std::fill(grad, grad + np, 0.0);
    double loss = 0.0;
    Acts A;
    for (int s = 0; s < n; ++s) {
        forward(params, in6 + 6 * s, emulate_bf16 != 0, A);
        const double w[3] = {dirs[3 * s], dirs[3 * s + 1], dirs[3 * s + 2]};
        double dy[kHid];
        double g[kOut];
        output_grad(A.x[kLayers], w, (double)weights[s] / n, g, &loss);
        for (int o = 0; o < kOut; ++o) dy[o] = g[o];
        for (int l = kLayers - 1; l >= 0; --l) {
            if (emulate_bf16)   // the GPU feeds gradients to the matrix cores as bf16
                for (int nn = 0; nn < kDimOut[l]; ++nn) dy[nn] = bf16(dy[nn]);
            const float* W = params + layer_offset(l);
            double* gW = grad + layer_offset(l);
            double* gb = gW + kDimIn[l] * kDimOut[l];
            for (int nn = 0; nn < kDimOut[l]; ++nn) {
                gb[nn] += dy[nn];
                for (int k = 0; k < kDimIn[l]; ++k) gW[k * kDimOut[l] + nn] += A.x[l][k] * dy[nn];
            }
            if (l == 0) break;
            double dx[kHid];
            for (int k = 0; k < kDimIn[l]; ++k) {
                double acc = 0.0;
                for (int nn = 0; nn < kDimOut[l]; ++nn) {
                    const double wv = emulate_bf16 ? bf16(W[k * kDimOut[l] + nn]) : (double)W[k * kDimOut[l] + nn];
                    acc += wv * dy[nn];
                }
                dx[k] = A.x[l][k] > 0.0 ? acc : 0.0;   // ReLU of the previous layer
            }
            std::memcpy(dy, dx, sizeof(double) * kDimIn[l]);
        }
    }
    return loss;
}

// One Adam step (bias-corrected), t = 1-based step count.
void or_adam_step(float* params, float* m, float* v, const float* grad, int n, int t, float lr,
                  float beta1, float beta2, float eps) {
    const double c1 = 1.0 - std::pow((double)beta1, t), c2 = 1.0 - std::pow((double)beta2, t);
    for (int i = 0; i < n; ++i) {
        const double g = grad[i];
        m[i] = (float)(beta1 * (double)m[i] + (1.0 - beta1) * g);
        v[i] = (float)(beta2 * (double)v[i] + (1.0 - beta2) * g * g);
        const double mh = m[i] / c1, vh = v[i] / c2;
        params[i] = (float)((double)params[i] - lr * mh / (std::sqrt(vh) + eps));
    }
}

}  // extern "C"
