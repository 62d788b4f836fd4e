// ORACLE — test infrastructure only (see oracle_trace.cpp).
//
// FP64 CPU statement of the learned guiding proposal (the reference has no
// path guiding; the model is the project's own and is specified in
// yaoray_amd/csrc/yrh_guide.h):
//   grid configuration   resolutions floor(n_min (n_max/n_min)^(l/7) + 0.5),
//                        dense when (res + 1)^3 <= table size
//   encoding             trilinear hash-grid lookup, 8 levels x 2 features
//   network              32 -> 64 -> 64 -> 64 -> 48, ReLU, no biases
//   head                 8 vMF lobes: softmax weights, exp(clamped) kappa,
//                        normalised mean directions
//   loss                 mean_j -(w_j / w_norm) log q(wi_j)
//   gradients            reverse mode by hand, in double
// With `bf16` set, inputs, weights, hidden activations and backward layer
// gradients are rounded to BF16 (nearest even) where the device's MFMA
// kernels round them, so the device can be compared closely; without it the
// model is evaluated in plain double.
// Parity: the double gradients are pinned by central finite differences of
// this file's own loss (tests/test_oracle_guide.py); pdf normalisation and
// sampling against quadrature and histograms.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/yrh.h"

namespace {

constexpr int L = 8;
constexpr int IN = 32;
constexpr int H = 64;
constexpr int OUT = 48;
constexpr int K = 8;
constexpr int W0 = 0, W1 = W0 + IN * H, W2 = W1 + H * H, W3 = W2 + H * H, NMLP = W3 + H * OUT;
constexpr double kPi = 3.14159265358979323846;
constexpr double kLogKappaMin = -3.0, kLogKappaMax = 6.2;

double Bf16(double x) {
    float f = static_cast<float>(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

std::vector<float> Weights(const float* mlp, bool bf16) {
    std::vector<float> w(mlp, mlp + NMLP);
    if (bf16) {
        for (float& v : w) v = static_cast<float>(Bf16(v));
    }
    return w;
}

struct Cfg {
    double bmin[3], inv_extent[3];
    int res[L];
    int dense[L];
    uint32_t table;
};

Cfg MakeCfg(const yrh_guide_desc& d) {
    Cfg c{};
    c.table = 1u << d.log2_table_size;
    for (int a = 0; a < 3; ++a) {
        c.bmin[a] = d.bounds_min[a];
        const double ext = static_cast<double>(d.bounds_max[a]) - static_cast<double>(d.bounds_min[a]);
        c.inv_extent[a] = ext > 0.0 ? 1.0 / ext : 0.0;
    }
    const double ratio = static_cast<double>(d.max_resolution) / static_cast<double>(d.min_resolution);
    for (int l = 0; l < L; ++l) {
        c.res[l] = static_cast<int>(std::floor(d.min_resolution * std::pow(ratio, l / 7.0) + 0.5));
        const double corners = std::pow(static_cast<double>(c.res[l]) + 1.0, 3.0);
        c.dense[l] = corners <= static_cast<double>(c.table) ? 1 : 0;
    }
    return c;
}

uint32_t Index(const Cfg& c, int l, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t r = static_cast<uint32_t>(c.res[l]) + 1u;
    if (c.dense[l]) return x + r * (y + r * z);
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (c.table - 1u);
}

struct Corner {
    uint32_t idx[8];
    double w[8];
};

Corner Locate(const Cfg& c, int l, const float* p) {
    uint32_t base[3];
    double frac[3];
    const double res = c.res[l];
    for (int a = 0; a < 3; ++a) {
        double u = (static_cast<double>(p[a]) - c.bmin[a]) * c.inv_extent[a];
        u = std::clamp(u, 0.0, 1.0);
        const double s = u * res;
        double f = std::floor(s);
        if (f > res - 1.0) f = res - 1.0;
        base[a] = static_cast<uint32_t>(f);
        frac[a] = s - f;
    }
    Corner out{};
    for (int k = 0; k < 8; ++k) {
        out.idx[k] = Index(c, l, base[0] + (k & 1), base[1] + ((k >> 1) & 1), base[2] + ((k >> 2) & 1));
        out.w[k] = ((k & 1) ? frac[0] : 1.0 - frac[0]) * ((k & 2) ? frac[1] : 1.0 - frac[1]) *
                   ((k & 4) ? frac[2] : 1.0 - frac[2]);
    }
    return out;
}

struct Forward {
    double x[IN];
    double h0[H], h1[H], h2[H];
    double o[OUT];
    Corner corners[L];
};

void Dense(const float* w, const double* in, int nin, int nout, double* out, bool relu, bool bf16) {
    for (int j = 0; j < nout; ++j) out[j] = 0.0;
    for (int i = 0; i < nin; ++i) {
        const double a = in[i];
        if (a == 0.0) continue;
        for (int j = 0; j < nout; ++j) out[j] += a * static_cast<double>(w[i * nout + j]);
    }
    if (relu) {
        for (int j = 0; j < nout; ++j) out[j] = out[j] > 0.0 ? (bf16 ? Bf16(out[j]) : out[j]) : 0.0;
    }
}

void RunForward(const Cfg& c, const float* mlp, const float* grid, const float* row, Forward& f, bool bf16 = false) {
    std::fill(f.x, f.x + IN, 0.0);
    for (int l = 0; l < L; ++l) {
        f.corners[l] = Locate(c, l, row);
        const float* table = grid + static_cast<size_t>(l) * c.table * 2;
        for (int k = 0; k < 8; ++k) {
            f.x[2 * l] += f.corners[l].w[k] * table[2 * f.corners[l].idx[k]];
            f.x[2 * l + 1] += f.corners[l].w[k] * table[2 * f.corners[l].idx[k] + 1];
        }
    }
    for (int a = 0; a < 3; ++a) {
        f.x[16 + a] = row[3 + a];
        f.x[19 + a] = row[6 + a];
    }
    f.x[22] = 1.0;
    if (bf16) {
        for (int i = 0; i < IN; ++i) f.x[i] = Bf16(f.x[i]);
    }
    Dense(mlp + W0, f.x, IN, H, f.h0, true, bf16);
    Dense(mlp + W1, f.h0, H, H, f.h1, true, bf16);
    Dense(mlp + W2, f.h1, H, H, f.h2, true, bf16);
    Dense(mlp + W3, f.h2, H, OUT, f.o, false, bf16);
}

struct Head {
    double pi[K], s[K], kappa[K], mu[K][3], len[K];
    bool clamped[K];
};

Head MakeHead(const double* o) {
    Head h{};
    double mx = o[0];
    for (int k = 1; k < K; ++k) mx = std::max(mx, o[5 * k]);
    double sum = 0.0;
    for (int k = 0; k < K; ++k) {
        h.pi[k] = std::exp(o[5 * k] - mx);
        sum += h.pi[k];
    }
    for (int k = 0; k < K; ++k) {
        h.pi[k] /= sum;
        const double s = o[5 * k + 1];
        h.clamped[k] = s < kLogKappaMin || s > kLogKappaMax;
        h.kappa[k] = std::exp(std::clamp(s, kLogKappaMin, kLogKappaMax));
        const double x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        h.len[k] = std::sqrt(x * x + y * y + z * z);
        if (h.len[k] > 1e-6) {
            h.mu[k][0] = x / h.len[k];
            h.mu[k][1] = y / h.len[k];
            h.mu[k][2] = z / h.len[k];
        } else {
            h.mu[k][0] = 0.0;
            h.mu[k][1] = 0.0;
            h.mu[k][2] = 1.0;
        }
    }
    return h;
}

double Vmf(double kappa, double c) { return kappa / (2.0 * kPi * (-std::expm1(-2.0 * kappa))) * std::exp(kappa * (c - 1.0)); }

double PdfHead(const Head& h, const double* w) {
    double q = 0.0;
    for (int k = 0; k < K; ++k) q += h.pi[k] * Vmf(h.kappa[k], h.mu[k][0] * w[0] + h.mu[k][1] * w[1] + h.mu[k][2] * w[2]);
    return q;
}

// d(-c log q)/d(outputs).
void HeadGradient(const double* o, const double* wi, double c, double* d_o, double* loss) {
    const Head h = MakeHead(o);
    double v[K], q = 0.0;
    for (int k = 0; k < K; ++k) {
        const double cm = h.mu[k][0] * wi[0] + h.mu[k][1] * wi[1] + h.mu[k][2] * wi[2];
        v[k] = Vmf(h.kappa[k], cm);
        q += h.pi[k] * v[k];
    }
    *loss += -c * std::log(q);
    std::fill(d_o, d_o + OUT, 0.0);
    for (int k = 0; k < K; ++k) {
        const double r = h.pi[k] * v[k] / q;
        d_o[5 * k] = -c * (r - h.pi[k]);
        const double kap = h.kappa[k];
        const double cm = h.mu[k][0] * wi[0] + h.mu[k][1] * wi[1] + h.mu[k][2] * wi[2];
        if (!h.clamped[k]) {
            const double dlogv_dkappa = 1.0 / kap - 2.0 / std::expm1(2.0 * kap) + (cm - 1.0);
            d_o[5 * k + 1] = -c * r * dlogv_dkappa * kap;
        }
        if (h.len[k] > 1e-6) {
            for (int a = 0; a < 3; ++a) d_o[5 * k + 2 + a] = -c * r * kap * (wi[a] - cm * h.mu[k][a]) / h.len[k];
        }
    }
}

// d_in of a hidden layer is masked by its ReLU and, with bf16, rounded (it is
// the next MFMA operand); the input gradient (no mask) stays unrounded.
void DenseBack(const float* w, const double* in, const double* d_out, int nin, int nout, double* gw, double* d_in,
               const double* in_act_for_mask, bool bf16) {
    for (int i = 0; i < nin; ++i) {
        double acc = 0.0;
        for (int j = 0; j < nout; ++j) {
            gw[i * nout + j] += in[i] * d_out[j];
            acc += d_out[j] * static_cast<double>(w[i * nout + j]);
        }
        if (d_in != nullptr) {
            if (in_act_for_mask == nullptr) {
                d_in[i] = acc;
            } else {
                d_in[i] = in_act_for_mask[i] > 0.0 ? (bf16 ? Bf16(acc) : acc) : 0.0;
            }
        }
    }
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) void oracle_guide_config(const yrh_guide_desc* d, int32_t* res, int32_t* dense) {
    const Cfg c = MakeCfg(*d);
    for (int l = 0; l < L; ++l) {
        res[l] = c.res[l];
        dense[l] = c.dense[l];
    }
}

// Rows of 9 floats (position, normal, outgoing) -> 40 floats (pi, kappa, mu).
__attribute__((visibility("default"))) void oracle_guide_forward(const yrh_guide_desc* d, const float* mlp,
                                                                 const float* grid, int64_t n, const float* rows,
                                                                 int32_t bf16, float* mixture) {
    const Cfg c = MakeCfg(*d);
    const std::vector<float> w = Weights(mlp, bf16 != 0);
    Forward f;
    for (int64_t r = 0; r < n; ++r) {
        RunForward(c, w.data(), grid, rows + 9 * r, f, bf16 != 0);
        const Head h = MakeHead(f.o);
        float* m = mixture + 40 * r;
        for (int k = 0; k < K; ++k) {
            m[k] = static_cast<float>(h.pi[k]);
            m[8 + k] = static_cast<float>(h.kappa[k]);
            for (int a = 0; a < 3; ++a) m[16 + 3 * k + a] = static_cast<float>(h.mu[k][a]);
        }
    }
}

// Rows of 13 floats (position, normal, outgoing, sampled direction, weight).
__attribute__((visibility("default"))) void oracle_guide_gradients(const yrh_guide_desc* d, const float* mlp,
                                                                   const float* grid, int64_t n, const float* rows,
                                                                   double w_norm, int32_t bf16, double* grad_mlp,
                                                                   double* grad_grid, double* loss) {
    const Cfg c = MakeCfg(*d);
    const bool q = bf16 != 0;
    const std::vector<float> wq = Weights(mlp, q);
    mlp = wq.data();
    std::fill(grad_mlp, grad_mlp + NMLP, 0.0);
    std::fill(grad_grid, grad_grid + static_cast<size_t>(L) * c.table * 2, 0.0);
    *loss = 0.0;
    Forward f;
    double d_o[OUT], d2[H], d1[H], d0[H], dx[IN];
    for (int64_t r = 0; r < n; ++r) {
        const float* row = rows + 13 * r;
        RunForward(c, mlp, grid, row, f, q);
        const double wi[3] = {row[9], row[10], row[11]};
        const double cw = static_cast<double>(row[12]) / w_norm / static_cast<double>(n);
        HeadGradient(f.o, wi, cw, d_o, loss);
        if (q) {
            for (int j = 0; j < OUT; ++j) d_o[j] = Bf16(d_o[j]);
        }
        DenseBack(mlp + W3, f.h2, d_o, H, OUT, grad_mlp + W3, d2, f.h2, q);
        DenseBack(mlp + W2, f.h1, d2, H, H, grad_mlp + W2, d1, f.h1, q);
        DenseBack(mlp + W1, f.h0, d1, H, H, grad_mlp + W1, d0, f.h0, q);
        DenseBack(mlp + W0, f.x, d0, IN, H, grad_mlp + W0, dx, nullptr, q);
        for (int l = 0; l < L; ++l) {
            double* g = grad_grid + static_cast<size_t>(l) * c.table * 2;
            for (int k = 0; k < 8; ++k) {
                g[2 * f.corners[l].idx[k]] += f.corners[l].w[k] * dx[2 * l];
                g[2 * f.corners[l].idx[k] + 1] += f.corners[l].w[k] * dx[2 * l + 1];
            }
        }
    }
}

// Loss only (for finite differences of the gradients above).
__attribute__((visibility("default"))) double oracle_guide_loss(const yrh_guide_desc* d, const float* mlp,
                                                                const float* grid, int64_t n, const float* rows,
                                                                double w_norm) {
    const Cfg c = MakeCfg(*d);
    Forward f;
    double loss = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        const float* row = rows + 13 * r;
        RunForward(c, mlp, grid, row, f);
        const Head h = MakeHead(f.o);
        const double wi[3] = {row[9], row[10], row[11]};
        loss += -(static_cast<double>(row[12]) / w_norm / static_cast<double>(n)) * std::log(PdfHead(h, wi));
    }
    return loss;
}

// Mixture rows (40 floats) -> pdf at dirs; samples from u as (dir, pdf).
__attribute__((visibility("default"))) void oracle_guide_mixture_eval(int64_t n, const float* mixture,
                                                                      const float* dirs, const float* u, float* pdf,
                                                                      float* samples) {
    for (int64_t r = 0; r < n; ++r) {
        const float* m = mixture + 40 * r;
        Head h{};
        for (int k = 0; k < K; ++k) {
            h.pi[k] = m[k];
            h.kappa[k] = m[8 + k];
            for (int a = 0; a < 3; ++a) h.mu[k][a] = m[16 + 3 * k + a];
        }
        const double w[3] = {dirs[3 * r], dirs[3 * r + 1], dirs[3 * r + 2]};
        pdf[r] = static_cast<float>(PdfHead(h, w));
        // Lobe by CDF, then Jakob's inversion around the lobe axis.
        const double u0 = u[3 * r], u1 = u[3 * r + 1], u2 = u[3 * r + 2];
        int k = 0;
        double acc = h.pi[0];
        while (k < K - 1 && u0 >= acc) {
            ++k;
            acc += h.pi[k];
        }
        const double kap = h.kappa[k];
        double ct = 1.0 + std::log(std::max(u1 + (1.0 - u1) * std::exp(-2.0 * kap), 1e-30)) / kap;
        ct = std::clamp(ct, -1.0, 1.0);
        const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        const double phi = 2.0 * kPi * u2;
        // Any orthonormal frame around mu gives the same distribution; this is
        // the frame of Duff et al. 2017, as the device uses.
        const double nx = h.mu[k][0], ny = h.mu[k][1], nz = h.mu[k][2];
        const double sign = nz >= 0.0 ? 1.0 : -1.0;
        const double a = -1.0 / (sign + nz);
        const double b = nx * ny * a;
        const double t[3] = {1.0 + sign * nx * nx * a, sign * b, -sign * nx};
        const double bt[3] = {b, sign + ny * ny * a, -ny};
        double o[3];
        for (int i = 0; i < 3; ++i) o[i] = t[i] * st * std::cos(phi) + bt[i] * st * std::sin(phi) + h.mu[k][i] * ct;
        const double len = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        for (int i = 0; i < 3; ++i) o[i] /= len;
        for (int i = 0; i < 3; ++i) samples[4 * r + i] = static_cast<float>(o[i]);
        samples[4 * r + 3] = static_cast<float>(PdfHead(h, o));
    }
}

}  // extern "C"
