// ORACLE — test infrastructure only (see oracle.h).
//
// The learned guiding proposal has no counterpart in the reference; this file
// restates what the device's guided shading computes, for same-stream
// comparison with tpt_render_guided:
//   - the network's forward pass with its operands rounded to bf16 where the
//     device feeds them to MFMA (as oracle/guide_oracle.py forward(emulate=True),
//     which finite differences and the device's own gradients pin), the outputs
//     then rounded to bf16 (the device keeps them so between its inference and
//     its shading);
//   - the 8-lobe von Mises-Fisher mixture: decode, density, and sampling by
//     inversion (Jakob 2012, "Numerically stable sampling of the von
//     Mises-Fisher distribution on S2").
// Parameters: the padded fp32 layout of include/tpt.h (W1 64x48, W2 64x64,
// W3 64x64, W4 48x64, b1, b2, b3, b4 (48), then the mix logit).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle.h"

namespace oracle {

namespace {

constexpr int kW1 = 0, kW2 = 64 * 48, kW3 = kW2 + 64 * 64, kW4 = kW3 + 64 * 64;
constexpr int kB1 = kW4 + 48 * 64, kB2 = kB1 + 64, kB3 = kB2 + 64, kB4 = kB3 + 64, kNParam = kB4 + 48;
constexpr float kLogKMin = -3.0f, kLogKMax = 7.0f;

// Dense layer on bf16 operands with a float64 sum: out[o] = b[o] + sum_i W[o][i] x[i].
void layer(const float *P, int w_off, int b_off, int n_out, int n_in, int stride, const double *x, double *out) {
  for (int o = 0; o < n_out; ++o) {
    double acc = P[b_off + o];
    for (int i = 0; i < n_in; ++i) acc += static_cast<double>(bf16_round(P[w_off + o * stride + i])) * x[i];
    out[o] = acc;
  }
}

}  // namespace

float bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

void guide_forward(const float *P, const float in9[9], double out[40]) {
  // Encoding: band 0 is sin/cos(pi p) rounded to float; band b+1 follows by
  // the float doublings sin 2x = (2 sin x) cos x, cos 2x = 1 - (2 sin x) sin x.
  double x[48] = {0};
  for (int c = 0; c < 3; ++c) {
    float s = static_cast<float>(std::sin(M_PI * static_cast<double>(in9[c])));
    float co = static_cast<float>(std::cos(M_PI * static_cast<double>(in9[c])));
    for (int b = 0; b < 6; ++b) {
      x[12 * c + 2 * b] = s;
      x[12 * c + 2 * b + 1] = co;
      const float s2 = 2.0f * s;
      const float ns = s2 * co;
      co = 1.0f - s2 * s;
      s = ns;
    }
  }
  for (int k = 0; k < 6; ++k) x[36 + k] = in9[3 + k];
  for (double &v : x) v = bf16_round(static_cast<float>(v));
  double h1[64], h2[64], h3[64];
  layer(P, kW1, kB1, 64, 48, 48, x, h1);
  for (double &v : h1) v = bf16_round(static_cast<float>(v > 0 ? v : 0.0));
  layer(P, kW2, kB2, 64, 64, 64, h1, h2);
  for (double &v : h2) v = bf16_round(static_cast<float>(v > 0 ? v : 0.0));
  layer(P, kW3, kB3, 64, 64, 64, h2, h3);
  for (double &v : h3) v = bf16_round(static_cast<float>(v > 0 ? v : 0.0));
  layer(P, kW4, kB4, 40, 64, 64, h3, out);
}

float guide_alpha(const float *P) { return 1.0f / (1.0f + std::exp(-P[kNParam])); }

bool guide_decode(const float *P, const float in9[9], guide_mix *m) {
  double raw[40];
  guide_forward(P, in9, raw);
  float o[40];
  bool finite = true;
  for (int k = 0; k < 40; ++k) {
    o[k] = bf16_round(static_cast<float>(raw[k]));
    finite = finite && std::fabs(o[k]) < 3.0e38f;
  }
  float mx = -3.0e38f;
  for (int i = 0; i < 8; ++i) mx = std::fmax(mx, o[5 * i + 4]);
  float z = 0.0f;
  for (int i = 0; i < 8; ++i) {
    const float a = o[5 * i], b = o[5 * i + 1], c = o[5 * i + 2];
    const float n2 = a * a + b * b + c * c;
    if (n2 > 1e-20f) {
      const float inv = 1.0f / std::sqrt(n2);
      m->mu[i][0] = a * inv, m->mu[i][1] = b * inv, m->mu[i][2] = c * inv;
    } else {
      m->mu[i][0] = 0.0f, m->mu[i][1] = 0.0f, m->mu[i][2] = 1.0f;
    }
    const float lk = std::fmin(std::fmax(o[5 * i + 3], kLogKMin), kLogKMax);
    m->kappa[i] = std::exp(lk);
    m->w[i] = std::exp(o[5 * i + 4] - mx);
    z += m->w[i];
  }
  const float iz = 1.0f / z;
  for (int i = 0; i < 8; ++i) m->w[i] *= iz;
  // w_i times the lobe's normaliser kappa / (2 pi (1 - e^{-2 kappa}))
  for (int i = 0; i < 8; ++i)
    m->wn[i] = m->w[i] * (m->kappa[i] / (2.0f * static_cast<float>(M_PI) * (1.0f - std::exp(-2.0f * m->kappa[i]))));
  return finite;
}

float guide_mix_pdf(const guide_mix &m, const vec3 &w) {
  float g = 0.0f;
  for (int i = 0; i < 8; ++i) {
    const float c = m.mu[i][0] * w.x() + m.mu[i][1] * w.y() + m.mu[i][2] * w.z();
    g += m.wn[i] * std::exp(m.kappa[i] * (c - 1.0f));
  }
  return g;
}

vec3 guide_mix_sample(const guide_mix &m, float u0, float u1, float u2) {
  // lobe by the running sum of the weights; the last lobe takes the remainder
  int lobe = 7;
  float acc = 0.0f;
  for (int i = 0; i < 7; ++i) {
    acc += m.w[i];
    if (u0 < acc) {
      lobe = i;
      break;
    }
  }
  const float kap = m.kappa[lobe];
  const vec3 mu(m.mu[lobe][0], m.mu[lobe][1], m.mu[lobe][2]);
  const float ct = 1.0f + std::log(u1 + (1.0f - u1) * std::exp(-2.0f * kap)) / kap;
  const float c = std::fmin(std::fmax(ct, -1.0f), 1.0f);
  const float st = std::sqrt(std::fmax(0.0f, 1.0f - c * c));
  const float phi = 2.0f * static_cast<float>(M_PI) * u2;
  const float s = std::sin(phi), co = std::cos(phi);
  const vec3 helper = std::fabs(mu.x()) > 0.9f ? vec3(0, 1, 0) : vec3(1, 0, 0);
  const vec3 t1 = unit_vector(cross(mu, helper));
  const vec3 t2 = cross(mu, t1);
  return (st * co) * t1 + (st * s) * t2 + c * mu;
}

}  // namespace oracle
