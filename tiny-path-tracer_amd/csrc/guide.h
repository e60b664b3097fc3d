// Learned path-guiding proposal: network layout, frequency encoding, the
// 8-lobe von Mises–Fisher mixture, and the BF16 MFMA forward pass.
//
// Network (no counterpart in the reference; it is an addition):
//   x = [sin/cos(2^b π p̂_c) for c in xyz, b in 0..5] (36) ++ n (3) ++ wo (3)   -> 42 inputs
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)      64 wide
//   o  = W4 h3 + b4                                                        40 outputs
//   lobe i: mean direction normalize(o[5i..5i+2]), κ = exp(clamp(o[5i+3])),
//   weight softmax(o[5i+4] over i).
// p̂ is the hit point mapped to [0,1]^3 by the scene box; n the shading normal;
// wo = -unit(ray direction).  13,672 parameters; 26,880 FLOP per inference.
//
// MFMA mapping (v_mfma_f32_16x16x16bf16_1k; lane layout verified on gfx950):
// layer outputs are computed transposed, Yᵀ[feature][sample] = W·Xᵀ, with
// A = W (16x16 tiles, one 8-byte LDS read per lane) and B = Xᵀ.  The D tile
// of one layer (lane l, register r = feature 4·(l/16)+r of sample l%16) is
// then exactly the B operand of the next layer after bias, ReLU and the
// bf16 conversion, so activations never leave registers.  Inputs are padded
// to 48 (3 K-tiles) and outputs to 48 (3 M-tiles).
#ifndef TPT_CSRC_GUIDE_H
#define TPT_CSRC_GUIDE_H

#include "dev_common.h"

namespace tptd {

constexpr int G_IN = 42, G_IN_PAD = 48, G_HID = 64, G_OUT = 40, G_OUT_PAD = 48, G_LOBES = 8;
constexpr int G_BIAS = G_HID * 3 + G_OUT_PAD;                          // 240 floats
// fp32 master parameters, padded: W1 64x48, W2 64x64, W3 64x64, W4 48x64, then biases.
constexpr int G_W1 = 0, G_W2 = G_W1 + 64 * 48, G_W3 = G_W2 + 64 * 64, G_W4 = G_W3 + 64 * 64;
constexpr int G_B1 = G_W4 + 48 * 64, G_B2 = G_B1 + 64, G_B3 = G_B2 + 64, G_B4 = G_B3 + 64;
constexpr int G_NPARAM_PAD = G_B4 + 48;  // 14,576 including padding
constexpr float G_LOGK_MIN = -3.0f, G_LOGK_MAX = 7.0f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  // round to nearest even (inputs are finite here)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Operand packing uses the float -> __bf16 conversion, which gfx950 does in
// hardware two values at a time (v_cvt_pk_bf16_f32, round to nearest even:
// the same bits as to_bf16 for every finite value).
__device__ __forceinline__ short bf16_bits(float f) { return __builtin_bit_cast(short, static_cast<__bf16>(f)); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = bf16_bits(a);
  v[1] = bf16_bits(b);
  v[2] = bf16_bits(c);
  v[3] = bf16_bits(d);
  return v;
}

// Frequency band b of coordinate p: sin and cos of 2^b π p, from sinpi/cospi
// of p by b angle doublings, sin 2x = (2 sin x) cos x, cos 2x = 1 - (2 sin x) sin x
// (each a single rounding in this order).
__device__ __forceinline__ void enc_trig(float p, int b, float *sn, float *cs) {
  float s = sinpif(p), c = cospif(p);
  for (int k = 0; k < b; ++k) {
    const float s2 = 2.0f * s;
    const float ns = s2 * c;
    c = 1.0f - s2 * s;
    s = ns;
  }
  *sn = s;
  *cs = c;
}

// Feature f (0..47) of an input (p̂, n, wo): f = 12 c + 2 b + e for the bands
// (e = 0 sin, 1 cos of 2^b π p̂_c), then n (36..38), wo (39..41), zeros.
__device__ __forceinline__ float guide_feature(int f, const float *in9) {
  if (f < 36) {
    float sn, cs;
    enc_trig(in9[f / 12], (f % 12) >> 1, &sn, &cs);
    return (f & 1) ? cs : sn;
  }
  if (f < 42) return in9[3 + (f - 36)];
  return 0.0f;
}

// ------------------------------------------- inference layout (16x16x32) ---
// Inference uses v_mfma_f32_16x16x32_bf16 (gfx950): lane l holds 8 bf16 of
// A row l%16 and of B column l%16, for K slots (l/16, q), q = 0..7; D is laid
// out as in the 16x16x16 form (lane l, register r = row 4(l/16) + r, column
// l%16).  The K order of a layer is free (it is a sum), so each layer's
// weight columns are stored permuted such that
//  - hidden layers: K step s, lane group g, register q reads neuron
//    16(2s + q/4) + 4g + q%4, which is register q%4 of D tile 2s + q/4 of the
//    same lane: the previous layer's D tiles are the next B operand as they are;
//  - first layer: lane group g computes its own 16 features (xslot_feature),
//    one coordinate's 12 bands from one sinpi/cospi, so nothing is exchanged.
// Tiles: W1 4 M x 2 K, W2 and W3 4 x 2, W4 3 x 2: 30 tiles of 1 KB.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int GX_TILES = 30, GX_TILE_WORDS = 256;
constexpr int GX_BASE[4] = {0, 8, 16, 24};

// Feature index held by first-layer K slot t (0..15) of lane group g, or -1.
__host__ __device__ inline int xslot_feature(int g, int t) {
  if (g < 3 && t < 12) return 12 * g + t;
  if (g == 0 && t >= 12) return 36 + (t - 12);  // n, wo.x
  if (g == 1 && t < 14) return 40 + (t - 12);   // wo.y, wo.z
  return -1;
}
// Inverse: first-layer K slot (g, t) of feature f < G_IN.
__host__ __device__ inline void feature_xslot(int f, int *g, int *t) {
  if (f < 36) *g = f / 12, *t = f % 12;
  else if (f < 40) *g = 0, *t = 12 + (f - 36);
  else *g = 1, *t = 12 + (f - 40);
}

__device__ __forceinline__ bf16x8 pack8(f32x4 a, f32x4 b) {
  bf16x8 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v[r] = static_cast<__bf16>(a[r]);
    v[4 + r] = static_cast<__bf16>(b[r]);
  }
  return v;
}

// First-layer B operands of lane group g for one sample: p = its coordinate g
// (g < 3), nwo = (n, wo).
__device__ __forceinline__ void guide_xfeatures(int g, float p, const float *nwo, bf16x8 xb[2]) {
  float v[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) v[t] = 0.0f;
  if (g < 3) {
    float s = sinpif(p), c = cospif(p);
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      v[2 * b] = s;
      v[2 * b + 1] = c;
      const float s2 = 2.0f * s;
      const float ns = s2 * c;
      c = 1.0f - s2 * s;
      s = ns;
    }
    if (g == 0) v[12] = nwo[0], v[13] = nwo[1], v[14] = nwo[2], v[15] = nwo[3];
    if (g == 1) v[12] = nwo[4], v[13] = nwo[5];
  }
  xb[0] = pack8(f32x4{v[0], v[1], v[2], v[3]}, f32x4{v[4], v[5], v[6], v[7]});
  xb[1] = pack8(f32x4{v[8], v[9], v[10], v[11]}, f32x4{v[12], v[13], v[14], v[15]});
}

// ReLU of 8 packed bf16 values: as 16-bit integers a negative bf16 (sign bit
// set, -0 included) is below 0 and a positive one orders like its value, so an
// integer max with 0 is the ReLU (v_pk_max_i16, two values per instruction).
// Rounding to bf16 keeps signs and order, so ReLU after the conversion gives
// the same bits as before it.
typedef short s16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 relu_bf16x8(bf16x8 v) {
  const s16x8 x = __builtin_bit_cast(s16x8, v);
  return __builtin_bit_cast(bf16x8, __builtin_elementwise_max(x, s16x8{0, 0, 0, 0, 0, 0, 0, 0}));
}

// Forward pass of 16 samples (one per lane column l%16) in the 16x16x32
// layout: xb = this lane's first-layer B operands, xt = LDS inference tiles,
// bias = LDS biases.  out[mt][r] = output 16 mt + 4 (l/16) + r.  The
// accumulators start at the biases (no separate bias add).
__device__ __forceinline__ void guide_forward_x32(const bf16x8 xb_in[2], const uint32_t *xt, const float *bias,
                                                  f32x4 out[3]) {
  const int l = __lane_id();
  const int fg = l >> 4;
  const bf16x8 *tiles = reinterpret_cast<const bf16x8 *>(xt);
  const f32x4 *bias4 = reinterpret_cast<const f32x4 *>(bias);  // bias4[4 t + fg]: rows 16 t + 4 fg .. + 3
  bf16x8 xb[2] = {xb_in[0], xb_in[1]};
#pragma unroll
  for (int layer = 0; layer < 3; ++layer) {
    f32x4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      acc[mt] = bias4[4 * (4 * layer + mt) + fg];
#pragma unroll
      for (int s = 0; s < 2; ++s)
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tiles[(GX_BASE[layer] + 2 * mt + s) * 64 + l], xb[s],
                                                          acc[mt], 0, 0, 0);
    }
    xb[0] = relu_bf16x8(pack8(acc[0], acc[1]));
    xb[1] = relu_bf16x8(pack8(acc[2], acc[3]));
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int mt = 0; mt < 3; ++mt) {
    out[mt] = bias4[4 * (12 + mt) + fg];
#pragma unroll
    for (int s = 0; s < 2; ++s)
      out[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tiles[(GX_BASE[3] + 2 * mt + s) * 64 + l], xb[s], out[mt], 0,
                                                        0, 0);
  }
}

// ---------------------------------------------------------------- vMF mix ---
struct vmf_mix {
  float mu[G_LOBES][3];
  float kappa[G_LOBES];
  float w[G_LOBES];
  float wn[G_LOBES];  // w_i times the lobe's normaliser (mix_norms), for mix_pdf
};

// Decodes 40 raw outputs.
__device__ __forceinline__ void decode_mix(const float *o, vmf_mix *m) {
  float mx = -3.0e38f;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) mx = fmaxf(mx, o[5 * i + 4]);
  float z = 0.0f;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) {
    const float a = o[5 * i], b = o[5 * i + 1], c = o[5 * i + 2];
    const float n2 = a * a + b * b + c * c;
    if (n2 > 1e-20f) {
      const float inv = rsqrtf(n2);
      m->mu[i][0] = a * inv, m->mu[i][1] = b * inv, m->mu[i][2] = c * inv;
    } else {
      m->mu[i][0] = 0.0f, m->mu[i][1] = 0.0f, m->mu[i][2] = 1.0f;
    }
    const float lk = fminf(fmaxf(o[5 * i + 3], G_LOGK_MIN), G_LOGK_MAX);
    m->kappa[i] = __expf(lk);
    m->w[i] = __expf(o[5 * i + 4] - mx);
    z += m->w[i];
  }
  const float iz = 1.0f / z;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) m->w[i] *= iz;
}

// decode_mix from 40 bf16 outputs (an LDS row), reading each output where it
// is used; false if an output is not finite.
__device__ __forceinline__ bool decode_mix_bf16(const uint16_t *o, vmf_mix *m) {
  float lg[G_LOBES];
  float mx = -3.0e38f;
  bool finite = true;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) {
    lg[i] = from_bf16(o[5 * i + 4]);
    finite = finite && fabsf(lg[i]) < 3.0e38f;
    mx = fmaxf(mx, lg[i]);
  }
  float z = 0.0f;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) {
    const float a = from_bf16(o[5 * i]), b = from_bf16(o[5 * i + 1]), c = from_bf16(o[5 * i + 2]);
    const float lk0 = from_bf16(o[5 * i + 3]);
    finite = finite && fabsf(a) < 3.0e38f && fabsf(b) < 3.0e38f && fabsf(c) < 3.0e38f && fabsf(lk0) < 3.0e38f;
    const float n2 = a * a + b * b + c * c;
    if (n2 > 1e-20f) {
      const float inv = rsqrtf(n2);
      m->mu[i][0] = a * inv, m->mu[i][1] = b * inv, m->mu[i][2] = c * inv;
    } else {
      m->mu[i][0] = 0.0f, m->mu[i][1] = 0.0f, m->mu[i][2] = 1.0f;
    }
    const float lk = fminf(fmaxf(lk0, G_LOGK_MIN), G_LOGK_MAX);
    m->kappa[i] = __expf(lk);
    m->w[i] = __expf(lg[i] - mx);
    z += m->w[i];
  }
  const float iz = 1.0f / z;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) m->w[i] *= iz;
  return finite;
}

// vMF density exp(κ(μ·ω − 1)) · κ / (2π (1 − e^{−2κ})) on the unit sphere
// (κ ≥ e^{−3}, the log-κ clamp: 1 − e^{−2κ} ≥ 0.095 loses nothing to the
// plain difference).
__device__ __forceinline__ float vmf_pdf(float kappa, float cosang) {
  const float norm = __fdividef(kappa, 2.0f * kPiF * (1.0f - __expf(-2.0f * kappa)));
  return norm * __expf(kappa * (cosang - 1.0f));
}

// Fills wn: w_i κ_i / (2π (1 − e^{−2κ_i})).  κ ≥ e^{−3} (the log-κ clamp), so
// 1 − e^{−2κ} ≥ 0.095 and the plain difference loses nothing that matters.
__device__ __forceinline__ void mix_norms(vmf_mix *m) {
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i)
    m->wn[i] = m->w[i] * __fdividef(m->kappa[i], 2.0f * kPiF * (1.0f - __expf(-2.0f * m->kappa[i])));
}

// Mixture density at unit direction w (needs mix_norms).
__device__ __forceinline__ float mix_pdf(const vmf_mix &m, v3 w) {
  float g = 0.0f;
#pragma unroll
  for (int i = 0; i < G_LOBES; ++i) {
    const float c = m.mu[i][0] * w.x + m.mu[i][1] * w.y + m.mu[i][2] * w.z;
    g += m.wn[i] * __expf(m.kappa[i] * (c - 1.0f));
  }
  return g;
}

// Samples a unit direction: lobe by weight (u0), then the vMF inversion
// (Wenzel Jakob, "Numerically stable sampling of the von Mises Fisher
// distribution on S2", 2012): cosθ = 1 + log(u1 + (1 − u1) e^{−2κ}) / κ.
__device__ __forceinline__ v3 mix_sample(const vmf_mix &m, float u0, float u1, float u2) {
  // lobe by the running sum of weights (the last lobe takes any rounding
  // remainder); selected without dynamic indexing so the mixture stays in
  // registers
  float acc = 0.0f, kap = m.kappa[G_LOBES - 1];
  v3 mu{m.mu[G_LOBES - 1][0], m.mu[G_LOBES - 1][1], m.mu[G_LOBES - 1][2]};
  bool found = false;
#pragma unroll
  for (int i = 0; i < G_LOBES - 1; ++i) {
    acc += m.w[i];
    const bool take = !found && u0 < acc;
    kap = take ? m.kappa[i] : kap;
    mu.x = take ? m.mu[i][0] : mu.x;
    mu.y = take ? m.mu[i][1] : mu.y;
    mu.z = take ? m.mu[i][2] : mu.z;
    found = found || take;
  }
  const float ct = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * kap)) / kap;
  const float c = fminf(fmaxf(ct, -1.0f), 1.0f);
  const float st = sqrtf(fmaxf(0.0f, 1.0f - c * c));
  float s, co;
  __sincosf(2.0f * kPiF * u2, &s, &co);
  const v3 helper = fabsf(mu.x) > 0.9f ? v3{0.0f, 1.0f, 0.0f} : v3{1.0f, 0.0f, 0.0f};
  const v3 t1 = unit(cross(mu, helper));
  const v3 t2 = cross(mu, t1);
  return (st * co) * t1 + (st * s) * t2 + c * mu;
}

}  // namespace tptd

#endif  // TPT_CSRC_GUIDE_H
