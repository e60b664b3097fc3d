// Guiding network kernels: weight packing, inference, fused forward/backward
// training step (BF16 MFMA, FP32 accumulation) and the reduce + Adam update.
//
// Training step, per block of 4 waves and 64 samples (16 per wave):
//   1. forward (guide.h layout: activations stay in the MFMA D/B layout) and
//      copies of x, h1, h2, h3 as bf16 into LDS rows [sample][feature];
//   2. one lane per sample evaluates the vMF mixture at the stored direction
//      and the gradient of -(c/n) log g with respect to the 40 outputs;
//   3. backward-data through W4ᵀ, W3ᵀ, W2ᵀ with MFMA on pre-transposed weight
//      tiles, gated by ReLU'; the gradients dz go to LDS as bf16 rows;
//   4. weight gradients dW = dz · hᵀ over the 64 samples: the 56 dW tiles are
//      split 14 per wave and accumulated in registers across steps; the
//      transposed operands are read straight from the LDS rows.
// Each block then writes its partial gradient; k_guide_adam sums the partials
// and applies Adam to the fp32 master weights, re-packing the bf16 tiles.
#ifndef TPT_CSRC_GUIDE_TRAIN_H
#define TPT_CSRC_GUIDE_TRAIN_H

#include "guide.h"

namespace tptd {

constexpr int G_TT_TILES = 12 + 16 + 16;  // transposed tiles for W4ᵀ, W3ᵀ, W2ᵀ
constexpr int G_PARTIAL = G_NPARAM_PAD + 4;  // + alpha grad, loss sum, count, pad
// LDS row layout for one step: [64 samples][480 features] bf16
constexpr int A_X = 0, A_H1 = 48, A_H2 = 112, A_H3 = 176, A_DO = 240, A_DZ3 = 288, A_DZ2 = 352, A_DZ1 = 416,
              A_W = 480;

// Layer geometry: fp32 offset, rows (out), padded cols (in), M/K tiles, tile base.
__host__ __device__ inline void layer_geom(int L, int *off, int *rows, int *cols, int *mts, int *kts, int *base) {
  const int o[4] = {G_W1, G_W2, G_W3, G_W4}, r[4] = {64, 64, 64, 48}, c[4] = {48, 64, 64, 64};
  const int b[4] = {0, 12, 28, 44};
  *off = o[L], *rows = r[L], *cols = c[L], *mts = r[L] / 16, *kts = c[L] / 16, *base = b[L];
}

// True where a padded-layout parameter is real (not padding).
__host__ __device__ inline bool param_live(int p) {
  if (p < G_W2) return (p % 48) < G_IN;               // W1 cols 42..47 are padding
  if (p < G_W4) return true;
  if (p < G_B1) return (p - G_W4) / 64 < G_OUT;       // W4 rows 40..47
  if (p < G_B4) return true;
  if (p < G_NPARAM_PAD) return (p - G_B4) < G_OUT;    // b4 40..47
  return true;
}

// bf16 tiles from fp32 master: training forward tiles (A = W), transposed
// tiles (A = Wᵀ, for layers 4, 3, 2) and the inference tiles (16x16x32
// layout, permuted K order), one thread per (layer, out, in).  Inference-tile
// slots that hold no weight are never written (they stay zero).
__global__ void k_guide_pack(const float *__restrict__ w, uint16_t *__restrict__ tiles, uint16_t *__restrict__ ttiles,
                             uint16_t *__restrict__ xtiles) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G_B1) return;
  int L = t < G_W2 ? 0 : (t < G_W3 ? 1 : (t < G_W4 ? 2 : 3));
  int off, rows, cols, mts, kts, base;
  layer_geom(L, &off, &rows, &cols, &mts, &kts, &base);
  const int o = (t - off) / cols, i = (t - off) % cols;
  const uint16_t v = to_bf16(param_live(t) ? w[t] : 0.0f);
  // forward: A[o][i], tile (o/16, i/16), lane (o%16) + 16*((i%16)/4), reg i%4
  const int tile = base + (o / 16) * kts + (i / 16);
  tiles[(tile * 64 + (o % 16) + 16 * ((i % 16) / 4)) * 4 + (i % 4)] = v;
  if (L >= 1) {
    // transposed: A[i][o], tiles M = cols/16, K = rows/16; order W4ᵀ, W3ᵀ, W2ᵀ
    const int tb = L == 3 ? 0 : (L == 2 ? 12 : 28);
    const int tkts = rows / 16;
    const int tt = tb + (i / 16) * tkts + (o / 16);
    ttiles[(tt * 64 + (i % 16) + 16 * ((o % 16) / 4)) * 4 + (o % 4)] = v;
  }
  // inference: K slot (s, g, q) of the input column i
  int g, s, q;
  if (L == 0) {
    if (!param_live(t)) return;
    int slot;
    feature_xslot(i, &g, &slot);
    s = slot / 8, q = slot % 8;
  } else {
    const int mt = i / 16;
    g = (i % 16) / 4, s = mt / 2, q = 4 * (mt % 2) + (i % 4);
  }
  const int xt = GX_BASE[L] + (o / 16) * 2 + s;
  xtiles[(xt * 64 + (o % 16) + 16 * g) * 8 + q] = v;
}

__device__ __forceinline__ void guide_stage_weights(uint32_t *dst, const uint32_t *src, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

// Raw outputs for inputs in9 (n x 9) -> out (n x 40), on the inference tiles.
// 256 threads, 64 inputs per block-iteration (16 per wave).
__global__ void __launch_bounds__(256) k_guide_infer(const uint32_t *__restrict__ xtiles, const float *__restrict__ params,
                                                     const float *__restrict__ in, uint64_t n, float *__restrict__ out) {
  __shared__ uint32_t wt[GX_TILES * GX_TILE_WORDS];
  __shared__ float bias[G_BIAS];
  guide_stage_weights(wt, xtiles, GX_TILES * GX_TILE_WORDS);
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) bias[i] = params[G_B1 + i];
  __syncthreads();
  const int l = __lane_id(), wave = threadIdx.x / 64, fg = l >> 4;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * 64; base < n; base += static_cast<uint64_t>(gridDim.x) * 64) {
    const uint64_t s = base + wave * 16 + (l & 15);
    float in9[9];
    for (int k = 0; k < 9; ++k) in9[k] = s < n ? in[s * 9 + k] : 0.0f;
    bf16x8 xb[2];
    guide_xfeatures(fg, fg == 0 ? in9[0] : (fg == 1 ? in9[1] : in9[2]), in9 + 3, xb);
    f32x4 o[3];
    guide_forward_x32(xb, wt, bias, o);
    if (s < n) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int f = 16 * mt + 4 * fg + r;
          if (f < G_OUT) out[s * G_OUT + f] = o[mt][r];
        }
    }
  }
}

// Loss gradient for one sample (one lane): outputs o[40] -> dout[40] (in
// place of o), returns c·(-log g) contribution; adds the mix-weight gradient.
__device__ __forceinline__ float guide_loss_grad(float *o, const float *row, float inv_n, float alpha, float *dalpha) {
  const float c = row[12];
  const float q = row[13];
  const v3 w{row[9], row[10], row[11]};
  if (!(c > 0.0f) || !(c < 3.0e38f)) {
    for (int k = 0; k < G_OUT; ++k) o[k] = 0.0f;
    return 0.0f;
  }
  vmf_mix m;
  decode_mix(o, &m);
  float dv[G_LOBES], cs[G_LOBES];
  float g = 0.0f;
  for (int i = 0; i < G_LOBES; ++i) {
    cs[i] = m.mu[i][0] * w.x + m.mu[i][1] * w.y + m.mu[i][2] * w.z;
    dv[i] = vmf_pdf(m.kappa[i], cs[i]);
    g += m.w[i] * dv[i];
  }
  g = fmaxf(g, 1e-30f);
  const float mm = alpha * g + (1.0f - alpha) * q;
  if (mm > 0.0f) *dalpha += -c * inv_n * (g - q) * alpha * (1.0f - alpha) / mm;
  const float loss = -c * __logf(g);
  const float f = -c * inv_n / g;
  for (int i = 0; i < G_LOBES; ++i) {
    const float a = o[5 * i], b = o[5 * i + 1], cc = o[5 * i + 2];
    const float nrm = sqrtf(a * a + b * b + cc * cc);
    const float kap = m.kappa[i];
    const float lk = o[5 * i + 3];
    const float wd = m.w[i] * dv[i];
    // d/d raw mean: κ (ω − (μ·ω) μ) / |raw|
    const float s = nrm > 1e-10f ? f * wd * kap / nrm : 0.0f;
    o[5 * i] = s * (w.x - cs[i] * m.mu[i][0]);
    o[5 * i + 1] = s * (w.y - cs[i] * m.mu[i][1]);
    o[5 * i + 2] = s * (w.z - cs[i] * m.mu[i][2]);
    // d/d log κ: 1 − 2κ/(e^{2κ} − 1) + κ(μ·ω − 1), zero where the clamp is active
    const bool live = lk > G_LOGK_MIN && lk < G_LOGK_MAX;
    o[5 * i + 3] = live ? f * wd * (1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (cs[i] - 1.0f)) : 0.0f;
    // d/d logit: w_i (D_i − g)
    o[5 * i + 4] = f * m.w[i] * (dv[i] - g);
  }
  return loss;
}

__device__ __forceinline__ bf16x4 lds_col4(const uint16_t *act, int s0, int col) {
  // act[(s0 + r) * A_W + col], r = 0..3
  bf16x4 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = static_cast<short>(act[(s0 + r) * A_W + col]);
  return v;
}

// One backward-data layer: dh[i][s] = Σ_o W[o][i] dz[o][s] via transposed
// tiles (tt_base, M = in/16 = 4 tiles, K = out/16 tiles), gated by ReLU'(h).
__device__ __forceinline__ void guide_bwd_data(const bf16x4 *tt, int tt_base, int ktiles, const bf16x4 dz[4],
                                               const f32x4 h[4], f32x4 out[4]) {
  const int l = __lane_id();
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) out[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) {
    if (kt < ktiles) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        out[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tt[(tt_base + mt * ktiles + kt) * 64 + l], dz[kt], out[mt], 0,
                                                           0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[mt][r] = h[mt][r] > 0.0f ? out[mt][r] : 0.0f;
}

// Adds a D-layout feature block (4 tiles or 3) summed over the 16 samples of
// the wave into LDS bias accumulators db[off + feature].
__device__ __forceinline__ void guide_bias_acc(float *db, int off, const f32x4 *v, int tiles) {
  const int l = __lane_id(), fg = l >> 4;
  for (int mt = 0; mt < tiles; ++mt) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float x = v[mt][r];
      x += __shfl_xor(x, 1);
      x += __shfl_xor(x, 2);
      x += __shfl_xor(x, 4);
      x += __shfl_xor(x, 8);
      if ((l & 15) == 0) atomicAdd(&db[off + 16 * mt + 4 * fg + r], x);
    }
  }
}

__device__ __forceinline__ void guide_store_rows(uint16_t *act, int col0, int tiles, const f32x4 *v) {
  const int l = __lane_id(), fg = l >> 4, s = (threadIdx.x / 64) * 16 + (l & 15);
  for (int mt = 0; mt < tiles; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) act[s * A_W + col0 + 16 * mt + 4 * fg + r] = to_bf16(v[mt][r]);
}

// Fused forward/backward over rows [0, n).  Dynamic LDS: see k_guide_train_lds.
constexpr size_t k_guide_train_lds() {
  return G_TILES * 512 + G_TT_TILES * 512 + G_BIAS * 4 + 64 * A_W * 2 + 4 * 16 * 48 * 4 + G_BIAS * 4 + 64;
}

__global__ void __launch_bounds__(256) k_guide_train(const uint32_t *__restrict__ tiles, const uint32_t *__restrict__ ttiles,
                                                     const float *__restrict__ params, const float *__restrict__ rows,
                                                     uint64_t n, uint64_t row0, uint64_t n_total,
                                                     float *__restrict__ partial) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t *wt = reinterpret_cast<uint32_t *>(lds);
  uint32_t *wtt = wt + G_TILES * G_TILE_WORDS;
  float *bias = reinterpret_cast<float *>(wtt + G_TT_TILES * G_TILE_WORDS);
  uint16_t *act = reinterpret_cast<uint16_t *>(bias + G_BIAS);
  float *outs = reinterpret_cast<float *>(act + 64 * A_W);
  float *db = outs + 4 * 16 * 48;
  float *scal = db + G_BIAS;  // [0] dalpha, [1] loss, [2] count
  guide_stage_weights(wt, tiles, G_TILES * G_TILE_WORDS);
  guide_stage_weights(wtt, ttiles, G_TT_TILES * G_TILE_WORDS);
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) {
    bias[i] = params[G_B1 + i];
    db[i] = 0.0f;
  }
  if (threadIdx.x < 4) scal[threadIdx.x] = 0.0f;
  __syncthreads();
  const int l = __lane_id(), wave = threadIdx.x / 64, fg = l >> 4, sl = l & 15;
  const float inv_n = 1.0f / static_cast<float>(n);
  const float alpha = 1.0f / (1.0f + __expf(-params[G_NPARAM_PAD]));
  const bf16x4 *T = reinterpret_cast<const bf16x4 *>(wt);
  const bf16x4 *TT = reinterpret_cast<const bf16x4 *>(wtt);
  // this wave's 14 dW tiles: global tile index j = 14*wave + t over the list
  // (layer 1: 4x3, layer 2: 4x4, layer 3: 4x4, layer 4: 3x4), accumulated across steps
  f32x4 dw[14];
#pragma unroll
  for (int t = 0; t < 14; ++t) dw[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float dalpha = 0.0f, loss = 0.0f;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * 64; base < n; base += static_cast<uint64_t>(gridDim.x) * 64) {
    const uint64_t sidx = base + wave * 16 + sl;
    const bool valid = sidx < n;
    float row[16];
    const float *rp = rows + ((row0 + sidx) % n_total) * TPT_GUIDE_ROW;
#pragma unroll
    for (int k = 0; k < 16; ++k) row[k] = valid ? rp[k] : 0.0f;
    // ---- forward (keeps activations)
    bf16x4 x[4];
    f32x4 xf[3];
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
      const int f = 16 * kt + 4 * fg;
#pragma unroll
      for (int r = 0; r < 4; ++r) xf[kt][r] = guide_feature(f + r, row);
      x[kt] = pack4(xf[kt][0], xf[kt][1], xf[kt][2], xf[kt][3]);
    }
    guide_store_rows(act, A_X, 3, xf);
    f32x4 h[3][4];
    int tile = 0;
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
      const int kts = layer == 0 ? 3 : 4;
      f32x4 acc[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
        if (kt < kts) {
#pragma unroll
          for (int mt = 0; mt < 4; ++mt)
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(T[(tile + mt * kts + kt) * 64 + l], x[kt], acc[mt], 0, 0, 0);
        }
      tile += 4 * kts;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float z = acc[mt][r] + bias[64 * layer + 16 * mt + 4 * fg + r];
          h[layer][mt][r] = z > 0.0f ? z : 0.0f;
        }
        x[mt] = pack4(h[layer][mt][0], h[layer][mt][1], h[layer][mt][2], h[layer][mt][3]);
      }
      guide_store_rows(act, layer == 0 ? A_H1 : (layer == 1 ? A_H2 : A_H3), 4, h[layer]);
    }
    f32x4 o[3];
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) o[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
      for (int mt = 0; mt < 3; ++mt)
        o[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(T[(tile + mt * 4 + kt) * 64 + l], x[kt], o[mt], 0, 0, 0);
    float *myo = outs + (wave * 16 + sl) * 48;
#pragma unroll
    for (int mt = 0; mt < 3; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) myo[16 * mt + 4 * fg + r] = o[mt][r] + bias[192 + 16 * mt + 4 * fg + r];
    __syncthreads();
    // ---- loss gradient, one lane per sample
    if (fg == 0) {
      if (valid) loss += guide_loss_grad(myo, row, inv_n, alpha, &dalpha);
      else
        for (int k = 0; k < 48; ++k) myo[k] = 0.0f;
      for (int k = G_OUT; k < 48; ++k) myo[k] = 0.0f;
    }
    __syncthreads();
    // ---- backward
    f32x4 dout[3];
    bf16x4 dz[4];
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
#pragma unroll
      for (int r = 0; r < 4; ++r) dout[mt][r] = myo[16 * mt + 4 * fg + r];
      dz[mt] = pack4(dout[mt][0], dout[mt][1], dout[mt][2], dout[mt][3]);
    }
    guide_store_rows(act, A_DO, 3, dout);
    guide_bias_acc(db, 192, dout, 3);
    f32x4 dh[4];
    guide_bwd_data(TT, 0, 3, dz, h[2], dh);  // through W4ᵀ, gated by h3
    guide_store_rows(act, A_DZ3, 4, dh);
    guide_bias_acc(db, 128, dh, 4);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) dz[mt] = pack4(dh[mt][0], dh[mt][1], dh[mt][2], dh[mt][3]);
    guide_bwd_data(TT, 12, 4, dz, h[1], dh);  // through W3ᵀ, gated by h2
    guide_store_rows(act, A_DZ2, 4, dh);
    guide_bias_acc(db, 64, dh, 4);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) dz[mt] = pack4(dh[mt][0], dh[mt][1], dh[mt][2], dh[mt][3]);
    guide_bwd_data(TT, 28, 4, dz, h[0], dh);  // through W2ᵀ, gated by h1
    guide_store_rows(act, A_DZ1, 4, dh);
    guide_bias_acc(db, 0, dh, 4);
    __syncthreads();
    // ---- weight gradients over the block's 64 samples (4 K-steps of 16)
#pragma unroll
    for (int t = 0; t < 14; ++t) {
      const int j = 14 * wave + t;
      int L, mo, ni;
      if (j < 12) L = 0, mo = j / 3, ni = j % 3;
      else if (j < 28) L = 1, mo = (j - 12) / 4, ni = (j - 12) % 4;
      else if (j < 44) L = 2, mo = (j - 28) / 4, ni = (j - 28) % 4;
      else L = 3, mo = (j - 44) / 4, ni = (j - 44) % 4;
      const int dzc = L == 0 ? A_DZ1 : (L == 1 ? A_DZ2 : (L == 2 ? A_DZ3 : A_DO));
      const int inc = L == 0 ? A_X : (L == 1 ? A_H1 : (L == 2 ? A_H2 : A_H3));
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x4 a = lds_col4(act, 16 * ks + 4 * fg, dzc + 16 * mo + sl);
        const bf16x4 b = lds_col4(act, 16 * ks + 4 * fg, inc + 16 * ni + sl);
        dw[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dw[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // ---- partial gradient of this block
  float *P = partial + static_cast<size_t>(blockIdx.x) * G_PARTIAL;
#pragma unroll
  for (int t = 0; t < 14; ++t) {
    const int j = 14 * wave + t;
    int L, mo, ni;
    if (j < 12) L = 0, mo = j / 3, ni = j % 3;
    else if (j < 28) L = 1, mo = (j - 12) / 4, ni = (j - 12) % 4;
    else if (j < 44) L = 2, mo = (j - 28) / 4, ni = (j - 28) % 4;
    else L = 3, mo = (j - 44) / 4, ni = (j - 44) % 4;
    int off, rows_, cols, mts, kts, tb;
    layer_geom(L, &off, &rows_, &cols, &mts, &kts, &tb);
#pragma unroll
    for (int r = 0; r < 4; ++r) P[off + (16 * mo + 4 * fg + r) * cols + 16 * ni + sl] = dw[t][r];
  }
  for (int k = 0; k < 4; ++k) {
    dalpha += __shfl_xor(dalpha, 1 << k);
    loss += __shfl_xor(loss, 1 << k);
  }
  if (l == 0) {
    atomicAdd(&scal[0], dalpha);
    atomicAdd(&scal[1], loss);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) P[G_B1 + i] = db[i];
  if (threadIdx.x == 0) {
    P[G_NPARAM_PAD] = scal[0];
    P[G_NPARAM_PAD + 1] = scal[1];
  }
}

// Sums block partials; mode 0 writes the gradient to grad, mode 1 applies
// Adam (step `t`, learning rate lr) to the master parameters.
__global__ void k_guide_reduce(const float *__restrict__ partial, int blocks, float *__restrict__ params,
                               float *__restrict__ adam_m, float *__restrict__ adam_v, float *__restrict__ grad,
                               float *__restrict__ loss_out, int mode, float lr, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > G_NPARAM_PAD + 1) return;
  float g = 0.0f;
  for (int b = 0; b < blocks; ++b) g += partial[static_cast<size_t>(b) * G_PARTIAL + p];
  if (p == G_NPARAM_PAD + 1) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (!param_live(p)) g = 0.0f;
  if (mode == 0) {
    grad[p] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * adam_m[p] + (1.0f - b1) * g;
  const float v = b2 * adam_v[p] + (1.0f - b2) * g * g;
  adam_m[p] = m;
  adam_v[p] = v;
  const float mh = m / (1.0f - powf(b1, static_cast<float>(step)));
  const float vh = v / (1.0f - powf(b2, static_cast<float>(step)));
  if (param_live(p)) params[p] -= lr * mh / (sqrtf(vh) + eps);
  if (p == G_NPARAM_PAD) params[p] = fminf(fmaxf(params[p], -3.0f), 3.0f);  // mix weight in [0.05, 0.95]
}

}  // namespace tptd

#endif  // TPT_CSRC_GUIDE_TRAIN_H
