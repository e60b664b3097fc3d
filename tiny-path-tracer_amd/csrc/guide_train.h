// Guiding network kernels: weight packing, inference, the fused
// forward/backward training step (BF16 MFMA, FP32 accumulation) and the
// gradient reduction + Adam update.
//
// Training step (k_guide_train).  Each wave works on 16 samples at a time
// (lane l: sample l%16, lane group g = l/16):
//   1. forward in the inference layout (16x16x32, permuted K order, guide.h):
//      the D tiles, "sample-major" (a lane holds neurons 16mt + 4g + r of one
//      sample), feed the next layer directly.  The output rows are permuted so
//      that lane group g holds lobes 2g and 2g+1 of its sample;
//   2. the loss gradient: each lane evaluates its two lobes; the softmax and
//      the mixture density are completed across the sample's four lanes by
//      shuffles, so all 64 lanes work;
//   3. backward data, sample-major, through W4ᵀ and W3ᵀ (transposed tiles in
//      the same permuted layout) to chain the deltas;
//   4. the weight gradients need "neuron-major" operands (a lane holds one
//      neuron of 4 samples).  Swapping the MFMA operands (B·A instead of A·B)
//      gives exactly that layout from the same registers: activations from
//      (activation, W) products, deltas from (delta, Wᵀ) products, and the
//      inputs and output deltas from products with identity tiles;
//   5. dW += δ·hᵀ with v_mfma_f32_16x16x16_bf16.  The 56 dW tiles are split
//      over the block's 8 waves (7 each, at most 2 per layer): each wave
//      publishes a layer's neuron-major deltas and activations in LDS, and
//      after one block barrier every wave adds the products of all 8 waves'
//      samples into the tiles it owns.  28 accumulator registers per wave
//      instead of 224 let two waves share each SIMD (512-thread blocks, one
//      66 KB copy of the weight tiles per CU); the exchange is double
//      buffered, so one barrier per layer suffices.
// Each wave then writes its own tiles into the block partial (every tile has
// one owner); k_guide_reduce adds the partials into an accumulator and
// k_guide_adam turns it into the gradient or the Adam update (and clears it).
#ifndef TPT_CSRC_GUIDE_TRAIN_H
#define TPT_CSRC_GUIDE_TRAIN_H

#include "guide.h"

namespace tptd {

constexpr int G_PARTIAL = G_NPARAM_PAD + 4;  // + mix-logit gradient, loss sum, 2 pad
// Training tile set (1 KB tiles, 16x16x32 layout): forward W1..W4 (W4 rows in
// lobe order), transposed W4ᵀ, W3ᵀ, W2ᵀ, and the identity tiles.
constexpr int GR_FWD = 30, GR_BWD = 24, GR_ID = 12, GR_TILES = GR_FWD + GR_BWD + GR_ID;
constexpr int GR_BWD_BASE = GR_FWD;           // W4ᵀ, then W3ᵀ, then W2ᵀ: 8 tiles each
constexpr int GR_IDX_BASE = GR_FWD + GR_BWD;  // inputs: 3 column tiles x 2 K steps
constexpr int GR_IDO_BASE = GR_IDX_BASE + 6;  // outputs: 3 column tiles x 2 K steps

// Output row order of the training forward: physical row p = 16 mt + 4 g + r
// holds output u = 4 mt + r (u < 10) of lane group g, i.e. logical output
// 10 g + u (lobes 2g, 2g+1); rows with u >= 10 are padding.
__host__ __device__ inline int out_phys(int o) {
  const int g = o / 10, u = o % 10;
  return 16 * (u / 4) + 4 * g + (u % 4);
}
__host__ __device__ inline int out_logical(int p) {
  const int u = 4 * (p / 16) + (p % 4);
  return u < 10 ? 10 * ((p % 16) / 4) + u : -1;
}

// Layer geometry: fp32 offset, rows (out), padded cols (in).
__host__ __device__ inline void layer_geom(int L, int *off, int *rows, int *cols) {
  const int o[4] = {G_W1, G_W2, G_W3, G_W4}, r[4] = {64, 64, 64, 48}, c[4] = {48, 64, 64, 64};
  *off = o[L], *rows = r[L], *cols = c[L];
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

// K slot (s, g, q) of a 16x16x32 operand that holds hidden/output neuron n in
// the permuted order n = 16 (2 s + q/4) + 4 g + q%4 (a lane's own D tiles).
__host__ __device__ inline void neuron_kslot(int n, int *s, int *g, int *q) {
  const int mt = n / 16;
  *s = mt / 2, *g = (n % 16) / 4, *q = 4 * (mt % 2) + (n % 4);
}

// bf16 tiles from the fp32 master weights, one thread per (layer, out, in):
// the inference tiles (natural output order) and the training tiles.  Slots
// that hold no weight are never written (they stay zero).  Threads below
// 12 * 512 also write the identity tiles.
__global__ void k_guide_pack(const float *__restrict__ w, uint16_t *__restrict__ xtiles, uint16_t *__restrict__ rtiles) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < GR_ID * 512) {
    const int tile = t / 512, lane = (t / 8) % 64, q = t % 8, j = lane % 16, g = lane / 16;
    const int ni = (tile % 6) / 2, s = tile % 2;
    bool one;
    if (tile < 6) {
      one = g == ni && j == 8 * s + q;  // input slot (g, 8 s + q) -> column 16 g + 8 s + q
    } else {
      one = 16 * (2 * s + q / 4) + 4 * g + q % 4 == 16 * ni + j;  // output row -> column
    }
    rtiles[(GR_IDX_BASE + tile) * 512 + lane * 8 + q] = one ? 0x3F80u : 0u;
  }
  if (t >= G_B1) return;
  const int L = t < G_W2 ? 0 : (t < G_W3 ? 1 : (t < G_W4 ? 2 : 3));
  int off, rows, cols;
  layer_geom(L, &off, &rows, &cols);
  const int o = (t - off) / cols, i = (t - off) % cols;
  if (!param_live(t)) return;
  const uint16_t v = to_bf16(w[t]);
  // input column i -> K slot
  int g, s, q;
  if (L == 0) {
    int slot;
    feature_xslot(i, &g, &slot);
    s = slot / 8, q = slot % 8;
  } else {
    neuron_kslot(i, &s, &g, &q);
  }
  // inference: natural output rows
  xtiles[((GX_BASE[L] + (o / 16) * 2 + s) * 64 + (o % 16) + 16 * g) * 8 + q] = v;
  // training forward: W4 rows in lobe order
  const int p = L == 3 ? out_phys(o) : o;
  rtiles[((GX_BASE[L] + (p / 16) * 2 + s) * 64 + (p % 16) + 16 * g) * 8 + q] = v;
  // training transposed (A = Wᵀ: rows = inputs i, K = outputs p in the permuted order)
  if (L >= 1) {
    int ks, kg, kq;
    neuron_kslot(p, &ks, &kg, &kq);
    const int tb = GR_BWD_BASE + 8 * (3 - L);  // W4ᵀ at 0, W3ᵀ at 8, W2ᵀ at 16
    rtiles[((tb + (i / 16) * 2 + ks) * 64 + (i % 16) + 16 * kg) * 8 + kq] = v;
  }
}

__device__ __forceinline__ void guide_stage_weights(uint32_t *dst, const uint32_t *src, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

// Raw outputs for inputs in9 (n x 9) -> out (n x 40), on the inference tiles.
// 256 threads, 64 inputs per block-iteration (16 per wave).
__global__ void __launch_bounds__(256) k_guide_infer(const uint32_t *__restrict__ xtiles, const float *__restrict__ params,
                                                     const float *__restrict__ in, uint64_t n, float *__restrict__ out) {
  __shared__ uint32_t wt[GX_TILES * GX_TILE_WORDS];
  __shared__ __align__(16) float bias[G_BIAS];
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

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4v(f32x4 v) { return pack4(v[0], v[1], v[2], v[3]); }

// One lobe of a sample's mixture, as held by the lane that owns it.
struct lobe {
  float mu[3], nrm, kap, lk, cs, d, w;
};

// raw = (mx, my, mz, log κ); om = the sampled direction ω.
__device__ __forceinline__ void lobe_eval(float a, float b, float c, float lk, v3 om, lobe *L) {
  const float n2 = a * a + b * b + c * c;
  L->nrm = sqrtf(n2);
  if (n2 > 1e-20f) {
    const float inv = rsqrtf(n2);
    L->mu[0] = a * inv, L->mu[1] = b * inv, L->mu[2] = c * inv;
  } else {
    L->mu[0] = 0.0f, L->mu[1] = 0.0f, L->mu[2] = 1.0f;
  }
  L->lk = lk;
  L->kap = __expf(fminf(fmaxf(lk, G_LOGK_MIN), G_LOGK_MAX));
  L->cs = L->mu[0] * om.x + L->mu[1] * om.y + L->mu[2] * om.z;
  L->d = vmf_pdf(L->kap, L->cs);
}

// The lobe's 5 output gradients of -(c/n) log g, given f = -c / (n g) and g.
__device__ __forceinline__ void lobe_grad(const lobe &L, v3 om, float f, float g, float *o5) {
  const float wd = L.w * L.d;
  // d/d raw mean: κ (ω − (μ·ω) μ) / |raw|
  const float s = L.nrm > 1e-10f ? __fdividef(f * wd * L.kap, L.nrm) : 0.0f;
  o5[0] = s * (om.x - L.cs * L.mu[0]);
  o5[1] = s * (om.y - L.cs * L.mu[1]);
  o5[2] = s * (om.z - L.cs * L.mu[2]);
  // d/d log κ: 1 − 2κ/(e^{2κ} − 1) + κ(μ·ω − 1), zero where the clamp is active
  // (2κ ≥ 0.0996: e^{2κ} − 1 ≥ 0.105 keeps ~6 digits; e^{2κ} = inf gives 0)
  const bool live = L.lk > G_LOGK_MIN && L.lk < G_LOGK_MAX;
  o5[3] = live ? f * wd * (1.0f - __fdividef(2.0f * L.kap, __expf(2.0f * L.kap) - 1.0f) + L.kap * (L.cs - 1.0f))
               : 0.0f;
  // d/d logit: w_i (D_i − g)
  o5[4] = f * L.w * (L.d - g);
}

// Block of the training step: 8 waves, 128 samples per iteration.
constexpr int GT_WAVES = 8, GT_THREADS = 64 * GT_WAVES, GT_ROWS = 16 * GT_WAVES;
// dW exchange: 2 buffers x 8 waves x 8 operands (dn 0..3, hn 0..3) x 64 lanes
// of bf16x4, operand-major so that a wave's 64 lanes touch consecutive words.
constexpr int GT_XCH = 2 * GT_WAVES * 8 * 64;

constexpr size_t k_guide_train_lds() {
  return GR_TILES * 1024 + G_BIAS * 4 + GT_XCH * 8 + (G_BIAS + 4) * 4;
}

// dW tile owned by slot k (0..7) of wave w: layer L = 3 - k/2, index
// q = base + 8 (k%2) within the layer (row-major over (mo, ni)), base = w
// except for W1 (L = 0), where (w + 4) % 8 evens out the 12-tile layers.
// False if the slot holds no tile (a 12-tile layer's second slot).
__device__ __forceinline__ bool owned_tile(int w, int k, int *L, int *mo, int *ni) {
  const int l = 3 - k / 2;
  const int nt = (l == 3 || l == 0) ? 12 : 16, nc = l == 0 ? 3 : 4;
  const int q = (l == 0 ? (w + 4) % 8 : w) + 8 * (k % 2);
  *L = l, *mo = q / nc, *ni = q % nc;
  return q < nt;
}

// Fused forward/backward over rows [0, n) (row k read at (row0 + k) % n_total).
// GT_THREADS threads per block; each wave takes groups of 16 rows, the block
// 128 per iteration.  Writes this block's partial gradient (natural parameter
// layout) to partial.
__global__ void __launch_bounds__(GT_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
k_guide_train(const uint32_t *__restrict__ rtiles, const float *__restrict__ params, const float *__restrict__ rows,
              uint64_t n, uint64_t row0, uint64_t n_total, float *__restrict__ partial) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t *wt = reinterpret_cast<uint32_t *>(lds);
  float *bias = reinterpret_cast<float *>(wt + GR_TILES * 256);  // b1 b2 b3 natural, b4 in lobe order
  bf16x4 *xch = reinterpret_cast<bf16x4 *>(bias + G_BIAS);       // dW exchange (GT_XCH entries)
  float *sums = reinterpret_cast<float *>(xch + GT_XCH);         // biases, mix-logit gradient, loss
  guide_stage_weights(wt, rtiles, GR_TILES * 256);
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) {
    float b = 0.0f;
    if (i < 192) b = params[G_B1 + i];
    else if (out_logical(i - 192) >= 0) b = params[G_B4 + out_logical(i - 192)];
    bias[i] = b;
  }
  for (int i = threadIdx.x; i < G_BIAS + 4; i += blockDim.x) sums[i] = 0.0f;
  __syncthreads();
  const bf16x8 *T = reinterpret_cast<const bf16x8 *>(wt);
  const int l = __lane_id(), wave = threadIdx.x / 64, g = l >> 4, j = l & 15;
  const float inv_n = 1.0f / static_cast<float>(n);
  const float alpha = 1.0f / (1.0f + __expf(-params[G_NPARAM_PAD]));
  // the dW tiles this wave owns (owned_tile), accumulated over the launch
  f32x4 dw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) dw[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  int own_mo[8], own_ni[8];
  bool own[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int L;
    own[k] = owned_tile(wave, k, &L, &own_mo[k], &own_ni[k]);
  }
  int xbuf = 0;
  float db[15];  // per lane: neuron 16 t + j of b1 (0..3), b2 (4..7), b3 (8..11), b4 lobe-order rows (12..14)
#pragma unroll
  for (int k = 0; k < 15; ++k) db[k] = 0.0f;
  float dalpha = 0.0f, loss = 0.0f;
  // every wave of a block runs the same iterations (the dW exchange has block
  // barriers); a wave past the end works on zero-weight rows
  for (uint64_t bbase = static_cast<uint64_t>(blockIdx.x) * GT_ROWS; bbase < n;
       bbase += static_cast<uint64_t>(gridDim.x) * GT_ROWS) {
    const uint64_t base = bbase + 16 * wave;
    // an opaque zero offset keeps the weight-tile loads inside the loop (hoisted
    // out of it they would hold the whole tile set in registers)
    int zoff = 0;
    asm volatile("" : "+v"(zoff));
    const bf16x8 *Tl = T + zoff;
    const uint64_t sidx = base + j;
    const bool valid = sidx < n;
    float row[16];
    const float *rp = rows + ((row0 + (valid ? sidx : 0)) % n_total) * TPT_GUIDE_ROW;
#pragma unroll
    for (int k = 0; k < 16; ++k) row[k] = valid ? rp[k] : 0.0f;
    // ---- 1. forward, sample-major
    bf16x8 xa[2], a1[2], a2[2], a3[2];
    guide_xfeatures(g, g == 0 ? row[0] : (g == 1 ? row[1] : row[2]), row + 3, xa);
    uint32_t m2 = 0, m3 = 0;  // sample-major ReLU masks of layers 2, 3, bit 4 mt + r
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
      const bf16x8 *in = layer == 0 ? xa : (layer == 1 ? a1 : a2);
      f32x4 acc[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[mt] = mfma32(Tl[(GX_BASE[layer] + 2 * mt + s) * 64 + l], in[s], acc[mt]);
      }
      uint32_t m = 0;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float z = acc[mt][r] + bias[64 * layer + 16 * mt + 4 * g + r];
          m |= z > 0.0f ? (1u << (4 * mt + r)) : 0u;
          acc[mt][r] = z > 0.0f ? z : 0.0f;
        }
      bf16x8 *outp = layer == 0 ? a1 : (layer == 1 ? a2 : a3);
      outp[0] = pack8(acc[0], acc[1]);
      outp[1] = pack8(acc[2], acc[3]);
      if (layer == 1) m2 = m;
      else if (layer == 2) m3 = m;
    }
    f32x4 o[3];
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      o[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int s = 0; s < 2; ++s) o[mt] = mfma32(Tl[(GX_BASE[3] + 2 * mt + s) * 64 + l], a3[s], o[mt]);
#pragma unroll
      for (int r = 0; r < 4; ++r) o[mt][r] += bias[192 + 16 * mt + 4 * g + r];
    }
    // ---- 2. loss gradient: this lane holds lobes 2g (o0.xyzw, o1.x) and 2g+1 (o1.yzw, o2.xy)
    const v3 om{row[9], row[10], row[11]};
    const float c = row[12], qd = row[13];
    const bool use = valid && c > 0.0f && c < 3.0e38f;
    lobe A, B;
    lobe_eval(o[0][0], o[0][1], o[0][2], o[0][3], om, &A);
    lobe_eval(o[1][1], o[1][2], o[1][3], o[2][0], om, &B);
    const float la = o[1][0], lb = o[2][1];
    float mx = fmaxf(la, lb);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float ea = __expf(la - mx), eb = __expf(lb - mx);
    float z = ea + eb;
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    const float iz = __fdividef(1.0f, z);
    A.w = ea * iz, B.w = eb * iz;
    float gp = A.w * A.d + B.w * B.d;
    gp += __shfl_xor(gp, 16);
    gp += __shfl_xor(gp, 32);
    gp = fmaxf(gp, 1e-30f);
    const float f = use ? __fdividef(-c * inv_n, gp) : 0.0f;
    float ga[5], gb[5];
    lobe_grad(A, om, f, gp, ga);
    lobe_grad(B, om, f, gp, gb);
    f32x4 d[3];
    d[0] = use ? f32x4{ga[0], ga[1], ga[2], ga[3]} : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    d[1] = use ? f32x4{ga[4], gb[0], gb[1], gb[2]} : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    d[2] = use ? f32x4{gb[3], gb[4], 0.0f, 0.0f} : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (use && g == 0) {
      loss += -c * __logf(gp);
      const float mm = alpha * gp + (1.0f - alpha) * qd;
      if (mm > 0.0f) dalpha += __fdividef(-c * inv_n * (gp - qd) * alpha * (1.0f - alpha), mm);
    }
    // ---- 3..5, layer by layer from the top, keeping few operands live:
    // neuron-major operands (lane: neuron 16 t + j; registers: samples 4g + r)
    // come from swapped products; dW += δ hᵀ over these 16 samples.
    const bf16x8 dob[2] = {pack8(d[0], d[1]), pack8(d[2], f32x4{0.0f, 0.0f, 0.0f, 0.0f})};
    // Weight layer L = 3 (W4) .. 0 (W1): its output delta δ_L and its input
    // activations h_L, neuron-major, give dW_L += δ_L h_Lᵀ.  `dup` is δ_{L+1}
    // sample-major (δ_3 = dout) and `dmask` the neuron-major ReLU' of the
    // activations computed in the previous iteration (those of δ_L's layer).
    bf16x8 dup[2] = {dob[0], dob[1]};
    uint32_t dmask = 0;
#pragma unroll
    for (int L = 3; L >= 0; --L) {
      // h_L: L > 0 from the swapped forward of layer L-1 (input a_{L-1}), L = 0 the inputs
      bf16x4 hn[4];
      uint32_t nm = 0;
      if (L > 0) {
        const bf16x8 *hin = L == 3 ? a2 : (L == 2 ? a1 : xa);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s = 0; s < 2; ++s) acc = mfma32(hin[s], Tl[(GX_BASE[L - 1] + 2 * t + s) * 64 + l], acc);
          const float b = bias[64 * (L - 1) + 16 * t + j];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float zz = acc[r] + b;
            nm |= zz > 0.0f ? (1u << (4 * t + r)) : 0u;
            acc[r] = zz > 0.0f ? zz : 0.0f;
          }
          hn[t] = pack4v(acc);
        }
      } else {
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s = 0; s < 2; ++s) acc = mfma32(xa[s], Tl[(GR_IDX_BASE + 2 * t + s) * 64 + l], acc);
          hn[t] = pack4v(acc);
        }
        hn[3] = bf16x4{0, 0, 0, 0};
      }
      // δ_L, neuron-major: dout through the identity tiles, else
      // (W_{L+1}ᵀ δ_{L+1}) gated by ReLU' of layer L's activations
      bf16x4 dn[4];
      const int nt = L == 3 ? 3 : 4;
      const int tb = L == 3 ? GR_IDO_BASE : GR_BWD_BASE + 8 * (2 - L);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t < nt) {
          f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s = 0; s < 2; ++s) acc = mfma32(dup[s], Tl[(tb + 2 * t + s) * 64 + l], acc);
          if (L < 3) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = ((dmask >> (4 * t + r)) & 1u) ? acc[r] : 0.0f;
          }
          db[(L == 3 ? 12 : 4 * L) + t] += acc[0] + acc[1] + acc[2] + acc[3];
          dn[t] = pack4v(acc);
        } else {
          dn[t] = bf16x4{0, 0, 0, 0};
        }
      }
      // publish δ_L and h_L (neuron-major, this wave's 16 samples), then add
      // every wave's products into the owned tiles of layer L
      bf16x4 *xw = xch + (xbuf * GT_WAVES + wave) * 8 * 64 + l;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        xw[64 * t] = dn[t];
        xw[64 * (4 + t)] = hn[t];
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int k = 2 * (3 - L) + e;
        if (own[k]) {
          const bf16x4 *xa_ = xch + xbuf * GT_WAVES * 8 * 64 + 64 * own_mo[k] + l;
          const bf16x4 *xb_ = xch + xbuf * GT_WAVES * 8 * 64 + 64 * (4 + own_ni[k]) + l;
#pragma unroll
          for (int src = 0; src < GT_WAVES; ++src)
            dw[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa_[src * 8 * 64], xb_[src * 8 * 64], dw[k], 0, 0, 0);
        }
      }
      xbuf ^= 1;
      // δ_L sample-major, for the next iteration: (W_{L+1}ᵀ δ_{L+1}) ⊙ ReLU'
      if (L == 2 || L == 1) {
        const uint32_t sm = L == 2 ? m3 : m2;
        f32x4 acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s = 0; s < 2; ++s) acc[mt] = mfma32(Tl[(tb + 2 * mt + s) * 64 + l], dup[s], acc[mt]);
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[mt][r] = ((sm >> (4 * mt + r)) & 1u) ? acc[mt][r] : 0.0f;
        }
        dup[0] = pack8(acc[0], acc[1]);
        dup[1] = pack8(acc[2], acc[3]);
      }
      dmask = nm;
    }
  }
  // biases, mix logit, loss: over the lane groups, then the waves
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    float v = db[k];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (g == 0) {
      const int idx = k < 12 ? 64 * (k / 4) + 16 * (k % 4) + j : 192 + 16 * (k - 12) + j;  // b4: lobe-order row
      atomicAdd(&sums[idx], v);
    }
  }
  for (int k = 0; k < 6; ++k) {
    dalpha += __shfl_xor(dalpha, 1 << k);
    loss += __shfl_xor(loss, 1 << k);
  }
  if (l == 0) {
    atomicAdd(&sums[G_BIAS], dalpha);
    atomicAdd(&sums[G_BIAS + 1], loss);
  }
  __syncthreads();
  float *P = partial + static_cast<size_t>(blockIdx.x) * G_PARTIAL;
  // this wave's tiles: tile (L, mo, ni), lane (g, j), register r holds
  // dW[16 mo + 4 g + r][column 16 ni + j]
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (!own[k]) continue;
    const int L = 3 - k / 2, mo = own_mo[k], ni = own_ni[k];
    int off, rws, cols;
    layer_geom(L, &off, &rws, &cols);
    const int col = L == 0 ? xslot_feature(ni, j) : 16 * ni + j;  // identity column 16 g + t = input slot (g, t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = 16 * mo + 4 * g + r;
      const int row_ = L == 3 ? out_logical(rr) : rr;
      if (row_ >= 0 && col >= 0) P[off + row_ * cols + col] = dw[k][r];
    }
  }
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) {
    if (i < 192) P[G_B1 + i] = sums[i];
    else if (out_logical(i - 192) >= 0) P[G_B4 + out_logical(i - 192)] = sums[i];
  }
  if (threadIdx.x == 0) {
    P[G_NPARAM_PAD] = sums[G_BIAS];
    P[G_NPARAM_PAD + 1] = sums[G_BIAS + 1];
  }
}

// acc[p] += the sum of the block partials of this y-chunk (PART_CHUNK of them).
constexpr int PART_CHUNK = 16;
__global__ void k_guide_reduce(const float *__restrict__ partial, int blocks, float *__restrict__ acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > G_NPARAM_PAD + 1) return;
  const int b0 = blockIdx.y * PART_CHUNK, b1 = min(blocks, b0 + PART_CHUNK);
  float s = 0.0f;
  for (int b = b0; b < b1; ++b) s += partial[static_cast<size_t>(b) * G_PARTIAL + p];
  if (b1 > b0) atomicAdd(&acc[p], s);
}

// Consumes acc (and clears it for the next step): mode 0 writes the gradient
// to grad, mode 1 applies Adam (step `step`, learning rate lr).
__global__ void k_guide_adam(float *__restrict__ acc, float *__restrict__ params, float *__restrict__ adam_m,
                             float *__restrict__ adam_v, float *__restrict__ grad, float *__restrict__ loss_out,
                             int mode, float lr, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > G_NPARAM_PAD + 1) return;
  float g = acc[p];
  acc[p] = 0.0f;
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
