// Image files for image_texture: baseline JPEG and PPM (see image_load.h).
//
// The JPEG decoder follows ITU-T T.81 (the JPEG standard) directly: marker
// parsing (Annex B), Huffman tables and the DECODE / RECEIVE / EXTEND
// procedures (Annex F.2.2), dequantisation in zig-zag order, the 8x8 inverse
// DCT of A.3.3 evaluated separably in double, level shift, and the JFIF
// YCbCr -> RGB conversion.  Sequential Huffman frames (SOF0/SOF1) with 8-bit
// samples, 1 or 3 components, any sampling factors (chroma is upsampled by
// replication), interleaved or single-component scans and restart intervals.
#include "image_load.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace tpt {
namespace {

// Largest image either decoder accepts (2^28 pixels, 768 MiB of RGB8).
constexpr uint64_t kMaxPixels = 1ull << 28;

// Zig-zag order: natural (row-major) index of the k-th coefficient.
constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct huffman {
  bool defined = false;
  int mincode[17], maxcode[18], valptr[17];
  uint8_t vals[256];
};

struct component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;  // Huffman tables of the current scan
  int bw = 0, bh = 0;  // blocks across / down in the component's plane (whole MCUs)
  long long dc_pred = 0;  // a crafted file may sum many large differences
  std::vector<uint8_t> plane;  // bw*8 x bh*8 samples
};

struct jpeg_error {
  std::string what;
};

class decoder {
 public:
  decoder(const uint8_t *d, size_t n) : d_(d), n_(n) {}

  void run(std::vector<uint8_t> *rgb, int *w, int *h) {
    if (n_ < 4 || d_[0] != 0xFF || d_[1] != 0xD8) fail("not a JPEG file");
    pos_ = 2;
    bool frame = false;
    for (;;) {
      const int m = next_marker();
      if (m == 0xD9) break;  // EOI
      const int len = u16();
      if (len < 2 || pos_ + len - 2 > n_) fail("truncated segment");
      const size_t end = pos_ + len - 2;
      if (m == 0xDB) {
        read_dqt(end);
      } else if (m == 0xC4) {
        read_dht(end);
      } else if (m == 0xC0 || m == 0xC1) {
        read_sof(end);
        frame = true;
      } else if ((m >= 0xC2 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
        fail("only baseline / extended sequential Huffman JPEG is supported");
      } else if (m == 0xDD) {
        restart_ = u16();
      } else if (m == 0xDA) {
        if (!frame) fail("scan before frame header");
        read_sos(end);
        pos_ = end;
        decode_scan();
        continue;
      }
      pos_ = end;
    }
    if (!frame) fail("no frame");
    convert(rgb);
    *w = width_;
    *h = height_;
  }

 private:
  const uint8_t *d_;
  size_t n_, pos_ = 0;
  int width_ = 0, height_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0, restart_ = 0;
  uint16_t q_[4][64] = {};
  huffman dc_[4], ac_[4];
  std::vector<component> comps_;
  std::vector<int> scan_;  // component indices of the current scan
  // entropy-coded segment reader
  uint32_t bits_ = 0;
  int nbits_ = 0;
  bool hit_marker_ = false;

  [[noreturn]] static void fail(const std::string &m) { throw jpeg_error{m}; }

  int u8() {
    if (pos_ >= n_) fail("unexpected end of file");
    return d_[pos_++];
  }
  int u16() {
    const int a = u8();
    return (a << 8) | u8();
  }
  int next_marker() {
    if (u8() != 0xFF) fail("marker expected");
    int m = u8();
    while (m == 0xFF) m = u8();  // fill bytes
    return m;
  }

  void read_dqt(size_t end) {
    while (pos_ < end) {
      const int pq_tq = u8();
      const int pq = pq_tq >> 4, tq = pq_tq & 15;
      if (tq > 3) fail("bad quantisation table id");
      for (int k = 0; k < 64; ++k) q_[tq][k] = static_cast<uint16_t>(pq ? u16() : u8());  // zig-zag order
    }
  }

  // Annex C: code lengths -> canonical codes; F.2.2.3 decoding tables.
  void read_dht(size_t end) {
    while (pos_ < end) {
      const int tc_th = u8();
      const int tc = tc_th >> 4, th = tc_th & 15;
      if (tc > 1 || th > 3) fail("bad Huffman table id");
      huffman &t = tc == 0 ? dc_[th] : ac_[th];
      int counts[17] = {0};
      int total = 0;
      for (int l = 1; l <= 16; ++l) total += counts[l] = u8();
      if (total > 256) fail("bad Huffman table");
      for (int i = 0; i < total; ++i) t.vals[i] = static_cast<uint8_t>(u8());
      int code = 0, k = 0;
      for (int l = 1; l <= 16; ++l) {
        t.valptr[l] = k;
        t.mincode[l] = code;
        code += counts[l];
        k += counts[l];
        t.maxcode[l] = counts[l] ? code - 1 : -1;
        code <<= 1;
      }
      t.maxcode[17] = 0x7FFFFFFF;
      t.defined = true;
    }
  }

  void read_sof(size_t end) {
    if (u8() != 8) fail("only 8-bit samples are supported");
    height_ = u16();
    width_ = u16();
    const int nf = u8();
    if (width_ <= 0 || height_ <= 0) fail("bad image size");
    if (static_cast<uint64_t>(width_) * static_cast<uint64_t>(height_) > kMaxPixels) fail("image too large");
    if (nf != 1 && nf != 3) fail("only 1 or 3 components are supported");
    comps_.assign(nf, component{});
    for (auto &c : comps_) {
      c.id = u8();
      const int hv = u8();
      c.h = hv >> 4;
      c.v = hv & 15;
      c.tq = u8();
      if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail("bad component");
      hmax_ = std::max(hmax_, c.h);
      vmax_ = std::max(vmax_, c.v);
    }
    mcux_ = (width_ + 8 * hmax_ - 1) / (8 * hmax_);
    mcuy_ = (height_ + 8 * vmax_ - 1) / (8 * vmax_);
    for (auto &c : comps_) {
      c.bw = mcux_ * c.h;
      c.bh = mcuy_ * c.v;
      c.plane.assign(static_cast<size_t>(c.bw) * 8 * c.bh * 8, 0);
    }
    if (pos_ != end) fail("bad frame header");
  }

  void read_sos(size_t end) {
    const int ns = u8();
    scan_.clear();
    for (int i = 0; i < ns; ++i) {
      const int cid = u8(), tdta = u8();
      int idx = -1;
      for (size_t k = 0; k < comps_.size(); ++k)
        if (comps_[k].id == cid) idx = static_cast<int>(k);
      if (idx < 0) fail("scan of an unknown component");
      comps_[idx].td = tdta >> 4;
      comps_[idx].ta = tdta & 15;
      if (comps_[idx].td > 3 || comps_[idx].ta > 3) fail("bad scan table id");
      scan_.push_back(idx);
    }
    pos_ += 3;  // Ss, Se, Ah/Al (0, 63, 0 in sequential mode)
    if (scan_.empty() || pos_ != end) fail("bad scan header");
  }

  // F.2.2.5: next bit of the entropy-coded segment (0xFF00 is a data 0xFF; a
  // marker ends the data, after which zeros are supplied).
  int bit() {
    if (nbits_ == 0) {
      int b = 0;
      if (!hit_marker_ && pos_ < n_) {
        b = d_[pos_];
        if (b == 0xFF) {
          const int nx = pos_ + 1 < n_ ? d_[pos_ + 1] : 0;
          if (nx == 0x00) {
            pos_ += 2;
          } else {
            hit_marker_ = true;
            b = 0;
          }
        } else {
          ++pos_;
        }
      }
      bits_ = static_cast<uint32_t>(b);
      nbits_ = 8;
    }
    --nbits_;
    return static_cast<int>((bits_ >> nbits_) & 1u);
  }
  int receive(int s) {
    int v = 0;
    for (int i = 0; i < s; ++i) v = (v << 1) | bit();
    return v;
  }
  static int extend(int v, int t) { return t == 0 ? 0 : (v < (1 << (t - 1)) ? v - (1 << t) + 1 : v); }
  int decode(const huffman &t) {
    if (!t.defined) fail("undefined Huffman table");
    int code = bit();
    int l = 1;
    while (code > t.maxcode[l]) {
      code = (code << 1) | bit();
      if (++l > 16) fail("bad Huffman code");
    }
    return t.vals[t.valptr[l] + code - t.mincode[l]];
  }

  void restart_marker() {
    nbits_ = 0;
    hit_marker_ = false;
    // skip to and over RSTn
    while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)) ++pos_;
    if (pos_ + 1 < n_) pos_ += 2;
    for (auto &c : comps_) c.dc_pred = 0;
  }

  // One 8x8 block into the component plane at block (bx, by).
  void block(component &c, int bx, int by) {
    double coef[64] = {0};
    const uint16_t *q = q_[c.tq];
    const int t = decode(dc_[c.td]);
    if (t > 11) fail("bad DC magnitude category");  // F.1.2.1.1: at most 11 for 8-bit samples
    c.dc_pred += extend(receive(t), t);
    coef[0] = static_cast<double>(c.dc_pred) * q[0];
    for (int k = 1; k < 64;) {
      const int rs = decode(ac_[c.ta]);
      const int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r != 15) break;  // EOB
        k += 16;             // ZRL
        continue;
      }
      if (s > 10) fail("bad AC magnitude category");  // F.1.2.2.1: at most 10 for 8-bit samples
      k += r;
      if (k > 63) fail("bad AC run");
      coef[kZigzag[k]] = static_cast<double>(extend(receive(s), s)) * q[k];
      ++k;
    }
    idct_store(coef, &c.plane[(static_cast<size_t>(by) * 8) * (c.bw * 8) + static_cast<size_t>(bx) * 8], c.bw * 8);
  }

  // A.3.3: s(y,x) = 1/4 sum_u sum_v C(u) C(v) S(v,u) cos((2x+1)u pi/16) cos((2y+1)v pi/16), + 128.
  static void idct_store(const double *in, uint8_t *out, int stride) {
    static double basis[8][8];
    static bool init = false;
    if (!init) {
      for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u)
          basis[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0);
      init = true;
    }
    double tmp[64];
    for (int v = 0; v < 8; ++v)  // rows of coefficients -> horizontal samples
      for (int x = 0; x < 8; ++x) {
        double s = 0;
        for (int u = 0; u < 8; ++u) s += basis[x][u] * in[v * 8 + u];
        tmp[v * 8 + x] = s;
      }
    for (int y = 0; y < 8; ++y)
      for (int x = 0; x < 8; ++x) {
        double s = 0;
        for (int v = 0; v < 8; ++v) s += basis[y][v] * tmp[v * 8 + x];
        const long r = std::lround(s / 4.0 + 128.0);
        out[y * stride + x] = static_cast<uint8_t>(r < 0 ? 0 : (r > 255 ? 255 : r));
      }
  }

  void decode_scan() {
    nbits_ = 0;
    hit_marker_ = false;
    for (auto &c : comps_) c.dc_pred = 0;
    int todo = restart_;
    long left = 0;  // MCUs after the current one
    auto tick = [&]() {
      if (restart_ && --todo == 0 && left > 0) {
        restart_marker();
        todo = restart_;
      }
    };
    if (scan_.size() == 1) {
      // non-interleaved: the component's own blocks, ceil(w_c/8) x ceil(h_c/8)
      component &c = comps_[scan_[0]];
      const int cw = (width_ * c.h + hmax_ - 1) / hmax_, ch = (height_ * c.v + vmax_ - 1) / vmax_;
      const int nbx = (cw + 7) / 8, nby = (ch + 7) / 8;
      left = static_cast<long>(nbx) * nby;
      for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
          --left;
          block(c, bx, by);
          tick();
        }
    } else {
      left = static_cast<long>(mcux_) * mcuy_;
      for (int my = 0; my < mcuy_; ++my)
        for (int mx = 0; mx < mcux_; ++mx) {
          --left;
          for (int idx : scan_) {
            component &c = comps_[idx];
            for (int v = 0; v < c.v; ++v)
              for (int h = 0; h < c.h; ++h) block(c, mx * c.h + h, my * c.v + v);
          }
          tick();
        }
    }
    // continue after the entropy-coded data: find the next marker
    while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] != 0x00 && !(d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)))
      ++pos_;
  }

  // Chroma upsampled by replication; JFIF YCbCr -> RGB.
  void convert(std::vector<uint8_t> *rgb) const {
    rgb->assign(static_cast<size_t>(width_) * height_ * 3, 0);
    auto sample = [&](const component &c, int x, int y) {
      const int sx = x * c.h / hmax_, sy = y * c.v / vmax_;
      return static_cast<double>(c.plane[static_cast<size_t>(sy) * (c.bw * 8) + sx]);
    };
    auto clamp8 = [](double v) {
      const long r = std::lround(v);
      return static_cast<uint8_t>(r < 0 ? 0 : (r > 255 ? 255 : r));
    };
    for (int y = 0; y < height_; ++y)
      for (int x = 0; x < width_; ++x) {
        uint8_t *o = &(*rgb)[(static_cast<size_t>(y) * width_ + x) * 3];
        if (comps_.size() == 1) {
          o[0] = o[1] = o[2] = static_cast<uint8_t>(sample(comps_[0], x, y));
          continue;
        }
        const double Y = sample(comps_[0], x, y), cb = sample(comps_[1], x, y) - 128.0,
                     cr = sample(comps_[2], x, y) - 128.0;
        o[0] = clamp8(Y + 1.402 * cr);
        o[1] = clamp8(Y - 0.344136 * cb - 0.714136 * cr);
        o[2] = clamp8(Y + 1.772 * cb);
      }
  }
};

bool read_file(const std::string &path, std::vector<uint8_t> *out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  out->assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  return true;
}

// P6 (binary) and P3 (ASCII) with maxval 255.
bool decode_ppm(const std::vector<uint8_t> &f, std::vector<uint8_t> *rgb, int *w, int *h, std::string *err) {
  size_t p = 2;
  const bool ascii = f[1] == '3';
  auto token = [&](long *v) {
    for (;;) {
      while (p < f.size() && std::isspace(f[p])) ++p;
      if (p < f.size() && f[p] == '#') {
        while (p < f.size() && f[p] != '\n') ++p;
        continue;
      }
      break;
    }
    if (p >= f.size() || !std::isdigit(f[p])) return false;
    long x = 0;
    while (p < f.size() && std::isdigit(f[p])) {
      x = x * 10 + (f[p++] - '0');
      if (x > 1000000) return false;  // no field of an accepted file is this large
    }
    *v = x;
    return true;
  };
  long W, H, M;
  if (!token(&W) || !token(&H) || !token(&M) || W <= 0 || H <= 0 || M != 255) {
    *err = "bad PPM header (8-bit P3/P6 only)";
    return false;
  }
  if (W > 65535 || H > 65535 || static_cast<uint64_t>(W) * static_cast<uint64_t>(H) > kMaxPixels) {
    *err = "PPM image too large";
    return false;
  }
  const size_t n = static_cast<size_t>(W) * H * 3;
  rgb->resize(n);
  if (ascii) {
    for (size_t i = 0; i < n; ++i) {
      long v;
      if (!token(&v) || v > 255) {
        *err = "truncated P3 data";
        return false;
      }
      (*rgb)[i] = static_cast<uint8_t>(v);
    }
  } else {
    ++p;  // the single whitespace after maxval
    if (p + n > f.size()) {
      *err = "truncated P6 data";
      return false;
    }
    std::memcpy(rgb->data(), &f[p], n);
  }
  *w = static_cast<int>(W);
  *h = static_cast<int>(H);
  return true;
}

}  // namespace

bool load_image_rgb8(const std::string &path, std::vector<uint8_t> *rgb, int *width, int *height, std::string *err) {
  std::vector<uint8_t> f;
  if (!read_file(path, &f)) {
    *err = "cannot open " + path;
    return false;
  }
  try {
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) {
      try {
        decoder(f.data(), f.size()).run(rgb, width, height);
        return true;
      } catch (const jpeg_error &e) {
        *err = path + ": " + e.what;
        return false;
      }
    }
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == '6' || f[1] == '3')) return decode_ppm(f, rgb, width, height, err);
  } catch (const std::exception &e) {  // std::bad_alloc, std::length_error
    *err = path + ": " + e.what();
    return false;
  }
  *err = path + ": not a JPEG or PPM file";
  return false;
}

}  // namespace tpt
