// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
// g++; no device code is built): every named scene is built, flattened and
// released through the C ABI, the oracle intersects and renders each, and the
// image loader behind load_image_texture (the stb_image replacement,
// src/utils.cc:236-240) reads seed files and seeded mutations of them.  The
// build stops at the first report (-fno-sanitize-recover=all).
//
//   san_main [iterations] [seed files...]
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "../../include/tpt.h"

extern "C" {
struct oracle_scene;
int oracle_scene_build(const char *name, uint32_t seed, oracle_scene **out);
void oracle_scene_destroy(oracle_scene *s);
int oracle_make_rays(const oracle_scene *s, int kind, uint64_t n, uint64_t seed, float *rays, int nthreads);
int oracle_trace(const oracle_scene *s, const float *rays, uint64_t n, float tmin, float tmax, int32_t *leaf,
                 float *t, float *rec, int32_t *mat, int nthreads, const float *time);
int oracle_render(const oracle_scene *s, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth, uint64_t seed,
                  uint32_t sample_offset, float background, int rng_mode, int nthreads, float *sum, float *sumsq,
                  int media_scatter);
}

namespace {

std::vector<std::string> split(const std::string &s) {
  std::vector<std::string> out;
  size_t a = 0;
  while (a <= s.size()) {
    const size_t b = s.find(',', a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

int scenes() {
  int built = 0;
  for (std::string name : split(tpt_scene_names())) {
    if (name == "random_spheres_n<N>") name = "random_spheres_n3000";
    tpt_host_scene *hs = nullptr;
    if (tpt_host_scene_build(name.c_str(), 5489, &hs) != TPT_OK) {
      std::printf("scene %-24s refused: %s\n", name.c_str(), tpt_last_error());
      continue;
    }
    tpt_flat_scene flat;
    tpt_camera cam;
    tpt_scene_info info;
    if (tpt_host_scene_flat(hs, &flat) != TPT_OK || tpt_host_scene_camera(hs, &cam) != TPT_OK ||
        tpt_host_scene_info(hs, &info) != TPT_OK) {
      std::printf("scene %s: host API failed: %s\n", name.c_str(), tpt_last_error());
      return 1;
    }
    tpt_host_scene_destroy(hs);
    oracle_scene *os = nullptr;
    if (oracle_scene_build(name.c_str(), 5489, &os) != 0) {
      std::printf("scene %s: oracle build failed\n", name.c_str());
      return 1;
    }
    const uint64_t n = 2048;
    std::vector<float> rays(6 * n), t(n), rec(8 * n), sum(12 * 12 * 3, 0.0f), sq(12 * 12 * 3, 0.0f);
    std::vector<int32_t> leaf(n), mat(n);
    for (int kind = 0; kind < 4; ++kind) {
      oracle_make_rays(os, kind, n, 7 + kind, rays.data(), 4);
      oracle_trace(os, rays.data(), n, 0.001f, 3.4e38f, leaf.data(), t.data(), rec.data(), mat.data(), 4, nullptr);
    }
    oracle_render(os, 12, 12, 2, 50, 3, 0, 1.0f, 0, 4, sum.data(), sq.data(), 0);
    oracle_scene_destroy(os);
    std::printf("scene %-24s %u slots, %u leaves: built, flattened, traced, rendered\n", name.c_str(), flat.n_slots,
                flat.n_leaves);
    ++built;
  }
  return built > 0 ? 0 : 1;
}

// Seeded mutations: byte overwrites, truncations, insertions, and 16-bit
// fields set to extreme values (lengths and sizes in headers).
std::vector<uint8_t> mutate(const std::vector<uint8_t> &base, std::mt19937 &g) {
  std::vector<uint8_t> d = base;
  const int edits = 1 + static_cast<int>(g() % 6);
  for (int e = 0; e < edits && !d.empty(); ++e) {
    const size_t k = g() % d.size();
    switch (g() % 4) {
      case 0:
        d[k] = static_cast<uint8_t>(g());
        break;
      case 1:
        d.resize(std::max<size_t>(2, k));
        break;
      case 2:
        d.insert(d.begin() + static_cast<long>(k), static_cast<uint8_t>(g()));
        break;
      default:
        if (k + 1 < d.size()) {
          const uint8_t v[4] = {0x00, 0xFF, 0x7F, 0x80};
          d[k] = v[g() % 4];
          d[k + 1] = v[g() % 4];
        }
    }
  }
  return d;
}

int fuzz(int iters, const std::vector<std::string> &files) {
  std::vector<std::vector<uint8_t>> seeds;
  for (const auto &f : files) {
    std::ifstream in(f, std::ios::binary);
    if (!in) continue;
    seeds.emplace_back(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  }
  // a PPM of its own: 5x7, P6
  std::vector<uint8_t> ppm;
  const std::string hdr = "P6\n5 7\n255\n";
  ppm.assign(hdr.begin(), hdr.end());
  for (int i = 0; i < 5 * 7 * 3; ++i) ppm.push_back(static_cast<uint8_t>(i * 37));
  seeds.push_back(ppm);
  const std::string path = "san_fuzz_input.bin";
  std::mt19937 g(20261203);
  int decoded = 0, refused = 0;
  for (int it = 0; it < iters; ++it) {
    const auto &base = seeds[static_cast<size_t>(it) % seeds.size()];
    const std::vector<uint8_t> d = it < static_cast<int>(seeds.size()) ? base : mutate(base, g);
    {
      std::ofstream out(path, std::ios::binary | std::ios::trunc);
      out.write(reinterpret_cast<const char *>(d.data()), static_cast<std::streamsize>(d.size()));
    }
    uint8_t *px = nullptr;
    int w = 0, h = 0;
    if (tpt_image_load(path.c_str(), &px, &w, &h) == TPT_OK) {
      // touch every decoded byte: a short buffer is an ASan report
      unsigned sum = 0;
      for (size_t i = 0; i < static_cast<size_t>(w) * h * 3; ++i) sum += px[i];
      (void)sum;
      tpt_image_free(px);
      ++decoded;
    } else {
      ++refused;
    }
  }
  std::remove(path.c_str());
  std::printf("image loader: %d inputs from %zu seeds: %d decoded, %d refused, no sanitizer report\n", iters,
              seeds.size(), decoded, refused);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
  std::vector<std::string> files;
  for (int i = 2; i < argc; ++i) files.push_back(argv[i]);
  if (scenes() != 0) return 1;
  return fuzz(iters, files);
}
