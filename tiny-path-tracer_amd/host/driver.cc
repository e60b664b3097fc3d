// tpt_render_main: the reference executable's driver (main.cpp) on the device
// path.  It reads the same config.ini keys (main.cpp:29-58, config.ini:1-13),
// renders through the C ABI and writes the same files: img.ppm (main.cpp:69,
// 176-189) and, with allow_bonus_pic, the progressive snapshots img_k.ppm
// taken every sample/bonus_pic samples (main.cpp:111-133, 191-215).
//
//   tpt_render_main [config.ini] [--dry-run]
//
// Extensions (section [TPT], optional): scene = a name from tpt_scene_names()
// (default ref_cornell_box, the reference driver's scene and light list),
// seed = Philox key (default 1), queue_capacity.  The camera is the scene's
// (look-from, look-at, focus distance) with the configured fov, aspect
// width/height, aperture and shutter, as main.cpp:87-91 builds it.  The
// reference's ImageMagick merge step (main.cpp:222-238) is not run.
#include <hip/hip_runtime.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tpt.h"

namespace {

using ini_t = std::map<std::string, std::map<std::string, std::string>>;

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

// [section] / key = value / ';' or '#' comment lines, as the reference's ini
// reader accepts them.
ini_t parse_ini(std::istream &in) {
  ini_t ini;
  std::string line, section;
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t.empty() || t[0] == ';' || t[0] == '#') continue;
    if (t.front() == '[' && t.back() == ']') {
      section = trim(t.substr(1, t.size() - 2));
      continue;
    }
    const size_t eq = t.find('=');
    if (eq == std::string::npos) continue;
    ini[section][trim(t.substr(0, eq))] = trim(t.substr(eq + 1));
  }
  return ini;
}

// Overwrites *v when the key is present and parses completely (inipp::extract).
template <class T>
void extract(const ini_t &ini, const char *sec, const char *key, T *v) {
  const auto s = ini.find(sec);
  if (s == ini.end()) return;
  const auto k = s->second.find(key);
  if (k == s->second.end()) return;
  std::istringstream is(k->second);
  T x;
  if (is >> x && (is >> std::ws).eof()) *v = x;
}

struct config {
  int nx = 400, ny = 200, ns = 10;  // main.cpp:43-51 defaults
  float aperture = 0.2f, time0 = 0.0f, time1 = 1.0f;
  int depth = 50;
  float fov = 90.0f;
  int allow_bonus_pic = 0, bonus_pic = 10;
  std::string scene = "ref_cornell_box";
  unsigned long long seed = 1;
  unsigned queue_capacity = 0;
};

config read_config(const ini_t &ini) {
  config c;
  extract(ini, "DEFAULT", "width", &c.nx);
  extract(ini, "DEFAULT", "height", &c.ny);
  extract(ini, "DEFAULT", "sample", &c.ns);
  extract(ini, "DEFAULT", "recur_depth", &c.depth);
  extract(ini, "DEFAULT", "fov", &c.fov);
  extract(ini, "DEFAULT", "bonus_pic", &c.bonus_pic);
  extract(ini, "DEFAULT", "allow_bonus_pic", &c.allow_bonus_pic);
  extract(ini, "BLUR", "aperture", &c.aperture);
  extract(ini, "CAM_MOTION", "start_time", &c.time0);
  extract(ini, "CAM_MOTION", "end_time", &c.time1);
  extract(ini, "TPT", "scene", &c.scene);
  extract(ini, "TPT", "seed", &c.seed);
  extract(ini, "TPT", "queue_capacity", &c.queue_capacity);
  return c;
}

int clamp_rgb(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

// int(255.99f * sqrt(sum / n)) per channel (main.cpp:135-139); a non-finite
// value (the reference's int conversion of it is undefined) is written as 0.
int to_byte(float sum, float n) {
  const float c = std::sqrt(sum / n);
  const float q = 255.99f * c;
  if (!std::isfinite(q)) return 0;
  return clamp_rgb(static_cast<int>(q));
}

// img.ppm: rows top first, one pixel per line (main.cpp:176-189); the bonus
// snapshots put every pixel on one line (main.cpp:206-209).
bool write_ppm(const std::string &path, const std::vector<float> &sum, int nx, int ny, float n, bool one_line) {
  std::ofstream out(path);
  if (!out) return false;
  out << "P3\n" << nx << " " << ny << "\n255\n";
  for (int j = ny - 1; j >= 0; --j)
    for (int i = 0; i < nx; ++i) {
      const float *c = &sum[3 * (static_cast<size_t>(j) * nx + i)];
      out << to_byte(c[0], n) << ' ' << to_byte(c[1], n) << ' ' << to_byte(c[2], n) << ' ';
      if (!one_line) out << '\n';
    }
  return static_cast<bool>(out);
}

int fail(const char *what) {
  std::fprintf(stderr, "tpt_render_main: %s: %s\n", what, tpt_last_error());
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  std::string path = "config.ini";
  bool dry = false;
  for (int a = 1; a < argc; ++a) {
    if (std::strcmp(argv[a], "--dry-run") == 0) dry = true;
    else path = argv[a];
  }
  std::ifstream cf(path);
  const ini_t ini = parse_ini(cf);
  const config c = read_config(ini);
  for (const auto &sec : ini) {
    std::cout << "[" << sec.first << "]\n";
    for (const auto &kv : sec.second) std::cout << kv.first << "=" << kv.second << "\n";
  }
  std::cout << "<===========>\nscene " << c.scene << ", " << c.nx << "x" << c.ny << ", " << c.ns
            << " spp, depth " << c.depth << ", fov " << c.fov << ", aperture " << c.aperture << ", shutter ["
            << c.time0 << ", " << c.time1 << "], bonus " << (c.allow_bonus_pic ? c.bonus_pic : 0) << "\n<===========>"
            << std::endl;
  if (c.nx <= 0 || c.ny <= 0 || c.ns <= 0 || c.depth < 0) {
    std::fprintf(stderr, "tpt_render_main: width, height and sample must be positive\n");
    return 1;
  }
  const int slice = c.allow_bonus_pic ? (c.bonus_pic > 0 ? c.ns / c.bonus_pic : 0) : c.ns;
  if (c.allow_bonus_pic && slice <= 0) {
    std::fprintf(stderr, "tpt_render_main: bonus_pic must be in [1, sample]\n");
    return 1;
  }
  if (dry) return 0;

  tpt_host_scene *hs = nullptr;
  if (tpt_host_scene_build(c.scene.c_str(), 5489, &hs) != TPT_OK) return fail("scene");
  tpt_flat_scene flat;
  tpt_scene_info info;
  if (tpt_host_scene_flat(hs, &flat) != TPT_OK || tpt_host_scene_info(hs, &info) != TPT_OK) return fail("scene");
  tpt_camera cam;
  const float vup[3] = {0.0f, 1.0f, 0.0f};
  if (tpt_camera_make(info.lookfrom, info.lookat, vup, c.fov, static_cast<float>(c.nx) / static_cast<float>(c.ny),
                      c.aperture, info.focus, c.time0, c.time1, &cam) != TPT_OK)
    return fail("camera");
  tpt_scene *scene = nullptr;
  if (tpt_scene_create(&flat, &scene) != TPT_OK) return fail("tpt_scene_create");

  const size_t n = 3ull * c.nx * c.ny;
  float *d_sum = nullptr;
  if (hipMalloc(&d_sum, n * sizeof(float)) != hipSuccess || hipMemset(d_sum, 0, n * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "tpt_render_main: device allocation failed\n");
    return 1;
  }
  std::vector<float> host(n);
  tpt_render_desc desc;
  std::memset(&desc, 0, sizeof(desc));
  desc.width = static_cast<uint32_t>(c.nx);
  desc.height = static_cast<uint32_t>(c.ny);
  desc.max_depth = static_cast<uint32_t>(c.depth);
  desc.seed = c.seed;
  desc.background = info.background;
  desc.queue_capacity = c.queue_capacity;
  const auto t0 = std::chrono::high_resolution_clock::now();
  // progressive passes: one per snapshot, then the rest of the samples
  const int snaps = c.allow_bonus_pic ? c.bonus_pic : 0;
  int done = 0;
  for (int k = 0; k <= snaps; ++k) {
    const int upto = k < snaps ? slice * (k + 1) : c.ns;
    if (upto > done) {
      desc.sample_offset = static_cast<uint32_t>(done);
      desc.spp = static_cast<uint32_t>(upto - done);
      if (tpt_render(scene, &cam, &desc, d_sum, nullptr, nullptr, nullptr) != TPT_OK) return fail("tpt_render");
      done = upto;
    }
    if (k < snaps) {
      if (hipMemcpy(host.data(), d_sum, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return fail("copy");
      const std::string file = "img_" + std::to_string(k) + ".ppm";
      if (!write_ppm(file, host, c.nx, c.ny, static_cast<float>(upto), true)) return fail(file.c_str());
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
      std::cout << "snapshot " << file << " at " << upto << " spp, " << ms / 1000.0 << " s" << std::endl;
    }
  }
  if (hipMemcpy(host.data(), d_sum, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return fail("copy");
  if (!write_ppm("img.ppm", host, c.nx, c.ny, static_cast<float>(c.ns), false)) return fail("img.ppm");
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
  std::cout << "time: " << ms / 1000.0 << " s" << std::endl;
  (void)hipFree(d_sum);
  tpt_scene_destroy(scene);
  tpt_host_scene_destroy(hs);
  return 0;
}
