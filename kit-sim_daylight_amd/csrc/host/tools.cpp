// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The two command-line tools over the library: `daylight` and, built with VIEWS_CLI, `views`. Their contract is
// reference src/main.cpp's: the flags and rules of cli.h, exit codes 0 / 1 (usage, with the usage text) / 2 (runtime),
// values on stdout or in --out FILE.txt ("%.10g", one row per element) or FILE.npy (float64 [N x C]) with FILE.json the
// manifest, and --bench NSENS NTRIS: the reference's synthetic benchmark scenes, generated by the same steps from the
// same seeded draws so that its published throughputs compare with this library's on the same scene.
#include "glibc_guard.h"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cli.h"
#include "core.h"

using namespace daylight;
#ifdef VIEWS_CLI
namespace tool = daylight::views::cli;
#else
namespace tool = daylight::cli;
#endif

namespace {

using Clk = std::chrono::steady_clock;
double seconds_since(Clk::time_point t0) { return std::chrono::duration<double>(Clk::now() - t0).count(); }

int usage_error(const std::string& msg) {
  std::fprintf(stderr, "error: %s\n%s", msg.c_str(), tool::kUsage);
  return 1;
}
int runtime_error(const std::string& msg) {
  std::fprintf(stderr, "error: %s\n", msg.c_str());
  return 2;
}

std::string lowercase_extension(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  std::string e = dot == std::string::npos ? std::string() : path.substr(dot);
  std::transform(e.begin(), e.end(), e.begin(), [](unsigned char c) { return char(std::tolower(c)); });
  return e;
}

// A mesh file's meshes appended to `into` with the given role.
bool load_meshes(const std::string& path, bool sensor, bool verbose, std::vector<SimulationMesh>& into, std::string& err) {
  const auto t0 = Clk::now();
  const size_t before = into.size();
  const std::string e = lowercase_extension(path);
  bool ok = false;
  if (e == ".obj") ok = read_obj(path, into, err);
  else if (e == ".dmesh") ok = read_dmesh(path, into, err);
  else if (e == ".usd" || e == ".usda" || e == ".usdc" || e == ".usdz") err = path + ": USD is not read by this library; export .obj or .dmesh";
  else err = path + ": unknown extension; .obj or .dmesh";
  if (!ok) return false;
  size_t faces = 0, points = 0;
  for (size_t k = before; k < into.size(); ++k) {
    into[k].sensor = sensor;
    faces += into[k].numFaces();
    points += into[k].numPoints();
  }
  if (verbose)
    std::fprintf(stderr, "info: read %s: %zu meshes, %zu faces, %zu points, %.3f s\n", path.c_str(), into.size() - before, faces, points, seconds_since(t0));
  return true;
}

// ------------------------------------------------------------------------------------------------ --bench scenes
// Appends boxes and grids of quads to a mesh.
struct MeshWriter {
  SimulationMesh& m;
  // the box [lo, hi] facing out: 6 quads, or 12 triangles (each quad (0 1 2) (0 2 3)); corner k has the hi side on
  // axis a when bit a of k is set
  void box(const double lo[3], const double hi[3], bool triangles) {
    static const int kFaces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}};
    const int base = int(m.numPoints());
    for (int k = 0; k < 8; ++k)
      for (int a = 0; a < 3; ++a) m.points.push_back(((k >> a) & 1) ? hi[a] : lo[a]);
    for (const auto& q : kFaces) {
      const int c[4] = {base + q[0], base + q[1], base + q[2], base + q[3]};
      if (triangles) {
        m.counts.insert(m.counts.end(), {3, 3});
        m.indices.insert(m.indices.end(), {c[0], c[1], c[2], c[0], c[2], c[3]});
      } else {
        m.counts.push_back(4);
        m.indices.insert(m.indices.end(), c, c + 4);
      }
    }
  }
  // nu x nv quads over o + s u + t v (s, t in [0, 1]), facing u x v; points row by row along u
  void grid(const double o[3], const double u[3], const double v[3], int nu, int nv) {
    const int base = int(m.numPoints());
    for (int j = 0; j <= nv; ++j)
      for (int i = 0; i <= nu; ++i)
        for (int a = 0; a < 3; ++a) m.points.push_back(o[a] + u[a] * i / nu + v[a] * j / nv);
    for (int j = 0; j < nv; ++j)
      for (int i = 0; i < nu; ++i) {
        const int p = base + j * (nu + 1) + i;
        m.counts.push_back(4);
        m.indices.insert(m.indices.end(), {p, p + 1, p + nu + 2, p + nu + 1});
      }
  }
};

SimulationMesh named(const char* name, bool sensor) {
  SimulationMesh m;
  m.name = name;
  m.sensor = sensor;
  return m;
}

#ifdef VIEWS_CLI
// The views benchmark city (reference main.cpp:91-146): B x B blocks of 40 m between 20 m streets, a building of height
// U(10, 60) m per block (mt19937 seed 42) with its four facades cut into ~square quads (the ~nSens receivers), every
// 8th block a park (a lawn 5 cm up, 16 trees), street trees every 15 m, a 50 m water strip south, the roofs and a ground
// grid (filling the scene to ~nTris triangles) as context; targets park and water.
void views_bench_city(size_t nSens, size_t nTris, std::vector<SimulationMesh>& meshes, std::vector<views::Target>& targets) {
  const int B = std::max(1, int(std::ceil(std::sqrt(double(nSens) / 1225))));
  const double side = 60.0 * B + 20;
  std::mt19937 rng(42);
  std::uniform_real_distribution<double> storey(10, 60);
  std::vector<double> height(size_t(B) * B, 0.0);   // 0: a park
  double facadeArea = 0;
  for (size_t k = 0; k < height.size(); ++k)
    if (k % 8 != 7) {
      height[k] = storey(rng);
      facadeArea += 160 * height[k];
    }
  const double quad = std::sqrt(facadeArea / double(std::max<size_t>(nSens, 1)));   // the receivers' edge
  SimulationMesh facades = named("facades", true), roofs = named("roofs", false), park = named("park", false), trees = named("trees", false),
                 water = named("water", false), ground = named("ground", false);
  MeshWriter F{facades}, R{roofs}, Pk{park}, Tr{trees}, Wa{water}, Gr{ground};
  for (int bj = 0; bj < B; ++bj)
    for (int bi = 0; bi < B; ++bi) {
      const double x0 = 20 + 60.0 * bi, y0 = 20 + 60.0 * bj, H = height[size_t(bj) * B + bi];
      if (H == 0) {
        const double o[3] = {x0, y0, 0.05}, u[3] = {40, 0, 0}, v[3] = {0, 40, 0};
        Pk.grid(o, u, v, 1, 1);
        for (int j = 0; j < 4; ++j)
          for (int i = 0; i < 4; ++i) {
            const double lo[3] = {x0 + 3 + 10 * i, y0 + 3 + 10 * j, 2}, hi[3] = {x0 + 7 + 10 * i, y0 + 7 + 10 * j, 8};
            Tr.box(lo, hi, false);
          }
        continue;
      }
      const int cols = std::max(1, int(std::lround(40 / quad))), rows = std::max(1, int(std::lround(H / quad)));
      const double corner[5][2] = {{x0, y0}, {x0 + 40, y0}, {x0 + 40, y0 + 40}, {x0, y0 + 40}, {x0, y0}};
      for (int w = 0; w < 4; ++w) {
        const double o[3] = {corner[w][0], corner[w][1], 0}, u[3] = {corner[w + 1][0] - corner[w][0], corner[w + 1][1] - corner[w][1], 0},
                     v[3] = {0, 0, H};
        F.grid(o, u, v, cols, rows);
      }
      const double o[3] = {x0, y0, H}, u[3] = {40, 0, 0}, v[3] = {0, 40, 0};
      R.grid(o, u, v, 1, 1);
    }
  for (int s = 0; s <= B; ++s)   // street trees along both axes of every street
    for (double t = 7.5; t + 4 < side; t += 15) {
      const double m = 60.0 * s + 10;
      const double a0[3] = {t, m - 2, 2}, a1[3] = {t + 4, m + 2, 8}, b0[3] = {m - 2, t, 2}, b1[3] = {m + 2, t + 4, 8};
      Tr.box(a0, a1, false);
      Tr.box(b0, b1, false);
    }
  const double wo[3] = {0, -60, 0.05}, wu[3] = {side, 0, 0}, wv[3] = {0, 50, 0};
  Wa.grid(wo, wu, wv, 1, 1);
  size_t tris = 0;
  for (const SimulationMesh* m : {&facades, &roofs, &park, &trees, &water}) tris += 2 * m->numFaces();
  const int G = std::max(1, int(std::sqrt(double(nTris > tris ? nTris - tris : 0) / 2)));
  const double go[3] = {-100, -160, 0}, gu[3] = {side + 200, 0, 0}, gv[3] = {0, side + 260, 0};
  Gr.grid(go, gu, gv, G, G);
  if (park.numFaces()) targets.push_back({"park", "park"});
  targets.push_back({"water", "water"});
  for (SimulationMesh* m : {&facades, &roofs, &park, &trees, &water, &ground})
    if (m->numFaces()) meshes.push_back(std::move(*m));
  size_t total = 0;
  for (const SimulationMesh& m : meshes) total += 2 * m.numFaces();
  std::fprintf(stderr, "bench: %zu facade sensors, %zu triangles, %d x %d blocks\n", meshes[0].numFaces(), total, B, B);
}
#else
// The daylight benchmark scene (reference main.cpp:148-192): nSens quads of a flat 1 km grid and nTris / 12 boxes
// (footprint U(5, 30) m, height U(5, 60) m, 0.1 m up, mt19937 seed 42) of 12 triangles as context, faces in 16 optics
// slots.
void daylight_bench_scene(size_t nSens, size_t nTris, std::vector<SimulationMesh>& meshes) {
  const size_t side = size_t(std::ceil(std::sqrt(double(nSens))));
  const double extent = 1000.0, cell = extent / double(side);
  SimulationMesh grid = named("bench_grid", true);
  for (size_t j = 0; j <= side; ++j)
    for (size_t i = 0; i <= side; ++i) grid.points.insert(grid.points.end(), {-extent / 2 + cell * double(i), -extent / 2 + cell * double(j), 0.0});
  const int row = int(side + 1);
  for (size_t j = 0; j < side && grid.counts.size() < nSens; ++j)
    for (size_t i = 0; i < side && grid.counts.size() < nSens; ++i) {
      const int a = int(j * (side + 1) + i);
      grid.counts.push_back(4);
      grid.indices.insert(grid.indices.end(), {a, a + 1, a + row + 1, a + row});
    }
  meshes.push_back(std::move(grid));
  SimulationMesh boxes = named("bench_boxes", false);
  MeshWriter B{boxes};
  std::mt19937 rng(42);
  std::uniform_real_distribution<double> at(-extent / 2, extent / 2), foot(5.0, 30.0), tall(5.0, 60.0);
  for (size_t b = 0; b < nTris / 12; ++b) {
    const double x0 = at(rng), y0 = at(rng), wx = foot(rng), wy = foot(rng), h = tall(rng);   // in the reference's draw order
    const double lo[3] = {x0, y0, 0.1}, hi[3] = {x0 + wx, y0 + wy, 0.1 + h};
    B.box(lo, hi, true);
  }
  boxes.slot.resize(boxes.counts.size());
  for (size_t f = 0; f < boxes.slot.size(); ++f) boxes.slot[f] = uint16_t(f % 16);
  meshes.push_back(std::move(boxes));
}
#endif

// ------------------------------------------------------------------------------------------------ the values
// A buffered writer of text rows: each value as "%.10g", columns space-separated.
class RowWriter {
 public:
  explicit RowWriter(std::FILE* f) : f_(f), buf_(size_t(1) << 16) {}
  bool row(const double* v, size_t n) {
    for (size_t c = 0; c < n; ++c) {
      if (buf_.size() - at_ < 40 && !flush()) return false;
      if (c) buf_[at_++] = ' ';
      at_ = size_t(std::to_chars(buf_.data() + at_, buf_.data() + buf_.size(), v[c], std::chars_format::general, 10).ptr - buf_.data());
    }
    if (buf_.size() - at_ < 2 && !flush()) return false;
    buf_[at_++] = '\n';
    return true;
  }
  bool flush() {
    const bool ok = std::fwrite(buf_.data(), 1, at_, f_) == at_;
    at_ = 0;
    return ok;
  }

 private:
  std::FILE* f_;
  std::vector<char> buf_;
  size_t at_ = 0;
};

int write_values(const Result& res, const std::string& out, const std::string& manifest, bool verbose) {
  const auto t0 = Clk::now();
  const bool npy = lowercase_extension(out) == ".npy";
  std::FILE* f = out.empty() ? stdout : std::fopen(out.c_str(), npy ? "wb" : "w");
  if (!f) return runtime_error("cannot write " + out);
  const size_t C = res.columns.size(), N = C ? res.values.size() / C : 0;
  bool ok = true;
  if (npy) ok = npy_write_header(f, N, C, "<f8") && std::fwrite(res.values.data(), 8, res.values.size(), f) == res.values.size();
  else {
    RowWriter w(f);
    for (size_t e = 0; e < N && ok; ++e) ok = w.row(res.values.data() + e * C, C);
    ok = ok && w.flush();
  }
  if (f == stdout) std::fflush(f);
  else {
    ok = std::fclose(f) == 0 && ok;
    if (ok) {   // the manifest beside the file
      std::FILE* m = std::fopen((out + ".json").c_str(), "w");
      ok = m && std::fwrite(manifest.data(), 1, manifest.size(), m) == manifest.size();
      if (m) ok = std::fclose(m) == 0 && ok;
    }
  }
  if (!ok) return runtime_error("cannot write " + out);
  if (verbose) std::fprintf(stderr, "info: output %.3f s\n", seconds_since(t0));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  tool::Args a;
  std::string err;
  if (tool::parse(argc, argv, a, err)) return usage_error(err);
  if (a.help) { std::fputs(tool::kUsage, stdout); return 0; }
  if (a.sensors.empty() && !a.bench) return usage_error("no mesh given");
#ifndef VIEWS_CLI
  Weather weather;
  bool haveWeather = false;
  if (const int rc = cli::prepare(a, weather, haveWeather, err)) return rc == 1 ? usage_error(err) : runtime_error(err);
#endif
  std::vector<SimulationMesh> meshes;
  if (a.bench) {
#ifdef VIEWS_CLI
    views_bench_city(a.benchSens, a.benchTris, meshes, a.opt.targets);
#else
    daylight_bench_scene(a.benchSens, a.benchTris, meshes);
    if (haveWeather) a.opt.quantities = {Quantity::Radiation, Quantity::SunHours};
    else {   // no weather: sun hours at Boston unless a site is given
      a.opt.quantities = {Quantity::SunHours};
      if (std::isnan(a.opt.lat)) { a.opt.lat = 42.35; a.opt.lon = -71.07; a.opt.tz = -5; }
    }
    std::fprintf(stderr, "bench: %zu sensors, %zu triangles\n", a.benchSens, a.benchTris);
#endif
    a.opt.verbose = true;
  } else {
    for (const std::string& p : a.sensors)
      if (!load_meshes(p, true, a.opt.verbose, meshes, err)) return runtime_error(err);
    for (const std::string& p : a.context)
      if (!load_meshes(p, false, a.opt.verbose, meshes, err)) return runtime_error(err);
  }
  const std::vector<SimulationMeshView> mv = views_of(meshes);
  std::vector<char> open;
  std::vector<SlotOptics> slots;
  if (!resolve_optics(mv, a.opt.optics, slots, err) || !resolve_no_occlude(mv, a.opt.noOcclude, open, err)) return usage_error(err);
#ifdef VIEWS_CLI
  std::vector<std::string> groups;
  std::vector<uint8_t> cls;
  if (!resolve_targets(mv, a.opt.targets, a.opt.selfOcclude, open, groups, cls, err)) return usage_error(err);
#endif
  const auto t0 = Clk::now();
  Result res;
#ifdef VIEWS_CLI
  if (!views::run(mv, a.opt, res, err)) return runtime_error(err);
  const char* what = "scene, GPU, columns";
#else
  if (!run(mv, haveWeather ? &weather : nullptr, a.opt, res, err)) return runtime_error(err);
  const char* what = "sky model, GPU, output";
#endif
  if (a.opt.verbose) std::fprintf(stderr, "info: run %.2f s total (incl. %s)\n", seconds_since(t0), what);
  if (a.bench) return 0;
#ifdef VIEWS_CLI
  views::Field f;
  views::field(mv, a.opt, f, err);
  return write_values(res, a.out, tool::manifest_json(res, a.opt, f), a.opt.verbose);
#else
  return write_values(res, a.out, tool::manifest_json(res), a.opt.verbose);
#endif
}
