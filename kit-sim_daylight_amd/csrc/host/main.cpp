// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The command-line tools (reference src/main.cpp): `daylight` (meshes + place + time in, one value row per element out)
// and, built with VIEWS_CLI, `views` (meshes + field of view in). Flags, rules and manifests are cli.cpp's; the runs are
// daylight::run / views::run on the GPU. Outputs: stdout or --out FILE.txt ("%.10g", one row per element, columns
// space-separated) or FILE.npy (float64 [N x C]), with FILE.json the manifest beside a file.
// --bench NSENS NTRIS runs the reference's synthetic benchmark scenes (a sensor grid among random boxes; views: a city
// of box buildings with facade sensors, parks, street trees and water), built as the reference describes them, so that
// its published throughputs compare with this build's on the same kind of scene.
#include "glibc_guard.h"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iterator>
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

int usage_error(const std::string& msg) {
  std::fprintf(stderr, "error: %s\n%s", msg.c_str(), tool::kUsage);
  return 1;
}

std::string ext_of(const std::string& path) {
  std::string e = path.substr(std::min(path.find_last_of('.'), path.size()));
  for (char& c : e) c = char(std::tolower(static_cast<unsigned char>(c)));
  return e;
}

double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool load(const std::string& path, bool sensor, bool verbose, std::vector<SimulationMesh>& meshes, std::string& err) {
  const size_t first = meshes.size();
  const std::string e = ext_of(path);
  const auto t0 = std::chrono::steady_clock::now();
  bool ok = false;
  if (e == ".obj") ok = read_obj(path, meshes, err);
  else if (e == ".dmesh") ok = read_dmesh(path, meshes, err);
  else if (e == ".usd" || e == ".usda" || e == ".usdc" || e == ".usdz") err = path + ": USD is not read by this build; export .obj or .dmesh";
  else err = path + ": unknown extension; .obj or .dmesh";
  if (!ok) return false;
  size_t faces = 0, points = 0;
  for (size_t k = first; k < meshes.size(); ++k) {
    meshes[k].sensor = sensor;
    faces += meshes[k].numFaces();
    points += meshes[k].numPoints();
  }
  if (verbose)
    std::fprintf(stderr, "info: read %s: %zu meshes, %zu faces, %zu points, %.3f s\n", path.c_str(), meshes.size() - first, faces, points,
                 secs_since(t0));
  return true;
}

// quads of a box [lo, hi] facing out, or its 12 triangles
void add_box(SimulationMesh& m, const double lo[3], const double hi[3], bool triangles) {
  const int b = int(m.numPoints());
  for (int k = 0; k < 8; ++k)
    for (int a = 0; a < 3; ++a) m.points.push_back(((k >> a) & 1) ? hi[a] : lo[a]);
  // corner k: bit a set = the hi side on axis a
  static const int q[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}};
  for (const auto& f : q) {
    if (triangles) {
      m.counts.insert(m.counts.end(), {3, 3});
      m.indices.insert(m.indices.end(), {b + f[0], b + f[1], b + f[2], b + f[0], b + f[2], b + f[3]});
    } else {
      m.counts.push_back(4);
      m.indices.insert(m.indices.end(), {b + f[0], b + f[1], b + f[2], b + f[3]});
    }
  }
}

// nu x nv quads over o + s u + t v, normal u x v
void add_grid(SimulationMesh& m, const double o[3], const double u[3], const double v[3], int nu, int nv) {
  const int b = int(m.numPoints());
  for (int j = 0; j <= nv; ++j)
    for (int i = 0; i <= nu; ++i)
      for (int a = 0; a < 3; ++a) m.points.push_back(o[a] + u[a] * i / nu + v[a] * j / nv);
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nu; ++i) {
      const int p = b + j * (nu + 1) + i;
      m.counts.push_back(4);
      m.indices.insert(m.indices.end(), {p, p + 1, p + nu + 2, p + nu + 1});
    }
}

#ifdef VIEWS_CLI
// The views benchmark city (reference main.cpp:91-146): B x B blocks of 40 m between 20 m streets, a box building of
// U(10, 60) m per block (mt19937 seed 42) whose facades, cut into quads, are the ~nSens receivers; every 8th block a park
// (lawn 5 cm up, trees on a 10 m grid); street trees every 15 m; a 50 m water strip along the south; roofs and a ground
// grid (filling the scene up to ~nTris triangles) as context; targets park and water.
void bench_city(size_t nSens, size_t nTris, std::vector<SimulationMesh>& meshes, std::vector<views::Target>& targets) {
  const int B = std::max(1, int(std::ceil(std::sqrt(double(nSens) / 1225))));
  const double W = 60.0 * B + 20;
  std::mt19937 rng(42);
  std::uniform_real_distribution<double> height(10, 60);
  std::vector<double> h(size_t(B) * B, 0.0);
  double area = 0;
  for (size_t k = 0; k < h.size(); ++k)
    if (k % 8 != 7) { h[k] = height(rng); area += 160 * h[k]; }
  const double q = std::sqrt(area / double(std::max<size_t>(nSens, 1)));
  SimulationMesh facades, roofs, park, trees, water, ground;
  facades.name = "facades"; roofs.name = "roofs"; park.name = "park"; trees.name = "trees"; water.name = "water"; ground.name = "ground";
  roofs.sensor = park.sensor = trees.sensor = water.sensor = ground.sensor = false;
  for (int bj = 0; bj < B; ++bj)
    for (int bi = 0; bi < B; ++bi) {
      const double x0 = 20 + 60.0 * bi, y0 = 20 + 60.0 * bj, H = h[size_t(bj) * B + bi];
      if (H == 0) {
        const double o[3] = {x0, y0, 0.05}, u[3] = {40, 0, 0}, v[3] = {0, 40, 0};
        add_grid(park, o, u, v, 1, 1);
        for (int j = 0; j < 4; ++j)
          for (int i = 0; i < 4; ++i) {
            const double lo[3] = {x0 + 3 + 10 * i, y0 + 3 + 10 * j, 2}, hi[3] = {x0 + 7 + 10 * i, y0 + 7 + 10 * j, 8};
            add_box(trees, lo, hi, false);
          }
        continue;
      }
      const int cols = std::max(1, int(std::lround(40 / q))), rows = std::max(1, int(std::lround(H / q)));
      const double c[5][2] = {{x0, y0}, {x0 + 40, y0}, {x0 + 40, y0 + 40}, {x0, y0 + 40}, {x0, y0}};
      for (int w = 0; w < 4; ++w) {
        const double o[3] = {c[w][0], c[w][1], 0}, u[3] = {c[w + 1][0] - c[w][0], c[w + 1][1] - c[w][1], 0}, v[3] = {0, 0, H};
        add_grid(facades, o, u, v, cols, rows);
      }
      const double o[3] = {x0, y0, H}, u[3] = {40, 0, 0}, v[3] = {0, 40, 0};
      add_grid(roofs, o, u, v, 1, 1);
    }
  for (int s = 0; s <= B; ++s)
    for (double t = 7.5; t + 4 < W; t += 15) {
      const double m = 60.0 * s + 10;
      const double a0[3] = {t, m - 2, 2}, a1[3] = {t + 4, m + 2, 8}, b0[3] = {m - 2, t, 2}, b1[3] = {m + 2, t + 4, 8};
      add_box(trees, a0, a1, false);
      add_box(trees, b0, b1, false);
    }
  const double wo[3] = {0, -60, 0.05}, wu[3] = {W, 0, 0}, wv[3] = {0, 50, 0};
  add_grid(water, wo, wu, wv, 1, 1);
  size_t made = 0;
  for (const SimulationMesh* m : {&facades, &roofs, &park, &trees, &water}) made += 2 * m->numFaces();
  const int G = std::max(1, int(std::sqrt(double(nTris > made ? nTris - made : 0) / 2)));
  const double go[3] = {-100, -160, 0}, gu[3] = {W + 200, 0, 0}, gv[3] = {0, W + 260, 0};
  add_grid(ground, go, gu, gv, G, G);
  if (park.numFaces()) targets.push_back({"park", "park"});
  targets.push_back({"water", "water"});
  for (SimulationMesh* m : {&facades, &roofs, &park, &trees, &water, &ground})
    if (m->numFaces()) meshes.push_back(std::move(*m));
  size_t tris = 0;
  for (const SimulationMesh& m : meshes) tris += 2 * m.numFaces();
  std::fprintf(stderr, "bench: %zu facade sensors, %zu triangles, %d x %d blocks\n", meshes[0].numFaces(), tris, B, B);
}
#else
// The daylight benchmark scene (reference main.cpp:148-192): nSens quads on a flat 1 km grid and nTris / 12 random boxes
// (footprint U(5, 30) m, height U(5, 60) m, 0.1 m up; mt19937 seed 42) as 12-triangle context, faces in 16 optics slots.
void bench_scene(size_t nSens, size_t nTris, std::vector<SimulationMesh>& meshes) {
  const size_t side = size_t(std::ceil(std::sqrt(double(nSens))));
  const double extent = 1000.0, cell = extent / double(side);
  SimulationMesh grid;
  grid.name = "bench_grid";
  for (size_t j = 0; j <= side; ++j)
    for (size_t i = 0; i <= side; ++i) grid.points.insert(grid.points.end(), {-extent / 2 + cell * double(i), -extent / 2 + cell * double(j), 0.0});
  const int w = int(side + 1);
  for (size_t j = 0; j < side && grid.counts.size() < nSens; ++j)
    for (size_t i = 0; i < side && grid.counts.size() < nSens; ++i) {
      const int a = int(j * (side + 1) + i);
      grid.counts.push_back(4);
      grid.indices.insert(grid.indices.end(), {a, a + 1, a + w + 1, a + w});
    }
  meshes.push_back(std::move(grid));
  SimulationMesh boxes;
  boxes.name = "bench_boxes";
  boxes.sensor = false;
  std::mt19937 rng(42);
  std::uniform_real_distribution<double> pos(-extent / 2, extent / 2), foot(5.0, 30.0), height(5.0, 60.0);
  for (size_t b = 0; b < nTris / 12; ++b) {
    const double x0 = pos(rng), y0 = pos(rng), wx = foot(rng), wy = foot(rng), hz = height(rng);
    const double lo[3] = {x0, y0, 0.1}, hi[3] = {x0 + wx, y0 + wy, 0.1 + hz};
    add_box(boxes, lo, hi, true);
  }
  for (size_t f = 0; f < boxes.counts.size(); ++f) boxes.slot.push_back(uint16_t(f % 16));
  meshes.push_back(std::move(boxes));
}
#endif

// the values: stdout or FILE.txt rows of "%.10g" columns, or FILE.npy float64 [N x C]; FILE.json the manifest
int write_values(const Result& res, const std::string& out, const std::string& manifest, bool verbose) {
  const auto t0 = std::chrono::steady_clock::now();
  const bool npy = ext_of(out) == ".npy";
  std::FILE* f = out.empty() ? stdout : std::fopen(out.c_str(), npy ? "wb" : "w");
  if (!f) { std::fprintf(stderr, "error: cannot write %s\n", out.c_str()); return 2; }
  const size_t C = res.columns.size(), N = C ? res.values.size() / C : 0;
  bool ok = true;
  if (npy) ok = npy_write_header(f, N, C, "<f8") && std::fwrite(res.values.data(), 8, res.values.size(), f) == res.values.size();
  else {
    std::vector<char> buf(size_t(1) << 16);
    char* p = buf.data();
    char* const end = buf.data() + buf.size();
    for (size_t e = 0; e < N && ok; ++e) {
      for (size_t c = 0; c < C; ++c) {
        if (end - p < 40) { ok = std::fwrite(buf.data(), 1, size_t(p - buf.data()), f) == size_t(p - buf.data()); p = buf.data(); }
        if (c) *p++ = ' ';
        p = std::to_chars(p, end, res.values[e * C + c], std::chars_format::general, 10).ptr;   // as "%.10g"
      }
      *p++ = '\n';
    }
    ok = ok && std::fwrite(buf.data(), 1, size_t(p - buf.data()), f) == size_t(p - buf.data());
  }
  if (f == stdout) {
    std::fflush(f);
  } else {
    ok = std::fclose(f) == 0 && ok;
    if (ok) {
      std::FILE* m = std::fopen((out + ".json").c_str(), "w");
      ok = m && std::fwrite(manifest.data(), 1, manifest.size(), m) == manifest.size();
      if (m) ok = std::fclose(m) == 0 && ok;
    }
  }
  if (!ok) { std::fprintf(stderr, "error: cannot write %s\n", out.c_str()); return 2; }
  if (verbose) std::fprintf(stderr, "info: output %.3f s\n", secs_since(t0));
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
  if (const int rc = cli::prepare(a, weather, haveWeather, err)) {
    if (rc == 1) return usage_error(err);
    std::fprintf(stderr, "error: %s\n", err.c_str());
    return rc;
  }
#endif
  std::vector<SimulationMesh> meshes;
  if (a.bench) {
#ifdef VIEWS_CLI
    bench_city(a.benchSens, a.benchTris, meshes, a.opt.targets);
#else
    bench_scene(a.benchSens, a.benchTris, meshes);
    if (!haveWeather) {
      a.opt.quantities = {Quantity::SunHours};
      if (std::isnan(a.opt.lat)) { a.opt.lat = 42.35; a.opt.lon = -71.07; a.opt.tz = -5; }
    } else a.opt.quantities = {Quantity::Radiation, Quantity::SunHours};
    std::fprintf(stderr, "bench: %zu sensors, %zu triangles\n", a.benchSens, a.benchTris);
#endif
    a.opt.verbose = true;
  } else {
    for (const std::string& p : a.sensors)
      if (!load(p, true, a.opt.verbose, meshes, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 2; }
    for (const std::string& p : a.context)
      if (!load(p, false, a.opt.verbose, meshes, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 2; }
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
  const auto t0 = std::chrono::steady_clock::now();
  Result res;
#ifdef VIEWS_CLI
  if (!views::run(mv, a.opt, res, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 2; }
  const char* what = "scene, GPU, columns";
#else
  if (!run(mv, haveWeather ? &weather : nullptr, a.opt, res, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 2; }
  const char* what = "sky model, GPU, output";
#endif
  if (a.opt.verbose) std::fprintf(stderr, "info: run %.2f s total (incl. %s)\n", secs_since(t0), what);
  if (a.bench) return 0;
#ifdef VIEWS_CLI
  views::Field f;
  views::field(mv, a.opt, f, err);
  return write_values(res, a.out, tool::manifest_json(res, a.opt, f), a.opt.verbose);
#else
  return write_values(res, a.out, tool::manifest_json(res), a.opt.verbose);
#endif
}
