// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The options that name meshes, resolved against a run's meshes: --optics rows per global slot (precedence * < NAME <
// NAME/SLOT), --no-occlude, --target. The ranges, the precedence, the global slot numbering (every mesh's slots in
// input order) and the messages are reference src/mesh.cpp's (optics_rows_ok, optics_ok, resolve_optics,
// resolve_no_occlude, resolve_targets).
#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "core.h"
#include "dl_math.h"

namespace daylight {
namespace {

// The meshes of a run by name; a name borne by several meshes is ambiguous.
class MeshNames {
 public:
  explicit MeshNames(const std::vector<SimulationMeshView>& meshes) {
    for (size_t k = 0; k < meshes.size(); ++k) {
      const auto [it, fresh] = at_.emplace(meshes[k].name, k);
      if (!fresh) it->second = kAmbiguous;
    }
  }
  // the mesh named `name`, or SIZE_MAX with err = prefix + why not
  size_t find(const std::string& name, const std::string& prefix, std::string& err) const {
    const auto it = at_.find(name);
    if (it == at_.end()) err = prefix + "no mesh of that name";
    else if (it->second == kAmbiguous) err = prefix + "several meshes have that name";
    else return it->second;
    return SIZE_MAX;
  }

 private:
  static constexpr size_t kAmbiguous = SIZE_MAX;
  std::unordered_map<std::string, size_t> at_;
};

std::string optics_label(const Optics& o) { return "--optics " + o.mesh + (o.slot >= 0 ? "/" + std::to_string(o.slot) : "") + ": "; }

bool ends_with(const std::string& s, const char* tail) {
  const size_t n = std::strlen(tail);
  return s.size() >= n && s.compare(s.size() - n, n, tail) == 0;
}

}  // namespace

bool optics_rows_ok(const std::vector<Optics>& rows, std::string& err) {
  for (const Optics& o : rows) {
    const auto bad = [&](const char* what) { err = optics_label(o) + what; return false; };
    if (o.mesh.empty() || o.slot < -1 || o.slot > 65535 || (o.mesh == "*" && o.slot >= 0)) return bad("NAME (or *) with an optional /SLOT 0..65535");
    if (!(o.rho >= 0 && o.rho <= 1 && o.tau >= 0 && o.tau <= 1)) return bad("RHO and TAU must be in 0..1");
    if (o.rho + o.tau > 1 + 1e-9) return bad("RHO + TAU must not exceed 1");
    if (o.glass && o.rho != 0) return bad("glass reflects by angle: RHO must be 0");
    if (o.glass && !(o.tau > 0 && o.tau <= 0.918)) return bad("glass TAU must be in (0, 0.918] (n 1.52 passes at most 0.918)");
  }
  return true;
}

bool optics_ok(const Options& opt, std::string& err) {
  if (!optics_rows_ok(opt.optics, err)) return false;
  const bool ladybug = opt.mode == Mode::Ladybug;
  const auto no = [&](const char* msg) { err = msg; return false; };
  if (opt.transmission && ladybug) return no("--transmission on needs --mode physical (ladybug: every surface opaque and black)");
  if (opt.bounces < 0 || opt.bounces > 20) return no("--bounces must be 0..20");
  if (opt.bounceSamples < 64 || opt.bounceSamples > 65536) return no("--bounce-samples must be 64..65536");
  if (opt.bounces > 0 && ladybug) return no("--bounces needs --mode physical (ladybug: every surface opaque and black)");
  return true;
}

bool resolve_optics(const std::vector<SimulationMeshView>& meshes, const std::vector<Optics>& rows, std::vector<SlotOptics>& out, std::string& err) {
  // the global slots: mesh k's slots 0 .. slot_count - 1 at first[k]
  std::vector<size_t> first(meshes.size() + 1, 0);
  for (size_t k = 0; k < meshes.size(); ++k) {
    const SimulationMesh& m = *meshes[k].mesh;
    if (!m.slot.empty() && m.slot.size() != m.numFaces()) {
      err = "mesh " + meshes[k].name + ": " + std::to_string(m.slot.size()) + " slots for " + std::to_string(m.numFaces()) + " faces";
      return false;
    }
    first[k + 1] = first[k] + slot_count(m);
  }
  // a row may name a mesh (or a mesh's slot) once
  std::vector<std::pair<std::string, int>> named;
  for (const Optics& o : rows) named.emplace_back(o.mesh, o.slot);
  std::sort(named.begin(), named.end());
  if (const auto twice = std::adjacent_find(named.begin(), named.end()); twice != named.end()) {
    err = "--optics " + twice->first + (twice->second >= 0 ? "/" + std::to_string(twice->second) : "") + " given twice";
    return false;
  }
  const MeshNames names(meshes);
  out.assign(first.back(), SlotOptics{});
  // rows in rising precedence, later ones overwriting: * (every slot), NAME (its mesh's), NAME/SLOT (one)
  const auto rank = [](const Optics& o) { return o.mesh == "*" ? 0 : o.slot < 0 ? 1 : 2; };
  for (int level = 0; level < 3; ++level)
    for (const Optics& o : rows) {
      if (rank(o) != level) continue;
      size_t b = 0, e = out.size();
      if (level > 0) {
        const size_t k = names.find(o.mesh, optics_label(o), err);
        if (k == SIZE_MAX) return false;
        b = first[k];
        e = first[k + 1];
        if (level == 2) {
          if (size_t(o.slot) >= e - b) { err = optics_label(o) + "slot out of range (the mesh has slots 0.." + std::to_string(e - b - 1) + ")"; return false; }
          b += size_t(o.slot);
          e = b + 1;
        }
      }
      const SlotOptics so{float(o.rho), float(o.tau), o.glass ? float(dlm::tn_from_Tn(o.tau)) : 0.f, o.glass ? DLG_SLOT_GLASS : 0u};
      std::fill(out.begin() + std::ptrdiff_t(b), out.begin() + std::ptrdiff_t(e), so);
    }
  return true;
}

bool resolve_no_occlude(const std::vector<SimulationMeshView>& meshes, const std::vector<std::string>& names, std::vector<char>& open, std::string& err) {
  open.assign(meshes.size(), 0);
  if (names.empty()) return true;
  const MeshNames index(meshes);
  for (const std::string& n : names) {
    const std::string prefix = "--no-occlude " + n + ": ";
    const size_t k = index.find(n, prefix, err);
    if (k == SIZE_MAX) return false;
    if (!meshes[k].sensor) { err = prefix + "a context mesh (leave it out to drop its shade)"; return false; }
    open[k] = 1;
  }
  return true;
}

bool resolve_targets(const std::vector<SimulationMeshView>& meshes, const std::vector<views::Target>& targets, bool selfOcclude,
                     const std::vector<char>& open, std::vector<std::string>& groups, std::vector<uint8_t>& cls, std::string& err) {
  groups.clear();
  cls.assign(meshes.size(), 2);
  if (targets.empty()) return true;
  const MeshNames index(meshes);
  static const char* const kWordChars = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789_-";
  for (const views::Target& t : targets) {
    const std::string prefix = "--target " + t.mesh + "=" + t.group + ": ";
    const auto bad = [&](const std::string& what) { err = prefix + what; return false; };
    const std::string& g = t.group;
    const size_t slash = t.mesh.rfind('/');
    if (slash != std::string::npos && slash + 1 < t.mesh.size() && t.mesh.find_first_not_of("0123456789", slash + 1) == std::string::npos)
      return bad("face-slot targets (NAME/SLOT) are not supported, target the whole mesh");
    if (g.empty() || g.find_first_not_of(kWordChars) != std::string::npos) return bad("GROUP takes letters, digits, _ and - only");
    const bool reserved = g == "sky" || g == "ground" || g == "obstruction" || g == "glazing" || g == "volume" || g == "area" ||
                          g.rfind("distance", 0) == 0 || g.rfind("reach", 0) == 0 || ends_with(g, "_distance") || ends_with(g, "_distance_min");
    if (reserved) return bad("GROUP is reserved (sky, ground, obstruction, glazing, volume, area, distance*, reach*, *_distance, *_distance_min)");
    const size_t k = index.find(t.mesh, prefix, err);
    if (k == SIZE_MAX) return false;
    if (cls[k] != 2) return bad("the mesh is a target twice");
    if (meshes[k].sensor && (!selfOcclude || open[k])) return bad("the mesh blocks no rays (--no-self-occlude or --no-occlude): none can end on it");
    size_t gi = size_t(std::find(groups.begin(), groups.end(), g) - groups.begin());
    if (gi == groups.size()) {
      if (groups.size() == 16) return bad("more than 16 target groups");
      groups.push_back(g);
    }
    cls[k] = uint8_t(3 + gi);
  }
  return true;
}

}  // namespace daylight
