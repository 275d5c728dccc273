// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The command-line contract of both tools as data: each flag is a row of a table (its name, whether it takes a value,
// and what the value does to the options), scanned by one loop that both tools share. The flag names, value syntax,
// defaults, checks, error strings and exit codes are those of reference src/cli.cpp (cli.h's parse / prepare /
// manifest_json), which scripts and the C ABI rely on; the manifests are written byte for byte as there, since
// consumers parse them and the reference's C ABI test compares them with the CLI's.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <string_view>

#include "cli.h"
#include "core.h"

namespace daylight::cli {

const char* const kUsage =
    "usage: daylight <mesh>... [flags]            (MI355X build: HIP kernels on gfx950)\n"
    "  meshes: .obj (one o/g group per mesh) or .dmesh (raw arrays); world metres, Z up; positional = sensor meshes\n"
    "  --context FILE           occluder-only mesh (repeatable)\n"
    "  --on faces|vertices      sensor placement (default faces)\n"
    "  --mode physical|ladybug  default physical\n"
    "  --quantity Q[,Q...]      radiation|direct|diffuse|irradiance|illuminance|sunhours (default radiation)\n"
    "  --sides front|back|both  receiving side(s) of each element (default front)\n"
    "  --cull-backfaces         occluders block only from their front side (default two-sided)\n"
    "  --epw FILE               weather + location; --lat --lon override its location\n"
    "  --clear-sky              synthetic cloudless weather (Ineichen) at --lat --lon instead of --epw\n"
    "  --turbidity TL  --altitude M   for --clear-sky (defaults 3, 0)\n"
    "  --tz H                   clock of the timestamps (default: the EPW's zone, else round(lon / 15))\n"
    "  --time [YYYY-]MM-DDTHH:MM | --from TS --to TS [--step MIN]   default: annual over the EPW rows\n"
    "  --hours HH:MM-HH:MM      with --from/--to: only these hours of every day\n"
    "  --sky-density 1|2|4  --patch-samples K  --ground RHO  --north DEG  --offset M  --no-self-occlude\n"
    "  --no-occlude NAME[;NAME...]  sensor meshes analysed but casting no shade\n"
    "  --virtual-ground on|off  ground below the horizon reflecting --ground RHO (default on)\n"
    "  --optics SPEC[;SPEC...]  NAME[/SLOT]=RHO[,TAU[,glass|porous]]\n"
    "  --transmission off|on  --fresnel on|off  --bounces N  --bounce-samples M\n"
    "  --out FILE[.npy]         values (text rows, or float64 [N x C] NumPy when .npy) + FILE.json manifest\n"
    "  --timeseries FILE.npy  --dump-bits FILE.npy  --dump-sun FILE\n"
    "  --bench [NSENS NTRIS]  -v          exit code: 1 usage, 2 runtime error\n";

// ------------------------------------------------------------------------------------------------ value readers
namespace vals {

// s split at sep into its pieces (an empty s is one empty piece)
std::vector<std::string> pieces(std::string_view s, char sep) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    const size_t e = s.find(sep, b);
    out.emplace_back(s.substr(b, e == std::string_view::npos ? std::string_view::npos : e - b));
    if (e == std::string_view::npos) return out;
    b = e + 1;
  }
}

// the whole of s as a double (strtod's grammar); false on an empty or partly read s
bool strict(const std::string& s, double& x) {
  char* end = nullptr;
  x = std::strtod(s.c_str(), &end);
  return !s.empty() && !*end;
}

// every comma-separated piece of s as a number
bool list(const std::string& s, std::vector<double>& out) {
  out.clear();
  for (const std::string& p : pieces(s, ',')) {
    double x;
    const bool ok = strict(p, x);
    out.push_back(x);
    if (!ok) return false;
  }
  return true;
}

// v as one of the words (its index), or -1
int word(const char* v, std::initializer_list<const char*> words) {
  int k = 0;
  for (const char* w : words) {
    if (!std::strcmp(v, w)) return k;
    ++k;
  }
  return -1;
}

}  // namespace vals

// s as JSON string text: quote, backslash and control bytes escaped, other bytes as given
std::string json_str(const std::string& s) {
  std::string o(1, '"');
  for (const char ch : s) {
    if (ch == '"' || ch == '\\') o += '\\';
    if (static_cast<unsigned char>(ch) < 0x20) {
      char u[8];
      std::snprintf(u, sizeof u, "\\u%04x", unsigned(ch));
      o += u;
    } else o += ch;
  }
  return o + '"';
}

// the manifest's "meshes" array: {name, first row, row count} per sensor mesh, in input order
std::string manifest_meshes(const Result& r) {
  std::string m = "[";
  size_t first = 0;
  for (size_t i = 0; i < r.ranges.size(); ++i) {
    m += i ? ",\n    {\"name\": " : "\n    {\"name\": ";
    m += json_str(r.ranges[i].first);
    m += ", \"first\": " + std::to_string(first) + ", \"count\": " + std::to_string(r.ranges[i].second) + "}";
    first += r.ranges[i].second;
  }
  return m + "\n  ]";
}

// ------------------------------------------------------------------------------------------------ the scanner
// A flag of a tool. A flag with a value gets it as `v`; a switch gets null. apply returns the error message of a bad
// value, or an empty string.
struct Flag {
  const char* name;
  bool value;
  std::function<std::string(const char* v)> apply;
};
using Table = std::vector<Flag>;
// a word no table knows that looks like a flag: the reference's tool hook (cli.h parse_common); -1 unknown, 0 taken
// (i moved past its value), > 0 the failure's exit code with err set
using Fallback = std::function<int(int& i, const std::string& flag)>;

std::string parse_optics_spec(const std::string& s, std::vector<Optics>& out);

// the flags both tools take (reference cli.cpp's parse_common)
Table common_flags(CommonArgs& a, RunOptions& o) {
  const auto onoff = [](bool& to, const char* msg) {
    return [&to, msg](const char* v) -> std::string {
      const int k = vals::word(v, {"on", "off"});
      if (k < 0) return msg;
      to = k == 0;
      return {};
    };
  };
  const auto finite = [](double& to, const char* msg) {
    return [&to, msg](const char* v) -> std::string {
      double x;
      if (!vals::strict(v, x) || !std::isfinite(x)) return msg;
      to = x;
      return {};
    };
  };
  return {
      {"-v", false, [&o](const char*) { o.verbose = true; return std::string(); }},
      {"--context", true, [&a](const char* v) { a.context.push_back(v); return std::string(); }},
      {"--on", true,
       [&o](const char* v) -> std::string {
         const int k = vals::word(v, {"faces", "vertices"});
         if (k < 0) return "--on faces|vertices";
         o.onVertices = k == 1;
         return {};
       }},
      {"--sides", true,
       [&o](const char* v) -> std::string {
         const int k = vals::word(v, {"front", "back", "both"});
         if (k < 0) return "--sides front|back|both";
         o.sides = k == 0 ? Sides::Front : k == 1 ? Sides::Back : Sides::Both;
         return {};
       }},
      {"--cull-backfaces", false, [&o](const char*) { o.cullBackfaces = true; return std::string(); }},
      {"--no-self-occlude", false, [&o](const char*) { o.selfOcclude = false; return std::string(); }},
      {"--no-occlude", true,
       [&o](const char* v) -> std::string {
         for (const std::string& n : vals::pieces(v, ';')) {
           if (n.empty()) return std::string("--no-occlude ") + v + ": empty NAME";
           o.noOcclude.push_back(n);
         }
         return {};
       }},
      {"--north", true, finite(o.north, "--north DEG")},
      {"--offset", true, finite(o.offset, "--offset M")},
      {"--virtual-ground", true, onoff(o.virtualGround, "--virtual-ground on|off")},
      {"--optics", true, [&o](const char* v) { return parse_optics_spec(v, o.optics); }},
      {"--transmission", true, onoff(o.transmission, "--transmission off|on")},
      {"--out", true, [&a](const char* v) { a.out = v; return std::string(); }},
  };
}

// Scans argv against the tables: -h / --help stops the scan (a.help); --bench takes two optional whole numbers; a
// word that looks like a flag (a dash and no digit after it) must be in a table (or taken by the fallback); anything
// else is a sensor mesh path.
int scan(int argc, char** argv, const std::vector<const Table*>& tables, CommonArgs& a, std::string& err, const Fallback& fallback = {}) {
  for (int i = 1; i < argc; ++i) {
    const std::string_view arg = argv[i];
    if (arg == "-h" || arg == "--help") { a.help = true; return 0; }
    if (arg == "--bench") {
      a.bench = true;
      const auto digit = [&](int k) { return k < argc && std::isdigit(static_cast<unsigned char>(argv[k][0])); };
      if (digit(i + 1) && digit(i + 2)) {
        a.benchSens = size_t(std::atoll(argv[i + 1]));
        a.benchTris = size_t(std::atoll(argv[i + 2]));
        i += 2;
      }
      continue;
    }
    const Flag* f = nullptr;
    for (const Table* t : tables)
      for (const Flag& row : *t)
        if (arg == row.name) f = &row;
    if (!f) {
      if (arg.size() > 1 && arg[0] == '-' && !std::isdigit(static_cast<unsigned char>(arg[1]))) {
        const int rc = fallback ? fallback(i, std::string(arg)) : -1;
        if (rc < 0) { err = "unknown flag " + std::string(arg); return 1; }
        if (rc) return rc;
        continue;
      }
      a.sensors.emplace_back(arg);
      continue;
    }
    const char* v = nullptr;
    if (f->value) {
      if (i + 1 >= argc) { err = std::string("missing value for ") + f->name; return 1; }
      v = argv[++i];
    }
    if (std::string e = f->apply(v); !e.empty()) { err = std::move(e); return 1; }
  }
  return 0;
}

// SPEC[;SPEC...], SPEC = NAME[/SLOT]=RHO[,TAU[,glass|porous]]: the last '=' splits; /SLOT only as a trailing '/' and
// digits (a name may hold a '/'); porous unless the kind is given. Ranges are checked by optics_rows_ok.
std::string parse_optics_spec(const std::string& s, std::vector<Optics>& out) {
  for (const std::string& spec : vals::pieces(s, ';')) {
    const auto bad = [&](const std::string& what) { return "--optics " + spec + ": " + what; };
    const size_t eq = spec.rfind('=');
    if (eq == std::string::npos || eq == 0) return bad("NAME[/SLOT]=RHO[,TAU[,glass|porous]]");
    Optics o;
    o.mesh = spec.substr(0, eq);
    const size_t sl = o.mesh.rfind('/');
    const bool slotted = sl != std::string::npos && sl + 1 < o.mesh.size() && o.mesh.find_first_not_of("0123456789", sl + 1) == std::string::npos;
    if (slotted) {
      unsigned long n = 0;
      const auto r = std::from_chars(o.mesh.data() + sl + 1, o.mesh.data() + o.mesh.size(), n);
      if (r.ec != std::errc() || n > 65535) return bad("SLOT 0..65535");
      o.slot = int(n);
      o.mesh.erase(sl);
      if (o.mesh.empty() || o.mesh == "*") return bad("/SLOT needs a mesh NAME");
    }
    const std::vector<std::string> parts = vals::pieces(std::string_view(spec).substr(eq + 1), ',');
    if (parts.size() > 3) return bad("NAME[/SLOT]=RHO[,TAU[,glass|porous]]");
    double x[2] = {0, 0};
    for (size_t k = 0; k < parts.size() && k < 2; ++k)
      if (!vals::strict(parts[k], x[k])) return bad("bad number '" + parts[k] + "'");
    if (parts.size() == 3 && parts[2] != "glass" && parts[2] != "porous") return bad("kind glass or porous");
    o.rho = x[0];
    o.tau = x[1];
    o.glass = parts.size() == 3 && parts[2] == "glass";
    out.push_back(o);
  }
  return {};
}

int parse_common(int argc, char** argv, CommonArgs& a, RunOptions& o, std::string& err,
                 const std::function<int(int& i, const std::string& flag)>& tool) {
  const Table common = common_flags(a, o);
  return scan(argc, argv, {&common}, a, err, tool);
}

int parse(int argc, char** argv, Args& a, std::string& err) {
  Options& o = a.opt;
  const auto number = [](double& to) { return [&to](const char* v) { to = std::atof(v); return std::string(); }; };
  const auto path = [](std::string& to) { return [&to](const char* v) { to = v; return std::string(); }; };
  const auto whole = [](int& to, const char* flag) {
    return [&to, flag](const char* v) -> std::string {
      const char* e = v + std::strlen(v);
      const auto r = std::from_chars(v, e, to);
      if (r.ec != std::errc() || r.ptr != e) return std::string(flag) + " needs a whole number";
      return {};
    };
  };
  const Table tool = {
      {"--mode", true,
       [&o](const char* v) -> std::string {
         const int k = vals::word(v, {"physical", "ladybug"});
         if (k < 0) return "--mode physical|ladybug";
         o.mode = k == 0 ? Mode::Physical : Mode::Ladybug;
         return {};
       }},
      {"--quantity", true,
       [&o](const char* v) -> std::string {
         o.quantities.clear();
         for (const std::string& q : vals::pieces(v, ',')) {
           const auto* at = std::find_if(std::begin(kQuantityNames), std::end(kQuantityNames), [&](const char* n) { return q == n; });
           if (at == std::end(kQuantityNames)) return "unknown quantity";
           o.quantities.push_back(static_cast<Quantity>(at - std::begin(kQuantityNames)));
         }
         return {};
       }},
      {"--epw", true, path(a.epw)},
      {"--clear-sky", false, [&a](const char*) { a.clearSky = true; return std::string(); }},
      {"--turbidity", true, number(a.turbidity)},
      {"--altitude", true, number(a.altitude)},
      {"--lat", true, number(o.lat)},
      {"--lon", true, number(o.lon)},
      {"--tz", true, number(o.tz)},
      {"--time", true, path(a.timeStr)},
      {"--from", true, path(a.fromStr)},
      {"--to", true, path(a.toStr)},
      {"--step", true,
       [&o](const char* v) -> std::string {
         o.stepMin = std::atoi(v);
         return o.stepMin > 0 ? std::string() : "--step must be positive minutes";
       }},
      {"--hours", true,
       [&o](const char* v) -> std::string {
         int h0, m0, h1, m1;
         char extra;
         const bool ok = std::sscanf(v, "%d:%d-%d:%d%c", &h0, &m0, &h1, &m1, &extra) == 4;
         const int from = h0 * 60 + m0, to = h1 * 60 + m1;
         if (!ok || h0 < 0 || m0 < 0 || m0 > 59 || m1 < 0 || m1 > 59 || from >= to || to > 1440) return "--hours HH:MM-HH:MM with start < end <= 24:00";
         o.dayFrom = from;
         o.dayTo = to;
         return {};
       }},
      {"--sky-density", true, [&o](const char* v) { o.mf = std::atoi(v); return std::string(); }},
      {"--patch-samples", true, [&o](const char* v) { o.patchSamples = std::atoi(v); return std::string(); }},
      {"--ground", true, number(o.ground)},
      {"--fresnel", true,
       [&o](const char* v) -> std::string {
         const int k = vals::word(v, {"on", "off"});
         if (k < 0) return "--fresnel on|off";
         o.fresnel = k == 0;
         return {};
       }},
      {"--bounces", true, whole(o.bounces, "--bounces")},
      {"--bounce-samples", true, whole(o.bounceSamples, "--bounce-samples")},
      {"--timeseries", true, path(o.seriesPath)},
      {"--dump-bits", true, path(o.dumpBitsPath)},
      {"--dump-sky", true, path(o.dumpSky)},
      {"--dump-sun", true, path(o.dumpSun)},
  };
  const Table common = common_flags(a, o);
  if (const int rc = scan(argc, argv, {&common, &tool}, a, err); rc || a.help) return rc;
  return optics_ok(o, err) ? 0 : 1;
}

int prepare(Args& a, Weather& weather, bool& haveWeather, std::string& err) {
  Options& o = a.opt;
  const auto usage = [&](const char* msg) { err = msg; return 1; };
  if (a.clearSky && !a.epw.empty()) return usage("--clear-sky and --epw are exclusive");
  if (a.clearSky && (std::isnan(o.lat) || std::isnan(o.lon))) return usage("--clear-sky needs --lat and --lon");
  // without a weather file the clock defaults to the longitude's zone
  if (!std::isnan(o.lat) && !std::isnan(o.lon) && std::isnan(o.tz) && a.epw.empty() && !haveWeather) {
    o.tz = std::round(o.lon / 15);
    if (o.verbose) std::fprintf(stderr, "info: tz defaulted to %g\n", o.tz);
  }
  if (!a.epw.empty() || a.clearSky) {
    if (haveWeather) return usage(a.clearSky ? "--clear-sky cannot be combined with weather supplied by the caller"
                                             : "--epw cannot be combined with weather supplied by the caller");
    if (a.clearSky) {
      if (!clear_sky_weather(o.lat, o.lon, o.tz, a.turbidity, a.altitude, weather, err)) return 1;
    } else {
      if (!read_epw(a.epw, weather, err)) return 2;
      if (o.verbose && weather.missing) std::fprintf(stderr, "warn: %d EPW rows with missing radiation treated as 0\n", weather.missing);
    }
    haveWeather = true;
  }
  // the time window: one instant, or a range (with an optional window of hours every day), in the weather's year
  const bool leap = haveWeather && weather.leap;
  const bool hasTime = !a.timeStr.empty(), hasFrom = !a.fromStr.empty(), hasTo = !a.toStr.empty();
  if (hasTime && (hasFrom || hasTo)) return usage("--time cannot be combined with --from/--to");
  if (hasFrom != hasTo) return usage(hasFrom ? "--from needs --to" : "--to needs --from");
  if (hasTime && !parse_hoy(a.timeStr, leap, o.t0, err)) return 1;
  if (hasFrom && (!parse_hoy(a.fromStr, leap, o.t0, err) || !parse_hoy(a.toStr, leap, o.t1, err))) return 1;
  if ((o.dayFrom > 0 || o.dayTo < 1440) && !hasFrom) return usage("--hours needs --from/--to");
  return 0;
}

std::string manifest_json(const Result& r, const std::string& extraFields) {
  std::string m = "{\n  \"columns\": [";
  for (size_t c = 0; c < r.columns.size(); ++c) m += (c ? ", \"" : "\"") + r.columns[c] + "\"";
  char head[160];
  std::snprintf(head, sizeof head, "],\n  \"duration_h\": %.6f,\n  \"dirs\": {\"sky\": %zu, \"ground\": %zu, \"sun\": %zu},\n  \"meshes\": ",
                r.durationHours, r.nSky, r.nSky, r.nSun);
  m += head;
  m += manifest_meshes(r);
  if (!extraFields.empty()) m += ",\n  " + extraFields;
  return m + "\n}\n";
}

}  // namespace daylight::cli

namespace daylight::views::cli {

using daylight::cli::Flag;
using daylight::cli::Table;

const char* const kUsage =
    "usage: views <mesh>... [flags]               (MI355X build: HIP kernels on gfx950)\n"
    "  meshes: .obj (one o/g group per mesh) or .dmesh (raw arrays); world metres, Z up; positional = sensor meshes\n"
    "  --context FILE           occluder-only mesh (repeatable)\n"
    "  --target NAME=GROUP[;NAME=GROUP...]  rays ending on mesh NAME count for GROUP (at most 16 groups)\n"
    "  --fov H[,LO,HI]          degrees: H wide about the normal (up to 180) or 360 all around, elevation LO..HI\n"
    "                           (default 120,-70,50; H,0,0: horizontal rays only, the 2D isovist)\n"
    "  --cell DEG               degrees per ray, rows and columns (default 2)\n"
    "  --jitter on|off          each ray at a seeded random point of its cell (default on)\n"
    "  --distance STAT[,STAT...]  isovist|mean|min|group|none (default: isovist with --fov 360, else mean,min)\n"
    "  --far M                  what open counts as, and the cap of every distance (default 1000)\n"
    "  --virtual-ground on|off  a ground plane ends downward rays (default on)   --ground-z Z  its height\n"
    "  --reach D[,D...]|en17037 share of the field that sees at least D metres (at most 8)\n"
    "  --on faces|vertices  --sides front|back|both  --cull-backfaces  --north DEG  --offset M  --no-self-occlude\n"
    "  --no-occlude NAME[;NAME...]  --optics SPEC[;SPEC...]  --transmission off|on\n"
    "  --out FILE[.npy]         values + FILE.json manifest    --dump-rays F.npy  classes uint8 + F.npy.depth.npy\n"
    "  --bench [NSENS NTRIS]  -v          exit code: 1 usage, 2 runtime error\n";

int parse(int argc, char** argv, Args& a, std::string& err) {
  Options& o = a.opt;
  const auto one = [](double& to, const char* msg) {
    return [&to, msg](const char* v) -> std::string {
      std::vector<double> x;
      if (!daylight::cli::vals::list(v, x) || x.size() != 1) return msg;
      to = x[0];
      return {};
    };
  };
  const Table tool = {
      {"--target", true,
       [&o](const char* v) -> std::string {
         for (const std::string& spec : daylight::cli::vals::pieces(v, ';')) {
           if (spec.empty()) return std::string("--target ") + v + ": empty NAME=GROUP (a stray ;)";
           const size_t eq = spec.rfind('=');
           if (eq == std::string::npos || eq == 0 || eq + 1 == spec.size()) return "--target " + spec + ": NAME=GROUP";
           o.targets.push_back({spec.substr(0, eq), spec.substr(eq + 1)});
         }
         return {};
       }},
      {"--fov", true,
       [&o](const char* v) -> std::string {
         std::vector<double> x;
         if (!daylight::cli::vals::list(v, x) || (x.size() != 1 && x.size() != 3)) return "--fov H[,LO,HI] in degrees";
         o.fovH = x[0];
         if (x.size() == 3) { o.elevLo = x[1]; o.elevHi = x[2]; }
         return {};
       }},
      {"--cell", true, one(o.cell, "--cell DEG")},
      {"--jitter", true,
       [&o](const char* v) -> std::string {
         const int k = daylight::cli::vals::word(v, {"on", "off"});
         if (k < 0) return "--jitter on|off";
         o.jitter = k == 0;
         return {};
       }},
      {"--distance", true,
       [&o](const char* v) -> std::string {
         o.distance = 0;
         if (!std::strcmp(v, "none")) return {};
         for (const std::string& t : daylight::cli::vals::pieces(v, ',')) {
           const int k = daylight::cli::vals::word(t.c_str(), {"isovist", "mean", "min", "group"});
           if (k < 0) return "--distance isovist|mean|min|group[,...] or none";
           o.distance |= k == 0 ? unsigned(DistIsovist) : k == 1 ? unsigned(DistMean) : k == 2 ? unsigned(DistMin) : unsigned(DistGroup);
         }
         return {};
       }},
      {"--far", true, one(o.farDist, "--far M")},
      {"--ground-z", true, one(o.groundZ, "--ground-z Z")},
      {"--reach", true,
       [&o](const char* v) -> std::string {
         if (!std::strcmp(v, "en17037")) { o.reach = {6, 20, 50}; return {}; }
         return daylight::cli::vals::list(v, o.reach) ? std::string() : "--reach D[,D...] in metres, or en17037";
       }},
      {"--dump-rays", true, [&o](const char* v) { o.dumpPath = v; return std::string(); }},
  };
  const Table common = daylight::cli::common_flags(a, o);
  if (const int rc = daylight::cli::scan(argc, argv, {&common, &tool}, a, err); rc || a.help) return rc;
  Field fd;
  return field({}, o, fd, err) ? 0 : 1;
}

std::string manifest_json(const Result& r, const Options& opt, const Field& f) {
  const auto num = [](double v) {
    if (!std::isfinite(v)) return std::string("null");
    char s[32];
    std::snprintf(s, sizeof s, "%.10g", v);
    return std::string(s);
  };
  const auto flag = [](bool b) { return std::string(b ? "true" : "false"); };
  std::string m = "{\n  \"columns\": [";
  for (size_t c = 0; c < r.columns.size(); ++c) m += (c ? ", \"" : "\"") + r.columns[c] + "\"";
  m += "],\n  \"view\": {\"fov\": [" + num(opt.fovH) + ", " + num(opt.elevLo) + ", " + num(opt.elevHi) + "], \"cell\": " + num(opt.cell);
  m += ", \"grid\": [" + std::to_string(f.nEl) + ", " + std::to_string(f.nAz) + "], \"jitter\": " + flag(opt.jitter);
  m += ", \"north\": " + num(opt.north) + ", \"offset\": " + num(opt.offset) + ", \"far\": " + num(opt.farDist);
  m += ", \"near\": 0.1, \"virtual_ground\": " + flag(opt.virtualGround) + ", \"ground_z\": " + num(f.groundZ);
  if (opt.transmission) m += ", \"transmission\": true";
  m += ", \"reach\": [";
  for (size_t k = 0; k < opt.reach.size(); ++k) m += (k ? ", " : "") + num(opt.reach[k]);
  m += "], \"classes\": [\"sky\", \"ground\", \"obstruction\"";
  std::vector<std::string> groups;   // in first-appearance order
  for (const Target& t : opt.targets)
    if (std::find(groups.begin(), groups.end(), t.group) == groups.end()) {
      groups.push_back(t.group);
      m += ", \"" + t.group + "\"";
    }
  m += "],\n    \"image\": \"rows from HI down, columns clockwise from the field's left edge (360: from north)\"},\n  \"targets\": [";
  for (size_t k = 0; k < opt.targets.size(); ++k)
    m += std::string(k ? ", " : "") + "{\"mesh\": " + daylight::cli::json_str(opt.targets[k].mesh) + ", \"group\": \"" + opt.targets[k].group + "\"}";
  m += "],\n  \"meshes\": " + daylight::cli::manifest_meshes(r) + "\n}\n";
  return m;
}

}  // namespace daylight::views::cli
