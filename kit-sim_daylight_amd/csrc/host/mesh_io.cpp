// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The mesh files of the path and the mesh checks. Formats and messages are reference src/mesh.cpp's (read_obj,
// read_dmesh, daylight.h:129-133):
//   .obj    one mesh per o / g group (before the first: the file's stem), v x y z, f with 1-based or negative corner
//           indices (texture / normal references after '/' ignored), '#' comments; a mesh keeps the points its faces
//           use, in file order; groups without faces are dropped.
//   .dmesh  "DLM1", u32 flags (bit 0: a u16 slot per face follows), u64 points, faces, indices, then points (f64
//           xyz), counts (i32), indices (i32) and the slots.
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>

#include "core.h"

namespace daylight {
namespace {

bool slurp(const std::string& path, std::string& buf) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::rewind(f);
  buf.assign(n > 0 ? size_t(n) : 0, '\0');
  const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  return ok;
}

// the file name without its directories and its last extension
std::string file_stem(const std::string& path) {
  const size_t slash = path.find_last_of("/\\"), from = slash == std::string::npos ? 0 : slash + 1, dot = path.rfind('.');
  return path.substr(from, dot == std::string::npos || dot < from ? std::string::npos : dot - from);
}

// A reader of one .obj file's text: a cursor over one line at a time.
class ObjReader {
 public:
  ObjReader(const std::string& path, const std::string& text, std::vector<SimulationMesh>& out) : path_(path), out_(out), s_(text.data()), end_(text.data() + text.size()) {
    cur_.name = file_stem(path);
  }

  bool read(std::string& err) {
    for (line_ = 1; s_ < end_; ++line_) {
      const char* nl = static_cast<const char*>(std::memchr(s_, '\n', size_t(end_ - s_)));
      eol_ = nl ? nl : end_;
      blanks();
      const char tag = s_[0];
      const bool spaced = s_ + 1 < eol_ && (s_[1] == ' ' || s_[1] == '\t');
      if (tag == 'v' && spaced && !vertex()) return fail(err, "bad vertex");
      if (tag == 'f' && spaced) {
        if (const char* what = face()) return fail(err, what);
      }
      if ((tag == 'o' || tag == 'g') && spaced) group();
      s_ = eol_ + 1;
    }
    close_mesh();
    return true;
  }

 private:
  const std::string& path_;
  std::vector<SimulationMesh>& out_;
  const char* s_;
  const char* const end_;
  const char* eol_ = nullptr;
  size_t line_ = 0;
  std::vector<double> xyz_;          // every point of the file
  std::vector<int> local_;           // per file point: its index in the current mesh, -1 unused there
  std::vector<int> used_;            // the file points the current mesh uses, in first-use order
  SimulationMesh cur_;

  bool fail(std::string& err, const char* what) const {
    err = path_ + ":" + std::to_string(line_) + ": " + what;
    return false;
  }
  void blanks() { while (s_ < eol_ && (*s_ == ' ' || *s_ == '\t' || *s_ == '\r')) ++s_; }
  void word_end() { while (s_ < eol_ && *s_ && *s_ != ' ' && *s_ != '\t' && *s_ != '\r') ++s_; }

  bool vertex() {
    ++s_;
    for (int a = 0; a < 3; ++a) {
      blanks();
      if (s_ < eol_ && *s_ == '+') ++s_;
      double x;
      const auto r = std::from_chars(s_, eol_, x);
      if (r.ec != std::errc()) return false;
      s_ = r.ptr;
      xyz_.push_back(x);
    }
    local_.push_back(-1);
    return true;
  }

  const char* face() {
    ++s_;
    int corners = 0;
    for (;;) {
      blanks();
      if (s_ >= eol_ || *s_ == '#') break;
      if (*s_ == '+') ++s_;
      long k;
      const auto r = std::from_chars(s_, eol_, k);
      if (r.ec != std::errc()) return "bad face token";
      s_ = r.ptr;
      word_end();   // "/vt/vn"
      const long n = long(local_.size());
      k = k < 0 ? n + k : k - 1;
      if (k < 0 || k >= n) return "vertex index out of range";
      if (local_[size_t(k)] < 0) {
        local_[size_t(k)] = int(used_.size());
        used_.push_back(int(k));
      }
      cur_.indices.push_back(int(k));   // a file index until the mesh closes
      ++corners;
    }
    if (corners) cur_.counts.push_back(corners);
    return nullptr;
  }

  void group() {
    close_mesh();
    s_ += 2;
    blanks();
    const char* w = s_;
    word_end();
    cur_.name = s_ > w ? std::string(w, s_) : file_stem(path_);
  }

  // the current mesh with its points in file order and its indices renumbered to them; then a fresh one
  void close_mesh() {
    if (!cur_.counts.empty()) {
      std::sort(used_.begin(), used_.end());
      cur_.points.resize(3 * used_.size());
      for (size_t k = 0; k < used_.size(); ++k) {
        local_[size_t(used_[k])] = int(k);
        std::copy_n(xyz_.data() + 3 * size_t(used_[k]), 3, cur_.points.data() + 3 * k);
      }
      for (int& i : cur_.indices) i = local_[size_t(i)];
      out_.push_back(std::move(cur_));
    }
    for (const int p : used_) local_[size_t(p)] = -1;
    used_.clear();
    cur_ = SimulationMesh();
  }
};

struct DmeshHeader {
  char magic[4];
  uint32_t flags;
  uint64_t nPoints, nFaces, nIndices;
};
static_assert(sizeof(DmeshHeader) == 32, "the .dmesh header is 32 bytes");

bool decode_dmesh(const char* data, size_t size, const std::string& name, const std::string& label, std::vector<SimulationMesh>& out, std::string& err) {
  const auto bad = [&](const std::string& what) { err = label + ": " + what; return false; };
  DmeshHeader h{};
  if (size < sizeof h) return bad("truncated");
  std::memcpy(&h, data, sizeof h);
  if (std::memcmp(h.magic, "DLM1", 4)) return bad("not a .dmesh file (magic DLM1)");
  if (h.flags & ~1u) return bad("unsupported header flags " + std::to_string(h.flags) + " (only bit 0, the slot block, is defined)");
  // sizes checked one by one before any product can overflow
  if (h.nPoints > size / 24 || h.nFaces > size / 4 || h.nIndices > size / 4) return bad("truncated");
  const uint64_t body = 24 * h.nPoints + 4 * h.nFaces + 4 * h.nIndices;
  if (32 + body > size) return bad("truncated");
  const uint64_t nSlots = (h.flags & 1) ? h.nFaces : 0;
  if (32 + body + 2 * nSlots > size) return bad("truncated slot block");
  SimulationMesh m;
  m.name = name;
  const char* p = data + sizeof h;
  const auto take = [&p](auto& v, uint64_t n) {
    v.resize(size_t(n));
    if (n) std::memcpy(v.data(), p, sizeof(v[0]) * size_t(n));
    p += sizeof(v[0]) * size_t(n);
  };
  take(m.points, 3 * h.nPoints);
  take(m.counts, h.nFaces);
  take(m.indices, h.nIndices);
  take(m.slot, nSlots);
  if (!mesh_ok(m, label, err)) return false;
  if (h.nFaces) out.push_back(std::move(m));
  return true;
}

}  // namespace

bool mesh_ok(const SimulationMesh& m, const std::string& label, std::string& err) {
  const auto bad = [&](const char* what) { err = label + ": " + what; return false; };
  uint64_t corners = 0;
  for (const int c : m.counts) {
    if (c < 3) return bad("face with fewer than 3 vertices");
    corners += uint64_t(c);
  }
  if (corners != m.indices.size()) return bad("sum(counts) != n_indices");
  if (!m.slot.empty() && m.slot.size() != m.counts.size()) return bad("one optics slot per face needed");
  const uint64_t n = m.numPoints();
  if (std::any_of(m.indices.begin(), m.indices.end(), [n](int i) { return i < 0 || uint64_t(i) >= n; })) return bad("vertex index out of range");
  return true;
}

bool read_obj(const std::string& path, std::vector<SimulationMesh>& out, std::string& err) {
  std::string text;
  if (!slurp(path, text)) { err = "cannot open " + path; return false; }
  return ObjReader(path, text, out).read(err);
}

bool read_dmesh(const char* data, size_t size, const std::string& name, std::vector<SimulationMesh>& out, std::string& err) {
  return decode_dmesh(data, size, name, name, out, err);
}

bool read_dmesh(const std::string& path, std::vector<SimulationMesh>& out, std::string& err) {
  std::string bytes;
  if (!slurp(path, bytes)) { err = "cannot open " + path; return false; }
  return decode_dmesh(bytes.data(), bytes.size(), file_stem(path), path, out, err);
}

}  // namespace daylight
