// meshio.cpp -- triangle-mesh files for the resident-mesh cache:
// igl::read_triangle_mesh (read_triangle_mesh.cpp:117-190) for .off and .obj.
//
//   OFF  readOFF.cpp:22-170: header OFF / NOFF / COFF (prefix match), comment
//        and blank lines before the counts, vertex lines with >= 3 numbers
//        ('#' comments between them), faces "valence i0 i1 ..." with anything
//        after the last index on the line ignored.
//   OBJ  readOBJ.cpp:68-276: "v" lines with >= 3 coordinates (only x y z
//        kept, read_triangle_mesh.cpp:146-151), "f" lines of words i, i/t,
//        i/t/n or i//n, 1-based, negative indices counted back from the
//        vertices read so far; vt / vn / usemtl / mtllib / g / s / comments
//        and unknown lines are skipped.
//   Polygons become fans (p0, pc, pc+1), c = 1 .. n-2, in file order
//   (polygons_to_triangles.cpp:16-31).
// Host code only; the arrays are handed to gwn_soup_build / the server.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "gwn_internal.h"

namespace gwn {
namespace {

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

void read_off(FILE* fp, std::vector<double>& V, std::vector<std::vector<int64_t>>& polys) {
  char header[1000] = {0};
  if (std::fscanf(fp, "%999s\n", header) != 1)
    throw Error{GWN_ERR_INVALID, "OFF: missing header"};
  const std::string h(header);
  GWN_CHECK(starts_with(h, "OFF") || starts_with(h, "NOFF") || starts_with(h, "COFF"), GWN_ERR_INVALID,
            "OFF: first line should be OFF, NOFF or COFF, not " + h);
  char line[1000];
  do {
    if (!std::fgets(line, sizeof line, fp)) throw Error{GWN_ERR_INVALID, "OFF: missing counts"};
  } while (line[0] == '#' || line[0] == '\n');
  long nv = 0, nf = 0, ne = 0;
  GWN_CHECK(std::sscanf(line, "%ld %ld %ld", &nv, &nf, &ne) >= 2 && nv >= 0 && nf >= 0, GWN_ERR_INVALID,
            "OFF: bad counts line");
  V.assign(3 * (size_t)nv, 0.0);
  polys.assign((size_t)nf, {});
  for (long i = 0; i < nv;) {
    if (!std::fgets(line, sizeof line, fp)) throw Error{GWN_ERR_INVALID, "OFF: fewer vertices than announced"};
    double x[6];
    if (std::sscanf(line, "%lg %lg %lg %lg %lg %lg", &x[0], &x[1], &x[2], &x[3], &x[4], &x[5]) >= 3) {
      V[3 * i] = x[0];
      V[3 * i + 1] = x[1];
      V[3 * i + 2] = x[2];
      ++i;
    }
    // anything else (a comment, a blank line) is skipped
  }
  for (long i = 0; i < nf;) {
    long valence = 0;
    if (std::fscanf(fp, "%ld", &valence) == 1) {
      GWN_CHECK(valence >= 0 && valence < (1l << 20), GWN_ERR_INVALID, "OFF: bad face valence");
      std::vector<int64_t>& face = polys[(size_t)i];
      face.resize((size_t)valence);
      for (long j = 0; j < valence; ++j) {
        long idx = 0;
        GWN_CHECK(std::fscanf(fp, "%ld", &idx) == 1, GWN_ERR_INVALID, "OFF: truncated face");
        face[(size_t)j] = idx;
      }
      // the rest of the line (e.g. a face colour) is ignored
      int c;
      while ((c = std::fgetc(fp)) != EOF && c != '\n') {
      }
      ++i;
    } else {
      char hash;
      if (std::fscanf(fp, " %c", &hash) == 1 && hash == '#') {
        int c;
        while ((c = std::fgetc(fp)) != EOF && c != '\n') {
        }
      } else {
        throw Error{GWN_ERR_INVALID, "OFF: bad face line"};
      }
    }
  }
}

void read_obj(FILE* fp, std::vector<double>& V, std::vector<std::vector<int64_t>>& polys) {
  std::vector<char> buf(1 << 16);
  int64_t ntc = 0, nn = 0, line_no = 0;
  std::string line;
  auto next_line = [&](std::string& out) -> bool {
    out.clear();
    while (std::fgets(buf.data(), (int)buf.size(), fp)) {
      out += buf.data();
      if (!out.empty() && out.back() == '\n') return true;
    }
    return !out.empty();
  };
  while (next_line(line)) {
    ++line_no;
    std::istringstream ls(line);
    std::string type;
    if (!(ls >> type)) continue;  // empty line
    if (type == "v") {
      std::vector<double> c;
      double x;
      while (ls >> x) c.push_back(x);
      GWN_CHECK(c.size() >= 3, GWN_ERR_INVALID,
                "OBJ: vertex on line " + std::to_string(line_no) + " should have at least 3 coordinates");
      V.push_back(c[0]);
      V.push_back(c[1]);
      V.push_back(c[2]);
    } else if (type == "vt") {
      ++ntc;
    } else if (type == "vn") {
      ++nn;
    } else if (type == "f") {
      const int64_t nvtx = (int64_t)(V.size() / 3);
      auto shift = [](long i, int64_t n) -> int64_t { return i < 0 ? i + n : i - 1; };
      std::vector<int64_t> face;
      size_t with_t = 0, with_n = 0;
      std::string word;
      while (ls >> word) {
        long i = 0, it = 0, in = 0;
        int used = 0;
        if (std::sscanf(word.c_str(), "%ld/%ld/%ld%n", &i, &it, &in, &used) == 3) {
          ++with_t;
          ++with_n;
        } else if (std::sscanf(word.c_str(), "%ld//%ld%n", &i, &in, &used) == 2) {
          ++with_n;
        } else if (std::sscanf(word.c_str(), "%ld/%ld%n", &i, &it, &used) == 2) {
          ++with_t;
        } else if (std::sscanf(word.c_str(), "%ld%n", &i, &used) == 1) {
        } else {
          throw Error{GWN_ERR_INVALID, "OBJ: face on line " + std::to_string(line_no) + " has an invalid element"};
        }
        face.push_back(shift(i, nvtx));
      }
      // readOBJ.cpp:217-222: texture / normal indices on all corners or none
      GWN_CHECK(!face.empty() && (with_t == 0 || with_t == face.size()) && (with_n == 0 || with_n == face.size()),
                GWN_ERR_INVALID, "OBJ: face on line " + std::to_string(line_no) + " has an invalid format");
      polys.push_back(std::move(face));
    }
    // usemtl, mtllib, g, s, o, comments and unknown lines: skipped
  }
  (void)ntc;
  (void)nn;
}

}  // namespace

void read_triangle_mesh(const char* path, std::vector<double>& V, std::vector<int32_t>& F) {
  GWN_CHECK(path != nullptr, GWN_ERR_INVALID, "null path");
  std::string p(path), ext;
  const size_t dot = p.find_last_of('.'), slash = p.find_last_of("/\\");
  if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) ext = p.substr(dot + 1);
  std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  GWN_CHECK(ext == "off" || ext == "obj", GWN_ERR_INVALID, p + " is not a recognized mesh file format (.off, .obj)");
  FILE* fp = std::fopen(path, "rb");
  GWN_CHECK(fp != nullptr, GWN_ERR_INVALID, p + " could not be opened");
  FileCloser closer{fp};
  std::vector<std::vector<int64_t>> polys;
  V.clear();
  if (ext == "off")
    read_off(fp, V, polys);
  else
    read_obj(fp, V, polys);
  // polygons_to_triangles.cpp:16-31: fans (p0, pc, pc+1)
  F.clear();
  for (const auto& poly : polys)
    for (size_t c = 1; c + 1 < poly.size(); ++c) {
      for (int64_t idx : {poly[0], poly[c], poly[c + 1]}) {
        GWN_CHECK(idx >= INT32_MIN && idx <= INT32_MAX, GWN_ERR_INVALID, "face index out of int32 range");
        F.push_back((int32_t)idx);
      }
    }
}

}  // namespace gwn
