// 3DGS .ply files (src/splat/gaussian_cloud.cpp:95-365): the header parsed
// as the reference parses it, binary little-endian or ASCII vertex records
// read into a cloud's arrays, and a cloud written in the standard 3DGS binary
// layout. Host code: a file is read once, in chunks, and each record is
// scattered into the arrays through a table of (offset, type) per value,
// built once from the header, instead of a lookup per property per vertex.
#include "ply.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace vsplat {

namespace {

// PLY scalar types (gaussian_cloud.cpp:34-58).
enum class PlyType : uint8_t { I8, U8, I16, U16, I32, U32, F32, F64 };

bool parse_type(const std::string &t, PlyType &out) {
  if (t == "char" || t == "int8") out = PlyType::I8;
  else if (t == "uchar" || t == "uint8") out = PlyType::U8;
  else if (t == "short" || t == "int16") out = PlyType::I16;
  else if (t == "ushort" || t == "uint16") out = PlyType::U16;
  else if (t == "int" || t == "int32") out = PlyType::I32;
  else if (t == "uint" || t == "uint32") out = PlyType::U32;
  else if (t == "float" || t == "float32") out = PlyType::F32;
  else if (t == "double" || t == "float64") out = PlyType::F64;
  else return false;
  return true;
}

size_t type_size(PlyType t) {
  switch (t) {
  case PlyType::I8:
  case PlyType::U8: return 1;
  case PlyType::I16:
  case PlyType::U16: return 2;
  case PlyType::F64: return 8;
  default: return 4;
  }
}

// A little-endian value of the given type at p, as a float (gaussian_cloud.cpp:60-81).
float read_value(PlyType t, const unsigned char *p) {
  switch (t) {
  case PlyType::I8: return static_cast<float>(static_cast<int8_t>(p[0]));
  case PlyType::U8: return static_cast<float>(p[0]);
  case PlyType::I16: return static_cast<float>(static_cast<int16_t>(static_cast<uint16_t>(p[0] | (p[1] << 8))));
  case PlyType::U16: return static_cast<float>(static_cast<uint16_t>(p[0] | (p[1] << 8)));
  case PlyType::I32:
  case PlyType::U32:
  case PlyType::F32: {
    const uint32_t u = static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) | (static_cast<uint32_t>(p[2]) << 16) |
                       (static_cast<uint32_t>(p[3]) << 24);
    if (t == PlyType::I32) return static_cast<float>(static_cast<int32_t>(u));
    if (t == PlyType::U32) return static_cast<float>(u);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  }
  case PlyType::F64: {
    uint64_t u = 0;
    for (int k = 7; k >= 0; --k) u = (u << 8) | p[k];
    double d;
    std::memcpy(&d, &u, 8);
    return static_cast<float>(d);
  }
  }
  return 0.0f;
}

struct Property {
  std::string name;
  PlyType type;
  size_t offset;
};

struct Header {
  bool binary = true;
  size_t vertexCount = 0;
  std::vector<Property> props;
  size_t stride = 0;
};

[[noreturn]] void fail(const std::string &path, const std::string &what) { throw Error(VSPLAT_RUNTIME_ERROR, path + ": " + what); }

// gaussian_cloud.cpp:96-170.
Header parse_header(std::istream &in, const std::string &path) {
  std::string line;
  if (!std::getline(in, line) || line.substr(0, 3) != "ply") fail(path, "not a PLY file");
  Header h;
  bool inVertex = false, sawVertex = false, sawFormat = false;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string word;
    ls >> word;
    if (word == "end_header") break;
    if (word == "comment" || word == "obj_info" || word.empty()) continue;
    if (word == "format") {
      std::string format;
      ls >> format;
      if (format == "binary_little_endian") h.binary = true;
      else if (format == "ascii") h.binary = false;
      else fail(path, "unsupported PLY format '" + format + "'");
      sawFormat = true;
    } else if (word == "element") {
      std::string name;
      size_t count = 0;
      ls >> name >> count;
      if (name == "vertex") {
        if (sawVertex) fail(path, "more than one vertex element");
        inVertex = sawVertex = true;
        h.vertexCount = count;
      } else {
        if (!sawVertex && count > 0) fail(path, "element '" + name + "' before the vertex element is not supported");
        inVertex = false;
      }
    } else if (word == "property") {
      if (!inVertex) continue; // properties of elements after "vertex" are never read
      std::string typeName, name;
      ls >> typeName;
      if (typeName == "list") fail(path, "list properties on vertices are not supported");
      ls >> name;
      PlyType type;
      if (!parse_type(typeName, type)) fail(path, "unknown property type '" + typeName + "'");
      h.props.push_back({name, type, h.stride});
      h.stride += type_size(type);
    }
  }
  if (!sawFormat || !sawVertex) fail(path, "PLY header has no format or vertex element");
  return h;
}

// Where each value of a Gaussian comes from in a record: the cloud's
// positions (3), scales (3), rotations (4), opacity, then the SH in the
// cloud's layout (coefficient-major, channels inner), from f_dc_* and the
// channel-major f_rest_* (gaussian_cloud.cpp:236-275, 289-305).
struct Layout {
  int degree = 0;
  uint32_t coefficients = 1;
  std::vector<uint32_t> source; // property index per value
};

Layout layout_of(const Header &h, const std::string &path) {
  std::unordered_map<std::string, uint32_t> index;
  for (uint32_t i = 0; i < h.props.size(); ++i) index[h.props[i].name] = i;
  const auto require = [&](const std::string &name) {
    const auto it = index.find(name);
    if (it == index.end()) fail(path, "missing property '" + name + "' (is this a Gaussian splat file?)");
    return it->second;
  };
  size_t rest = 0;
  while (index.count("f_rest_" + std::to_string(rest)) != 0) ++rest;
  Layout l;
  switch (rest) {
  case 0: l.degree = 0; break;
  case 9: l.degree = 1; break;
  case 24: l.degree = 2; break;
  case 45: l.degree = 3; break;
  default: fail(path, "unexpected number of f_rest properties (" + std::to_string(rest) + ")");
  }
  l.coefficients = static_cast<uint32_t>((l.degree + 1) * (l.degree + 1));
  for (const char *name : {"x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity"}) {
    l.source.push_back(require(name));
  }
  for (uint32_t k = 0; k < l.coefficients; ++k) {
    for (uint32_t c = 0; c < 3; ++c) {
      l.source.push_back(k == 0 ? require("f_dc_" + std::to_string(c)) : index.at("f_rest_" + std::to_string(c * (l.coefficients - 1) + (k - 1))));
    }
  }
  return l;
}

// One record's values into the cloud's arrays at Gaussian v.
void store(const std::vector<float> &record, const Layout &l, size_t v, const PlyCloud &out) {
  const uint32_t *s = l.source.data();
  for (int a = 0; a < 3; ++a) out.positions[v * 3 + a] = record[s[a]];
  for (int a = 0; a < 3; ++a) out.scales[v * 3 + a] = record[s[3 + a]];
  for (int a = 0; a < 4; ++a) out.rotations[v * 4 + a] = record[s[6 + a]];
  out.opacities[v] = record[s[10]];
  const size_t shFloats = static_cast<size_t>(l.coefficients) * 3;
  for (size_t j = 0; j < shFloats; ++j) out.sh[v * shFloats + j] = record[s[11 + j]];
}

struct Opened {
  std::ifstream in;
  Header header;
  Layout layout;
};

void open_ply(const std::string &path, Opened &f) {
  f.in.open(path, std::ios::binary);
  if (!f.in) throw Error(VSPLAT_RUNTIME_ERROR, "Cannot open '" + path + "'");
  f.header = parse_header(f.in, path);
  f.layout = layout_of(f.header, path);
}

} // namespace

PlyInfo ply_info(const std::string &path) {
  Opened f;
  open_ply(path, f);
  return PlyInfo{f.header.vertexCount, f.layout.degree};
}

// load_ply (gaussian_cloud.cpp:236-327), into caller arrays sized from ply_info.
void read_ply(const std::string &path, const PlyCloud &out) {
  Opened f;
  open_ply(path, f);
  const Header &h = f.header;
  std::vector<float> record(h.props.size());
  if (h.binary) {
    constexpr size_t kChunk = 65536;
    std::vector<unsigned char> buffer(h.stride * kChunk);
    for (size_t first = 0; first < h.vertexCount; first += kChunk) {
      const size_t n = std::min(kChunk, h.vertexCount - first);
      f.in.read(reinterpret_cast<char *>(buffer.data()), static_cast<std::streamsize>(n * h.stride));
      if (static_cast<size_t>(f.in.gcount()) != n * h.stride) fail(path, "file is truncated");
      for (size_t i = 0; i < n; ++i) {
        const unsigned char *rec = buffer.data() + i * h.stride;
        for (size_t p = 0; p < h.props.size(); ++p) record[p] = read_value(h.props[p].type, rec + h.props[p].offset);
        store(record, f.layout, first + i, out);
      }
    }
  } else {
    for (size_t v = 0; v < h.vertexCount; ++v) {
      for (size_t p = 0; p < h.props.size(); ++p) {
        double d = 0.0;
        if (!(f.in >> d)) fail(path, "file is truncated");
        record[p] = static_cast<float>(d);
      }
      store(record, f.layout, v, out);
    }
  }
}

// save_ply (gaussian_cloud.cpp:330-365): binary little-endian, the standard
// 3DGS property order, zero normals, f_rest channel-major.
void write_ply(const std::string &path, const ConstPlyCloud &c) {
  if (c.shDegree < 0 || c.shDegree > 3) throw Error(VSPLAT_RUNTIME_ERROR, "GaussianCloud: SH degree must be 0..3");
  std::ofstream out(path, std::ios::binary);
  if (!out) throw Error(VSPLAT_RUNTIME_ERROR, "Cannot write '" + path + "'");
  const size_t n = c.count, coeffs = static_cast<size_t>((c.shDegree + 1) * (c.shDegree + 1)), rest = 3 * (coeffs - 1);
  std::string header = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
  for (const char *name : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"}) header += std::string("property float ") + name + "\n";
  for (size_t i = 0; i < rest; ++i) header += "property float f_rest_" + std::to_string(i) + "\n";
  for (const char *name : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
    header += std::string("property float ") + name + "\n";
  header += "end_header\n";
  out.write(header.data(), static_cast<std::streamsize>(header.size()));
  const size_t per = 9 + rest + 8;
  constexpr size_t kChunk = 65536;
  std::vector<float> records(per * std::min(n, kChunk));
  for (size_t first = 0; first < n; first += kChunk) {
    const size_t m = std::min(kChunk, n - first);
    for (size_t i = 0; i < m; ++i) {
      const size_t v = first + i;
      float *r = records.data() + i * per;
      size_t k = 0;
      for (int a = 0; a < 3; ++a) r[k++] = c.positions[v * 3 + a];
      for (int a = 0; a < 3; ++a) r[k++] = 0.0f; // normals
      for (int a = 0; a < 3; ++a) r[k++] = c.sh[v * coeffs * 3 + a];
      for (size_t ch = 0; ch < 3 && coeffs > 1; ++ch)
        for (size_t j = 1; j < coeffs; ++j) r[k++] = c.sh[(v * coeffs + j) * 3 + ch];
      r[k++] = c.opacities[v];
      for (int a = 0; a < 3; ++a) r[k++] = c.scales[v * 3 + a];
      for (int a = 0; a < 4; ++a) r[k++] = c.rotations[v * 4 + a];
    }
    // This library is built for little-endian hosts only (x86-64 with gfx950).
    out.write(reinterpret_cast<const char *>(records.data()), static_cast<std::streamsize>(m * per * sizeof(float)));
  }
  if (!out) throw Error(VSPLAT_RUNTIME_ERROR, "Failed writing '" + path + "'");
}

} // namespace vsplat
