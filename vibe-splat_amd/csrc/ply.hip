// 3DGS .ply files. The rules (header keywords, accepted scalar types, the
// property names, channel-major f_rest, the written layout) and the error
// messages are the reference's load_ply / save_ply
// (src/splat/gaussian_cloud.cpp:95-365); the code is organised its own way:
//
// * Reading maps the file and works on its bytes in place: the header is cut
//   into lines and words without copies, and interpreting it yields a schema
//   (format, vertex count, the vertex record's properties with their byte
//   offsets and scalar types). A plan then names, for every float of a
//   Gaussian in GaussianCloud's layout, the property it comes from, and the
//   records are decoded through that plan, split over host threads for large
//   binary files. Truncation is known before any record is decoded.
// * Writing describes the output as a table of columns (name, source array,
//   stride, component); the same table makes the header lines and fills the
//   records, a chunk of vertices at a time.
#include "ply.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <string_view>
#include <thread>
#include <unordered_map>

namespace vsplat {

namespace {

// A scalar type's names, size and conversion to float. The library is built
// for x86-64 only, whose byte order is the file's (little-endian).
using Converter = float (*)(const unsigned char *);

template <class T> float convert(const unsigned char *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return static_cast<float>(v);
}

struct ScalarType {
  std::string_view name, alias;
  uint32_t bytes;
  Converter decode;
};

constexpr ScalarType kScalarTypes[] = {
    {"char", "int8", 1, convert<int8_t>},       {"uchar", "uint8", 1, convert<uint8_t>},
    {"short", "int16", 2, convert<int16_t>},    {"ushort", "uint16", 2, convert<uint16_t>},
    {"int", "int32", 4, convert<int32_t>},      {"uint", "uint32", 4, convert<uint32_t>},
    {"float", "float32", 4, convert<float>},    {"double", "float64", 8, convert<double>},
};

const ScalarType *scalar_type(std::string_view name) {
  for (const ScalarType &t : kScalarTypes)
    if (name == t.name || name == t.alias) return &t;
  return nullptr;
}

[[noreturn]] void fail(const std::string &path, const std::string &what) { throw Error(VSPLAT_RUNTIME_ERROR, path + ": " + what); }

// A read-only mapping of a whole file (empty for an empty or irregular file).
class MappedFile {
public:
  explicit MappedFile(const std::string &path) {
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) throw Error(VSPLAT_RUNTIME_ERROR, "Cannot open '" + path + "'");
    struct stat st {};
    if (::fstat(fd_, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
      void *p = ::mmap(nullptr, static_cast<size_t>(st.st_size), PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) throw Error(VSPLAT_RUNTIME_ERROR, "Cannot open '" + path + "'");
      data_ = p;
      size_ = static_cast<size_t>(st.st_size);
      (void)::madvise(data_, size_, MADV_SEQUENTIAL);
    }
  }
  MappedFile(const MappedFile &) = delete;
  MappedFile &operator=(const MappedFile &) = delete;
  ~MappedFile() {
    if (data_ != nullptr) ::munmap(data_, size_);
    if (fd_ >= 0) ::close(fd_);
  }
  std::string_view text() const { return {static_cast<const char *>(data_), size_}; }
  const unsigned char *bytes() const { return static_cast<const unsigned char *>(data_); }

private:
  int fd_ = -1;
  void *data_ = nullptr;
  size_t size_ = 0;
};

constexpr std::string_view kBlank = " \t\r\v\f";

// Up to `most` whitespace-separated words of a line.
size_t split_words(std::string_view line, std::string_view *words, size_t most) {
  size_t n = 0, at = 0;
  while (n < most) {
    const size_t begin = line.find_first_not_of(kBlank, at);
    if (begin == std::string_view::npos) break;
    const size_t end = std::min(line.find_first_of(kBlank, begin), line.size());
    words[n++] = line.substr(begin, end - begin);
    at = end;
  }
  return n;
}

struct VertexProperty {
  std::string name;
  const ScalarType *type;
  size_t offset; // in the binary record
};

struct Schema {
  bool ascii = false;
  uint64_t vertices = 0;
  std::vector<VertexProperty> properties;
  size_t stride = 0; // bytes per binary record
  size_t body = 0;   // where the records start
};

// The header's lines, interpreted in order with the reference's rules: only
// the vertex element's properties are kept, an element with entries before it
// or a list property on it is refused, and the first offending line decides
// the message.
Schema read_schema(std::string_view file, const std::string &path) {
  size_t at = 0;
  const auto next_line = [&](std::string_view &line) {
    if (at >= file.size()) return false;
    const size_t end = std::min(file.find('\n', at), file.size());
    line = file.substr(at, end - at);
    at = end < file.size() ? end + 1 : end;
    return true;
  };
  std::string_view line;
  if (!next_line(line) || line.substr(0, 3) != "ply") fail(path, "not a PLY file");
  Schema s;
  enum class Section { None, Vertex, Other } section = Section::None;
  bool haveFormat = false, haveVertex = false;
  while (next_line(line)) {
    std::string_view w[3];
    const size_t words = split_words(line, w, 3);
    const std::string_view key = words > 0 ? w[0] : std::string_view{};
    if (key == "end_header") break;
    if (key.empty() || key == "comment" || key == "obj_info") continue;
    if (key == "format") {
      if (w[1] == "ascii") s.ascii = true;
      else if (w[1] == "binary_little_endian") s.ascii = false;
      else fail(path, "unsupported PLY format '" + std::string(w[1]) + "'");
      haveFormat = true;
    } else if (key == "element") {
      uint64_t entries = 0;
      if (words > 2) (void)std::from_chars(w[2].data(), w[2].data() + w[2].size(), entries);
      if (w[1] == "vertex") {
        if (haveVertex) fail(path, "more than one vertex element");
        haveVertex = true;
        section = Section::Vertex;
        s.vertices = entries;
      } else {
        if (!haveVertex && entries > 0) fail(path, "element '" + std::string(w[1]) + "' before the vertex element is not supported");
        section = Section::Other;
      }
    } else if (key == "property" && section == Section::Vertex) {
      if (w[1] == "list") fail(path, "list properties on vertices are not supported");
      const ScalarType *type = scalar_type(w[1]);
      if (type == nullptr) fail(path, "unknown property type '" + std::string(w[1]) + "'");
      s.properties.push_back({std::string(w[2]), type, s.stride});
      s.stride += type->bytes;
    }
  }
  if (!haveFormat || !haveVertex) fail(path, "PLY header has no format or vertex element");
  s.body = at;
  return s;
}

// Which property each float of a Gaussian comes from, in GaussianCloud's
// order: position (3), log-scale (3), rotation (4), logit opacity, then the SH
// coefficient-major with channels inner, from f_dc_* and the channel-major
// f_rest_*. The SH degree follows from how many f_rest_i there are.
struct Plan {
  int degree = 0;
  uint32_t coefficients = 1;
  std::vector<uint32_t> from;
  uint32_t floats() const { return static_cast<uint32_t>(from.size()); }
};

Plan plan_for(const Schema &s, const std::string &path) {
  std::unordered_map<std::string_view, uint32_t> where; // a repeated name: its last property
  for (uint32_t i = 0; i < s.properties.size(); ++i) where[s.properties[i].name] = i;
  const auto property = [&](const std::string &name) {
    const auto found = where.find(name);
    if (found == where.end()) fail(path, "missing property '" + name + "' (is this a Gaussian splat file?)");
    return found->second;
  };
  uint32_t rest = 0;
  while (where.count("f_rest_" + std::to_string(rest)) != 0) ++rest;
  Plan p;
  constexpr std::pair<uint32_t, int> kDegreeOfRest[] = {{0u, 0}, {9u, 1}, {24u, 2}, {45u, 3}};
  const auto degree = std::find_if(std::begin(kDegreeOfRest), std::end(kDegreeOfRest), [&](const auto &d) { return d.first == rest; });
  if (degree == std::end(kDegreeOfRest)) fail(path, "unexpected number of f_rest properties (" + std::to_string(rest) + ")");
  p.degree = degree->second;
  p.coefficients = static_cast<uint32_t>((p.degree + 1) * (p.degree + 1));
  for (const char *name : {"x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity"})
    p.from.push_back(property(name));
  std::vector<uint32_t> dc;
  for (int c = 0; c < 3; ++c) dc.push_back(property("f_dc_" + std::to_string(c)));
  const uint32_t perChannel = p.coefficients - 1;
  for (uint32_t k = 0; k < p.coefficients; ++k)
    for (uint32_t c = 0; c < 3; ++c) p.from.push_back(k == 0 ? dc[c] : where.at("f_rest_" + std::to_string(c * perChannel + k - 1)));
  return p;
}

// The destination of float j of Gaussian v: one of the cloud's five arrays.
struct Destination {
  float *base;
  uint32_t width; // floats per Gaussian in that array
  uint32_t index; // which of them
};

std::vector<Destination> destinations(const PlyCloud &out, const Plan &p) {
  std::vector<Destination> d;
  for (uint32_t a = 0; a < 3; ++a) d.push_back({out.positions, 3, a});
  for (uint32_t a = 0; a < 3; ++a) d.push_back({out.scales, 3, a});
  for (uint32_t a = 0; a < 4; ++a) d.push_back({out.rotations, 4, a});
  d.push_back({out.opacities, 1, 0});
  const uint32_t shFloats = 3 * p.coefficients;
  for (uint32_t j = 0; j < shFloats; ++j) d.push_back({out.sh, shFloats, j});
  return d;
}

// Binary records [first, last) through the plan.
void decode_binary(const unsigned char *records, const Schema &s, const Plan &p, const std::vector<Destination> &to, uint64_t first,
                   uint64_t last) {
  const uint32_t floats = p.floats();
  std::vector<Converter> decode(floats);
  std::vector<size_t> offset(floats);
  for (uint32_t j = 0; j < floats; ++j) {
    decode[j] = s.properties[p.from[j]].type->decode;
    offset[j] = s.properties[p.from[j]].offset;
  }
  for (uint64_t v = first; v < last; ++v) {
    const unsigned char *record = records + v * s.stride;
    for (uint32_t j = 0; j < floats; ++j) to[j].base[v * to[j].width + to[j].index] = decode[j](record + offset[j]);
  }
}

// The next number of an ASCII body, as iostream extraction of a double reads
// it (optional sign, digits, point, exponent), then narrowed to float.
bool next_number(const char *&at, const char *end, float &value) {
  while (at < end && std::isspace(static_cast<unsigned char>(*at))) ++at;
  if (at == end) return false;
  const char *start = at;
  if (*start == '+') ++start;
  if (start == end || !(std::isdigit(static_cast<unsigned char>(*start)) || *start == '.' || *start == '-')) return false;
  double d = 0.0;
  const auto parsed = std::from_chars(start, end, d);
  if (parsed.ec != std::errc()) return false;
  at = parsed.ptr;
  value = static_cast<float>(d);
  return true;
}

struct Opened {
  MappedFile file;
  Schema schema;
  Plan plan;
  Opened(const std::string &path) : file(path), schema(read_schema(file.text(), path)), plan(plan_for(schema, path)) {}
};

} // namespace

PlyInfo ply_info(const std::string &path) {
  const Opened f(path);
  return PlyInfo{static_cast<size_t>(f.schema.vertices), f.plan.degree};
}

void read_ply(const std::string &path, const PlyCloud &out) {
  const Opened f(path);
  const Schema &s = f.schema;
  const std::vector<Destination> to = destinations(out, f.plan);
  const std::string_view text = f.file.text();
  if (!s.ascii) {
    const uint64_t available = text.size() - s.body;
    if (s.stride != 0 && s.vertices > available / s.stride) fail(path, "file is truncated");
    const unsigned char *records = f.file.bytes() + s.body;
    // Large files are decoded by several host threads, each a contiguous range of records.
    const uint64_t perThread = uint64_t{1} << 16;
    const uint32_t threads = static_cast<uint32_t>(std::min<uint64_t>({8, std::max(1u, std::thread::hardware_concurrency()),
                                                                       (s.vertices + perThread - 1) / perThread}));
    if (threads <= 1) {
      decode_binary(records, s, f.plan, to, 0, s.vertices);
      return;
    }
    std::vector<std::thread> pool;
    const uint64_t share = (s.vertices + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
      const uint64_t first = std::min(s.vertices, t * share), last = std::min(s.vertices, first + share);
      pool.emplace_back(decode_binary, records, std::cref(s), std::cref(f.plan), std::cref(to), first, last);
    }
    for (std::thread &t : pool) t.join();
    return;
  }
  // ASCII: every property of a record is read (kept or not), in order.
  const char *at = text.data() + s.body, *end = text.data() + text.size();
  std::vector<float> record(s.properties.size());
  for (uint64_t v = 0; v < s.vertices; ++v) {
    for (float &value : record)
      if (!next_number(at, end, value)) fail(path, "file is truncated");
    for (uint32_t j = 0; j < f.plan.floats(); ++j) to[j].base[v * to[j].width + to[j].index] = record[f.plan.from[j]];
  }
}

void write_ply(const std::string &path, const ConstPlyCloud &c) {
  if (c.shDegree < 0 || c.shDegree > 3) throw Error(VSPLAT_RUNTIME_ERROR, "GaussianCloud: SH degree must be 0..3");
  const uint32_t coefficients = static_cast<uint32_t>((c.shDegree + 1) * (c.shDegree + 1)), shFloats = 3 * coefficients;
  // The standard 3DGS record: position, zero normals, f_dc, f_rest channel-major
  // (all of red's higher coefficients, then green's, then blue's), opacity,
  // log-scales, rotation; every property a float.
  struct Column {
    std::string name;
    const float *base; // value of vertex v: base[v * width + index]
    uint32_t width, index;
  };
  static const float kZero = 0.0f; // the normals: width 0 reads it for every vertex
  std::vector<Column> columns;
  for (uint32_t a = 0; a < 3; ++a) columns.push_back({std::string(1, "xyz"[a]), c.positions, 3, a});
  for (uint32_t a = 0; a < 3; ++a) columns.push_back({std::string("n") + "xyz"[a], &kZero, 0, 0});
  for (uint32_t ch = 0; ch < 3; ++ch) columns.push_back({"f_dc_" + std::to_string(ch), c.sh, shFloats, ch});
  for (uint32_t ch = 0; ch < 3; ++ch)
    for (uint32_t k = 1; k < coefficients; ++k)
      columns.push_back({"f_rest_" + std::to_string(ch * (coefficients - 1) + k - 1), c.sh, shFloats, k * 3 + ch});
  columns.push_back({"opacity", c.opacities, 1, 0});
  for (uint32_t a = 0; a < 3; ++a) columns.push_back({"scale_" + std::to_string(a), c.scales, 3, a});
  for (uint32_t a = 0; a < 4; ++a) columns.push_back({"rot_" + std::to_string(a), c.rotations, 4, a});

  std::string header = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(c.count) + "\n";
  for (const Column &col : columns) header += "property float " + col.name + "\n";
  header += "end_header\n";

  std::FILE *file = std::fopen(path.c_str(), "wb");
  if (file == nullptr) throw Error(VSPLAT_RUNTIME_ERROR, "Cannot write '" + path + "'");
  bool ok = std::fwrite(header.data(), 1, header.size(), file) == header.size();
  const size_t width = columns.size();
  // The columns' sources as (array, component), the arrays' rows found once per vertex.
  const float *arrays[] = {c.positions, &kZero, c.sh, c.opacities, c.scales, c.rotations};
  const uint32_t widths[] = {3, 0, shFloats, 1, 3, 4};
  std::vector<std::pair<uint32_t, uint32_t>> from;
  for (const Column &col : columns) {
    const uint32_t a = static_cast<uint32_t>(std::find(std::begin(arrays), std::end(arrays), col.base) - std::begin(arrays));
    from.push_back({a, col.index});
  }
  constexpr size_t kVerticesPerChunk = size_t{1} << 14;
  std::vector<float> chunk(width * std::min(c.count, kVerticesPerChunk));
  for (size_t first = 0; ok && first < c.count; first += kVerticesPerChunk) {
    const size_t n = std::min(kVerticesPerChunk, c.count - first);
    float *record = chunk.data();
    for (size_t v = first; v < first + n; ++v) {
      const float *row[6];
      for (int a = 0; a < 6; ++a) row[a] = arrays[a] + v * widths[a];
      for (const auto &f : from) *record++ = row[f.first][f.second];
    }
    ok = std::fwrite(chunk.data(), sizeof(float), n * width, file) == n * width;
  }
  ok = (std::fclose(file) == 0) && ok;
  if (!ok) throw Error(VSPLAT_RUNTIME_ERROR, "Failed writing '" + path + "'");
}

} // namespace vsplat
