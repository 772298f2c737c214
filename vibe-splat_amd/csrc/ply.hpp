// 3DGS .ply reading and writing (src/splat/gaussian_cloud.hpp:60-67).
#pragma once

#include <cstddef>
#include <string>

#include "trainer.hpp"

namespace vsplat {

struct PlyInfo {
  size_t count;
  int shDegree;
};

// Caller arrays in GaussianCloud's layout, sized for PlyInfo's count and degree.
struct PlyCloud {
  float *positions, *scales, *rotations, *opacities, *sh;
};
struct ConstPlyCloud {
  const float *positions, *scales, *rotations, *opacities, *sh;
  size_t count;
  int shDegree;
};

// The vertex count and SH degree of a file (its header only).
PlyInfo ply_info(const std::string &path);
void read_ply(const std::string &path, const PlyCloud &out);
void write_ply(const std::string &path, const ConstPlyCloud &cloud);

} // namespace vsplat
