// Oracle internals.  Test infrastructure only (see orc_common.hpp).
#pragma once
#include <cstdint>
#include <functional>
#include <tuple>
#include <unordered_map>
#include <vector>
#include "orc_api.h"
#include "orc_common.hpp"

namespace orc {

struct Lod {
  uint64_t size[3];    // LODInfo::m_iLODPixelSize
  uint64_t bricks[3];  // LODInfo::m_iLODBrickCount
  uint64_t offset;     // LODInfo::m_iLoDOffset
  double aspect[3];    // LODInfo::m_vAspect
};

struct TocEntry {
  uint64_t coords[4];
  uint64_t size[3];
  uint64_t offset, bytes;
  double minmax[2];
};

struct Octree {
  uint64_t dims[3] = {0, 0, 0};
  uint64_t brick[3] = {0, 0, 0};
  uint64_t overlap = 0;
  uint32_t bpv = 1;
  std::vector<Lod> lods;
  std::vector<TocEntry> toc;
  std::vector<uint8_t> data;

  void compute_metadata();
  void brick_size(const uint64_t c[4], uint64_t out[3]) const;
  uint64_t coords_to_index(const uint64_t c[4]) const;
  void index_to_coords(uint64_t index, uint64_t c[4]) const;
  void get_brick(const uint64_t c[4], uint8_t* out) const;
  void set_brick(const uint64_t c[4], const uint8_t* in);
  void convert(const uint8_t* in, bool clamp, bool median, bool lod0_only = false);
};

// Brick.h:57-74
typedef std::tuple<size_t, size_t, size_t> Key;
struct KeyHash {
  size_t operator()(const Key& bk) const {
    size_t ts = std::hash<size_t>()(std::get<0>(bk));
    size_t h_lod = std::hash<size_t>()(std::get<1>(bk));
    size_t brick = std::hash<size_t>()(std::get<2>(bk));
    size_t seed = h_lod;
    seed ^= brick + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    seed ^= ts + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    return seed;
  }
};
struct BrickMD {
  V3 center, extents;
  uint32_t n_voxels[3];
};
typedef std::unordered_map<Key, BrickMD, KeyHash> Table;

}  // namespace orc

struct orc_ds {
  orc::Octree tree;
  float scale[3] = {1, 1, 1};
  orc::Table table;
  void build_table();
  const orc::BrickMD& md_at(int lod, uint64_t lin) const;
  void texcoords(int lod, uint64_t lin, bool pow2, orc::V3& tmin, orc::V3& tmax) const;
};
