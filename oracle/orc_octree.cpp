// Oracle: ExtendedOctree bricking, LOD pyramid, per-brick stats, brick table.
// Test infrastructure only (see orc_common.hpp).
//
// This is a literal, procedural restatement of the reference conversion so
// that the product's closed-form device bricking can be checked against it:
//   ComputeMetadata      ExtendedOctree.cpp:188-243
//   ComputeBrickSize     ExtendedOctree.cpp:276-285
//   GetInputBrick        ExtendedOctreeConverter.cpp:288-373
//   ClampToEdge          ExtendedOctreeConverter.cpp:376-462
//   PermuteInputData     ExtendedOctreeConverter.cpp:1381-1414
//   DownsampleBricktoBrick / DownsampleBrick / ComputeHierarchy
//                        ExtendedOctreeConverter.inc:1-398
//   VolumeTools::Filter  VolumeTools.h:168-262
//   FillOverlap          ExtendedOctreeConverter.cpp:1203-1372
//   CopyBrickToBrick     ExtendedOctreeConverter.cpp:1153-1181
//   ComputeBrickStats    ExtendedOctreeConverter.inc:408-444
//   ComputeMetadataTOC   uvfDataset.cpp:242-288
//   GetTextCoords        uvfDataset.cpp:1747-1774
#include "orc_internal.hpp"
#include <cassert>
#include <limits>
#include <stdexcept>
#include <vector>

namespace orc {

void Octree::compute_metadata() {
  // ExtendedOctree.cpp:188-243
  uint64_t vs[3] = {dims[0], dims[1], dims[2]};
  double asp[3] = {1.0, 1.0, 1.0};
  uint64_t usable[3];
  for (int i = 0; i < 3; ++i) {
    if (brick[i] <= 2 * overlap) throw std::runtime_error("brick size must exceed 2*overlap");
    usable[i] = brick[i] - 2 * overlap;
  }
  lods.clear();
  do {
    Lod l;
    for (int i = 0; i < 3; ++i) l.size[i] = vs[i];
    if (!lods.empty()) {
      for (int i = 0; i < 3; ++i) {
        if (vs[i] > 1) {
          l.size[i] = (uint64_t)std::ceil(vs[i] / 2.0);
          asp[i] *= (vs[i] % 2) ? (float(vs[i]) / float(l.size[i])) : 2;
        }
      }
      double mx = std::max(asp[0], std::max(asp[1], asp[2]));
      for (int i = 0; i < 3; ++i) asp[i] /= mx;
      for (int i = 0; i < 3; ++i) vs[i] = l.size[i];
    }
    for (int i = 0; i < 3; ++i) {
      l.aspect[i] = asp[i];
      l.bricks[i] = (uint64_t)std::ceil(vs[i] / double(usable[i]));
    }
    l.offset = 0;
    lods.push_back(l);
  } while (vs[0] > 1 || vs[1] > 1 || vs[2] > 1);
  lods[0].offset = 0;
  for (size_t i = 1; i < lods.size(); ++i)
    lods[i].offset = lods[i - 1].offset + lods[i - 1].bricks[0] * lods[i - 1].bricks[1] * lods[i - 1].bricks[2];
}

void Octree::brick_size(const uint64_t c[4], uint64_t out[3]) const {
  // ExtendedOctree.cpp:276-285 + IsLastBrick :380-392
  const Lod& l = lods[c[3]];
  for (int i = 0; i < 3; ++i) {
    bool last = c[i] >= l.bricks[i] - 1;
    uint64_t core = brick[i] - 2 * overlap;
    out[i] = (last && (l.size[i] % core)) ? (2 * overlap + (l.size[i] % core)) : brick[i];
  }
}

uint64_t Octree::coords_to_index(const uint64_t c[4]) const {
  // ExtendedOctree.cpp:400-407
  const Lod& l = lods[c[3]];
  return l.offset + c[0] + c[1] * l.bricks[0] + c[2] * l.bricks[0] * l.bricks[1];
}

void Octree::index_to_coords(uint64_t index, uint64_t c[4]) const {
  // ExtendedOctree.cpp:414-428
  uint64_t w = 0;
  for (size_t i = 0; i < lods.size(); ++i) {
    if (lods[i].offset <= index) w = i; else break;
  }
  index -= lods[w].offset;
  const uint64_t* bc = lods[w].bricks;
  c[0] = index % bc[0];
  c[1] = (index / bc[0]) % bc[1];
  c[2] = index / (bc[0] * bc[1]);
  c[3] = w;
}

namespace {

struct Conv {
  Octree& t;
  const uint8_t* in;
  bool clamp;
  size_t vox;
  explicit Conv(Octree& tree, const uint8_t* input, bool c)
      : t(tree), in(input), clamp(c), vox(tree.bpv) {}

  uint64_t ov() const { return t.overlap; }

  // ExtendedOctreeConverter.cpp:376-462
  void clamp_to_edge(std::vector<uint8_t>& d, bool xs, bool ys, bool zs, bool xe, bool ye,
                     bool ze, const uint64_t bs[3]) {
    const uint64_t o = ov();
    if (xs)
      for (uint64_t z = 0; z < bs[2]; z++)
        for (uint64_t y = 0; y < bs[1]; y++) {
          size_t src = vox * (o + y * bs[0] + z * bs[0] * bs[1]);
          size_t tgt = src;
          for (uint64_t k = 0; k < o; k++) { tgt -= vox; std::memcpy(&d[tgt], &d[src], vox); }
        }
    if (xe)
      for (uint64_t z = 0; z < bs[2]; z++)
        for (uint64_t y = 0; y < bs[1]; y++) {
          size_t src = vox * (((bs[0] - 1) - o) + y * bs[0] + z * bs[0] * bs[1]);
          size_t tgt = src;
          for (uint64_t k = 0; k < o; k++) { tgt += vox; std::memcpy(&d[tgt], &d[src], vox); }
        }
    if (ys)
      for (uint64_t z = 0; z < bs[2]; z++) {
        size_t src = vox * (o * bs[0] + z * bs[0] * bs[1]);
        size_t tgt = src;
        for (uint64_t k = 0; k < o; k++) { tgt -= vox * bs[0]; std::memcpy(&d[tgt], &d[src], vox * bs[0]); }
      }
    if (ye)
      for (uint64_t z = 0; z < bs[2]; z++) {
        size_t src = vox * (((bs[1] - 1) - o) * bs[0] + z * bs[0] * bs[1]);
        size_t tgt = src;
        for (uint64_t k = 0; k < o; k++) { tgt += vox * bs[0]; std::memcpy(&d[tgt], &d[src], vox * bs[0]); }
      }
    if (zs)
      for (uint64_t y = 0; y < bs[1]; y++) {
        size_t src = vox * (o * (bs[0] * bs[1]) + y * bs[0]);
        size_t tgt = src;
        for (uint64_t k = 0; k < o; k++) { tgt -= vox * (bs[0] * bs[1]); std::memcpy(&d[tgt], &d[src], vox * bs[0]); }
      }
    if (ze)
      for (uint64_t y = 0; y < bs[1]; y++) {
        size_t src = vox * (((bs[2] - 1) - o) * (bs[0] * bs[1]) + y * bs[0]);
        size_t tgt = src;
        for (uint64_t k = 0; k < o; k++) { tgt += vox * (bs[0] * bs[1]); std::memcpy(&d[tgt], &d[src], vox * bs[0]); }
      }
  }

  // ExtendedOctreeConverter.cpp:288-373
  void get_input_brick(std::vector<uint8_t>& d, const uint64_t c[4]) {
    uint64_t bs[3];
    t.brick_size(c, bs);
    size_t bytes = (size_t)(bs[0] * bs[1] * bs[2] * vox);
    if (d.size() != bytes) d.resize(bytes);
    if (!clamp) std::memset(d.data(), 0, bytes);
    const uint64_t o = ov();
    const uint64_t* b0 = t.lods[0].bricks;
    const uint64_t xS = (c[0] == 0) ? o : 0, yS = (c[1] == 0) ? o : 0, zS = (c[2] == 0) ? o : 0;
    const uint64_t yE = bs[1] - ((c[1] == b0[1] - 1) ? o : 0);
    const uint64_t zE = bs[2] - ((c[2] == b0[2] - 1) ? o : 0);
    const int64_t X = (int64_t)t.dims[0], Y = (int64_t)t.dims[1];
    for (uint64_t z = 0; z < zE - zS; z++) {
      for (uint64_t y = 0; y < yE - yS; y++) {
        int64_t vx = -(int64_t)(o - xS) + (int64_t)(c[0] * (t.brick[0] - o * 2));
        int64_t vy = (int64_t)y - (int64_t)(o - yS) + (int64_t)(c[1] * (t.brick[1] - o * 2));
        int64_t vz = (int64_t)z - (int64_t)(o - zS) + (int64_t)(c[2] * (t.brick[2] - o * 2));
        int64_t inoff = (int64_t)vox * (vx + vy * X + vz * X * Y);
        size_t outoff = vox * (size_t)(xS + (y + yS) * bs[0] + (z + zS) * bs[0] * bs[1]);
        uint64_t line = bs[0] * vox;
        if (c[0] == 0) line -= o * vox;
        if (c[0] == b0[0] - 1) line -= o * vox;
        // The reference reads the line from the raw file.  When the last brick
        // along an axis has a core thinner than the overlap, the overlap of the
        // brick before it reaches past the end of the volume and that read
        // comes back short; model it as a short read (nothing copied) instead
        // of touching memory outside the input.
        if (vy >= Y || vz >= (int64_t)t.dims[2]) continue;
        if (vx + (int64_t)(line / vox) > X) line = (uint64_t)(X - vx) * vox;
        std::memcpy(&d[outoff], in + inoff, (size_t)line);
      }
    }
    if (clamp)
      clamp_to_edge(d, c[0] == 0, c[1] == 0, c[2] == 0, c[0] == b0[0] - 1, c[1] == b0[1] - 1,
                    c[2] == b0[2] - 1, bs);
  }

  // ExtendedOctreeConverter.cpp:1381-1414
  void permute_input() {
    std::vector<uint8_t> d;
    const uint64_t* bb = t.lods[0].bricks;
    for (uint64_t z = 0; z < bb[2]; z++)
      for (uint64_t y = 0; y < bb[1]; y++)
        for (uint64_t x = 0; x < bb[0]; x++) {
          uint64_t c[4] = {x, y, z, 0};
          get_input_brick(d, c);
          t.set_brick(c, d.data());
        }
  }

  template <class T, bool Median>
  static T filter8(T a, T b, T c, T d, T e, T f, T g, T h) {
    // VolumeTools.h:230-256
    if (Median) {
      auto order = [](T& p, T& q) { if (p > q) std::swap(p, q); };
      auto ins = [&](T& A, T& B, T& C, T& D, T& P) {
        if (P > C) { order(D, P); }
        else {
          if (P < B) { D = C; C = B; B = P; order(A, B); }
          else { D = C; C = P; }
        }
      };
      order(a, b); order(c, d); order(a, c); order(b, d); order(b, c);
      ins(a, b, c, d, e);
      ins(a, b, c, d, f);
      return std::max(std::min(d, g), c);
    }
    return T((double(a) + double(b) + double(c) + double(d) + double(e) + double(f) +
              double(g) + double(h)) / double(8));
  }
  template <class T, bool Median>
  static T filter4(T a, T b, T c, T d) {
    // VolumeTools.h:190-202
    if (Median) {
      auto order = [](T& p, T& q) { if (p > q) std::swap(p, q); };
      order(a, b); order(b, c);
      return std::max(a, b);
    }
    return T((double(a) + double(b) + double(c) + double(d)) / double(4));
  }
  template <class T, bool Median>
  static T filter2(T a, T b) {
    // VolumeTools.h:170-175
    if (Median) return a;
    return T((double(a) + double(b)) / double(2));
  }

  // ExtendedOctreeConverter.inc:1-248 (single component)
  template <class T, bool Median>
  void downsample_brick_to_brick(T* pData, const uint64_t ts[3], T* src, const uint64_t sc[4],
                                 const uint64_t toff[3]) {
    uint64_t ss[3];
    t.brick_size(sc, ss);
    t.get_brick(sc, (uint8_t*)src);
    const uint64_t o = ov();
    const uint64_t ex = (ss[0] - 2 * o) / 2, ey = (ss[1] - 2 * o) / 2, ez = (ss[2] - 2 * o) / 2;
    auto S = [&](uint64_t x, uint64_t y, uint64_t z) -> T& {
      return src[x + y * ss[0] + z * ss[0] * ss[1]];
    };
    auto D = [&](uint64_t x, uint64_t y, uint64_t z) -> T& {
      return pData[(x + o + toff[0]) + (y + o + toff[1]) * ts[0] + (z + o + toff[2]) * ts[0] * ts[1]];
    };
    for (uint64_t z = 0; z < ez; z++)
      for (uint64_t y = 0; y < ey; y++)
        for (uint64_t x = 0; x < ex; x++) {
          uint64_t X = 2 * x + o, Y = 2 * y + o, Z = 2 * z + o;
          // p0..p7 ordering of the reference: p0=(x,y,z) p1=(x,y,z+1) p2=(x,y+1,z)
          // p3=(x,y+1,z+1) p4..p7 = same with x+1
          D(x, y, z) = filter8<T, Median>(S(X, Y, Z), S(X, Y, Z + 1), S(X, Y + 1, Z), S(X, Y + 1, Z + 1),
                                         S(X + 1, Y, Z), S(X + 1, Y, Z + 1), S(X + 1, Y + 1, Z),
                                         S(X + 1, Y + 1, Z + 1));
        }
    if (ss[0] % 2)
      for (uint64_t z = 0; z < ez; z++)
        for (uint64_t y = 0; y < ey; y++) {
          uint64_t X = 2 * ex + o, Y = 2 * y + o, Z = 2 * z + o;
          D(ex, y, z) = filter4<T, Median>(S(X, Y, Z), S(X, Y, Z + 1), S(X, Y + 1, Z), S(X, Y + 1, Z + 1));
        }
    if (ss[1] % 2)
      for (uint64_t z = 0; z < ez; z++)
        for (uint64_t x = 0; x < ex; x++) {
          uint64_t X = 2 * x + o, Y = 2 * ey + o, Z = 2 * z + o;
          D(x, ey, z) = filter4<T, Median>(S(X, Y, Z), S(X, Y, Z + 1), S(X + 1, Y, Z), S(X + 1, Y, Z + 1));
        }
    if (ss[2] % 2)
      for (uint64_t y = 0; y < ey; y++)
        for (uint64_t x = 0; x < ex; x++) {
          uint64_t X = 2 * x + o, Y = 2 * y + o, Z = 2 * ez + o;
          D(x, y, ez) = filter4<T, Median>(S(X, Y, Z), S(X, Y + 1, Z), S(X + 1, Y, Z), S(X + 1, Y + 1, Z));
        }
    if ((ss[0] % 2) && (ss[1] % 2))
      for (uint64_t z = 0; z < ez; z++) {
        uint64_t X = 2 * ex + o, Y = 2 * ey + o, Z = 2 * z + o;
        D(ex, ey, z) = filter2<T, Median>(S(X, Y, Z), S(X, Y, Z + 1));
      }
    if ((ss[1] % 2) && (ss[2] % 2))
      for (uint64_t x = 0; x < ex; x++) {
        uint64_t X = 2 * x + o, Y = 2 * ey + o, Z = 2 * ez + o;
        D(x, ey, ez) = filter2<T, Median>(S(X, Y, Z), S(X + 1, Y, Z));
      }
    if ((ss[0] % 2) && (ss[2] % 2))
      for (uint64_t y = 0; y < ey; y++) {
        uint64_t X = 2 * ex + o, Y = 2 * y + o, Z = 2 * ez + o;
        D(ex, y, ez) = filter2<T, Median>(S(X, Y, Z), S(X, Y + 1, Z));
      }
    if ((ss[0] % 2) && (ss[1] % 2) && (ss[2] % 2)) {
      uint64_t X = 2 * ex + o, Y = 2 * ey + o, Z = 2 * ez + o;
      D(ex, ey, ez) = S(X, Y, Z);
    }
  }

  // ExtendedOctreeConverter.inc:250-356
  template <class T, bool Median>
  void downsample_brick(const uint64_t c[4], T* pData, T* src) {
    uint64_t ts[3];
    t.brick_size(c, ts);
    if (!clamp) std::memset(pData, 0, (size_t)(ts[0] * ts[1] * ts[2] * vox));
    const uint64_t* lower = t.lods[c[3] - 1].bricks;
    const bool R = c[0] * 2 + 1 < lower[0];
    const bool B = c[1] * 2 + 1 < lower[1];
    const bool K = c[2] * 2 + 1 < lower[2];
    const uint64_t o = ov();
    const uint64_t sp[3] = {(uint64_t)std::ceil((t.brick[0] - 2 * o) / 2.0),
                            (uint64_t)std::ceil((t.brick[1] - 2 * o) / 2.0),
                            (uint64_t)std::ceil((t.brick[2] - 2 * o) / 2.0)};
    const uint64_t L = c[3] - 1;
    auto run = [&](uint64_t dx, uint64_t dy, uint64_t dz) {
      uint64_t sc[4] = {c[0] * 2 + dx, c[1] * 2 + dy, c[2] * 2 + dz, L};
      uint64_t off[3] = {dx ? sp[0] : 0, dy ? sp[1] : 0, dz ? sp[2] : 0};
      downsample_brick_to_brick<T, Median>(pData, ts, src, sc, off);
    };
    run(0, 0, 0);
    if (R) run(1, 0, 0);
    if (B) run(0, 1, 0);
    if (K) run(0, 0, 1);
    if (R && B) run(1, 1, 0);
    if (R && K) run(1, 0, 1);
    if (B && K) run(0, 1, 1);
    if (R && B && K) run(1, 1, 1);
    t.set_brick(c, (uint8_t*)pData);
  }

  // ExtendedOctreeConverter.cpp:1153-1181
  void copy_b2b(const std::vector<uint8_t>& s, const uint64_t sbs[3], std::vector<uint8_t>& d,
                const uint64_t tbs[3], const uint64_t so[3], const uint64_t to[3], const uint64_t rs[3]) {
    for (uint64_t z = 0; z < rs[2]; z++)
      for (uint64_t y = 0; y < rs[1]; y++) {
        size_t sp = vox * (size_t)(so[0] + (so[1] + y) * sbs[0] + (so[2] + z) * sbs[0] * sbs[1]);
        size_t tp = vox * (size_t)(to[0] + (to[1] + y) * tbs[0] + (to[2] + z) * tbs[0] * tbs[1]);
        std::memcpy(&d[tp], &s[sp], (size_t)(vox * rs[0]));
      }
  }

  // ExtendedOctreeConverter.cpp:1203-1372
  void fill_overlap(uint64_t lod) {
    const uint64_t* bb = t.lods[lod].bricks;
    const uint64_t o = ov();
    size_t maxb = (size_t)(t.brick[0] * t.brick[1] * t.brick[2] * vox);
    std::vector<uint8_t> T(maxb), S(maxb);
    for (uint64_t z = 0; z < bb[2]; z++) {
      bool Fr = z > 0, Bk = z < bb[2] - 1;
      for (uint64_t y = 0; y < bb[1]; y++) {
        bool Tp = y > 0, Bt = y < bb[1] - 1;
        for (uint64_t x = 0; x < bb[0]; x++) {
          bool Lf = x > 0, Rt = x < bb[0] - 1;
          uint64_t c[4] = {x, y, z, lod}, tbs[3];
          t.brick_size(c, tbs);
          t.get_brick(c, T.data());
          auto from = [&](int64_t dx, int64_t dy, int64_t dz, uint64_t sx, uint64_t sy, uint64_t sz,
                          uint64_t tx, uint64_t ty, uint64_t tz, int rx, int ry, int rz) {
            uint64_t sc[4] = {x + dx, y + dy, z + dz, lod}, sbs[3];
            t.brick_size(sc, sbs);
            t.get_brick(sc, S.data());
            uint64_t so[3] = {sx, sy, sz}, to[3] = {tx, ty, tz};
            // region extents: -1 means "full source extent along that axis"
            uint64_t rs[3] = {rx < 0 ? sbs[0] : (uint64_t)rx, ry < 0 ? sbs[1] : (uint64_t)ry,
                              rz < 0 ? sbs[2] : (uint64_t)rz};
            copy_b2b(S, sbs, T, tbs, so, to, rs);
          };
          const int OV = (int)o;
          if (Rt) from(1, 0, 0, o, 0, 0, tbs[0] - o, 0, 0, OV, -1, -1);
          if (Bt) from(0, 1, 0, 0, o, 0, 0, tbs[1] - o, 0, -1, OV, -1);
          if (Bk) from(0, 0, 1, 0, 0, o, 0, 0, tbs[2] - o, -1, -1, OV);
          if (Lf) {
            uint64_t sc[4] = {x - 1, y, z, lod}, sbs[3];
            t.brick_size(sc, sbs);
            from(-1, 0, 0, sbs[0] - o * 2, 0, 0, 0, 0, 0, OV, -1, -1);
          }
          if (Tp) {
            uint64_t sc[4] = {x, y - 1, z, lod}, sbs[3];
            t.brick_size(sc, sbs);
            from(0, -1, 0, 0, sbs[1] - o * 2, 0, 0, 0, 0, -1, OV, -1);
          }
          if (Fr) {
            uint64_t sc[4] = {x, y, z - 1, lod}, sbs[3];
            t.brick_size(sc, sbs);
            from(0, 0, -1, 0, 0, sbs[2] - o * 2, 0, 0, 0, -1, -1, OV);
          }
          if (Bt && Rt) from(1, 1, 0, o, o, 0, tbs[0] - o, tbs[1] - o, 0, OV, OV, -1);
          if (Rt && Bk) from(1, 0, 1, o, 0, o, tbs[0] - o, 0, tbs[2] - o, OV, -1, OV);
          if (Bt && Bk) from(0, 1, 1, 0, o, o, 0, tbs[1] - o, tbs[2] - o, -1, OV, OV);
          if (Rt && Bt && Bk) from(1, 1, 1, o, o, o, tbs[0] - o, tbs[1] - o, tbs[2] - o, OV, OV, OV);
          if (clamp) clamp_to_edge(T, !Lf, !Tp, !Fr, !Rt, !Bt, !Bk, tbs);
          t.set_brick(c, T.data());
        }
      }
    }
  }

  // ExtendedOctreeConverter.inc:358-398
  template <class T, bool Median>
  void compute_hierarchy() {
    size_t n = (size_t)(t.brick[0] * t.brick[1] * t.brick[2]);
    std::vector<T> a(n), b(n);
    for (size_t lod = 1; lod < t.lods.size(); lod++) {
      const uint64_t* bb = t.lods[lod].bricks;
      for (uint64_t z = 0; z < bb[2]; z++)
        for (uint64_t y = 0; y < bb[1]; y++)
          for (uint64_t x = 0; x < bb[0]; x++) {
            uint64_t c[4] = {x, y, z, lod};
            downsample_brick<T, Median>(c, a.data(), b.data());
          }
      fill_overlap(lod);
    }
  }
};

template <class T>
void brick_stats(const uint8_t* p, size_t bytes, double mm[2]) {
  // ExtendedOctreeConverter.inc:408-444 (one component)
  mm[0] = std::numeric_limits<double>::max();
  mm[1] = -std::numeric_limits<double>::max();
  const T* e = reinterpret_cast<const T*>(p);
  size_t n = bytes / sizeof(T);
  for (size_t i = 0; i < n; ++i) {
    double cur = (double)e[i];
    mm[0] = cur < mm[0] ? cur : mm[0];
    mm[1] = cur > mm[1] ? cur : mm[1];
  }
}

}  // namespace

void Octree::get_brick(const uint64_t c[4], uint8_t* out) const {
  const TocEntry& e = toc[coords_to_index(c)];
  std::memcpy(out, data.data() + e.offset, (size_t)e.bytes);
}
void Octree::set_brick(const uint64_t c[4], const uint8_t* in) {
  const TocEntry& e = toc[coords_to_index(c)];
  std::memcpy(data.data() + e.offset, in, (size_t)e.bytes);
}

void Octree::convert(const uint8_t* in, bool clamp, bool median, bool lod0_only) {
  compute_metadata();
  // TOC in index order; every brick stored uncompressed and back to back.
  toc.clear();
  uint64_t off = 0;
  const size_t nl = lod0_only ? 1 : lods.size();
  for (size_t lod = 0; lod < nl; ++lod) {
    const uint64_t* bb = lods[lod].bricks;
    for (uint64_t z = 0; z < bb[2]; z++)
      for (uint64_t y = 0; y < bb[1]; y++)
        for (uint64_t x = 0; x < bb[0]; x++) {
          TocEntry e;
          e.coords[0] = x; e.coords[1] = y; e.coords[2] = z; e.coords[3] = lod;
          brick_size(e.coords, e.size);
          e.bytes = e.size[0] * e.size[1] * e.size[2] * bpv;
          e.offset = off;
          off += e.bytes;
          toc.push_back(e);
        }
  }
  data.assign((size_t)off, 0);
  Conv cv(*this, in, clamp);
  cv.permute_input();
  if (lod0_only) {
  } else if (bpv == 1) {
    if (median) cv.compute_hierarchy<uint8_t, true>(); else cv.compute_hierarchy<uint8_t, false>();
  } else {
    if (median) cv.compute_hierarchy<uint16_t, true>(); else cv.compute_hierarchy<uint16_t, false>();
  }
  for (auto& e : toc) {
    if (bpv == 1) brick_stats<uint8_t>(data.data() + e.offset, (size_t)e.bytes, e.minmax);
    else brick_stats<uint16_t>(data.data() + e.offset, (size_t)e.bytes, e.minmax);
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_octree_meta(const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                    uint64_t* lod_size, uint64_t* lod_bricks, uint64_t* lod_offset,
                    double* lod_aspect, int maxl) {
  Octree t;
  for (int i = 0; i < 3; ++i) { t.dims[i] = dims[i]; t.brick[i] = brick[i]; }
  t.overlap = overlap;
  try { t.compute_metadata(); } catch (...) { return -1; }
  int n = (int)t.lods.size();
  for (int l = 0; l < n && l < maxl; ++l) {
    for (int i = 0; i < 3; ++i) {
      lod_size[l * 3 + i] = t.lods[l].size[i];
      lod_bricks[l * 3 + i] = t.lods[l].bricks[i];
      lod_aspect[l * 3 + i] = t.lods[l].aspect[i];
    }
    lod_offset[l] = t.lods[l].offset;
  }
  return n;
}

orc_ds* orc_ds_create(const void* voxels, int bpv, const uint64_t dims[3], const uint32_t brick[3],
                      uint32_t overlap, int clamp, int median, const float scale[3]) {
  if (bpv != 1 && bpv != 2) return nullptr;
  auto* ds = new orc_ds;
  Octree& t = ds->tree;
  for (int i = 0; i < 3; ++i) { t.dims[i] = dims[i]; t.brick[i] = brick[i]; ds->scale[i] = scale ? scale[i] : 1.f; }
  t.overlap = overlap;
  t.bpv = (uint32_t)bpv;
  try {
    t.convert((const uint8_t*)voxels, clamp != 0, median != 0);
  } catch (...) {
    delete ds;
    return nullptr;
  }
  ds->build_table();
  return ds;
}

orc_ds* orc_ds_create_lod0(const void* voxels, int bpv, const uint64_t dims[3], const uint32_t brick[3],
                           uint32_t overlap, int clamp, const float scale[3]) {
  if (bpv != 1 && bpv != 2) return nullptr;
  auto* ds = new orc_ds;
  Octree& t = ds->tree;
  for (int i = 0; i < 3; ++i) { t.dims[i] = dims[i]; t.brick[i] = brick[i]; ds->scale[i] = scale ? scale[i] : 1.f; }
  t.overlap = overlap;
  t.bpv = (uint32_t)bpv;
  try {
    if (voxels) t.convert((const uint8_t*)voxels, clamp != 0, false, true);
    else t.compute_metadata();   // metadata and brick table only (no voxels)
  } catch (...) {
    delete ds;
    return nullptr;
  }
  ds->build_table();
  return ds;
}

void orc_ds_destroy(orc_ds* ds) { delete ds; }
int orc_ds_lod_count(const orc_ds* ds) { return (int)ds->tree.lods.size(); }
uint64_t orc_ds_brick_count(const orc_ds* ds) { return ds->tree.toc.size(); }
void orc_ds_brick_info(const orc_ds* ds, uint64_t i, uint64_t coords[4], uint32_t size[3],
                       uint64_t* off, double mm[2]) {
  const TocEntry& e = ds->tree.toc[i];
  for (int k = 0; k < 4; ++k) coords[k] = e.coords[k];
  for (int k = 0; k < 3; ++k) size[k] = (uint32_t)e.size[k];
  *off = e.offset;
  mm[0] = e.minmax[0]; mm[1] = e.minmax[1];
}
const void* orc_ds_brick_data(const orc_ds* ds, uint64_t i) {
  return ds->tree.data.data() + ds->tree.toc[i].offset;
}
uint64_t orc_ds_total_bytes(const orc_ds* ds) { return ds->tree.data.size(); }
const void* orc_ds_all_data(const orc_ds* ds) { return ds->tree.data.data(); }

void orc_ds_brick_md(const orc_ds* ds, int lod, uint64_t lin, float center[3], float extents[3],
                     uint32_t nvox[3]) {
  const BrickMD& m = ds->md_at(lod, lin);
  for (int i = 0; i < 3; ++i) { center[i] = m.center[i]; extents[i] = m.extents[i]; nvox[i] = m.n_voxels[i]; }
}

void orc_ds_texcoords(const orc_ds* ds, int lod, uint64_t lin, int pow2, float tmin[3], float tmax[3]) {
  V3 a, b;
  ds->texcoords(lod, lin, pow2 != 0, a, b);
  for (int i = 0; i < 3; ++i) { tmin[i] = a[i]; tmax[i] = b[i]; }
}

double orc_ds_range_max(const orc_ds* ds) {
  double m = -std::numeric_limits<double>::max();
  for (const auto& e : ds->tree.toc)
    if (e.coords[3] == 0) m = std::max(m, e.minmax[1]);
  return m;
}

float orc_round_fp16(float f) { return round_fp16(f); }

// VolumeTools::Filter as the pyramid builder applies it: n groups of `taps`
// values (2, 4 or 8), mean (double, truncated) or median; bpv 1 or 2.
void orc_filter(int taps, int median, int bpv, const void* in, uint64_t n, void* out) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    const T* s = static_cast<const T*>(in);
    T* d = static_cast<T*>(out);
    for (uint64_t i = 0; i < n; ++i) {
      const T* v = s + i * (uint64_t)taps;
      if (taps == 8)
        d[i] = median ? Conv::filter8<T, true>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7])
                      : Conv::filter8<T, false>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
      else if (taps == 4)
        d[i] = median ? Conv::filter4<T, true>(v[0], v[1], v[2], v[3]) : Conv::filter4<T, false>(v[0], v[1], v[2], v[3]);
      else
        d[i] = median ? Conv::filter2<T, true>(v[0], v[1]) : Conv::filter2<T, false>(v[0], v[1]);
    }
  };
  if (bpv == 1) run(uint8_t()); else run(uint16_t());
}

// iteration order of the dataset's brick table: (lod, 1D index) per brick
uint64_t orc_ds_table_order(const orc_ds* ds, uint64_t* out, uint64_t cap) {
  uint64_t i = 0;
  for (auto it = ds->table.begin(); it != ds->table.end(); ++it, ++i)
    if (i < cap) {
      out[2 * i] = std::get<1>(it->first);
      out[2 * i + 1] = std::get<2>(it->first);
    }
  return i;
}

}  // extern "C"

void orc_ds::build_table() {
  // uvfDataset.cpp:242-288.  Insertion order (x outer, y, z inner) is kept so
  // that an unordered_map with the same hash sees the same history.
  const Octree& t = tree;
  size_t n = 0;
  for (auto& l : t.lods) n += (size_t)(l.bricks[0] * l.bricks[1] * l.bricks[2]);
  table.rehash((size_t)std::ceil(float(n) / table.max_load_factor()));
  for (size_t j = 0; j < t.lods.size(); ++j) {
    const uint64_t* bc = t.lods[j].bricks;
    V3 corner(0, 0, 0);
    BrickMD bmd;
    for (uint64_t x = 0; x < bc[0]; x++) {
      corner.y = 0;
      for (uint64_t y = 0; y < bc[1]; y++) {
        corner.z = 0;
        for (uint64_t z = 0; z < bc[2]; z++) {
          uint64_t c[4] = {x, y, z, (uint64_t)j};
          Key k(0, j, (size_t)(z * bc[0] * bc[1] + y * bc[0] + x));
          V3 asp((float)t.lods[j].aspect[0], (float)t.lods[j].aspect[1], (float)t.lods[j].aspect[2]);
          V3 nds = V3((float)t.lods[j].size[0], (float)t.lods[j].size[1], (float)t.lods[j].size[2]) * asp;
          float maxVal = nds.maxVal();
          nds = nds / maxVal;
          uint64_t bs[3];
          t.brick_size(c, bs);
          V3 eff((float)(bs[0] - 2 * t.overlap), (float)(bs[1] - 2 * t.overlap), (float)(bs[2] - 2 * t.overlap));
          bmd.extents = eff * asp / maxVal;
          bmd.center = (corner + bmd.extents / 2.0f) - nds * 0.5f;
          bmd.n_voxels[0] = (uint32_t)bs[0]; bmd.n_voxels[1] = (uint32_t)bs[1]; bmd.n_voxels[2] = (uint32_t)bs[2];
          table.insert(std::make_pair(k, bmd));
          corner.z += bmd.extents.z;
        }
        corner.y += bmd.extents.y;
      }
      corner.x += bmd.extents.x;
    }
  }
}

const BrickMD& orc_ds::md_at(int lod, uint64_t lin) const {
  return table.at(Key(0, (size_t)lod, (size_t)lin));
}

void orc_ds::texcoords(int lod, uint64_t lin, bool pow2, V3& tmin, V3& tmax) const {
  // uvfDataset.cpp:1747-1774
  const BrickMD& m = md_at(lod, lin);
  const Lod& l = tree.lods[lod];
  V3 asp((float)l.aspect[0], (float)l.aspect[1], (float)l.aspect[2]);
  float ov = (float)tree.overlap;
  V3 nv((float)m.n_voxels[0], (float)m.n_voxels[1], (float)m.n_voxels[2]);
  if (pow2) {
    V3 real((float)next_pow2(m.n_voxels[0]), (float)next_pow2(m.n_voxels[1]), (float)next_pow2(m.n_voxels[2]));
    tmin = V3(ov, ov, ov) / real;
    tmax = V3(1, 1, 1) - tmin;
    tmax = tmax - (real - nv) / real;
  } else {
    tmin = V3(ov, ov, ov) / nv;
    tmax = V3(1, 1, 1) - tmin;
  }
  tmax = tmax * asp;
}


