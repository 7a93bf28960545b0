// UVF / ExtendedOctree file reader feeding the brick pool.
//
// The on-disk format as the reference reads it:
//   - "UVF-DATA", then GlobalHeader::GetHeaderFromFile (IO/UVF/GlobalHeader.cpp:38-48):
//     endianness byte, file version, checksum semantics, checksum length and
//     bytes, offset to the first data block (from the end of the header,
//     GlobalHeader::GetDataPos, :34-36);
//   - data blocks (UVF::ParseDataBlocks, IO/UVF/UVF.cpp:269-291;
//     DataBlock::GetHeaderFromFile, IO/UVF/DataBlock.cpp:52-70): block id
//     length and string, semantics, compression, offset to the next block (0
//     for the last);
//   - a TOC block (BS_TOC_BLOCK, UVFTables.h:24-36) holds an ExtendedOctree
//     right after its block header (TOCBlock::GetHeaderFromFile, IO/UVF/TOCBlock.cpp:48-59);
//   - ExtendedOctree::Open (IO/UVF/ExtendedOctree/ExtendedOctree.cpp:87-165),
//     in the writing machine's byte order (the octree reads and writes with
//     the host's order, LargeRAWFile.h:78-90, so little endian on x86):
//     component type and count, precomputed-normals flag, volume size, aspect,
//     maximum brick size, overlap, version (UVF 5), header+data size (version
//     >= 1), compression level (version >= 2), then one TOC entry per brick in
//     ExtendedOctree index order (offset from the octree, length, compression,
//     valid length, atlas size);
//   - ExtendedOctree::GetBrickData (:313-360): the brick at octree offset +
//     entry offset, uncompressed or a zlib stream.
// The bricks go to tvk_set_bricks / tvk_upload_brick (resident pool, or a
// budgeted pool that pages them into HBM): the file is read twice, once for
// every brick's min/max (MaxMinForKey) and once for the payloads.
#include <zlib.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../ctx.hpp"

using tvk::fail;

namespace {

constexpr uint64_t kBlockToc = 9;          // UVFTables::BS_TOC_BLOCK
constexpr uint32_t kCompNone = 0, kCompZlib = 1;   // COMPRESSION_TYPE CT_NONE, CT_ZLIB

struct File {
  FILE* f = nullptr;
  bool be = false;   // the file's byte order is big endian
  bool ok = true;
  ~File() { if (f) std::fclose(f); }
  void seek(uint64_t pos) { if (ok) ok = fseeko(f, (off_t)pos, SEEK_SET) == 0; }
  uint64_t tell() { return (uint64_t)ftello(f); }
  void raw(void* p, size_t n) { if (ok) ok = std::fread(p, 1, n, f) == n; }
  template <typename T> T get() {
    uint8_t b[sizeof(T)] = {};
    raw(b, sizeof(T));
    if (be) std::reverse(b, b + sizeof(T));   // LargeRAWFile::ReadData swaps to the host's order
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
  }
};

struct Octree {
  uint32_t comp_type = 0;
  uint64_t comp_count = 0;
  uint8_t normals = 0;
  uint64_t size[3] = {0, 0, 0};
  double aspect[3] = {1, 1, 1};
  uint64_t brick[3] = {0, 0, 0};
  uint32_t overlap = 0, version = 0, level = 0;
  uint64_t total = 0;
  struct Entry { uint64_t offset, length, valid; uint32_t comp, atlas[2]; };
  std::vector<Entry> toc;
};

}  // namespace

extern "C" int tvk_open_uvf(tvk_ctx* c, const char* path, uint64_t budget_bytes, tvk_uvf_info* info) {
  if (!c || !path) return TVK_ERR_ARG;
  File F;
  F.f = std::fopen(path, "rb");
  if (!F.f) return fail(c, TVK_ERR_ARG, std::string("cannot open ") + path);
  char magic[8];
  F.raw(magic, 8);
  if (!F.ok || std::memcmp(magic, "UVF-DATA", 8) != 0) return fail(c, TVK_ERR_ARG, "file magic not found");
  // GlobalHeader::GetHeaderFromFile
  F.be = F.get<uint8_t>() != 0;
  const uint64_t file_version = F.get<uint64_t>();
  (void)F.get<uint64_t>();                       // checksum semantics
  const uint64_t cs_len = F.get<uint64_t>();
  F.seek(F.tell() + cs_len);
  const uint64_t to_first = F.get<uint64_t>();
  if (!F.ok) return fail(c, TVK_ERR_ARG, "truncated UVF global header");
  if (file_version < 5) return fail(c, TVK_ERR_UNSUPPORTED, "UVF files before version 5 hold no ExtendedOctree");
  // GetDataPos: magic + header (1 + 4 * 8 + checksum) + offset to the first block
  uint64_t at = 8 + 1 + 4 * 8 + cs_len + to_first, octree_at = 0;
  bool found = false;
  for (int guard = 0; guard < 1 << 20; ++guard) {
    F.seek(at);
    const uint64_t idlen = F.get<uint64_t>();
    if (!F.ok || idlen > (1u << 20)) return fail(c, TVK_ERR_ARG, "bad UVF data block header");
    F.seek(F.tell() + idlen);
    const uint64_t sem = F.get<uint64_t>();
    (void)F.get<uint64_t>();                     // block compression
    const uint64_t next = F.get<uint64_t>();
    if (!F.ok) return fail(c, TVK_ERR_ARG, "truncated UVF data block header");
    if (sem == kBlockToc) {
      octree_at = F.tell();
      found = true;
      break;
    }
    if (next == 0) break;
    at += next;
  }
  if (!found) return fail(c, TVK_ERR_UNSUPPORTED, "the file holds no TOC block (ExtendedOctree)");
  // ExtendedOctree::Open: the host's byte order, not the file header's
  Octree O;
  F.be = false;
  F.seek(octree_at);
  O.comp_type = F.get<uint32_t>();
  O.comp_count = F.get<uint64_t>();
  O.normals = F.get<uint8_t>();
  for (int k = 0; k < 3; ++k) O.size[k] = F.get<uint64_t>();
  for (int k = 0; k < 3; ++k) O.aspect[k] = F.get<double>();
  for (int k = 0; k < 3; ++k) O.brick[k] = F.get<uint64_t>();
  O.overlap = F.get<uint32_t>();
  O.version = F.get<uint32_t>();
  if (O.version > 0) O.total = F.get<uint64_t>();
  if (O.version > 1) O.level = F.get<uint32_t>();
  if (!F.ok) return fail(c, TVK_ERR_ARG, "truncated ExtendedOctree header");
  if (O.version == 0) return fail(c, TVK_ERR_ARG, "ExtendedOctree version 0 in a version-5 file (corrupt)");
  if (O.comp_count != 1 || O.normals) return fail(c, TVK_ERR_UNSUPPORTED, "only scalar volumes are rendered");
  if (O.comp_type > 1) return fail(c, TVK_ERR_UNSUPPORTED, "only 8- and 16-bit unsigned voxels are rendered");
  const int bpv = O.comp_type == 0 ? 1 : 2;
  uint32_t brick[3];
  for (int k = 0; k < 3; ++k) {
    if (O.brick[k] == 0 || O.brick[k] > 0xffffffffull) return fail(c, TVK_ERR_ARG, "bad brick size");
    brick[k] = (uint32_t)O.brick[k];
  }
  // the brick table follows from the header (ExtendedOctree::ComputeMetadata)
  tvk::Layout L;
  std::string err;
  if (!L.build(O.size, brick, O.overlap, bpv, 0, &err)) return fail(c, TVK_ERR_ARG, err);
  const size_t nb = L.toc.size();
  O.toc.resize(nb);
  for (size_t i = 0; i < nb; ++i) {
    Octree::Entry& e = O.toc[i];
    e.offset = F.get<uint64_t>();
    e.length = F.get<uint64_t>();
    e.comp = F.get<uint32_t>();
    e.valid = F.get<uint64_t>();
    e.atlas[0] = F.get<uint32_t>();
    e.atlas[1] = F.get<uint32_t>();
  }
  if (!F.ok) return fail(c, TVK_ERR_ARG, "truncated ExtendedOctree TOC");
  uint64_t ncomp[2] = {0, 0};
  for (size_t i = 0; i < nb; ++i) {
    const Octree::Entry& e = O.toc[i];
    if (e.atlas[0] != 0 && e.atlas[1] != 0)
      return fail(c, TVK_ERR_UNSUPPORTED, "bricks stored as 2D atlases are not read");
    if (e.comp != kCompNone && e.comp != kCompZlib)
      return fail(c, TVK_ERR_UNSUPPORTED, "brick codec " + std::to_string(e.comp) + " (only none and zlib are read)");
    ++ncomp[e.comp];
  }
  // one brick, decompressed, into buf (the reference layout: x fastest, stored size)
  std::vector<uint8_t> buf, packed;
  auto load = [&](size_t t) -> int {
    const Octree::Entry& e = O.toc[t];
    const size_t want = (size_t)L.toc[t].real * (size_t)bpv;
    buf.resize(want);
    F.seek(octree_at + e.offset);
    if (e.comp == kCompNone) {
      if (e.length != want) return fail(c, TVK_ERR_ARG, "brick " + std::to_string(t) + " has the wrong length");
      F.raw(buf.data(), want);
    } else {
      packed.resize((size_t)e.length);
      F.raw(packed.data(), packed.size());
      uLongf out = (uLongf)want;
      if (F.ok && (uncompress(buf.data(), &out, packed.data(), (uLong)packed.size()) != Z_OK || out != want))
        return fail(c, TVK_ERR_ARG, "brick " + std::to_string(t) + ": zlib stream does not inflate to the brick");
    }
    if (!F.ok) return fail(c, TVK_ERR_ARG, "truncated brick data");
    return TVK_OK;
  };
  // pass 1: MaxMinForKey of every brick over its stored voxels
  std::vector<tvk_brick_key> keys(nb);
  for (size_t t = 0; t < nb; ++t) {
    const int rc = load(t);
    if (rc) return rc;
    uint32_t mn = 0xffffffffu, mx = 0u;
    if (bpv == 1) {
      for (uint8_t v : buf) { mn = std::min<uint32_t>(mn, v); mx = std::max<uint32_t>(mx, v); }
    } else {
      const size_t n = buf.size() / 2;
      for (size_t i = 0; i < n; ++i) {
        uint16_t v;
        std::memcpy(&v, buf.data() + 2 * i, 2);
        mn = std::min<uint32_t>(mn, v);
        mx = std::max<uint32_t>(mx, v);
      }
    }
    const uint64_t* co = L.toc[t].coords;
    const tvk::LodInfo& li = L.lods[co[3]];
    keys[t].lod = (uint32_t)co[3];
    keys[t].index = co[0] + li.bricks[0] * (co[1] + li.bricks[1] * co[2]);
    keys[t].minmax[0] = (double)mn;
    keys[t].minmax[1] = (double)mx;
  }
  const int dtype = bpv == 1 ? TVK_U8 : TVK_U16;
  int rc = tvk_set_bricks(c, dtype, O.size, brick, O.overlap, 0, keys.data(), nb, budget_bytes);
  if (rc) return rc;
  // pass 2: the payloads
  for (size_t t = 0; t < nb; ++t) {
    rc = load(t);
    if (rc) return rc;
    rc = tvk_upload_brick(c, keys[t].lod, keys[t].index, buf.data(), buf.size());
    if (rc) return rc;
  }
  // the dataset's voxel aspect (UVFDataset's domain scale)
  for (int k = 0; k < 3; ++k) c->L.scale[k] = (float)O.aspect[k];
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->file_version = file_version;
    info->octree_version = O.version;
    info->octree_offset = octree_at;
    for (int k = 0; k < 3; ++k) {
      info->dims[k] = O.size[k];
      info->brick_size[k] = brick[k];
      info->aspect[k] = O.aspect[k];
    }
    info->overlap = O.overlap;
    info->bit_width = 8 * bpv;
    info->lod_count = (int)L.lods.size();
    info->brick_count = nb;
    info->zlib_bricks = ncomp[kCompZlib];
  }
  return TVK_OK;
}
