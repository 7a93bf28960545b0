// A C++ client of the HIP raycaster written the way a Tuvok renderer subclass
// drives it (include/tvk_raycaster.hpp): the bricks of a dataset are handed
// over one by one with their keys and min/max (BrickedDataset::GetBrick /
// MaxMinForKey), then one frame is painted.
//
// usage: ingest_render <dir> [budget_bytes]
//   <dir>/meta.txt   dtype dims[3] brick[3] overlap nbricks width height eye[3] fov lit
//   <dir>/keys.bin   per brick: uint32 lod, uint64 index, double min, double max
//   <dir>/bricks.bin the payloads in key order (x fastest, stored size)
//   <dir>/tf1d.bin   256 RGBA8 entries
// writes <dir>/frame.f32 (height x width x RGBA32F, premultiplied, GL row order)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>
#include "tvk_raycaster.hpp"

static std::vector<char> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: ingest_render <dir> [budget_bytes]\n";
    return 2;
  }
  const std::string dir = argv[1];
  const uint64_t budget = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 0;
  try {
    std::ifstream meta(dir + "/meta.txt");
    int dtype = 0, lit = 0;
    uint64_t dims[3];
    uint32_t brick[3], overlap = 0, w = 0, h = 0;
    size_t nb = 0;
    float eye[3], fov = 45.f;
    meta >> dtype >> dims[0] >> dims[1] >> dims[2] >> brick[0] >> brick[1] >> brick[2] >> overlap >> nb >> w >> h >>
        eye[0] >> eye[1] >> eye[2] >> fov >> lit;
    if (!meta) throw std::runtime_error("bad meta.txt");
    const std::vector<char> kb = slurp(dir + "/keys.bin");
    const size_t rec = 4 + 8 + 8 + 8;
    if (kb.size() != nb * rec) throw std::runtime_error("keys.bin size");
    std::vector<tvk_brick_key> keys(nb);
    std::vector<size_t> bytes(nb);
    for (size_t i = 0; i < nb; ++i) {
      const char* r = kb.data() + i * rec;
      std::memcpy(&keys[i].lod, r, 4);
      std::memcpy(&keys[i].index, r + 4, 8);
      std::memcpy(&keys[i].minmax[0], r + 12, 8);
      std::memcpy(&keys[i].minmax[1], r + 20, 8);
    }
    const std::vector<char> payload = slurp(dir + "/bricks.bin");
    const std::vector<char> tfb = slurp(dir + "/tf1d.bin");

    tvk::Raycaster ren(0);
    ren.SetBricks(dtype, dims, brick, overlap, keys, 0, budget);
    // payload sizes follow from the layout: walk them in key order
    size_t off = 0;
    for (size_t i = 0; i < nb; ++i) {
      // the stored size of brick i: ask the library for its TOC entry
      uint64_t coords[4];
      uint32_t size[3];
      uint64_t pool_off = 0;
      tvk_lod_info(ren.handle(), (int)keys[i].lod, nullptr, nullptr, &pool_off, nullptr);
      const uint64_t toc = pool_off + keys[i].index;
      if (tvk_brick_info(ren.handle(), toc, coords, size, nullptr) != TVK_OK) throw std::runtime_error("brick_info");
      const size_t b = (size_t)size[0] * size[1] * size[2] * (size_t)dtype;
      if (off + b > payload.size()) throw std::runtime_error("bricks.bin too short");
      ren.UploadBrick(keys[i].lod, keys[i].index, payload.data() + off, b);
      off += b;
    }
    if (off != payload.size()) throw std::runtime_error("bricks.bin too long");
    ren.Set1DTrans(std::vector<unsigned char>(tfb.begin(), tfb.end()));
    ren.SetRendermode(TVK_RM_1DTRANS);
    ren.SetUseLighting(lit != 0);
    ren.SetViewParameters(fov, 0.01f, 1000.f);
    ren.SetViewPosition({eye[0], eye[1], eye[2]});
    ren.Resize(w, h);
    tvk_stats st{};
    const std::vector<float> img = ren.Paint(&st);
    std::ofstream out(dir + "/frame.f32", std::ios::binary);
    out.write(reinterpret_cast<const char*>(img.data()), (std::streamsize)(img.size() * sizeof(float)));
    const tvk_pool_stats ps = ren.PoolStats();
    std::printf("ingest_render: %zu bricks, %u x %u frame, %u bricks drawn, %llu slots, %u parts\n", nb, w, h,
                st.bricks_nonempty, (unsigned long long)ps.slots, ps.parts);
  } catch (const std::exception& e) {
    std::cerr << "ingest_render: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
