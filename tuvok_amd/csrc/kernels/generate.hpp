// Parameters of the seeded synthetic volume generators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tvk {

constexpr int kMaxBlobs = 32;
constexpr int kMaxShells = 64;

struct GenBlobs {
  int n;
  float c[kMaxBlobs][3];
  float amp[kMaxBlobs];
  float inv2s2[kMaxBlobs];  // 1 / (2 sigma^2)
};

struct GenShells {
  int n;
  int octaves;
  float base_freq;
  uint64_t seed;
  float c[kMaxShells][3];
  float r[kMaxShells];
  float half_thickness_vox;  // |r - r_s| * N < half thickness -> shell voxel
  float shell_value;
};

struct GenEllipsoid {
  float c[3];
  float a[3];    // semi-axes
  float amin;    // smallest semi-axis
  float value;   // phantom intensity
};

// host-side counter-based generator (splitmix64) for the object parameters
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform(double lo, double hi) { return lo + (hi - lo) * (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

hipError_t launch_blobs(const GenBlobs& g, const uint64_t d[3], int bpv, void* out, unsigned* d_tmp, uint64_t seed,
                        hipStream_t st);
hipError_t launch_fbm_shells(const GenShells& g, const uint64_t d[3], int bpv, void* out, hipStream_t st);
hipError_t launch_ellipsoid_sdf(const GenEllipsoid* d_e, int ne, const uint64_t d[3], void* out, hipStream_t st);
hipError_t launch_phantom(const GenEllipsoid* d_e, int ne, const uint64_t d[3], void* out, float sigma, uint64_t seed,
                          hipStream_t st);

}  // namespace tvk
