// Host-side check of the dispatch order of a multi-view render
// (tvk::group_costs, tvk::dispatch_order): the group costs against a brute-force
// sum over the bricks' pixel rectangles, and for both orders (single groups,
// rows) a permutation of the launch's groups in 8 contiguous ranges of about
// equal group count, each range in decreasing cost with ties in launch order.
// Prints "ok" and exits 0, or names the first failed check and exits 1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "host/host.hpp"

using namespace tvk;

static int fails = 0;
static void check(bool ok, const char* what, int a = 0, int b = 0) {
  if (!ok && fails++ == 0) std::printf("FAILED: %s (%d, %d)\n", what, a, b);
}

int main() {
  // three views of different sizes (partial 16x16 groups), bricks of two sizes
  const int WH[3][2] = {{200, 136}, {64, 48}, {333, 17}};
  std::vector<ViewConst> views(3);
  std::vector<FrameBrick> list;
  std::srand(7);
  uint32_t groups = 0;
  for (int v = 0; v < 3; ++v) {
    ViewConst& vc = views[v];
    vc = ViewConst{};
    vc.W = (uint32_t)WH[v][0];
    vc.H = (uint32_t)WH[v][1];
    vc.groups_x = (vc.W + 15) / 16;
    vc.groups = vc.groups_x * ((vc.H + 15) / 16);
    vc.group_offset = groups;
    groups += vc.groups;
    vc.list_offset = (uint32_t)list.size();
    for (int k = 0; k < 40; ++k) {
      FrameBrick f{};
      const int x0 = std::rand() % (int)vc.W - 20, y0 = std::rand() % (int)vc.H - 20;
      const int x1 = x0 + std::rand() % 90, y1 = y0 + std::rand() % 70;
      f.rect[0] = (int16_t)std::max(0, x0);
      f.rect[1] = (int16_t)std::max(0, y0);
      f.rect[2] = (int16_t)std::min((int)vc.W - 1, x1);
      f.rect[3] = (int16_t)std::min((int)vc.H - 1, y1);
      if (k % 9 == 0) { f.rect[0] = 1; f.rect[2] = 0; f.rect[1] = 1; f.rect[3] = 0; }   // off screen
      const float e = (k % 2) ? 0.25f : 0.125f;
      f.bmin[0] = f.bmin[1] = f.bmin[2] = 0.0f;
      f.bmax[0] = e; f.bmax[1] = e * 0.5f; f.bmax[2] = e;
      f.ray_step = 0.002f;
      list.push_back(f);
    }
    vc.nlist = (uint32_t)list.size() - vc.list_offset;
  }
  std::vector<float> cost(groups, 0.0f);
  for (int v = 0; v < 3; ++v) group_costs(views[v], list.data() + views[v].list_offset, cost.data() + views[v].group_offset);
  // brute force: the rectangles covering each group's pixels, weight = longest edge / ray step
  for (int v = 0; v < 3; ++v) {
    const ViewConst& vc = views[v];
    double unit = 0.0;
    for (uint32_t i = 0; i < vc.nlist; ++i) {
      const FrameBrick& f = list[vc.list_offset + i];
      if (f.rect[2] < f.rect[0] || f.rect[3] < f.rect[1]) continue;
      unit = std::max(unit, (double)(std::max(f.bmax[0] - f.bmin[0], std::max(f.bmax[1] - f.bmin[1], f.bmax[2] - f.bmin[2])) / f.ray_step));
    }
    for (uint32_t g = 0; g < vc.groups; ++g) {
      const int gx = (int)(g % vc.groups_x), gy = (int)(g / vc.groups_x);
      double want = 0.05 * unit;
      for (uint32_t i = 0; i < vc.nlist; ++i) {
        const FrameBrick& f = list[vc.list_offset + i];
        if (f.rect[2] < f.rect[0] || f.rect[3] < f.rect[1]) continue;
        if (f.rect[2] < gx * 16 || f.rect[0] > gx * 16 + 15 || f.rect[3] < gy * 16 || f.rect[1] > gy * 16 + 15) continue;
        want += (double)(std::max(f.bmax[0] - f.bmin[0], std::max(f.bmax[1] - f.bmin[1], f.bmax[2] - f.bmin[2])) / f.ray_step);
      }
      const float got = cost[vc.group_offset + g];
      check(std::fabs((double)got - want) <= 1e-3 * std::max(1.0, want), "group cost", (int)g, v);
    }
  }
  for (int mode = 1; mode <= 2; ++mode) {
    std::vector<uint32_t> order;
    uint32_t start[8], count[8];
    dispatch_order(views.data(), 3, cost, mode, order, start, count);
    check(order.size() == groups, "order size", (int)order.size(), (int)groups);
    std::vector<int> seen(groups, 0);
    for (uint32_t g : order) { check(g < groups, "group index", (int)g); if (g < groups) ++seen[g]; }
    for (uint32_t g = 0; g < groups; ++g) check(seen[g] == 1, "permutation", (int)g, seen[g]);
    uint32_t at = 0;
    for (int x = 0; x < 8; ++x) {
      check(start[x] == at, "contiguous ranges", x, (int)start[x]);
      at += count[x];
      // each range holds the groups between its first and last launch index (whole rows in mode 2)
      uint32_t lo = groups, hi = 0;
      for (uint32_t j = start[x]; j < start[x] + count[x]; ++j) { lo = std::min(lo, order[j]); hi = std::max(hi, order[j]); }
      if (count[x]) check(hi - lo + 1 == count[x], "range is a contiguous run of groups", x);
      // about an eighth of the groups (a row's worth of slack in mode 2)
      check(std::abs((int)count[x] - (int)(groups / 8)) <= (mode == 2 ? 24 : 2) || x == 7, "balanced count", x, (int)count[x]);
      if (mode == 1) {
        for (uint32_t j = start[x] + 1; j < start[x] + count[x]; ++j) {
          const float a = cost[order[j - 1]], b = cost[order[j]];
          // decreasing in 1024 levels of the range's top cost (ties in launch order)
          check(a >= b - 1e-3f * cost[order[start[x]]] || order[j - 1] < order[j], "decreasing cost", (int)j);
        }
      }
    }
    check(at == groups, "ranges cover the launch", (int)at, (int)groups);
  }
  if (fails) return 1;
  std::printf("ok\n");
  return 0;
}
