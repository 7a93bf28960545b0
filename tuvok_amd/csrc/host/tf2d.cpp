// 2D transfer function from swatches: TransferFunction2D::RenderTransferFunction8Bit
// (IO/TransferFunction2D.cpp:294-361) without Qt.  The reference paints, into a
// canvas of GetRenderSize() (TransferFunction2D.h:136-144), the 1D TF stretched
// over the whole canvas, then every swatch -- a polygon filled with a linear or
// radial gradient -- over it, and scales the canvas to the table size; under
// TUVOK_NO_QT it returns an empty table (:353-359).  This restatement fixes the
// painting rules Qt leaves to its raster engine:
//   - canvas: 8-bit RGBA, not premultiplied, cleared to 0; image scaling
//     (drawImage into the canvas rectangle, QImage::scaled to the table size)
//     takes the nearest source pixel of each destination pixel centre;
//   - polygons: vertices at int(normalised coordinate * canvas size)
//     (Normalized2Offscreen, :277-280), filled with the odd-even rule
//     (QPainter::drawPolygon's default) at pixel centres, no antialiasing;
//   - gradients: QLinearGradient from point 0 to point 1, t = the projection
//     of the pixel centre onto that segment; QRadialGradient centred at point
//     0 with radius |point 1 - point 0|, t = distance / radius; pad spread
//     (t clamped to [0, 1]); stops sorted by position, colours as
//     QColor(int(c * 255)) and interpolated per channel between the stops
//     around t;
//   - composition: source over, in floats, rounded to 8 bits per pixel.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "tvk_hip.h"

namespace {

struct RGBA {
  float r, g, b, a;   // 0..255, not premultiplied
};

// the colour of a gradient at t (pad spread)
RGBA gradient_color(const tvk_gradient_stop* stops, uint32_t n, float t) {
  RGBA c{0, 0, 0, 0};
  if (n == 0) return c;
  t = std::min(std::max(t, 0.0f), 1.0f);
  auto col = [&](uint32_t i) {
    // QColor(int(v * 255)) per channel
    return RGBA{(float)(int)(stops[i].rgba[0] * 255.0f), (float)(int)(stops[i].rgba[1] * 255.0f),
                (float)(int)(stops[i].rgba[2] * 255.0f), (float)(int)(stops[i].rgba[3] * 255.0f)};
  };
  if (t <= stops[0].pos) return col(0);
  if (t >= stops[n - 1].pos) return col(n - 1);
  uint32_t k = 0;
  while (k + 1 < n && stops[k + 1].pos < t) ++k;
  const float span = stops[k + 1].pos - stops[k].pos;
  const float f = span > 0.0f ? (t - stops[k].pos) / span : 0.0f;
  const RGBA a = col(k), b = col(k + 1);
  return RGBA{a.r + (b.r - a.r) * f, a.g + (b.g - a.g) * f, a.b + (b.b - a.b) * f, a.a + (b.a - a.a) * f};
}

// odd-even rule at the point (x, y)
bool inside(const std::vector<int>& px, const std::vector<int>& py, float x, float y) {
  bool in = false;
  const size_t n = px.size();
  for (size_t i = 0, j = n - 1; i < n; j = i++) {
    const float xi = (float)px[i], yi = (float)py[i], xj = (float)px[j], yj = (float)py[j];
    if ((yi > y) != (yj > y)) {
      const float xc = xi + (y - yi) * (xj - xi) / (yj - yi);
      if (x < xc) in = !in;
    }
  }
  return in;
}

uint8_t to8(float v) { return (uint8_t)std::min(255.0f, std::max(0.0f, std::floor(v + 0.5f))); }

}  // namespace

extern "C" int tvk_rasterize_tf2d(int w, int h, const uint8_t* tf1d, int n1d, const tvk_swatch* sw, int nsw,
                                  const float* points, size_t npoints, const tvk_gradient_stop* stops, size_t nstops,
                                  uint8_t* out) {
  if (w < 1 || h < 1 || w > 4096 || h > 4096 || !out || nsw < 0 || n1d < 0) return TVK_ERR_ARG;
  if (nsw > 0 && (!sw || !points || !stops)) return TVK_ERR_ARG;
  if (n1d > 0 && !tf1d) return TVK_ERR_ARG;
  for (int s = 0; s < nsw; ++s) {
    if ((size_t)sw[s].first_point + sw[s].n_points > npoints) return TVK_ERR_RANGE;
    if ((size_t)sw[s].first_stop + sw[s].n_stops > nstops) return TVK_ERR_RANGE;
  }
  // GetRenderSize
  size_t sx = std::min<size_t>((size_t)w, 4096), sy = std::min<size_t>((size_t)h, 4096);
  size_t RW, RH;
  if (sx > sy) { RW = sx; RH = (size_t)(sx / 2.0); } else { RW = sy * 2; RH = sy; }
  std::vector<RGBA> canvas(RW * RH, RGBA{0, 0, 0, 0});
  // the 1D TF image (min(width, size) entries) stretched over the canvas
  const size_t n1 = std::min<size_t>((size_t)w, (size_t)n1d);
  if (n1 > 0)
    for (size_t y = 0; y < RH; ++y)
      for (size_t x = 0; x < RW; ++x) {
        const size_t i = std::min(n1 - 1, (size_t)(((double)x + 0.5) * (double)n1 / (double)RW));
        canvas[y * RW + x] = RGBA{(float)tf1d[4 * i], (float)tf1d[4 * i + 1], (float)tf1d[4 * i + 2], (float)tf1d[4 * i + 3]};
      }
  for (int s = 0; s < nsw; ++s) {
    const tvk_swatch& S = sw[s];
    if (S.n_points == 0) continue;
    std::vector<int> px(S.n_points), py(S.n_points);
    int x0 = (int)RW, x1 = -1, y0 = (int)RH, y1 = -1;
    for (uint32_t k = 0; k < S.n_points; ++k) {
      px[k] = (int)(points[2 * (S.first_point + k)] * (float)(int)RW);
      py[k] = (int)(points[2 * (S.first_point + k) + 1] * (float)(int)RH);
      x0 = std::min(x0, px[k]); x1 = std::max(x1, px[k]);
      y0 = std::min(y0, py[k]); y1 = std::max(y1, py[k]);
    }
    // stops sorted by position (QGradient keeps them ordered)
    std::vector<tvk_gradient_stop> st(stops + S.first_stop, stops + S.first_stop + S.n_stops);
    std::stable_sort(st.begin(), st.end(), [](const tvk_gradient_stop& a, const tvk_gradient_stop& b) {
      return a.pos < b.pos;
    });
    const float g0x = (float)(int)(S.gradient[0] * (float)(int)RW), g0y = (float)(int)(S.gradient[1] * (float)(int)RH);
    const float g1x = (float)(int)(S.gradient[2] * (float)(int)RW), g1y = (float)(int)(S.gradient[3] * (float)(int)RH);
    const float dx = g1x - g0x, dy = g1y - g0y, len2 = dx * dx + dy * dy;
    const float radius = std::sqrt((float)((double)dx * dx + (double)dy * dy));
    for (int y = std::max(y0, 0); y <= std::min(y1, (int)RH - 1); ++y)
      for (int x = std::max(x0, 0); x <= std::min(x1, (int)RW - 1); ++x) {
        const float cx = (float)x + 0.5f, cy = (float)y + 0.5f;
        if (!inside(px, py, cx, cy)) continue;
        float t;
        if (S.radial) t = radius > 0.0f ? std::sqrt((cx - g0x) * (cx - g0x) + (cy - g0y) * (cy - g0y)) / radius : 1.0f;
        else t = len2 > 0.0f ? ((cx - g0x) * dx + (cy - g0y) * dy) / len2 : 0.0f;
        const RGBA src = gradient_color(st.data(), (uint32_t)st.size(), t);
        RGBA& d = canvas[(size_t)y * RW + (size_t)x];
        const float sa = src.a / 255.0f, da = d.a / 255.0f;
        const float oa = sa + da * (1.0f - sa);
        RGBA o{0, 0, 0, 0};
        if (oa > 0.0f) {
          o.r = (src.r * sa + d.r * da * (1.0f - sa)) / oa;
          o.g = (src.g * sa + d.g * da * (1.0f - sa)) / oa;
          o.b = (src.b * sa + d.b * da * (1.0f - sa)) / oa;
          o.a = oa * 255.0f;
        }
        d = RGBA{(float)to8(o.r), (float)to8(o.g), (float)to8(o.b), (float)to8(o.a)};
      }
  }
  // QImage::scaled to the table size (nearest), RGBA bytes as GetByteArray returns them
  for (int y = 0; y < h; ++y) {
    const size_t cy = std::min(RH - 1, (size_t)(((double)y + 0.5) * (double)RH / (double)h));
    for (int x = 0; x < w; ++x) {
      const size_t cx = std::min(RW - 1, (size_t)(((double)x + 0.5) * (double)RW / (double)w));
      const RGBA& c = canvas[cy * RW + cx];
      uint8_t* o = out + 4 * ((size_t)y * (size_t)w + (size_t)x);
      o[0] = (uint8_t)c.r; o[1] = (uint8_t)c.g; o[2] = (uint8_t)c.b; o[3] = (uint8_t)c.a;
    }
  }
  return TVK_OK;
}

extern "C" int tvk_set_tf2d_swatches(tvk_ctx* c, int w, int h, const uint8_t* tf1d, int n1d, const tvk_swatch* sw,
                                     int nsw, const float* points, size_t npoints, const tvk_gradient_stop* stops,
                                     size_t nstops) {
  if (!c) return TVK_ERR_ARG;
  if (w < 2 || h < 2 || w > 4096 || h > 4096) return TVK_ERR_ARG;
  std::vector<uint8_t> table((size_t)w * (size_t)h * 4);
  const int rc = tvk_rasterize_tf2d(w, h, tf1d, n1d, sw, nsw, points, npoints, stops, nstops, table.data());
  if (rc) return rc;
  return tvk_set_tf2d(c, table.data(), w, h);
}
