// Batched rendering: many requests (dataset + TF + parameters of a tvk_ctx, a
// view rotation and an image size each) rendered by persistent kernels that
// pull ray tiles from device work queues, one launch per kernel variant
// (mode, lighting, voxel type), replayed as one hipGraph per frame.
//
// The reference renders its RenderRegions one at a time inside
// AbstrRenderer::Paint (GLRenderer.cpp:571-667); a batch gives the same
// per-request frames (same planner, same march) from one graph launch.
#include "ctx.hpp"

using namespace tvk;

namespace tvk {
int batch_occupancy(int mode, bool lit, int bpv, bool count, int learn, int* groups_per_cu);
}

struct tvk_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  struct Req {
    tvk_ctx* ctx;
    Mat4 rot;
    uint32_t W, H;
  };
  std::vector<Req> reqs;
  tvk_tuning tune{};
  // step proposal for the TF-mode requests (nullptr: fixed step)
  struct tvk_stepnet* net = nullptr;
  LearnConst learn{};
  // per-request results of the last successful render: first pixel in d_out
  // and the image size it was rendered at (empty after clear, add or a failure)
  struct Frame {
    uint64_t offset;
    uint32_t W, H;
  };
  std::vector<Frame> frames;
  std::vector<Frame> planned_frames;   // of the render in progress
  // device buffers
  BatchRequest* d_req = nullptr;
  size_t req_cap = 0;
  FrameBrick* d_list = nullptr;
  size_t list_cap = 0;
  float* d_out = nullptr;
  size_t out_cap = 0;
  unsigned int* d_queues = nullptr;
  unsigned long long* d_counter = nullptr;
  uint32_t* d_touched = nullptr;
  size_t touched_cap = 0;
  bool count = false;
  // pinned staging
  BatchRequest* h_req = nullptr;
  size_t h_req_cap = 0;
  FrameBrick* h_list = nullptr;
  size_t h_list_cap = 0;
  // one launch per kernel variant
  struct Launch {
    int mode;
    bool lit;
    int bpv;
    int learn;
    BatchArgs args;
    int groups;
  };
  std::vector<Launch> launches;
  // the captured frame
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> graph_key;
  hipEvent_t ev[2] = {nullptr, nullptr};
  uint64_t last_samples = 0;
  uint32_t last_bricks = 0;
  float last_ms = 0.f;
  int cu_count = 0;
};

namespace {

int bfail(tvk_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}
#define BCHK(b, x)                                                                                    \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess)                                                                             \
      return bfail(b, e_ == hipErrorOutOfMemory ? TVK_ERR_NOMEM : TVK_ERR_HIP,                        \
                   std::string(#x) + ": " + hipGetErrorString(e_));                                   \
  } while (0)

template <typename T>
int bensure(tvk_batch* b, T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return TVK_OK;
  if (*p) { (void)hipFree(*p); *p = nullptr; *cap = 0; }
  const size_t want = std::max<size_t>(n + n / 4, 1);
  BCHK(b, hipMalloc((void**)p, want * sizeof(T)));
  *cap = want;
  return TVK_OK;
}
template <typename T>
int bensure_pinned(tvk_batch* b, T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return TVK_OK;
  if (*p) { (void)hipHostFree(*p); *p = nullptr; *cap = 0; }
  const size_t want = std::max<size_t>(n + n / 4, 16);
  BCHK(b, hipHostMalloc((void**)p, want * sizeof(T), hipHostMallocDefault));
  *cap = want;
  return TVK_OK;
}

void drop_graph(tvk_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

bool valid_tile(uint32_t w, uint32_t h) {
  return w * h == 64 && (w == 1 || w == 2 || w == 4 || w == 8 || w == 16 || w == 32 || w == 64);
}

// Plans every request, fills the staging buffers and the launch list.
int plan_batch(tvk_batch* b, uint32_t* bricks_total) {
  const uint32_t tw = b->tune.tile_w, th = b->tune.tile_h;
  struct Planned {
    BatchRequest br;
    int mode;
    bool lit;
    int bpv;
    uint32_t ntoc;
    std::vector<FrameBrick> list;
  };
  std::vector<Planned> P(b->reqs.size());
  b->frames.clear();
  b->planned_frames.assign(b->reqs.size(), tvk_batch::Frame{0, 0, 0});
  uint64_t pixels = 0;
  std::vector<PlannedBrick> planned;
  for (size_t i = 0; i < b->reqs.size(); ++i) {
    const tvk_batch::Req& r = b->reqs[i];
    tvk_ctx* c = r.ctx;
    if (!c->has_ds) return bfail(b, TVK_ERR_STATE, "request " + std::to_string(i) + ": context has no dataset");
    if (!c->has_view) return bfail(b, TVK_ERR_STATE, "request " + std::to_string(i) + ": context has no view");
    if (budgeted(c) || c->ingest_pending)
      return bfail(b, TVK_ERR_UNSUPPORTED, "request " + std::to_string(i) +
                                               ": a batch needs every brick of its datasets resident (pool budget 0)");
    ViewState V = c->V;
    V.rot = r.rot;
    if (r.W != V.W || r.H != V.H) {
      if (!c->proj_from_fov)
        return bfail(b, TVK_ERR_ARG, "request " + std::to_string(i) +
                                         ": image size differs from a context whose projection was set explicitly");
      V.W = r.W;
      V.H = r.H;
      V.proj = perspective(V.fov, (float)r.W / (float)r.H, V.znear, V.zfar);   // GLRenderer.cpp:889-920
    }
    if (compat_active(c) && b->net && c->P.mode != TVK_RM_ISOSURFACE)
      return bfail(b, TVK_ERR_UNSUPPORTED, "request " + std::to_string(i) +
                                               ": the learned step reads the dataset's own bricks (compatibility "
                                               "textures set)");
    int lod = 0;
    int rc = frame_lod(c, V, &lod);
    if (!rc) rc = ensure_compat(c);
    if (!rc) {
      PlanParams pp;
      rc = plan_params(c, lod, pp);
      if (!rc) {
        plan_frame(c->L, c->minmax, c->range_max, V, pp, planned);
        make_frame_bricks(c->L, V, pp, planned, P[i].list, compat_dims(c), compat_offsets(c));
        if (b->net && c->P.mode != TVK_RM_ISOSURFACE) {
          rc = learn_prepare(c);
          if (!rc)
            for (FrameBrick& fb : P[i].list) brick_tf_stats(c, fb);
        }
      }
    }
    if (rc) return bfail(b, rc, "request " + std::to_string(i) + ": " + c->err);
    BatchRequest& br = P[i].br;
    br = BatchRequest{};
    fill_shade(c, br.S);
    make_view_const(V, br.V);
    br.V.nlist = (uint32_t)P[i].list.size();
    br.V.out_offset = (uint32_t)pixels;
    br.tiles_x = (r.W + tw - 1) / tw;
    br.ntiles = br.tiles_x * ((r.H + th - 1) / th);
    P[i].mode = c->P.mode;
    P[i].lit = c->P.lighting != 0 && c->P.mode != TVK_RM_ISOSURFACE;
    P[i].bpv = render_bpv(c);
    P[i].ntoc = (uint32_t)c->L.toc.size();
    if (P[i].mode == TVK_RM_1DTRANS && !c->d_tf1f)
      return bfail(b, TVK_ERR_STATE, "request " + std::to_string(i) + ": no 1D transfer function");
    if (P[i].mode == TVK_RM_2DTRANS && !c->d_tf2f)
      return bfail(b, TVK_ERR_STATE, "request " + std::to_string(i) + ": no 2D transfer function");
    b->planned_frames[i] = tvk_batch::Frame{pixels, r.W, r.H};
    pixels += (uint64_t)r.W * r.H;
  }
  if (pixels >= (1ull << 32)) return bfail(b, TVK_ERR_ARG, "too many pixels in one batch");
  // group by kernel variant (stable: requests keep their order inside a group)
  std::vector<size_t> order(P.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  auto key = [&](size_t i) { return P[i].mode * 4 + (P[i].lit ? 2 : 0) + (P[i].bpv == 2 ? 1 : 0); };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key(x) < key(y); });
  size_t nlist = 0, ntouched = 0;
  for (const Planned& p : P) { nlist += p.list.size(); ntouched += p.ntoc; }
  // the previous frame's uploads read the staging buffers: wait for them
  BCHK(b, hipStreamSynchronize(b->stream));
  int rc = bensure_pinned(b, &b->h_req, &b->h_req_cap, P.size());
  if (!rc) rc = bensure_pinned(b, &b->h_list, &b->h_list_cap, std::max<size_t>(nlist, 1));
  if (!rc) rc = bensure(b, &b->d_req, &b->req_cap, P.size());
  if (!rc) rc = bensure(b, &b->d_list, &b->list_cap, std::max<size_t>(nlist, 1));
  if (!rc) rc = bensure(b, &b->d_out, &b->out_cap, (size_t)pixels * 4);
  if (!rc && b->count) rc = bensure(b, &b->d_touched, &b->touched_cap, std::max<size_t>(ntouched, 1));
  if (rc) return rc;
  b->launches.clear();
  size_t lpos = 0, tpos = 0;
  for (size_t k = 0; k < order.size();) {
    const int kk = key(order[k]);
    tvk_batch::Launch L;
    std::memset(&L, 0, sizeof(L));   // the launch list is compared bytewise (graph key)
    L.mode = P[order[k]].mode;
    L.lit = P[order[k]].lit;
    L.bpv = P[order[k]].bpv;
    L.learn = (b->net && L.mode != TVK_RM_ISOSURFACE) ? 1 : 0;
    const size_t first = k;
    uint32_t tiles = 0;
    for (; k < order.size() && key(order[k]) == kk; ++k) {
      Planned& p = P[order[k]];
      p.br.V.list_offset = (uint32_t)lpos;
      p.br.tile_offset = tiles;
      p.br.touched_offset = (uint32_t)tpos;
      std::copy(p.list.begin(), p.list.end(), b->h_list + lpos);
      lpos += p.list.size();
      tpos += p.ntoc;
      tiles += p.br.ntiles;
      b->h_req[k] = p.br;
    }
    BatchArgs& a = L.args;
    if (L.learn) {
      a.learn = b->learn;
      if (!stepnet_device(b->net, &a.learn.net)) return bfail(b, TVK_ERR_STATE, "no step network");
    }
    a.req = b->d_req + first;
    a.list = b->d_list;
    a.out = b->d_out;
    a.queues = b->d_queues;
    a.counter = b->count ? b->d_counter : nullptr;
    a.touched = b->count ? b->d_touched : nullptr;
    a.nreq = (int)(k - first);
    a.tiles_total = tiles;
    a.tile_w = tw;
    a.tile_h = th;
    int per_cu = b->tune.groups_per_cu;
    if (per_cu <= 0) {
      hipError_t e = (hipError_t)batch_occupancy(L.mode, L.lit, L.bpv, b->count, L.learn, &per_cu);
      if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    }
    L.groups = ((b->cu_count * per_cu) + 7) & ~7;
    b->launches.push_back(L);
  }
  *bricks_total = (uint32_t)nlist;
  BCHK(b, hipMemcpyAsync(b->d_req, b->h_req, P.size() * sizeof(BatchRequest), hipMemcpyHostToDevice, b->stream));
  if (nlist)
    BCHK(b, hipMemcpyAsync(b->d_list, b->h_list, nlist * sizeof(FrameBrick), hipMemcpyHostToDevice, b->stream));
  if (b->count) {
    BCHK(b, hipMemsetAsync(b->d_counter, 0, sizeof(unsigned long long), b->stream));
    BCHK(b, hipMemsetAsync(b->d_touched, 0, std::max<size_t>(ntouched, 1) * sizeof(uint32_t), b->stream));
  }
  return TVK_OK;
}

int launch_all(tvk_batch* b) {
  for (const tvk_batch::Launch& L : b->launches)
    BCHK(b, launch_batch(L.args, L.mode, L.lit, L.bpv, L.groups, b->stream, L.learn));
  return TVK_OK;
}

// the launches of a frame as bytes: a captured graph is reused while they match
std::vector<uint8_t> launch_key(const tvk_batch* b) {
  std::vector<uint8_t> k;
  for (const tvk_batch::Launch& L : b->launches) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&L);
    k.insert(k.end(), p, p + sizeof(L));
  }
  return k;
}

}  // namespace

extern "C" {

int tvk_default_tuning(tvk_tuning* t) {
  if (!t) return TVK_ERR_ARG;
  std::memset(t, 0, sizeof(*t));
  t->tile_w = 8;
  t->tile_h = 8;
  t->groups_per_cu = 0;
  t->use_graph = 1;
  return TVK_OK;
}

int tvk_batch_create(tvk_batch** out, int device) {
  if (!out) return TVK_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TVK_ERR_HIP;
  if (device < 0 || device >= n) return TVK_ERR_ARG;
  tvk_batch* b = new (std::nothrow) tvk_batch;
  if (!b) return TVK_ERR_NOMEM;
  b->device = device;
  tvk_default_tuning(&b->tune);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_queues, 8 * sizeof(unsigned int));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_counter, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipEventCreate(&b->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&b->ev[1]);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    tvk_batch_destroy(b);
    return TVK_ERR_HIP;
  }
  *out = b;
  return TVK_OK;
}

void tvk_batch_destroy(tvk_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (void* p : {(void*)b->d_req, (void*)b->d_list, (void*)b->d_out, (void*)b->d_queues, (void*)b->d_counter,
                  (void*)b->d_touched})
    if (p) (void)hipFree(p);
  if (b->h_req) (void)hipHostFree(b->h_req);
  if (b->h_list) (void)hipHostFree(b->h_list);
  for (hipEvent_t e : b->ev) if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* tvk_batch_last_error(const tvk_batch* b) { return b ? b->err.c_str() : "null batch"; }

int tvk_batch_clear(tvk_batch* b) {
  if (!b) return TVK_ERR_ARG;
  b->reqs.clear();
  b->frames.clear();
  return TVK_OK;
}

int tvk_batch_add(tvk_batch* b, tvk_ctx* c, const float rot[16], uint32_t w, uint32_t h, int* id) {
  if (!b) return TVK_ERR_ARG;
  if (!c) return bfail(b, TVK_ERR_ARG, "null context");
  if (w == 0 || h == 0 || w > 32767 || h > 32767) return bfail(b, TVK_ERR_ARG, "bad image size");
  if (c->device != b->device) return bfail(b, TVK_ERR_ARG, "context and batch are on different devices");
  tvk_batch::Req r;
  r.ctx = c;
  r.rot = rot ? Mat4::from(rot) : c->V.rot;
  r.W = w;
  r.H = h;
  b->reqs.push_back(r);
  b->frames.clear();   // the last render's frames no longer match the request list
  if (id) *id = (int)b->reqs.size() - 1;
  return TVK_OK;
}

int tvk_batch_set_tuning(tvk_batch* b, const tvk_tuning* t) {
  if (!b || !t) return TVK_ERR_ARG;
  if (!valid_tile((uint32_t)t->tile_w, (uint32_t)t->tile_h))
    return bfail(b, TVK_ERR_ARG, "tile_w * tile_h must be 64 with power-of-two sides");
  if (t->groups_per_cu < 0 || t->groups_per_cu > 8) return bfail(b, TVK_ERR_ARG, "groups_per_cu must be 0..8");
  b->tune = *t;
  return TVK_OK;
}

int tvk_batch_get_tuning(tvk_batch* b, tvk_tuning* t) {
  if (!b || !t) return TVK_ERR_ARG;
  *t = b->tune;
  return TVK_OK;
}

int tvk_batch_set_step_proposal(tvk_batch* b, tvk_stepnet* net, const tvk_proposal_params* p) {
  if (!b) return TVK_ERR_ARG;
  if (net && !p) return bfail(b, TVK_ERR_ARG, "null proposal parameters");
  if (p && p->fixed_log2 >= 0) return bfail(b, TVK_ERR_UNSUPPORTED, "a batch takes the network (fixed_log2 < 0)");
  b->net = net;
  b->learn = LearnConst{};
  if (net) {
    if (p->max_log2 < 0 || p->max_log2 > 3) return bfail(b, TVK_ERR_ARG, "max_log2 must be 0..3");
    b->learn.margin = p->margin;
    const double q = p->skip_threshold;
    b->learn.skip_logit = q >= 1.0 ? INFINITY : (q <= 0.0 ? -INFINITY : (float)std::log(q / (1.0 - q)));
    b->learn.max_log2 = p->max_log2;
    b->learn.fallback = p->fallback_bound > 0.0f ? p->fallback_bound : INFINITY;
    b->learn.fallback_tv = p->fallback_tv > 0.0f ? p->fallback_tv : INFINITY;
    b->learn.skip_bound = p->skip_bound;
  }
  return TVK_OK;
}

int tvk_batch_set_count_samples(tvk_batch* b, int on) {
  if (!b) return TVK_ERR_ARG;
  b->count = on != 0;
  return TVK_OK;
}

int tvk_batch_render(tvk_batch* b, tvk_stats* st) {
  if (!b) return TVK_ERR_ARG;
  if (b->reqs.empty()) return bfail(b, TVK_ERR_STATE, "empty batch");
  BCHK(b, hipSetDevice(b->device));
  uint32_t bricks = 0;
  int rc = plan_batch(b, &bricks);
  if (rc) return rc;
  BCHK(b, hipEventRecord(b->ev[0], b->stream));
  if (b->tune.use_graph) {
    std::vector<uint8_t> key = launch_key(b);
    if (!b->exec || key != b->graph_key) {
      drop_graph(b);
      BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = launch_all(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return bfail(b, TVK_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      b->graph = g;
      BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_key = key;
    }
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
  } else {
    rc = launch_all(b);
    if (rc) return rc;
  }
  BCHK(b, hipEventRecord(b->ev[1], b->stream));
  unsigned long long ns = 0;
  if (b->count) BCHK(b, hipMemcpyAsync(&ns, b->d_counter, sizeof(ns), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float ms = 0.f;
  BCHK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->last_samples = ns;
  b->last_bricks = bricks;
  b->last_ms = ms;
  b->frames = b->planned_frames;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->samples = ns;
    st->bricks_nonempty = bricks;
    st->bricks_in_list = bricks;
    st->ms = ms;
  }
  return TVK_OK;
}

int tvk_batch_frame(tvk_batch* b, int id, float* host_out) {
  if (!b || !host_out) return TVK_ERR_ARG;
  if (id < 0 || (size_t)id >= b->frames.size()) return bfail(b, TVK_ERR_RANGE, "no such request in the last render");
  const tvk_batch::Frame& f = b->frames[(size_t)id];
  const size_t floats = (size_t)f.W * f.H * 4;
  if (!b->d_out || 4 * f.offset + floats > b->out_cap) return bfail(b, TVK_ERR_STATE, "frame outside the output buffer");
  BCHK(b, hipMemcpyAsync(host_out, b->d_out + 4 * f.offset, floats * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return TVK_OK;
}

int tvk_batch_autotune(tvk_batch* b, int reps, tvk_tuning* best, float* best_ms) {
  if (!b) return TVK_ERR_ARG;
  if (b->reqs.empty()) return bfail(b, TVK_ERR_STATE, "empty batch");
  if (reps < 1) reps = 1;
  const tvk_tuning keep = b->tune;
  static const int shapes[][2] = {{8, 8}, {16, 4}, {4, 16}, {32, 2}};
  static const int per_cu[] = {0, 1, 2, 4};
  tvk_tuning win = keep;
  float win_ms = 0.f;
  bool have = false;
  const bool count = b->count;
  b->count = false;
  int rc = TVK_OK;
  for (const auto& s : shapes) {
    for (int g : per_cu) {
      tvk_tuning t = keep;
      t.tile_w = s[0];
      t.tile_h = s[1];
      t.groups_per_cu = g;
      b->tune = t;
      rc = tvk_batch_render(b, nullptr);   // warm-up (and graph capture)
      if (rc) break;
      float tot = 0.f;
      for (int r = 0; r < reps && !rc; ++r) {
        rc = tvk_batch_render(b, nullptr);
        tot += b->last_ms;
      }
      if (rc) break;
      const float ms = tot / (float)reps;
      if (!have || ms < win_ms) { win = t; win_ms = ms; have = true; }
    }
    if (rc) break;
  }
  b->count = count;
  b->tune = rc ? keep : win;
  if (rc) return rc;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return TVK_OK;
}

int tvk_batch_frame_device_ptr(tvk_batch* b, int id, void** p) {
  if (!b || !p) return TVK_ERR_ARG;
  if (id < 0 || (size_t)id >= b->frames.size()) return bfail(b, TVK_ERR_RANGE, "no such request in the last render");
  *p = b->d_out + 4 * b->frames[(size_t)id].offset;
  return TVK_OK;
}

}  // extern "C"
