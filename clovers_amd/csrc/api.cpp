/*
 * api.cpp -- the C ABI (include/clovers_hip.h) over the HIP renderer.
 *
 * clv_render is the drop-in for draw_cpu::draw (clovers-cli/src/draw_cpu.rs:26-98,
 * called at clovers-cli/src/render.rs:149).  A render is a sequence of
 * wavefront passes; pass p covers samples [p*k, p*k+k) of every pixel, with k
 * chosen so that a pass keeps ~4M paths in flight (enough waves for 256 CUs,
 * ~350 MB of path state).  Per pass:
 *
 *   generate                       camera rays of all paths of the pass
 *   repeat max_depth+1 times:      reset counters; extend; shade
 *   resolve                        per-sample estimator, per-pixel sums
 *
 * All launches go to the context's stream; the host never waits inside a
 * render (queue sizes stay on the device, kernels read them), so the launch
 * sequence of a pass is fixed and the driver pipelines it.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/clovers_hip.h"
#include "host_build.h"
#include "kernels.h"

using namespace clv;

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct SceneRes {
  DevScene base;
  std::vector<DevBuf> bufs;
  clv_camera_init camera;
  uint32_t n_nodes = 0, n_prims = 0, prim_bytes = 0, depth = 0;
  uint32_t prim_units = 0;   /* float4 units of primitive records */
  double build_ms = 0.0;
};

}  // namespace

struct clv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<uint32_t, SceneRes *> scenes;
  uint32_t next_scene = 1;
  /* path state */
  uint32_t cap_paths = 0;
  PathBuf pb;
  uint32_t *queue[2] = {nullptr, nullptr};
  uint32_t *shadeq[MC_COUNT] = {nullptr};
  /* accumulators */
  uint32_t cap_pixels = 0;
  float *acc32 = nullptr;
  double *acc64 = nullptr, *accsq = nullptr;
  float *out = nullptr, *out_sq = nullptr;
  uint32_t *ids = nullptr;
  float *ts = nullptr;
  /* misc */
  uint32_t *counts = nullptr;
  uint64_t *stats = nullptr;
  uint32_t *spill = nullptr;      /* traversal stack overflow, one slice per launched thread */
  size_t spill_words = 0;         /* size of the spill buffer */
  size_t lds_block_bytes = 64512; /* dynamic LDS budget of one block of the LDS-cached traversal (64 KB less static use) */
  uint32_t trace_blocks = 0;      /* resident blocks of the persistent traversal */
  int trace_variant = TRACE_LDS;
  bool full_grid = false;         /* one thread per possible ray instead of the resident grid */
  float4 *cie = nullptr;
  std::vector<hipEvent_t> events;
  clv_counters counters;
  uint32_t target_paths = 1u << 22;
  uint32_t grid_cap = 2048;
};

namespace {

clv_status fail(clv_ctx *c, clv_status s, const std::string &msg) {
  if (c) c->err = msg;
  return s;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      return fail((ctx), e_ == hipErrorOutOfMemory ? CLV_ENOMEM : CLV_EDEVICE,              \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    }                                                                                       \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, n * sizeof(T) + 16);
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

void free_paths(clv_ctx *c) {
  dfree(c->pb.ray_o); dfree(c->pb.ray_d); dfree(c->pb.thr); dfree(c->pb.rad);
  dfree(c->pb.rng); dfree(c->pb.meta); dfree(c->pb.hit_t); dfree(c->pb.hit_p);
  for (int i = 0; i < 2; i++) dfree(c->queue[i]);
  for (uint32_t k = 0; k < MC_COUNT; k++) dfree(c->shadeq[k]);
  memset(&c->pb, 0, sizeof(c->pb));
  c->queue[0] = c->queue[1] = nullptr;
  for (uint32_t k = 0; k < MC_COUNT; k++) c->shadeq[k] = nullptr;
  c->cap_paths = 0;
}

clv_status ensure_spill(clv_ctx *c, size_t words) {
  if (words <= c->spill_words) return CLV_OK;
  dfree(c->spill);
  c->spill = nullptr;
  c->spill_words = 0;
  HIPCHK(c, dalloc(&c->spill, words));
  c->spill_words = words;
  return CLV_OK;
}

/* traversal launch geometry of a scene for the configured variant */
struct TraceGeom {
  uint32_t blocks = 0;
  size_t lds_bytes = 0;
  uint32_t lds_nodes = 0, lds_units = 0;
};

clv_status trace_geometry(clv_ctx *c, const SceneRes *s, uint32_t full_grid_paths, TraceGeom *g) {
  if (c->trace_variant == TRACE_LDS) {
    size_t budget = c->lds_block_bytes > lds_trace_stack_bytes() ? c->lds_block_bytes - lds_trace_stack_bytes() : 0;
    g->lds_nodes = (uint32_t)std::min<size_t>(s->n_nodes, budget / sizeof(Bvh4Node));
    budget -= (size_t)g->lds_nodes * sizeof(Bvh4Node);
    g->lds_units = (uint32_t)std::min<size_t>(s->prim_units, budget / 16u);
    g->lds_bytes = lds_trace_bytes(g->lds_nodes, g->lds_units);
    g->blocks = lds_trace_grid_blocks(g->lds_bytes);
    return ensure_spill(c, lds_trace_spill_words(g->blocks));
  }
  g->blocks = c->trace_blocks;
  if (full_grid_paths && c->full_grid && c->trace_variant != TRACE_PERSISTENT) g->blocks = (full_grid_paths + 255u) / 256u;
  return ensure_spill(c, trace_spill_words(g->blocks));
}

clv_status ensure_paths(clv_ctx *c, uint32_t n) {
  if (n <= c->cap_paths) return CLV_OK;
  free_paths(c);
  HIPCHK(c, dalloc(&c->pb.ray_o, n));
  HIPCHK(c, dalloc(&c->pb.ray_d, n));
  HIPCHK(c, dalloc(&c->pb.thr, n));
  HIPCHK(c, dalloc(&c->pb.rad, n));
  HIPCHK(c, dalloc(&c->pb.rng, n));
  HIPCHK(c, dalloc(&c->pb.meta, n));
  HIPCHK(c, dalloc(&c->pb.hit_t, n));
  HIPCHK(c, dalloc(&c->pb.hit_p, n));
  for (int i = 0; i < 2; i++) HIPCHK(c, dalloc(&c->queue[i], n));
  for (uint32_t k = 0; k < MC_COUNT; k++) HIPCHK(c, dalloc(&c->shadeq[k], n));
  c->cap_paths = n;
  return CLV_OK;
}

void free_pixels(clv_ctx *c) {
  dfree(c->acc32); dfree(c->acc64); dfree(c->accsq); dfree(c->out); dfree(c->out_sq); dfree(c->ids); dfree(c->ts);
  c->acc32 = nullptr; c->acc64 = c->accsq = nullptr; c->out = c->out_sq = nullptr; c->ids = nullptr; c->ts = nullptr;
  c->cap_pixels = 0;
}

clv_status ensure_pixels(clv_ctx *c, uint32_t n) {
  if (n <= c->cap_pixels) return CLV_OK;
  free_pixels(c);
  HIPCHK(c, dalloc(&c->acc32, 3 * (size_t)n));
  HIPCHK(c, dalloc(&c->acc64, 3 * (size_t)n));
  HIPCHK(c, dalloc(&c->accsq, 3 * (size_t)n));
  HIPCHK(c, dalloc(&c->out, 3 * (size_t)n));
  HIPCHK(c, dalloc(&c->out_sq, 3 * (size_t)n));
  HIPCHK(c, dalloc(&c->ids, (size_t)n));
  HIPCHK(c, dalloc(&c->ts, (size_t)n));
  c->cap_pixels = n;
  return CLV_OK;
}

void free_scene(SceneRes *s) {
  for (auto &b : s->bufs) dfree(b.p);
  delete s;
}

clv_status upload(clv_ctx *c, SceneRes *s, const void *src, size_t bytes, const void **dst) {
  DevBuf b;
  b.bytes = bytes ? bytes : 16;
  HIPCHK(c, hipMalloc(&b.p, b.bytes));
  s->bufs.push_back(b);
  if (bytes) HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  *dst = b.p;
  return CLV_OK;
}

clv_status check_opts(clv_ctx *c, const clv_render_opts *o) {
  if (!o) return fail(c, CLV_EINVAL, "opts is NULL");
  if (o->width == 0 || o->height == 0) return fail(c, CLV_EINVAL, "width and height must be positive");
  if ((uint64_t)o->width * o->height > (1ull << 27)) return fail(c, CLV_EINVAL, "image larger than 2^27 pixels");
  if (o->max_depth > 1000) return fail(c, CLV_EINVAL, "max_depth above 1000");
  if (o->guiding != CLV_GUIDE_OFF && o->guiding != CLV_GUIDE_ON) return fail(c, CLV_EINVAL, "unknown guiding mode");
  return CLV_OK;
}

DevScene scene_for(const SceneRes *s, const clv_render_opts *o) {
  DevScene d = s->base;
  d.cam = make_camera(s->camera, (float)o->width / (float)o->height);
  return d;
}

}  // namespace

extern "C" {

uint32_t clv_abi_version(void) { return CLV_ABI_VERSION; }

float clv_spectral_power(const float xyz[3], uint32_t lambda) {
  if (!xyz) return 0.0f;
  return clv::spectral_power(xyz, (int)lambda);
}

clv_status clv_wavelength_into_xyz(uint32_t lambda, float xyz_out[3]) {
  if (!xyz_out || lambda < CLV_LAMBDA_MIN || lambda >= CLV_LAMBDA_MIN + CLV_SPECTRUM) return CLV_EINVAL;
  float table[CLV_SPECTRUM * 4];
  cie_xyz_table(table);
  for (int k = 0; k < 3; k++) xyz_out[k] = table[4 * (lambda - CLV_LAMBDA_MIN) + k];
  return CLV_OK;
}

clv_status clv_create(int hip_device, clv_ctx **out) {
  if (!out) return CLV_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CLV_EDEVICE;
  if (hip_device < 0 || hip_device >= n) return CLV_EINVAL;
  clv_ctx *c = new (std::nothrow) clv_ctx();
  if (!c) return CLV_ENOMEM;
  memset(&c->pb, 0, sizeof(c->pb));
  memset(&c->counters, 0, sizeof(c->counters));
  c->device = hip_device;
  if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return CLV_EDEVICE;
  }
  c->trace_blocks = trace_grid_blocks();
  if (const char *e = getenv("CLV_TRACE_BLOCKS")) c->trace_blocks = (uint32_t)strtoul(e, nullptr, 10);
  if (const char *e = getenv("CLV_TRACE_KERNEL"))
    c->trace_variant = strcmp(e, "persistent") == 0 ? TRACE_PERSISTENT
                       : strcmp(e, "ww") == 0       ? TRACE_WW
                       : strcmp(e, "simple") == 0   ? TRACE_SIMPLE
                                                    : TRACE_LDS;
  if (const char *e = getenv("CLV_TRACE_GRID")) c->full_grid = strcmp(e, "full") == 0;
  if (c->trace_blocks < 1) c->trace_blocks = 1;
  if (const char *e = getenv("CLV_LDS_BYTES")) c->lds_block_bytes = (size_t)strtoul(e, nullptr, 10);
  if (dalloc(&c->counts, CNT_TOTAL) != hipSuccess || dalloc(&c->stats, 4) != hipSuccess ||
      dalloc(&c->cie, CLV_SPECTRUM) != hipSuccess) {
    clv_destroy(c);
    return CLV_ENOMEM;
  }
  float table[CLV_SPECTRUM * 4];
  cie_xyz_table(table);
  if (hipMemcpy(c->cie, table, sizeof(table), hipMemcpyHostToDevice) != hipSuccess) {
    clv_destroy(c);
    return CLV_EDEVICE;
  }
  if (const char *e = getenv("CLV_TARGET_PATHS")) c->target_paths = (uint32_t)strtoul(e, nullptr, 10);
  if (const char *e = getenv("CLV_GRID_CAP")) c->grid_cap = (uint32_t)strtoul(e, nullptr, 10);
  if (c->target_paths < 65536) c->target_paths = 65536;
  if (c->grid_cap < 1) c->grid_cap = 1;
  *out = c;
  return CLV_OK;
}

void clv_destroy(clv_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->scenes) free_scene(kv.second);
  free_paths(c);
  free_pixels(c);
  dfree(c->counts);
  dfree(c->stats);
  dfree(c->spill);
  dfree(c->cie);
  for (auto e : c->events) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *clv_last_error(const clv_ctx *c) { return c ? c->err.c_str() : "no context"; }

clv_status clv_scene_upload(clv_ctx *c, const clv_scene_desc *desc, uint32_t *scene_id) {
  if (!c || !scene_id) return CLV_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HostScene hs;
  std::string err;
  int st = build_host_scene(desc, &hs, &err);
  if (st != CLV_OK) return fail(c, (clv_status)st, err);
  SceneRes *s = new (std::nothrow) SceneRes();
  if (!s) return fail(c, CLV_ENOMEM, "host allocation failed");
  memset(&s->base, 0, sizeof(s->base));
  clv_status r = CLV_OK;
  const void *p;
#define UP(vec, field, T)                                                                      \
  if (r == CLV_OK) {                                                                           \
    r = upload(c, s, hs.vec.data(), hs.vec.size() * sizeof(hs.vec[0]), &p);                   \
    s->base.field = (const T *)p;                                                              \
  }
  UP(nodes, nodes, Bvh4Node);
  UP(prims, prims, float4);
  UP(prim_info, info, PrimInfo);
  UP(xforms, xforms, Xform);
  UP(materials, mats, DevMaterial);
  UP(spd, spd, float);
  UP(lights, lights, DevLight);
  UP(light_nodes, lnodes, DevLightNode);
  UP(light_weights, lweights, float);
#undef UP
  if (r == CLV_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) r = fail(c, CLV_EDEVICE, hipGetErrorString(e));
  }
  if (r != CLV_OK) {
    free_scene(s);
    return r;
  }
  s->base.n_lights = (uint32_t)hs.lights.size();
  s->base.light_root = hs.light_root;
  s->base.bg_spd = hs.background_spd;
  s->base.has_xforms = hs.xforms.empty() ? 0u : 1u;
  s->camera = hs.camera;
  s->n_nodes = (uint32_t)hs.nodes.size();
  s->prim_units = (uint32_t)(hs.prims.size() / 4);
  s->n_prims = hs.n_prims;
  s->prim_bytes = hs.n_prims ? (uint32_t)(hs.prims.size() * 4 / hs.n_prims) : 0;
  s->depth = hs.bvh_depth;
  s->build_ms = hs.build_ms;
  c->counters.build_ms = hs.build_ms;
  uint32_t id = c->next_scene++;
  c->scenes[id] = s;
  *scene_id = id;
  return CLV_OK;
}

clv_status clv_scene_validate(const clv_scene_desc *desc, clv_scene_info *info) {
  if (!info) return CLV_EINVAL;
  memset(info, 0, sizeof(*info));
  HostScene hs;
  std::string err;
  int st = build_host_scene(desc, &hs, &err);
  if (st != CLV_OK) {
    snprintf(info->error, sizeof(info->error), "%s", err.c_str());
    return (clv_status)st;
  }
  info->n_objects = hs.n_objects;
  info->n_prims = hs.n_prims;
  info->n_lights = (uint32_t)hs.lights.size();
  info->bvh_nodes = (uint32_t)hs.nodes.size();
  info->bvh_depth = hs.bvh_depth;
  info->light_nodes = (uint32_t)hs.light_nodes.size();
  info->device_bytes = hs.nodes.size() * sizeof(Bvh4Node) + hs.prims.size() * sizeof(float) +
                       hs.prim_info.size() * sizeof(PrimInfo) + hs.xforms.size() * sizeof(Xform) +
                       hs.materials.size() * sizeof(DevMaterial) + hs.spd.size() * sizeof(float) +
                       hs.lights.size() * sizeof(DevLight) + hs.light_nodes.size() * sizeof(DevLightNode) +
                       hs.light_weights.size() * sizeof(float);
  info->build_ms = hs.build_ms;
  return CLV_OK;
}

clv_status clv_scene_free(clv_ctx *c, uint32_t scene_id) {
  if (!c) return CLV_EINVAL;
  auto it = c->scenes.find(scene_id);
  if (it == c->scenes.end()) return fail(c, CLV_EINVAL, "unknown scene id");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_scene(it->second);
  c->scenes.erase(it);
  return CLV_OK;
}

clv_status clv_render(clv_ctx *c, uint32_t scene_id, const clv_render_opts *o, float *xyz_out, float *xyz_sq_out) {
  if (!c) return CLV_EINVAL;
  clv_status st = check_opts(c, o);
  if (st != CLV_OK) return st;
  if (!xyz_out) return fail(c, CLV_EINVAL, "xyz_out is NULL");
  if (o->samples == 0) return fail(c, CLV_EINVAL, "samples must be positive");
  if (o->guiding == CLV_GUIDE_ON) return fail(c, CLV_EUNSUPPORTED, "guiding is not available in this build");
  auto it = c->scenes.find(scene_id);
  if (it == c->scenes.end()) return fail(c, CLV_EINVAL, "unknown scene id");
  const SceneRes *s = it->second;
  HIPCHK(c, hipSetDevice(c->device));
  const uint32_t W = o->width, H = o->height, WH = W * H;
  uint32_t k = c->target_paths / WH;
  if (k < 1) k = 1;
  if (k > o->samples) k = o->samples;
  if ((st = ensure_paths(c, WH * k)) != CLV_OK) return st;
  if ((st = ensure_pixels(c, WH)) != CLV_OK) return st;
  TraceGeom geom;
  if ((st = trace_geometry(c, s, WH * k, &geom)) != CLV_OK) return st;
  const DevScene ds = scene_for(s, o);
  const bool timing = (o->flags & CLV_FLAG_TIME_KERNELS) != 0;
  const bool counting = (o->flags & CLV_FLAG_COUNT) != 0;
  const uint32_t passes = (o->samples + k - 1) / k;
  const uint32_t extends = passes * (o->max_depth + 1);
  size_t need_events = 2 + (timing ? 2 * (size_t)extends : 0);
  while (c->events.size() < need_events) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->events.push_back(e);
  }
  ResolveParams rp;
  memset(&rp, 0, sizeof(rp));
  rp.pixels = WH;
  rp.fp64 = o->fp64_accum ? 1u : 0u;
  rp.acc32 = c->acc32;
  rp.acc64 = c->acc64;
  rp.acc_sq = c->accsq;
  ShadeQueues sq;
  for (uint32_t q = 0; q < MC_COUNT; q++) sq.q[q] = c->shadeq[q];
  sq.counts = c->counts + CNT_SHADE;
  HIPCHK(c, hipMemsetAsync(c->stats, 0, 4 * sizeof(uint64_t), c->stream));
  launch_zero_accum(c->stream, rp);
  HIPCHK(c, hipEventRecord(c->events[0], c->stream));
  size_t ev = 2;
  uint32_t done = 0;
  for (uint32_t p = 0; p < passes; p++) {
    uint32_t kk = o->samples - done < k ? o->samples - done : k;
    GenParams gp;
    memset(&gp, 0, sizeof(gp));
    gp.width = W;
    gp.height = H;
    gp.pixels = WH;
    gp.paths = WH * kk;
    gp.sample_base = o->sample_offset + done;
    gp.seed = o->seed;
    launch_set_count(c->stream, c->counts, CNT_QA, gp.paths);
    launch_generate(c->stream, ds, c->pb, gp, c->queue[0]);
    for (uint32_t depth = 0; depth <= o->max_depth; depth++) {
      uint32_t in = depth & 1u, out = in ^ 1u;
      uint32_t in_slot = in ? CNT_QB : CNT_QA, out_slot = out ? CNT_QB : CNT_QA;
      launch_reset_counts(c->stream, c->counts, in_slot, out_slot, c->stats);
      if (timing) HIPCHK(c, hipEventRecord(c->events[ev], c->stream));
      TraceArgs ta;
      memset(&ta, 0, sizeof(ta));
      ta.qin = c->queue[in];
      ta.counts = c->counts;
      ta.in_slot = in_slot;
      ta.sq = sq;
      ta.spill = c->spill;
      ta.stats = (unsigned long long *)c->stats;
      ta.lds_nodes = geom.lds_nodes;
      ta.lds_units = geom.lds_units;
      if (c->trace_variant == TRACE_LDS)
        launch_trace_lds(c->stream, ds, c->pb, ta, geom.blocks, geom.lds_bytes, counting, 0);
      else
        launch_extend(c->stream, ds, c->pb, ta, geom.blocks, counting, c->trace_variant);
      if (timing) HIPCHK(c, hipEventRecord(c->events[ev + 1], c->stream));
      ev += timing ? 2 : 0;
      launch_shade(c->stream, ds, c->pb, sq, c->queue[out], c->counts, out_slot, o->max_depth, gp.paths, c->grid_cap);
    }
    rp.per_pixel = kk;
    launch_resolve(c->stream, c->pb, rp, c->cie);
    done += kk;
  }
  launch_finalize(c->stream, rp, c->out, xyz_sq_out ? c->out_sq : nullptr, o->samples);
  HIPCHK(c, hipEventRecord(c->events[1], c->stream));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(xyz_out, c->out, 3 * sizeof(float) * WH, hipMemcpyDeviceToHost, c->stream));
  if (xyz_sq_out) HIPCHK(c, hipMemcpyAsync(xyz_sq_out, c->out_sq, 3 * sizeof(float) * WH, hipMemcpyDeviceToHost, c->stream));
  uint64_t stats[4];
  HIPCHK(c, hipMemcpyAsync(stats, c->stats, sizeof(stats), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.0f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->events[0], c->events[1]));
  clv_counters &cn = c->counters;
  cn.samples = (uint64_t)WH * o->samples;
  cn.rays = stats[2];
  cn.node_visits = counting ? stats[0] : 0;
  cn.prim_tests = counting ? stats[1] : 0;
  cn.render_ms = ms;
  cn.extend_ms = 0.0;
  if (timing) {
    for (size_t e = 2; e < ev; e += 2) {
      float m = 0.0f;
      HIPCHK(c, hipEventElapsedTime(&m, c->events[e], c->events[e + 1]));
      cn.extend_ms += m;
    }
  }
  cn.extend_launches = extends;
  cn.passes = passes;
  cn.bvh_nodes = s->n_nodes;
  cn.bvh_width = 4;
  cn.node_bytes = sizeof(Bvh4Node);
  cn.prim_bytes = s->prim_bytes;
  cn.n_prims = s->n_prims;
  cn.bvh_depth = s->depth;
  cn.build_ms = s->build_ms;
  return CLV_OK;
}

clv_status clv_primary_hits(clv_ctx *c, uint32_t scene_id, const clv_render_opts *o, uint32_t *ids, float *ts) {
  if (!c) return CLV_EINVAL;
  clv_status st = check_opts(c, o);
  if (st != CLV_OK) return st;
  if (!ids) return fail(c, CLV_EINVAL, "prim_id_out is NULL");
  auto it = c->scenes.find(scene_id);
  if (it == c->scenes.end()) return fail(c, CLV_EINVAL, "unknown scene id");
  HIPCHK(c, hipSetDevice(c->device));
  const uint32_t W = o->width, H = o->height, WH = W * H;
  if ((st = ensure_paths(c, WH)) != CLV_OK) return st;
  if ((st = ensure_pixels(c, WH)) != CLV_OK) return st;
  const DevScene ds = scene_for(it->second, o);
  GenParams gp;
  memset(&gp, 0, sizeof(gp));
  gp.width = W;
  gp.height = H;
  gp.pixels = WH;
  gp.paths = WH;
  gp.sample_base = o->sample_offset;
  gp.seed = o->seed;
  launch_generate(c->stream, ds, c->pb, gp, c->queue[0]);
  launch_set_count(c->stream, c->counts, CNT_FETCH, 0u);
  TraceArgs ta;
  memset(&ta, 0, sizeof(ta));
  ta.counts = c->counts;
  ta.n_direct = WH;
  ta.spill = c->spill;
  ta.ids = c->ids;
  ta.ts = c->ts;
  TraceGeom geom;
  if ((st = trace_geometry(c, it->second, 0, &geom)) != CLV_OK) return st;
  ta.lds_nodes = geom.lds_nodes;
  ta.lds_units = geom.lds_units;
  if (c->trace_variant == TRACE_LDS)
    launch_trace_lds(c->stream, ds, c->pb, ta, geom.blocks, geom.lds_bytes, false, 1);
  else
    launch_primary(c->stream, ds, c->pb, ta, geom.blocks, c->trace_variant);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(ids, c->ids, sizeof(uint32_t) * WH, hipMemcpyDeviceToHost, c->stream));
  if (ts) HIPCHK(c, hipMemcpyAsync(ts, c->ts, sizeof(float) * WH, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLV_OK;
}

clv_status clv_counters_get(const clv_ctx *c, clv_counters *out) {
  if (!c || !out) return CLV_EINVAL;
  *out = c->counters;
  return CLV_OK;
}

}  // extern "C"
