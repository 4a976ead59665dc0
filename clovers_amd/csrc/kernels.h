/* kernels.h -- device-side views of the scene and path state, and the kernel
 * launchers (implemented in kernels.hip). */
#ifndef CLV_KERNELS_H
#define CLV_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clovers_hip.h"
#include "clv_layout.h"

namespace clv {

/* slots of the per-render counter array */
#define CNT_QA 0u       /* ray queue A size */
#define CNT_QB 1u       /* ray queue B size */
#define CNT_SHADE 2u    /* MC_COUNT shading queue sizes start here */
#define CNT_FETCH 7u    /* work counter of the persistent traversal */
#define CNT_TOTAL 8u

struct DevScene {
  const Bvh4Node *nodes;
  const float4 *prims;
  const PrimInfo *info;
  const Xform *xforms;
  const DevMaterial *mats;
  const float *spd;
  const DevLight *lights;
  const DevLightNode *lnodes;
  const float *lweights;
  uint32_t n_lights;
  uint32_t light_root;
  int32_t bg_spd;
  uint32_t has_xforms;
  DevCamera cam;
};

/* structure-of-arrays path state, one slot per in-flight sample */
struct PathBuf {
  float4 *ray_o;    /* origin */
  float4 *ray_d;    /* unit direction */
  float4 *thr;      /* throughput, 4 hero wavelengths */
  float4 *rad;      /* accumulated radiance, 4 hero wavelengths */
  uint64_t *rng;    /* PCG32 state */
  uint32_t *meta;   /* hero wavelength index (9 bits) | depth << 9 */
  float *hit_t;
  uint32_t *hit_p;  /* primitive record offset of the closest hit, ~0u on a miss */
};

struct ShadeQueues {
  uint32_t *q[MC_COUNT];
  uint32_t *counts; /* = counter array + CNT_SHADE */
};

/* arguments of the persistent traversal kernel */
struct TraceArgs {
  const uint32_t *qin;         /* extend mode: ray queue */
  uint32_t *counts;            /* counter array (queue sizes, fetch counter) */
  uint32_t in_slot;            /* extend mode: slot of the ray-queue size */
  uint32_t n_direct;           /* primary mode: number of rays (paths 0..n-1) */
  ShadeQueues sq;              /* extend mode: shading queues */
  uint32_t *spill;             /* per-thread overflow of the traversal stack */
  unsigned long long *stats;   /* counting mode: node visits, primitive tests */
  uint32_t *ids;               /* primary mode outputs */
  float *ts;
  uint32_t lds_nodes;          /* LDS variant: nodes [0, lds_nodes) are cached in LDS */
  uint32_t lds_units;          /* LDS variant: primitive float4 units [0, lds_units) are cached */
};

struct GenParams {
  uint32_t width, height, pixels, paths;  /* paths = pixels * samples in this pass */
  uint32_t sample_base;
  uint32_t pad;
  uint64_t seed;
};

struct ResolveParams {
  uint32_t pixels;
  uint32_t per_pixel;  /* samples of this pass */
  uint32_t fp64;
  uint32_t pad;
  float *acc32;        /* 3 per pixel */
  double *acc64;       /* 3 per pixel */
  double *acc_sq;      /* 3 per pixel */
};

/* stats: 4 x u64 = node visits, primitive tests, rays traced, spare */
void launch_reset_counts(hipStream_t st, uint32_t *counts, uint32_t in_slot, uint32_t out_slot, uint64_t *stats);
void launch_set_count(hipStream_t st, uint32_t *counts, uint32_t slot, uint32_t value);
void launch_generate(hipStream_t st, const DevScene &sc, const PathBuf &pb, const GenParams &gp, uint32_t *qout);
/* persistent traversal: grid = resident blocks of the device; spill buffer of
 * trace_spill_words(blocks) words */
uint32_t trace_grid_blocks();
size_t trace_spill_words(uint32_t blocks);
/* traversal kernel variants: one ray per thread over a static grid-stride
 * loop, or persistent waves refilling lanes from an atomic work counter */
#define TRACE_SIMPLE 0
#define TRACE_PERSISTENT 1
#define TRACE_WW 2
#define TRACE_LDS 3
/* LDS-cached traversal: dynamic LDS bytes for a cached prefix of `nodes` wide
 * nodes and `units` primitive float4, resident blocks for that size, and the
 * spill words its grid needs; mode 0 = extend, 1 = primary */
size_t lds_trace_bytes(uint32_t nodes, uint32_t units);
uint32_t lds_trace_stack_bytes();
size_t lds_trace_spill_words(uint32_t blocks);
uint32_t lds_trace_grid_blocks(size_t lds_bytes);
void launch_trace_lds(hipStream_t st, const DevScene &sc, const PathBuf &pb, const TraceArgs &ta, uint32_t blocks,
                      size_t lds_bytes, bool counting, int mode);
void launch_extend(hipStream_t st, const DevScene &sc, const PathBuf &pb, const TraceArgs &ta, uint32_t blocks,
                   bool counting, int variant);
void launch_primary(hipStream_t st, const DevScene &sc, const PathBuf &pb, const TraceArgs &ta, uint32_t blocks,
                    int variant);
void launch_shade(hipStream_t st, const DevScene &sc, const PathBuf &pb, const ShadeQueues &sq, uint32_t *qout,
                  uint32_t *counts, uint32_t out_slot, uint32_t max_depth, uint32_t max_paths, uint32_t grid_cap);
void launch_resolve(hipStream_t st, const PathBuf &pb, const ResolveParams &rp, const float4 *cie);
void launch_finalize(hipStream_t st, const ResolveParams &rp, float *out, float *out_sq, uint32_t spp);
void launch_zero_accum(hipStream_t st, const ResolveParams &rp);

}  // namespace clv
#endif
