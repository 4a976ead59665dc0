/*
 * bvh_build.h -- scene BVH built on the GPU.
 *
 * The reference builds its BVH on the host, one thread, top-down
 * (clovers/src/bvh/build/surface_area_heuristic.rs:18-156); this library's
 * host builder does the same with binned SAH.  The device builder produces the
 * same 4-wide node and leaf-record layout in HBM directly:
 *
 *   1. 63-bit Morton codes of the primitives' box centres (21 bits per axis
 *      over the centres' bounding cube), sorted with a stable LSD radix sort
 *      (8-bit digits; per-tile histograms, one scan, a stable scatter ranked
 *      with wave ballots);
 *   2. agglomerative clustering over the sorted order (PLOC: Meister and
 *      Bittner 2018, "Parallel Locally-Ordered Clustering for Bounding Volume
 *      Hierarchy Construction"): every cluster finds, among the 16 on either
 *      side, the one whose union box has the least surface area; mutual
 *      nearest pairs merge into a new binary node; survivors are compacted in
 *      order, until one cluster is left.  Each new node gets its primitive
 *      count, record size and the decisions of the cost-optimal collapse (the
 *      host builder's: D(n, k), the least wide-tree cost of its subtree as at
 *      most k parts, from its two children's, which are final by then);
 *   3. collapse into 4-wide nodes one level at a time: every wide node takes
 *      the parts its binary node's decisions give (subtrees of up to 8
 *      primitives may be leaves); node indices are allocated by prefix sums
 *      per level, so the array is breadth-first (the layout the LDS-cached
 *      traversal relies on) and the result is deterministic.  Leaf records
 *      are copied into place, in scene order, in the same pass.
 *
 * Everything after the host's per-primitive record packing runs on the GPU;
 * the host reads back one counter per clustering iteration and per level.
 */
#ifndef CLV_BVH_BUILD_H
#define CLV_BVH_BUILD_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "clv_layout.h"

namespace clv {

struct DeviceBvhInput {
  const float4 *box;     /* per primitive: culling box lo.xyz, hi.xyz (2 float4) */
  const uint32_t *unit;  /* per primitive: offset of its record in src, float4 units */
  const float4 *src;     /* records in primitive order (headers as in the arena) */
  uint32_t n;            /* primitives, >= 1 */
  float clo[3], chi[3];  /* bounds of the box centres */
};

struct DeviceBvhResult {
  uint32_t n_nodes = 0;  /* 4-wide nodes written */
  uint32_t depth = 0;    /* deepest 4-wide level (root = 0) */
  uint32_t units = 0;    /* record units written */
  uint32_t iterations = 0;  /* clustering iterations */
};

/* Build into nodes[node_base ...] and dst[unit_base ...]: child words carry
 * the bases.  nodes must hold node_base + max(n - 1, 1) entries, dst
 * unit_base + (record units of all primitives).  Fails (CLV_EUNSUPPORTED)
 * when the tree is deeper than max_depth 4-wide levels.  Synchronises the
 * stream; returns a clv_status. */
int device_build_bvh(hipStream_t st, const DeviceBvhInput &in, Bvh4Node *nodes, uint32_t node_base, float4 *dst,
                     uint32_t unit_base, uint32_t max_depth, DeviceBvhResult *res, std::string *err);

}  // namespace clv
#endif
