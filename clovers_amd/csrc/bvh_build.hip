/*
 * bvh_build.hip -- the scene BVH built on the GPU: Morton sort, PLOC
 * clustering, breadth-first collapse into 4-wide nodes.  The algorithm and
 * its rules are described in bvh_build.h.
 */
#include <float.h>

#include <algorithm>
#include <utility>

#include "../../include/clovers_hip.h"
#include "bvh_build.h"

namespace clv {
namespace {

#define BB_LEAF 0x80000000u   /* child id of a primitive (its index); else a binary node */
#define BB_NONE 0xFFFFFFFFu
#define PL_R 16               /* clustering search radius */
#define RS_THREADS 256
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define SC_THREADS 256
#define SC_ITEMS 4
#define SC_TILE (SC_THREADS * SC_ITEMS)

__device__ __forceinline__ float bb_area(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}
__device__ __forceinline__ float4 bb_min(float4 a, float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ __forceinline__ float4 bb_max(float4 a, float4 b) {
  return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

/* ------------------------------------------------------------ Morton --- */
/* 21 bits of x spread to every third bit of a 63-bit word */
__device__ __forceinline__ uint64_t expand21(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ uint32_t quant21(float c, float lo, float s) {
  const float q = fminf(fmaxf((c - lo) * s, 0.0f), 2097151.0f);
  return (uint32_t)q;
}

/* 63-bit Morton code of the box centre, as low and high words */
__global__ __launch_bounds__(256) void k_bb_morton(uint32_t n, const float4 *__restrict__ box, float lx, float ly,
                                                   float lz, float sx, float sy, float sz, uint32_t *__restrict__ klo,
                                                   uint32_t *__restrict__ khi, uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 lo = box[2u * i], hi = box[2u * i + 1u];
  const uint32_t qx = quant21(0.5f * (lo.x + hi.x), lx, sx), qy = quant21(0.5f * (lo.y + hi.y), ly, sy),
                 qz = quant21(0.5f * (lo.z + hi.z), lz, sz);
  const uint64_t key = (expand21(qx) << 2) | (expand21(qy) << 1) | expand21(qz);
  klo[i] = (uint32_t)key;
  khi[i] = (uint32_t)(key >> 32);
  vals[i] = i;
}

/* ------------------------------------------------------ exclusive scan --- */
/* three kernels: per-tile scan (4 consecutive elements per thread) with tile
 * sums, one block scanning the tile sums (and writing the total), add-back */
template <typename T>
__global__ __launch_bounds__(SC_THREADS) void k_scan_local(T *a, uint32_t m, T *bsum) {
  __shared__ T s[SC_THREADS];
  const uint32_t tid = threadIdx.x, base = blockIdx.x * SC_TILE + tid * SC_ITEMS;
  T v[SC_ITEMS], sum = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; k++) {
    v[k] = base + k < m ? a[base + k] : (T)0;
    sum += v[k];
  }
  s[tid] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < SC_THREADS; off <<= 1) {
    const T t = tid >= off ? s[tid - off] : (T)0;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  T run = s[tid] - sum;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; k++) {
    if (base + k < m) a[base + k] = run;
    run += v[k];
  }
  if (tid == SC_THREADS - 1) bsum[blockIdx.x] = s[tid];
}

template <typename T>
__global__ __launch_bounds__(1024) void k_scan_blocks(T *bsum, uint32_t nb, T *total) {
  __shared__ T s[1024];
  const uint32_t tid = threadIdx.x;
  T carry = 0;
  for (uint32_t start = 0; start < nb; start += 1024u) {
    const uint32_t idx = start + tid;
    const T v = idx < nb ? bsum[idx] : (T)0;
    s[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
      const T t = tid >= off ? s[tid - off] : (T)0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + s[tid] - v;
    const T chunk = s[1023];
    __syncthreads();
    carry += chunk;
  }
  if (tid == 0) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(SC_THREADS) void k_scan_add(T *a, uint32_t m, const T *bsum) {
  const uint32_t base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
  const T add = bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SC_ITEMS; k++)
    if (base + k < m) a[base + k] += add;
}

/* ---------------------------------------------------------- radix sort --- */
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(uint32_t n, const uint32_t *__restrict__ keys, uint32_t shift,
                                                        uint32_t nb, uint32_t *__restrict__ hist) {
  __shared__ uint32_t cnt[256];
  const uint32_t tid = threadIdx.x;
  cnt[tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RS_TILE;
  for (uint32_t k = 0; k < RS_ITEMS; k++) {
    const uint32_t i = base + k * RS_THREADS + tid;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[tid * nb + blockIdx.x] = cnt[tid]; /* digit-major: one scan gives every (digit, tile) offset */
}

/* stable scatter of (key low word, key high word, value): elements of a tile
 * are ranked in index order, a chunk of 256 at a time; within a wave, lanes
 * with the same digit find each other with 8 ballots */
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(uint32_t n, const uint32_t *__restrict__ lin,
                                                           const uint32_t *__restrict__ hin, const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ lout, uint32_t *__restrict__ hout,
                                                           uint32_t *__restrict__ vout, uint32_t high, uint32_t shift,
                                                           uint32_t nb, const uint32_t *__restrict__ hist) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wc[RS_THREADS / 64][256];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  base[tid] = hist[tid * nb + blockIdx.x];
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t k = 0; k < RS_ITEMS; k++) {
    for (uint32_t q = tid; q < (RS_THREADS / 64) * 256u; q += RS_THREADS) (&wc[0][0])[q] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * RS_TILE + k * RS_THREADS + tid;
    const bool valid = i < n;
    const uint32_t kl = valid ? lin[i] : 0u, kh = valid ? hin[i] : 0u, val = valid ? vin[i] : 0u;
    const uint32_t d = ((high ? kh : kl) >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + rank;
      for (uint32_t ww = 0; ww < w; ww++) pos += wc[ww][d];
      lout[pos] = kl;
      hout[pos] = kh;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t ww = 0; ww < RS_THREADS / 64; ww++) add += wc[ww][tid];
    base[tid] += add;
    __syncthreads();
  }
}

/* ------------------------------------------------------------- PLOC --- */
struct BMeta {
  uint32_t count, units;
  float cost;
  uint32_t leafy;
};
/* a child id's count, record units, SAH cost and leaf flag: binary nodes
 * carry them in meta (x count, y units, z cost bits, w leaf); a primitive is
 * a one-primitive leaf of cost 1 */
__device__ __forceinline__ BMeta child_meta(uint32_t id, const uint4 *__restrict__ meta, const uint32_t *__restrict__ unit,
                                            const float4 *__restrict__ src) {
  BMeta r;
  if (id & BB_LEAF) {
    const float4 f0 = src[unit[id & ~BB_LEAF]];
    r.count = 1u;
    r.units = PREC_UNITS(__float_as_uint(f0.w), __float_as_uint(f0.z));
    r.cost = 1.0f;
    r.leafy = 1u;
  } else {
    const uint4 m = meta[id];
    r.count = m.x;
    r.units = m.y;
    r.cost = __uint_as_float(m.z);
    r.leafy = m.w;
  }
  return r;
}

__global__ __launch_bounds__(256) void k_bb_init(uint32_t n, const uint32_t *__restrict__ order,
                                                 const float4 *__restrict__ box, float4 *__restrict__ cb,
                                                 uint32_t *__restrict__ cid) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  cb[2u * i] = box[2u * p];
  cb[2u * i + 1u] = box[2u * p + 1u];
  cid[i] = BB_LEAF | p;
}

/* nearest neighbour of every cluster among the PL_R on either side: least
 * union area, ties broken by a key of the pair alone -- the index distance,
 * then the parity of the lower index, then the lower index.  A strict order
 * on pairs makes every chain of choices end in a mutual pair (no longer
 * cycles), and among equal clusters it pairs (0,1), (2,3), ..., so that
 * coincident primitives still halve per iteration instead of forming a chain. */
__global__ __launch_bounds__(256) void k_pl_nn(uint32_t cnt, const float4 *__restrict__ cb, uint32_t *__restrict__ nn) {
  __shared__ float4 slo[256 + 2 * PL_R], shi[256 + 2 * PL_R];
  const int s = (int)(blockIdx.x * 256u) - PL_R;
  for (int k = threadIdx.x; k < 256 + 2 * PL_R; k += 256) {
    const int g = s + k;
    if (g >= 0 && g < (int)cnt) {
      slo[k] = cb[2 * g];
      shi[k] = cb[2 * g + 1];
    }
  }
  __syncthreads();
  const int i = (int)(blockIdx.x * 256u + threadIdx.x);
  if (i >= (int)cnt) return;
  const float4 lo = slo[threadIdx.x + PL_R], hi = shi[threadIdx.x + PL_R];
  float best = FLT_MAX;
  int bj = -1;
  uint32_t bkey = 0xFFFFFFFFu;
  for (int d = -PL_R; d <= PL_R; d++) {
    const int j = i + d;
    if (d == 0 || j < 0 || j >= (int)cnt) continue;
    const int k = (int)threadIdx.x + PL_R + d;
    const float a = bb_area(bb_min(lo, slo[k]), bb_max(hi, shi[k]));
    const uint32_t lo_i = (uint32_t)min(i, j);
    /* |d| <= 16 (5 bits), parity (1 bit), lower index (26 bits, wraps only
     * beyond 2^26 clusters, where it still orders the pairs of one window) */
    const uint32_t key = ((uint32_t)abs(d) << 27) | ((lo_i & 1u) << 26) | (lo_i & 0x03FFFFFFu);
    if (bj < 0 || a < best || (a == best && key < bkey)) {
      best = a;
      bj = j;
      bkey = key;
    }
  }
  nn[i] = (uint32_t)bj;
}

/* per cluster: kept (1 in the low word) and starts a merge (1 in the high word) */
__global__ __launch_bounds__(256) void k_pl_flags(uint32_t cnt, const uint32_t *__restrict__ nn, uint64_t *__restrict__ fl) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  const uint64_t keep = (mutual && j < i) ? 0u : 1u, merge = (mutual && i < j) ? 1u : 0u;
  fl[i] = keep | (merge << 32);
}

/* wide-tree cost of a primitive test and of a 4-wide node visit, per unit of
 * surface area (the host collapse uses the same) */
#define DP_C_PRIM 1.0f
#define DP_C_NODE 1.0f

/* D(c, 1..4) of a child: a primitive is a one-primitive leaf in every case */
__device__ __forceinline__ void child_D(uint32_t id, float area, const float4 *__restrict__ bD, float d[4]) {
  if (id & BB_LEAF) {
    d[0] = d[1] = d[2] = d[3] = DP_C_PRIM * area;
  } else {
    const float4 v = bD[id];
    d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
  }
}

/* merge mutual pairs into new binary nodes (numbered node0 + merge prefix) and
 * compact the survivors in order; pre = exclusive prefix of the flags.  Each
 * new node gets its count, record units and the cost-optimal collapse's
 * decisions: D(n, k) the least wide-tree cost of its subtree as at most k
 * parts, F(n) = min(a leaf of at most LEAF_MAX_PRIMS primitives, a wide node
 * W(n) = C_NODE A(n) + min_j D(left, j) + D(right, 4 - j)), and
 * D(n, k) = min(F(n), min_j D(left, j) + D(right, k - j)).  bdec holds the j
 * of W (bits 0-1), the j of D(n, 2..4) (bits 2-7, 0 = F whole) and the leaf
 * flag (bit 8). */
__global__ __launch_bounds__(256) void k_pl_compact(uint32_t cnt, const uint32_t *__restrict__ nn,
                                                    const uint64_t *__restrict__ pre, const float4 *__restrict__ cbA,
                                                    const uint32_t *__restrict__ cidA, float4 *__restrict__ cbB,
                                                    uint32_t *__restrict__ cidB, uint32_t node0, float4 *__restrict__ bbox,
                                                    uint2 *__restrict__ bch, uint4 *__restrict__ bmeta,
                                                    float4 *__restrict__ bD, uint32_t *__restrict__ bdec,
                                                    const uint32_t *__restrict__ unit, const float4 *__restrict__ src) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  if (mutual && j < i) return;
  const uint64_t p = pre[i];
  const uint32_t dst = (uint32_t)p;
  float4 lo = cbA[2u * i], hi = cbA[2u * i + 1u];
  uint32_t id = cidA[i];
  if (mutual) {
    const uint32_t node = node0 + (uint32_t)(p >> 32);
    const float4 lo2 = cbA[2u * j], hi2 = cbA[2u * j + 1u];
    const uint32_t a = id, b = cidA[j];
    const BMeta ma = child_meta(a, bmeta, unit, src), mb = child_meta(b, bmeta, unit, src);
    const float4 ulo = bb_min(lo, lo2), uhi = bb_max(hi, hi2);
    const float A = bb_area(ulo, uhi), Aa = bb_area(lo, hi), Ab = bb_area(lo2, hi2);
    const uint32_t count = ma.count + mb.count;
    float da[4], db[4], dn[4];
    child_D(a, Aa, bD, da);
    child_D(b, Ab, bD, db);
    float bw = FLT_MAX;
    uint32_t dec = 1u;
#pragma unroll
    for (int jj = 1; jj <= 3; jj++) {
      const float c = da[jj - 1] + db[3 - jj];
      if (c < bw) {
        bw = c;
        dec = (uint32_t)jj;
      }
    }
    const float W = DP_C_NODE * A + bw;
    const float L = count <= LEAF_MAX_PRIMS ? DP_C_PRIM * A * (float)count : FLT_MAX;
    const bool leafy = L <= W;
    const float F = leafy ? L : W;
    dn[0] = F;
#pragma unroll
    for (int k = 2; k <= 4; k++) {
      float dk = F;
      uint32_t sk = 0;
#pragma unroll
      for (int jj = 1; jj < k; jj++) {
        const float c = da[jj - 1] + db[k - jj - 1];
        if (c < dk) {
          dk = c;
          sk = (uint32_t)jj;
        }
      }
      dn[k - 1] = dk;
      dec |= sk << (2 * (k - 1));
    }
    dec |= leafy ? 0x100u : 0u;
    bbox[2u * node] = ulo;
    bbox[2u * node + 1u] = uhi;
    bch[node] = make_uint2(a, b);
    bmeta[node] = make_uint4(count, ma.units + mb.units, __float_as_uint(F), leafy ? 1u : 0u);
    bD[node] = make_float4(dn[0], dn[1], dn[2], dn[3]);
    bdec[node] = dec;
    lo = ulo;
    hi = uhi;
    id = node;
  }
  cbB[2u * dst] = lo;
  cbB[2u * dst + 1u] = hi;
  cidB[dst] = id;
}

/* ------------------------------------------------------------ collapse --- */
__device__ __forceinline__ bool is_leafy(uint32_t id, const uint4 *__restrict__ bmeta) {
  return (id & BB_LEAF) || bmeta[id].w;
}

/* the children of one 4-wide node: the cost-optimal collapse's parts of its
 * binary node (read back from the merge decisions); count the interior ones
 * (next level) and the leaves' records */
__global__ __launch_bounds__(256) void k_cl_plan(uint32_t m, const uint32_t *__restrict__ F, const uint2 *__restrict__ bch,
                                                 const uint4 *__restrict__ bmeta, const uint32_t *__restrict__ bdec,
                                                 const uint32_t *__restrict__ unit, const float4 *__restrict__ src,
                                                 uint4 *__restrict__ plan, uint64_t *__restrict__ cntv) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= m) return;
  const uint32_t f = F[i];
  uint32_t ch[4] = {BB_NONE, BB_NONE, BB_NONE, BB_NONE};
  int nc = 0;
  if (is_leafy(f, bmeta)) { /* only a root can be a leaf */
    ch[nc++] = f;
  } else {
    /* split (node, k) pairs: k parts of node's subtree, as D(node, k) chose */
    uint32_t sn[8];
    int sk[8], sp = 0;
    const uint32_t w = bdec[f] & 3u;
    const uint2 c = bch[f];
    sn[sp] = c.y; sk[sp] = 4 - (int)w; sp++;
    sn[sp] = c.x; sk[sp] = (int)w; sp++;
    while (sp > 0) {
      sp--;
      const uint32_t n = sn[sp];
      const int k = sk[sp];
      const uint32_t j = ((n & BB_LEAF) || k == 1) ? 0u : (bdec[n] >> (2 * (k - 1))) & 3u;
      if (j == 0u) {
        if (nc < 4) ch[nc++] = n;
      } else {
        const uint2 c2 = bch[n];
        sn[sp] = c2.y; sk[sp] = k - (int)j; sp++;
        sn[sp] = c2.x; sk[sp] = (int)j; sp++;
      }
    }
  }
  uint32_t nint = 0, units = 0;
  for (int k = 0; k < nc; k++) {
    if (is_leafy(ch[k], bmeta)) units += child_meta(ch[k], bmeta, unit, src).units;
    else nint++;
  }
  plan[i] = make_uint4(ch[0], ch[1], ch[2], ch[3]);
  cntv[i] = (uint64_t)nint | ((uint64_t)units << 32);
}

/* write the level's 4-wide nodes, the next level's frontier and the leaf
 * records (a leaf's primitives in left-to-right order) */
__global__ __launch_bounds__(256) void k_cl_emit(uint32_t m, const uint4 *__restrict__ plan, const uint64_t *__restrict__ pre,
                                                 uint32_t lvl_base, uint32_t next_base, uint32_t unit_run, uint32_t node_base,
                                                 uint32_t unit_base, const float4 *__restrict__ bbox,
                                                 const uint2 *__restrict__ bch, const uint4 *__restrict__ bmeta,
                                                 const float4 *__restrict__ pbox, const uint32_t *__restrict__ unit,
                                                 const float4 *__restrict__ src, Bvh4Node *__restrict__ nodes,
                                                 float4 *__restrict__ dst, uint32_t *__restrict__ Fn) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= m) return;
  const uint4 pl = plan[i];
  const uint32_t ch[4] = {pl.x, pl.y, pl.z, pl.w};
  const uint64_t p = pre[i];
  uint32_t int_off = (uint32_t)p, u_off = unit_run + (uint32_t)(p >> 32);
  float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
  uint32_t cw[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t id = ch[k];
    if (id == BB_NONE) {
      lx[k] = ly[k] = lz[k] = FLT_MAX;
      hx[k] = hy[k] = hz[k] = -FLT_MAX;
      cw[k] = CHILD_EMPTY;
      continue;
    }
    float4 lo, hi;
    if (id & BB_LEAF) {
      lo = pbox[2u * (id & ~BB_LEAF)];
      hi = pbox[2u * (id & ~BB_LEAF) + 1u];
    } else {
      lo = bbox[2u * id];
      hi = bbox[2u * id + 1u];
    }
    lx[k] = lo.x; ly[k] = lo.y; lz[k] = lo.z;
    hx[k] = hi.x; hy[k] = hi.y; hz[k] = hi.z;
    if (is_leafy(id, bmeta)) {
      /* the leaf's primitives (at most LEAF_MAX_PRIMS), copied in scene order:
       * the sides of a Boxy that share a leaf are tested in the reference's
       * side order */
      const uint32_t u0 = u_off;
      uint32_t stk[LEAF_MAX_PRIMS + 1], sp = 0, cnt = 0, ps[LEAF_MAX_PRIMS];
      stk[sp++] = id;
      while (sp) {
        const uint32_t x = stk[--sp];
        if (x & BB_LEAF) {
          uint32_t q = x & ~BB_LEAF, j = cnt++;
          for (; j > 0 && ps[j - 1] > q; j--) ps[j] = ps[j - 1];
          ps[j] = q;
        } else {
          const uint2 c = bch[x];
          stk[sp++] = c.y;
          stk[sp++] = c.x;
        }
      }
      for (uint32_t t = 0; t < cnt; t++) {
        const uint32_t s0 = unit[ps[t]];
        const float4 f0 = src[s0];
        const uint32_t nu = PREC_UNITS(__float_as_uint(f0.w), __float_as_uint(f0.z));
        for (uint32_t u = 0; u < nu; u++) dst[unit_base + u_off + u] = src[s0 + u];
        u_off += nu;
      }
      cw[k] = CHILD_LEAF | ((cnt - 1u) << 28) | (unit_base + u0);
    } else {
      cw[k] = node_base + next_base + int_off;
      Fn[int_off] = id;
      int_off++;
    }
  }
  float4 *o = reinterpret_cast<float4 *>(nodes + node_base + lvl_base + i);
  o[0] = make_float4(lx[0], lx[1], lx[2], lx[3]);
  o[1] = make_float4(hx[0], hx[1], hx[2], hx[3]);
  o[2] = make_float4(ly[0], ly[1], ly[2], ly[3]);
  o[3] = make_float4(hy[0], hy[1], hy[2], hy[3]);
  o[4] = make_float4(lz[0], lz[1], lz[2], lz[3]);
  o[5] = make_float4(hz[0], hz[1], hz[2], hz[3]);
  o[6] = make_float4(__uint_as_float(cw[0]), __uint_as_float(cw[1]), __uint_as_float(cw[2]), __uint_as_float(cw[3]));
  o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

/* ---------------------------------------------------------------- host --- */
inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

template <typename T>
hipError_t scan(hipStream_t st, T *a, uint32_t m, T *bsum, T *total) {
  const uint32_t nb = cdiv(m, SC_TILE);
  hipLaunchKernelGGL(k_scan_local<T>, dim3(nb), dim3(SC_THREADS), 0, st, a, m, bsum);
  hipLaunchKernelGGL(k_scan_blocks<T>, dim3(1), dim3(1024), 0, st, bsum, nb, total);
  hipLaunchKernelGGL(k_scan_add<T>, dim3(nb), dim3(SC_THREADS), 0, st, a, m, (const T *)bsum);
  return hipGetLastError();
}

/* carve 256-B aligned pieces out of one allocation */
struct Carve {
  char *base = nullptr;
  size_t off = 0;
  template <typename T>
  T *take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T *p = reinterpret_cast<T *>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

}  // namespace

int device_build_bvh(hipStream_t st, const DeviceBvhInput &in, Bvh4Node *nodes, uint32_t node_base, float4 *dst,
                     uint32_t unit_base, uint32_t max_depth, DeviceBvhResult *res, std::string *err) {
  const uint32_t n = in.n;
  if (n == 0 || n >= BB_LEAF) {
    *err = "device BVH build: primitive count out of range";
    return CLV_EINVAL;
  }
  const uint32_t nbR = cdiv(n, RS_TILE);
  const uint64_t scan_max = std::max<uint64_t>(n, 256ull * nbR);
  const uint32_t nbS = cdiv(scan_max, SC_TILE);
  /* sizes of the pieces, counted once to allocate and again to carve */
  auto layout = [&](Carve &c) {
    struct P {
      uint32_t *k0, *h0, *v0, *k1, *h1, *v1, *hist, *nn, *cid0, *cid1, *F0, *F1;
      uint64_t *bsum, *tot, *fl;
      float4 *cb0, *cb1, *bbox, *bD;
      uint2 *bch;
      uint4 *bmeta, *plan;
      uint32_t *bdec;
    } p;
    p.k0 = c.take<uint32_t>(n);
    p.h0 = c.take<uint32_t>(n);
    p.v0 = c.take<uint32_t>(n);
    p.k1 = c.take<uint32_t>(n);
    p.h1 = c.take<uint32_t>(n);
    p.v1 = c.take<uint32_t>(n);
    p.hist = c.take<uint32_t>(256ull * nbR);
    p.nn = c.take<uint32_t>(n);
    p.cid0 = c.take<uint32_t>(n);
    p.cid1 = c.take<uint32_t>(n);
    p.F0 = c.take<uint32_t>(n);
    p.F1 = c.take<uint32_t>(n);
    p.bsum = c.take<uint64_t>(nbS + 1);
    p.tot = c.take<uint64_t>(2);
    p.fl = c.take<uint64_t>(n);
    p.cb0 = c.take<float4>(2ull * n);
    p.cb1 = c.take<float4>(2ull * n);
    p.bbox = c.take<float4>(2ull * n);
    p.bch = c.take<uint2>(n);
    p.bmeta = c.take<uint4>(n);
    p.plan = c.take<uint4>(n);
    p.bD = c.take<float4>(n);
    p.bdec = c.take<uint32_t>(n);
    return p;
  };
  Carve sizing;
  (void)layout(sizing);
  char *mem = nullptr;
  hipError_t e = hipMalloc((void **)&mem, sizing.off + 256);
  if (e != hipSuccess) {
    *err = std::string("device BVH build: ") + hipGetErrorString(e);
    return CLV_ENOMEM;
  }
  Carve carve;
  carve.base = mem;
  auto P = layout(carve);
  int status = CLV_OK;
#define BBCHK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      *err = std::string("device BVH build: ") + hipGetErrorString(e_);      \
      status = CLV_EDEVICE;                                                   \
      goto done;                                                              \
    }                                                                         \
  } while (0)
  {
    /* 1. 63-bit Morton codes (one scale for all axes: the centres' largest
     * extent), stable radix sort (8 passes, result back in k0/h0/v0) */
    float s[3];
    float ext = 0.0f;
    for (int a = 0; a < 3; a++) ext = fmaxf(ext, in.chi[a] - in.clo[a]);
    for (int a = 0; a < 3; a++) s[a] = ext > 0.0f ? 2097152.0f / ext : 0.0f;
    const uint32_t g = cdiv(n, 256);
    hipLaunchKernelGGL(k_bb_morton, dim3(g), dim3(256), 0, st, n, in.box, in.clo[0], in.clo[1], in.clo[2], s[0], s[1],
                       s[2], P.k0, P.h0, P.v0);
    BBCHK(hipGetLastError());
    uint32_t *lin = P.k0, *hin = P.h0, *vin = P.v0, *lout = P.k1, *hout = P.h1, *vout = P.v1;
    for (uint32_t pass = 0; pass < 8; pass++) {
      const uint32_t high = pass >> 2, shift = 8u * (pass & 3u);
      hipLaunchKernelGGL(k_rs_hist, dim3(nbR), dim3(RS_THREADS), 0, st, n, (const uint32_t *)(high ? hin : lin), shift,
                         nbR, P.hist);
      BBCHK(hipGetLastError());
      BBCHK(scan<uint32_t>(st, P.hist, 256u * nbR, reinterpret_cast<uint32_t *>(P.bsum),
                           reinterpret_cast<uint32_t *>(P.tot)));
      hipLaunchKernelGGL(k_rs_scatter, dim3(nbR), dim3(RS_THREADS), 0, st, n, (const uint32_t *)lin,
                         (const uint32_t *)hin, (const uint32_t *)vin, lout, hout, vout, high, shift, nbR,
                         (const uint32_t *)P.hist);
      BBCHK(hipGetLastError());
      std::swap(lin, lout);
      std::swap(hin, hout);
      std::swap(vin, vout);
    }
    /* 2. clustering */
    hipLaunchKernelGGL(k_bb_init, dim3(g), dim3(256), 0, st, n, (const uint32_t *)vin,
                       in.box, P.cb0, P.cid0);
    BBCHK(hipGetLastError());
    float4 *cbA = P.cb0, *cbB = P.cb1;
    uint32_t *cidA = P.cid0, *cidB = P.cid1;
    uint32_t cnt = n, node0 = 0, iters = 0;
    while (cnt > 1) {
      const uint32_t gc = cdiv(cnt, 256);
      hipLaunchKernelGGL(k_pl_nn, dim3(gc), dim3(256), 0, st, cnt, (const float4 *)cbA, P.nn);
      hipLaunchKernelGGL(k_pl_flags, dim3(gc), dim3(256), 0, st, cnt, (const uint32_t *)P.nn, P.fl);
      BBCHK(hipGetLastError());
      BBCHK(scan<uint64_t>(st, P.fl, cnt, P.bsum, P.tot));
      hipLaunchKernelGGL(k_pl_compact, dim3(gc), dim3(256), 0, st, cnt, (const uint32_t *)P.nn,
                         (const uint64_t *)P.fl, (const float4 *)cbA, (const uint32_t *)cidA, cbB, cidB, node0, P.bbox,
                         P.bch, P.bmeta, P.bD, P.bdec, in.unit, in.src);
      BBCHK(hipGetLastError());
      uint64_t tot = 0;
      BBCHK(hipMemcpyAsync(&tot, P.tot, 8, hipMemcpyDeviceToHost, st));
      BBCHK(hipStreamSynchronize(st));
      const uint32_t kept = (uint32_t)tot, merged = (uint32_t)(tot >> 32);
      if (merged == 0 || kept != cnt - merged) {
        *err = "device BVH build: clustering made no progress";
        status = CLV_EINTERNAL;
        goto done;
      }
      node0 += merged;
      cnt = kept;
      std::swap(cbA, cbB);
      std::swap(cidA, cidB);
      iters++;
    }
    if (node0 != n - 1) {
      *err = "device BVH build: wrong binary node count";
      status = CLV_EINTERNAL;
      goto done;
    }
    /* 3. collapse, one 4-wide level at a time; the root cluster id seeds it */
    BBCHK(hipMemcpyAsync(P.F0, cidA, 4, hipMemcpyDeviceToDevice, st));
    uint32_t *F = P.F0, *Fn = P.F1;
    uint32_t m = 1, lvl_base = 0, unit_run = 0, level = 0;
    while (m > 0) {
      if (level > max_depth) {
        *err = "device-built BVH deeper than the traversal stack allows (use the host builder)";
        status = CLV_EUNSUPPORTED;
        goto done;
      }
      const uint32_t gm = cdiv(m, 256);
      hipLaunchKernelGGL(k_cl_plan, dim3(gm), dim3(256), 0, st, m, (const uint32_t *)F, (const uint2 *)P.bch,
                         (const uint4 *)P.bmeta, (const uint32_t *)P.bdec, in.unit, in.src, P.plan, P.fl);
      BBCHK(hipGetLastError());
      BBCHK(scan<uint64_t>(st, P.fl, m, P.bsum, P.tot));
      hipLaunchKernelGGL(k_cl_emit, dim3(gm), dim3(256), 0, st, m, (const uint4 *)P.plan, (const uint64_t *)P.fl,
                         lvl_base, lvl_base + m, unit_run, node_base, unit_base, (const float4 *)P.bbox,
                         (const uint2 *)P.bch, (const uint4 *)P.bmeta, in.box, in.unit, in.src, nodes, dst, Fn);
      BBCHK(hipGetLastError());
      uint64_t tot = 0;
      BBCHK(hipMemcpyAsync(&tot, P.tot, 8, hipMemcpyDeviceToHost, st));
      BBCHK(hipStreamSynchronize(st));
      unit_run += (uint32_t)(tot >> 32);
      lvl_base += m;
      m = (uint32_t)tot;
      std::swap(F, Fn);
      level++;
    }
    res->n_nodes = lvl_base;
    res->depth = level - 1;
    res->units = unit_run;
    res->iterations = iters;
  }
done:
#undef BBCHK
  (void)hipStreamSynchronize(st);
  (void)hipFree(mem);
  return status;
}

}  // namespace clv
