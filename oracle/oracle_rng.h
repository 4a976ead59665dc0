/*
 * ORACLE -- test infrastructure only.  Not part of the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Deterministic per-path random streams shared (by definition, not by code)
 * with the GPU renderer.  The reference seeds two SmallRng generators per row
 * from the OS (clovers-cli/src/draw_cpu.rs:56,62), which no two runs
 * reproduce; this build instead keys one PCG32 stream (O'Neill 2014,
 * "PCG: A Family of Simple Fast Space-Efficient Statistically Good Algorithms
 * for Random Number Generation", pcg32 = 64-bit LCG + XSH-RR output) per
 * (seed, pixel, sample).  The sampler draws first, then the tracer, in the
 * reference's order (sampler/random.rs:20-32, trace.rs:21-122).
 *
 * The unit-disc and unit-sphere samplers follow the published rejection
 * methods that rand_distr 0.5.1 (UnitDisc, UnitSphere) implements: rejection
 * in the square [-1,1)^2, and Marsaglia 1972 for the sphere.
 */
#ifndef ORACLE_RNG_H
#define ORACLE_RNG_H
#include <stdint.h>
#include <math.h>

typedef struct { uint64_t state; } orc_pcg;

static inline uint64_t orc_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

#define ORC_PCG_MULT 6364136223846793005ull
#define ORC_PCG_INC 1442695040888963407ull

static inline uint32_t orc_pcg_next(orc_pcg *r) {
  uint64_t old = r->state;
  r->state = old * ORC_PCG_MULT + ORC_PCG_INC;
  uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

/* stream for (seed, pixel index x + y*W of the reference, sample index) */
static inline orc_pcg orc_pcg_for_path(uint64_t seed, uint32_t pixel, uint32_t sample) {
  orc_pcg r;
  uint64_t key = ((uint64_t)sample << 32) | (uint64_t)pixel;
  r.state = orc_splitmix64(seed ^ orc_splitmix64(key));
  (void)orc_pcg_next(&r);
  return r;
}

/* uniform f32 in [0,1): top 24 bits */
static inline float orc_uniform(orc_pcg *r) { return (float)(orc_pcg_next(r) >> 8) * 0x1p-24f; }
/* fair coin */
static inline int orc_bool(orc_pcg *r) { return (int)(orc_pcg_next(r) >> 31); }
/* uniform f32 in [-1,1) */
static inline float orc_uniform_pm1(orc_pcg *r) { return 2.0f * orc_uniform(r) - 1.0f; }
/* uniform integer wavelength in [380,780) */
static inline int orc_wavelength(orc_pcg *r) {
  return 380 + (int)(((uint64_t)orc_pcg_next(r) * 400ull) >> 32);
}

#endif
