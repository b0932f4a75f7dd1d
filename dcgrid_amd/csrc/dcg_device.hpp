// dcg_device.hpp — device-side helpers shared by the gfx950 kernels.
//
// Float expressions are written in the evaluation order of the reference's kernels and the
// library is compiled with -ffp-contract=off, so each float result is the IEEE result of the
// same operation sequence on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dcgrid_amd.h"

namespace dcg {

static constexpr uint32_t NFU = 0xFFFFFFFFu;  // "not found" block / subblock index
static constexpr uint8_t FREE = 0xFF;         // blockLevels value of a free block
static constexpr uint8_t MOVED = 1, REFINED = 2;  // DCGrid::blockMoved / blockRefined

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Morton index of a local cell (x most significant bit of each triple); BB = log2(BW).
template <int BB>
__device__ __forceinline__ uint32_t loc(int x, int y, int z) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < BB; b++)
    r |= ((uint32_t)((x >> b) & 1) << (3 * b + 2)) | ((uint32_t)((y >> b) & 1) << (3 * b + 1)) |
         ((uint32_t)((z >> b) & 1) << (3 * b));
  return r;
}
template <int BB>
__device__ __forceinline__ void locPos(uint32_t l, int& x, int& y, int& z) {
  x = y = z = 0;
#pragma unroll
  for (int b = 0; b < BB; b++) {
    x |= (int)((l >> (3 * b + 2)) & 1) << b;
    y |= (int)((l >> (3 * b + 1)) & 1) << b;
    z |= (int)((l >> (3 * b)) & 1) << b;
  }
}

// Fixed-point accumulation in LDS for the adjoint's scatter tiles. LDS float atomics run at
// about 1/35 of the rate of integer ones on this hardware (measured: 0.2 against 7.4 T
// lane-ops/s), so a tile holds 64-bit integers on a scale 2^E of its own, E = 50 - (binary
// exponent of the largest |lam| the tile's workgroup scatters): a contribution w lam (w <= 1) is
// rounded to a multiple of 2^-E, i.e. to 2^-50 of that largest value, up to 2^13 such
// contributions fit in one entry, and the integer sums do not depend on their order.
struct FixScale {
  double up, down;
};
__device__ __forceinline__ FixScale fixScale(uint32_t maxBits) {
  int ex = 0;
  frexpf(__uint_as_float(maxBits), &ex);
  return FixScale{ldexp(1.0, 50 - ex), ldexp(1.0, ex - 50)};
}
__device__ __forceinline__ void fixAdd(unsigned long long* p, float v, const FixScale& sc) {
  atomicAdd(p, (unsigned long long)__double2ll_rn((double)v * sc.up));
}
__device__ __forceinline__ float fixGet(unsigned long long v, const FixScale& sc) {
  return (float)((double)(long long)v * sc.down);
}
__device__ __forceinline__ void ldsMax(uint32_t* p, float v) { atomicMax(p, __float_as_uint(fabsf(v))); }

// Inlet / outside tests of velocityBndCond and densityBndCond (src/utils/sim_utils.cu:24-55).
__device__ __forceinline__ bool inletHit(const dcg_params& P, int x, int y, int z, int scale) {
  if (y >= 0) return false;
  const float a = (float)(x * scale) - .5f * (float)P.gx;
  const float b = (float)(z * scale) - .5f * (float)P.gz;
  return __builtin_sqrtf(a * a + b * b) < P.emission_radius * P.rdx;
}
__device__ __forceinline__ bool outside(const dcg_params& P, int x, int y, int z, int scale) {
  return x < 0 || y < 0 || z < 0 || x * scale >= P.gx || y * scale >= P.gy || z * scale >= P.gz;
}
// Deterministic exp: Cody-Waite reduction by ln 2 (split in a float-exact high part and a low
// part) and the degree-7 Taylor polynomial of exp on |r| <= ln2/2, scaled by 2^n. The same
// float operations run on the host, so both produce the same bits.
__device__ __forceinline__ float expDet(float x) {
  const float n = floorf(x * 1.44269504f + .5f);
  const float r = (x - n * 0.693145752f) - n * 1.42860677e-06f;
  const float p =
      1.f + r * (1.f + r * (.5f + r * (1.f / 6.f + r * (1.f / 24.f + r * (1.f / 120.f + r * (1.f / 720.f + r * (1.f / 5040.f)))))));
  return ldexpf(p, (int)n);
}
// Moist atmosphere (extension): ambient temperature at height hy (finest cells), saturation
// vapor (Magnus form), ambient vapor profile.
__device__ __forceinline__ float ambientT(const dcg_params& P, float hy) {
  return P.moist ? P.t_ambient - P.lapse_rate * hy : P.t_ambient;
}
__device__ __forceinline__ float qsat(const dcg_params& P, float T) {
  return P.qsat_scale * expDet(17.27f * T / (T + 237.3f));
}
__device__ __forceinline__ float ambientVapor(const dcg_params& P, float hy) {
  return hy < P.moist_layer ? P.humidity * qsat(P, ambientT(P, hy)) : 0.f;
}
// Height (finest cells) of the centre of level-local cell row y.
__device__ __forceinline__ float cellHeight(int y, int scale) { return ((float)y + .5f) * (float)scale; }

__device__ __forceinline__ float3 velocityBC(const dcg_params& P, float3 v, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return make_float3(0.f, P.velocity_emission_rate, 0.f);
  if (outside(P, x, y, z, scale)) {
    // lateral boundaries carry the ambient wind (zero unless configured); floor and lid are walls
    if (y >= 0 && y * scale < P.gy) return make_float3(P.wind[0], P.wind[1], P.wind[2]);
    return make_float3(0.f, 0.f, 0.f);
  }
  return v;
}
__device__ __forceinline__ float densityBC(const dcg_params& P, float q, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return P.density_emission_rate;
  if (outside(P, x, y, z, scale)) return 0.f;
  return q;
}
__device__ __forceinline__ float temperatureBC(const dcg_params& P, float q, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return P.t_ambient;
  if (outside(P, x, y, z, scale)) return ambientT(P, cellHeight(y, scale));
  return q;
}
__device__ __forceinline__ float vaporBC(const dcg_params& P, float q, int x, int y, int z, int scale) {
  if (outside(P, x, y, z, scale)) return ambientVapor(P, cellHeight(y, scale));
  return q;
}
// Saturation adjustment of one cell: condense a fraction of the supersaturation (or evaporate
// cloud water into an undersaturated cell), with latent heating.
__device__ __forceinline__ void saturationAdjust(const dcg_params& P, float& qv, float& qc, float& T) {
  const float ex = qv - qsat(P, T);
  float c = P.cond_rate * ex;
  if (c < -qc) c = -qc;
  qv = qv - c;
  qc = qc + c;
  T = T + P.latent_heat * c;
}

// getCellFluidity (src/utils/sim_utils.cu:11-22) with the sphere of src/sdf.cuh:8-18, or with a
// terrain heightfield (extension): signed distance = height of the cell centre above the
// terrain.
__device__ __forceinline__ float terrainHeight(const dcg_params& P, float x, float z);
__device__ __forceinline__ float cellFluidity(const dcg_params& P, int x, int y, int z, int scale) {
  const float sqrt3 = 1.73205f;
  if (!P.enable_additional_solids) return 1.f;
  const float px = ((float)x + .5f) * (float)scale;
  const float py = ((float)y + .5f) * (float)scale;
  const float pz = ((float)z + .5f) * (float)scale;
  float d;
  if (P.solid_mode == DCG_SOLID_TERRAIN) {
    d = py - terrainHeight(P, px, pz);
  } else {
    const float cx = (float)P.gx * .5f, cy = (float)P.gy * .45f, cz = (float)P.gz * .5f;
    const float r = 2000.f * P.rdx;
    const float dx = px - cx, dy = py - cy, dz = pz - cz;
    d = __builtin_sqrtf(dx * dx + dy * dy + dz * dz) - r;
  }
  const float overlap = clampf(.5f - d / ((float)scale * sqrt3), 0.f, 1.f);
  return 1.f - overlap;
}

// ---- seeded synthetic inputs: counter-based hash and value noise ----
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ float hash01(uint32_t seed, int x, int y, int z, int w) {
  uint32_t h = mix32(seed ^ 0x9e3779b9u);
  h = mix32(h ^ (uint32_t)x);
  h = mix32(h ^ (uint32_t)y ^ 0x68e31da4u);
  h = mix32(h ^ (uint32_t)z ^ 0xb5297a4du);
  h = mix32(h ^ (uint32_t)w ^ 0x1b56c4e9u);
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float valueNoise(float x, float y, float z, uint32_t seed) {
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
  const float fx = x - xf, fy = y - yf, fz = z - zf;
  const float sx = fx * fx * (3.f - 2.f * fx);
  const float sy = fy * fy * (3.f - 2.f * fy);
  const float sz = fz * fz * (3.f - 2.f * fz);
  float v[8];
#pragma unroll
  for (int c = 0; c < 8; c++)
    v[c] = 2.f * hash01(seed, ix + ((c >> 2) & 1), iy + ((c >> 1) & 1), iz + (c & 1), 0) - 1.f;
  const float a00 = v[0] + sx * (v[4] - v[0]);
  const float a10 = v[2] + sx * (v[6] - v[2]);
  const float a01 = v[1] + sx * (v[5] - v[1]);
  const float a11 = v[3] + sx * (v[7] - v[3]);
  const float b0 = a00 + sy * (a10 - a00);
  const float b1 = a01 + sy * (a11 - a01);
  return b0 + sz * (b1 - b0);
}
__device__ __forceinline__ float fbm3(float x, float y, float z, float period, uint32_t seed) {
  float n = 0.f, amp = 1.f, freq = 1.f / period;
  for (int o = 0; o < 3; o++) {
    n = n + amp * valueNoise(x * freq, y * freq, z * freq, seed + 1013u * (uint32_t)o);
    amp = amp * .5f;
    freq = freq * 2.f;
  }
  return n / 1.75f;
}
// Terrain heightfield (finest cells): 5-octave fBm of value noise mapped to [0, terrain_height].
__device__ __forceinline__ float terrainHeight(const dcg_params& P, float x, float z) {
  float n = 0.f, amp = 1.f, freq = 1.f / P.terrain_period;
  const uint32_t seed = P.seed ^ 0x3c6ef372u;
  for (int o = 0; o < 5; o++) {
    n = n + amp * valueNoise(x * freq, 0.f, z * freq, seed + 7919u * (uint32_t)o);
    amp = amp * .5f;
    freq = freq * 2.f;
  }
  return P.terrain_height * (.5f + .5f * (n / 1.9375f));
}
__device__ __forceinline__ float3 initVelocity(const dcg_params& P, float cx, float cy, float cz) {
  float u[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (P.init_velocity_noise == 0.f) {
      u[c] = 0.f;
      continue;
    }
    const uint32_t s = P.seed ^ (0x51ed270bu * (uint32_t)(c + 1));
    u[c] = P.init_velocity_noise * fbm3(cx, cy, cz, P.noise_period, s);
  }
  return make_float3(u[0] + P.wind[0], u[1] + P.wind[1], u[2] + P.wind[2]);
}
__device__ __forceinline__ bool inSource(const dcg_params& P, float cx, float cy, float cz) {
  const float dx = cx - P.src_center[0], dy = cy - P.src_center[1], dz = cz - P.src_center[2];
  return dx * dx + dy * dy + dz * dz < P.src_radius * P.src_radius;
}
__device__ __forceinline__ float sourceTemperature(const dcg_params& P, int x, int y, int z, int level) {
  const float u = hash01(P.seed ^ 0x2545f491u, x, y, z, level);
  return P.src_temperature + P.src_noise * (2.f * u - 1.f);
}

// float3 helpers with explicit evaluation order
__device__ __forceinline__ float3 f3mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }

// XCD-aware remap of a linear workgroup id: consecutive ids are dispatched round-robin over the
// 8 XCDs of MI355X; this maps them so each XCD works on a contiguous range (shared L2 reuse).
__device__ __forceinline__ uint32_t xcdRemap(uint32_t id, uint32_t n) {
  const uint32_t xcd = id & 7u, k = id >> 3;
  const uint32_t per = n >> 3, rem = n & 7u;
  // XCDs 0..rem-1 own per+1 ids, the others own per ids
  const uint32_t base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  const uint32_t own = xcd < rem ? per + 1 : per;
  return k < own ? base + k : id;  // ids beyond a full round keep their position
}

}  // namespace dcg
