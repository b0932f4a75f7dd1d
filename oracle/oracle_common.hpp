// oracle_common.hpp — shared pieces of the CPU oracle (test infrastructure only; see oracle.h).
//
// Everything here restates reference semantics with the evaluation order of the reference's
// float expressions, so that the float instantiation is a bit-for-bit twin of IEEE float code
// compiled without contraction (-ffp-contract=off).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

namespace orc {

static constexpr size_t NF = SIZE_MAX;  // DCGrid::notFound, src/dcgrid/dcgrid.h:69
static constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;  // DCGrid::hashEmpty, dcgrid.h:70

template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
inline V3<R> mk3(R a, R b, R c) {
  return V3<R>{a, b, c};
}
template <class R>
inline V3<R> add(const V3<R>& a, const V3<R>& b) {
  return V3<R>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
inline V3<R> sub(const V3<R>& a, const V3<R>& b) {
  return V3<R>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
inline V3<R> mul(const V3<R>& a, R s) {
  return V3<R>{a.x * s, a.y * s, a.z * s};
}

// int iDivUp, src/utils/grid_math.cuh:5-7
inline int iDivUp(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// mipmapCells / mipmapIdx, src/utils/grid_math.cuh:24-52
inline size_t mipmapCells(size_t w, size_t h, size_t d) {
  size_t n = 0;
  for (size_t s = 1; w % s == 0 && h % s == 0 && d % s == 0; s *= 2) n += (w * h * d) / (s * s * s);
  return n;
}
inline size_t mipmapIdx(int x, int y, int z, size_t w, size_t h, size_t d, size_t scale) {
  size_t off = 0;
  for (size_t s = 1; s < scale; s *= 2) off += (w * h * d) / (s * s * s);
  return off + ((size_t)z * (h / scale) + (size_t)y) * (w / scale) + (size_t)x;
}

// bitSpread3 / gridHash, src/dcgrid/dcgrid_utils.cuh:83-96 (bit i of v -> bit 3i, 11 bits)
inline uint32_t bitSpread3(uint32_t data) {
  uint32_t res = 0;
  for (uint32_t mask = 1u; mask; mask <<= 3, data <<= 2) res |= data & mask;
  return res;
}
inline uint32_t gridHash(int x, int y, int z, int bw) {
  return (bitSpread3((uint32_t)(x / bw)) << 2) | (bitSpread3((uint32_t)(y / bw)) << 1) |
         bitSpread3((uint32_t)(z / bw));
}

// Local cell index inside a block: 3-D Morton code with x the most significant bit of each
// triple. For BW=4 this is [xh yh zh xl yl zl] (dcgrid_utils.cuh:50-58,67-81; interior fill
// order dcgrid_structure.cu:18-27).
inline uint32_t localIndex(int x, int y, int z, int bb) {
  uint32_t r = 0;
  for (int b = 0; b < bb; b++)
    r |= ((uint32_t)((x >> b) & 1) << (3 * b + 2)) | ((uint32_t)((y >> b) & 1) << (3 * b + 1)) |
         ((uint32_t)((z >> b) & 1) << (3 * b));
  return r;
}
inline void localPos(uint32_t l, int bb, int& x, int& y, int& z) {
  x = y = z = 0;
  for (int b = 0; b < bb; b++) {
    x |= (int)((l >> (3 * b + 2)) & 1) << b;
    y |= (int)((l >> (3 * b + 1)) & 1) << b;
    z |= (int)((l >> (3 * b)) & 1) << b;
  }
}

// velocityBndCond / densityBndCond, src/utils/sim_utils.cu:24-55. The geometric tests are
// float expressions as in the reference.
inline bool inletHit(const dcg_params& P, int x, int y, int z, int scale) {
  if (y >= 0) return false;
  const float a = (float)(x * scale) - .5f * (float)P.gx;
  const float b = (float)(z * scale) - .5f * (float)P.gz;
  return sqrtf(a * a + b * b) < P.emission_radius * P.rdx;
}
inline bool outside(const dcg_params& P, int x, int y, int z, int scale) {
  return x < 0 || y < 0 || z < 0 || x * scale >= P.gx || y * scale >= P.gy || z * scale >= P.gz;
}
// Deterministic exp (extension helper, identical float operations on the GPU): Cody-Waite
// reduction by ln 2 and the degree-7 Taylor polynomial on |r| <= ln2/2, scaled by 2^n.
inline float expDet(float x) {
  const float n = floorf(x * 1.44269504f + .5f);
  const float r = (x - n * 0.693145752f) - n * 1.42860677e-06f;
  const float p =
      1.f + r * (1.f + r * (.5f + r * (1.f / 6.f + r * (1.f / 24.f + r * (1.f / 120.f + r * (1.f / 720.f + r * (1.f / 5040.f)))))));
  return ldexpf(p, (int)n);
}
// Moist atmosphere (extension; no counterpart in the reference).
inline float ambientT(const dcg_params& P, float hy) { return P.moist ? P.t_ambient - P.lapse_rate * hy : P.t_ambient; }
inline float qsat(const dcg_params& P, float T) { return P.qsat_scale * expDet(17.27f * T / (T + 237.3f)); }
inline float ambientVapor(const dcg_params& P, float hy) {
  return hy < P.moist_layer ? P.humidity * qsat(P, ambientT(P, hy)) : 0.f;
}
inline float cellHeight(int y, int scale) { return ((float)y + .5f) * (float)scale; }

template <class R>
inline V3<R> velocityBC(const dcg_params& P, const V3<R>& v, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return V3<R>{(R)0, (R)P.velocity_emission_rate, (R)0};
  if (outside(P, x, y, z, scale)) {
    // lateral boundaries carry the ambient wind (zero unless configured); floor and lid are walls
    if (y >= 0 && y * scale < P.gy) return V3<R>{(R)P.wind[0], (R)P.wind[1], (R)P.wind[2]};
    return V3<R>{(R)0, (R)0, (R)0};
  }
  return v;
}
template <class R>
inline R densityBC(const dcg_params& P, R q, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return (R)P.density_emission_rate;
  if (outside(P, x, y, z, scale)) return (R)0;
  return q;
}
// Temperature has no counterpart in the reference; it behaves like density except that
// the inlet value is the ambient temperature and outside cells hold the ambient profile.
template <class R>
inline R temperatureBC(const dcg_params& P, R q, int x, int y, int z, int scale) {
  if (inletHit(P, x, y, z, scale)) return (R)P.t_ambient;
  if (outside(P, x, y, z, scale)) return (R)ambientT(P, cellHeight(y, scale));
  return q;
}
template <class R>
inline R vaporBC(const dcg_params& P, R q, int x, int y, int z, int scale) {
  if (outside(P, x, y, z, scale)) return (R)ambientVapor(P, cellHeight(y, scale));
  return q;
}
// Saturation adjustment of one cell (condensation / evaporation with latent heating).
template <class R>
inline void saturationAdjust(const dcg_params& P, R& qv, R& qc, R& T) {
  const R ex = qv - (R)qsat(P, (float)T);
  R c = (R)P.cond_rate * ex;
  if (c < -qc) c = -qc;
  qv = qv - c;
  qc = qc + c;
  T = T + (R)P.latent_heat * c;
}

// getCellFluidity, src/utils/sim_utils.cu:11-22, with sphereSDF src/sdf.cuh:8-18, or with the
// terrain heightfield (extension): signed distance = height of the cell centre above the terrain.
inline float terrainHeight(const dcg_params& P, float x, float z);
inline float cellFluidity(const dcg_params& P, int x, int y, int z, int scale) {
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
    d = sqrtf(dx * dx + dy * dy + dz * dz) - r;
  }
  const float overlap = clampf(.5f - d / ((float)scale * sqrt3), 0.f, 1.f);
  return 1.f - overlap;
}

// ---- Seeded synthetic inputs (extension; identical definition in the HIP library) ----
inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
inline uint32_t hash4(uint32_t seed, int x, int y, int z, int w) {
  uint32_t h = mix32(seed ^ 0x9e3779b9u);
  h = mix32(h ^ (uint32_t)x);
  h = mix32(h ^ (uint32_t)y ^ 0x68e31da4u);
  h = mix32(h ^ (uint32_t)z ^ 0xb5297a4du);
  h = mix32(h ^ (uint32_t)w ^ 0x1b56c4e9u);
  return h;
}
inline float hash01(uint32_t seed, int x, int y, int z, int w) {
  return (float)(hash4(seed, x, y, z, w) >> 8) * (1.0f / 16777216.0f);
}
// One octave of value noise: lattice values in [-1,1), smoothstep-weighted trilinear blend.
inline float valueNoise(float x, float y, float z, uint32_t seed) {
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
  const float fx = x - xf, fy = y - yf, fz = z - zf;
  const float sx = fx * fx * (3.f - 2.f * fx);
  const float sy = fy * fy * (3.f - 2.f * fy);
  const float sz = fz * fz * (3.f - 2.f * fz);
  float v[8];
  for (int c = 0; c < 8; c++)
    v[c] = 2.f * hash01(seed, ix + ((c >> 2) & 1), iy + ((c >> 1) & 1), iz + (c & 1), 0) - 1.f;
  // v index bits: [x y z]
  const float a00 = v[0] + sx * (v[4] - v[0]);
  const float a10 = v[2] + sx * (v[6] - v[2]);
  const float a01 = v[1] + sx * (v[5] - v[1]);
  const float a11 = v[3] + sx * (v[7] - v[3]);
  const float b0 = a00 + sy * (a10 - a00);
  const float b1 = a01 + sy * (a11 - a01);
  return b0 + sz * (b1 - b0);
}
// 3-octave fBm normalised to [-1,1] (amplitudes 1, 1/2, 1/4).
inline float fbm3(float x, float y, float z, float period, uint32_t seed) {
  float n = 0.f, amp = 1.f, freq = 1.f / period;
  for (int o = 0; o < 3; o++) {
    n = n + amp * valueNoise(x * freq, y * freq, z * freq, seed + 1013u * (uint32_t)o);
    amp = amp * .5f;
    freq = freq * 2.f;
  }
  return n / 1.75f;
}
// Terrain heightfield (finest cells): 5-octave fBm of value noise mapped to [0, terrain_height].
inline float terrainHeight(const dcg_params& P, float x, float z) {
  float n = 0.f, amp = 1.f, freq = 1.f / P.terrain_period;
  const uint32_t seed = P.seed ^ 0x3c6ef372u;
  for (int o = 0; o < 5; o++) {
    n = n + amp * valueNoise(x * freq, 0.f, z * freq, seed + 7919u * (uint32_t)o);
    amp = amp * .5f;
    freq = freq * 2.f;
  }
  return P.terrain_height * (.5f + .5f * (n / 1.9375f));
}
// Initial velocity of the seeded scene at a cell centre (finest-cell units).
inline void initVelocity(const dcg_params& P, float cx, float cy, float cz, float out[3]) {
  for (int c = 0; c < 3; c++) {
    if (P.init_velocity_noise == 0.f) {
      out[c] = 0.f;
      continue;
    }
    const uint32_t s = P.seed ^ (0x51ed270bu * (uint32_t)(c + 1));
    out[c] = P.init_velocity_noise * fbm3(cx, cy, cz, P.noise_period, s);
  }
  for (int c = 0; c < 3; c++) out[c] = out[c] + P.wind[c];
}
// Sphere source membership of a cell centre (finest-cell units).
inline bool inSource(const dcg_params& P, float cx, float cy, float cz) {
  const float dx = cx - P.src_center[0], dy = cy - P.src_center[1], dz = cz - P.src_center[2];
  return dx * dx + dy * dy + dz * dz < P.src_radius * P.src_radius;
}
inline float sourceTemperature(const dcg_params& P, int x, int y, int z, int level) {
  const float u = hash01(P.seed ^ 0x2545f491u, x, y, z, level);
  return P.src_temperature + P.src_noise * (2.f * u - 1.f);
}

// ---------------------------------------------------------------- field storage formats
// Value of v after a store into a field of storage format `storage` (dcg_params.storage, an
// extension): IEEE binary16 or bfloat16, round to nearest even, read back exactly as float.
// Restated from the formats' definitions: the quantum of v is 2^(E - mantissa bits) with E the
// exponent of v clamped to the format's minimum normal exponent (subnormals share that quantum);
// std::nearbyint rounds v / quantum to nearest even (default rounding mode) and both scalings
// by a power of two are exact. binary16 overflows to infinity above 65504 after rounding.
inline float roundStorage(float v, int storage) {
  if (storage == DCG_STORAGE_FP32 || v == 0.f || !std::isfinite(v)) return v;
  int e;
  std::frexp(v, &e);  // |v| = m 2^e, m in [0.5, 1)
  const int mant = storage == DCG_STORAGE_FP16 ? 10 : 7;
  const int emin = storage == DCG_STORAGE_FP16 ? -14 : -126;
  const int E = e - 1 < emin ? emin : e - 1;
  const float q = std::ldexp(1.f, E - mant);
  const float r = std::nearbyint(v / q) * q;
  if (storage == DCG_STORAGE_FP16 && std::fabs(r) > 65504.f) return std::copysign(INFINITY, v);
  return r;
}

struct Sim {
  dcg_params P;
  std::string err;
  virtual ~Sim() {}
  virtual int kind() const = 0;
  virtual int init() = 0;
  virtual int reset() = 0;
  virtual int adaptTopology() = 0;
  virtual int advectVelocity() = 0;
  virtual int project() = 0;
  virtual int projectLocal() = 0;
  virtual int advectDensity() = 0;
  virtual int debugStats(double* out) = 0;
  virtual int seedScene() = 0;
  virtual int info(dcg_info* out) = 0;
  virtual int fieldSize(int field, uint64_t* n) = 0;
  virtual int exportField(int field, double* dst, uint64_t n) = 0;
  virtual int importField(int field, const double* src, uint64_t n) = 0;
  virtual int exportTables(dcg_tables* t) = 0;
  virtual int runStage(int stage, int arg) = 0;
  // Ray-marched image (FluidSimulation::render), RGBA float per pixel, row-major.
  virtual int render(const dcg_render_params& r, float* out) { return fail(DCG_E_UNSUPPORTED, "render"); }
  // getBlockIndexDeep (DCGrid only): finest allocated block at level >= `level` containing the
  // level-local cell (x,y,z); -1 when not applicable.
  virtual int64_t lookup(int x, int y, int z, int level) { return -1; }
  // Test hooks (DCGrid only): the scores the next adaptation starts from; replacing the block
  // tables, loads and move limits.
  virtual int scores(float* sub, float* blk) { return fail(DCG_E_UNSUPPORTED, "scores: DCGrid only"); }
  virtual int importTables(const dcg_tables* t, const uint64_t* loads, const uint64_t* moveLimits) {
    return fail(DCG_E_UNSUPPORTED, "import_tables: DCGrid only");
  }
  // Projection with the level systems solved to solve_tol (extension; DCGrid only).
  virtual int projectSolve() { return fail(DCG_E_UNSUPPORTED, "the projection solve is implemented for the DCGrid"); }
  virtual int solveReport(dcg_solve_info* out) {
    return fail(DCG_E_UNSUPPORTED, "the projection solve is implemented for the DCGrid");
  }
  int projectByMode() {
    if (P.project_mode == DCG_PROJECT_LOCAL) return projectLocal();
    if (P.project_mode == DCG_PROJECT_SOLVE) return projectSolve();
    return project();
  }
  // Differentiable rollout (extension): record n steps (on the DCGrid with the block topology
  // frozen), then pull the adjoint of the final state back to the recorded initial state. Null
  // temperature pointers mean a zero seed / no output.
  virtual int tapeForward(int n) = 0;
  virtual int tapeBackward(const double* gden, const double* gvel, const double* gtem, double* oden,
                           double* ovel, double* otem) = 0;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

}  // namespace orc
