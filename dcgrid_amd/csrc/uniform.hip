// uniform.hip — gfx950 kernels and host driver of the dense (uniform) grid back-end, the
// counterpart of FluidSimulationUniform (src/uniformgrid/fluid_simulation_uniform.cu) and its
// kernels (src/uniformgrid/uniformgrid_fluid.cu, uniformgrid_structure.cu).
//
// Layout: x-fastest SoA scalars, float3 AoS velocity (src/utils/grid_math.cuh:10); pressure,
// fluidity and divergence carry the mip pyramid (grid_math.cuh:24-52). One thread per cell,
// 256-thread workgroups over the linear index so every wavefront reads 256 contiguous bytes.
#include <algorithm>
#include <cstring>

#include "dcg_device.hpp"
#include "dcg_internal.hpp"
#include "render.hpp"

namespace dcg {

struct UGrid {
  int gx, gy, gz;
  float* den;
  float* tem;
  float3* vel;
  float* flu;  // mip
  float* pre;  // mip
  float* tpre; // mip
  float* div;  // mip
  float* tden;
  float* ttem;
  float3* tvel;
};

// Ray-march samples of the uniform grid (src/uniformgrid/uniformgrid_rendering.cu:5-47):
// coarse = the cell of the truncated position, precise = trilinear with the reference's
// edge handling (the lower corner clamped, the upper one repeated at the far faces).
struct UMarchSampler {
  UGrid g;
  __device__ void sample(const dcg_params& P, int precise, float3 pos, MarchAux& a) const;
  __device__ float3 velocity(const MarchAux& a) const { return marchInterp3(g.vel, a); }
  __device__ float fluidity(const MarchAux& a) const { return marchInterp(g.flu, a); }
};

struct UDims {
  int w, h, d;
  size_t off;
};

__device__ __forceinline__ size_t uidx(const UGrid& g, int x, int y, int z) {
  return ((size_t)z * g.gy + y) * g.gx + x;
}
__device__ void UMarchSampler::sample(const dcg_params& P, int precise, float3 pos, MarchAux& a) const {
  if (!precise) {
    const int ix = clampi((int)pos.x, 0, g.gx - 1), iy = clampi((int)pos.y, 0, g.gy - 1);
    const int iz = clampi((int)pos.z, 0, g.gz - 1);
    a.ipos = make_int3(ix, iy, iz);
    const uint32_t c = (uint32_t)uidx(g, ix, iy, iz);
#pragma unroll
    for (int k = 0; k < 8; k++) a.cell[k] = c;
    a.density = g.den[c];
    return;
  }
  const float x = pos.x - .5f, y = pos.y - .5f, z = pos.z - .5f;
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  a.ipos = make_int3((int)xf, (int)yf, (int)zf);
  a.frac = make_float3(x - xf, y - yf, z - zf);
  const int x0 = clampi(a.ipos.x, 0, g.gx - 1), y0 = clampi(a.ipos.y, 0, g.gy - 1), z0 = clampi(a.ipos.z, 0, g.gz - 1);
  const uint32_t xs = a.ipos.x < g.gx - 1 ? 1u : 0u;
  const uint32_t ys = a.ipos.y < g.gy - 1 ? (uint32_t)g.gx : 0u;
  const uint32_t zs = a.ipos.z < g.gz - 1 ? (uint32_t)g.gx * (uint32_t)g.gy : 0u;
  a.cell[0] = (uint32_t)uidx(g, x0, y0, z0);
  a.cell[1] = a.cell[0] + zs;
  a.cell[2] = a.cell[0] + ys;
  a.cell[3] = a.cell[1] + ys;
  a.cell[4] = a.cell[0] + xs;
  a.cell[5] = a.cell[4] + zs;
  a.cell[6] = a.cell[4] + ys;
  a.cell[7] = a.cell[6] + zs;
  a.density = marchInterp(g.den, a);
}

// Coordinates of linear index i of a grid w x h x ... (x fastest) in 32-bit arithmetic: grids
// have fewer than 2^32 cells (checked at creation), and a 64-bit division is a long software
// sequence on the GPU.
__device__ __forceinline__ void ucoord(size_t i, int w, int h, int& x, int& y, int& z) {
  const uint32_t ii = (uint32_t)i, uw = (uint32_t)w, uh = (uint32_t)h;
  const uint32_t q = ii / uw;
  x = (int)(ii - q * uw);
  y = (int)(q % uh);
  z = (int)(q / uh);
}

// k_uniform_set_solidity_ratio, uniformgrid_structure.cu:23-31
__global__ void __launch_bounds__(256) k_u_fluidity(dcg_params P, UGrid g, UDims L, int scale) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  g.flu[L.off + i] = cellFluidity(P, x, y, z, scale);
}

// INIT_SAMPLE (uniformgrid_fluid.cu:7-48): 8 clamped corner indices and normalised weights.
struct USample {
  size_t i[8];
  float w[8];
  int x0, y0, z0;
  bool zero;
};
template <bool FLUIDITY>
__device__ __forceinline__ void usample(const UGrid& g, float px, float py, float pz, USample& s) {
  const float x = px - .5f, y = py - .5f, z = pz - .5f;
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
  const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  const int x0c = clampi(x0, 0, g.gx - 1), y0c = clampi(y0, 0, g.gy - 1), z0c = clampi(z0, 0, g.gz - 1);
  const int x1c = clampi(x1, 0, g.gx - 1), y1c = clampi(y1, 0, g.gy - 1), z1c = clampi(z1, 0, g.gz - 1);
  s.i[0] = uidx(g, x0c, y0c, z0c);
  s.i[1] = uidx(g, x0c, y0c, z1c);
  s.i[2] = uidx(g, x0c, y1c, z0c);
  s.i[3] = uidx(g, x0c, y1c, z1c);
  s.i[4] = uidx(g, x1c, y0c, z0c);
  s.i[5] = uidx(g, x1c, y0c, z1c);
  s.i[6] = uidx(g, x1c, y1c, z0c);
  s.i[7] = uidx(g, x1c, y1c, z1c);
  const float dx = x - xf, dy = y - yf, dz = z - zf;
  const float Dx = 1.f - dx, Dy = 1.f - dy, Dz = 1.f - dz;
  float w[8];
  if (FLUIDITY) {
    w[0] = g.flu[s.i[0]] * Dx * Dy * Dz;
    w[1] = g.flu[s.i[1]] * Dx * Dy * dz;
    w[2] = g.flu[s.i[2]] * Dx * dy * Dz;
    w[3] = g.flu[s.i[3]] * Dx * dy * dz;
    w[4] = g.flu[s.i[4]] * dx * Dy * Dz;
    w[5] = g.flu[s.i[5]] * dx * Dy * dz;
    w[6] = g.flu[s.i[6]] * dx * dy * Dz;
    w[7] = g.flu[s.i[7]] * dx * dy * dz;
  } else {  // fluidity == 1 exactly: 1*a == a, so the products are the same floats
    w[0] = Dx * Dy * Dz;
    w[1] = Dx * Dy * dz;
    w[2] = Dx * dy * Dz;
    w[3] = Dx * dy * dz;
    w[4] = dx * Dy * Dz;
    w[5] = dx * Dy * dz;
    w[6] = dx * dy * Dz;
    w[7] = dx * dy * dz;
  }
  const float wAcc = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
  const float wInv = 1.f / wAcc;
#pragma unroll
  for (int c = 0; c < 8; c++) s.w[c] = w[c] * wInv;
  s.zero = wAcc < 1e-6f;
  s.x0 = x0;
  s.y0 = y0;
  s.z0 = z0;
}

// k_uniform_advect_velocity (uniformgrid_fluid.cu:50-67,88-95) + buoyancy (extension).
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_u_advect_velocity(dcg_params P, UGrid g) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
  const float3 v = g.vel[i];
  const float px = ((float)x + .5f) - v.x * P.dt * P.rdx;
  const float py = ((float)y + .5f) - v.y * P.dt * P.rdx;
  const float pz = ((float)z + .5f) - v.z * P.dt * P.rdx;
  USample s;
  usample<FLUIDITY>(g, px, py, pz, s);
  float3 acc = make_float3(0.f, 0.f, 0.f);
  if (!s.zero) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float3 vc = velocityBC(P, g.vel[s.i[c]], s.x0 + ((c >> 2) & 1), s.y0 + ((c >> 1) & 1), s.z0 + (c & 1), 1);
      acc = c == 0 ? f3mul(vc, s.w[0]) : f3add(acc, f3mul(vc, s.w[c]));
    }
  }
  if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT) {
    const float b = P.dt * (P.buoy_beta * (g.tem[i] - P.t_ambient) - P.buoy_alpha * g.den[i]);
    acc.y = acc.y + b;
  }
  g.tvel[i] = acc;
}

// k_uniform_advect_density (uniformgrid_fluid.cu:69-86,97-105) + temperature and the sphere
// source (extension).
template <bool FLUIDITY, bool SRC>
__global__ void __launch_bounds__(256) k_u_advect_density(dcg_params P, UGrid g) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
  const float3 v = g.vel[i];
  const float px = ((float)x + .5f) - v.x * P.dt * P.rdx;
  const float py = ((float)y + .5f) - v.y * P.dt * P.rdx;
  const float pz = ((float)z + .5f) - v.z * P.dt * P.rdx;
  USample s;
  usample<FLUIDITY>(g, px, py, pz, s);
  float q = 0.f, t = 0.f;
  if (!s.zero) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int cx = s.x0 + ((c >> 2) & 1), cy = s.y0 + ((c >> 1) & 1), cz = s.z0 + (c & 1);
      const float qc = densityBC(P, g.den[s.i[c]], cx, cy, cz, 1);
      q = c == 0 ? qc * s.w[0] : q + qc * s.w[c];
      if (SRC) {
        const float tc = temperatureBC(P, g.tem[s.i[c]], cx, cy, cz, 1);
        t = c == 0 ? tc * s.w[0] : t + tc * s.w[c];
      }
    }
  }
  if (SRC) {
    if (inSource(P, (float)x + .5f, (float)y + .5f, (float)z + .5f)) {
      q = P.src_density;
      t = sourceTemperature(P, x, y, z, 0);
    }
    g.ttem[i] = t;
  }
  g.tden[i] = q;
}

// k_uniform_calc_divergence, uniformgrid_fluid.cu:107-132
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_u_divergence(dcg_params P, UGrid g) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
  const size_t pl = (size_t)g.gx * g.gy;
  const size_t il = x > 0 ? i - 1 : i, ir = x < g.gx - 1 ? i + 1 : i;
  const size_t id = y > 0 ? i - g.gx : i, iu = y < g.gy - 1 ? i + g.gx : i;
  const size_t ib = z > 0 ? i - pl : i, iF = z < g.gz - 1 ? i + pl : i;
  g.pre[i] = 0.f;
  g.tpre[i] = 0.f;
  const float wl = FLUIDITY ? g.flu[il] : 1.f, wr = FLUIDITY ? g.flu[ir] : 1.f;
  const float wd = FLUIDITY ? g.flu[id] : 1.f, wu = FLUIDITY ? g.flu[iu] : 1.f;
  const float wb = FLUIDITY ? g.flu[ib] : 1.f, wf = FLUIDITY ? g.flu[iF] : 1.f;
  const float3 vl = velocityBC(P, g.vel[il], x - 1, y, z, 1);
  const float3 vr = velocityBC(P, g.vel[ir], x + 1, y, z, 1);
  const float3 vd = velocityBC(P, g.vel[id], x, y - 1, z, 1);
  const float3 vu = velocityBC(P, g.vel[iu], x, y + 1, z, 1);
  const float3 vb = velocityBC(P, g.vel[ib], x, y, z - 1, 1);
  const float3 vf = velocityBC(P, g.vel[iF], x, y, z + 1, 1);
  g.div[i] = (.5f * P.rdx) * (wr * vr.x - wl * vl.x + wu * vu.y - wd * vd.y + wf * vf.z - wb * vb.z);
}

// k_uniform_restrict, uniformgrid_fluid.cu:134-160
__device__ __forceinline__ void uRestrictCell(float* __restrict__ pre, float* __restrict__ tpre, float* __restrict__ div,
                                              const UDims& L, const UDims& Lc, size_t i) {
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  const size_t w = Lc.w, h = Lc.h;
  const size_t i000 = Lc.off + ((size_t)(2 * z) * h + 2 * y) * w + 2 * x;
  const size_t i001 = i000 + 1, i010 = i000 + w, i011 = i010 + 1;
  const size_t i100 = i000 + w * h, i101 = i100 + 1, i110 = i100 + w, i111 = i110 + 1;
  const size_t o = L.off + i;
  pre[o] = 0.f;
  tpre[o] = 0.f;
  div[o] = .125f * (div[i000] + div[i001] + div[i010] + div[i011] + div[i100] + div[i101] + div[i110] + div[i111]);
}
__global__ void __launch_bounds__(256) k_u_restrict(UGrid g, UDims L, UDims Lc) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uRestrictCell(g.pre, g.tpre, g.div, L, Lc, i);
}

// calcPressure / k_uniform_jacobi(_inv), uniformgrid_fluid.cu:162-204. The workgroup covers a
// 64x4 tile of an x-y plane so the y neighbours are in the same L1 lines.
__device__ __forceinline__ void uJacobiCell(const float* __restrict__ in, float* __restrict__ out,
                                            const float* __restrict__ div, const UDims& L, float alpha, size_t i) {
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  const size_t pl = (size_t)L.w * L.h;
  const size_t il = x > 0 ? i - 1 : i, ir = x < L.w - 1 ? i + 1 : i;
  const size_t id = y > 0 ? i - L.w : i, iu = y < L.h - 1 ? i + L.w : i;
  const size_t ib = z > 0 ? i - pl : i, iF = z < L.d - 1 ? i + pl : i;
  const float* p = in + L.off;
  out[L.off + i] = (p[il] + p[ir] + p[id] + p[iu] + p[ib] + p[iF] - alpha * div[L.off + i]) / 6.f;
}
__global__ void __launch_bounds__(256) k_u_jacobi(const float* __restrict__ in, float* __restrict__ out,
                                                  const float* __restrict__ div, UDims L, float alpha) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uJacobiCell(in, out, div, L, alpha, i);
}

// k_uniform_prolongate, uniformgrid_fluid.cu:206-237
__device__ __forceinline__ void uProlongCell(float* __restrict__ pre, const UDims& L, const UDims& Lp, size_t i) {
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  const int w = L.w, h = L.h, d = L.d;
  const size_t i000 = Lp.off + ((size_t)(z / 2) * Lp.h + (y / 2)) * Lp.w + (x / 2);
  const int sx = (x == 0 || x == w - 1) ? 0 : 2 * (x % 2) - 1;
  const int sy = (y == 0 || y == h - 1) ? 0 : 2 * (y % 2) - 1;
  const int sz = (z == 0 || z == d - 1) ? 0 : 2 * (z % 2) - 1;
  const size_t i001 = i000 + sx;
  const size_t i010 = i000 + sy * w / 2;
  const size_t i011 = i010 + sx;
  const size_t i100 = i000 + sz * (w / 2) * (h / 2);
  const size_t i101 = i100 + sx;
  const size_t i110 = i100 + sy * (w / 2);
  const size_t i111 = i110 + sx;
  pre[L.off + i] = (27.f * pre[i000] + 9.f * (pre[i001] + pre[i010] + pre[i100]) +
                    3.f * (pre[i011] + pre[i101] + pre[i110]) + pre[i111]) /
                   64.f;
}
__global__ void __launch_bounds__(256) k_u_prolongate(float* __restrict__ pre, UDims L, UDims Lp) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uProlongCell(pre, L, Lp, i);
}

// The levels of the pyramid small enough for one workgroup (<= kCoarseCells cells each),
// finest first: their part of project() (restriction into them, the top level's sweeps, and
// prolongation and sweeps back down to the first of them) in one launch. Their pressure,
// temporary pressure and divergence live in LDS for the whole sequence (at most
// kCoarseCells * 8/7 cells), phase after phase with a workgroup barrier between; each cell is
// computed by the same per-cell code as the per-level kernels, and the results are written back.
// Saves a launch per restriction, half-sweep and prolongation on levels whose work is far below
// one launch's latency.
constexpr size_t kCoarseCells = 4096;
constexpr int kCoarseLds = 4096 + 512 + 64 + 8 + 1 + 1;  // >= sum of a geometric pyramid from 4096 cells
struct ULevels {
  UDims d[DCG_MAX_LEVELS];
};
__device__ __forceinline__ size_t ucells(const UDims& D) { return (size_t)D.w * D.h * D.d; }
// Dimensions of levels Ls.. with offsets relative to level Ls (their LDS layout).
__device__ __forceinline__ void coarseLocal(const ULevels& D, int L, int Ls, ULevels& S, size_t& total) {
  total = 0;
  for (int l = Ls; l < L; l++) {
    S.d[l] = D.d[l];
    S.d[l].off = D.d[l].off - D.d[Ls].off;
    total = S.d[l].off + ucells(D.d[l]);
  }
}
__global__ void __launch_bounds__(1024) k_u_coarse(UGrid g, ULevels D, int L, int Ls, int topPairs, int levelPairs,
                                                   float dx) {
  __shared__ float sp[kCoarseLds], st[kCoarseLds], sd[kCoarseLds];
  __shared__ ULevels S;  // the level table in LDS: a per-thread copy indexed by level goes to scratch
  __shared__ size_t sTotal;
  if (threadIdx.x == 0) coarseLocal(D, L, Ls, S, sTotal);
  __syncthreads();
  const size_t total = sTotal;
  const size_t base = D.d[Ls].off;
  for (size_t i = threadIdx.x; i < total; i += blockDim.x) {
    sp[i] = g.pre[base + i];
    st[i] = g.tpre[base + i];
    sd[i] = g.div[base + i];
  }
  __syncthreads();
  for (int l = Ls + 1; l < L; l++) {
    for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uRestrictCell(sp, st, sd, S.d[l], S.d[l - 1], i);
    __syncthreads();
  }
  auto sweeps = [&](int l, int pairs) {
    const int scale = 1 << l;
    const float alpha = dx * dx * (float)scale * (float)scale;  // as jacobiPairs
    for (int p = 0; p < pairs; p++) {
      for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uJacobiCell(sp, st, sd, S.d[l], alpha, i);
      __syncthreads();
      for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uJacobiCell(st, sp, sd, S.d[l], alpha, i);
      __syncthreads();
    }
  };
  sweeps(L - 1, topPairs);
  for (int l = L - 2; l >= Ls; l--) {
    for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uProlongCell(sp, S.d[l], S.d[l + 1], i);
    __syncthreads();
    sweeps(l, levelPairs);
  }
  for (size_t i = threadIdx.x; i < total; i += blockDim.x) {
    g.pre[base + i] = sp[i];
    g.tpre[base + i] = st[i];
    g.div[base + i] = sd[i];
  }
}

// k_uniform_apply_pressure, uniformgrid_fluid.cu:239-260
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_u_apply_pressure(dcg_params P, UGrid g) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
  const size_t pl = (size_t)g.gx * g.gy;
  const size_t il = x > 0 ? i - 1 : i, ir = x < g.gx - 1 ? i + 1 : i;
  const size_t id = y > 0 ? i - g.gx : i, iu = y < g.gy - 1 ? i + g.gx : i;
  const size_t ib = z > 0 ? i - pl : i, iF = z < g.gz - 1 ? i + pl : i;
  const float alpha = .5f * P.rdx;
  const float p = g.pre[i];
  const float fl = FLUIDITY ? g.flu[il] : 1.f, fr = FLUIDITY ? g.flu[ir] : 1.f;
  const float fd = FLUIDITY ? g.flu[id] : 1.f, fu = FLUIDITY ? g.flu[iu] : 1.f;
  const float fb = FLUIDITY ? g.flu[ib] : 1.f, ff = FLUIDITY ? g.flu[iF] : 1.f;
  float3 v = g.vel[i];
  v.x = v.x - alpha * (fr * (g.pre[ir] - p) + fl * (p - g.pre[il]));
  v.y = v.y - alpha * (fu * (g.pre[iu] - p) + fd * (p - g.pre[id]));
  v.z = v.z - alpha * (ff * (g.pre[iF] - p) + fb * (p - g.pre[ib]));
  g.vel[i] = v;
}

// k_uniform_debug_stats, uniformgrid_structure.cu:33-42: one bin of 256 cells per thread.
__global__ void __launch_bounds__(256) k_u_stats(UGrid g, float* __restrict__ stats, size_t numBins) {
  const size_t b = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= numBins) return;
  float s = 0.f;
  for (size_t k = 0; k < 256; k++) s = s + g.den[b * 256 + k] * g.flu[b * 256 + k];
  stats[b] = s;
}

// Seeded initial condition (extension).
__global__ void __launch_bounds__(256) k_u_seed(dcg_params P, UGrid g) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
  const float cx = (float)x + .5f, cy = (float)y + .5f, cz = (float)z + .5f;
  g.vel[i] = initVelocity(P, cx, cy, cz);
  if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT && inSource(P, cx, cy, cz)) {
    g.den[i] = P.src_density;
    g.tem[i] = sourceTemperature(P, x, y, z, 0);
  } else {
    g.den[i] = 0.f;
    g.tem[i] = P.t_ambient;
  }
}

}  // namespace dcg

#include "uniform_adjoint.hpp"

namespace dcg {

static unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

struct UniformSim : SimBase {
  int gx = 0, gy = 0, gz = 0, L = 0;
  size_t N = 0, MC = 0;
  UGrid g{};
  std::vector<UDims> dims;  // per mip level used by the solver (and one more for allocation)
  float* stats = nullptr;
  float* hstats = nullptr;
  bool customFluidity = false;
  std::vector<char> fluKey;  // parameters the fluidity pyramid was last computed with
  uint64_t cellStepCount = 0;

  // Differentiable rollout: checkpoints of the state (density, temperature, velocity) every
  // tapeK steps plus the final state, a segment of tapeK + 1 recomputed states, and the
  // adjoint fields.
  int tapeN = 0, tapeK = 0, tapeCk = 0, tapeSeg = 0;
  bool tapeValid = false;
  float* ckBuf = nullptr;   // tapeCk states of 5N floats
  float* segBuf = nullptr;  // tapeSeg states of 5N floats
  int* winBuf = nullptr;    // window radius per segment state
  int* adjErr = nullptr;
  float *hq = nullptr, *ht = nullptr;  // ring slots of the transposed advection's tiles
  float3* hu = nullptr;
  float *lq = nullptr, *lt = nullptr, *lq0 = nullptr, *lt0 = nullptr;
  float3 *lu = nullptr, *lu0 = nullptr;
  float *lp = nullptr, *ltp = nullptr, *ld = nullptr;

  int kind() const override { return DCG_KIND_UNIFORM; }

  // FluidSimulationUniform ctor, fluid_simulation_uniform.cu:6-56
  int create(const int32_t size[3], const dcg_params& p) {
    P = p;
    gx = size[0];
    gy = size[1];
    gz = size[2];
    if (gx <= 0 || gy <= 0 || gz <= 0 || gx != P.gx || gy != P.gy || gz != P.gz)
      return fail(DCG_E_ARG, "size must be positive and equal params gx,gy,gz");
    if (P.moist) return fail(DCG_E_UNSUPPORTED, "the moist atmosphere is implemented for the DCGrid");
    const int minDim = std::min(gx, std::min(gy, gz));
    int cellSize = 2;
    L = 1;
    while (gx % cellSize == 0 && gy % cellSize == 0 && gz % cellSize == 0 && cellSize * 4 <= minDim) {
      L++;
      cellSize *= 2;
    }
    N = (size_t)gx * gy * gz;
    if (N >= 0xFFFFFFFFull) return fail(DCG_E_UNSUPPORTED, "the uniform grid is limited to 2^32 - 1 cells");
    MC = 0;
    for (size_t s = 1; gx % s == 0 && gy % s == 0 && gz % s == 0; s *= 2) MC += N / (s * s * s);
    size_t off = 0;
    for (int l = 0; l < L; l++) {
      dims.push_back(UDims{gx >> l, gy >> l, gz >> l, off});
      off += N >> (3 * l);
    }
    DCG_TRY(createStream());
    g.gx = gx;
    g.gy = gy;
    g.gz = gz;
    DCG_TRY(alloc(&g.den, N));
    DCG_TRY(alloc(&g.tem, N));
    DCG_TRY(alloc(&g.tden, N));
    DCG_TRY(alloc(&g.ttem, N));
    DCG_TRY(alloc(&g.vel, N));
    DCG_TRY(alloc(&g.tvel, N));
    DCG_TRY(alloc(&g.flu, MC));
    DCG_TRY(alloc(&g.pre, MC));
    DCG_TRY(alloc(&g.tpre, MC));
    DCG_TRY(alloc(&g.div, MC));
    const size_t nb = N / 256;
    DCG_TRY(alloc(&stats, nb > 0 ? nb : 1));
    DCG_TRY(fill32(g.div, 0, MC));
    DCG_TRY(fill32(g.tpre, 0, MC));
    return reset();
  }

  bool fluidityOn() const { return P.enable_additional_solids || customFluidity; }

  // reset(), fluid_simulation_uniform.cu:81-88
  int reset() override {
    DCG_TRY(fill32(g.den, 0, N));
    DCG_TRY(fill32(g.tem, 0, N));
    DCG_TRY(fill32(g.vel, 0, 3 * N));
    DCG_TRY(fill32(g.pre, 0, MC));
    DCG_TRY(fill32(g.flu, 0, MC));
    customFluidity = false;
    fluKey.clear();
    dropGraph();  // a captured step may rely on the fluidity computed before this reset
    cellStepCount = 0;
    return init();
  }
  // init(), fluid_simulation_uniform.cu:76-79
  int init() override {
    DCG_TRY(adaptTopology());
    DCG_TRY(fill32(g.den, 0, N));
    DCG_TRY(fill32(g.vel, 0, 3 * N));
    uint32_t ta;
    std::memcpy(&ta, &P.t_ambient, 4);
    DCG_TRY(fill32(g.tem, ta, N));
    return DCG_OK;
  }
  // adaptTopology(): fluidity per level, fluid_simulation_uniform.cu:143-147. The pyramid is a
  // function of the parameters alone, so it is recomputed only when they changed since its last
  // computation, after reset, or when a fluidity field was imported (which this call overwrites,
  // as the reference does); otherwise the array already holds exactly these values.
  int adaptTopology() override {
    const char* pb = (const char*)&P;
    if (customFluidity || fluKey.size() != sizeof(P) || std::memcmp(fluKey.data(), pb, sizeof(P)) != 0) {
      for (int l = 0; l < L; l++) {
        const UDims& D = dims[l];
        const size_t n = (size_t)D.w * D.h * D.d;
        DCG_LAUNCH(DCG_K_OTHER, "k_u_fluidity", (k_u_fluidity<<<nblk(n), 256, 0, stream>>>(P, g, D, 1 << l)));
      }
      fluKey.assign(pb, pb + sizeof(P));
    }
    customFluidity = false;
    return DCG_OK;
  }
  int advectVelocity() override {
    if (fluidityOn())
      DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_u_advect_velocity", (k_u_advect_velocity<true><<<nblk(N), 256, 0, stream>>>(P, g)));
    else
      DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_u_advect_velocity", (k_u_advect_velocity<false><<<nblk(N), 256, 0, stream>>>(P, g)));
    std::swap(g.vel, g.tvel);  // replaces the D2D copy of uniform.cu:91-92
    return DCG_OK;
  }
  int advectDensity() override {
    const bool src = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
    const bool fl = fluidityOn();
#define U_ADV_D(F, S) DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_u_advect_density", (k_u_advect_density<F, S><<<nblk(N), 256, 0, stream>>>(P, g)))
    if (fl && src) U_ADV_D(true, true);
    else if (fl) U_ADV_D(true, false);
    else if (src) U_ADV_D(false, true);
    else U_ADV_D(false, false);
#undef U_ADV_D
    std::swap(g.den, g.tden);
    if (src) std::swap(g.tem, g.ttem);
    return DCG_OK;
  }
  int divergence() {
    if (fluidityOn())
      DCG_LAUNCH(DCG_K_DIVERGENCE, "k_u_divergence", (k_u_divergence<true><<<nblk(N), 256, 0, stream>>>(P, g)));
    else
      DCG_LAUNCH(DCG_K_DIVERGENCE, "k_u_divergence", (k_u_divergence<false><<<nblk(N), 256, 0, stream>>>(P, g)));
    return DCG_OK;
  }
  int jacobiPairs(int level, int pairs) {
    const UDims& D = dims[level];
    const size_t n = (size_t)D.w * D.h * D.d;
    const int scale = 1 << level;
    // alpha = dx*dx*scale*scale, uniformgrid_fluid.cu:171
    const float alpha = P.dx * P.dx * (float)scale * (float)scale;
    if (pairs <= 0) return DCG_OK;
    Pending batch;  // the run of back-to-back sweeps is timed as one batch
    DCG_TRY(beginTimed(DCG_K_JACOBI, &batch));
    batchK = DCG_K_JACOBI;
    for (int i = 0; i < pairs; i++) {
      DCG_LAUNCH(DCG_K_JACOBI, "k_u_jacobi", (k_u_jacobi<<<nblk(n), 256, 0, stream>>>(g.pre, g.tpre, g.div, D, alpha)));
      DCG_LAUNCH(DCG_K_JACOBI, "k_u_jacobi", (k_u_jacobi<<<nblk(n), 256, 0, stream>>>(g.tpre, g.pre, g.div, D, alpha)));
    }
    batchK = -1;
    return endTimed(DCG_K_JACOBI, &batch, 2 * pairs);
  }
  int countStep() override { return DCG_OK; }
  void countHost(uint64_t steps) override { cellStepCount += steps * N; }
  bool graphCapable() const override { return true; }
  void graphKeyBytes(std::vector<char>& out) const override {
    out.clear();
    auto put = [&](const void* p, size_t n) { out.insert(out.end(), (const char*)p, (const char*)p + n); };
    put(&g, sizeof(g));
    put(&P, sizeof(P));
    put(&customFluidity, sizeof(customFluidity));
  }
  bool fluiditySettled() const {
    return !customFluidity && fluKey.size() == sizeof(P) && std::memcmp(fluKey.data(), &P, sizeof(P)) == 0;
  }
  bool captureUnsettled() const override { return !fluiditySettled(); }
  void hostSnapshot(std::vector<char>& out) const override {
    out.assign((const char*)&g, (const char*)&g + sizeof(g));
    out.push_back(customFluidity ? 1 : 0);
    out.insert(out.end(), fluKey.begin(), fluKey.end());
  }
  void hostRestore(const std::vector<char>& in) override {
    std::memcpy(&g, in.data(), sizeof(g));
    customFluidity = in[sizeof(g)] != 0;
    fluKey.assign(in.begin() + sizeof(g) + 1, in.end());
  }

  int cellSteps(uint64_t* out) override {
    *out = cellStepCount;
    return DCG_OK;
  }
  int applyPressure() {
    if (fluidityOn())
      DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_u_apply_pressure", (k_u_apply_pressure<true><<<nblk(N), 256, 0, stream>>>(P, g)));
    else
      DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_u_apply_pressure", (k_u_apply_pressure<false><<<nblk(N), 256, 0, stream>>>(P, g)));
    return DCG_OK;
  }
  // project(), fluid_simulation_uniform.cu:96-124
  // First level from which every coarser level has at most kCoarseCells cells (L if none).
  // (levels stored contiguously from Ls on, in total at most kCoarseLds cells: the LDS arrays)
  int coarseStart() const {
    int Ls = L;
    size_t total = 0;
    while (Ls > 0) {
      const size_t n = (size_t)dims[Ls - 1].w * dims[Ls - 1].h * dims[Ls - 1].d;
      if (n > kCoarseCells || total + n > (size_t)kCoarseLds) break;
      total += n;
      Ls--;
    }
    return Ls;
  }
  ULevels levels() const {
    ULevels u{};
    for (int l = 0; l < L && l < DCG_MAX_LEVELS; l++) u.d[l] = dims[l];
    return u;
  }
  // project(), fluid_simulation_uniform.cu:96-124; the levels from coarseStart() up run in one
  // workgroup (k_u_coarse), the others level by level.
  int project() override {
    DCG_TRY(divergence());
    const int Ls = coarseStart();
    for (int l = 1; l < L && l <= Ls; l++) {
      const UDims& D = dims[l];
      DCG_LAUNCH(DCG_K_RESTRICT, "k_u_restrict", (k_u_restrict<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(g, D, dims[l - 1])));
    }
    if (Ls < L) {
      DCG_LAUNCH(DCG_K_JACOBI, "k_u_coarse",
                 (k_u_coarse<<<1, 1024, 0, stream>>>(g, levels(), L, Ls, P.project_top_pairs, P.project_level_pairs, P.dx)));
    } else {
      DCG_TRY(jacobiPairs(L - 1, P.project_top_pairs));
    }
    for (int l = std::min(Ls, L - 1) - 1; l >= 0; l--) {
      const UDims& D = dims[l];
      DCG_LAUNCH(DCG_K_PROLONGATE, "k_u_prolongate", (k_u_prolongate<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(g.pre, D, dims[l + 1])));
      DCG_TRY(jacobiPairs(l, P.project_level_pairs));
    }
    return applyPressure();
  }
  // projectLocal(), fluid_simulation_uniform.cu:126-135
  int projectLocal() override {
    DCG_TRY(divergence());
    DCG_TRY(jacobiPairs(0, P.project_local_pairs));
    return applyPressure();
  }
  // debugStats(), fluid_simulation_uniform.cu:160-175
  int debugStats(double* out) override {
    const size_t nb = N / 256;
    float sum = 0.f;
    if (nb > 0) {
      DCG_LAUNCH(DCG_K_OTHER, "k_u_stats", (k_u_stats<<<nblk(nb), 256, 0, stream>>>(g, stats, nb)));
      std::vector<float> h(nb);
      DCG_HIP(hipMemcpyAsync(h.data(), stats, nb * sizeof(float), hipMemcpyDeviceToHost, stream));
      DCG_TRY(sync());
      for (size_t b = 0; b < nb; b++) sum = sum + h[b];
    }
    *out = (double)sum;
    return DCG_OK;
  }
  int seedScene() override {
    DCG_LAUNCH(DCG_K_OTHER, "k_u_seed", (k_u_seed<<<nblk(N), 256, 0, stream>>>(P, g)));
    return DCG_OK;
  }
  int info(dcg_info* o) override {
    std::memset(o, 0, sizeof(*o));
    o->kind = DCG_KIND_UNIFORM;
    o->size[0] = gx;
    o->size[1] = gy;
    o->size[2] = gz;
    o->levels = L;
    o->num_cells = N;
    o->active_cells = N;
    return DCG_OK;
  }
  int fieldSize(int f, uint64_t* n) override {
    switch (f) {
      case DCG_FIELD_DENSITY:
      case DCG_FIELD_TEMPERATURE: *n = N; return DCG_OK;
      case DCG_FIELD_VELOCITY: *n = 3 * N; return DCG_OK;
      case DCG_FIELD_PRESSURE:
      case DCG_FIELD_T_PRESSURE:
      case DCG_FIELD_DIVERGENCE:
      case DCG_FIELD_FLUIDITY: *n = MC; return DCG_OK;
      default: return fail(DCG_E_ARG, "field not available on the uniform grid");
    }
  }
  int render(const dcg_render_params& r, float* rgba) override {
    const dim3 grid((unsigned)((r.width + 15) / 16), (unsigned)((r.height + 15) / 16));
    UMarchSampler smp{g};
    DCG_LAUNCH(DCG_K_OTHER, "k_render", (k_render<<<grid, 256, 0, stream>>>(smp, P, r, (float4*)rgba)));
    return DCG_OK;
  }

  // State: the exported fields (pressure, fluidity and divergence with their mip pyramids).
  void stateSections(std::vector<StateSection>& out) override {
    out.clear();
    for (int f = 0; f < DCG_NUM_FIELDS; f++) {
      uint64_t n;
      if (!fieldPtr(f) || fieldSize(f, &n) != DCG_OK) continue;
      out.push_back({(uint32_t)(DCG_STATE_FIELD + f), 4, fieldPtr(f), n * 4});
    }
  }
  uint32_t stateFlags() const override { return customFluidity ? 1u : 0u; }
  void afterStateLoad(uint32_t flags) override {
    customFluidity = (flags & 1u) != 0;
    fluKey.clear();  // the pyramid is recomputed from the loaded parameters at the next adaptation
    tapeValid = false;
    dropGraph();
  }
  float* fieldPtr(int f) {
    switch (f) {
      case DCG_FIELD_DENSITY: return g.den;
      case DCG_FIELD_TEMPERATURE: return g.tem;
      case DCG_FIELD_VELOCITY: return (float*)g.vel;
      case DCG_FIELD_PRESSURE: return g.pre;
      case DCG_FIELD_T_PRESSURE: return g.tpre;
      case DCG_FIELD_DIVERGENCE: return g.div;
      case DCG_FIELD_FLUIDITY: return g.flu;
      default: return nullptr;
    }
  }
  int exportField(int f, float* dst, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    DCG_HIP(hipMemcpyAsync(dst, fieldPtr(f), n * sizeof(float), hipMemcpyDeviceToHost, stream));
    return sync();
  }
  int importField(int f, const float* src, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need) return fail(DCG_E_ARG, "size mismatch");
    DCG_HIP(hipMemcpyAsync(fieldPtr(f), src, n * sizeof(float), hipMemcpyHostToDevice, stream));
    if (f == DCG_FIELD_FLUIDITY) customFluidity = true;
    return sync();
  }
  int exportTables(dcg_tables*) override { return fail(DCG_E_ARG, "uniform grid has no block tables"); }
  int exportFieldDevice(int f, float* dst, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need || !dst) return fail(DCG_E_ARG, "size mismatch");
    DCG_HIP(hipMemcpyAsync(dst, fieldPtr(f), n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    return sync();
  }
  int importFieldDevice(int f, const float* src, uint64_t n) override {
    uint64_t need;
    DCG_TRY(fieldSize(f, &need));
    if (n != need || !src) return fail(DCG_E_ARG, "size mismatch");
    DCG_HIP(hipMemcpyAsync(fieldPtr(f), src, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    if (f == DCG_FIELD_FLUIDITY) customFluidity = true;
    return sync();
  }

  // ---------------------------------------------------------------- differentiable rollout
  float* stateDen(float* base, int k) { return base + (size_t)k * 5 * N; }
  float* stateTem(float* base, int k) { return base + (size_t)k * 5 * N + N; }
  float3* stateVel(float* base, int k) { return (float3*)(base + (size_t)k * 5 * N + 2 * N); }
  // Temperature changes only with the sphere source (it is advected then); otherwise it is the
  // same in every state and is neither saved nor restored.
  int saveState(float* base, int k) {
    DCG_HIP(hipMemcpyAsync(stateDen(base, k), g.den, N * 4, hipMemcpyDeviceToDevice, stream));
    if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT)
      DCG_HIP(hipMemcpyAsync(stateTem(base, k), g.tem, N * 4, hipMemcpyDeviceToDevice, stream));
    DCG_HIP(hipMemcpyAsync(stateVel(base, k), g.vel, N * 12, hipMemcpyDeviceToDevice, stream));
    return DCG_OK;
  }
  int loadState(float* base, int k) {
    DCG_HIP(hipMemcpyAsync(g.den, stateDen(base, k), N * 4, hipMemcpyDeviceToDevice, stream));
    if (P.source_mode == DCG_SOURCE_SPHERE_BUOYANT)
      DCG_HIP(hipMemcpyAsync(g.tem, stateTem(base, k), N * 4, hipMemcpyDeviceToDevice, stream));
    DCG_HIP(hipMemcpyAsync(g.vel, stateVel(base, k), N * 12, hipMemcpyDeviceToDevice, stream));
    return DCG_OK;
  }
  // One step of the differentiable rollout: the dcg_step order without the cell-step count.
  int plainStep() {
    DCG_TRY(advectVelocity());
    DCG_TRY(adaptTopology());
    DCG_TRY(P.project_mode == DCG_PROJECT_LOCAL ? projectLocal() : project());
    return advectDensity();
  }
  int tapeForward(int steps, int K) override {
    if (steps < 1 || K < 1) return fail(DCG_E_ARG, "steps and checkpoint interval must be positive");
    if (P.project_mode == DCG_PROJECT_SOLVE) return fail(DCG_E_UNSUPPORTED, "the adjoint follows the Jacobi schedules");
    if (K > steps) K = steps;
    const int ck = (steps + K - 1) / K + 1;  // states 0, K, 2K, ... below steps, and the final one
    if (ck != tapeCk || K + 1 != tapeSeg) {
      release(ckBuf);
      release(segBuf);
      release(winBuf);
      ckBuf = segBuf = nullptr;
      winBuf = nullptr;
      tapeCk = tapeSeg = 0;
      DCG_TRY(alloc(&ckBuf, (size_t)ck * 5 * N));
      DCG_TRY(alloc(&segBuf, (size_t)(K + 1) * 5 * N));
      DCG_TRY(alloc(&winBuf, (size_t)(K + 1)));
      tapeCk = ck;
      tapeSeg = K + 1;
    }
    if (!lq) {
      DCG_TRY(alloc(&adjErr, 1));
      DCG_TRY(alloc(&lq, N));
      DCG_TRY(alloc(&lt, N));
      DCG_TRY(alloc(&lq0, N));
      DCG_TRY(alloc(&lt0, N));
      DCG_TRY(alloc(&lu, N));
      DCG_TRY(alloc(&lu0, N));
      DCG_TRY(alloc(&lp, MC));
      DCG_TRY(alloc(&ltp, MC));
      DCG_TRY(alloc(&ld, MC));
      const size_t ring = (size_t)adjTiles(g.gx, g.gy, g.gz) * ADJ_SRC;
      DCG_TRY(alloc(&hq, ring));
      DCG_TRY(alloc(&ht, ring));
      DCG_TRY(alloc(&hu, ring));
    }
    tapeValid = false;
    DCG_TRY(saveState(ckBuf, 0));
    for (int k = 0; k < steps; k++) {
      DCG_TRY(plainStep());
      if ((k + 1) % K == 0 && k + 1 < steps) DCG_TRY(saveState(ckBuf, (k + 1) / K));
    }
    DCG_TRY(saveState(ckBuf, ck - 1));
    tapeN = steps;
    tapeK = K;
    tapeValid = true;
    return DCG_OK;
  }
  int window(int slot) {
    DCG_TRY(fill32(winBuf + slot, 1, 1));  // every state's radius is at least 1
    DCG_LAUNCH(DCG_K_OTHER, "k_ua_window", (k_ua_window<<<nblk(N), 256, 0, stream>>>(P, g, stateVel(segBuf, slot), winBuf + slot)));
    return DCG_OK;
  }
  // Adjoint of one step between segment states s (before) and s + 1 (after).
  int adjointStep(int s) {
    const bool src = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
    const bool fl = fluidityOn();
    const float* den0 = stateDen(segBuf, s);
    const float* tem0 = stateTem(segBuf, s);
    const float3* vel0 = stateVel(segBuf, s);
    const float3* vel1 = stateVel(segBuf, s + 1);
    // advectDensity: lq -> lq0 (and lt -> lt0 with the source), foot terms into lu
#define UA_ADV_D(F, S) DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_ua_advect_density", (k_ua_advect_density<F, S><<<adjTiles(g.gx, g.gy, g.gz), 256, 0, stream>>>(P, g, den0, tem0, vel1, lq, lt, lq0, lt0, lu, hq, ht, winBuf + s + 1, adjErr)))
    if (fl && src) UA_ADV_D(true, true);
    else if (fl) UA_ADV_D(true, false);
    else if (src) UA_ADV_D(false, true);
    else UA_ADV_D(false, false);
#undef UA_ADV_D
    DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_ua_halo_gather", (k_ua_halo_gather<float><<<nblk(N), 256, 0, stream>>>(g, lq0, hq)));
    if (src)
      DCG_LAUNCH(DCG_K_ADVECT_DENSITY, "k_ua_halo_gather", (k_ua_halo_gather<float><<<nblk(N), 256, 0, stream>>>(g, lt0, ht)));
    // projection: lu (after) -> lu (before)
    if (fl)
      DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_ua_grad_T", (k_ua_grad_T<true><<<nblk(N), 256, 0, stream>>>(P, g, lu, lp)));
    else
      DCG_LAUNCH(DCG_K_APPLY_PRESSURE, "k_ua_grad_T", (k_ua_grad_T<false><<<nblk(N), 256, 0, stream>>>(P, g, lu, lp)));
    DCG_TRY(fill32(ld, 0, MC));
    auto pairsT = [&](int level, int pairs) -> int {
      const UDims& D = dims[level];
      const size_t n = (size_t)D.w * D.h * D.d;
      const int scale = 1 << level;
      const float alpha = P.dx * P.dx * (float)scale * (float)scale;
      for (int i = 0; i < pairs; i++) {
        DCG_LAUNCH(DCG_K_JACOBI, "k_ua_jacobi", (k_ua_jacobi<<<nblk(n), 256, 0, stream>>>(lp, ltp, ld, D, alpha)));
        DCG_LAUNCH(DCG_K_JACOBI, "k_ua_jacobi", (k_ua_jacobi<<<nblk(n), 256, 0, stream>>>(ltp, lp, ld, D, alpha)));
      }
      return DCG_OK;
    };
    if (P.project_mode == DCG_PROJECT_LOCAL) {
      DCG_TRY(pairsT(0, P.project_local_pairs));
    } else {
      const int Ls = coarseStart();  // levels from Ls up: one workgroup (k_ua_coarse)
      for (int l = 0; l <= L - 2 && l < Ls; l++) {
        DCG_TRY(pairsT(l, P.project_level_pairs));
        const UDims& C = dims[l + 1];
        DCG_LAUNCH(DCG_K_PROLONGATE, "k_ua_prolongate_T", (k_ua_prolongate_T<<<prolTTiles(C), 256, 0, stream>>>(lp, dims[l], C)));
      }
      if (Ls < L)
        DCG_LAUNCH(DCG_K_JACOBI, "k_ua_coarse",
                   (k_ua_coarse<<<1, 1024, 0, stream>>>(lp, ltp, ld, levels(), L, Ls, P.project_top_pairs, P.project_level_pairs, P.dx)));
      else
        DCG_TRY(pairsT(L - 1, P.project_top_pairs));
      for (int l = std::min(Ls, L - 1); l >= 1; l--) {
        const UDims& D = dims[l - 1];
        DCG_LAUNCH(DCG_K_RESTRICT, "k_ua_restrict_T", (k_ua_restrict_T<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(ld, D, dims[l])));
      }
    }
    if (fl)
      DCG_LAUNCH(DCG_K_DIVERGENCE, "k_ua_div_T", (k_ua_div_T<true><<<nblk(N), 256, 0, stream>>>(P, g, ld, lu)));
    else
      DCG_LAUNCH(DCG_K_DIVERGENCE, "k_ua_div_T", (k_ua_div_T<false><<<nblk(N), 256, 0, stream>>>(P, g, ld, lu)));
    // advectVelocity: lu -> lu0, buoyancy into lq0 / lt0
    float* ltOut = src ? lt0 : lt;  // without the source temperature is not advected
#define UA_ADV_V(F, B) DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_ua_advect_velocity", (k_ua_advect_velocity<F, B><<<adjTiles(g.gx, g.gy, g.gz), 256, 0, stream>>>(P, g, vel0, lu, lu0, lq0, ltOut, hu, winBuf + s, adjErr)))
    if (fl && src) UA_ADV_V(true, true);
    else if (fl) UA_ADV_V(true, false);
    else if (src) UA_ADV_V(false, true);
    else UA_ADV_V(false, false);
#undef UA_ADV_V
    DCG_LAUNCH(DCG_K_ADVECT_VELOCITY, "k_ua_halo_gather", (k_ua_halo_gather<float3><<<nblk(N), 256, 0, stream>>>(g, lu0, hu)));
    std::swap(lq, lq0);
    std::swap(lu, lu0);
    if (src) std::swap(lt, lt0);
    return DCG_OK;
  }
  int tapeBackward(const float* gden, const float* gvel, const float* gtem, float* oden, float* ovel,
                   float* otem) override {
    if (!tapeValid) return fail(DCG_E_ARG, "no recorded rollout (dcg_tape_forward)");
    DCG_HIP(hipMemcpyAsync(lq, gden, N * 4, hipMemcpyDeviceToDevice, stream));
    DCG_HIP(hipMemcpyAsync(lu, gvel, N * 12, hipMemcpyDeviceToDevice, stream));
    if (gtem)
      DCG_HIP(hipMemcpyAsync(lt, gtem, N * 4, hipMemcpyDeviceToDevice, stream));
    else
      DCG_TRY(fill32(lt, 0, N));
    DCG_TRY(fill32(adjErr, 0, 1));
    const int segs = (tapeN + tapeK - 1) / tapeK;
    for (int sg = segs - 1; sg >= 0; sg--) {
      const int k0 = sg * tapeK, k1 = std::min(tapeN, k0 + tapeK);
      DCG_TRY(loadState(ckBuf, sg));
      DCG_TRY(saveState(segBuf, 0));
      DCG_TRY(window(0));
      for (int k = k0; k < k1; k++) {
        if (k + 1 == k1) {  // the segment's last state is the next checkpoint: no step to re-run
          DCG_HIP(hipMemcpyAsync(stateDen(segBuf, k - k0 + 1), stateDen(ckBuf, sg + 1), (size_t)5 * N * 4,
                                 hipMemcpyDeviceToDevice, stream));
        } else {
          DCG_TRY(plainStep());
          DCG_TRY(saveState(segBuf, k - k0 + 1));
        }
        DCG_TRY(window(k - k0 + 1));
      }
      for (int k = k1 - 1; k >= k0; k--) DCG_TRY(adjointStep(k - k0));
    }
    DCG_HIP(hipMemcpyAsync(oden, lq, N * 4, hipMemcpyDeviceToDevice, stream));
    DCG_HIP(hipMemcpyAsync(ovel, lu, N * 12, hipMemcpyDeviceToDevice, stream));
    if (otem) DCG_HIP(hipMemcpyAsync(otem, lt, N * 4, hipMemcpyDeviceToDevice, stream));
    DCG_TRY(loadState(ckBuf, tapeCk - 1));  // the live state is the rollout's final state again
    int herr = 0;
    DCG_HIP(hipMemcpyAsync(&herr, adjErr, 4, hipMemcpyDeviceToHost, stream));
    DCG_TRY(sync());
    if (herr) return fail(DCG_E_UNSUPPORTED, "a foot travels more than 4 cells in one step; the adjoint gather supports at most 4");
    return DCG_OK;
  }

  int runStage(int stage, int arg) override {
    const bool perLevel = stage == DCG_STAGE_JACOBI || stage == DCG_STAGE_JACOBI_INV || stage == DCG_STAGE_PROLONGATE;
    if (perLevel && (arg < 0 || arg >= L || (stage == DCG_STAGE_PROLONGATE && arg >= L - 1)))
      return fail(DCG_E_ARG, "level out of range");
    if (stage == DCG_STAGE_DIVERGENCE) return divergence();
    if (stage == DCG_STAGE_RESTRICT) {
      for (int l = 1; l < L; l++) {
        const UDims& D = dims[l];
        DCG_LAUNCH(DCG_K_RESTRICT, "k_u_restrict", (k_u_restrict<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(g, D, dims[l - 1])));
      }
      return DCG_OK;
    }
    if (stage == DCG_STAGE_JACOBI || stage == DCG_STAGE_JACOBI_INV) {
      const UDims& D = dims[arg];
      const int scale = 1 << arg;
      const float alpha = P.dx * P.dx * (float)scale * (float)scale;
      const bool fwd = stage == DCG_STAGE_JACOBI;
      DCG_LAUNCH(DCG_K_JACOBI, "k_u_jacobi",
                 (k_u_jacobi<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(fwd ? g.pre : g.tpre, fwd ? g.tpre : g.pre, g.div, D, alpha)));
      return DCG_OK;
    }
    if (stage == DCG_STAGE_PROLONGATE) {
      const UDims& D = dims[arg];
      DCG_LAUNCH(DCG_K_PROLONGATE, "k_u_prolongate", (k_u_prolongate<<<nblk((size_t)D.w * D.h * D.d), 256, 0, stream>>>(g.pre, D, dims[arg + 1])));
      return DCG_OK;
    }
    if (stage == DCG_STAGE_APPLY_PRESSURE) return applyPressure();
    return fail(DCG_E_ARG, "stage not available on the uniform grid");
  }
};

SimBase* makeUniform(const int32_t size[3], const dcg_params& p, int* status, std::string* err) {
  auto* s = new UniformSim();
  int st = s->create(size, p);
  *status = st;
  if (st != DCG_OK) {
    *err = s->err;
    delete s;
    return nullptr;
  }
  return s;
}

}  // namespace dcg
