// render.hpp — gfx950 ray marcher over the simulation fields: FluidSimulation::render
// (src/fluid_simulation.h:19) for both grids, following src/raymarching.cuh:63-275 with the
// samplers of src/dcgrid/dcgrid_rendering.cu:6-64 and src/uniformgrid/uniformgrid_rendering.cu:5-47.
//
// One thread per pixel in 16x16 tiles. Every float operation is written out in the order of the
// reference's helper_math expressions, with correctly rounded division and square root and the
// library's own deterministic exp, so the image equals the oracle's restatement bit for bit.
// Deliberate differences: normalize divides by the length (helper_math multiplies by rsqrtf); the
// exp of the Beer-Lambert terms is expDet, 0 below -87; the apron coordinates of a precise sample
// are clamped into the block's apron (the reference indexes them unchecked; they only leave it
// for positions outside the domain); with terrain solids the scene distance is half the height
// above the terrain (the reference's scene is the sphere alone).
#pragma once

#include <float.h>

#include "dcg_device.hpp"

namespace dcg {

// MarchAux (src/raymarching.cuh:18-41): the last sample's cell, trilinear fraction, the 8
// corner cells and the density there.
struct MarchAux {
  int3 ipos;
  float3 frac;
  uint32_t cell[8];
  float density;
};

template <class T>
__device__ __forceinline__ float marchInterp(const T* __restrict__ v, const MarchAux& a) {
  const float dx = a.frac.x, dxi = 1.f - dx;
  const float c00 = (float)v[a.cell[0]] * dxi + (float)v[a.cell[4]] * dx;
  const float c01 = (float)v[a.cell[1]] * dxi + (float)v[a.cell[5]] * dx;
  const float c10 = (float)v[a.cell[2]] * dxi + (float)v[a.cell[6]] * dx;
  const float c11 = (float)v[a.cell[3]] * dxi + (float)v[a.cell[7]] * dx;
  const float dy = a.frac.y, dyi = 1.f - dy;
  const float c0 = c00 * dyi + c10 * dy;
  const float c1 = c01 * dyi + c11 * dy;
  return c0 * (1.f - a.frac.z) + c1 * a.frac.z;
}
template <class V>
__device__ __forceinline__ float3 marchInterp3(const V* __restrict__ v, const MarchAux& a) {
  float3 c[8];
#pragma unroll
  for (int k = 0; k < 8; k++) c[k] = v[a.cell[k]];
  const float dx = a.frac.x, dxi = 1.f - dx, dy = a.frac.y, dyi = 1.f - dy, dz = a.frac.z, dzi = 1.f - dz;
  auto lerp3 = [](float3 p, float pw, float3 q, float qw) {
    return make_float3(p.x * pw + q.x * qw, p.y * pw + q.y * qw, p.z * pw + q.z * qw);
  };
  const float3 c00 = lerp3(c[0], dxi, c[4], dx), c01 = lerp3(c[1], dxi, c[5], dx);
  const float3 c10 = lerp3(c[2], dxi, c[6], dx), c11 = lerp3(c[3], dxi, c[7], dx);
  const float3 c0 = lerp3(c00, dyi, c10, dy), c1 = lerp3(c01, dyi, c11, dy);
  return lerp3(c0, dzi, c1, dz);
}

__device__ __forceinline__ float3 v3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 v3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 v3mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float v3len(float3 a) { return __builtin_sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 v3norm(float3 a) {
  const float l = v3len(a);
  return make_float3(a.x / l, a.y / l, a.z / l);
}
// exp of a Beer-Lambert exponent (__expf in the reference).
__device__ __forceinline__ float marchExp(float x) { return x < -87.f ? 0.f : expDet(x); }
// densityToAlpha (src/raymarching.cuh:43-46)
__device__ __forceinline__ float densityToAlpha(const dcg_params& P, float distance, float density) {
  return 1.f - marchExp(-distance * P.dx * density);
}
// mix (src/utils/colors.cuh:5-14): front-to-back compositing.
__device__ __forceinline__ void marchMix(float3 src, float srcA, float3& dst, float& dstA) {
  if (srcA > 1e-6f) {
    const float impact = (1.f - dstA) * srcA;
    const float den = dstA + impact;
    dst = make_float3((dst.x * dstA + src.x * impact) / den, (dst.y * dstA + src.y * impact) / den,
                      (dst.z * dstA + src.z * impact) / den);
    dstA = dstA + impact;
  }
}
// rayDomainIntersections (src/raymarching.cuh:48-78)
__device__ __forceinline__ bool rayDomain(const dcg_params& P, float& dNear, float& dFar, float3 o, float3 inv) {
  const float cx = (float)P.gx * .5f, cy = (float)P.gy * .5f, cz = (float)P.gz * .5f;
  const float nx = inv.x * (o.x - cx), ny = inv.y * (o.y - cy), nz = inv.z * (o.z - cz);
  const float kx = fabsf(inv.x) * cx, ky = fabsf(inv.y) * cy, kz = fabsf(inv.z) * cz;
  const float t1x = -nx - kx, t1y = -ny - ky, t1z = -nz - kz;
  const float t2x = -nx + kx, t2y = -ny + ky, t2z = -nz + kz;
  dNear = fmaxf(fmaxf(t1x, t1y), t1z);
  dFar = fminf(fminf(t2x, t2y), t2z);
  return dNear < dFar && dFar > 0.f;
}
// sceneSDF (src/sdf.cuh:8-22): the sphere of the additional solids, or half the height above
// the terrain (extension).
__device__ __forceinline__ float sceneSDF(const dcg_params& P, float3 p) {
  if (P.solid_mode == DCG_SOLID_TERRAIN) return .5f * (p.y - terrainHeight(P, p.x, p.z));
  const float cx = (float)P.gx * .5f, cy = (float)P.gy * .45f, cz = (float)P.gz * .5f;
  const float r = 2000.f * P.rdx;
  const float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
  return __builtin_sqrtf(dx * dx + dy * dy + dz * dz) - r;
}
// estimateNormal (src/sdf.cuh:24-33)
__device__ __forceinline__ float3 sceneNormal(const dcg_params& P, float3 p) {
  const float e = 1e-4f;
  return v3norm(make_float3(sceneSDF(P, make_float3(p.x + e, p.y, p.z)) - sceneSDF(P, make_float3(p.x - e, p.y, p.z)),
                            sceneSDF(P, make_float3(p.x, p.y + e, p.z)) - sceneSDF(P, make_float3(p.x, p.y - e, p.z)),
                            sceneSDF(P, make_float3(p.x, p.y, p.z + e)) - sceneSDF(P, make_float3(p.x, p.y, p.z - e))));
}

// lightmarch (src/raymarching.cuh:80-121)
template <class S>
__device__ float lightMarch(const S& smp, const dcg_params& P, const dcg_render_params& R, float3 origin, MarchAux& aux,
                            float3 sun, float3 sunInv) {
  float lt = 1.f;
  if (R.solids && P.enable_additional_solids) {
    float dScene = 0.f;
    for (int i = 0; i < 100; i++) {
      const float sdf = sceneSDF(P, v3add(origin, v3mul(sun, dScene)));
      lt = fminf(lt, (sdf / dScene) / .06f);
      if (lt <= 1e-2f) return 0.f;
      dScene = dScene + sdf;
    }
  }
  if (R.channel == DCG_RENDER_FLUID) {
    float dNear, dFar;
    if (rayDomain(P, dNear, dFar, origin, sunInv)) {
      float3 pos = origin;
      if (dNear > 0.f) {
        pos = v3add(pos, v3mul(sun, dNear));
        dFar = dFar - dNear;
      }
      const float step = fmaxf(dFar / 100.f, 1e-1f);
      float density = 0.f;
      for (float d = step; d < dFar && density < 5.f; d = d + step) {
        pos = v3add(pos, v3mul(sun, step));
        smp.sample(P, R.precise, pos, aux);
        density = density + aux.density * P.dx * step;
      }
      lt = lt * marchExp(-density);
    }
  }
  return lt;
}

// raymarch (src/raymarching.cuh:123-275); returns (color, alpha).
template <class S>
__device__ float4 rayMarch(const S& smp, const dcg_params& P, const dcg_render_params& R, float3 origin, float3 dir,
                           float3 sun) {
  enum { NONE, SCENE, FLOOR, DEBUG };
  const float ambientInv = 1.f - R.ambient;
  const float floorIllum = fmaxf(0.f, sun.y);
  const float sliceZ = (float)P.gz * .5f;
  const float3 dirInv = make_float3(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
  const float3 sunInv = make_float3(1.f / sun.x, 1.f / sun.y, 1.f / sun.z);
  float dCurrent = 0.f;
  float3 color = make_float3(0.f, 0.f, 0.f);
  float alpha = 0.f, illum;
  MarchAux aux;
  aux.ipos = make_int3(0, 0, 0);
  aux.frac = make_float3(0.f, 0.f, 0.f);
  aux.density = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) aux.cell[k] = 0;
  float dSolid = FLT_MAX;
  int what = NONE;
  if (R.channel != DCG_RENDER_FLUID && (sliceZ - origin.z) * dir.z > 0.f) {
    const float dDebug = (sliceZ - origin.z) / dir.z;
    const float dx = origin.x + dDebug * dir.x, dy = origin.y + dDebug * dir.y;
    if (dx >= 0.f && dy >= 0.f && dx < (float)P.gx && dy < (float)P.gy && dDebug < dSolid) {
      dSolid = dDebug;
      what = DEBUG;
    }
  }
  if (R.solids && (0.f - origin.y) * dir.y > 0.f) {
    const float dFloor = (0.f - origin.y) / dir.y;
    if (dFloor < dSolid) {
      dSolid = dFloor;
      what = FLOOR;
    }
  }
  if (R.solids && P.enable_additional_solids) {
    float sdf = FLT_MAX, dScene = dCurrent;
    for (int i = 0; i < 100; i++) {
      sdf = sceneSDF(P, v3add(origin, v3mul(dir, dScene)));
      if (sdf < 1e-1f || dScene >= dSolid) break;
      dScene = dScene + sdf;
    }
    if (sdf < 1e-1f && dScene < dSolid) {
      dSolid = dScene;
      what = SCENE;
    }
  }
  if (R.channel == DCG_RENDER_FLUID) {
    float dNear, dFar;
    if (rayDomain(P, dNear, dFar, origin, dirInv)) {
      dCurrent = fmaxf(dCurrent, fmaxf(dNear, 10.f));  // Z_NEAR
      float3 pos = v3add(origin, v3mul(dir, dCurrent));
      const float dMax = fminf(dFar, dSolid);
      while (dCurrent < dMax && alpha <= .99f) {
        const float3 prev = pos;
        dCurrent = fminf(fmaxf(dCurrent * 1.f, dCurrent + .5f), dMax);  // STEP_MULTIPLIER 1
        pos = v3add(origin, v3mul(dir, dCurrent));
        const float3 mid = v3mul(v3add(prev, pos), .5f);
        smp.sample(P, R.precise, mid, aux);
        if (aux.density > 1e-6f) {
          const float sa = densityToAlpha(P, aux.density, v3len(v3sub(pos, prev)));
          illum = 1.f;
          if (R.shadows) illum = illum * lightMarch(smp, P, R, mid, aux, sun, sunInv);
          illum = R.ambient + illum * ambientInv;
          marchMix(make_float3(illum * R.smoke_color[0], illum * R.smoke_color[1], illum * R.smoke_color[2]), sa, color,
                   alpha);
        }
      }
    }
  }
  float3 pos = v3add(origin, v3mul(dir, dSolid));
  const float bgAlpha = densityToAlpha(P, 1e-6f, dSolid);
  if (what == SCENE) {
    const float3 n = sceneNormal(P, pos);
    illum = fmaxf(0.f, n.x * sun.x + n.y * sun.y + n.z * sun.z);
    if (R.shadows) illum = illum * lightMarch(smp, P, R, pos, aux, sun, sunInv);
    illum = R.ambient + illum * ambientInv;
    marchMix(make_float3(illum * R.scene_color[0], illum * R.scene_color[1], illum * R.scene_color[2]), 1.f - bgAlpha,
             color, alpha);
  }
  if (what == FLOOR) {
    pos.y = 0.f;
    illum = floorIllum;
    if (R.shadows) illum = illum * lightMarch(smp, P, R, pos, aux, sun, sunInv);
    illum = R.ambient + illum * ambientInv;
    illum = illum * (1.f + .1f * fmodf(fabsf(floorf(pos.x / (float)P.gx) + floorf(pos.z / (float)P.gx)), 2.f));
    marchMix(make_float3(illum * R.floor_color[0], illum * R.floor_color[1], illum * R.floor_color[2]), 1.f - bgAlpha,
             color, alpha);
  }
  if (what == DEBUG) {
    pos.z = sliceZ;
    smp.sample(P, R.precise, pos, aux);
    if (R.channel == DCG_RENDER_DENSITY) {
      const float v = aux.density * 1e2f;
      marchMix(make_float3(v, 0.f, 0.f), 1.f, color, alpha);
    }
    if (R.channel == DCG_RENDER_VELOCITY) {
      const float v = v3len(smp.velocity(aux)) * 1e-1f;
      marchMix(make_float3(v - 2.f, v - 1.f, v), 1.f, color, alpha);
    }
    if (R.channel == DCG_RENDER_RESOLUTION) {
      const float c = (float)((aux.ipos.x / 4 + aux.ipos.y / 4) % 2);
      marchMix(make_float3(c, c, c), 1.f, color, alpha);
    }
    if (R.channel == DCG_RENDER_FLUIDITY) {
      const float f = smp.fluidity(aux);
      marchMix(make_float3(f, f, f), 1.f, color, alpha);
    }
  }
  return make_float4(color.x, color.y, color.z, alpha);
}

// k_dcgrid_render / k_uniform_render: per pixel, aa_samples^2 rays averaged.
template <class S>
__global__ void __launch_bounds__(256) k_render(S smp, dcg_params P, dcg_render_params R, float4* __restrict__ out) {
  const int x = blockIdx.x * 16 + (int)(threadIdx.x % 16);
  const int y = blockIdx.y * 16 + (int)(threadIdx.x / 16);
  if (x >= R.width || y >= R.height) return;
  const float cx = .5f * (float)R.width, cy = .5f * (float)R.height;
  const float s = R.fov / fmaxf(cx, cy);
  const float step = 1.f / R.aa_samples;
  const float3 origin = make_float3(R.eye_origin[0] * P.rdx, R.eye_origin[1] * P.rdx, R.eye_origin[2] * P.rdx);
  const float3 sun = make_float3(R.sun_dir[0], R.sun_dir[1], R.sun_dir[2]);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (float dx = 0.f; dx < 1.f; dx = dx + step)
    for (float dy = 0.f; dy < 1.f; dy = dy + step) {
      const float3 view = v3norm(make_float3(((float)x + dx - cx) * s, ((float)y + dy - cy) * s, 1.f));
      const float3 dir = make_float3(view.x, view.z * R.eye_dir[1] - view.y * R.eye_dir[2],
                                     view.y * R.eye_dir[1] + view.z * R.eye_dir[2]);
      const float4 c = rayMarch(smp, P, R, origin, dir, sun);
      acc = make_float4(acc.x + c.x, acc.y + c.y, acc.z + c.z, acc.w + c.w);
    }
  const float w = step * step;
  out[(size_t)y * R.width + x] = make_float4(acc.x * w, acc.y * w, acc.z * w, acc.w * w);
}

}  // namespace dcg
