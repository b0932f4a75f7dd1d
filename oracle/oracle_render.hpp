// oracle_render.hpp — CPU restatement of the reference's ray marcher (test infrastructure
// only): src/raymarching.cuh:18-275 with the samplers of src/dcgrid/dcgrid_rendering.cu:6-64
// and src/uniformgrid/uniformgrid_rendering.cu:5-47, and the pixel loop of k_dcgrid_render
// (src/dcgrid/dcgrid_rendering.cu:66-92). FP32, every expression in the reference's order, with
// the deliberate differences the library documents (normalize divides by the length; the
// Beer-Lambert exp is expDet, 0 below -87; precise apron coordinates are clamped into the
// block's apron; terrain solids have half the height above the terrain as scene distance).
#pragma once

#include <float.h>

#include <cmath>

#include "oracle_common.hpp"

namespace orc {

struct F3 {
  float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len3(F3 a) { return sqrtf(dot3(a, a)); }
inline F3 norm3(F3 a) {
  const float l = len3(a);
  return f3(a.x / l, a.y / l, a.z / l);
}

// MarchAux (raymarching.cuh:18-41)
struct Aux {
  int ipos[3] = {0, 0, 0};
  float frac[3] = {0.f, 0.f, 0.f};
  size_t cell[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float density = 0.f;
};
// MarchAux::interpolate for a scalar array (values widened to float)
template <class T>
float interp(const std::vector<T>& v, const Aux& a) {
  const float dx = a.frac[0], dxInv = 1.f - dx;
  const float c00 = (float)v[a.cell[0]] * dxInv + (float)v[a.cell[4]] * dx;
  const float c01 = (float)v[a.cell[1]] * dxInv + (float)v[a.cell[5]] * dx;
  const float c10 = (float)v[a.cell[2]] * dxInv + (float)v[a.cell[6]] * dx;
  const float c11 = (float)v[a.cell[3]] * dxInv + (float)v[a.cell[7]] * dx;
  const float dy = a.frac[1], dyInv = 1.f - dy;
  const float c0 = c00 * dyInv + c10 * dy;
  const float c1 = c01 * dyInv + c11 * dy;
  return c0 * (1.f - a.frac[2]) + c1 * a.frac[2];
}
// ... for float3 values: component by component
template <class V>
F3 interp3(const std::vector<V>& v, const Aux& a) {
  F3 c[8];
  for (int k = 0; k < 8; k++) c[k] = f3((float)v[a.cell[k]].x, (float)v[a.cell[k]].y, (float)v[a.cell[k]].z);
  const float dx = a.frac[0], dxInv = 1.f - dx, dy = a.frac[1], dyInv = 1.f - dy;
  const F3 c00 = c[0] * dxInv + c[4] * dx, c01 = c[1] * dxInv + c[5] * dx;
  const F3 c10 = c[2] * dxInv + c[6] * dx, c11 = c[3] * dxInv + c[7] * dx;
  const F3 c0 = c00 * dyInv + c10 * dy, c1 = c01 * dyInv + c11 * dy;
  return c0 * (1.f - a.frac[2]) + c1 * a.frac[2];
}

inline float marchExp(float x) { return x < -87.f ? 0.f : expDet(x); }
// densityToAlpha (raymarching.cuh:43-46)
inline float densityToAlpha(const dcg_params& P, float distance, float density) {
  return 1.f - marchExp(-distance * P.dx * density);
}
// mix (src/utils/colors.cuh:5-14)
inline void mix(F3 src, float srcA, F3& dst, float& dstA) {
  if (srcA > 1e-6f) {
    const float impact = (1.f - dstA) * srcA;
    dst = f3((dst.x * dstA + src.x * impact) / (dstA + impact), (dst.y * dstA + src.y * impact) / (dstA + impact),
             (dst.z * dstA + src.z * impact) / (dstA + impact));
    dstA += impact;
  }
}
// rayDomainIntersections (raymarching.cuh:48-78)
inline bool rayDomain(const dcg_params& P, float& dNear, float& dFar, F3 o, F3 inv) {
  const float cx = P.gx * .5f, cy = P.gy * .5f, cz = P.gz * .5f;
  const float nx = inv.x * (o.x - cx), ny = inv.y * (o.y - cy), nz = inv.z * (o.z - cz);
  const float kx = fabsf(inv.x) * cx, ky = fabsf(inv.y) * cy, kz = fabsf(inv.z) * cz;
  dNear = fmaxf(fmaxf(-nx - kx, -ny - ky), -nz - kz);
  dFar = fminf(fminf(-nx + kx, -ny + ky), -nz + kz);
  return dNear < dFar && dFar > 0.f;
}
// sceneSDF, estimateNormal (src/sdf.cuh:8-33)
inline float sceneSDF(const dcg_params& P, F3 p) {
  if (P.solid_mode == DCG_SOLID_TERRAIN) return .5f * (p.y - terrainHeight(P, p.x, p.z));
  const float dx = p.x - P.gx * .5f, dy = p.y - P.gy * .45f, dz = p.z - P.gz * .5f;
  return sqrtf(dx * dx + dy * dy + dz * dz) - 2000.f * P.rdx;
}
inline F3 sceneNormal(const dcg_params& P, F3 p) {
  const float e = 1e-4f;
  return norm3(f3(sceneSDF(P, f3(p.x + e, p.y, p.z)) - sceneSDF(P, f3(p.x - e, p.y, p.z)),
                  sceneSDF(P, f3(p.x, p.y + e, p.z)) - sceneSDF(P, f3(p.x, p.y - e, p.z)),
                  sceneSDF(P, f3(p.x, p.y, p.z + e)) - sceneSDF(P, f3(p.x, p.y, p.z - e))));
}

// lightmarch (raymarching.cuh:80-121)
template <class S>
float lightmarch(const S& smp, const dcg_params& P, const dcg_render_params& R, F3 origin, Aux& aux, F3 sun,
                 F3 sunInv) {
  float transport = 1.f;
  if (R.solids && P.enable_additional_solids) {
    float dScene = 0.f;
    for (int i = 0; i < 100; i++) {
      const float sdf = sceneSDF(P, origin + sun * dScene);
      transport = fminf(transport, (sdf / dScene) / .06f);
      if (transport <= 1e-2f) return 0.f;
      dScene += sdf;
    }
  }
  if (R.channel == DCG_RENDER_FLUID) {
    float dNear, dFar;
    if (rayDomain(P, dNear, dFar, origin, sunInv)) {
      F3 pos = origin;
      if (dNear > 0.f) {
        pos = pos + sun * dNear;
        dFar -= dNear;
      }
      const float stepSize = fmaxf(dFar / 100.f, 1e-1f);
      float density = 0.f;
      for (float d = stepSize; d < dFar && density < 5.f; d += stepSize) {
        pos = pos + sun * stepSize;
        smp.sample(P, R.precise, pos, aux);
        density += aux.density * P.dx * stepSize;
      }
      transport *= marchExp(-density);
    }
  }
  return transport;
}

struct RGBA {
  float r, g, b, a;
};

// raymarch (raymarching.cuh:123-275)
template <class S>
RGBA raymarch(const S& smp, const dcg_params& P, const dcg_render_params& R, F3 origin, F3 dir, F3 sun) {
  enum { NONE, SCENE, FLOOR, DEBUG };
  const float ambientInv = 1.f - R.ambient;
  const float floorIllumination = fmaxf(0.f, sun.y);
  const float sliceZ = P.gz * .5f;
  const F3 dirInv = f3(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
  const F3 sunInv = f3(1.f / sun.x, 1.f / sun.y, 1.f / sun.z);
  float dCurrent = 0.f;
  F3 color = f3(0.f, 0.f, 0.f);
  float alpha = 0.f, illumination;
  Aux aux;
  float dSolid = FLT_MAX;
  int renderable = NONE;
  if (R.channel != DCG_RENDER_FLUID && (sliceZ - origin.z) * dir.z > 0.f) {
    const float dDebug = (sliceZ - origin.z) / dir.z;
    const float dx = origin.x + dDebug * dir.x, dy = origin.y + dDebug * dir.y;
    if (dx >= 0.f && dy >= 0.f && dx < P.gx && dy < P.gy && dDebug < dSolid) {
      dSolid = dDebug;
      renderable = DEBUG;
    }
  }
  if (R.solids && (0.f - origin.y) * dir.y > 0.f) {
    const float dFloor = (0.f - origin.y) / dir.y;
    if (dFloor < dSolid) {
      dSolid = dFloor;
      renderable = FLOOR;
    }
  }
  if (R.solids && P.enable_additional_solids) {
    float sdfScene = FLT_MAX, dScene = dCurrent;
    for (int i = 0; i < 100; i++) {
      sdfScene = sceneSDF(P, origin + dir * dScene);
      if (sdfScene < 1e-1f || dScene >= dSolid) break;
      dScene += sdfScene;
    }
    if (sdfScene < 1e-1f && dScene < dSolid) {
      dSolid = dScene;
      renderable = SCENE;
    }
  }
  if (R.channel == DCG_RENDER_FLUID) {
    float dNear, dFar;
    if (rayDomain(P, dNear, dFar, origin, dirInv)) {
      dCurrent = fmaxf(dCurrent, fmaxf(dNear, 10.f));
      F3 pos = origin + dir * dCurrent, prevPos = pos;
      const float dMax = fminf(dFar, dSolid);
      while (dCurrent < dMax && alpha <= .99f) {
        prevPos = pos;
        dCurrent = fminf(fmaxf(dCurrent * 1.f, dCurrent + .5f), dMax);
        pos = origin + dir * dCurrent;
        const F3 mid = (prevPos + pos) * .5f;
        smp.sample(P, R.precise, mid, aux);
        if (aux.density > 1e-6f) {
          const float sampleAlpha = densityToAlpha(P, aux.density, len3(pos - prevPos));
          illumination = 1.f;
          if (R.shadows) illumination *= lightmarch(smp, P, R, mid, aux, sun, sunInv);
          illumination = R.ambient + illumination * ambientInv;
          mix(f3(illumination * R.smoke_color[0], illumination * R.smoke_color[1], illumination * R.smoke_color[2]),
              sampleAlpha, color, alpha);
        }
      }
    }
  }
  F3 pos = origin + dir * dSolid;
  const float backgroundAlpha = densityToAlpha(P, 1e-6f, dSolid);
  if (renderable == SCENE) {
    illumination = fmaxf(0.f, dot3(sceneNormal(P, pos), sun));
    if (R.shadows) illumination *= lightmarch(smp, P, R, pos, aux, sun, sunInv);
    illumination = R.ambient + illumination * ambientInv;
    mix(f3(illumination * R.scene_color[0], illumination * R.scene_color[1], illumination * R.scene_color[2]),
        1.f - backgroundAlpha, color, alpha);
  }
  if (renderable == FLOOR) {
    pos.y = 0.f;
    illumination = floorIllumination;
    if (R.shadows) illumination *= lightmarch(smp, P, R, pos, aux, sun, sunInv);
    illumination = R.ambient + illumination * ambientInv;
    illumination *= 1.f + .1f * fmodf(fabsf(floorf(pos.x / P.gx) + floorf(pos.z / P.gx)), 2.f);
    mix(f3(illumination * R.floor_color[0], illumination * R.floor_color[1], illumination * R.floor_color[2]),
        1.f - backgroundAlpha, color, alpha);
  }
  if (renderable == DEBUG) {
    pos.z = sliceZ;
    smp.sample(P, R.precise, pos, aux);
    if (R.channel == DCG_RENDER_DENSITY) mix(f3(aux.density * 1e2f, 0.f, 0.f), 1.f, color, alpha);
    if (R.channel == DCG_RENDER_VELOCITY) {
      const float v = len3(smp.velocity(aux)) * 1e-1f;
      mix(f3(v - 2.f, v - 1.f, v), 1.f, color, alpha);
    }
    if (R.channel == DCG_RENDER_RESOLUTION) {
      const float c = (float)((aux.ipos[0] / 4 + aux.ipos[1] / 4) % 2);
      mix(f3(c, c, c), 1.f, color, alpha);
    }
    if (R.channel == DCG_RENDER_FLUIDITY) {
      const float f = smp.fluidity(aux);
      mix(f3(f, f, f), 1.f, color, alpha);
    }
  }
  return RGBA{color.x, color.y, color.z, alpha};
}

// k_dcgrid_render / k_uniform_render pixel loop (dcgrid_rendering.cu:66-92); out: RGBA per
// pixel, row y, column x.
template <class S>
void renderImage(const S& smp, const dcg_params& P, const dcg_render_params& R, float* out) {
  const float cx = .5f * R.width, cy = .5f * R.height;
  const float s = R.fov / fmaxf(cx, cy);
  const float step = 1.f / R.aa_samples;
  const F3 origin = f3(R.eye_origin[0] * P.rdx, R.eye_origin[1] * P.rdx, R.eye_origin[2] * P.rdx);
  const F3 sun = f3(R.sun_dir[0], R.sun_dir[1], R.sun_dir[2]);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < R.height; y++)
    for (int x = 0; x < R.width; x++) {
      RGBA acc{0.f, 0.f, 0.f, 0.f};
      for (float dx = 0; dx < 1.f; dx += step)
        for (float dy = 0; dy < 1.f; dy += step) {
          const F3 view = norm3(f3((x + dx - cx) * s, (y + dy - cy) * s, 1.f));
          const F3 dir = f3(view.x, view.z * R.eye_dir[1] - view.y * R.eye_dir[2],
                            view.y * R.eye_dir[1] + view.z * R.eye_dir[2]);
          const RGBA c = raymarch(smp, P, R, origin, dir, sun);
          acc = RGBA{acc.r + c.r, acc.g + c.g, acc.b + c.b, acc.a + c.a};
        }
      const float w = step * step;
      float* o = out + 4 * ((size_t)y * R.width + x);
      o[0] = acc.r * w;
      o[1] = acc.g * w;
      o[2] = acc.b * w;
      o[3] = acc.a * w;
    }
}

}  // namespace orc
