// uniform_adjoint.hpp — gfx950 adjoint (reverse-mode) kernels of the uniform-grid step on a
// fixed fluidity field. Extension: the reference has no derivatives. Included by uniform.hip
// after the forward kernels, whose sampling (usample) and boundary conditions they reuse.
//
// The step is u* = A(u; u) (+ buoyancy), u' = u* - G P D u* (the fixed multilevel Jacobi
// schedule P is linear), q' = A(q; u'). Its adjoint:
//  * transposed advection: lam_q[j] = sum over sources i of w_c(i) lam_q'[i] where corner c of
//    the trilinear sample of i is cell j. Written as a gather per destination j over the window
//    of sources whose foot can reach j (radius R from the largest cell offset of any foot,
//    measured per state): one pass, no atomics, a fixed summation order;
//  * foot term: lam_u[i] += d(sample)/d(u(i)) . lam[i] (local to the source);
//  * projection: G^T, the Jacobi half-sweeps reversed (the clamped 7-point stencil is symmetric,
//    so each reversed sweep is the same stencil with the div term accumulated), prolongation and
//    restriction transposed, D^T.
#pragma once

namespace dcg {

constexpr int ADJ_MAX_WINDOW = 4;  // largest supported foot offset (cells) of a recorded state

__device__ __forceinline__ void ufoot(const dcg_params& P, int x, int y, int z, const float3& v, float& px, float& py,
                                      float& pz) {
  // the expressions of k_u_advect_velocity / k_u_advect_density
  px = ((float)x + .5f) - v.x * P.dt * P.rdx;
  py = ((float)y + .5f) - v.y * P.dt * P.rdx;
  pz = ((float)z + .5f) - v.z * P.dt * P.rdx;
}
__device__ __forceinline__ int ubase(float p) { return (int)floorf(p - .5f); }

// Largest |offset| (and |offset+1|) between a cell and the base corner of its foot; one int
// per state, reduced with atomicMax on a global word.
__global__ void __launch_bounds__(256) k_ua_window(dcg_params P, UGrid g, const float3* __restrict__ vel,
                                                   int* __restrict__ out) {
  __shared__ int red[256];
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  int r = 0;
  if (i < n) {
    const int x = (int)(i % g.gx), y = (int)((i / g.gx) % g.gy), z = (int)(i / ((size_t)g.gx * g.gy));
    float px, py, pz;
    ufoot(P, x, y, z, vel[i], px, py, pz);
    const int o[3] = {ubase(px) - x, ubase(py) - y, ubase(pz) - z};
#pragma unroll
    for (int a = 0; a < 3; a++) r = max(r, max(abs(o[a]), abs(o[a] + 1)));
  }
  red[threadIdx.x] = r;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// Unnormalised corner weights f_c t_c of a foot and d t_c / d d_a.
struct UFootD {
  USample s;
  float f[8], t[8], dtd[8][3], W;
};
template <bool FLUIDITY>
__device__ __forceinline__ void ufootD(const UGrid& g, float px, float py, float pz, UFootD& F) {
  usample<FLUIDITY>(g, px, py, pz, F.s);
  const float xs = px - .5f, ys = py - .5f, zs = pz - .5f;
  const float d[3] = {xs - floorf(xs), ys - floorf(ys), zs - floorf(zs)};
  F.W = 0.f;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int bit[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
    float tau[3];
#pragma unroll
    for (int a = 0; a < 3; a++) tau[a] = bit[a] ? d[a] : 1.f - d[a];
    F.t[c] = tau[0] * tau[1] * tau[2];
    F.dtd[c][0] = (bit[0] ? 1.f : -1.f) * tau[1] * tau[2];
    F.dtd[c][1] = (bit[1] ? 1.f : -1.f) * tau[0] * tau[2];
    F.dtd[c][2] = (bit[2] ? 1.f : -1.f) * tau[0] * tau[1];
    F.f[c] = FLUIDITY ? g.flu[F.s.i[c]] : 1.f;
    F.W += F.f[c] * F.t[c];
  }
}
// d(sample)/d(u_a) for BC-applied corner values vc (sample = sum f t vc / W), scaled by lam.
__device__ __forceinline__ float3 ufootGrad(const dcg_params& P, const UFootD& F, const float vc[8], float lam) {
  const float k = -P.dt * P.rdx;
  float val = 0.f;
#pragma unroll
  for (int c = 0; c < 8; c++) val += F.f[c] * F.t[c] * vc[c];
  val /= F.W;
  float g[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float num = 0.f, den = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      num += F.f[c] * vc[c] * F.dtd[c][a];
      den += F.f[c] * F.dtd[c][a];
    }
    g[a] = k * (num - val * den) / F.W * lam;
  }
  return make_float3(g[0], g[1], g[2]);
}

__device__ __forceinline__ bool uinDomain(const UGrid& g, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < g.gx && y < g.gy && z < g.gz;
}

// Weight with which source (xi,yi,zi) under velocity v samples cell (x,y,z); 0 if it does not.
template <bool FLUIDITY>
__device__ __forceinline__ float utransWeight(const dcg_params& P, const UGrid& g, int xi, int yi, int zi,
                                              const float3& v, int x, int y, int z) {
  float px, py, pz;
  ufoot(P, xi, yi, zi, v, px, py, pz);
  const int bx = x - ubase(px), by = y - ubase(py), bz = z - ubase(pz);
  if ((unsigned)bx > 1u || (unsigned)by > 1u || (unsigned)bz > 1u) return 0.f;
  USample s;
  usample<FLUIDITY>(g, px, py, pz, s);
  if (s.zero) return 0.f;
  return s.w[(bx << 2) | (by << 1) | bz];
}

__device__ __forceinline__ int windowRadius(const int* __restrict__ win, int* __restrict__ errFlag) {
  int R = *win;
  if (R > ADJ_MAX_WINDOW) {
    if (threadIdx.x == 0) atomicOr(errFlag, 1);
    R = ADJ_MAX_WINDOW;
  }
  return R;
}

// Adjoint of advectDensity under velocity vel1 = u_{k+1}: lq0 = A^T lq (and lt0 = A^T lt with
// the source), sources masked inside the emitting sphere; lu += foot terms (den0/tem0 = state k).
template <bool FLUIDITY, bool SRC>
__global__ void __launch_bounds__(256) k_ua_advect_density(dcg_params P, UGrid g, const float* __restrict__ den0,
                                                           const float* __restrict__ tem0,
                                                           const float3* __restrict__ vel1,
                                                           const float* __restrict__ lq, const float* __restrict__ lt,
                                                           float* __restrict__ lq0, float* __restrict__ lt0,
                                                           float3* __restrict__ lu, const int* __restrict__ win,
                                                           int* __restrict__ errFlag) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int R = windowRadius(win, errFlag);
  if (j >= n) return;
  const int x = (int)(j % g.gx), y = (int)((j / g.gx) % g.gy), z = (int)(j / ((size_t)g.gx * g.gy));
  // transposed gather
  float aq = 0.f, at = 0.f;
  for (int zi = max(z - R, 0); zi <= min(z + R, g.gz - 1); zi++)
    for (int yi = max(y - R, 0); yi <= min(y + R, g.gy - 1); yi++)
      for (int xi = max(x - R, 0); xi <= min(x + R, g.gx - 1); xi++) {
        const size_t i = uidx(g, xi, yi, zi);
        if (SRC && inSource(P, (float)xi + .5f, (float)yi + .5f, (float)zi + .5f)) continue;
        const float w = utransWeight<FLUIDITY>(P, g, xi, yi, zi, vel1[i], x, y, z);
        if (w != 0.f) {
          aq += w * lq[i];
          if (SRC) at += w * lt[i];
        }
      }
  lq0[j] = aq;
  if (SRC) lt0[j] = at;
  // foot term of cell j
  if (SRC && inSource(P, (float)x + .5f, (float)y + .5f, (float)z + .5f)) return;
  float px, py, pz;
  ufoot(P, x, y, z, vel1[j], px, py, pz);
  UFootD F;
  ufootD<FLUIDITY>(g, px, py, pz, F);
  if (F.s.zero) return;
  float qc[8], tc[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int cx = F.s.x0 + ((c >> 2) & 1), cy = F.s.y0 + ((c >> 1) & 1), cz = F.s.z0 + (c & 1);
    qc[c] = densityBC(P, den0[F.s.i[c]], cx, cy, cz, 1);
    if (SRC) tc[c] = temperatureBC(P, tem0[F.s.i[c]], cx, cy, cz, 1);
  }
  float3 d = ufootGrad(P, F, qc, lq[j]);
  if (SRC) d = f3add(d, ufootGrad(P, F, tc, lt[j]));
  lu[j] = f3add(lu[j], d);
}

// Adjoint of advectVelocity with velocity vel0 = u_k (self-advection) and buoyancy:
// lu0 = A^T lus + foot terms; lq0 += -dt*alpha*lus.y, lt0 += dt*beta*lus.y.
template <bool FLUIDITY, bool BUOY>
__global__ void __launch_bounds__(256) k_ua_advect_velocity(dcg_params P, UGrid g, const float3* __restrict__ vel0,
                                                            const float3* __restrict__ lus,
                                                            float3* __restrict__ lu0, float* __restrict__ lq0,
                                                            float* __restrict__ lt0, const int* __restrict__ win,
                                                            int* __restrict__ errFlag) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int R = windowRadius(win, errFlag);
  if (j >= n) return;
  const int x = (int)(j % g.gx), y = (int)((j / g.gx) % g.gy), z = (int)(j / ((size_t)g.gx * g.gy));
  float3 acc = make_float3(0.f, 0.f, 0.f);
  for (int zi = max(z - R, 0); zi <= min(z + R, g.gz - 1); zi++)
    for (int yi = max(y - R, 0); yi <= min(y + R, g.gy - 1); yi++)
      for (int xi = max(x - R, 0); xi <= min(x + R, g.gx - 1); xi++) {
        const size_t i = uidx(g, xi, yi, zi);
        const float w = utransWeight<FLUIDITY>(P, g, xi, yi, zi, vel0[i], x, y, z);
        if (w != 0.f) acc = f3add(acc, f3mul(lus[i], w));
      }
  const float3 lam = lus[j];
  float px, py, pz;
  ufoot(P, x, y, z, vel0[j], px, py, pz);
  UFootD F;
  ufootD<FLUIDITY>(g, px, py, pz, F);
  if (!F.s.zero) {
    float vx[8], vy[8], vz[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float3 v = velocityBC(P, vel0[F.s.i[c]], F.s.x0 + ((c >> 2) & 1), F.s.y0 + ((c >> 1) & 1), F.s.z0 + (c & 1), 1);
      vx[c] = v.x;
      vy[c] = v.y;
      vz[c] = v.z;
    }
    acc = f3add(acc, ufootGrad(P, F, vx, lam.x));
    acc = f3add(acc, ufootGrad(P, F, vy, lam.y));
    acc = f3add(acc, ufootGrad(P, F, vz, lam.z));
  }
  lu0[j] = acc;
  if (BUOY) {
    lq0[j] = lq0[j] + (-P.dt * P.buoy_alpha) * lam.y;
    lt0[j] = lt0[j] + (P.dt * P.buoy_beta) * lam.y;
  }
}

// G^T: lambda of the level-0 pressure from lambda of the velocity after applyPressure.
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_ua_grad_T(dcg_params P, UGrid g, const float3* __restrict__ lu,
                                                   float* __restrict__ lp) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int x = (int)(j % g.gx), y = (int)((j / g.gx) % g.gy), z = (int)(j / ((size_t)g.gx * g.gy));
  const size_t pl = (size_t)g.gx * g.gy;
  const float a = .5f * P.rdx;
  const float fj = FLUIDITY ? g.flu[j] : 1.f;
  float s = 0.f;
  // x: neighbours j-1 (its right face is j) and j+1 (its left face is j), and j itself
  if (x > 0) s -= fj * lu[j - 1].x;
  if (x < g.gx - 1) s += fj * lu[j + 1].x;
  {
    const float fr = x < g.gx - 1 ? (FLUIDITY ? g.flu[j + 1] : 1.f) : 0.f;
    const float fl = x > 0 ? (FLUIDITY ? g.flu[j - 1] : 1.f) : 0.f;
    s += (fr - fl) * lu[j].x;
  }
  if (y > 0) s -= fj * lu[j - g.gx].y;
  if (y < g.gy - 1) s += fj * lu[j + g.gx].y;
  {
    const float fu = y < g.gy - 1 ? (FLUIDITY ? g.flu[j + g.gx] : 1.f) : 0.f;
    const float fd = y > 0 ? (FLUIDITY ? g.flu[j - g.gx] : 1.f) : 0.f;
    s += (fu - fd) * lu[j].y;
  }
  if (z > 0) s -= fj * lu[j - pl].z;
  if (z < g.gz - 1) s += fj * lu[j + pl].z;
  {
    const float ff = z < g.gz - 1 ? (FLUIDITY ? g.flu[j + pl] : 1.f) : 0.f;
    const float fb = z > 0 ? (FLUIDITY ? g.flu[j - pl] : 1.f) : 0.f;
    s += (ff - fb) * lu[j].z;
  }
  lp[j] = a * s;
}

// One reversed Jacobi half-sweep on a level: out = S in (clamped 7-point sum / 6, symmetric),
// ld += -alpha/6 * in.
__global__ void __launch_bounds__(256) k_ua_jacobi(const float* __restrict__ in, float* __restrict__ out,
                                                   float* __restrict__ ld, UDims L, float alpha) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % L.w), y = (int)((i / L.w) % L.h), z = (int)(i / ((size_t)L.w * L.h));
  const size_t pl = (size_t)L.w * L.h;
  const size_t il = x > 0 ? i - 1 : i, ir = x < L.w - 1 ? i + 1 : i;
  const size_t id = y > 0 ? i - L.w : i, iu = y < L.h - 1 ? i + L.w : i;
  const size_t ib = z > 0 ? i - pl : i, iF = z < L.d - 1 ? i + pl : i;
  const float* p = in + L.off;
  out[L.off + i] = (p[il] + p[ir] + p[id] + p[iu] + p[ib] + p[iF]) / 6.f;
  ld[L.off + i] = ld[L.off + i] - alpha * (p[i] / 6.f);
}

// Weights with which fine coordinate x (of a level of width w) prolongates from coarse X:
// 3/4 from x/2, 1/4 from x/2 + s (s = 0 at the two edges, else -1 for even x, +1 for odd x).
__device__ __forceinline__ float prolW(int x, int w, int X) {
  const int s = (x == 0 || x == w - 1) ? 0 : 2 * (x & 1) - 1;
  float r = 0.f;
  if (x / 2 == X) r += .75f;
  if (x / 2 + s == X) r += .25f;
  return r;
}
// Prolongation transposed: lambda of the coarse level (Lc) from lambda of the fine level (L).
__global__ void __launch_bounds__(256) k_ua_prolongate_T(float* __restrict__ lp, UDims L, UDims Lc) {
  const size_t n = (size_t)Lc.w * Lc.h * Lc.d;
  const size_t I = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (I >= n) return;
  const int X = (int)(I % Lc.w), Y = (int)((I / Lc.w) % Lc.h), Z = (int)(I / ((size_t)Lc.w * Lc.h));
  float s = 0.f;
  for (int z = max(2 * Z - 1, 0); z <= min(2 * Z + 2, L.d - 1); z++) {
    const float wz = prolW(z, L.d, Z);
    if (wz == 0.f) continue;
    for (int y = max(2 * Y - 1, 0); y <= min(2 * Y + 2, L.h - 1); y++) {
      const float wy = prolW(y, L.h, Y);
      if (wy == 0.f) continue;
      for (int x = max(2 * X - 1, 0); x <= min(2 * X + 2, L.w - 1); x++) {
        const float wx = prolW(x, L.w, X);
        if (wx == 0.f) continue;
        s += wz * wy * wx * lp[L.off + ((size_t)z * L.h + y) * L.w + x];
      }
    }
  }
  lp[Lc.off + I] = s;
}

// Restriction transposed: every child cell of level L gains 1/8 of its parent's lambda (Lp).
__global__ void __launch_bounds__(256) k_ua_restrict_T(float* __restrict__ ld, UDims L, UDims Lp) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % L.w), y = (int)((i / L.w) % L.h), z = (int)(i / ((size_t)L.w * L.h));
  const size_t par = Lp.off + ((size_t)(z / 2) * Lp.h + (y / 2)) * Lp.w + (x / 2);
  ld[L.off + i] = ld[L.off + i] + .125f * ld[par];
}

// D^T: lu += (d div / d u)^T ld on level 0.
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_ua_div_T(dcg_params P, UGrid g, const float* __restrict__ ld,
                                                  float3* __restrict__ lu) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int x = (int)(j % g.gx), y = (int)((j / g.gx) % g.gy), z = (int)(j / ((size_t)g.gx * g.gy));
  const size_t pl = (size_t)g.gx * g.gy;
  const float af = (.5f * P.rdx) * (FLUIDITY ? g.flu[j] : 1.f);
  float3 v = lu[j];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (x > 0) sx += ld[j - 1];
  if (x < g.gx - 1) sx -= ld[j + 1];
  if (y > 0) sy += ld[j - g.gx];
  if (y < g.gy - 1) sy -= ld[j + g.gx];
  if (z > 0) sz += ld[j - pl];
  if (z < g.gz - 1) sz -= ld[j + pl];
  v.x = v.x + af * sx;
  v.y = v.y + af * sy;
  v.z = v.z + af * sz;
  lu[j] = v;
}

}  // namespace dcg
