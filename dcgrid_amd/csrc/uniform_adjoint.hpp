// uniform_adjoint.hpp — gfx950 adjoint (reverse-mode) kernels of the uniform-grid step on a
// fixed fluidity field. Extension: the reference has no derivatives. Included by uniform.hip
// after the forward kernels, whose sampling (usample) and boundary conditions they reuse.
//
// The step is u* = A(u; u) (+ buoyancy), u' = u* - G P D u* (the fixed multilevel Jacobi
// schedule P is linear), q' = A(q; u'). Its adjoint:
//  * transposed advection: lam_q[j] = sum over sources i of w_c(i) lam_q'[i] where corner c of
//    the trilinear sample of i is cell j. With every foot within one cell of its own (radius 1,
//    measured per state) each cell scatters into an LDS tile in 64-bit fixed point (integer
//    atomics, order independent), the tiles' rings are exchanged through a halo buffer and added
//    in a fixed order; wider windows gather per destination over the window of sources whose
//    foot can reach it;
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
// per state, reduced with atomicMax on a global word (skipped when the word already holds at
// least the workgroup's maximum: thousands of workgroups otherwise serialise on one address).
__global__ void __launch_bounds__(256) k_ua_window(dcg_params P, UGrid g, const float3* __restrict__ vel,
                                                   int* __restrict__ out) {
  __shared__ int red[256];
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  int r = 0;
  if (i < n) {
    int x, y, z;
  ucoord(i, g.gx, g.gy, x, y, z);
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
  if (threadIdx.x == 0 && red[0] > *(volatile const int*)out) atomicMax(out, red[0]);
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
  // the corner's weight picked with static indices (a data-dependent index into the register
  // array would place it in scratch, set up on every launch of the kernel)
  const int c = (bx << 2) | (by << 1) | bz;
  float w = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) w = k == c ? s.w[k] : w;
  return w;
}

__device__ __forceinline__ int windowRadius(const int* __restrict__ win, int* __restrict__ errFlag) {
  int R = *win;
  if (R > ADJ_MAX_WINDOW) {
    if (threadIdx.x == 0) atomicOr(errFlag, 1);
    R = ADJ_MAX_WINDOW;
  }
  return R;
}

// Tiled transposed gather. A workgroup owns a tile of ADJ_TX x ADJ_TY x ADJ_TZ destination
// cells. When every foot of the state stays within one cell of its own (window radius 1, the
// case for CFL < 1), the sources that can reach the tile are the tile grown by one cell: each is
// sampled once per workgroup into LDS (weights, base-corner offset, its lambda) and every
// destination then walks its 27 candidate sources in the same order as the untiled gather, so
// the sums are the same floats. Larger windows take the untiled per-destination gather.
constexpr int ADJ_TX = 8, ADJ_TY = 8, ADJ_TZ = 4;  // 256 destinations per workgroup
constexpr int ADJ_SX = ADJ_TX + 2, ADJ_SY = ADJ_TY + 2, ADJ_SZ = ADJ_TZ + 2;
constexpr int ADJ_SRC = ADJ_SX * ADJ_SY * ADJ_SZ;  // sources of a tile at window radius 1
static_assert(ADJ_TX * ADJ_TY * ADJ_TZ == 256, "one destination per thread");

__host__ __device__ inline unsigned adjTiles(int gx, int gy, int gz) {
  return (unsigned)(((gx + ADJ_TX - 1) / ADJ_TX) * ((gy + ADJ_TY - 1) / ADJ_TY) * ((gz + ADJ_TZ - 1) / ADJ_TZ));
}
// Destination cell of this thread and the tile origin.
__device__ __forceinline__ void adjTileCell(const UGrid& g, int& X0, int& Y0, int& Z0, int& x, int& y, int& z) {
  const int ntx = (g.gx + ADJ_TX - 1) / ADJ_TX, nty = (g.gy + ADJ_TY - 1) / ADJ_TY;
  const int b = (int)blockIdx.x;
  X0 = (b % ntx) * ADJ_TX;
  Y0 = ((b / ntx) % nty) * ADJ_TY;
  Z0 = (b / (ntx * nty)) * ADJ_TZ;
  const int t = (int)threadIdx.x;
  x = X0 + t % ADJ_TX;
  y = Y0 + (t / ADJ_TX) % ADJ_TY;
  z = Z0 + t / (ADJ_TX * ADJ_TY);
}
// Base-corner offset (ubase(p) - source coordinate, in [-2, 1] at radius 1) packed per axis.
__device__ __forceinline__ int adjPack(int ox, int oy, int oz) { return (ox + 8) | ((oy + 8) << 4) | ((oz + 8) << 8); }
// Corner of the source's sample that is destination (x,y,z), or -1.
__device__ __forceinline__ int adjCorner(int packed, int xi, int yi, int zi, int x, int y, int z) {
  const int bx = x - (xi + (packed & 15) - 8), by = y - (yi + ((packed >> 4) & 15) - 8),
            bz = z - (zi + ((packed >> 8) & 15) - 8);
  if ((unsigned)bx > 1u || (unsigned)by > 1u || (unsigned)bz > 1u) return -1;
  return (bx << 2) | (by << 1) | bz;
}
// Region index of source (xi,yi,zi) in the tile grown by one cell.
__device__ __forceinline__ int adjSlot(int X0, int Y0, int Z0, int xi, int yi, int zi) {
  return ((zi - Z0 + 1) * ADJ_SY + (yi - Y0 + 1)) * ADJ_SX + (xi - X0 + 1);
}
// Grown-tile coordinates of slot k (the tile grown by one cell).
__device__ __forceinline__ void adjSlotCell(int X0, int Y0, int Z0, int k, int& xi, int& yi, int& zi) {
  xi = X0 - 1 + k % ADJ_SX;
  yi = Y0 - 1 + (k / ADJ_SX) % ADJ_SY;
  zi = Z0 - 1 + k / (ADJ_SX * ADJ_SY);
}
__device__ __forceinline__ bool adjOwn(int X0, int Y0, int Z0, int xi, int yi, int zi) {
  return xi >= X0 && xi < X0 + ADJ_TX && yi >= Y0 && yi < Y0 + ADJ_TY && zi >= Z0 && zi < Z0 + ADJ_TZ;
}

// Transposed trilinear gathers as scatters (window radius 1, CFL < 1): every cell of a
// workgroup's tile samples its own foot once and adds w_c lam to the corner entries of an LDS
// tile of the tile grown by one cell, in 64-bit fixed point (fixAdd: integer LDS atomics, order
// independent); the same sample gives the cell's foot term. The tile is then written out: its
// own cells stored, the ring into the tile's slots of a halo buffer that k_ua_halo_gather adds
// to the neighbouring tiles' cells in a fixed order. A launch whose states move a foot further
// (window radius > 1) gathers per destination over the window instead and leaves its halo slots
// zero.

// Adjoint of advectDensity under velocity vel1 = u_{k+1}: lq0 = A^T lq (and lt0 = A^T lt with
// the source), sources masked inside the emitting sphere; lu += foot terms (den0/tem0 = state k).
template <bool FLUIDITY, bool SRC>
__global__ void __launch_bounds__(256) k_ua_advect_density(dcg_params P, UGrid g, const float* __restrict__ den0,
                                                           const float* __restrict__ tem0,
                                                           const float3* __restrict__ vel1,
                                                           const float* __restrict__ lq, const float* __restrict__ lt,
                                                           float* __restrict__ lq0, float* __restrict__ lt0,
                                                           float3* __restrict__ lu, float* __restrict__ hq,
                                                           float* __restrict__ ht, const int* __restrict__ win,
                                                           int* __restrict__ errFlag) {
  __shared__ unsigned long long tq[ADJ_SRC];
  __shared__ unsigned long long tt[SRC ? ADJ_SRC : 1];
  __shared__ uint32_t mq, mt;
  const int R = windowRadius(win, errFlag);
  int X0, Y0, Z0, x, y, z;
  adjTileCell(g, X0, Y0, Z0, x, y, z);
  const bool inside = x < g.gx && y < g.gy && z < g.gz;
  const size_t j = inside ? uidx(g, x, y, z) : 0;
  const bool masked = SRC && inSource(P, (float)x + .5f, (float)y + .5f, (float)z + .5f);
  const float gq = inside ? lq[j] : 0.f, gt = (SRC && inside) ? lt[j] : 0.f;
  float* hqt = hq + (size_t)blockIdx.x * ADJ_SRC;
  float* htt = SRC ? ht + (size_t)blockIdx.x * ADJ_SRC : nullptr;
  if (R <= 1) {  // uniform over the launch
    for (int k = threadIdx.x; k < ADJ_SRC; k += 256) {
      tq[k] = 0ull;
      if (SRC) tt[k] = 0ull;
    }
    if (threadIdx.x == 0) mq = mt = 0u;
    __syncthreads();
    if (inside && !masked) {
      ldsMax(&mq, gq);
      if (SRC) ldsMax(&mt, gt);
    }
    __syncthreads();
    const FixScale sq = fixScale(mq), sc = fixScale(SRC ? mt : 0u);
    if (inside && !masked && (gq != 0.f || gt != 0.f)) {
      float px, py, pz;
      ufoot(P, x, y, z, vel1[j], px, py, pz);
      UFootD F;
      ufootD<FLUIDITY>(g, px, py, pz, F);
      if (!F.s.zero) {
        float qc[8], tc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const int cx = F.s.x0 + ((c >> 2) & 1), cy = F.s.y0 + ((c >> 1) & 1), cz = F.s.z0 + (c & 1);
          qc[c] = densityBC(P, den0[F.s.i[c]], cx, cy, cz, 1);
          if (SRC) tc[c] = temperatureBC(P, tem0[F.s.i[c]], cx, cy, cz, 1);
          if (uinDomain(g, cx, cy, cz)) {  // the BC passes the stored value through
            const int k = adjSlot(X0, Y0, Z0, cx, cy, cz);
            fixAdd(&tq[k], F.s.w[c] * gq, sq);
            if (SRC) fixAdd(&tt[k], F.s.w[c] * gt, sc);
          }
        }
        float3 d = ufootGrad(P, F, qc, gq);
        if (SRC) d = f3add(d, ufootGrad(P, F, tc, gt));
        lu[j] = f3add(lu[j], d);
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ADJ_SRC; k += 256) {
      int xi, yi, zi;
      adjSlotCell(X0, Y0, Z0, k, xi, yi, zi);
      if (!uinDomain(g, xi, yi, zi)) continue;
      const float v = fixGet(tq[k], sq), vt = SRC ? fixGet(tt[k], sc) : 0.f;
      if (adjOwn(X0, Y0, Z0, xi, yi, zi)) {
        const size_t i = uidx(g, xi, yi, zi);
        lq0[i] = v;
        if (SRC) lt0[i] = vt;
      } else {
        hqt[k] = v;
        if (SRC) htt[k] = vt;
      }
    }
    return;
  }
  for (int k = threadIdx.x; k < ADJ_SRC; k += 256) {
    hqt[k] = 0.f;
    if (SRC) htt[k] = 0.f;
  }
  if (!inside) return;
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
  if (masked) return;
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
                                                            float* __restrict__ lt0, float3* __restrict__ hu,
                                                            const int* __restrict__ win, int* __restrict__ errFlag) {
  __shared__ unsigned long long tu[3][ADJ_SRC];
  __shared__ float3 selfs[256];
  __shared__ uint32_t mu;
  const int R = windowRadius(win, errFlag);
  int X0, Y0, Z0, x, y, z;
  adjTileCell(g, X0, Y0, Z0, x, y, z);
  const bool inside = x < g.gx && y < g.gy && z < g.gz;
  const size_t j = inside ? uidx(g, x, y, z) : 0;
  const float3 lam = inside ? lus[j] : make_float3(0.f, 0.f, 0.f);
  float3* hut = hu + (size_t)blockIdx.x * ADJ_SRC;
  // foot term of cell j (every path)
  auto footTerm = [&](const UFootD& F) {
    float vx[8], vy[8], vz[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float3 v = velocityBC(P, vel0[F.s.i[c]], F.s.x0 + ((c >> 2) & 1), F.s.y0 + ((c >> 1) & 1), F.s.z0 + (c & 1), 1);
      vx[c] = v.x;
      vy[c] = v.y;
      vz[c] = v.z;
    }
    float3 acc = ufootGrad(P, F, vx, lam.x);
    acc = f3add(acc, ufootGrad(P, F, vy, lam.y));
    return f3add(acc, ufootGrad(P, F, vz, lam.z));
  };
  if (inside && BUOY) {
    lq0[j] = lq0[j] + (-P.dt * P.buoy_alpha) * lam.y;
    lt0[j] = lt0[j] + (P.dt * P.buoy_beta) * lam.y;
  }
  if (R <= 1) {  // uniform over the launch
    for (int k = threadIdx.x; k < ADJ_SRC; k += 256) tu[0][k] = tu[1][k] = tu[2][k] = 0ull;
    if (threadIdx.x == 0) mu = 0u;
    selfs[threadIdx.x] = make_float3(0.f, 0.f, 0.f);
    __syncthreads();
    if (inside) ldsMax(&mu, fmaxf(fabsf(lam.x), fmaxf(fabsf(lam.y), fabsf(lam.z))));
    __syncthreads();
    const FixScale su = fixScale(mu);
    if (inside && (lam.x != 0.f || lam.y != 0.f || lam.z != 0.f)) {
      float px, py, pz;
      ufoot(P, x, y, z, vel0[j], px, py, pz);
      UFootD F;
      ufootD<FLUIDITY>(g, px, py, pz, F);
      if (!F.s.zero) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const int cx = F.s.x0 + ((c >> 2) & 1), cy = F.s.y0 + ((c >> 1) & 1), cz = F.s.z0 + (c & 1);
          if (!uinDomain(g, cx, cy, cz)) continue;
          const int k = adjSlot(X0, Y0, Z0, cx, cy, cz);
          fixAdd(&tu[0][k], lam.x * F.s.w[c], su);
          fixAdd(&tu[1][k], lam.y * F.s.w[c], su);
          fixAdd(&tu[2][k], lam.z * F.s.w[c], su);
        }
        selfs[threadIdx.x] = footTerm(F);
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ADJ_SRC; k += 256) {
      int xi, yi, zi;
      adjSlotCell(X0, Y0, Z0, k, xi, yi, zi);
      if (!uinDomain(g, xi, yi, zi)) continue;
      float3 v = make_float3(fixGet(tu[0][k], su), fixGet(tu[1][k], su), fixGet(tu[2][k], su));
      if (adjOwn(X0, Y0, Z0, xi, yi, zi)) {
        const int t = ((zi - Z0) * ADJ_TY + (yi - Y0)) * ADJ_TX + (xi - X0);  // adjTileCell's thread
        lu0[uidx(g, xi, yi, zi)] = f3add(v, selfs[t]);
      } else {
        hut[k] = v;
      }
    }
    return;
  }
  for (int k = threadIdx.x; k < ADJ_SRC; k += 256) hut[k] = make_float3(0.f, 0.f, 0.f);
  if (!inside) return;
  float3 acc = make_float3(0.f, 0.f, 0.f);
  for (int zi = max(z - R, 0); zi <= min(z + R, g.gz - 1); zi++)
    for (int yi = max(y - R, 0); yi <= min(y + R, g.gy - 1); yi++)
      for (int xi = max(x - R, 0); xi <= min(x + R, g.gx - 1); xi++) {
        const size_t i = uidx(g, xi, yi, zi);
        const float w = utransWeight<FLUIDITY>(P, g, xi, yi, zi, vel0[i], x, y, z);
        if (w != 0.f) acc = f3add(acc, f3mul(lus[i], w));
      }
  float px, py, pz;
  ufoot(P, x, y, z, vel0[j], px, py, pz);
  UFootD F;
  ufootD<FLUIDITY>(g, px, py, pz, F);
  if (!F.s.zero) acc = f3add(acc, footTerm(F));
  lu0[j] = acc;
}

__device__ __forceinline__ float uaddv(float a, float b) { return a + b; }
__device__ __forceinline__ float3 uaddv(float3 a, float3 b) { return f3add(a, b); }
// Adds to each cell the ring entries of the neighbouring tiles that stand for it (the tiles on
// the sides where the cell lies on its tile's border), in a fixed order.
template <class T>
__global__ void __launch_bounds__(256) k_ua_halo_gather(UGrid g, T* __restrict__ dst, const T* __restrict__ halo) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  int x, y, z;
  ucoord(j, g.gx, g.gy, x, y, z);
  const int ntx = (g.gx + ADJ_TX - 1) / ADJ_TX, nty = (g.gy + ADJ_TY - 1) / ADJ_TY, ntz = (g.gz + ADJ_TZ - 1) / ADJ_TZ;
  const int tx = x / ADJ_TX, ty = y / ADJ_TY, tz = z / ADJ_TZ;
  const int bx = x % ADJ_TX == 0 ? -1 : (x % ADJ_TX == ADJ_TX - 1 ? 1 : 0);
  const int by = y % ADJ_TY == 0 ? -1 : (y % ADJ_TY == ADJ_TY - 1 ? 1 : 0);
  const int bz = z % ADJ_TZ == 0 ? -1 : (z % ADJ_TZ == ADJ_TZ - 1 ? 1 : 0);
  if (!(bx | by | bz)) return;
  T v = dst[j];
  for (int m = 1; m < 8; m++) {
    const int dx = m & 4 ? bx : 0, dy = m & 2 ? by : 0, dz = m & 1 ? bz : 0;
    if ((dx == 0 && (m & 4)) || (dy == 0 && (m & 2)) || (dz == 0 && (m & 1))) continue;
    const int ux = tx + dx, uy = ty + dy, uz = tz + dz;
    if (ux < 0 || uy < 0 || uz < 0 || ux >= ntx || uy >= nty || uz >= ntz) continue;
    const int k = adjSlot(ux * ADJ_TX, uy * ADJ_TY, uz * ADJ_TZ, x, y, z);
    v = uaddv(v, halo[(size_t)((uz * nty + uy) * ntx + ux) * ADJ_SRC + k]);
  }
  dst[j] = v;
}

// G^T: lambda of the level-0 pressure from lambda of the velocity after applyPressure.
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_ua_grad_T(dcg_params P, UGrid g, const float3* __restrict__ lu,
                                                   float* __restrict__ lp) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  int x, y, z;
  ucoord(j, g.gx, g.gy, x, y, z);
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
__device__ __forceinline__ void uaJacobiCell(const float* __restrict__ in, float* __restrict__ out,
                                             float* __restrict__ ld, const UDims& L, float alpha, size_t i) {
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  const size_t pl = (size_t)L.w * L.h;
  const size_t il = x > 0 ? i - 1 : i, ir = x < L.w - 1 ? i + 1 : i;
  const size_t id = y > 0 ? i - L.w : i, iu = y < L.h - 1 ? i + L.w : i;
  const size_t ib = z > 0 ? i - pl : i, iF = z < L.d - 1 ? i + pl : i;
  const float* p = in + L.off;
  out[L.off + i] = (p[il] + p[ir] + p[id] + p[iu] + p[ib] + p[iF]) / 6.f;
  ld[L.off + i] = ld[L.off + i] - alpha * (p[i] / 6.f);
}
__global__ void __launch_bounds__(256) k_ua_jacobi(const float* __restrict__ in, float* __restrict__ out,
                                                   float* __restrict__ ld, UDims L, float alpha) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uaJacobiCell(in, out, ld, L, alpha, i);
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
// The trilinear weights are separable, so a workgroup stages the fine region of a tile of
// PT_X x PT_Y x PT_Z coarse cells in LDS and contracts it one axis at a time (x, then y, then z):
// about 1.3 global reads per fine cell instead of 8.
constexpr int PT_X = 8, PT_Y = 8, PT_Z = 4;
constexpr int PF_X = 2 * PT_X + 2, PF_Y = 2 * PT_Y + 2, PF_Z = 2 * PT_Z + 2;
static_assert(PT_X * PT_Y * PT_Z == 256, "the z pass maps one coarse cell of the tile to each thread");
__host__ __device__ inline unsigned prolTTiles(const UDims& Lc) {
  return (unsigned)(((Lc.w + PT_X - 1) / PT_X) * ((Lc.h + PT_Y - 1) / PT_Y) * ((Lc.d + PT_Z - 1) / PT_Z));
}
__global__ void __launch_bounds__(256) k_ua_prolongate_T(float* __restrict__ lp, UDims L, UDims Lc) {
  __shared__ float fine[PF_Z * PF_Y * PF_X];
  __shared__ float ax[PF_Z * PF_Y * PT_X];
  __shared__ float ay[PF_Z * PT_Y * PT_X];
  const int ntx = (Lc.w + PT_X - 1) / PT_X, nty = (Lc.h + PT_Y - 1) / PT_Y;
  const int b = (int)blockIdx.x;
  const int X0 = (b % ntx) * PT_X, Y0 = ((b / ntx) % nty) * PT_Y, Z0 = (b / (ntx * nty)) * PT_Z;
  const int fx0 = 2 * X0 - 1, fy0 = 2 * Y0 - 1, fz0 = 2 * Z0 - 1;
  for (int k = threadIdx.x; k < PF_Z * PF_Y * PF_X; k += 256) {
    const int x = fx0 + k % PF_X, y = fy0 + (k / PF_X) % PF_Y, z = fz0 + k / (PF_X * PF_Y);
    const bool in = x >= 0 && y >= 0 && z >= 0 && x < L.w && y < L.h && z < L.d;
    fine[k] = in ? lp[L.off + ((size_t)z * L.h + y) * L.w + x] : 0.f;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < PF_Z * PF_Y * PT_X; k += 256) {  // x: fine x -> coarse X
    const int cx = k % PT_X, fy = (k / PT_X) % PF_Y, fz = k / (PT_X * PF_Y);
    const int X = X0 + cx;
    float s = 0.f;
    for (int d = 0; d < 4; d++) {
      const int x = 2 * X - 1 + d;
      if (x < 0 || x >= L.w) continue;
      const float w = prolW(x, L.w, X);
      if (w != 0.f) s += w * fine[(fz * PF_Y + fy) * PF_X + (x - fx0)];
    }
    ax[k] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < PF_Z * PT_Y * PT_X; k += 256) {  // y
    const int cx = k % PT_X, cy = (k / PT_X) % PT_Y, fz = k / (PT_X * PT_Y);
    const int Y = Y0 + cy;
    float s = 0.f;
    for (int d = 0; d < 4; d++) {
      const int y = 2 * Y - 1 + d;
      if (y < 0 || y >= L.h) continue;
      const float w = prolW(y, L.h, Y);
      if (w != 0.f) s += w * ax[(fz * PF_Y + (y - fy0)) * PT_X + cx];
    }
    ay[k] = s;
  }
  __syncthreads();
  {  // z, one coarse cell per thread
    const int k = threadIdx.x;
    const int cx = k % PT_X, cy = (k / PT_X) % PT_Y, cz = k / (PT_X * PT_Y);
    const int X = X0 + cx, Y = Y0 + cy, Z = Z0 + cz;
    if (X >= Lc.w || Y >= Lc.h || Z >= Lc.d) return;
    float s = 0.f;
    for (int d = 0; d < 4; d++) {
      const int z = 2 * Z - 1 + d;
      if (z < 0 || z >= L.d) continue;
      const float w = prolW(z, L.d, Z);
      if (w != 0.f) s += w * ay[((z - fz0) * PT_Y + cy) * PT_X + cx];
    }
    lp[Lc.off + ((size_t)Z * Lc.h + Y) * Lc.w + X] = s;
  }
}

// Restriction transposed: every child cell of level L gains 1/8 of its parent's lambda (Lp).
__device__ __forceinline__ void uaRestrictTCell(float* __restrict__ ld, const UDims& L, const UDims& Lp, size_t i) {
  int x, y, z;
  ucoord(i, L.w, L.h, x, y, z);
  const size_t par = Lp.off + ((size_t)(z / 2) * Lp.h + (y / 2)) * Lp.w + (x / 2);
  ld[L.off + i] = ld[L.off + i] + .125f * ld[par];
}
__global__ void __launch_bounds__(256) k_ua_restrict_T(float* __restrict__ ld, UDims L, UDims Lp) {
  const size_t n = (size_t)L.w * L.h * L.d;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uaRestrictTCell(ld, L, Lp, i);
}

// Prolongation transposed for one coarse cell I of level Lc (direct 4x4x4 gather; used on the
// small levels handled by k_ua_coarse).
__device__ __forceinline__ void uaProlongTCell(float* __restrict__ lp, const UDims& L, const UDims& Lc, size_t I) {
  int X, Y, Z;
  ucoord(I, Lc.w, Lc.h, X, Y, Z);
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

// The adjoint of k_u_coarse's part of project() on the small levels (from Ls up), in one
// workgroup with the levels' lambdas in LDS: reversed sweeps and transposed prolongation going
// up, the top level's reversed sweeps, transposed restriction coming down to Ls.
__global__ void __launch_bounds__(1024) k_ua_coarse(float* __restrict__ lp, float* __restrict__ ltp,
                                                    float* __restrict__ ld, ULevels D, int L, int Ls, int topPairs,
                                                    int levelPairs, float dx) {
  __shared__ float sp[kCoarseLds], st[kCoarseLds], sd[kCoarseLds];
  __shared__ ULevels S;  // the level table in LDS: a per-thread copy indexed by level goes to scratch
  __shared__ size_t sTotal;
  if (threadIdx.x == 0) coarseLocal(D, L, Ls, S, sTotal);
  __syncthreads();
  const size_t total = sTotal;
  const size_t base = D.d[Ls].off;
  for (size_t i = threadIdx.x; i < total; i += blockDim.x) {
    sp[i] = lp[base + i];
    st[i] = ltp[base + i];
    sd[i] = ld[base + i];
  }
  __syncthreads();
  auto sweepsT = [&](int l, int pairs) {
    const int scale = 1 << l;
    const float alpha = dx * dx * (float)scale * (float)scale;
    for (int p = 0; p < pairs; p++) {
      for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uaJacobiCell(sp, st, sd, S.d[l], alpha, i);
      __syncthreads();
      for (size_t i = threadIdx.x; i < ucells(S.d[l]); i += blockDim.x) uaJacobiCell(st, sp, sd, S.d[l], alpha, i);
      __syncthreads();
    }
  };
  for (int l = Ls; l <= L - 2; l++) {
    sweepsT(l, levelPairs);
    for (size_t I = threadIdx.x; I < ucells(S.d[l + 1]); I += blockDim.x) uaProlongTCell(sp, S.d[l], S.d[l + 1], I);
    __syncthreads();
  }
  sweepsT(L - 1, topPairs);
  for (int l = L - 1; l >= Ls + 1; l--) {
    for (size_t i = threadIdx.x; i < ucells(S.d[l - 1]); i += blockDim.x) uaRestrictTCell(sd, S.d[l - 1], S.d[l], i);
    __syncthreads();
  }
  for (size_t i = threadIdx.x; i < total; i += blockDim.x) {
    lp[base + i] = sp[i];
    ltp[base + i] = st[i];
    ld[base + i] = sd[i];
  }
}

// D^T: lu += (d div / d u)^T ld on level 0.
template <bool FLUIDITY>
__global__ void __launch_bounds__(256) k_ua_div_T(dcg_params P, UGrid g, const float* __restrict__ ld,
                                                  float3* __restrict__ lu) {
  const size_t n = (size_t)g.gx * g.gy * g.gz;
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  int x, y, z;
  ucoord(j, g.gx, g.gy, x, y, z);
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
