// dcgrid_adjoint.hpp — gfx950 adjoint (reverse-mode) kernels of the DCGrid step on a frozen
// block topology: advectVelocity, the Jacobi projection (project / projectLocal), advectDensity,
// with the block tables, neighbour descriptors and fluidity fixed over the rollout. Extension:
// the reference has no derivatives. Included by dcgrid.hip after the forward kernels, whose
// boundary conditions, foot point and parent stencil they reuse.
//
// Every forward kernel `out = f(in)` has a transposed kernel `lam_in += (df/din)^T lam_out`
// (then `lam_out = 0` where the output overwrote its previous value), run in reverse order.
// Two forms, by what the forward stencil reads:
//  * Stencils between the cells of one level (Jacobi half-sweeps, divergence, gradient) are
//    symmetric between same-level neighbours: a same-level neighbour's apron entry toward a
//    block is that block's cell. Their transposes are gathers: a workgroup stages its block's
//    adjoint values and those of its same-level face neighbours in LDS (as the forward kernels
//    stage their inputs) and each cell sums what its neighbours' forward stencils took from it,
//    and stores it. What a cell's stencil reads elsewhere goes the other way: an out-of-domain
//    apron entry (the block's own clamped cell; the reference's form for placed blocks is not
//    always the cell itself) is added in LDS to that cell of the block, and a coarser ghost cell
//    gets a global FP32 atomic (global_atomic_add_f32). Levels are launched from the finest up,
//    so a cell is never stored and atomically updated in the same launch.
//  * The semi-Lagrangian advection reads arbitrary corners: its transpose scatters. A workgroup
//    takes one block (four at BW 4, as the forward stencils), adds its cells' contributions into
//    an LDS tile of the block's apron with LDS atomics, and writes the tile out with few global
//    atomics (interior stored, same-level halo entries through a halo buffer and a gather; see
//    the section below).
//  * The prolongation reads a fixed parent stencil: its transpose is a fixed 4-tap gather per
//    axis over the child block's cells, staged in LDS; the parent region entries it produces are
//    shared between neighbouring child blocks and are added with global atomics.
//  * Restriction hands a parent cell's lam/8 to each of its 8 children (a gather).
// Atomic sums are formed in an order that can change from run to run: the gradient is
// reproducible to FP32 rounding, not bitwise.
#pragma once

namespace dcg {

__device__ __forceinline__ void atomicAddF(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomicAddF3(float3* p, float3 v) {
  float* q = reinterpret_cast<float*>(p);
  if (v.x != 0.f) unsafeAtomicAdd(q, v.x);
  if (v.y != 0.f) unsafeAtomicAdd(q + 1, v.y);
  if (v.z != 0.f) unsafeAtomicAdd(q + 2, v.z);
}
// LDS tile entries (apron coordinates (i,j,k) in [0,BW+1]^3).
template <int BW>
__device__ __forceinline__ int tix(int i, int j, int k) {
  constexpr int A = Geo<BW>::A;
  return (i * A + j) * A + k;
}
__device__ __forceinline__ void ldsAdd(float* p, float v) { atomicAdd(p, v); }
__device__ __forceinline__ void ldsAdd3(float3* p, float3 v) {
  float* q = reinterpret_cast<float*>(p);
  atomicAdd(q, v.x);
  atomicAdd(q + 1, v.y);
  atomicAdd(q + 2, v.z);
}
template <int BW, class T>
__device__ __forceinline__ void zeroTile(T* tile, int t) {
  for (int e = t; e < Geo<BW>::A3; e += Geo<BW>::BV) tile[e] = zeroOf<T>();
}
// Derivative of a normalised fluidity-weighted trilinear sample with respect to the advecting
// velocity: from the fractions d of the foot in the sampled level's cells and the corner
// fluidities f, the unnormalised weights f_c t_c (sum W), d t_c / d d_a, and
// k = d d / d u = -dt rdx / scale of the sampled level.
struct FootGrad {
  float f[8], t[8], dtd[8][3], W, k;
  __device__ __forceinline__ void set(const float d[3], float kk) {
    W = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bit[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
      float tau[3];
#pragma unroll
      for (int a = 0; a < 3; a++) tau[a] = bit[a] ? d[a] : 1.f - d[a];
      t[c] = tau[0] * tau[1] * tau[2];
      dtd[c][0] = (bit[0] ? 1.f : -1.f) * tau[1] * tau[2];
      dtd[c][1] = (bit[1] ? 1.f : -1.f) * tau[0] * tau[2];
      dtd[c][2] = (bit[2] ? 1.f : -1.f) * tau[0] * tau[1];
      W += f[c] * t[c];
    }
    k = kk;
  }
  // d(sample)/d(u) of the interpolant of the BC-applied corner values vc.
  __device__ __forceinline__ float3 grad(const float vc[8]) const {
    float val = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) val += f[c] * t[c] * vc[c];
    val /= W;
    float out[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float num = 0.f, den = 0.f;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        num += f[c] * vc[c] * dtd[c][a];
        den += f[c] * dtd[c][a];
      }
      out[a] = k * (num - val * den) / W;
    }
    return make_float3(out[0], out[1], out[2]);
  }
};
// Fractions of foot (px,py,pz) in the cells of a level of scale `scale` (the sample's).
__device__ __forceinline__ void footFractions(float px, float py, float pz, int scale, float d[3]) {
  const float si = 1.f / (float)scale;
  const float x = px * si - .5f, y = py * si - .5f, z = pz * si - .5f;
  d[0] = x - floorf(x);
  d[1] = y - floorf(y);
  d[2] = z - floorf(z);
}

// Foot sample through the block tables (the sample of CrossSample, same expressions: block sb
// containing the foot, corners through sb's apron from (ai,aj,ak), normalised fluidity-weighted
// weights) with its derivative.
template <int BW, bool FLUIDITY>
struct FootD {
  uint32_t sb, i[8];
  int x0, y0, z0, scale;
  bool zero;
  float w[8];
  FootGrad G;
  __device__ __forceinline__ void init(const DG& g, const dcg_params& P, float px, float py, float pz) {
    const int ix = clampi((int)floorf(px), 0, g.gx - 1);
    const int iy = clampi((int)floorf(py), 0, g.gy - 1);
    const int iz = clampi((int)floorf(pz), 0, g.gz - 1);
    sb = findBlock<BW>(g, ix, iy, iz, 0);
    const int sl = g.lev[sb];
    const int3 sp = g.pos[sb];
    scale = 1 << sl;
    const float scaleInv = 1.f / (float)scale;
    const float x = px * scaleInv - .5f, y = py * scaleInv - .5f, z = pz * scaleInv - .5f;
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    const float dx = x - xf, dy = y - yf, dz = z - zf;
    x0 = (int)xf;
    y0 = (int)yf;
    z0 = (int)zf;
    const int ai = clampi(x0 + 1 - sp.x, 0, BW), aj = clampi(y0 + 1 - sp.y, 0, BW), ak = clampi(z0 + 1 - sp.z, 0, BW);
#pragma unroll
    for (int c = 0; c < 8; c++) i[c] = apronCell<BW>(g, sb, ai + ((c >> 2) & 1), aj + ((c >> 1) & 1), ak + (c & 1));
    const float Dx = 1.f - dx, Dy = 1.f - dy, Dz = 1.f - dz;
    float ww[8];
    ww[0] = Dx * Dy * Dz;
    ww[1] = Dx * Dy * dz;
    ww[2] = Dx * dy * Dz;
    ww[3] = Dx * dy * dz;
    ww[4] = dx * Dy * Dz;
    ww[5] = dx * Dy * dz;
    ww[6] = dx * dy * Dz;
    ww[7] = dx * dy * dz;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      G.f[c] = FLUIDITY ? (float)g.flu[i[c]] : 1.f;
      if (FLUIDITY) ww[c] = G.f[c] * ww[c];
    }
    const float wAcc = ww[0] + ww[1] + ww[2] + ww[3] + ww[4] + ww[5] + ww[6] + ww[7];
    const float wInv = 1.f / wAcc;
#pragma unroll
    for (int c = 0; c < 8; c++) w[c] = ww[c] * wInv;
    zero = wAcc < 1e-6f;
    const float d[3] = {dx, dy, dz};
    G.set(d, -(P.dt * P.rdx) * scaleInv);
  }
  __device__ __forceinline__ void corner(int c, int& x, int& y, int& z) const {
    x = x0 + ((c >> 2) & 1);
    y = y0 + ((c >> 1) & 1);
    z = z0 + (c & 1);
  }
};
// The derivative of a TileSample (corner fluidities from the block's fluidity tile).
template <int BW, bool FLUIDITY>
__device__ __forceinline__ void tileFootGrad(const TileSample<BW, FLUIDITY>& ts, const dcg_params& P, float px, float py,
                                             float pz, const float* fluTile, FootGrad& G) {
#pragma unroll
  for (int c = 0; c < 8; c++) G.f[c] = FLUIDITY ? fluTile[ts.a0 + TileSample<BW, FLUIDITY>::off(c)] : 1.f;
  float d[3];
  footFractions(px, py, pz, ts.scale, d);
  G.set(d, -(P.dt * P.rdx) / (float)ts.scale);
}

// restrictAll reversed for one level (launched from the coarsest restriction down): each parent
// cell's adjoint is shared equally by its 8 children, then cleared. The groups of k_restrict.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_da_restrict_T(DGBase g, int level, T* __restrict__ lam) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, GPB = BV / 8;
  const uint32_t groups = g.ctl[CTL_LOADS + level] * GPB;
  for (uint32_t gi = blockIdx.x * 256 + threadIdx.x; gi < groups; gi += gridDim.x * 256) {
    const uint32_t b = g.offs[level] + gi / GPB;
    const uint32_t ps = g.par[b];
    if (ps == NFU) continue;
    const uint32_t grp = gi % GPB;
    const uint32_t parentCell = (ps / 8) * BV + ((ps % 8) << SBB) + grp;
    const T G = mulv(lam[parentCell], .125f);
    T* dst = lam + (size_t)b * BV + 8 * grp;
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = addv(dst[i], G);
    lam[parentCell] = zeroOf<T>();
  }
}

// ---- scatter tiles of the transposed advection
// A workgroup's block b adds the contributions of its cells' feet whose corners lie in b's apron
// box on b's level (the foot's block on b's level: then b's apron entry at a corner's position
// is the corner cell) into an LDS tile of the apron. The tile is then written out without
// global atomics where possible: the interior is stored to the destination (the block owns those
// cells), entries toward a same-level neighbour go to the block's slot of a halo buffer that
// the neighbour gathers from (k_da_halo_gather), and entries toward a coarser block are added
// atomically to a far buffer, as are the corners of feet outside the apron box. Out-of-domain
// entries stay empty (corners outside the domain take the BCs' constants). The gather then adds,
// per cell, the far buffer (clearing it for the next use) and the halo entries that the
// same-level neighbours hold for the cell.
__device__ __forceinline__ bool nonzero(float v) { return v != 0.f; }
__device__ __forceinline__ bool nonzero(float3 v) { return v.x != 0.f || v.y != 0.f || v.z != 0.f; }
__device__ __forceinline__ void ldsAddT(float* p, float v) { ldsAdd(p, v); }
__device__ __forceinline__ void ldsAddT(float3* p, float3 v) { ldsAdd3(p, v); }
__device__ __forceinline__ void atomicAddT(float* p, float v) { atomicAddF(p, v); }
__device__ __forceinline__ void atomicAddT(float3* p, float3 v) { atomicAddF3(p, v); }
template <int BW>
__device__ __forceinline__ bool interiorEntry(int i, int j, int k) {
  return i > 0 && i <= BW && j > 0 && j <= BW && k > 0 && k <= BW;
}
// The apronCell rule with block b's 27 neighbour descriptors staged in LDS (nd27).
template <int BW>
__device__ __forceinline__ uint32_t apronCellStaged(const int4* nd27, uint32_t b, int ai, int aj, int ak) {
  constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
  const int dx = ai == 0 ? -1 : (ai == BW + 1 ? 1 : 0);
  const int dy = aj == 0 ? -1 : (aj == BW + 1 ? 1 : 0);
  const int dz = ak == 0 ? -1 : (ak == BW + 1 ? 1 : 0);
  if ((dx | dy | dz) == 0) return b * BV + loc<BB>(ai - 1, aj - 1, ak - 1);
  const int4 d = nd27[dirIndex(dx, dy, dz)];
  const uint32_t sh = ndShift(d);
  if (sh == ND_OUT) {
    const int ci = clampi(ai, 1, BW), cj = clampi(aj, 1, BW), ck = clampi(ak, 1, BW);
    if (d.y) return b * BV + loc<BB>(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1));
    return b * BV + loc<BB>(ci - 1, cj - 1, ck - 1);
  }
  return ndBlock(d) * BV + loc<BB>((ai - 1 - d.y) >> sh, (aj - 1 - d.z) >> sh, (ak - 1 - d.w) >> sh);
}
// forApron (one cell per thread) with the descriptors from LDS.
template <int BW, class Op>
__device__ __forceinline__ void forApronStaged(const int4* nd27, uint32_t b, int t, Op op) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3;
  op(apronIndexOfLocal<BW>(t), b * BV + t);
  for (int a = t; a < A3; a += BV) {
    const int ai = a / (A * A), aj = (a / A) % A, ak = a % A;
    if (ai > 0 && ai <= BW && aj > 0 && aj <= BW && ak > 0 && ak <= BW) continue;
    op(a, apronCellStaged<BW>(nd27, b, ai, aj, ak));
  }
}
template <int BW, class T, class Val>
__device__ __forceinline__ void storeTile(const int4* nd27, uint32_t b, Val val, T* __restrict__ dst, T* __restrict__ halo,
                                          T* __restrict__ far, int t) {
  constexpr int A = Geo<BW>::A, A3 = Geo<BW>::A3, BV = Geo<BW>::BV, BB = Geo<BW>::BB;
  for (int e = t; e < A3; e += BV) {
    const int i = e / (A * A), j = (e / A) % A, k = e % A;
    if (interiorEntry<BW>(i, j, k)) {
      dst[b * BV + loc<BB>(i - 1, j - 1, k - 1)] = val(e);
      continue;
    }
    const int dx = i == 0 ? -1 : (i == BW + 1 ? 1 : 0), dy = j == 0 ? -1 : (j == BW + 1 ? 1 : 0);
    const int dz = k == 0 ? -1 : (k == BW + 1 ? 1 : 0);
    const uint32_t sh = ndShift(nd27[dirIndex(dx, dy, dz)]);
    if (sh == ND_OUT) continue;
    const T v = val(e);
    if (sh == 0)
      halo[(size_t)b * A3 + e] = v;
    else if (nonzero(v))
      atomicAddT(&far[apronCellStaged<BW>(nd27, b, i, j, k)], v);
  }
}
// Whether all corners of foot F lie in the apron box of block b (origin p0, level lb), and the
// tile entry of corner 000 there.
template <int BW, bool FLUIDITY>
__device__ __forceinline__ bool cornersInApron(const DGBase& g, const FootD<BW, FLUIDITY>& F, int3 p0, int lb, int& e0) {
  if (g.lev[F.sb] != lb) return false;
  const int ri = F.x0 + 1 - p0.x, rj = F.y0 + 1 - p0.y, rk = F.z0 + 1 - p0.z;
  e0 = tix<BW>(ri, rj, rk);
  return ri >= 0 && ri <= BW && rj >= 0 && rj <= BW && rk >= 0 && rk <= BW;
}
template <int BW>
__device__ __forceinline__ int cornerOffset(int c) {
  return tix<BW>((c >> 2) & 1, (c >> 1) & 1, c & 1);
}

// The gather of the scatter tiles (see above), one thread per cell of the allocated blocks.
template <int BW, class T>
__global__ void __launch_bounds__(256) k_da_halo_gather(DGBase g, T* __restrict__ dst, const T* __restrict__ halo,
                                                        T* __restrict__ far) {
  constexpr int BV = Geo<BW>::BV, A3 = Geo<BW>::A3, BB = Geo<BW>::BB;
  const uint32_t total = slotTotal(g, -1) * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
    int bl;
    const uint32_t b = slotBlock(g, -1, i / BV, bl);
    const int t = (int)(i % BV);
    const uint32_t cell = b * BV + t;
    T v = dst[cell];
    const T f = far[cell];
    if (nonzero(f)) {
      v = addv(v, f);
      far[cell] = zeroOf<T>();
    }
    int lx, ly, lz;
    locPos<BB>((uint32_t)t, lx, ly, lz);
    const int bx = lx == 0 ? -1 : (lx == BW - 1 ? 1 : 0), by = ly == 0 ? -1 : (ly == BW - 1 ? 1 : 0);
    const int bz = lz == 0 ? -1 : (lz == BW - 1 ? 1 : 0);
    if (bx | by | bz) {
      for (int m = 1; m < 8; m++) {  // the face, edge and corner directions the cell touches
        const int dx = m & 4 ? bx : 0, dy = m & 2 ? by : 0, dz = m & 1 ? bz : 0;
        if ((dx == 0 && (m & 4)) || (dy == 0 && (m & 2)) || (dz == 0 && (m & 1))) continue;
        const int4 d = g.nd[b * 27 + dirIndex(dx, dy, dz)];
        if (ndShift(d) != 0) continue;
        const int e = tix<BW>(lx - dx * BW + 1, ly - dy * BW + 1, lz - dz * BW + 1);
        v = addv(v, halo[(size_t)ndBlock(d) * A3 + e]);
      }
    }
    dst[cell] = v;
  }
}

// advectDensity reversed (leaf cells): lq, lt (after) scatter into lq0, lt0 (before; every
// allocated cell is written through the tiles, see above); the foot terms of both go into lu,
// the adjoint of the advecting velocity. Cells inside the sphere source were overwritten by its
// constants and pass nothing on. As in the forward kernel, the block's apron values (and
// fluidity) are staged in LDS and a foot in the block itself is sampled from them (TileSample);
// other feet sample through the block tables.
template <int BW, bool FLUIDITY, bool SRC>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_advect_density_T(DG g, dcg_params P, const float* __restrict__ den0,
                                                                     const float* __restrict__ tem0,
                                                                     const float3* __restrict__ vel1,
                                                                     const float* __restrict__ lq,
                                                                     const float* __restrict__ lt,
                                                                     float* __restrict__ lq0, float* __restrict__ lt0,
                                                                     float3* __restrict__ lu, float* __restrict__ hq,
                                                                     float* __restrict__ ht, float* __restrict__ fq,
                                                                     float* __restrict__ ft) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW;
  __shared__ unsigned long long qtiles[BPW][A3];  // scatter tiles (fixed point)
  __shared__ unsigned long long ttiles[SRC ? BPW : 1][SRC ? A3 : 1];
  __shared__ uint32_t qmax[BPW], tmax[BPW];        // largest |lam| scattered, as float bits
  __shared__ float dstage[BPW][A3];  // staged apron values
  __shared__ float tstage[SRC ? BPW : 1][SRC ? A3 : 1];
  __shared__ float fstage[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ uint32_t chis[BPW][8];
  __shared__ int4 nd27s[BPW][27];
  const bool sphere = P.source_mode == DCG_SOURCE_SPHERE_BUOYANT;
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();  // the previous block's tiles are written out
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);
    unsigned long long* qt = qtiles[c.sub];
    unsigned long long* tt = SRC ? ttiles[c.sub] : nullptr;
    float* ds = dstage[c.sub];
    float* ts_ = SRC ? tstage[c.sub] : nullptr;
    float* fs = FLUIDITY ? fstage[c.sub] : nullptr;
    uint32_t* chi8 = chis[c.sub];
    int4* nd27 = nd27s[c.sub];
    for (int e = c.t; e < A3; e += BV) {
      qt[e] = 0ull;
      if (SRC) tt[e] = 0ull;
    }
    if (c.t == 0) {
      qmax[c.sub] = 0u;
      tmax[c.sub] = 0u;
    }
    int3 p0 = make_int3(0, 0, 0);
    if (c.valid) {
      p0 = g.pos[c.b];
      if (c.t < 8) chi8[c.t] = g.chi[8 * c.b + c.t];
      if (c.t < 27) nd27[c.t] = g.nd[c.b * 27 + c.t];
    }
    const uint32_t cell = c.b * BV + c.t;
    const float gq = c.valid ? lq[cell] : 0.f, gt = (SRC && c.valid) ? lt[cell] : 0.f;
    __syncthreads();
    if (c.valid) {
      forApronStaged<BW>(nd27, c.b, c.t, [&](int a, uint32_t src) {
        ds[a] = den0[src];
        if (SRC) ts_[a] = tem0[src];
        if (FLUIDITY) fs[a] = g.flu[src];
      });
      ldsMax(&qmax[c.sub], gq);
      if (SRC) ldsMax(&tmax[c.sub], gt);
    }
    __syncthreads();
    const FixScale qs = fixScale(qmax[c.sub]), tsc = fixScale(SRC ? tmax[c.sub] : 0u);
    if (c.valid && chi8[c.t >> SBB] == NFU && (gq != 0.f || gt != 0.f)) {
      float px, py, pz;
      int x, y, z;
      footPoint<BW>(P, p0, c.level, c.t, vel1[cell], px, py, pz, x, y, z);
      const float sc = (float)(1 << c.level);
      const bool fixed = SRC && sphere && inSource(P, ((float)x + .5f) * sc, ((float)y + .5f) * sc, ((float)z + .5f) * sc);
      TileSample<BW, FLUIDITY> ts;
      float3 du = make_float3(0.f, 0.f, 0.f);
      if (fixed) {
      } else if (ts.init(g, px, py, pz, chi8, p0, c.level, fs)) {
        if (!ts.zero) {
          FootGrad G;
          tileFootGrad<BW, FLUIDITY>(ts, P, px, py, pz, fs, G);
          float qc[8], tc[8];
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int a = ts.a0 + TileSample<BW, FLUIDITY>::off(k);
            const int cx = ts.x0 + ((k >> 2) & 1), cy = ts.y0 + ((k >> 1) & 1), cz = ts.z0 + (k & 1);
            qc[k] = densityBC(P, ds[a], cx, cy, cz, ts.scale);
            if (SRC) tc[k] = temperatureBC(P, ts_[a], cx, cy, cz, ts.scale);
            if (!outside(P, cx, cy, cz, ts.scale)) {  // the BC passes the stored value through
              fixAdd(&qt[a], ts.w[k] * gq, qs);
              if (SRC) fixAdd(&tt[a], ts.w[k] * gt, tsc);
            }
          }
          du = f3mul(G.grad(qc), gq);
          if (SRC) du = f3add(du, f3mul(G.grad(tc), gt));
        }
      } else {
        FootD<BW, FLUIDITY> F;
        F.init(g, P, px, py, pz);
        if (!F.zero) {
          int e0 = 0;
          const bool own = cornersInApron<BW>(g, F, p0, c.level, e0);
          float qc[8], tc[8];
#pragma unroll
          for (int k = 0; k < 8; k++) {
            int cx, cy, cz;
            F.corner(k, cx, cy, cz);
            const uint32_t i = F.i[k];
            qc[k] = densityBC(P, den0[i], cx, cy, cz, F.scale);
            if (SRC) tc[k] = temperatureBC(P, tem0[i], cx, cy, cz, F.scale);
            if (!outside(P, cx, cy, cz, F.scale)) {
              if (own) {
                fixAdd(&qt[e0 + cornerOffset<BW>(k)], F.w[k] * gq, qs);
                if (SRC) fixAdd(&tt[e0 + cornerOffset<BW>(k)], F.w[k] * gt, tsc);
              } else {
                atomicAddF(&fq[i], F.w[k] * gq);
                if (SRC) atomicAddF(&ft[i], F.w[k] * gt);
              }
            }
          }
          du = f3mul(F.G.grad(qc), gq);
          if (SRC) du = f3add(du, f3mul(F.G.grad(tc), gt));
        }
      }
      if (nonzero(du)) lu[cell] = f3add(lu[cell], du);
    }
    __syncthreads();
    if (c.valid) {
      storeTile<BW>(nd27, c.b, [&](int e) { return fixGet(qt[e], qs); }, lq0, hq, fq, c.t);
      if (SRC) storeTile<BW>(nd27, c.b, [&](int e) { return fixGet(tt[e], tsc); }, lt0, ht, ft, c.t);
    }
  }
}

// advectVelocity reversed (leaf cells): lu (after) scatters into lu0 (before; every allocated
// cell is written through the tiles), with the foot terms of the three components at the cell
// itself; the buoyancy term adds to lq0, lt0 (the density and temperature adjoints of the
// step's start). Feet in the block itself sample the staged apron, as in k_da_advect_density_T.
template <int BW, bool FLUIDITY, bool BUOY>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_advect_velocity_T(DG g, dcg_params P,
                                                                      const float3* __restrict__ vel0,
                                                                      const float3* __restrict__ lu,
                                                                      float3* __restrict__ lu0,
                                                                      float* __restrict__ lq0,
                                                                      float* __restrict__ lt0,
                                                                      float3* __restrict__ hu,
                                                                      float3* __restrict__ fu) {
  constexpr int BV = Geo<BW>::BV, SBB = Geo<BW>::SBB, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW;
  __shared__ unsigned long long tiles[BPW][3][A3];  // scatter tiles (fixed point), per component
  __shared__ uint32_t umax[BPW];                    // largest |lam| component scattered
  __shared__ float3 selfs[BPW][BV];                 // foot terms, by local cell
  __shared__ float3 vstage[BPW][A3];  // staged apron velocity
  __shared__ float fstage[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ uint32_t chis[BPW][8];
  __shared__ int4 nd27s[BPW][27];
  const uint32_t total = slotTotal(g, -1);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, -1, base, total);
    unsigned long long(*ut)[A3] = tiles[c.sub];
    float3* vs = vstage[c.sub];
    float* fs = FLUIDITY ? fstage[c.sub] : nullptr;
    uint32_t* chi8 = chis[c.sub];
    int4* nd27 = nd27s[c.sub];
    for (int e = c.t; e < A3; e += BV) ut[0][e] = ut[1][e] = ut[2][e] = 0ull;
    if (c.t == 0) umax[c.sub] = 0u;
    selfs[c.sub][c.t] = make_float3(0.f, 0.f, 0.f);
    int3 p0 = make_int3(0, 0, 0);
    if (c.valid) {
      p0 = g.pos[c.b];
      if (c.t < 8) chi8[c.t] = g.chi[8 * c.b + c.t];
      if (c.t < 27) nd27[c.t] = g.nd[c.b * 27 + c.t];
    }
    const uint32_t cell = c.b * BV + c.t;
    const float3 lam = c.valid ? lu[cell] : make_float3(0.f, 0.f, 0.f);
    __syncthreads();
    if (c.valid) {
      forApronStaged<BW>(nd27, c.b, c.t, [&](int a, uint32_t src) {
        vs[a] = vel0[src];
        if (FLUIDITY) fs[a] = g.flu[src];
      });
      ldsMax(&umax[c.sub], fmaxf(fabsf(lam.x), fmaxf(fabsf(lam.y), fabsf(lam.z))));
    }
    __syncthreads();
    const FixScale us = fixScale(umax[c.sub]);
    auto addCorner = [&](int e, float w) {
      fixAdd(&ut[0][e], lam.x * w, us);
      fixAdd(&ut[1][e], lam.y * w, us);
      fixAdd(&ut[2][e], lam.z * w, us);
    };
    if (c.valid && chi8[c.t >> SBB] == NFU) {
      if (BUOY) {
        lq0[cell] += -P.dt * P.buoy_alpha * lam.y;
        lt0[cell] += P.dt * P.buoy_beta * lam.y;
      }
      if (nonzero(lam)) {
        float px, py, pz;
        int x, y, z;
        footPoint<BW>(P, p0, c.level, c.t, vs[apronIndexOfLocal<BW>(c.t)], px, py, pz, x, y, z);
        TileSample<BW, FLUIDITY> ts;
        float3 self = make_float3(0.f, 0.f, 0.f);
        if (ts.init(g, px, py, pz, chi8, p0, c.level, fs)) {
          if (!ts.zero) {
            FootGrad G;
            tileFootGrad<BW, FLUIDITY>(ts, P, px, py, pz, fs, G);
            float vc[3][8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
              const int a = ts.a0 + TileSample<BW, FLUIDITY>::off(k);
              const int cx = ts.x0 + ((k >> 2) & 1), cy = ts.y0 + ((k >> 1) & 1), cz = ts.z0 + (k & 1);
              const float3 v = velocityBC(P, vs[a], cx, cy, cz, ts.scale);
              vc[0][k] = v.x;
              vc[1][k] = v.y;
              vc[2][k] = v.z;
              if (!outside(P, cx, cy, cz, ts.scale)) addCorner(a, ts.w[k]);
            }
            self = f3mul(G.grad(vc[0]), lam.x);
            self = f3add(self, f3mul(G.grad(vc[1]), lam.y));
            self = f3add(self, f3mul(G.grad(vc[2]), lam.z));
          }
        } else {
          FootD<BW, FLUIDITY> F;
          F.init(g, P, px, py, pz);
          if (!F.zero) {
            int e0 = 0;
            const bool own = cornersInApron<BW>(g, F, p0, c.level, e0);
            float vc[3][8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
              int cx, cy, cz;
              F.corner(k, cx, cy, cz);
              const uint32_t i = F.i[k];
              const float3 v = velocityBC(P, vel0[i], cx, cy, cz, F.scale);
              vc[0][k] = v.x;
              vc[1][k] = v.y;
              vc[2][k] = v.z;
              if (!outside(P, cx, cy, cz, F.scale)) {
                if (own)
                  addCorner(e0 + cornerOffset<BW>(k), F.w[k]);
                else
                  atomicAddF3(&fu[i], f3mul(lam, F.w[k]));
              }
            }
            self = f3mul(F.G.grad(vc[0]), lam.x);
            self = f3add(self, f3mul(F.G.grad(vc[1]), lam.y));
            self = f3add(self, f3mul(F.G.grad(vc[2]), lam.z));
          }
        }
        selfs[c.sub][c.t] = self;
      }
    }
    __syncthreads();
    if (c.valid)
      storeTile<BW>(nd27, c.b, [&](int e) {
        float3 v = make_float3(fixGet(ut[0][e], us), fixGet(ut[1][e], us), fixGet(ut[2][e], us));
        constexpr int A = Geo<BW>::A;
        const int i = e / (A * A), j = (e / A) % A, k = e % A;
        if (interiorEntry<BW>(i, j, k)) v = f3add(v, selfs[c.sub][loc<Geo<BW>::BB>(i - 1, j - 1, k - 1)]);
        return v;
      }, lu0, hu, fu, c.t);
  }
}

// Local index of a cell of its block.
template <int BW>
__device__ __forceinline__ int localOf(uint32_t cell) {
  return (int)(cell % (uint32_t)Geo<BW>::BV);
}
// Tile entries of the six face neighbours of tile entry a: -x, +x, -y, +y, -z, +z.
template <int BW>
__device__ __forceinline__ void faceEntries(int a, int e[6]) {
  constexpr int A = Geo<BW>::A;
  e[0] = a - A * A;
  e[1] = a + A * A;
  e[2] = a - A;
  e[3] = a + A;
  e[4] = a - 1;
  e[5] = a + 1;
}
// Kind of the apron entry of a cell in face direction k: inside the block, a same-level
// neighbour block, a coarser block or outside the domain.
enum : int { NB_INSIDE = 0, NB_SAME = 1, NB_COARSER = 2, NB_OUT = 3 };

// The neighbour descriptors of a block's six faces (-x, +x, -y, +y, -z, +z), staged in LDS once
// per block so that the stencil transposes resolve their face entries with integer arithmetic
// (the apronCell rule) instead of a descriptor load per entry.
template <int BW>
struct Faces {
  const int4* d;
  uint32_t b;
  static __device__ __forceinline__ int dirOf(int k) {
    const int s = k & 1 ? 1 : -1, ax = k >> 1;
    return dirIndex(ax == 0 ? s : 0, ax == 1 ? s : 0, ax == 2 ? s : 0);
  }
  static __device__ __forceinline__ void stage(const DGBase& g, int4* d, uint32_t b, int t) {
    if (t < 6) d[t] = g.nd[b * 27 + dirOf(t)];
  }
  __device__ __forceinline__ int kindOf(int k) const {
    const uint32_t sh = ndShift(d[k]);
    return sh == ND_OUT ? NB_OUT : (sh == 0 ? NB_SAME : NB_COARSER);
  }
  // Cell of apron entry (ai,aj,ak) on face k (the apronCell rule for that face's descriptor).
  __device__ __forceinline__ uint32_t cellOf(int k, int ai, int aj, int ak) const {
    constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
    const int4 q = d[k];
    const uint32_t sh = ndShift(q);
    if (sh == ND_OUT) {
      const int ci = clampi(ai, 1, BW), cj = clampi(aj, 1, BW), ck = clampi(ak, 1, BW);
      if (q.y) return b * BV + loc<BB>(ci & (BW - 1), cj & (BW - 1), ck & (BW - 1));
      return b * BV + loc<BB>(ci - 1, cj - 1, ck - 1);
    }
    return ndBlock(q) * BV + loc<BB>((ai - 1 - q.y) >> sh, (aj - 1 - q.z) >> sh, (ak - 1 - q.w) >> sh);
  }
  // Kind and cell of the face-k neighbour of local cell (lx,ly,lz).
  __device__ __forceinline__ int neighbour(int lx, int ly, int lz, int k, uint32_t& cell) const {
    constexpr int BB = Geo<BW>::BB, BV = Geo<BW>::BV;
    const int s = k & 1 ? 1 : -1, ax = k >> 1;
    const int nx = lx + (ax == 0 ? s : 0), ny = ly + (ax == 1 ? s : 0), nz = lz + (ax == 2 ? s : 0);
    if ((unsigned)nx < (unsigned)BW && (unsigned)ny < (unsigned)BW && (unsigned)nz < (unsigned)BW) {
      cell = b * BV + loc<BB>(nx, ny, nz);
      return NB_INSIDE;
    }
    cell = cellOf(k, nx + 1, ny + 1, nz + 1);
    return kindOf(k);
  }
  // Stage the block's values v(cell) and those of its same-level face neighbours into an apron
  // tile; face entries toward coarser blocks or outside the domain are zero. The stencil
  // transposes gather from these entries: a same-level neighbour's apron entry toward this
  // block is this block's cell, so its contribution is picked up here instead of scattered.
  template <class T, class F>
  __device__ __forceinline__ void stageSameLevel(T* tile, int t, F v) const {
    constexpr int BV = Geo<BW>::BV, FACE = Geo<BW>::FACE;
    tile[apronIndexOfLocal<BW>(t)] = v(b * BV + t);
    for (int fc = t; fc < FACE; fc += BV) {
      int ai, aj, ak;
      faceCell<BW>(fc, ai, aj, ak);
      const int k = fc / (BW * BW);
      tile[tix<BW>(ai, aj, ak)] = ndShift(d[k]) == 0 ? v(cellOf(k, ai, aj, ak)) : zeroOf<T>();
    }
  }
};

// applyPressure reversed, one level (launched from the finest level up): leaf cell c did
// v -= alpha (f_r (p_r - p_c) + f_l (p_c - p_l)) per axis, so lp[r] -= alpha f_r lam_c,
// lp[l] += alpha f_l lam_c and lp[c] += alpha (f_r - f_l) lam_c. A cell gathers the terms of its
// same-level neighbours (their lam is 0 unless they are leaves: the restriction's transpose
// cleared the refined cells), adds its own term and the terms its block's out-of-domain entries
// point at it (LDS), and stores the sum; a term for a coarser cell is added to it atomically
// (coarser levels are launched later, so no cell is stored and atomically updated in one launch).
template <int BW, bool FLUIDITY>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_grad_T(DG g, dcg_params P, int level, const float3* __restrict__ lu,
                                                           float* __restrict__ lp) {
  constexpr int BV = Geo<BW>::BV, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, BB = Geo<BW>::BB;
  __shared__ float3 ltiles[BPW][A3];
  __shared__ float ftiles[FLUIDITY ? BPW : 1][FLUIDITY ? A3 : 1];
  __shared__ float accs[BPW][BV];
  __shared__ int4 nds[BPW][6];
  const float alpha = .5f * P.rdx / (float)(1 << level);
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    float3* lt = ltiles[c.sub];
    float* ft = FLUIDITY ? ftiles[c.sub] : nullptr;
    float* acc = accs[c.sub];
    acc[c.t] = 0.f;
    if (c.valid) Faces<BW>::stage(g, nds[c.sub], c.b, c.t);
    __syncthreads();
    const Faces<BW> fa{nds[c.sub], c.b};
    if (c.valid) {
      fa.stageSameLevel(lt, c.t, [&](uint32_t i) { return lu[i]; });
      if (FLUIDITY)
        forApron<BW, 1>(g, c.b, c.t, [&](int a, uint32_t src) { ft[a] = g.flu[src]; });
    }
    __syncthreads();
    const int a = apronIndexOfLocal<BW>(c.t);
    const uint32_t cell = c.b * BV + c.t;
    float own = 0.f;
    if (c.valid) {
      float fw[6];
      tileFluidity<BW, FLUIDITY>(ft, a, fw);
      const float3 lam3 = lt[a];
      const float lam[3] = {lam3.x, lam3.y, lam3.z};
#pragma unroll
      for (int ax = 0; ax < 3; ax++) own += alpha * fw[2 * ax + 1] * lam[ax] - alpha * fw[2 * ax] * lam[ax];
      // terms of this cell for its coarser and out-of-domain neighbours
      if (lam3.x != 0.f || lam3.y != 0.f || lam3.z != 0.f) {
        int lx, ly, lz;
        locPos<BB>((uint32_t)c.t, lx, ly, lz);
#pragma unroll
        for (int k = 0; k < 6; k++) {
          uint32_t nb;
          const int kind = fa.neighbour(lx, ly, lz, k, nb);
          if (kind != NB_COARSER && kind != NB_OUT) continue;
          const float v = (k & 1 ? -alpha : alpha) * fw[k] * lam[k >> 1];
          if (kind == NB_OUT)
            ldsAdd(&acc[localOf<BW>(nb)], v);
          else
            atomicAddF(&lp[nb], v);
        }
      }
    }
    __syncthreads();
    if (c.valid) {
      // terms of the same-level neighbours for this cell: from the left neighbour (this cell is
      // its right one) -alpha f lam.x, from the right neighbour +alpha f lam.x, per axis
      int e[6];
      faceEntries<BW>(a, e);
      const float f = FLUIDITY ? ft[a] : 1.f;
      const float gx = lt[e[1]].x - lt[e[0]].x, gy = lt[e[3]].y - lt[e[2]].y, gz = lt[e[5]].z - lt[e[4]].z;
      lp[cell] = lp[cell] + (own + alpha * f * (gx + gy + gz) + acc[c.t]);
    }
  }
}

// jacobi(level, in, out) reversed: out = (sum of the 6 apron cells of in - alpha div) / 6 on
// every cell of the level's blocks, so lin[apron] += lout/6, ld -= alpha lout/6, lout = 0. A cell
// gathers lout/6 of its same-level neighbours (the stencil between cells of one level is
// symmetric) and of the out-of-domain entries of its block that point at it (LDS), and adds
// lout/6 to the coarser cells its apron reads atomically. lout is not cleared: within a level's
// run of sweeps the first launch adds to lin (accumulate = 1: lin holds what finer levels and
// later stages passed to it), the later ones store (their lin was the previous launch's lout,
// cleared in the forward sense); coarser cells always accumulate.
template <int BW>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_jacobi_T(DGBase g, int level, const float* __restrict__ lout,
                                                             float* __restrict__ lin, float* __restrict__ ld,
                                                             float alpha, int accumulate) {
  constexpr int BV = Geo<BW>::BV, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, BB = Geo<BW>::BB;
  __shared__ float tiles[BPW][A3];
  __shared__ float accs[BPW][BV];
  __shared__ int4 nds[BPW][6];
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    float* tile = tiles[c.sub];
    float* acc = accs[c.sub];
    acc[c.t] = 0.f;
    if (c.valid) Faces<BW>::stage(g, nds[c.sub], c.b, c.t);
    __syncthreads();
    const Faces<BW> fa{nds[c.sub], c.b};
    if (c.valid) fa.stageSameLevel(tile, c.t, [&](uint32_t i) { return lout[i] / 6.f; });
    __syncthreads();
    const int a = apronIndexOfLocal<BW>(c.t);
    const uint32_t cell = c.b * BV + c.t;
    if (c.valid) {
      const float G = tile[a];
      if (G != 0.f) {
        ld[cell] += -alpha * G;
        int lx, ly, lz;
        locPos<BB>((uint32_t)c.t, lx, ly, lz);
#pragma unroll
        for (int k = 0; k < 6; k++) {
          uint32_t nb;
          const int kind = fa.neighbour(lx, ly, lz, k, nb);
          if (kind == NB_OUT)
            ldsAdd(&acc[localOf<BW>(nb)], G);
          else if (kind == NB_COARSER)
            atomicAddF(&lin[nb], G);
        }
      }
    }
    __syncthreads();
    if (c.valid) {
      int e[6];
      faceEntries<BW>(a, e);
      const float sum = tile[e[0]] + tile[e[1]] + tile[e[2]] + tile[e[3]] + tile[e[4]] + tile[e[5]] + acc[c.t];
      lin[cell] = accumulate ? lin[cell] + sum : sum;
    }
  }
}

// k_da_jacobi_T on a dense level (l >= S): every face neighbour is a same-level cell or, outside
// the domain, the cell itself (dense blocks keep the clamped-own-cell form), so the transposed
// sweep is the forward stencil applied to lout/6, staged like the forward dense sweep (faces from
// the block index, no descriptor loads, no coarser ghosts).
template <int BW>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_jacobi_dense_T(DGBase g, int level, const float* __restrict__ lout,
                                                                   float* __restrict__ lin, float* __restrict__ ld,
                                                                   float alpha, int accumulate) {
  constexpr int BV = Geo<BW>::BV, A = Geo<BW>::A, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW;
  __shared__ float tiles[BPW][A3];
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    float* tile = tiles[c.sub];
    if (c.valid) stageScalarDense<BW>(g, level, lout, tile, c.b, c.t);
    __syncthreads();
    if (!c.valid) continue;
    const int a = apronIndexOfLocal<BW>(c.t);
    const uint32_t cell = c.b * BV + c.t;
    const float G = tile[a] / 6.f;
    if (G != 0.f) ld[cell] += -alpha * G;
    const float sum = (tile[a - A * A] + tile[a + A * A] + tile[a - A] + tile[a + A] + tile[a - 1] + tile[a + 1]) / 6.f;
    lin[cell] = accumulate ? lin[cell] + sum : sum;
  }
}

// prolongate(level) reversed: a child cell's pre = (27 q0 + 9 (q1+q2+q3) + 3 (q4+q5+q6) + q7) / 64
// of its parent stencil, i.e. per axis 3/4 of its parent cell and 1/4 of the parent cell's
// neighbour on its side. Transposed per axis, parent region entry r (the (BW/2 + 2)^3 region of
// the parent's apron that k_prolongate stages, r = 1 + x/2 for child x) receives from children
// 2r-3, 2r-2, 2r-1, 2r with weights 1/4, 3/4, 3/4, 1/4. A workgroup stages its child block's
// lp/64, gathers each region entry over that 4x4x4 window (no LDS atomics), adds the nonzero
// entries to the parent cells they stand for (global atomics: neighbouring child blocks share
// region entries) and clears lp on its cells.
template <int BW>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_prolongate_T(DGBase g, int level, float* __restrict__ lp) {
  constexpr int BV = Geo<BW>::BV, BB = Geo<BW>::BB, BPW = Geo<BW>::BPW, H = BW / 2 + 2, H3 = H * H * H;
  __shared__ float stage[BPW][BV];  // lp/64 by block-local raster position (x BW + y) BW + z
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    const uint32_t ps = c.valid ? g.par[c.b] : NFU;
    float* st = stage[c.sub];
    int lx, ly, lz;
    locPos<BB>((uint32_t)c.t, lx, ly, lz);
    if (ps != NFU) {
      const uint32_t cell = c.b * BV + c.t;
      const float v = lp[cell];
      st[(lx * BW + ly) * BW + lz] = v / 64.f;
      if (v != 0.f) lp[cell] = 0.f;
    }
    __syncthreads();
    if (ps == NFU) continue;
    const int3 p0 = g.pos[c.b];
    const int ox = (p0.x / 2) % BW, oy = (p0.y / 2) % BW, oz = (p0.z / 2) % BW;
    for (int e = c.t; e < H3; e += BV) {
      const int r[3] = {e / (H * H), (e / H) % H, e % H};
      float v = 0.f;
      for (int a = 0; a < 4; a++) {
        const int x = 2 * r[0] - 3 + a;
        if (x < 0 || x >= BW) continue;
        const float wx = (a == 1 || a == 2) ? 3.f : 1.f;
        for (int b2 = 0; b2 < 4; b2++) {
          const int y = 2 * r[1] - 3 + b2;
          if (y < 0 || y >= BW) continue;
          const float wxy = wx * ((b2 == 1 || b2 == 2) ? 3.f : 1.f);
          for (int d = 0; d < 4; d++) {
            const int z = 2 * r[2] - 3 + d;
            if (z < 0 || z >= BW) continue;
            v += wxy * ((d == 1 || d == 2) ? 3.f : 1.f) * st[(x * BW + y) * BW + z];
          }
        }
      }
      if (v != 0.f) atomicAddF(&lp[apronCell<BW>(g, ps / 8, ox + r[0], oy + r[1], oz + r[2])], v);
    }
  }
}

// calcDivergence reversed, one level (launched from the finest level up): leaf cell c had
// div = alpha (f_r v_r.x - f_l v_l.x + ...) with the neighbours' BC-applied velocities, so each
// in-domain neighbour's component gets +-alpha f lam_div. A cell gathers these terms from its
// same-level neighbours (lam_div is 0 on refined cells after the restriction's transpose) and
// adds to its own velocity adjoint; a term for a coarser neighbour is added to it atomically
// (launched later, as in k_da_grad_T). Out-of-domain neighbours get constants: no term.
template <int BW, bool FLUIDITY>
__global__ void __launch_bounds__(Geo<BW>::WG) k_da_div_T(DG g, dcg_params P, int level, const float* __restrict__ ld,
                                                          float3* __restrict__ lu) {
  constexpr int BV = Geo<BW>::BV, A3 = Geo<BW>::A3, BPW = Geo<BW>::BPW, BB = Geo<BW>::BB;
  __shared__ float tiles[BPW][A3];
  __shared__ int4 nds[BPW][6];
  const float alpha = .5f * P.rdx / (float)(1 << level);
  const uint32_t total = slotTotal(g, level);
  for (uint32_t base = firstSlot<BW>(); base < total; base += slotStride<BW>()) {
    __syncthreads();
    const BlockCtx<BW> c = slotCtx<BW>(g, level, base, total);
    float* tile = tiles[c.sub];
    if (c.valid) Faces<BW>::stage(g, nds[c.sub], c.b, c.t);
    __syncthreads();
    const Faces<BW> fa{nds[c.sub], c.b};
    if (c.valid) fa.stageSameLevel(tile, c.t, [&](uint32_t i) { return alpha * ld[i]; });
    __syncthreads();
    if (!c.valid) continue;
    const int a = apronIndexOfLocal<BW>(c.t);
    const uint32_t cell = c.b * BV + c.t;
    int e[6];
    faceEntries<BW>(a, e);
    const float G = tile[a];
    if (G != 0.f) {
      int lx, ly, lz;
      locPos<BB>((uint32_t)c.t, lx, ly, lz);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        uint32_t nb;
        if (fa.neighbour(lx, ly, lz, k, nb) != NB_COARSER) continue;
        const float f = FLUIDITY ? (float)g.flu[nb] : 1.f;
        float* comp = reinterpret_cast<float*>(&lu[nb]) + (k >> 1);
        atomicAddF(comp, (k & 1 ? 1.f : -1.f) * f * G);
      }
    }
    const float f = FLUIDITY ? (float)g.flu[cell] : 1.f;
    float3 u = lu[cell];
    u.x += f * (tile[e[0]] - tile[e[1]]);
    u.y += f * (tile[e[2]] - tile[e[3]]);
    u.z += f * (tile[e[4]] - tile[e[5]]);
    lu[cell] = u;
  }
}

// The saved state of the rollout: density, temperature (with a source) and velocity of the
// allocated blocks' cells, copied between the live fields and a state slot.
template <int BW>
__global__ void __launch_bounds__(256) k_da_copy_state(DGBase g, const float* __restrict__ d0, const float* __restrict__ t0,
                                                       const float3* __restrict__ v0, float* __restrict__ d1,
                                                       float* __restrict__ t1, float3* __restrict__ v1) {
  constexpr int BV = Geo<BW>::BV;
  const uint32_t total = slotTotal(g, -1) * BV;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
    int bl;
    const uint32_t cell = slotBlock(g, -1, i / BV, bl) * BV + i % BV;
    d1[cell] = d0[cell];
    if (t0) t1[cell] = t0[cell];
    v1[cell] = v0[cell];
  }
}

// Gradient entries of the cells of free blocks are zero (no step reads or defines those cells).
template <int BW>
__global__ void __launch_bounds__(256) k_da_mask_free(DGBase g, float* __restrict__ d, float3* __restrict__ v,
                                                      float* __restrict__ t, size_t C) {
  constexpr int BV = Geo<BW>::BV;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < C; i += (size_t)gridDim.x * 256) {
    if (g.lev[i / BV] != FREE) continue;
    d[i] = 0.f;
    v[i] = make_float3(0.f, 0.f, 0.f);
    if (t) t[i] = 0.f;
  }
}

}  // namespace dcg
